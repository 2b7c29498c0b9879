"""CPU oracle of the pattern preprocessing of util/processimg.cpp:11-109
(test infrastructure only).  numpy restatement: background subtraction
(add(d_sig, d_bkg, -1)), centroid of the pattern thresholded at half its
maximum (applyThreshold + findMiddle, src/post_gpu/misc.cc:27-42 with
multiplyx/multiplyy, src/gpu/cudaConfig.cu:1427-1437), crop
(cudaConfig.cu:348-358, float arithmetic of its offsets kept), shiftWave
(misc.cc:97-104 with multiplyShift, cudaConfig.cu:373-378) or shiftMiddle
(misc.cc:88-95), getReal and mergePixel (cudaConfig.cu:753-756,949-961).
FFTs in float64."""
import numpy as np

f32 = np.float32


def crop(src, r, c, sx=0, sy=0):
    R, C = src.shape
    x0, y0 = (R - r) // 2 + sx, (C - c) // 2 + sy
    return src[x0:x0 + r, y0:y0 + c].copy()


def centroid(s):
    """(midx, midy) in pixels as float32, from the thresholded image"""
    R, C = s.shape
    s32 = s.astype(f32)
    v = np.where(s32 > s32.max() / f32(2), s32, f32(0))
    wx = ((np.arange(R, dtype=f32) + f32(0.5)) / f32(R) - f32(0.5))[:, None]
    wy = ((np.arange(C, dtype=f32) + f32(0.5)) / f32(C) - f32(0.5))[None, :]
    sw = v.astype(np.float64).sum()
    sx = (v * wx).astype(np.float64).sum()
    sy = (v * wy).astype(np.float64).sum()
    cx, cy = (sx / sw, sy / sw) if sw != 0 else (0.0, 0.0)
    return f32(cx) * f32(R), f32(cy) * f32(C)


def shift_wave(z, shx, shy):
    n, m = z.shape
    F = np.fft.fftshift(np.fft.fft2(z))
    x = np.arange(n)[:, None]
    y = np.arange(m)[None, :]
    ph = -2 * np.pi * (shx * (x - n // 2) / n + shy * (y - m // 2) / m)
    return np.fft.ifft2(np.fft.ifftshift(F * np.exp(1j * ph)))


def shift_middle(z):
    F = np.abs(np.fft.fft2(np.fft.fftshift(z)))
    return np.fft.fftshift(np.fft.ifft2(F))


def merge_pixels(a, nm):
    r, c = a.shape
    return a[:r // nm * nm, :c // nm * nm].reshape(r // nm, nm, c // nm, nm).mean(axis=(1, 3))


def preprocess(sig, bkg=None, nmerge=1, centre=False, shift_row=0.0, shift_col=0.0, shift_mode=0):
    s = np.asarray(sig, f32)
    if bkg is not None:
        b = np.asarray(bkg, f32)
        if b.shape[0] != s.shape[0]:
            if b.shape[0] > s.shape[0]:
                b = crop(b, *s.shape)
            else:
                s = crop(s, *b.shape)
        s = s - b
    R, C = s.shape
    midx, midy = centroid(s)
    if centre:
        midx = f32(midx + f32(shift_row))
        midy = f32(midy - f32(shift_col))
    else:
        midx = midy = f32(0)
    step = nmerge * 4
    outrow = (R - int(abs(midx)) * 2) // step * step
    outcol = (C - int(abs(midy)) * 2) // step * step
    n = min(outrow, outcol)
    shx, shy = f32(int(midx)) - midx, f32(int(midy)) - midy
    sx = int(f32(R) * (midx / f32(R)))
    if sx + n // 2 > R // 2:
        sx = (R - n) // 2
    elif sx - n // 2 < -(R // 2):
        sx = (n - R) // 2
    sy = int(f32(C) * (midy / f32(C)))
    if sy + n // 2 > C // 2:
        sy = (R - n) // 2
    elif sy - n // 2 < -(C // 2):
        sy = (n - C) // 2
    z = crop(s, n, n, sx, sy).astype(np.complex128)
    z = shift_middle(z) if shift_mode == 1 else shift_wave(z, float(shx), float(shy))
    return merge_pixels(z.real, nmerge), (float(midx), float(midy))
