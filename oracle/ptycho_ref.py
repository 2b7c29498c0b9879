"""CPU oracle of the ptychographic scan-position loop (test infrastructure only).

A float64 numpy restatement of CDIxx's per-position update, run one
position after another in the caller's order, used to check the GPU sweep
(cdx_ptycho_sweep) and to pin the properties below.  Followed:

  util/ptycho.cpp:514-561   the loop: window, exit wave, FFT, projection with
                            bsnorm = 1/sqrt(N), IFFT, 1/sqrt(N), U = psi - psi',
                            object (and probe-step) update, window write-back
  src/gpu/ptycho.cu:17-30   getWindow (zero outside the object), updateWindow
                            (pixels outside the object are not written)
  src/gpu/ptycho.cu:58-71   rPIE, rPIE_step (ALPHA = 0.1, BETA = 0.5, :5-6)
  src/gpu/ptycho.cu:70-73   updateObject (stepsize 1)
  src/gpu/ptycho.cu:91-101  updateObjectAndProbeStep (objectStepsize 0.5; the
                            probe term is formed from the UPDATED object value)
  src/gpu/cdi.cu:351-388    applyModAccurate (4-argument form)
  src/gpu/cdi.cu:438-485    applyMod (5-argument form, bsnorm applied first)
  util/ptycho.cpp:328-347   createPattern without noise: I = |FFT(psi)|^2 / N

The FFTs are numpy's (unnormalised forward, as myFFT; the reference's
unnormalised IFFT is numpy's ifft2 times N).
"""
import numpy as np

ALPHA, BETA = 0.1, 0.5


def get_window(obj, sx, sy, n):
    """getWindow: w[x, y] = obj[x + sx, y + sy], zero outside the object"""
    w = np.zeros((n, n), dtype=np.complex128)
    R, C = obj.shape
    x0, x1 = max(sx, 0), min(sx + n, R)
    y0, y1 = max(sy, 0), min(sy + n, C)
    if x0 < x1 and y0 < y1:
        w[x0 - sx:x1 - sx, y0 - sy:y1 - sy] = obj[x0:x1, y0:y1]
    return w


def update_window(obj, sx, sy, w):
    """updateWindow: write w back where the window lies inside the object"""
    n = w.shape[0]
    R, C = obj.shape
    x0, x1 = max(sx, 0), min(sx + n, R)
    y0, y1 = max(sy, 0), min(sy + n, C)
    if x0 < x1 and y0 < y1:
        obj[x0:x1, y0:y1] = w[x0 - sx:x1 - sx, y0 - sy:y1 - sy]


def mod_accurate(F, I, bs, bsnorm, scale):
    """applyModAccurate(F, I, bs, bsnorm) on a whole array (cdi.cu:351-388)"""
    r = F * bsnorm
    mod2 = np.maximum(I, 0.0)
    src2 = np.abs(r) ** 2
    out = r.copy()
    blocked = (bs > 0.5) if bs is not None else np.zeros(I.shape, bool)
    keep = blocked | (src2 == 0) | ((mod2 >= 0.95 * scale) & (src2 >= mod2))
    s = np.sqrt(np.divide(mod2, src2, out=np.zeros_like(mod2), where=src2 > 0))
    out[~keep] = (s * r)[~keep]
    # src2 == 0: the reference leaves F unwritten (F = r / bsnorm there)
    z = (~blocked) & (src2 == 0)
    out[z] = F[z]
    return out


def mod_noise(F, I, bs, noise_level, bsnorm, scale, rcolor):
    """applyMod(F, I, bs, noiseLevel, bsnorm) on a whole array (cdi.cu:438-485)"""
    r = F * bsnorm
    mod2 = np.maximum(I, 0.0)
    src2 = np.abs(r) ** 2
    over = mod2 >= 0.95 * scale
    mod2 = np.where(over, np.maximum(mod2, src2), mod2)
    tol = np.sqrt(noise_level + mod2 * rcolor / scale) * scale / rcolor
    diff = mod2 - src2
    val = np.where(diff > tol, mod2 - tol, np.where(diff < -tol, mod2 + tol, mod2))
    s = np.sqrt(np.divide(np.maximum(val, 0), src2, out=np.zeros_like(val), where=src2 > 0))
    out = np.where(src2 == 0, val + 0j, s * r)
    if bs is not None:
        out = np.where(bs > 0.5, r, out)
    return out


def rpie(target, s, d, maxi, param, step):
    """rPIE (ptycho.cu:58-64): target - conj(s) d step / ((1-a)|s|^2 + a max(maxi, 1e-5))"""
    den = step / ((1 - param) * np.abs(s) ** 2 + param * max(maxi, 1e-5))
    return target - np.conj(s) * d * den


def rpie_step(s, d, maxi, param):
    """rPIE_step (ptycho.cu:65-71): -conj(s) d / ((1-a)|s|^2 + a max(maxi, 1e-5))"""
    den = 1.0 / ((1 - param) * np.abs(s) ** 2 + param * max(maxi, 1e-5))
    return -np.conj(s) * d * den


def position_update(obj, probe, I, bs, sx, sy, mode="accurate", noise_level=0, scale=1.0, rcolor=65535.0,
                    probe_max=None, obj_max=0.0, update_probe=False, obj_step=None, probe_step_size=0.0):
    """One position of ptycho.cpp:516-560 (no sub-pixel shift, no Fresnel
    phase, no position refinement).  Updates obj in place; returns the
    probe-step term (or None)."""
    n = probe.shape[0]
    N = n * n
    norm = 1.0 / np.sqrt(N)
    if probe_max is None:
        probe_max = float(np.max(np.abs(probe) ** 2))
    if obj_step is None:
        obj_step = 0.5 if update_probe else 1.0
    w = get_window(obj, sx, sy, n)
    psi = probe * w
    F = np.fft.fft2(psi)
    if mode == "accurate":
        F = mod_accurate(F, I, bs, norm, scale)
    else:
        F = mod_noise(F, I, bs, noise_level, norm, scale, rcolor)
    psi2 = np.fft.ifft2(F) * N * norm
    U = psi - psi2
    w_new = rpie(w, probe, U, probe_max, ALPHA, obj_step)
    term = rpie_step(w_new, U, obj_max, BETA) * probe_step_size if update_probe else None
    update_window(obj, sx, sy, w_new)
    return term


def sweep(obj, probe, patterns, bs, sx, sy, order, probe_step=None, **kw):
    """positions in `order`, one after another; obj (and probe_step) updated in place"""
    for i in order:
        term = position_update(obj, probe, patterns[i], bs, int(sx[i]), int(sy[i]), **kw)
        if term is not None:
            probe_step += term
    return obj


def simulate(obj, probe, sx, sy):
    """createPattern without noise (ptycho.cpp:328-347): |FFT(P w_i)|^2 / N"""
    n = probe.shape[0]
    return np.stack([np.abs(np.fft.fft2(probe * get_window(obj, int(a), int(b), n))) ** 2 / (n * n)
                     for a, b in zip(sx, sy)])


def data_residual(obj, probe, patterns, bs, sx, sy):
    """sum over positions and measured pixels of (|FFT(P w)|/sqrt(N) - sqrt(I))^2"""
    n = probe.shape[0]
    tot = 0.0
    for i in range(len(sx)):
        a = np.abs(np.fft.fft2(probe * get_window(obj, int(sx[i]), int(sy[i]), n))) / n
        m = np.ones(a.shape, bool) if bs is None else bs <= 0.5
        tot += float(np.sum((a - np.sqrt(np.maximum(patterns[i], 0)))[m] ** 2))
    return tot
