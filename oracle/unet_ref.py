"""CPU oracle of the learned warm-start network (test infrastructure only).

A numpy restatement of the U-Net the library trains on the GPU
(cdixx_amd/csrc/unet.hip), used to check its forward pass and its gradients.
The reference has no such network (its loop starts from random signs,
src/post_gpu/cdi.cc:222-223), so this oracle is written from the layer
definitions themselves:

  * activations NHWC; 3x3 'same' convolutions with zero padding (tap = 3 dy +
    dx, dy, dx in -1..1) or 1x1; weights w[co][k], k = tap * cin + ci, rows
    zero-padded to Kp; bias; ReLU;
  * per level two conv+ReLU, 2x2 average pooling down, nearest 2x upsampling
    and channel concatenation (upsampled first, skip second) up, 1x1 head
    with 8 outputs of which 0..2 are amplitude, phase, support logit;
  * loss = mean over pixels of (a - a*)^2 + 0.5 m* (1 - cos(p - p*)) +
    BCE(logit, m*).

With bf16=True every tensor the GPU stores in BF16 (weights, activations,
activation gradients) is rounded to BF16 (nearest even) at the same points,
and arithmetic is float64; with bf16=False everything is float64, which the
tests check against finite differences.
"""
import numpy as np


def round_bf16(x):
    """Round float values to the nearest BF16 (ties to even), as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


class UNetRef:
    def __init__(self, layers, size, levels, base, bf16=True):
        """layers: rows (cin, cout, taps, kp, w_off, b_off) as reported by the
        library (cdx_unet_layers)."""
        self.layers = [tuple(int(v) for v in r) for r in layers]
        self.size, self.levels, self.base, self.bf16 = size, levels, base, bf16

    def _r(self, x):
        return round_bf16(x) if self.bf16 else np.asarray(x, np.float64)

    # ---- parameters ------------------------------------------------------
    def unpack(self, flat):
        out = []
        for cin, cout, taps, kp, wo, bo in self.layers:
            w = np.asarray(flat[wo:wo + cout * kp], np.float64).reshape(cout, kp)[:, :taps * cin]
            out.append((w.reshape(cout, taps, cin), np.asarray(flat[bo:bo + cout], np.float64)))
        return out

    def pack_grads(self, grads, n):
        flat = np.zeros(n, np.float64)
        for (cin, cout, taps, kp, wo, bo), (gw, gb) in zip(self.layers, grads):
            blk = np.zeros((cout, kp))
            blk[:, :taps * cin] = gw.reshape(cout, taps * cin)
            flat[wo:wo + cout * kp] = blk.ravel()
            flat[bo:bo + cout] = gb
        return flat

    # ---- layer primitives --------------------------------------------------
    @staticmethod
    def _shifts(taps):
        return [(0, 0)] if taps == 1 else [(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]

    @staticmethod
    def _shift(x, dy, dx):
        """y[b,h,w] = x[b,h+dy,w+dx], zero outside"""
        y = np.zeros_like(x)
        H, W = x.shape[1], x.shape[2]
        hs, he = max(0, -dy), min(H, H - dy)
        ws, we = max(0, -dx), min(W, W - dx)
        y[:, hs:he, ws:we] = x[:, hs + dy:he + dy, ws + dx:we + dx]
        return y

    def conv(self, x, w, b, relu):
        cout, taps, cin = w.shape
        y = np.zeros(x.shape[:3] + (cout,))
        for t, (dy, dx) in enumerate(self._shifts(taps)):
            y += self._shift(x, dy, dx) @ w[:, t, :].T
        y += b
        return np.maximum(y, 0) if relu else y

    def conv_grads(self, x, w, dz):
        """weight, bias and input gradients of a conv with pre-activation grad dz"""
        cout, taps, cin = w.shape
        gw = np.zeros_like(w)
        dx = np.zeros(x.shape)
        for t, (dy, ddx) in enumerate(self._shifts(taps)):
            xs = self._shift(x, dy, ddx)
            gw[:, t, :] = np.einsum("bhwo,bhwi->oi", dz, xs)
            dx += self._shift(dz @ w[:, t, :], -dy, -ddx)
        return gw, dz.sum(axis=(0, 1, 2)), dx

    @staticmethod
    def pool(x):
        return 0.25 * (x[:, 0::2, 0::2] + x[:, 1::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 1::2])

    @staticmethod
    def up(x):
        return x.repeat(2, axis=1).repeat(2, axis=2)

    # ---- network -----------------------------------------------------------
    def _idx(self):
        L = self.levels
        enc = [(2 * l, 2 * l + 1) for l in range(L)]
        dec = [(2 * L + 2 * l, 2 * L + 2 * l + 1) for l in range(L - 1)]
        return enc, dec, 2 * L + 2 * (L - 1)

    def forward(self, x0, params):
        """x0: [B][H][W] network input channel 0.  Returns (head output
        [B][H][W][8] float64, cache)."""
        r = self._r
        P = [(r(w), b) for w, b in params]
        enc, dec, hd = self._idx()
        L = self.levels
        B, H = x0.shape[0], x0.shape[1]
        inp = np.zeros((B, H, H, 8))
        inp[..., 0] = r(x0)
        c = {"in": [inp], "ea": [], "e": [], "cat": {}, "da": {}, "d": {}, "P": P}
        x = inp
        for l in range(L):
            ea = r(self.conv(x, *P[enc[l][0]], True))
            e = r(self.conv(ea, *P[enc[l][1]], True))
            c["ea"].append(ea)
            c["e"].append(e)
            if l + 1 < L:
                x = r(self.pool(e))
                c["in"].append(x)
        low = c["e"][L - 1]
        for l in range(L - 2, -1, -1):
            cat = np.concatenate([self.up(low), c["e"][l]], axis=3)
            da = r(self.conv(cat, *P[dec[l][0]], True))
            d = r(self.conv(da, *P[dec[l][1]], True))
            c["cat"][l], c["da"][l], c["d"][l] = cat, da, d
            low = d
        top = c["d"][0] if L >= 2 else c["e"][0]
        c["top"] = top
        y = self.conv(top, *P[hd], False)
        if self.bf16:
            y = y.astype(np.float32).astype(np.float64)
        return y, c

    @staticmethod
    def loss(y, tgt, lp=0.5, ls=1.0):
        """mean loss and dLoss/dy (channels 0..2)"""
        a, p, s = y[..., 0], y[..., 1], y[..., 2]
        at, pt, mt = tgt[..., 0], tgt[..., 1], tgt[..., 2]
        M = a.size
        d = p - pt
        val = (a - at) ** 2 + lp * mt * (1 - np.cos(d)) + ls * (np.maximum(s, 0) - s * mt + np.log1p(np.exp(-np.abs(s))))
        g = np.zeros(y.shape)
        g[..., 0] = 2 * (a - at) / M
        g[..., 1] = lp * mt * np.sin(d) / M
        g[..., 2] = ls * (1 / (1 + np.exp(-s)) - mt) / M
        return val.sum() / M, g

    def backward(self, dy, c):
        """gradients [(gw, gb)] per layer for head-output gradient dy"""
        r = self._r
        P = c["P"]
        enc, dec, hd = self._idx()
        L = self.levels
        grads = [None] * len(P)
        dz = r(dy)
        gw, gb, dtop = self.conv_grads(c["top"], P[hd][0], dz)
        grads[hd] = (gw, gb)
        dtop = r(dtop)
        de = [None] * L
        dd = dtop
        for l in range(L - 1):
            dz = dd * (c["d"][l] > 0)
            gw, gb, dda = self.conv_grads(c["da"][l], P[dec[l][1]][0], dz)
            grads[dec[l][1]] = (gw, gb)
            dz = r(dda) * (c["da"][l] > 0)
            gw, gb, dcat = self.conv_grads(c["cat"][l], P[dec[l][0]][0], dz)
            grads[dec[l][0]] = (gw, gb)
            dcat = r(dcat)
            cu = dcat.shape[3] - c["e"][l].shape[3]
            de[l] = dcat[..., cu:].copy()
            dlow = dcat[..., :cu]
            dd = r(dlow[:, 0::2, 0::2] + dlow[:, 1::2, 0::2] + dlow[:, 0::2, 1::2] + dlow[:, 1::2, 1::2])
        if L >= 2:
            de[L - 1] = dd
        else:
            de[0] = dtop
        for l in range(L - 1, -1, -1):
            dz = de[l] * (c["e"][l] > 0)
            gw, gb, dea = self.conv_grads(c["ea"][l], P[enc[l][1]][0], dz)
            grads[enc[l][1]] = (gw, gb)
            dz = r(dea) * (c["ea"][l] > 0)
            gw, gb, dx = self.conv_grads(c["in"][l], P[enc[l][0]][0], dz)
            grads[enc[l][0]] = (gw, gb)
            if l > 0:
                dp = r(dx)
                de[l - 1] = r(de[l - 1] + 0.25 * self.up(dp))
        return grads
