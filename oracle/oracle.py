"""ctypes loader for the CPU oracle (liboracle_f64.so / liboracle_f32.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
RAAR, ER, POSER, HIO, POSHIO, FHIO, SHRINKWRAP, XCORRELATION, KKT, TV = range(10)


def build():
    subprocess.run(["sh", os.path.join(HERE, "build.sh")], check=True)


class OParams(ctypes.Structure):
    pass


def _params_type(real):
    class P(ctypes.Structure):
        _fields_ = [("simCCDbit", ctypes.c_int), ("useBS", ctypes.c_int), ("noiseLevel", ctypes.c_int),
                    ("rcolor", ctypes.c_int), ("shrinkMode", ctypes.c_int), ("startSign", ctypes.c_int),
                    ("scale", real), ("beta", real), ("relax", real), ("shrinkThreshold", real),
                    ("sigmaStart", real), ("sigmaMin", real), ("sigmaDecay", real),
                    ("freeBeamstop", ctypes.c_int)]
    return P


class Oracle:
    """One precision of the oracle; arrays are numpy, complex interleaved."""

    def __init__(self, precision="f64"):
        path = os.path.join(HERE, f"liboracle_{precision}.so")
        if not os.path.exists(path):
            build()
        self.lib = ctypes.CDLL(path)
        self.real = np.float64 if precision == "f64" else np.float32
        self.cplx = np.complex128 if precision == "f64" else np.complex64
        self.creal = ctypes.c_double if precision == "f64" else ctypes.c_float
        self.Params = _params_type(self.creal)
        L, vp, i, i64, u64 = self.lib, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64
        r = self.creal
        sig = {
            "orc_fft1d": [i, vp, i], "orc_fft2d": [i, i, vp, vp, i], "orc_dft2d_brute": [i, i, vp, vp, i],
            "orc_apply_mod_abs": [i64, vp, vp], "orc_apply_mod": [i64, vp, vp, vp, i, r, i, r],
            "orc_apply_support": [i64, vp, vp, i, vp, r, r, r], "orc_apply_norm": [i64, vp, r],
            "orc_get_mod2": [i64, vp, vp, r], "orc_gauss_mult_freq": [i, i, vp, vp, r],
            "orc_gauss_conv_real": [i, i, vp, vp, r], "orc_convert_fo_real": [i, i, vp],
            "orc_convert_fo_cpx": [i, i, vp], "orc_create_support_rect": [i, i, ctypes.c_double, vp],
            "orc_create_beamstop": [i, i, r, vp], "orc_create_circle": [i, i, r, vp],
            "orc_random_sign_start": [i, i, vp, vp, r, vp, vp],
            "orc_parser_expand": [ctypes.c_char_p, vp, i],
            "orc_rfactor": [i64, vp, vp, vp, vp, vp], "orc_residual": [i64, vp, vp, vp],
            "orc_philox4x64": [vp, vp, vp], "orc_uniform": [u64, u64, u64],
            "orc_sim_object": [i, r, u64, r, vp], "orc_sim_intensity": [i, i, vp, ctypes.c_double, u64, i, vp],
            "orc_poisson": [ctypes.c_double, u64, u64, u64],
            "orc_iteration": [i, i, vp, vp, vp, vp, vp, i, vp, r, vp],
            "orc_phase_retrieve": [i, i, ctypes.c_char_p, vp, vp, vp, vp, vp, vp, vp, vp, i, vp],
            "orc_phase_retrieve_prtf": [i, i, ctypes.c_char_p, vp, vp, vp, vp, vp, vp, vp, vp, i, vp, i, vp, vp, vp],
            "orc_add_phase": [i64, vp, vp, r], "orc_radial_sum": [i, i, vp, vp],
            "orc_shrink_wrap_freq": [i, i, vp, vp, r, vp, vp],
            "orc_create_wavefront": [i, i, vp, vp, i, i, vp], "orc_convert_fo_phase": [i, i, vp, r],
            "orc_poisson_noise_wo": [i64, vp, ctypes.c_double, r, r, u64], "orc_extend_to_complex": [i64, vp, vp],
        }
        for name, args in sig.items():
            getattr(L, name).argtypes = args
        L.orc_find_max.restype = r
        L.orc_find_sum.restype = r
        L.orc_residual.restype = ctypes.c_double
        L.orc_uniform.restype = ctypes.c_double
        L.orc_poisson.restype = ctypes.c_int64
        L.orc_parser_expand.restype = i
        L.orc_phase_retrieve.restype = i
        L.orc_phase_retrieve_prtf.restype = i

    # helpers
    def _c(self, a):
        return np.ascontiguousarray(a, dtype=self.cplx)

    def _r(self, a):
        return np.ascontiguousarray(a, dtype=self.real)

    @staticmethod
    def _p(a):
        return None if a is None else a.ctypes.data

    def params(self, **kw):
        p = self.Params()
        p.rcolor, p.scale, p.beta, p.relax, p.sigmaStart, p.sigmaMin, p.sigmaDecay = 65535, 1, 1, 1.9, 3, 2, 0.99
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def fft2(self, x, sign=-1):
        x = self._c(x)
        out = np.empty_like(x)
        self.lib.orc_fft2d(x.shape[0], x.shape[1], x.ctypes.data, out.ctypes.data, sign)
        return out

    def dft2_brute(self, x, sign=-1):
        x = self._c(x)
        out = np.empty_like(x)
        self.lib.orc_dft2d_brute(x.shape[0], x.shape[1], x.ctypes.data, out.ctypes.data, sign)
        return out

    def fft1(self, x, sign=-1):
        x = self._c(x).copy()
        self.lib.orc_fft1d(x.size, x.ctypes.data, sign)
        return x

    def apply_mod_abs(self, F, I):
        F = self._c(F).copy()
        I = self._r(I)
        self.lib.orc_apply_mod_abs(F.size, F.ctypes.data, I.ctypes.data)
        return F

    def apply_mod(self, F, I, bs=None, noise=0, scale=1.0, rcolor=65535, bsnorm=1.0):
        F = self._c(F).copy()
        I = self._r(I)
        bs = None if bs is None else self._r(bs)
        self.lib.orc_apply_mod(F.size, F.ctypes.data, I.ctypes.data, self._p(bs), noise, scale, rcolor, bsnorm)
        return F

    def apply_support(self, g, gp, algo, spt, threshold, beta=1.0, relax=1.9):
        g = self._c(g).copy()
        gp = self._c(gp)
        spt = self._r(spt)
        self.lib.orc_apply_support(g.size, g.ctypes.data, gp.ctypes.data, algo, spt.ctypes.data, threshold, beta,
                                   relax)
        return g

    def gauss_mult_freq(self, x, sigma):
        x = self._c(x)
        out = np.empty_like(x)
        self.lib.orc_gauss_mult_freq(x.shape[0], x.shape[1], x.ctypes.data, out.ctypes.data, sigma)
        return out

    def gauss_conv_real(self, x, sigma):
        x = self._r(x)
        out = np.empty_like(x)
        self.lib.orc_gauss_conv_real(x.shape[0], x.shape[1], x.ctypes.data, out.ctypes.data, sigma)
        return out

    def convert_fo(self, x):
        x = np.array(x, dtype=self.cplx if np.iscomplexobj(x) else self.real, copy=True)
        f = self.lib.orc_convert_fo_cpx if np.iscomplexobj(x) else self.lib.orc_convert_fo_real
        f(x.shape[0], x.shape[1], x.ctypes.data)
        return x

    def support_rect(self, rows, cols, os_=2.0):
        out = np.empty((rows, cols), self.real)
        self.lib.orc_create_support_rect(rows, cols, os_, out.ctypes.data)
        return out

    def beamstop(self, rows, cols, r):
        out = np.empty((rows, cols), self.real)
        self.lib.orc_create_beamstop(rows, cols, r, out.ctypes.data)
        return out

    def circle(self, rows, cols, r):
        out = np.empty((rows, cols), self.real)
        self.lib.orc_create_circle(rows, cols, r, out.ctypes.data)
        return out

    def expand(self, formula, cap=1 << 20):
        buf = np.zeros(cap, np.int32)
        n = self.lib.orc_parser_expand(formula.encode(), buf.ctypes.data, cap)
        return None if n < 0 else list(buf[:min(n, cap)])

    def rfactor(self, F, I, bs=None):
        F, I = self._c(F), self._r(I)
        bs = None if bs is None else self._r(bs)
        num, den = ctypes.c_double(), ctypes.c_double()
        self.lib.orc_rfactor(F.size, F.ctypes.data, I.ctypes.data, self._p(bs), ctypes.byref(num), ctypes.byref(den))
        return num.value / den.value if den.value > 0 else 0.0

    def residual(self, F, I, bs=None):
        F, I = self._c(F), self._r(I)
        bs = None if bs is None else self._r(bs)
        return self.lib.orc_residual(F.size, F.ctypes.data, I.ctypes.data, self._p(bs))

    def philox(self, ctr, key):
        c = np.array(ctr, np.uint64)
        k = np.array(key, np.uint64)
        o = np.zeros(4, np.uint64)
        self.lib.orc_philox4x64(c.ctypes.data, k.ctypes.data, o.ctypes.data)
        return o

    def uniform(self, seed, stream, idx):
        return self.lib.orc_uniform(seed, stream, idx)

    def poisson(self, lam, seed, stream, idx):
        return self.lib.orc_poisson(lam, seed, stream, idx)

    def sim_object(self, n, radius, seed, sigma=3.0):
        out = np.empty((n, n), self.cplx)
        self.lib.orc_sim_object(n, radius, seed, sigma, out.ctypes.data)
        return out

    def sim_intensity(self, obj, photons, seed, poisson=True):
        obj = self._c(obj)
        out = np.empty(obj.shape, self.real)
        self.lib.orc_sim_intensity(obj.shape[0], obj.shape[1], obj.ctypes.data, photons, seed, int(poisson),
                                   out.ctypes.data)
        return out

    def iteration(self, F, g, I, spt, algo, threshold, bs=None, **kw):
        F, g = self._c(F).copy(), self._c(g).copy()
        I, spt = self._r(I), self._r(spt)
        bs = None if bs is None else self._r(bs)
        gp = np.empty_like(g)
        p = self.params(**kw)
        self.lib.orc_iteration(F.shape[0], F.shape[1], F.ctypes.data, g.ctypes.data, I.ctypes.data, self._p(bs),
                               spt.ctypes.data, algo, ctypes.byref(p), threshold, gp.ctypes.data)
        return F, g

    def phase_retrieve(self, algorithm, F, g, I, spt, threshold=0.5, bs=None, rtrace_cap=0, prtf_iter=-1, **kw):
        F, g = self._c(F).copy(), self._c(g).copy()
        I, spt = self._r(I), self._r(spt).copy()
        bs = None if bs is None else self._r(bs)
        p = self.params(**kw)
        thr = self.creal(threshold)
        rt = np.zeros(max(rtrace_cap, 1), np.float64)
        nit, cnt = ctypes.c_int(), ctypes.c_int()
        nr = F.shape[0] // 2
        prtf, wts = np.zeros(nr, self.real), np.zeros(nr, self.real)
        rc = self.lib.orc_phase_retrieve_prtf(F.shape[0], F.shape[1], algorithm.encode(), F.ctypes.data,
                                              g.ctypes.data, I.ctypes.data, self._p(bs), spt.ctypes.data,
                                              ctypes.byref(p), ctypes.byref(thr),
                                              rt.ctypes.data if rtrace_cap else None, rtrace_cap, ctypes.byref(nit),
                                              prtf_iter, prtf.ctypes.data, wts.ctypes.data, ctypes.byref(cnt))
        if rc != 0:
            raise RuntimeError("oracle phase_retrieve failed")
        out = dict(F=F, g=g, spt=spt, threshold=thr.value, n_iter=nit.value, rtrace=rt[:rtrace_cap])
        if prtf_iter >= 0:
            out.update(prtf=prtf, weights=wts, prtf_count=cnt.value)
        return out

    def sw_state(self, sigma=3.0, avg=0.0, threshold=0.5, decay=0.99):
        """orc_sw_state {sigma, avg, threshold, decay} of the frequency-form shrink-wrap"""
        return (self.creal * 4)(sigma, avg, threshold, decay)

    def shrink_wrap_freq(self, F, shrink_threshold, state):
        """one frequency-form shrink-wrap step on the Fourier estimate F:
        returns the new (non-binary) support; `state` (sw_state) is updated"""
        F = self._c(F)
        spt = np.empty(F.shape, self.real)
        scratch = np.empty_like(F)
        self.lib.orc_shrink_wrap_freq(F.shape[0], F.shape[1], F.ctypes.data, spt.ctypes.data, shrink_threshold,
                                      ctypes.cast(state, ctypes.c_void_p), scratch.ctypes.data)
        return spt

    def create_wavefront(self, rows, cols, I=None, P=None, shape=None):
        """createWaveFront(I, P, out, objrow, objcol): the object image centred in a rows x cols frame"""
        shp = shape if shape is not None else (I if I is not None else P).shape
        I = None if I is None else self._r(I)
        P = None if P is None else self._r(P)
        out = np.empty((rows, cols), self.cplx)
        self.lib.orc_create_wavefront(rows, cols, self._p(I), self._p(P), shp[0], shp[1], out.ctypes.data)
        return out

    def convert_fo_phase(self, x, norm):
        x = self._c(x).copy()
        self.lib.orc_convert_fo_phase(x.shape[0], x.shape[1], x.ctypes.data, norm)
        return x

    def poisson_noise_wo(self, w, noise, scale, rcolor=65535, seed=0):
        w = self._r(w).copy()
        self.lib.orc_poisson_noise_wo(w.size, w.ctypes.data, noise, scale, rcolor, seed)
        return w

    def extend_to_complex(self, a):
        a = self._r(a)
        out = np.empty(a.shape, self.cplx)
        self.lib.orc_extend_to_complex(a.size, a.ctypes.data, out.ctypes.data)
        return out

    def add_phase(self, a, b, phi):
        a, b = self._c(a).copy(), self._c(b)
        self.lib.orc_add_phase(a.size, a.ctypes.data, b.ctypes.data, phi)
        return a

    def radial_sum(self, img):
        img = self._r(img)
        out = np.zeros(img.shape[0] // 2, self.real)
        self.lib.orc_radial_sum(img.shape[0], img.shape[1], img.ctypes.data, out.ctypes.data)
        return out
