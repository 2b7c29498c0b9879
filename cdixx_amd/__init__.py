"""cdixx_amd -- MI355X-native CDI phase-retrieval hot path (Python host mirror).

Thin ctypes layer over libcdixx_amd.so (include/cdx_abi.h).  It mirrors the
reference's operator surface for this path (CDIxx include/cdi.hpp,
include/cudaConfig.hpp): the same operations, argument meaning and failure
behaviour, with an explicit context instead of hidden globals.  There is no
CPU fallback: if the HIP library is missing, importing the module raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcdixx_amd.so")

F32, F64 = 0, 1
RAAR, ER, POSER, HIO, POSHIO, FHIO, SHRINKWRAP, XCORRELATION, KKT, TV = range(10)
TOKENS = {"RAAR": RAAR, "ER": ER, "POSER": POSER, "HIO": HIO, "POSHIO": POSHIO, "FHIO": FHIO,
          "shrinkWrap": SHRINKWRAP, "XCORRELATION": XCORRELATION, "KKT": KKT, "TV": TV}
FORWARD, INVERSE = -1, 1
STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_SCHEDULE", 4: "ERR_UNSUPPORTED",
          5: "ERR_DATA", 6: "ERR_IO", 7: "ERR_CONFIG"}


class CdxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python cdixx_amd/build.py` "
                          "(there is no CPU fallback for this path)")
    return ctypes.CDLL(LIB_PATH)


lib = _load()


class Params(ctypes.Structure):
    """cdx_params (include/cdx_abi.h)."""
    _fields_ = [("simCCDbit", ctypes.c_int), ("useBS", ctypes.c_int), ("freeBeamstop", ctypes.c_int),
                ("noiseLevel", ctypes.c_int), ("rcolor", ctypes.c_int), ("exposure", ctypes.c_double),
                ("beta", ctypes.c_double), ("relax", ctypes.c_double), ("shrinkThreshold", ctypes.c_double),
                ("shrinkMode", ctypes.c_int), ("sigma0", ctypes.c_double), ("sigmaMin", ctypes.c_double),
                ("sigmaDecay", ctypes.c_double), ("recordR", ctypes.c_int), ("useGraph", ctypes.c_int),
                ("prtfIter", ctypes.c_int)]

    @classmethod
    def default(cls, **kw):
        p = cls()
        lib.cdx_params_default(ctypes.byref(p))
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
        return p


_vp, _i, _d, _u64, _i64, _sz = (ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_uint64,
                                ctypes.c_int64, ctypes.c_size_t)
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_SIGS = {
    "cdx_last_error": ([], ctypes.c_char_p), "cdx_version": ([], ctypes.c_char_p),
    "cdx_size_supported": ([_i], _i),
    "cdx_malloc": ([ctypes.POINTER(_vp), _sz], _i), "cdx_free": ([_vp], _i),
    "cdx_memcpy_h2d": ([_vp, _vp, _sz], _i), "cdx_memcpy_d2h": ([_vp, _vp, _sz], _i),
    "cdx_memcpy_d2d": ([_vp, _vp, _sz], _i), "cdx_memset0": ([_vp, _sz], _i),
    "cdx_device_synchronize": ([], _i),
    "cdx_create": ([ctypes.POINTER(_vp), _i, _i, _i, _vp], _i), "cdx_destroy": ([_vp], _i),
    "cdx_synchronize": ([_vp], _i), "cdx_stream": ([_vp], _vp),
    "cdx_init_vars": ([_vp, _i, _d], _i), "cdx_set_threshold": ([_vp, _d], _i),
    "cdx_get_threshold": ([_vp, _dp], _i), "cdx_set_beta": ([_vp, _d], _i), "cdx_set_relax": ([_vp, _d], _i),
    "cdx_fft": ([_vp, _vp, _vp, _i], _i), "cdx_apply_mod_abs": ([_vp, _vp, _vp], _i),
    "cdx_apply_mod": ([_vp, _vp, _vp, _vp, _i, _d], _i),
    "cdx_apply_support": ([_vp, _vp, _vp, _i, _vp], _i), "cdx_apply_norm": ([_vp, _vp, _d], _i),
    "cdx_apply_norm_real": ([_vp, _vp, _d], _i), "cdx_get_mod2": ([_vp, _vp, _vp, _d], _i),
    "cdx_gauss_mult_freq": ([_vp, _vp, _vp, _d], _i), "cdx_gauss_conv": ([_vp, _vp, _vp, _d], _i),
    "cdx_convert_fo": ([_vp, _vp, _vp, _i], _i), "cdx_create_support": ([_vp, _vp, _d], _i),
    "cdx_create_beamstop": ([_vp, _vp, _d], _i), "cdx_create_circle": ([_vp, _vp, _d], _i),
    "cdx_find_max": ([_vp, _vp, _i64, _dp], _i), "cdx_find_sum": ([_vp, _vp, _i64, _dp], _i),
    "cdx_random_start": ([_vp, _vp, _vp, _u64, _i, _vp], _i),
    "cdx_residual": ([_vp, _vp, _vp, _vp, _dp], _i), "cdx_rfactor": ([_vp, _vp, _vp, _vp, _dp], _i),
    "cdx_params_default": ([ctypes.POINTER(Params)], None),
    "cdx_phase_retrieve": ([_vp, ctypes.c_char_p, ctypes.POINTER(Params), _vp, _vp, _vp, _vp, _vp, _dp, _i, _ip],
                           _i),
    "cdx_expand_schedule": ([ctypes.c_char_p, _ip, _i], _i),
    "cdx_time_iterations": ([_vp, ctypes.POINTER(Params), _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _dp, _dp, _dp],
                            _i),
    "cdx_time_passes": ([_vp, ctypes.POINTER(Params), _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _dp, _dp], _i),
    "cdx_sim_object": ([_vp, _d, _u64, _d, _vp], _i),
    "cdx_sim_intensity": ([_vp, _vp, _d, _u64, _i, _vp], _i),
    "cdx_read_config_json": ([ctypes.c_char_p, ctypes.c_char_p, _i], _i),
    "cdx_copy": ([_vp, _vp, _vp, _sz], _i),
    "cdx_prtf": ([_vp, _dp, _dp, _i, _ip], _i),
    "cdx_phase_retrieve_static": ([_i, _i, ctypes.c_char_p, ctypes.POINTER(Params), _vp, _vp, _vp, _vp, _vp, _dp,
                                   _vp, _ip], _i),
    "cdx_batch_create": ([ctypes.POINTER(_vp), _i, _ip, _i, _vp], _i), "cdx_batch_destroy": ([_vp], _i),
    "cdx_batch_frame": ([_vp, _i] + [ctypes.POINTER(_vp)] * 5 + [_ip], _i),
    "cdx_batch_context": ([_vp, _i], _vp), "cdx_batch_set_threshold": ([_vp, _d], _i),
    "cdx_batch_get_thresholds": ([_vp, _dp], _i), "cdx_batch_synchronize": ([_vp], _i),
    "cdx_batch_phase_retrieve": ([_vp, ctypes.c_char_p, ctypes.POINTER(Params), _dp, _i, _ip], _i),
    "cdx_batch_time_iterations": ([_vp, ctypes.POINTER(Params), _i, _i, _i, _dp], _i),
    "cdx_batch_set_tiling": ([_vp, _i], _i), "cdx_batch_get_tiling": ([_vp, _ip], _i),
    "cdx_batch_autotune": ([_vp, _i, _dp, _i], _i),
    "cdx_batch_set_schedule": ([_vp, _i, _ip], _i),
    "cdx_batch_get_schedule": ([_vp, _ip, _ip, ctypes.c_char_p, _i], _i),
    "cdx_set_tiling": ([_vp, _i, _i], _i), "cdx_get_tiling": ([_vp, _ip, _ip], _i),
    "cdx_set_timing_warmup": ([_vp, _i], _i),
    "cdx_autotune": ([_vp, _i, _dp, _i], _i),
    "cdx_convert_fo_phase": ([_vp, _vp, _d], _i), "cdx_mask_multiply": ([_vp, _vp, _vp, _i], _i),
    "cdx_poisson_noise_wo": ([_vp, _vp, _d, _u64], _i), "cdx_extend_to_complex": ([_vp, _vp, _vp], _i),
    "cdx_create_wavefront": ([_vp, _vp, _vp, _i, _i, _vp], _i),
    "cdx_bin_read_header": ([ctypes.c_char_p, _ip, _ip, _ip], _i),
    "cdx_bin_read": ([ctypes.c_char_p, _vp, _sz], _i),
    "cdx_bin_write": ([ctypes.c_char_p, _i, _i, _i, _vp], _i),
    "cdx_sim_object_shape": ([_vp, _i, _dp, _i, _u64, _d, _vp], _i),
    "cdx_preprocess_pattern": ([_vp, _i, _i, _vp, _i, _i, _i, _i, _d, _d, _i, _vp, _i, _ip, _dp], _i),
    "cdx_context_info": ([_vp, _ip, _ip, _ip], _i),
    "cdx_unet_create": ([ctypes.POINTER(_vp), _i, _i, _i, _i, _u64], _i),
    "cdx_unet_destroy": ([_vp], _i), "cdx_unet_num_params": ([_vp, ctypes.POINTER(_i64)], _i),
    "cdx_unet_layers": ([_vp, ctypes.POINTER(_i64), _i], _i),
    "cdx_unet_set_params": ([_vp, _vp, _i64], _i), "cdx_unet_get_params": ([_vp, _vp, _i64], _i),
    "cdx_unet_get_grads": ([_vp, _vp, _i64], _i), "cdx_unet_zero_grads": ([_vp], _i),
    "cdx_unet_set_batch": ([_vp, _i, _vp, _vp], _i), "cdx_unet_get_batch": ([_vp, _i, _vp, _vp], _i),
    "cdx_unet_forward": ([_vp, _i], _i), "cdx_unet_output": ([_vp, _i, _vp], _i),
    "cdx_unet_backward": ([_vp, _i, _dp], _i), "cdx_unet_adam": ([_vp, _d], _i),
    "cdx_unet_synth_batch": ([_vp, _i, _u64, _d, _d], _i),
    "cdx_unet_synth_frame": ([_vp, _u64, _d, _d, _vp, _vp], _i),
    "cdx_unet_train": ([_vp, _i, _i, _u64, _d, _d, _d, _dp], _i),
    "cdx_unet_warm_start": ([_vp, _vp, _vp, _d, _vp, _vp, _vp], _i),
    "cdx_unet_warm_start_scaled": ([_vp, _vp, _vp, _vp, _vp, _d, _i, _u64, _vp, _vp, _vp], _i),
    "cdx_unet_save": ([_vp, ctypes.c_char_p], _i), "cdx_unet_load": ([_vp, ctypes.c_char_p], _i),
    "cdx_ptycho_create": ([ctypes.POINTER(_vp), _i, _i, _i, _i, _ip, _ip, _vp], _i),
    "cdx_ptycho_destroy": ([_vp], _i), "cdx_ptycho_set_positions": ([_vp, _ip, _ip], _i),
    "cdx_ptycho_set_patterns": ([_vp, _vp, _vp], _i), "cdx_ptycho_params_default": ([_vp], _i),
    "cdx_ptycho_sweep": ([_vp, _ip, _i, _vp, _vp, _vp, _vp, _ip], _i),
    "cdx_ptycho_apply_probe_step": ([_vp, _vp, _vp], _i),
    "cdx_ptycho_simulate": ([_vp, _vp, _vp, _vp], _i), "cdx_ptycho_synchronize": ([_vp], _i),
    "cdx_ptycho_time_sweeps": ([_vp, _ip, _i, _vp, _vp, _vp, _vp, _i, _dp, _ip], _i),
    "cdx_ptycho_levels": ([_i, _i, _i, _i, _ip, _ip, _ip, _i, _i, _ip, _ip, _ip], _i),
}
for _name, (_args, _res) in _SIGS.items():
    if hasattr(lib, _name):
        _f = getattr(lib, _name)
        _f.argtypes = _args
        _f.restype = _res


def check(rc):
    if rc != 0:
        raise CdxError(rc, lib.cdx_last_error().decode())
    return rc


def last_error():
    return lib.cdx_last_error().decode()


def expand_schedule(algorithm):
    """AlgoParser expansion of an algorithm string (host only, no GPU needed)."""
    n = lib.cdx_expand_schedule(algorithm.encode(), None, 0)
    if n < 0:
        raise CdxError(-n, f"bad algorithm string {algorithm!r}")
    out = (ctypes.c_int * max(n, 1))()
    lib.cdx_expand_schedule(algorithm.encode(), out, n)
    return list(out[:n])


def preprocess_pattern(sig, bkg=None, nmerge=1, centre=False, shift_row=0.0, shift_col=0.0, shift_mode=0):
    """util/processimg.cpp: background subtraction, optional centring with a
    sub-pixel FFT shift (mode 0) or spectrum-modulus re-centring (mode 1),
    cropping to a multiple of 4*nmerge and nmerge x nmerge binning.
    Returns (image, (midx, midy))."""
    sig = np.ascontiguousarray(sig, np.float32)
    b = None if bkg is None else np.ascontiguousarray(bkg, np.float32)
    cap = sig.size
    out = np.zeros(cap, np.float32)
    n = ctypes.c_int()
    mid = (ctypes.c_double * 2)()
    check(lib.cdx_preprocess_pattern(sig.ctypes.data, sig.shape[0], sig.shape[1],
                                     None if b is None else b.ctypes.data, 0 if b is None else b.shape[0],
                                     0 if b is None else b.shape[1], nmerge, int(centre), shift_row, shift_col,
                                     shift_mode, out.ctypes.data, cap, ctypes.byref(n), mid))
    k = n.value
    return out[:k * k].reshape(k, k).copy(), (mid[0], mid[1])


_BIN_TYPES = {3: np.float32, 4: np.complex64, 5: np.float64, 6: np.complex128}


def bin_read_header(path):
    """(type, rows, cols) of a .bin image (struct imageFile, include/imageFile.hpp)"""
    t, r, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(lib.cdx_bin_read_header(path.encode(), ctypes.byref(t), ctypes.byref(r), ctypes.byref(c)))
    return t.value, r.value, c.value


def bin_read(path):
    """readImage_c of a .bin file (FLOAT, FLOAT2, DOUBLE, DOUBLE2) -> numpy array"""
    t, r, c = bin_read_header(path)
    out = np.empty((r, c), _BIN_TYPES[t])
    check(lib.cdx_bin_read(path.encode(), out.ctypes.data, out.nbytes))
    return out


def bin_write(path, a):
    """writeFloatImage / writeComplexImage (and their double forms) of a 2-D array"""
    a = np.ascontiguousarray(a)
    code = {np.dtype(v): k for k, v in _BIN_TYPES.items()}.get(a.dtype)
    if code is None or a.ndim != 2:
        raise ValueError("bin_write: 2-D float32/complex64/float64/complex128 array expected")
    check(lib.cdx_bin_write(path.encode(), code, a.shape[0], a.shape[1], a.ctypes.data))


def phase_retrieve_static(rows, cols, algorithm, params, F, g, I, spt, bs=None, threshold=0.5, stream=None):
    """cdx_phase_retrieve_static: the loop on FP32 frames up to 512 x 512
    with no device allocation (static workspace).  Returns (n_iter, threshold)."""
    thr = ctypes.c_double(threshold)
    n_iter = ctypes.c_int()
    check(lib.cdx_phase_retrieve_static(rows, cols, algorithm.encode(), ctypes.byref(params), F.ptr, g.ptr, I.ptr,
                                        bs.ptr if bs else None, spt.ptr, ctypes.byref(thr), stream,
                                        ctypes.byref(n_iter)))
    return n_iter.value, thr.value


def read_config(path):
    """readConfig of a libconfig file -> dict of every field (host only)."""
    import json
    buf = ctypes.create_string_buffer(1 << 16)
    check(lib.cdx_read_config_json(path.encode(), buf, len(buf)))
    return json.loads(buf.value.decode())


class DeviceBuffer:
    """Raw device allocation owned by Python (cdx_malloc / cdx_free)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = _vp()
        check(lib.cdx_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib.cdx_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes))

    def download(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(lib.cdx_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self):
        check(lib.cdx_memset0(self.ptr, self.nbytes))

    def free(self):
        if getattr(self, "ptr", None):
            lib.cdx_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """Explicit replacement of the reference's hidden state (image size, FFT
    plan, cudaVars) -- one per thread."""

    def __init__(self, rows, cols=None, precision=F32):
        cols = rows if cols is None else cols
        self.rows, self.cols, self.precision = rows, cols, precision
        self.rdtype = np.float64 if precision == F64 else np.float32
        self.cdtype = np.complex128 if precision == F64 else np.complex64
        h = _vp()
        check(lib.cdx_create(ctypes.byref(h), rows, cols, precision, None))
        self.h = h.value

    # buffers
    @property
    def shape(self):
        return (self.rows, self.cols)

    def real(self, a=None):
        b = DeviceBuffer(self.rows * self.cols * np.dtype(self.rdtype).itemsize)
        if a is not None:
            b.upload(np.asarray(a, dtype=self.rdtype))
        return b

    def cplx(self, a=None):
        b = DeviceBuffer(self.rows * self.cols * np.dtype(self.cdtype).itemsize)
        if a is not None:
            b.upload(np.asarray(a, dtype=self.cdtype))
        return b

    def get_real(self, b):
        self.sync()
        return b.download(self.rdtype, self.shape)

    def get_cplx(self, b):
        self.sync()
        return b.download(self.cdtype, self.shape)

    # tiling of the FFT passes
    def set_tiling(self, col_lines=0, row_lines=0):
        check(lib.cdx_set_tiling(self.h, col_lines, row_lines))

    def set_timing_warmup(self, iters):
        """untimed iterations time_iterations runs (at least) right before its
        timed ones (cdx_set_timing_warmup)"""
        check(lib.cdx_set_timing_warmup(self.h, int(iters)))

    def tiling(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        check(lib.cdx_get_tiling(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def autotune(self, reps=50):
        """Times every valid tiling of both passes and keeps the fastest.
        Returns [(pass, lines, us_per_launch), ...] (pass 0 = column, 1 = row)."""
        rep = (ctypes.c_double * 96)()
        check(lib.cdx_autotune(self.h, reps, rep, 96))
        out = [(int(rep[i]), int(rep[i + 1]), rep[i + 2]) for i in range(0, 96, 3) if rep[i + 1] > 0]
        return out

    # state
    def sync(self):
        check(lib.cdx_synchronize(self.h))

    def init_vars(self, rcolor=65535, scale=1.0):
        check(lib.cdx_init_vars(self.h, rcolor, scale))

    def set_threshold(self, t):
        check(lib.cdx_set_threshold(self.h, t))

    def threshold(self):
        v = ctypes.c_double()
        check(lib.cdx_get_threshold(self.h, ctypes.byref(v)))
        return v.value

    def set_beta(self, b):
        check(lib.cdx_set_beta(self.h, b))

    def set_relax(self, r):
        check(lib.cdx_set_relax(self.h, r))

    # free functions of the path
    def fft(self, src, dst=None, direction=FORWARD):
        dst = src if dst is None else dst
        check(lib.cdx_fft(self.h, src.ptr, dst.ptr, direction))
        return dst

    def apply_mod_abs(self, F, I):
        check(lib.cdx_apply_mod_abs(self.h, F.ptr, I.ptr))

    def apply_mod(self, F, I, bs=None, noise_level=0, bsnorm=1.0):
        check(lib.cdx_apply_mod(self.h, F.ptr, I.ptr, bs.ptr if bs else None, noise_level, bsnorm))

    def apply_support(self, g, gp, algo, spt):
        check(lib.cdx_apply_support(self.h, g.ptr, gp.ptr, algo, spt.ptr))

    def apply_norm(self, d, f):
        check(lib.cdx_apply_norm(self.h, d.ptr, f))

    def get_mod2(self, out, x, norm=1.0):
        check(lib.cdx_get_mod2(self.h, out.ptr, x.ptr, norm))

    def gauss_mult_freq(self, src, dst, sigma):
        check(lib.cdx_gauss_mult_freq(self.h, src.ptr, dst.ptr, sigma))

    def gauss_conv(self, src, dst, sigma):
        check(lib.cdx_gauss_conv(self.h, src.ptr, dst.ptr, sigma))

    def convert_fo(self, d, out=None, is_complex=True):
        check(lib.cdx_convert_fo(self.h, d.ptr, out.ptr if out else None, int(is_complex)))

    def create_support(self, spt, oversampling_spt=2.0):
        check(lib.cdx_create_support(self.h, spt.ptr, oversampling_spt))

    def create_beamstop(self, bs, radius):
        check(lib.cdx_create_beamstop(self.h, bs.ptr, radius))

    def create_circle(self, m, radius):
        check(lib.cdx_create_circle(self.h, m.ptr, radius))

    def find_max(self, x, n=0):
        v = ctypes.c_double()
        check(lib.cdx_find_max(self.h, x.ptr, n, ctypes.byref(v)))
        return v.value

    def find_sum(self, x, n=0):
        v = ctypes.c_double()
        check(lib.cdx_find_sum(self.h, x.ptr, n, ctypes.byref(v)))
        return v.value

    def random_start(self, I, bs, F, seed=0, mode=0):
        check(lib.cdx_random_start(self.h, I.ptr, bs.ptr if bs else None, seed, mode, F.ptr))

    def residual(self, F, I, bs=None):
        v = ctypes.c_double()
        check(lib.cdx_residual(self.h, F.ptr, I.ptr, bs.ptr if bs else None, ctypes.byref(v)))
        return v.value

    def rfactor(self, F, I, bs=None):
        v = ctypes.c_double()
        check(lib.cdx_rfactor(self.h, F.ptr, I.ptr, bs.ptr if bs else None, ctypes.byref(v)))
        return v.value

    def phase_retrieve(self, algorithm, params, F, g, I, spt, bs=None, rtrace_cap=0):
        """CDI::phaseRetrieve (src/post_gpu/cdi.cc:297-421). Returns (n_iter, rtrace)."""
        rt = (ctypes.c_double * max(rtrace_cap, 1))()
        n_iter = ctypes.c_int()
        check(lib.cdx_phase_retrieve(self.h, algorithm.encode(), ctypes.byref(params), F.ptr, g.ptr, I.ptr,
                                     bs.ptr if bs else None, spt.ptr, rt if rtrace_cap else None, rtrace_cap,
                                     ctypes.byref(n_iter)))
        return n_iter.value, np.array(rt[:rtrace_cap]) if rtrace_cap else None

    def prtf(self):
        """PRTF of the last phase_retrieve with params.prtfIter >= 0 (cdi.cc:423-440):
        (prtf sums, weights, count); prtf.txt holds prtf/weights."""
        m = self.rows // 2
        pr, wt, cnt = (ctypes.c_double * m)(), (ctypes.c_double * m)(), ctypes.c_int()
        check(lib.cdx_prtf(self.h, pr, wt, m, ctypes.byref(cnt)))
        return np.array(pr[:]), np.array(wt[:]), cnt.value

    def time_iterations(self, params, algo, F, g, I, spt, n_iter, bs=None, use_graph=True, per_kernel=True):
        tot, cm, rm = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib.cdx_time_iterations(self.h, ctypes.byref(params), algo, F.ptr, g.ptr, I.ptr,
                                      bs.ptr if bs else None, spt.ptr, n_iter, int(use_graph), ctypes.byref(tot),
                                      ctypes.byref(cm) if per_kernel else None,
                                      ctypes.byref(rm) if per_kernel else None))
        return tot.value, (cm.value if per_kernel else None), (rm.value if per_kernel else None)

    # setup steps of CDI::prepareIter / readObjectWave
    def create_wavefront(self, out, I=None, phase=None, obj_shape=None):
        """createWaveFront(I, phase, out, objrow, objcol): the object image
        centred in the frame (I, phase: device buffers of obj_shape or None)"""
        r, c = obj_shape
        check(lib.cdx_create_wavefront(self.h, I.ptr if I else None, phase.ptr if phase else None, r, c, out.ptr))

    def convert_fo_phase(self, d, norm):
        check(lib.cdx_convert_fo_phase(self.h, d.ptr, norm))

    def mask_multiply(self, d, mask, invert=False):
        check(lib.cdx_mask_multiply(self.h, d.ptr, mask.ptr, int(invert)))

    def poisson_noise_wo(self, d, noise_level, seed=0):
        check(lib.cdx_poisson_noise_wo(self.h, d.ptr, noise_level, seed))

    def extend_to_complex(self, a, b):
        check(lib.cdx_extend_to_complex(self.h, a.ptr, b.ptr))

    def time_passes(self, params, algo, F, g, I, spt, n, bs=None, mode=0):
        """ms per launch of the column and the row pass (cdx_time_passes):
        mode 0 back to back in a graph, mode 1 each launch between its own events"""
        cm, rm = ctypes.c_double(), ctypes.c_double()
        check(lib.cdx_time_passes(self.h, ctypes.byref(params), algo, F.ptr, g.ptr, I.ptr, bs.ptr if bs else None,
                                  spt.ptr, n, mode, ctypes.byref(cm), ctypes.byref(rm)))
        return cm.value, rm.value

    def sim_object(self, obj, radius, seed=0, smooth_sigma=3.0):
        check(lib.cdx_sim_object(self.h, radius, seed, smooth_sigma, obj.ptr))

    def sim_object_shape(self, obj, kind, params, seed=0, smooth_sigma=3.0):
        """kind 0 disk [r], 1 ellipse [a, b, rot], 2 polygon [angles..., radii...]"""
        prm = (ctypes.c_double * len(params))(*params)
        check(lib.cdx_sim_object_shape(self.h, kind, prm, len(params), seed, smooth_sigma, obj.ptr))

    def sim_intensity(self, obj, I, photons, seed=0, poisson=True):
        check(lib.cdx_sim_intensity(self.h, obj.ptr, photons, seed, int(poisson), I.ptr))

    def close(self):
        if getattr(self, "h", None):
            lib.cdx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class DeviceView:
    """A slice of memory owned elsewhere (e.g. a frame of a batch)."""

    def __init__(self, ptr, nbytes):
        self.ptr, self.nbytes = ptr, nbytes

    upload = DeviceBuffer.upload
    download = DeviceBuffer.download

    def zero(self):
        check(lib.cdx_memset0(self.ptr, self.nbytes))


class _Borrowed(Context):
    """A context owned by a batch (not destroyed here)."""

    def __init__(self, handle, n, precision):
        self.rows = self.cols = n
        self.precision = precision
        self.rdtype = np.float64 if precision == F64 else np.float32
        self.cdtype = np.complex128 if precision == F64 else np.complex64
        self.h = handle

    def close(self):
        self.h = None


class UNet:
    """Learned warm start (cdx_unet_*): U-Net on the centred log-intensity
    predicting amplitude, phase and a support logit."""

    def __init__(self, size=256, levels=4, base=32, max_batch=16, seed=1):
        self.size, self.levels, self.base, self.max_batch = size, levels, base, max_batch
        h = _vp()
        check(lib.cdx_unet_create(ctypes.byref(h), size, levels, base, max_batch, seed))
        self.h = h
        n = _i64()
        check(lib.cdx_unet_num_params(self.h, ctypes.byref(n)))
        self.nparam = n.value

    def layers(self):
        buf = (_i64 * (6 * 64))()
        nl = lib.cdx_unet_layers(self.h, buf, 6 * 64)
        return [tuple(buf[6 * i:6 * i + 6]) for i in range(nl)]

    def params(self):
        out = np.zeros(self.nparam, np.float32)
        check(lib.cdx_unet_get_params(self.h, out.ctypes.data, self.nparam))
        return out

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        check(lib.cdx_unet_set_params(self.h, flat.ctypes.data, self.nparam))

    def grads(self):
        out = np.zeros(self.nparam, np.float32)
        check(lib.cdx_unet_get_grads(self.h, out.ctypes.data, self.nparam))
        return out

    def zero_grads(self):
        check(lib.cdx_unet_zero_grads(self.h))

    def set_batch(self, x, targets=None):
        x = np.ascontiguousarray(x, np.float32)
        t = None if targets is None else np.ascontiguousarray(targets, np.float32)
        check(lib.cdx_unet_set_batch(self.h, x.shape[0], x.ctypes.data, None if t is None else t.ctypes.data))

    def get_batch(self, B):
        x = np.zeros((B, self.size, self.size), np.float32)
        t = np.zeros((B, self.size, self.size, 3), np.float32)
        check(lib.cdx_unet_get_batch(self.h, B, x.ctypes.data, t.ctypes.data))
        return x, t

    def forward(self, B):
        check(lib.cdx_unet_forward(self.h, B))
        out = np.zeros((B, self.size, self.size, 3), np.float32)
        check(lib.cdx_unet_output(self.h, B, out.ctypes.data))
        return out

    def backward(self, B):
        loss = ctypes.c_double()
        check(lib.cdx_unet_backward(self.h, B, ctypes.byref(loss)))
        return loss.value

    def adam(self, lr):
        check(lib.cdx_unet_adam(self.h, lr))

    def synth_batch(self, B, seed0, photons=(1e5, 1e8)):
        check(lib.cdx_unet_synth_batch(self.h, B, seed0, photons[0], photons[1]))

    def synth_frame(self, seed, obj, I, photons=(1e5, 1e8)):
        check(lib.cdx_unet_synth_frame(self.h, seed, photons[0], photons[1], obj.ptr, I.ptr))

    def train(self, steps, B, seed0, lr, photons=(1e5, 1e8)):
        losses = (ctypes.c_double * max(steps, 1))()
        check(lib.cdx_unet_train(self.h, steps, B, seed0, lr, photons[0], photons[1], losses))
        return np.array(losses[:steps])

    def warm_start(self, ctx, I, F, g, spt, threshold=0.5):
        check(lib.cdx_unet_warm_start(self.h, ctx.h, I.ptr, threshold, F.ptr, g.ptr, spt.ptr))

    def warm_start_scaled(self, ctx_small, ctx, I, bs, F, g, spt, threshold=0.5, dilate=0, seed=0):
        """Warm start of a frame f times larger than the network (per side):
        predicted phases on the central frequencies, random phases elsewhere,
        support dilated by `dilate` network pixels and upsampled about the
        centre (cdx_unet_warm_start_scaled)."""
        check(lib.cdx_unet_warm_start_scaled(self.h, ctx_small.h, ctx.h, I.ptr, bs.ptr if bs else None, threshold,
                                              int(dilate), seed, F.ptr, g.ptr, spt.ptr))

    def save(self, path):
        check(lib.cdx_unet_save(self.h, path.encode()))

    def load(self, path):
        check(lib.cdx_unet_load(self.h, path.encode()))

    def close(self):
        if getattr(self, "h", None):
            lib.cdx_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many frames of sizes 128/256/384/512 advanced together (cdx_batch_*)."""

    def __init__(self, sizes, precision=F32):
        self.sizes = [int(s) for s in sizes]
        self.precision = precision
        arr = (ctypes.c_int * len(self.sizes))(*self.sizes)
        h = _vp()
        check(lib.cdx_batch_create(ctypes.byref(h), len(self.sizes), arr, precision, None))
        self.h = h.value
        self.rdtype = np.float64 if precision == F64 else np.float32
        self.cdtype = np.complex128 if precision == F64 else np.complex64

    def __len__(self):
        return len(self.sizes)

    def frame(self, i):
        ptrs = [_vp() for _ in range(5)]
        n = ctypes.c_int()
        check(lib.cdx_batch_frame(self.h, i, *[ctypes.byref(p) for p in ptrs], ctypes.byref(n)))
        n = n.value
        cb, rb = n * n * np.dtype(self.cdtype).itemsize, n * n * np.dtype(self.rdtype).itemsize
        F, g, I, bs, spt = (DeviceView(p.value, b) for p, b in zip(ptrs, (cb, cb, rb, rb, rb)))
        return dict(F=F, g=g, I=I, bs=bs, spt=spt, n=n)

    def context(self, i):
        return _Borrowed(lib.cdx_batch_context(self.h, i), self.sizes[i], self.precision)

    def set_threshold(self, t):
        check(lib.cdx_batch_set_threshold(self.h, t))

    def thresholds(self):
        out = (ctypes.c_double * len(self.sizes))()
        check(lib.cdx_batch_get_thresholds(self.h, out))
        return np.array(out[:])

    def sync(self):
        check(lib.cdx_batch_synchronize(self.h))

    def get(self, i, name):
        self.sync()
        f = self.frame(i)
        dt = self.cdtype if name in ("F", "g") else self.rdtype
        return f[name].download(dt, (f["n"], f["n"]))

    def put(self, i, name, arr):
        f = self.frame(i)
        dt = self.cdtype if name in ("F", "g") else self.rdtype
        f[name].upload(np.asarray(arr, dtype=dt))

    def phase_retrieve(self, algorithm, params, rtrace_cap=0):
        nf = len(self.sizes)
        rt = (ctypes.c_double * max(nf * rtrace_cap, 1))()
        n_iter = ctypes.c_int()
        check(lib.cdx_batch_phase_retrieve(self.h, algorithm.encode(), ctypes.byref(params),
                                           rt if rtrace_cap else None, rtrace_cap, ctypes.byref(n_iter)))
        trace = np.array(rt[:nf * rtrace_cap]).reshape(nf, rtrace_cap) if rtrace_cap else None
        return n_iter.value, trace

    def time_iterations(self, params, algo, n, use_graph=True):
        ms = ctypes.c_double()
        check(lib.cdx_batch_time_iterations(self.h, ctypes.byref(params), algo, n, int(use_graph),
                                            ctypes.byref(ms)))
        return ms.value

    def set_tiling(self, threads):
        """threads per workgroup of both passes: 64, 128 or 256 (cdx_batch_set_tiling)"""
        check(lib.cdx_batch_set_tiling(self.h, int(threads)))

    def tiling(self):
        t = ctypes.c_int()
        check(lib.cdx_batch_get_tiling(self.h, ctypes.byref(t)))
        return t.value

    def set_schedule(self, threads, variants):
        """threads per workgroup and line variants (radix schedules) per frame
        size: variants = {128: v, 256: v, 384: v, 512: v}, v in {0, 1}
        (cdx_batch_set_schedule)"""
        arr = (ctypes.c_int * 4)(*[int(variants.get(n, 0)) for n in (128, 256, 384, 512)])
        check(lib.cdx_batch_set_schedule(self.h, int(threads), arr))

    def schedule(self):
        """(threads, {size: variant}, radices string) of the current schedule"""
        t, arr, buf = ctypes.c_int(), (ctypes.c_int * 4)(), ctypes.create_string_buffer(256)
        check(lib.cdx_batch_get_schedule(self.h, ctypes.byref(t), arr, buf, 256))
        return t.value, dict(zip((128, 256, 384, 512), arr[:])), buf.value.decode()

    def autotune(self, reps=20):
        """times HIO iterations for every schedule (64/128/256 threads per
        workgroup x line variants) and keeps the fastest; returns {threads: us
        per iteration of the best schedule at that size} of the sizes that fit"""
        us = (ctypes.c_double * 3)()
        check(lib.cdx_batch_autotune(self.h, int(reps), us, 3))
        return {64 << k: us[k] for k in range(3) if us[k] > 0}

    def simulate(self, radii, photons, seeds, bs_radius=0, support_scale=1.0, start_mode=1, poisson=True):
        """Fills every frame: seeded object of the given radius, Poisson
        intensities, beamstop, disk support of radius support_scale*radius,
        random start F and g = IFFT(F)/n (as prepareIter leaves objectWave)."""
        for i in range(len(self.sizes)):
            f = self.frame(i)
            ctx = self.context(i)
            ctx.sim_object(f["F"], radii[i], seed=int(seeds[i]))
            ctx.sim_intensity(f["F"], f["I"], photons[i], seed=int(seeds[i]), poisson=poisson)
            if bs_radius:
                ctx.create_beamstop(f["bs"], bs_radius)
            else:
                f["bs"].zero()
            ctx.create_circle(f["spt"], support_scale * radii[i])
            ctx.random_start(f["I"], f["bs"] if bs_radius else None, f["F"], seed=int(seeds[i]), mode=start_mode)
            ctx.fft(f["F"], f["g"], INVERSE)
            ctx.apply_norm(f["g"], 1.0 / f["n"])
        self.sync()

    def close(self):
        if getattr(self, "h", None):
            lib.cdx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PT_ACCURATE, PT_NOISE = 0, 1


class PtychoParams(ctypes.Structure):
    """cdx_ptycho_params (include/cdx_abi.h)."""
    _fields_ = [("mode", ctypes.c_int), ("noiseLevel", ctypes.c_int), ("scale", ctypes.c_double),
                ("rcolor", ctypes.c_double), ("probeMax", ctypes.c_double), ("objMax", ctypes.c_double),
                ("objStep", ctypes.c_double), ("probeStepSize", ctypes.c_double), ("updateProbe", ctypes.c_int),
                ("maxConcurrent", ctypes.c_int)]

    @classmethod
    def default(cls, **kw):
        p = cls()
        check(lib.cdx_ptycho_params_default(ctypes.byref(p)))
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
        return p


def ptycho_levels(n, obj_shape, sx, sy, order=None, max_concurrent=0):
    """Launch schedule of a sweep (host only): list of groups, each the scan
    indices launched together, in launch order (cdx_ptycho_levels)."""
    m = len(order) if order is not None else len(sx)
    ints = lambda v: (ctypes.c_int * max(len(v), 1))(*[int(x) for x in v])
    groups, lst, ng = (ctypes.c_int * max(m, 1))(), (ctypes.c_int * max(m, 1))(), ctypes.c_int()
    check(lib.cdx_ptycho_levels(int(n), int(obj_shape[0]), int(obj_shape[1]), len(sx), ints(sx), ints(sy),
                                ints(order) if order is not None else None, m, int(max_concurrent), groups, lst,
                                ctypes.byref(ng)))
    out, k = [], 0
    for g in groups[:ng.value]:
        out.append(list(lst[k:k + g]))
        k += g
    return out


class Ptycho:
    """The scan-position loop of CDIxx's ptychography (util/ptycho.cpp:514-561):
    n x n probe, obj_shape object, integer window offsets sx (rows), sy
    (columns) per scan position.  Object, probe, probe step and patterns are
    device buffers (DeviceBuffer) owned by the caller."""

    def __init__(self, n, obj_shape, sx, sy):
        self.n = int(n)
        self.obj_shape = (int(obj_shape[0]), int(obj_shape[1]))
        self.nscan = len(sx)
        h = _vp()
        check(lib.cdx_ptycho_create(ctypes.byref(h), self.n, self.obj_shape[0], self.obj_shape[1], self.nscan,
                                    self._ints(sx), self._ints(sy), None))
        self.h = h.value

    @staticmethod
    def _ints(v):
        v = [int(x) for x in v]
        return (ctypes.c_int * max(len(v), 1))(*v)

    def set_positions(self, sx, sy):
        check(lib.cdx_ptycho_set_positions(self.h, self._ints(sx), self._ints(sy)))

    def set_patterns(self, patterns, bs=None):
        """patterns: device f32 [nscan][n][n] (FFT order, divided by the exposure)"""
        check(lib.cdx_ptycho_set_patterns(self.h, patterns.ptr, bs.ptr if bs is not None else None))

    def sweep(self, params, obj, probe, probe_step=None, order=None):
        """one pass over `order` (default 0..nscan-1); returns the launch groups used"""
        ng = ctypes.c_int()
        o = self._ints(order) if order is not None else None
        check(lib.cdx_ptycho_sweep(self.h, o, len(order) if order is not None else 0, ctypes.byref(params),
                                   obj.ptr, probe.ptr, probe_step.ptr if probe_step is not None else None,
                                   ctypes.byref(ng)))
        return ng.value

    def apply_probe_step(self, probe, probe_step):
        check(lib.cdx_ptycho_apply_probe_step(self.h, probe.ptr, probe_step.ptr))

    def simulate(self, obj, probe, patterns):
        check(lib.cdx_ptycho_simulate(self.h, obj.ptr, probe.ptr, patterns.ptr))

    def sync(self):
        check(lib.cdx_ptycho_synchronize(self.h))

    def time_sweeps(self, params, obj, probe, n, probe_step=None, order=None):
        """(ms for n sweeps after one untimed sweep, launch groups per sweep)"""
        ms, ng = ctypes.c_double(), ctypes.c_int()
        o = self._ints(order) if order is not None else None
        check(lib.cdx_ptycho_time_sweeps(self.h, o, len(order) if order is not None else 0, ctypes.byref(params),
                                         obj.ptr, probe.ptr, probe_step.ptr if probe_step is not None else None,
                                         int(n), ctypes.byref(ms), ctypes.byref(ng)))
        return ms.value, ng.value

    def close(self):
        if getattr(self, "h", None):
            lib.cdx_ptycho_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
