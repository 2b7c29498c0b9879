"""Cross-check of the hand-written FFT against rocFFT (librocfft from the
ROCm install, called through its C API in this process on the same device
buffers).  rocFFT is a check only: the product path never calls it."""
import ctypes
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

ROCFFT = "/opt/rocm/lib/librocfft.so"
INPLACE, FORWARD, INVERSE, SINGLE = 0, 0, 1, 0  # rocfft_result_placement / transform_type / precision


def _rocfft(buf, rows, cols, kind):
    lib = ctypes.CDLL(ROCFFT)
    assert lib.rocfft_setup() == 0
    plan = ctypes.c_void_p()
    lengths = (ctypes.c_size_t * 2)(cols, rows)  # fastest dimension first
    assert lib.rocfft_plan_create(ctypes.byref(plan), INPLACE, kind, SINGLE, ctypes.c_size_t(2), lengths,
                                  ctypes.c_size_t(1), None) == 0
    ptrs = (ctypes.c_void_p * 1)(buf.ptr)
    assert lib.rocfft_execute(plan, ptrs, None, None) == 0
    import cdixx_amd as cdx
    cdx.check(cdx.lib.cdx_device_synchronize())
    lib.rocfft_plan_destroy(plan)


@pytest.mark.skipif(not os.path.exists(ROCFFT), reason="librocfft not installed")
@pytest.mark.parametrize("shape", [(256, 256), (384, 384), (1024, 1024), (2048, 2048), (512, 768)])
def test_fft_agrees_with_rocfft(shape):
    print(f"rocFFT cross-check {shape}", flush=True)
    import cdixx_amd as cdx
    rs = np.random.default_rng(7)
    x = (rs.standard_normal(shape) + 1j * rs.standard_normal(shape)).astype(np.complex64)
    ctx = cdx.Context(*shape, precision=cdx.F32)
    d = ctx.cplx(x)
    ctx.fft(d, d, cdx.FORWARD)
    ours = ctx.get_cplx(d)
    r = ctx.cplx(x)
    _rocfft(r, shape[0], shape[1], FORWARD)
    roc = ctx.get_cplx(r)
    # inverse of our forward result, both libraries (unnormalised)
    ctx.fft(d, d, cdx.INVERSE)
    back = ctx.get_cplx(d)
    r2 = ctx.cplx(ours)
    _rocfft(r2, shape[0], shape[1], INVERSE)
    roc_inv = ctx.get_cplx(r2)
    ctx.close()
    assert rel_l2(ours, roc) < 2e-6
    assert rel_l2(back, roc_inv) < 2e-6
    assert rel_l2(back / (shape[0] * shape[1]), x) < 2e-6
    # the same accuracy class as rocFFT's single precision (the reference's
    # myFFT is cuFFT in FP32): both against numpy's double transform
    exact = np.fft.fft2(x.astype(np.complex128))
    e_ours, e_roc = rel_l2(ours, exact), rel_l2(roc, exact)
    print(f"  error vs double: ours {e_ours:.3e}, rocFFT {e_roc:.3e}", flush=True)
    assert e_ours < 1.25 * e_roc
