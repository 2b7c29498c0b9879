"""Cross-check of the hand-written FFT against rocFFT (through torch.fft on
the GPU, which calls hipFFT/rocFFT on ROCm).  rocFFT is a check only: the
product path never calls it."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(256, 256), (384, 384), (1024, 1024), (2048, 2048), (512, 768)])
def test_fft_agrees_with_rocfft(shape):
    print(f"rocFFT cross-check {shape}", flush=True)
    import torch
    import cdixx_amd as cdx
    rs = np.random.default_rng(7)
    x = (rs.standard_normal(shape) + 1j * rs.standard_normal(shape)).astype(np.complex64)
    ctx = cdx.Context(*shape, precision=cdx.F32)
    d = ctx.cplx(x)
    ctx.fft(d, d, cdx.FORWARD)
    ours = ctx.get_cplx(d)
    ctx.fft(d, d, cdx.INVERSE)
    back = ctx.get_cplx(d)
    ctx.close()
    t = torch.from_numpy(x).to("cuda")
    roc = torch.fft.fft2(t).cpu().numpy()
    roc_inv = torch.fft.ifft2(torch.from_numpy(ours).to("cuda")).cpu().numpy() * (shape[0] * shape[1])
    torch.cuda.synchronize()
    assert rel_l2(ours, roc) < 2e-6
    assert rel_l2(back, roc_inv) < 2e-6
