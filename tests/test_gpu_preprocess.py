"""Pattern preprocessing (util/processimg.cpp) on the GPU against the numpy
oracle (oracle/preprocess_ref.py): background subtraction, centroid
centring with a sub-pixel FFT shift or spectrum-modulus re-centring,
cropping and binning."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _pattern(shape, cx, cy, seed=0):
    rs = np.random.default_rng(seed)
    R, C = shape
    x, y = np.mgrid[0:R, 0:C]
    r2 = (x - cx) ** 2 + (y - cy) ** 2
    sig = 1e4 * np.exp(-r2 / 50.0) * (1 + 0.5 * np.cos(0.3 * np.sqrt(r2))) + rs.uniform(0, 5, shape) + 3.0
    return sig.astype(np.float32)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("centre", [True, False])
def test_preprocess_matches_oracle(mode, centre):
    import cdixx_amd as cdx
    from oracle import preprocess_ref as pp
    print(f"preprocess mode={mode} centre={centre}", flush=True)
    R = C = 524 if centre else 516  # both crop to 512 (centre offset 3 / 2 px)
    sig = _pattern((R, C), R / 2 + 3.3, C / 2 - 2.6)
    bkg = (3.0 + np.random.default_rng(1).uniform(0, 0.1, (R + 8, C + 8))).astype(np.float32)
    got, mid = cdx.preprocess_pattern(sig, bkg, nmerge=2, centre=centre, shift_row=0.25, shift_col=-0.5,
                                      shift_mode=mode)
    ref, rmid = pp.preprocess(sig, bkg, nmerge=2, centre=centre, shift_row=0.25, shift_col=-0.5, shift_mode=mode)
    assert got.shape == ref.shape == (256, 256)
    assert abs(mid[0] - rmid[0]) < 1e-3 and abs(mid[1] - rmid[1]) < 1e-3
    assert rel_l2(got, ref) < 1e-5, rel_l2(got, ref)


def test_preprocess_rejects_unsupported_crop():
    import cdixx_amd as cdx
    sig = _pattern((200, 200), 100, 100)
    with pytest.raises(cdx.CdxError):
        cdx.preprocess_pattern(sig, None, nmerge=1)  # 200 is not an FFT length of this build
