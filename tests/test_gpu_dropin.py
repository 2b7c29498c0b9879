"""The drop-in boundary as a user of the reference drives it: the C++ class
of include/cdi.hpp through a caller shaped like util/cdi.cpp:8-50
(tests/dropin/cdi_caller.cpp, built by the package build), and the cdi_run
driver on a config file, both on a simulation from .bin object images
(runSim, simCCDbit, useBS, prtfIter, restart file).

Checks:
  - the simulated pattern CDI::prepareIter leaves in patternData
    (cdi.cc:161-200: createWaveFront, FFT, convertFOPhase, |.|^2, beamstop,
    applyPoissonNoise_WO in the centred layout) against the FP32 oracle's
    restatement of the same steps: without noise to 1e-5 relative L2; with
    noise, every pixel within one quantisation step and 99.9% exactly on the
    oracle's value;
  - saveState -> restart: the saved patternWave is the next start bit for bit;
  - cdi_run: exit status 0, the restart file (FLOAT2, the frame's size) and
    prtf.txt ("i, value" for i < rows/2, cdi.cc:435-439).
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_l2

pytestmark = pytest.mark.gpu

CALLER = os.path.join(ROOT, "tests", "dropin", "cdi_caller")
CDI_RUN = os.path.join(ROOT, "cdixx_amd", "cdi_run")
OBJ = 128            # object images; oversampling 2 -> 256 x 256 frame
EXPOSURE = 0.1
NOISE = 24
SEED = 11


def _write_inputs(cdx, tmp, sim_ccd, algorithm, prtf_iter, n_iter=0):
    rs = np.random.default_rng(5)
    y, x = np.mgrid[0:OBJ, 0:OBJ] - OBJ / 2
    disk = (x ** 2 + y ** 2) < (0.35 * OBJ) ** 2
    inten = (disk * rs.uniform(0.3, 1.0, (OBJ, OBJ))).astype(np.float32)
    phase = (disk * rs.uniform(0.2, 0.8, (OBJ, OBJ))).astype(np.float32)
    cdx.bin_write(str(tmp / "inten.bin"), inten)
    cdx.bin_write(str(tmp / "phase.bin"), phase)
    cfg = f"""name = "drop-in test";
InputImages = {{
  default = {{ Intensity = "inten.bin"; Phase = "phase.bin"; restart = "recon.bin"; }};
}};
Job = {{
  algorithm = "{algorithm}";
  oversampling = 2.; oversampling_spt = 2.;
  exposure = {EXPOSURE}; shrinkThreshold = 0.05;
  beamStopSize = 3; noiseLevel = {NOISE}; nIter = {n_iter}; prtfIter = {prtf_iter}; verbose = 0; seed = {SEED};
  phaseModulation = 1.;
  runSim = True; simCCDbit = {"True" if sim_ccd else "False"}; useBS = True; doIteration = True;
}};
"""
    (tmp / "cdi.cfg").write_text(cfg)
    return inten, phase


def _oracle_pattern(o32, o64, inten, phase, sim_ccd):
    """prepareIter's simulated pattern (FFT order) from the oracle's steps"""
    n = 2 * OBJ
    w = o32.create_wavefront(n, n, inten, phase)
    F = o32.convert_fo_phase(o32.fft2(w), 1 / np.sqrt(n * n))
    pat = (np.abs(F.astype(np.complex64)) ** 2).astype(np.float32)
    pat = pat * (1 - o32.beamstop(n, n, 3))
    if sim_ccd:
        # applyPoissonNoise_WO on the centred pattern, the first draw of the
        # seeded stream; vars.scale = 1/exposure
        c = o32.convert_fo(pat)
        c = o32.poisson_noise_wo(c, NOISE, 1 / EXPOSURE, 65535, seed=SEED)
        pat = o32.convert_fo(c)
    return pat.astype(np.float32)


@pytest.mark.parametrize("sim_ccd", [0, 1])
def test_cdi_class_caller(cdx, o32, o64, tmp_path, sim_ccd):
    print(f"drop-in caller simCCDbit={sim_ccd}", flush=True)
    assert os.path.exists(CALLER), "tests/dropin/cdi_caller is built by the package build"
    inten, phase = _write_inputs(cdx, tmp_path, sim_ccd, "2*(20*HIO+1*shrinkWrap)+10*ER", 45, n_iter=1)
    r = subprocess.run([CALLER, "cdi.cfg"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["rows"] == out["cols"] == 2 * OBJ and out["restart_bitexact"] is True
    assert np.isfinite(out["residual"]) and out["smallest_residual"] <= out["residual"] + 1e-12
    got = cdx.bin_read(str(tmp_path / "pattern.bin"))
    want = _oracle_pattern(o32, o64, inten, phase, sim_ccd)
    if not sim_ccd:
        e = rel_l2(got, want)
        print(f"  pattern vs oracle: rel-L2 {e:.2e}", flush=True)
        assert e < 1e-5
    else:
        step = (1 / EXPOSURE) / 65535
        d = np.abs(got - want)
        print(f"  pattern vs oracle: {np.mean(d == 0):.5f} exact, max {d.max() / step:.2f} steps", flush=True)
        assert d.max() <= 1.01 * step and np.mean(d == 0) > 0.999
    assert os.path.exists(tmp_path / "prtf.txt")


def test_cdi_run_driver(cdx, tmp_path):
    print("cdi_run on a config", flush=True)
    assert os.path.exists(CDI_RUN)
    _write_inputs(cdx, tmp_path, 1, "3*(20*RAAR+1*shrinkWrap)+20*ER", 50, n_iter=1)
    r = subprocess.run([CDI_RUN, "cdi.cfg"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert cdx.bin_read_header(str(tmp_path / "recon.bin")) == (4, 2 * OBJ, 2 * OBJ)
    wave = cdx.bin_read(str(tmp_path / "recon.bin"))
    assert np.isfinite(wave).all() and np.abs(wave).max() > 0
    lines = (tmp_path / "prtf.txt").read_text().strip().splitlines()
    assert len(lines) == OBJ
    for i, ln in enumerate(lines):
        a, b = ln.split(",")
        assert int(a) == i
        float(b)
    # a missing config file is an error exit, not a crash
    r = subprocess.run([CDI_RUN, "missing.cfg"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 7, (r.returncode, r.stderr)


@pytest.fixture(scope="module")
def cdx():
    import cdixx_amd
    return cdixx_amd
