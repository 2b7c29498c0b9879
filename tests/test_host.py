"""Host-side tests (no GPU): library exports, schedule grammar, config reader,
argument checks that fail before any device call."""
import ctypes
import glob
import json
import os
import re

import numpy as np
import pytest

from conftest import REFERENCE, ROOT

HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(cdx_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    import cdixx_amd
    syms = declared_symbols()
    assert len(syms) > 30
    missing = [s for s in sorted(syms) if not hasattr(cdixx_amd.lib, s)]
    assert not missing, missing


def test_version_and_sizes():
    import cdixx_amd
    assert b"gfx950" in cdixx_amd.lib.cdx_version()
    for n in (16, 256, 384, 512, 1024, 2048):
        assert cdixx_amd.lib.cdx_size_supported(n) == 1
    for n in (0, 100, 4096, 1000):
        assert cdixx_amd.lib.cdx_size_supported(n) == 0


def test_create_rejects_bad_arguments_before_touching_the_gpu():
    import cdixx_amd
    h = ctypes.c_void_p()
    assert cdixx_amd.lib.cdx_create(ctypes.byref(h), 1000, 1000, 0, None) == 1
    assert "unsupported size" in cdixx_amd.last_error()
    assert cdixx_amd.lib.cdx_create(ctypes.byref(h), 256, 256, 7, None) == 1
    assert cdixx_amd.lib.cdx_create(None, 256, 256, 0, None) == 1
    assert cdixx_amd.lib.cdx_phase_retrieve(None, b"1*HIO", None, None, None, None, None, None, None, 0, None) == 1


@pytest.mark.parametrize("formula", [
    "100*(3*POSHIO+20*RAAR+1*POSHIO+1*shrinkWrap)+20*RAAR+1*ER",
    "200*(20*POSHIO+1*POSHIO+1*shrinkWrap)+100*POSHIO",
    "3000*RAAR+20*ER", "2*(1*HIO+1*shrinkWrap)+1*ER", "1000*HIO",
    "10*(20*HIO+1*shrinkWrap)+2*(20*ER+1*shrinkWrap)+10*ER",
    "3*(2*(1*HIO+1*ER)+1*shrinkWrap)+1*FHIO", " 5 * RAAR + 2 * KKT ",
])
def test_schedule_matches_oracle(o64, formula):
    import cdixx_amd
    assert cdixx_amd.expand_schedule(formula) == o64.expand(formula)


def test_schedule_known_answers():
    import cdixx_amd
    assert len(cdixx_amd.expand_schedule("100*(3*POSHIO+20*RAAR+1*POSHIO+1*shrinkWrap)+20*RAAR+1*ER")) == 2521
    assert cdixx_amd.expand_schedule("2*(1*HIO+1*shrinkWrap)+1*ER") == [3, 6, 3, 6, 1]
    for bad in ("3*BOGUS", "HIO", "2*(1*HIO", ""):
        with pytest.raises(cdixx_amd.CdxError):
            cdixx_amd.expand_schedule(bad)


CFG = """name = "synthetic";
InputImages = {
  default = { Pattern = "pattern.bin"; Intensity = "obj.bin"; restart = "recon.bin"; };
  pupil = { Pattern = "p.png"; };
};
Job = {
  algorithm = "10*(20*HIO+1*shrinkWrap)" "+10*ER";  // adjacent literals
  oversampling = 2.; shrinkThreshold = 0.2; exposure = 1;   /* int into a float field */
  beamStopSize = 8.0;   // float into an int field: ignored (libconfig type rule)
  noiseLevel = 24; prtfIter = 0x10;
  useBS = True; simCCDbit = false; doIteration = TRUE;
  betaHIO = 0.9; unknownKey = 3;
};
"""


def test_config_reader_rules(tmp_path):
    import cdixx_amd
    p = tmp_path / "c.cfg"
    p.write_text(CFG)
    c = cdixx_amd.read_config(str(p))
    assert c["algorithm"] == "10*(20*HIO+1*shrinkWrap)+10*ER"
    assert c["oversampling"] == 2.0 and c["shrinkThreshold"] == 0.2 and c["exposure"] == 1.0
    assert c["beamStopSize"] == 5            # wrong type leaves the default
    assert c["noiseLevel"] == 24 and c["prtfIter"] == 16
    assert c["useBS"] is True and c["simCCDbit"] is False and c["doIteration"] is True
    assert c["betaHIO"] == 0.9 and c["relaxER"] == 1.9
    assert c["common"]["Pattern"] == "pattern.bin" and c["common"]["Phase"] is None
    assert c["pupil"]["Pattern"] == "p.png"
    assert c["verbose"] == 2 and c["photoncount"] == 1e9 and c["BSimage"] == ""


def test_config_errors(tmp_path):
    import cdixx_amd
    with pytest.raises(cdixx_amd.CdxError) as e:
        cdixx_amd.read_config(str(tmp_path / "missing.cfg"))
    assert e.value.code == 7
    p = tmp_path / "bad.cfg"
    p.write_text("Job = { algorithm = ; };")
    with pytest.raises(cdixx_amd.CdxError):
        cdixx_amd.read_config(str(p))


@pytest.mark.parametrize("rel,expect", [
    ("html/cdi.cfg", dict(shrinkThreshold=0.03, exposure=0.05, cropPattern=400, nIter=0, isFlip=True,
                          domnist=True, simCCDbit=True, doIteration=True, restart=True, oversampling=2.0,
                          algorithm="100*(3*POSHIO+20*RAAR+1*POSHIO+1*shrinkWrap)+20*RAAR+1*ER")),
    ("html/default/cdi.cfg", dict(shrinkThreshold=0.015, prtfIter=4401, noiseLevel=24, BSimage="mask512.png",
                                  useBS=True, isFresnel=False, verbose=4, oversampling=1.0, exposure=0.1)),
    ("html/default/pulse.cfg", dict(algorithm="3000*RAAR+20*ER", shrinkThreshold=0.23, nIter=1000,
                                    noiseLevel=100, exposure=0.3)),
])
def test_config_reads_reference_cfg_files(rel, expect):
    import cdixx_amd
    path = os.path.join(REFERENCE, rel)
    if not os.path.exists(path):
        pytest.skip("reference tree not mounted")
    c = cdixx_amd.read_config(path)
    for k, v in expect.items():
        assert c[k] == v, k


def test_reference_bin_file_read_in_place():
    """python/einstein.bin of the reference, read in place through the
    library's .bin reader (readImage_c, src/independant/imgio.c:23-43): the
    20-byte header (struct imageFile, include/imageFile.hpp:11-17) says FLOAT
    (type 3), 150 x 150, and the file holds 20 + 150*150*4 = 90,020 bytes."""
    import cdixx_amd
    path = os.path.join(REFERENCE, "python", "einstein.bin")
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    assert os.path.getsize(path) == 90020
    assert cdixx_amd.bin_read_header(path) == (3, 150, 150)
    img = cdixx_amd.bin_read(path)
    assert img.shape == (150, 150) and img.dtype == np.float32 and np.isfinite(img).all() and img.max() > 0


@pytest.mark.parametrize("dtype,code", [(np.float32, 3), (np.complex64, 4), (np.float64, 5), (np.complex128, 6)])
def test_bin_round_trip(tmp_path, dtype, code):
    """writeFloatImage / writeComplexImage (imgio.c:223-243) and the reader:
    the header carries the type code and shape, the data come back bit for
    bit; a truncated file and an unknown type code are errors."""
    import cdixx_amd
    rs = np.random.default_rng(code)
    a = rs.standard_normal((7, 12))
    if np.issubdtype(dtype, np.complexfloating):
        a = a + 1j * rs.standard_normal((7, 12))
    a = a.astype(dtype)
    path = str(tmp_path / "x.bin")
    cdixx_amd.bin_write(path, a)
    assert os.path.getsize(path) == 20 + a.nbytes
    assert cdixx_amd.bin_read_header(path) == (code, 7, 12)
    assert np.array_equal(cdixx_amd.bin_read(path), a)
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[:-4])
    with pytest.raises(cdixx_amd.CdxError):
        cdixx_amd.bin_read(path)
    open(path, "wb").write(bytes([9]) + raw[1:])
    with pytest.raises(cdixx_amd.CdxError):
        cdixx_amd.bin_read_header(path)
