"""cdx_phase_retrieve_static (the loop with a static workspace, for callers
that cannot allocate device memory) against cdx_phase_retrieve on the same
inputs: two calls in a row, the second on a smaller frame, each matching the
context entry point's F, g, support and threshold (same kernels; FP32, held
to the single-iteration bar of 1e-5 relative L2, threshold to 1e-6).  Also:
frames beyond the 512 x 512 workspace and PRTF requests are refused."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

ALG = "3*RAAR+1*shrinkWrap+2*ER"


@pytest.fixture(scope="module")
def cdx():
    import cdixx_amd
    return cdixx_amd


def _inputs(cdx, n, seed):
    ctx = cdx.Context(n)
    obj, I, bs, spt, F, g = ctx.cplx(), ctx.real(), ctx.real(), ctx.real(), ctx.cplx(), ctx.cplx()
    ctx.sim_object(obj, 0.16 * n, seed=seed)
    ctx.sim_intensity(obj, I, 1e7, seed=seed, poisson=True)
    ctx.create_beamstop(bs, 4)
    ctx.create_circle(spt, 0.2 * n)
    ctx.random_start(I, bs, F, seed=seed, mode=1)
    ctx.fft(F, g, cdx.INVERSE)
    ctx.apply_norm(g, 1.0 / n)
    out = {k: (ctx.get_cplx(v) if k in ("F", "g") else ctx.get_real(v))
           for k, v in dict(F=F, g=g, I=I, bs=bs, spt=spt).items()}
    ctx.close()
    return out


def test_static_entry_point_matches_context_entry_point(cdx):
    p = cdx.Params.default(beta=0.85, useBS=1, freeBeamstop=1, shrinkThreshold=0.05)
    for n, seed in ((512, 1), (256, 2)):
        print(f"static workspace n={n}", flush=True)
        d = _inputs(cdx, n, seed)
        ctx = cdx.Context(n)
        ctx.set_threshold(0.5)
        dF, dg, dI, ds, dbs = ctx.cplx(d["F"]), ctx.cplx(d["g"]), ctx.real(d["I"]), ctx.real(d["spt"]), ctx.real(d["bs"])
        nit_ref, _ = ctx.phase_retrieve(ALG, p, dF, dg, dI, ds, bs=dbs)
        ref = dict(F=ctx.get_cplx(dF), g=ctx.get_cplx(dg), spt=ctx.get_real(ds), thr=ctx.threshold())
        ctx.close()
        sF, sg = cdx.DeviceBuffer.from_array(d["F"]), cdx.DeviceBuffer.from_array(d["g"])
        sI, ss, sbs = (cdx.DeviceBuffer.from_array(d[k]) for k in ("I", "spt", "bs"))
        nit, thr = cdx.phase_retrieve_static(n, n, ALG, p, sF, sg, sI, ss, bs=sbs, threshold=0.5)
        assert nit == nit_ref == 6
        got = dict(F=sF.download(np.complex64, (n, n)), g=sg.download(np.complex64, (n, n)),
                   spt=ss.download(np.float32, (n, n)))
        same = all(np.array_equal(got[k], ref[k]) for k in got) and thr == ref["thr"]
        eF, eg, es = rel_l2(got["F"], ref["F"]), rel_l2(got["g"], ref["g"]), rel_l2(got["spt"], ref["spt"])
        print(f"  F {eF:.2e} g {eg:.2e} spt {es:.2e} threshold {thr:.6g} vs {ref['thr']:.6g}; bit-identical: {same}",
              flush=True)
        assert eF < 1e-5 and eg < 1e-5 and es < 1e-5
        assert abs(thr - ref["thr"]) <= 1e-6 * abs(ref["thr"])
        for b in (sF, sg, sI, ss, sbs):
            b.free()


def test_static_entry_point_refuses_what_it_cannot_hold(cdx):
    p = cdx.Params.default()
    b = cdx.DeviceBuffer(8)
    with pytest.raises(cdx.CdxError) as e:
        cdx.phase_retrieve_static(1024, 1024, "1*HIO", p, b, b, b, b)
    assert e.value.code == 1 and "static workspace" in str(e.value)
    with pytest.raises(cdx.CdxError) as e:
        cdx.phase_retrieve_static(256, 256, "1*HIO", cdx.Params.default(prtfIter=0), b, b, b, b)
    assert e.value.code == 1
    b.free()
