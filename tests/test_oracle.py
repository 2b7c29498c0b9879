"""Pins for the CPU oracle (oracle/oracle.c) -- runs without a GPU.

The reference ships no golden vectors for this path (no tests at all), so
each oracle function is pinned by what the mathematics fixes or by an
independent library routine, chosen so that a plausible mistake (dropped
term, wrong sign or index, transposed operand) fails:

  FFT (a5)              brute-force DFT from the definition (square, rectangular,
                        radix-3 sizes); numpy.fft (pocketfft) at full sizes;
                        delta/plane-wave closed forms; Parseval; inverse round trip.
  applyModAbs (a3)      |F'| = sqrt(max(I,0)), phase kept, F=0 -> (sqrt I, 0),
                        idempotent.
  applyMod (a4)         beamstop untouched; inside tolerance |F'|^2 = I; outside
                        |F'|^2 = I -+ tol (tol from the noise model); phase kept;
                        over-exposed pixels keep their modulus.
  support updates (a7)  RAAR(beta=1) == HIO; ER with relax=1 is the textbook
                        projection; POSER output Re >= 0 and zero outside; HIO
                        fixed point; each branch against its defining identity.
  shrink-wrap (a8)      frequency Gaussian == periodic spatial convolution with
                        sigma_x = rows/(2 pi sigma_f) (brute force at 64^2);
                        threshold rule from closed forms (constant field, plane
                        wave, delta with Parseval, first-call average).
  residual (a11)        exact modulus match -> 0, single/double perturbations,
                        over-exposed and beamstop pixels excluded.
  real blur (a8')       scipy.ndimage.convolve with the (2r+1)^2 kernel,
                        zero padding.
  masks (a10)           disk/rectangle extents and counts.
  AlgoParser (a2)       token counts of the reference's own .cfg schedules.
  Philox4x64-10         Random123 known-answer vector and numpy.random.Philox.
  Poisson               moments and chi-square against scipy.stats.poisson.
"""
import os

import numpy as np
import pytest

from conftest import REFERENCE, rel_l2

rng = np.random.default_rng(1234)


def crand(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(8, 8), (12, 20), (16, 24), (24, 36), (6, 10), (32, 16)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_matches_bruteforce_dft(o64, shape, sign):
    x = crand(*shape)
    assert rel_l2(o64.fft2(x, sign), o64.dft2_brute(x, sign)) < 1e-13


@pytest.mark.parametrize("n", [256, 384, 512, 1024])
def test_fft_matches_numpy(o64, n):
    x = crand(n, n)
    assert rel_l2(o64.fft2(x, -1), np.fft.fft2(x)) < 1e-13
    assert rel_l2(o64.fft2(x, +1), np.fft.ifft2(x) * n * n) < 1e-13


def test_fft_closed_forms(o64):
    n = 64
    d = np.zeros((n, n), complex)
    d[0, 0] = 1
    assert np.allclose(o64.fft2(d), np.ones((n, n)))
    kx, ky = 5, 11
    x, y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    pw = np.exp(2j * np.pi * (kx * x + ky * y) / n)
    F = o64.fft2(pw, -1)
    expect = np.zeros((n, n), complex)
    expect[kx, ky] = n * n
    assert np.abs(F - expect).max() < 1e-9
    z = crand(n, n)
    assert abs(np.sum(np.abs(o64.fft2(z)) ** 2) / (n * n) - np.sum(np.abs(z) ** 2)) < 1e-8 * np.sum(np.abs(z) ** 2)
    assert rel_l2(o64.fft2(o64.fft2(z, -1), +1) / (n * n), z) < 1e-14


def test_fft_f32_accuracy(o32, o64):
    x = crand(256, 256)
    assert rel_l2(o32.fft2(x), o64.fft2(x)) < 2e-6


# ---------------------------------------------------------------- projections
def test_mod_abs_properties(o64):
    F = crand(32, 32)
    F[3, 4] = 0
    I = rng.uniform(-1, 4, (32, 32))
    out = o64.apply_mod_abs(F, I)
    target = np.sqrt(np.maximum(I, 0))
    assert np.allclose(np.abs(out), target, rtol=1e-14, atol=1e-14)
    keep = (F != 0) & (I > 0)
    assert np.allclose(np.angle(out[keep]), np.angle(F[keep]), atol=1e-12)
    assert out[3, 4] == (np.sqrt(max(I[3, 4], 0)) + 0j)
    assert np.allclose(o64.apply_mod_abs(out, I), out, rtol=1e-14, atol=1e-15)


def test_mod_noise_branches(o64):
    n = 4096
    scale, rcolor, noise = 1.0 / 0.1, 65535, 24
    F = crand(n) * 3
    I = np.abs(F) ** 2 * rng.uniform(0.5, 1.5, n)
    I[:50] = scale * rng.uniform(0.96, 2.0, 50)       # over-exposed
    bs = np.zeros(n)
    bs[50:80] = 1
    out = o64.apply_mod(F, I, bs, noise, scale, rcolor)
    assert np.array_equal(out[50:80], F[50:80])       # beamstop untouched
    src2 = np.abs(F) ** 2
    m2 = np.where(I >= 0.95 * scale, np.maximum(I, src2), I)
    tol = np.sqrt(noise + m2 * rcolor / scale) * scale / rcolor
    d = m2 - src2
    got2 = np.abs(out) ** 2
    sel = np.ones(n, bool)
    sel[50:80] = False
    inside = sel & (np.abs(d) <= tol)
    above = sel & (d > tol)
    below = sel & (d < -tol)
    assert inside.sum() > 10 and above.sum() > 10 and below.sum() > 10
    assert np.allclose(got2[inside], m2[inside], rtol=1e-12)
    assert np.allclose(got2[above], (m2 - tol)[above], rtol=1e-12)
    assert np.allclose(got2[below], (m2 + tol)[below], rtol=1e-12)
    assert np.allclose(np.angle(out[sel]), np.angle(F[sel]), atol=1e-12)
    ox = np.arange(50)[I[:50] <= src2[:50]]
    assert np.allclose(np.abs(out[ox]), np.abs(F[ox]), rtol=1e-12)  # over-exposed: modulus kept


# ---------------------------------------------------------------- support updates
@pytest.fixture
def upd_data():
    g, gp = crand(64, 64), crand(64, 64)
    spt = rng.uniform(0, 1, (64, 64))
    return g, gp, spt, 0.5


def test_support_hio_raar(o64, upd_data):
    g, gp, spt, thr = upd_data
    inside = spt > thr
    hio = o64.apply_support(g, gp, 3, spt, thr, beta=0.9)
    assert np.allclose(hio[inside], gp[inside])
    assert np.allclose(hio[~inside], g[~inside] - 0.9 * gp[~inside])
    raar = o64.apply_support(g, gp, 0, spt, thr, beta=0.85)
    assert np.allclose(raar[~inside], 0.85 * g[~inside] + (1 - 1.7) * gp[~inside])
    # at beta = 1 (the reference's fixed value) RAAR and HIO are the same map
    assert np.allclose(o64.apply_support(g, gp, 0, spt, thr, 1.0), o64.apply_support(g, gp, 3, spt, thr, 1.0))


def test_support_er_poser_poshio_fhio(o64, upd_data):
    g, gp, spt, thr = upd_data
    inside = spt > thr
    er1 = o64.apply_support(g, gp, 1, spt, thr, relax=1.0)      # textbook ER
    assert np.allclose(er1, np.where(inside, gp, 0))
    er = o64.apply_support(g, gp, 1, spt, thr)                  # over-relaxed, 1.9
    assert np.allclose(er[inside], g[inside] + 1.9 * (gp[inside] - g[inside]))
    pos = o64.apply_support(g, gp, 2, spt, thr)
    assert (pos.real >= 0).all() and (pos[~inside] == 0).all()
    keep = inside & (gp.real > 0)
    assert np.allclose(pos[keep], gp[keep])
    ph = o64.apply_support(g, gp, 4, spt, thr, beta=1.0)
    assert np.allclose(ph.imag, g.imag - gp.imag)
    assert np.allclose(ph.real[keep], g.real[keep] + 1.9 * (gp.real[keep] - g.real[keep]))
    assert np.allclose(ph.real[~keep], g.real[~keep] - gp.real[~keep])
    fh = o64.apply_support(g, gp, 5, spt, thr)
    assert np.allclose(fh[~inside], g[~inside] - 1.2 * gp[~inside])
    # unknown tokens (KKT=8) take the FHIO branch, as the dispatcher's default
    assert np.allclose(o64.apply_support(g, gp, 8, spt, thr), fh)


def test_hio_fixed_point(o64):
    g = crand(32, 32)
    spt = np.ones((32, 32))
    assert np.allclose(o64.apply_support(g, g, 3, spt, 0.5), g)


# ---------------------------------------------------------------- shrink-wrap
def test_gauss_mult_is_periodic_convolution(o64):
    n = 64
    x = rng.uniform(0, 1, (n, n)) + 0j
    sf = 8.0
    out = o64.fft2(o64.gauss_mult_freq(o64.fft2(x, -1), sf), +1) / (n * n)
    sx = n / (2 * np.pi * sf)
    # brute-force periodic convolution with the sampled, periodised Gaussian
    d = np.arange(n)
    d = np.minimum(d, n - d)
    # the frequency window is the wrapped Gaussian exp(-k^2/2sf^2); its exact
    # inverse DFT is the kernel; compare against direct evaluation via sums
    k = np.where(np.arange(n) >= n // 2, np.arange(n) - n, np.arange(n))
    G1 = np.exp(-k ** 2 / (2 * sf ** 2))
    h1 = np.real(np.fft.ifft(G1))
    h = np.outer(h1, h1)
    ref = np.zeros((n, n))
    xr = x.real
    for a in range(n):
        for b in range(n):
            ref += h[a, b] * np.roll(np.roll(xr, a, 0), b, 1)
    assert np.abs(out.real - ref).max() < 1e-12
    # and the kernel is close to the continuous Gaussian of width rows/(2 pi sf)
    # (up to the truncation of the window at the Nyquist frequency)
    cont = np.exp(-d ** 2 / (2 * sx ** 2))
    cont /= cont.sum()
    assert np.abs(h1 - cont).max() < 1e-4


def _gauss_window(n, sigma_f):
    """exp(-(kx^2 + ky^2) / 2 sigma_f^2) on wrapped frequencies (applyGaussMult
    isFreq=1, cudaConfig.cu:380-395)"""
    k = np.where(np.arange(n) >= n // 2, np.arange(n) - n, np.arange(n))
    return np.exp(-(k[:, None] ** 2 + k[None, :] ** 2) / (2 * sigma_f ** 2))


def test_shrink_wrap_threshold_closed_forms(o64):
    """The frequency shrink-wrap of cdi.cc:326-341: S = |IFFT(F G)|^2 with
    sigma_f = (rows/4)/sigma; threshold = min(max S, 5 avg) * shrinkThreshold,
    avg = 4 sum S / N taken on the first call only; sigma *= 0.99 while > 2.
    Each case has an answer fixed without the oracle's own steps:
      - a constant field c (F = c N at DC) gives S = (c N)^2 everywhere, so
        the threshold is S t (max < 5 avg);
      - a plane wave at (5, -7) gives the constant S = (c N G(k))^2, which pins
        the window's width and its wrapped coordinates (unwrapped, -7 would be
        57 and S would underflow);
      - a delta at the origin (F = 1) gives S = |sum_k G_k e^{...}|^2: its peak
        is (sum G)^2 and, by Parseval, sum S = N sum G^2, so 5 avg = 20 sum G^2
        < max and the threshold is 20 sum G^2 t;
      - a second call keeps the first call's average: a constant field brighter
        than 5 avg is then clamped to 5 avg t."""
    n, t, c = 64, 0.3, 2.0
    N = n * n
    sf = (n / 4) / 3.0
    # constant field
    st = o64.sw_state(sigma=3.0)
    F = np.zeros((n, n), complex)
    F[0, 0] = c * N
    S = o64.shrink_wrap_freq(F, t, st)
    assert np.allclose(S, (c * N) ** 2, rtol=1e-12)
    assert abs(st[2] - (c * N) ** 2 * t) <= 1e-12 * (c * N) ** 2
    assert abs(st[1] - 4 * (c * N) ** 2) <= 1e-12 * (c * N) ** 2
    assert abs(st[0] - 3.0 * 0.99) < 1e-15
    # plane wave at wrapped frequency (5, -7)
    st = o64.sw_state(sigma=3.0)
    F = np.zeros((n, n), complex)
    F[5, n - 7] = c * N
    S = o64.shrink_wrap_freq(F, t, st)
    want = (c * N * np.exp(-(5 ** 2 + 7 ** 2) / (2 * sf ** 2))) ** 2
    assert np.allclose(S, want, rtol=1e-12)
    assert abs(st[2] - want * t) <= 1e-12 * want
    # delta at the origin
    st = o64.sw_state(sigma=3.0)
    G = _gauss_window(n, sf)
    S = o64.shrink_wrap_freq(np.ones((n, n), complex), t, st)
    assert abs(S[0, 0] - G.sum() ** 2) <= 1e-12 * G.sum() ** 2 and S.max() == S[0, 0]
    assert abs(S.sum() - N * (G ** 2).sum()) <= 1e-12 * S.sum()
    assert 20 * (G ** 2).sum() < G.sum() ** 2
    assert abs(st[2] - 20 * (G ** 2).sum() * t) <= 1e-12 * st[2]
    first_avg = 4 * (G ** 2).sum()
    assert abs(st[1] - first_avg) <= 1e-12 * first_avg
    # second call: the first call's average clamps a much brighter field
    F = np.zeros((n, n), complex)
    F[0, 0] = 1e3 * N
    S = o64.shrink_wrap_freq(F, t, st)
    assert S.min() > 5 * first_avg
    assert abs(st[1] - first_avg) <= 1e-12 * first_avg
    assert abs(st[2] - 5 * first_avg * t) <= 1e-12 * st[2]
    assert abs(st[0] - 3.0 * 0.99 * 0.99) < 1e-15
    # sigma stops decaying at 2
    st = o64.sw_state(sigma=2.0)
    o64.shrink_wrap_freq(F, t, st)
    assert st[0] == 2.0


def test_residual_closed_forms(o64):
    """residual of cdi.cc:413-421: sum over pixels with I < 0.99 of
    ((|F|^2 - I) (1 - bs))^2 (getMod2, addRemoveOE(., ., -1), invert(beamstop)
    and multiply, getMod2, findSum).  An exact modulus match gives 0; a
    perturbation d of |F|^2 at one pixel gives d^2, at two pixels the sum of
    squares; over-exposed pixels (I >= 0.99) and beamstop pixels do not count;
    a beamstop value of 0.5 weights the difference by 0.5."""
    n = 16
    rs = np.random.default_rng(8)
    I = rs.uniform(0, 0.9, (n, n))
    ph = np.exp(2j * np.pi * rs.uniform(size=(n, n)))
    F = np.sqrt(I) * ph
    assert o64.residual(F, I) < 1e-28
    F1 = F.copy()
    F1[3, 4] = np.sqrt(I[3, 4] + 0.25) * ph[3, 4]
    assert abs(o64.residual(F1, I) - 0.0625) < 1e-14
    F1[9, 2] = np.sqrt(I[9, 2] + 0.5) * ph[9, 2]
    assert abs(o64.residual(F1, I) - (0.0625 + 0.25)) < 1e-14
    # over-exposed pixel: excluded whatever F holds there
    I2 = I.copy()
    I2[9, 2] = 0.99
    assert abs(o64.residual(F1, I2) - 0.0625) < 1e-14
    # beamstop: 1 excludes, 0.5 halves the difference
    bs = np.zeros((n, n))
    bs[3, 4] = 1.0
    assert abs(o64.residual(F1, I, bs) - 0.25) < 1e-14
    bs[3, 4] = 0.5
    assert abs(o64.residual(F1, I, bs) - (0.25 * 0.0625 + 0.25)) < 1e-14


def test_gauss_conv_real_matches_scipy(o64):
    from scipy.ndimage import convolve
    x = rng.uniform(0, 1, (40, 56))
    for sigma in (3.0, 1.5, 2.2):
        r = int(np.floor(sigma * 6)) >> 1
        dx = np.arange(-r, r + 1)
        k = np.exp(-(dx[:, None] ** 2 + dx[None, :] ** 2) / (2 * sigma ** 2)) / (2 * np.pi * sigma ** 2)
        ref = convolve(x, k, mode="constant", cval=0.0)
        assert np.abs(o64.gauss_conv_real(x, sigma) - ref).max() < 1e-13


# ---------------------------------------------------------------- masks
def test_masks(o64):
    n = 256
    s = o64.support_rect(n, n, 2.0)
    # rect [n/4-1, 3n/4+1) in both axes (cdi.cc:242-245, rect::isInside)
    assert s.sum() == (n // 2 + 2) ** 2 and s[n // 4 - 1, n // 4 - 1] == 1 and s[n // 4 - 2, n // 4] == 0
    bs = o64.beamstop(n, n, 8)
    assert bs[0, 0] == 1 and bs[0, 7] == 1 and bs[0, 8] == 0 and bs[n - 7, 0] == 1
    kx = np.minimum(np.arange(n), n - np.arange(n))
    assert bs.sum() == np.sum((kx[:, None] ** 2 + kx[None, :] ** 2) < 64)
    c = o64.circle(n, n, 40)
    yy = np.arange(n) - n // 2
    assert c.sum() == np.sum((yy[:, None] ** 2 + yy[None, :] ** 2) < 1600)


def test_convert_fo_is_fftshift(o64):
    x = crand(16, 24)
    assert np.array_equal(o64.convert_fo(x), np.fft.fftshift(x))


# ---------------------------------------------------------------- AlgoParser
def test_parser_reference_configs(o64):
    # token counts of the schedules in the reference's own config files
    assert len(o64.expand("100*(3*POSHIO+20*RAAR+1*POSHIO+1*shrinkWrap)+20*RAAR+1*ER")) == 2521
    assert len(o64.expand("200*(20*POSHIO+1*POSHIO+1*shrinkWrap)+100*POSHIO")) == 4500
    assert len(o64.expand("3000*RAAR+20*ER")) == 3020
    assert o64.expand("2*(1*HIO+1*shrinkWrap)+1*ER") == [3, 6, 3, 6, 1]
    assert o64.expand("1000*HIO") == [3] * 1000
    assert o64.expand("3*BOGUS") is None


def test_parser_reads_reference_cfg_files(o64):
    """The schedules of the reference's three .cfg files (fixture written by
    tests/golden/make_cfg_schedules.py) expand to their known token counts in
    the oracle and in the library's host parser; where the reference tree is
    mounted, the fixture matches the files."""
    import json
    import re
    import cdixx_amd
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_schedules.json")))
    assert [e["tokens"] for e in fx] == [2521, 4500, 3020]
    for e in fx:
        toks = o64.expand(e["algorithm"])
        assert len(toks) == e["tokens"]
        assert cdixx_amd.expand_schedule(e["algorithm"]) == toks
        path = os.path.join(REFERENCE, e["file"])
        if os.path.exists(path):
            m = re.search(r'^\s*algorithm\s*=\s*"([^"]*)"', open(path).read(), re.M)
            assert m.group(1) == e["algorithm"]


# ---------------------------------------------------------------- RNG
def test_philox_known_answers(o64):
    # Random123 philox4x64-10 known-answer vector (counter 0, key 0)
    assert [int(v) for v in o64.philox([0, 0, 0, 0], [0, 0])] == [
        0x16554D9ECA36314C, 0xDB20FE9D672D0FDC, 0xD7E772CEE186176B, 0x7E68B68AEC7BA23B]
    # numpy's Philox is the same generator (it increments the counter before use)
    for key in ([0, 0], [12345, 678], [0xDEADBEEF12345, 99]):
        bg = np.random.Philox(counter=[0, 0, 0, 0], key=key)
        assert [int(v) for v in bg.random_raw(4)] == [int(v) for v in o64.philox([1, 0, 0, 0], key)]


def test_poisson_statistics(o64):
    from scipy import stats
    for lam in (0.3, 4.0, 9.9, 10.0, 37.5, 1e4):
        s = np.array([o64.poisson(lam, 7, 2, i) for i in range(20000)])
        assert abs(s.mean() - lam) < 5 * np.sqrt(lam / len(s))
        assert abs(s.var() / lam - 1) < 0.06
        if lam < 100:
            ks = np.arange(int(lam + 6 * np.sqrt(lam)) + 3)
            obs = np.array([(s == k).sum() for k in ks])
            exp = stats.poisson.pmf(ks, lam) * len(s)
            m = exp > 5
            chi = ((obs[m] - exp[m]) ** 2 / exp[m]).sum()
            assert chi < stats.chi2.ppf(0.999, m.sum() - 1)


def test_sim_object_ranges(o64):
    obj = o64.sim_object(128, 30, seed=3)
    yy = np.arange(128) - 64
    disk = (yy[:, None] ** 2 + yy[None, :] ** 2) < 900
    a = np.abs(obj)
    assert (a[~disk] == 0).all()
    assert a[disk].min() >= 0.2 - 1e-12 and a[disk].max() <= 1 + 1e-12
    I = o64.sim_intensity(obj, 1e6, seed=3, poisson=False)
    assert abs(I.sum() - 1e6) < 1e-6 * 1e6
    assert rel_l2(I, np.abs(np.fft.fft2(obj)) ** 2 * 1e6 / np.sum(np.abs(np.fft.fft2(obj)) ** 2)) < 1e-12


# ---------------------------------------------------------------- loop
def test_iteration_composition(o64):
    """One iteration equals projection -> IFFT/N -> update -> FFT built from
    numpy's FFT and the separately pinned pieces."""
    n = 64
    F, g = crand(n, n), crand(n, n)
    I = rng.uniform(0, 3, (n, n))
    spt = o64.support_rect(n, n)
    F1, g1 = o64.iteration(F, g, I, spt, 3, 0.5, beta=0.9)
    Fp = o64.apply_mod_abs(F, I)
    gp = np.fft.ifft2(Fp)
    g_ref = o64.apply_support(g, gp, 3, spt, 0.5, beta=0.9)
    assert rel_l2(g1, g_ref) < 1e-13
    assert rel_l2(F1, np.fft.fft2(g_ref)) < 1e-13


def test_phase_retrieve_converges(o64):
    n = 64
    obj = o64.sim_object(n, 14, seed=5)
    I = o64.sim_intensity(obj, 1e9, seed=5, poisson=False)
    spt = o64.circle(n, n, 16)
    rs = np.random.default_rng(0)
    F0 = np.sqrt(I) * np.exp(2j * np.pi * rs.uniform(size=I.shape))
    g0 = np.fft.ifft2(F0) * np.sqrt(n * n)
    out = o64.phase_retrieve("300*HIO+50*ER", F0, g0, I, spt, threshold=0.5, rtrace_cap=350, beta=0.9, relax=1.0)
    assert out["n_iter"] == 350
    r = out["rtrace"]
    assert r[0] < 1e-12            # the start has the measured modulus
    assert r[-1] < 0.05 and r[-1] < 0.1 * r[1]


# ---------------------------------------------------------------------------
# PRTF pieces (cdi.cc:355-380,423-440; cudaConfig.cu:876-887,1027-1063)
# ---------------------------------------------------------------------------
def test_radial_sum_counts_angles_and_is_centred(o64):
    # a constant image: every in-range sample is 1, so the sum counts the
    # 2*cols angles until the 2x2 stencil leaves the image
    out = o64.radial_sum(np.ones((64, 64)))
    assert np.all(out[:31] == 128.0)
    assert out[31] < 128.0
    rect = o64.radial_sum(np.ones((64, 32)))
    assert np.all(rect[:15] == 64.0) and np.all(rect[17:] < 64.0)
    # a disc of radius 12 about the centre (rows/2, cols/2): full count inside,
    # nothing outside -- an off-centre sampling grid breaks both
    x, y = np.mgrid[0:64, 0:64]
    disc = (((x - 32) ** 2 + (y - 32) ** 2) < 12.5 ** 2).astype(float)
    d = o64.radial_sum(disc)
    assert np.all(d[:11] == 128.0) and np.all(d[14:] == 0.0)
    # an affine image is reproduced exactly by bilinear sampling: the sum of
    # r*cos(theta) over a full circle of angles vanishes, the offset counts
    aff = o64.radial_sum(7.0 + (x - 32) + 0.5 * (y - 32))
    assert np.allclose(aff[:31], 7.0 * 128, rtol=0, atol=1e-9)


def test_add_phase_closed_form(o64):
    a = np.array([0, 1 + 1j, 5, 0], np.complex128)
    b = np.array([2j, 0, 1e-7, -3], np.complex128)
    out = o64.add_phase(a, b, np.pi / 2)
    # unit phasor of b rotated by pi/2, added where |b| > 1e-6
    assert np.allclose(out, [-1, 1 + 1j, 5, -1j], atol=1e-15)


def _fixed_point(o64, n=64, r=10, seed=5):
    """An exact fixed point of the loop: F = FFT(obj), I = |F|^2, g =
    IFFT(F)/sqrt(N) with obj inside the support."""
    rs = np.random.default_rng(seed)
    x, y = np.mgrid[0:n, 0:n]
    spt = (((x - n // 2) ** 2 + (y - n // 2) ** 2) < r * r).astype(float)
    obj = spt * (0.5 + rs.uniform(size=(n, n))) * np.exp(1j * rs.uniform(-1, 1, (n, n)))
    F = np.fft.fft2(obj)
    return F, np.fft.ifft2(F) * n, np.abs(F) ** 2, spt


def test_prtf_is_one_at_a_fixed_point(o64):
    F, g, I, spt = _fixed_point(o64)
    out = o64.phase_retrieve("12*ER", F, g, I, spt, 0.5, prtf_iter=4, relax=1.0)
    assert out["prtf_count"] == 12 - 4
    w, p = out["weights"], out["prtf"]
    assert w[0] > 0 and np.all(w[:30] > 100)
    ok = w > 0
    assert np.allclose(p[ok] / w[ok], 1.0, rtol=1e-12)


def test_prtf_counts_and_accumulates_partially(o64):
    F, g, I, spt = _fixed_point(o64, seed=6)
    # shrinkWrap tokens advance the counter but accumulate nothing; the
    # divisor is still iter - prtfIter (cdi.cc:425)
    out = o64.phase_retrieve("6*ER+2*shrinkWrap+4*ER", F, g, I, spt, 0.5, prtf_iter=4, relax=1.0,
                             shrinkThreshold=0.01)
    assert out["n_iter"] == 12 and out["prtf_count"] == 8
    ok = out["weights"] > 0
    # 6 of the 8 counted steps accumulate unit phasors of a fixed F
    assert np.allclose(out["prtf"][ok] / out["weights"][ok], 6 / 8, rtol=1e-9)
    # prtfIter beyond the schedule: nothing accumulated
    none = o64.phase_retrieve("5*ER", F, g, I, spt, 0.5, prtf_iter=10, relax=1.0)
    assert none["prtf_count"] == 0 and not none["prtf"].any() and not none["weights"].any()


# ---------------------------------------------------------------------------
# pattern preprocessing (util/processimg.cpp)
# ---------------------------------------------------------------------------
def test_preprocess_pieces_closed_forms():
    from oracle import preprocess_ref as pp
    # a band-limited periodic image shifted by a fraction of a pixel is the
    # same sum of plane waves evaluated at shifted coordinates
    n = 64
    x, y = np.mgrid[0:n, 0:n]
    f = lambda xx, yy: np.cos(2 * np.pi * (3 * xx + 5 * yy) / n) + 0.5 * np.sin(2 * np.pi * (7 * xx - 2 * yy) / n)
    out = pp.shift_wave(f(x, y).astype(complex), 0.3, -0.45)
    assert np.allclose(out.real, f(x - 0.3, y + 0.45), atol=1e-12) and np.allclose(out.imag, 0, atol=1e-12)
    # integer shifts are circular shifts
    z = np.random.default_rng(0).standard_normal((n, n))
    assert np.allclose(pp.shift_wave(z, 2, -3).real, np.roll(z, (2, -3), axis=(0, 1)), atol=1e-12)
    # binning of a ramp: the block means
    a = np.arange(36.0).reshape(6, 6)
    assert np.allclose(pp.merge_pixels(a, 2), [[3.5, 5.5, 7.5], [15.5, 17.5, 19.5], [27.5, 29.5, 31.5]])
    # the centroid of a symmetric blob at (x0, y0) sits at x0 + 0.5 - R/2
    R, C = 80, 96
    X, Y = np.mgrid[0:R, 0:C]
    blob = np.exp(-((X - 30) ** 2 + (Y - 61) ** 2) / 20.0)
    mx, my = pp.centroid(blob)
    assert abs(mx - (30 + 0.5 - R / 2)) < 1e-4 and abs(my - (61 + 0.5 - C / 2)) < 1e-4
    # shiftMiddle re-centres: a displaced symmetric blob comes back centred
    img = np.exp(-((x - 40) ** 2 + (y - 25) ** 2) / 8.0)
    c = pp.shift_middle(img.astype(complex)).real
    assert np.unravel_index(np.argmax(c), c.shape) == (n // 2, n // 2)


# ---------------------------------------------------------------------------
# setup steps of CDI::prepareIter / readObjectWave (cdi.cc:62-82,161-200)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("frame,obj", [((12, 16), (5, 6)), ((16, 16), (8, 8)), ((9, 14), (4, 14))])
def test_create_wavefront_is_a_centred_padding(o64, frame, obj):
    """createWaveFront (cudaConfig.cu:688-709) places sqrt(max(I,0)) e^{2 pi i
    (P - 0.5)} with zero margins of (frame - obj) // 2 before the image: the
    same array as numpy's pad of the complex object, odd margins included."""
    rs = np.random.default_rng(21)
    I = rs.uniform(-0.5, 4, obj)
    P = rs.uniform(0, 1, obj)
    (R, C), (r, c) = frame, obj
    pad = ((( R - r) // 2, R - r - (R - r) // 2), ((C - c) // 2, C - c - (C - c) // 2))
    want = np.pad(np.sqrt(np.maximum(I, 0)) * np.exp(2j * np.pi * (P - 0.5)), pad)
    assert np.abs(o64.create_wavefront(R, C, I, P) - want).max() < 1e-14
    assert np.array_equal(o64.create_wavefront(R, C, I, None), np.pad(np.sqrt(np.maximum(I, 0)), pad) + 0j)
    assert np.array_equal(o64.create_wavefront(R, C, None, None, shape=obj), np.pad(np.ones(obj), pad) + 0j)


def test_convert_fo_phase_is_the_spectrum_of_the_centred_object(o64):
    """convertFOPhase (cudaConfig.cu:861-870) multiplies by (-1)^(x+y) norm.
    For even sizes (-1)^(kx+ky) FFT(x) = FFT(fftshift(x)) (shift theorem), so
    prepareIter's convertFOPhase(FFT(obj), 1/sqrt(N)) is the unitary spectrum
    of the object moved from the frame centre to the origin."""
    rs = np.random.default_rng(22)
    x = rs.standard_normal((16, 24)) + 1j * rs.standard_normal((16, 24))
    N = x.size
    got = o64.convert_fo_phase(np.fft.fft2(x), 1 / np.sqrt(N))
    assert rel_l2(got, np.fft.fft2(np.fft.fftshift(x)) / np.sqrt(N)) < 1e-14
    # norm 1 flips the odd checkerboard only
    y = o64.convert_fo_phase(x, 1.0)
    X, Y = np.mgrid[0:16, 0:24]
    assert np.array_equal(y, np.where((X + Y) % 2 == 1, -x, x))


def test_poisson_noise_wo_quantises_and_adds_dark_counts(o64):
    """applyPoissonNoise_WO (cdi.cu:291-296): w = scale (int(w rcolor/scale) +
    D - noise) / rcolor, D ~ Poisson(noise).  Without noise it is truncation
    toward zero to steps of scale/rcolor; with noise the dark counts D are
    non-negative integers with mean and variance noise."""
    scale, rcolor = 2.0, 8.0
    w = np.array([0.30, -0.30, 0.25, 1.99, -1.99, 0.0, 5.0])
    want = scale * np.trunc(w * rcolor / scale) / rcolor
    assert np.array_equal(o64.poisson_noise_wo(w, 0.0, scale, rcolor), want)
    rs = np.random.default_rng(23)
    w = rs.uniform(0, 3, 200_000)
    out = o64.poisson_noise_wo(w, 24.0, scale, rcolor, seed=5)
    D = out * rcolor / scale - np.trunc(w * rcolor / scale) + 24.0
    assert np.allclose(D, np.round(D), atol=1e-9) and D.min() >= 0
    assert abs(D.mean() - 24) < 0.05 and abs(D.var() / 24 - 1) < 0.02


def test_extend_to_complex(o64):
    a = np.random.default_rng(24).standard_normal((8, 8))
    assert np.array_equal(o64.extend_to_complex(a), a + 0j)

