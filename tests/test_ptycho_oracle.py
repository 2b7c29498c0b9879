"""Pins of the ptychography oracle (oracle/ptycho_ref.py) by what the
mathematics fixes, independent of its own formulas:

  * the forward model against a brute-force DFT;
  * applyModAccurate's branches by hand-computed values, applyMod against the
    C oracle's (already pinned) applyMod;
  * the true object and probe are a fixed point of a sweep;
  * with a constant probe the object update makes the window's exit wave
    satisfy the measured modulus exactly (pins the rPIE direction, conjugate
    and denominator), and with a constant object and zero object step the
    probe step does the same for the probe;
  * a sweep from a perturbed object lowers the data residual;
  * windows partly or wholly outside the object gather zeros and never write
    outside.
"""
import numpy as np
import pytest

from oracle import ptycho_ref as pr

rng = np.random.default_rng(7)


def crand(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def smooth_object(shape, seed=0):
    rs = np.random.default_rng(seed)
    a = rs.uniform(0, 1, shape)
    ph = rs.uniform(-1, 1, shape)
    k = np.exp(-np.arange(-6, 7) ** 2 / 8.0)
    k /= k.sum()
    for ax in (0, 1):
        a = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), ax, a)
        ph = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), ax, ph)
    return (0.3 + a) * np.exp(1j * np.pi * ph)


def disk_probe(n, r):
    x, y = np.mgrid[0:n, 0:n] - n / 2
    q = np.exp(-(x ** 2 + y ** 2) / (2 * r * r)) * np.exp(1j * 0.002 * (x ** 2 + y ** 2))
    return q * ((x ** 2 + y ** 2) < (1.6 * r) ** 2)


def raster(obj_shape, n, step, margin=0):
    sx, sy = [], []
    for a in range(-margin, obj_shape[0] - n + margin + 1, step):
        for b in range(-margin, obj_shape[1] - n + margin + 1, step):
            sx.append(a)
            sy.append(b)
    return np.array(sx), np.array(sy)


def test_simulate_matches_bruteforce_dft():
    n = 8
    obj = crand(13, 11)
    probe = crand(n, n)
    sx, sy = np.array([0, 3, -2, 8]), np.array([1, -3, 5, 6])
    pats = pr.simulate(obj, probe, sx, sy)
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    for i in range(len(sx)):
        w = np.zeros((n, n), complex)
        for x in range(n):
            for y in range(n):
                X, Y = x + sx[i], y + sy[i]
                if 0 <= X < 13 and 0 <= Y < 11:
                    w[x, y] = obj[X, Y]
        F = W @ (probe * w) @ W.T
        assert np.allclose(pats[i], np.abs(F) ** 2 / (n * n), rtol=1e-12, atol=1e-12)


def test_mod_accurate_branches():
    F = np.array([3 + 4j, 0j, 4 + 0j, 1 + 0j, 3 + 4j, 2 - 1j])
    I = np.array([1.0, 1.0, 2.0, 2.0, 9.0, -1.0])
    bs = np.array([1.0, 0, 0, 0, 0, 0])
    out = pr.mod_accurate(F, I, bs, 1.0, 1.0)
    assert out[0] == 3 + 4j                      # beamstop, bsnorm 1: unchanged
    assert out[1] == 0                           # |F| = 0: left as is
    assert out[2] == 4                           # over-exposed (I >= 0.95 scale) and |F|^2 >= I: kept
    assert np.isclose(out[3], np.sqrt(2.0))      # over-exposed but |F|^2 < I: modulus set
    assert out[4] == 3 + 4j                      # I = 9 is over-exposed at scale 1, |F|^2 = 25 >= 9: kept
    assert out[5] == 0                           # I < 0 counts as 0
    big = pr.mod_accurate(F, I, bs, 1.0, 100.0)
    assert np.isclose(big[2], np.sqrt(2.0))      # not over-exposed at scale 100: |F| = 4 -> sqrt(2)
    assert np.isclose(big[4], 1.8 + 2.4j)        # |F| = 5 -> 3, phase kept
    half = pr.mod_accurate(F, I, bs, 0.5, 100.0)
    assert half[0] == 1.5 + 2j                   # beamstop pixel written back scaled by bsnorm
    assert np.isclose(half[4], 1.8 + 2.4j)       # the modulus target does not depend on bsnorm


def test_mod_noise_matches_c_oracle(o64):
    n = 2048
    scale, rcolor, noise = 10.0, 65535, 24
    F = crand(n) * 3
    I = np.abs(F) ** 2 * rng.uniform(0.5, 1.5, n)
    I[:40] = scale * rng.uniform(0.96, 2.0, 40)
    F[40:45] = 0
    bs = np.zeros(n)
    bs[100:130] = 1
    for b in (1.0, 0.37):
        got = pr.mod_noise(F, I, bs, noise, b, scale, rcolor)
        ref = o64.apply_mod(F * b, I, bs, noise, scale, rcolor)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


def _setup(n=32, shape=(70, 60), step=8, margin=6, seed=0, bs_radius=0):
    obj = smooth_object(shape, seed)
    probe = disk_probe(n, n / 6)
    sx, sy = raster(shape, n, step, margin)
    pats = pr.simulate(obj, probe, sx, sy)
    bs = None
    if bs_radius:
        k = np.fft.fftfreq(n) * n
        bs = ((k[:, None] ** 2 + k[None, :] ** 2) < bs_radius ** 2).astype(float)
    return obj, probe, sx, sy, pats, bs


@pytest.mark.parametrize("mode", ["accurate", "noise"])
def test_truth_is_a_fixed_point(mode):
    obj, probe, sx, sy, pats, bs = _setup(bs_radius=2)
    o = obj.copy()
    order = rng.permutation(len(sx))
    pr.sweep(o, probe, pats, bs, sx, sy, order, mode=mode, scale=1e6)
    assert np.allclose(o, obj, rtol=0, atol=1e-9)


def test_constant_probe_update_meets_the_modulus():
    n = 16
    c = 0.7 - 0.4j
    probe = np.full((n, n), c)
    obj = crand(30, 30)
    I = np.abs(np.fft.fft2(probe * pr.get_window(smooth_object((30, 30)), 5, 7, n))) ** 2 / (n * n)
    o = obj.copy()
    pr.position_update(o, probe, I, None, 5, 7, probe_max=abs(c) ** 2, scale=1e6)
    w = pr.get_window(o, 5, 7, n)
    a = np.abs(np.fft.fft2(c * w)) / n
    assert np.allclose(a, np.sqrt(I), rtol=1e-9, atol=1e-12)
    # outside the window the object is untouched
    mask = np.ones(o.shape, bool)
    mask[5:5 + n, 7:7 + n] = False
    assert np.array_equal(o[mask], obj[mask])


def test_rpie_denominator_weights():
    # max|P|^2 = 4|c|^2: denominator (1-0.1)|c|^2 + 0.1*4|c|^2 = 1.3|c|^2, so the
    # exit wave moves 1/1.3 of the way to its modulus projection
    n = 16
    c = 0.6 + 0.8j
    probe = np.full((n, n), c)
    obj = smooth_object((30, 30), seed=4)
    I = np.abs(np.fft.fft2(probe * pr.get_window(smooth_object((30, 30), seed=5), 2, 3, n))) ** 2 / (n * n)
    psi = c * pr.get_window(obj, 2, 3, n)
    F = np.fft.fft2(psi)
    proj = np.fft.ifft2(np.sqrt(I) * n * F / np.abs(F))
    o = obj.copy()
    pr.position_update(o, probe, I, None, 2, 3, probe_max=4 * abs(c) ** 2, scale=1e6)
    assert np.allclose(c * pr.get_window(o, 2, 3, n), psi - (psi - proj) / 1.3, rtol=1e-9, atol=1e-12)


def test_constant_object_probe_step_meets_the_modulus():
    n = 16
    c0 = 1.3 + 0.5j
    obj = np.full((40, 40), c0)
    probe = crand(n, n)
    I = np.abs(np.fft.fft2(disk_probe(n, 3) * c0)) ** 2 / (n * n)
    term = pr.position_update(obj.copy(), probe, I, None, 10, 12, scale=1e6, update_probe=True, obj_step=0.0,
                              obj_max=abs(c0) ** 2, probe_step_size=1.0)
    a = np.abs(np.fft.fft2((probe + term) * c0)) / n
    assert np.allclose(a, np.sqrt(I), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("update_probe", [False, True])
def test_sweep_lowers_the_residual(update_probe):
    obj, probe, sx, sy, pats, bs = _setup(bs_radius=1)
    o = obj * np.exp(0.3j * np.random.default_rng(3).standard_normal(obj.shape)) * 0.8
    r0 = pr.data_residual(o, probe, pats, bs, sx, sy)
    P = probe.copy()
    res = [r0]
    for k in range(4):
        step = np.zeros_like(P)
        pr.sweep(o, P, pats, bs, sx, sy, rng.permutation(len(sx)), probe_step=step, scale=1e6,
                 update_probe=update_probe, obj_max=float(np.max(np.abs(o) ** 2)), probe_step_size=1.0 / len(sx))
        P += step  # ptycho.cpp:563-566
        res.append(pr.data_residual(o, P, pats, bs, sx, sy))
    assert res[1] < 0.7 * r0 and res[-1] < 0.3 * r0, res


def test_windows_at_the_edges():
    obj = crand(20, 24)
    n = 8
    w = pr.get_window(obj, -3, 20, n)
    assert np.array_equal(w[3:, :4], obj[0:5, 20:24])
    assert np.all(w[:3] == 0) and np.all(w[:, 4:] == 0)
    assert np.all(pr.get_window(obj, 40, 0, n) == 0)
    o = obj.copy()
    pr.update_window(o, -3, 20, np.full((n, n), 5 + 0j))
    assert np.all(o[0:5, 20:24] == 5)
    mask = np.ones(o.shape, bool)
    mask[0:5, 20:24] = False
    assert np.array_equal(o[mask], obj[mask])
