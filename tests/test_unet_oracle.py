"""Pins of the U-Net oracle (oracle/unet_ref.py) on the CPU: BF16 rounding,
the convolution's tap orientation against a direct loop, and the backward
pass against central finite differences in float64."""
import numpy as np
import pytest

from oracle.unet_ref import UNetRef, round_bf16


def layer_table(size, levels, base):
    """The parameter layout of cdx_unet_create: per level (encoder a, b), per
    level below the bottom (decoder a, b), then the 1x1 head; weights
    [cout][Kp] (Kp = roundup(taps*cin, 32)) followed by biases padded to 8."""
    rows, off = [], 0

    def add(cin, cout, taps):
        nonlocal off
        kp = (taps * cin + 31) // 32 * 32
        w = off
        off += cout * kp
        b = off
        off += (cout + 7) // 8 * 8
        rows.append((cin, cout, taps, kp, w, b))

    for l in range(levels):
        add(8 if l == 0 else base << (l - 1), base << l, 9)
        add(base << l, base << l, 9)
    for l in range(levels - 1):
        add((base << (l + 1)) + (base << l), base << l, 9)
        add(base << l, base << l, 9)
    add(base, 8, 1)
    return rows, off


def test_round_bf16_ties_to_even():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5, 3.0e-3], np.float64)
    r = round_bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0            # tie -> even mantissa
    assert r[2] == 1 + 2 ** -6                     # tie -> even (up)
    assert r[3] == 1 + 2 ** -7 and r[4] == -2.5    # representable
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_conv_taps_against_direct_loop():
    rs = np.random.default_rng(0)
    net = UNetRef([], 8, 1, 8, bf16=False)
    x = rs.standard_normal((1, 5, 6, 2))
    w = rs.standard_normal((3, 9, 2))
    b = rs.standard_normal(3)
    y = net.conv(x, w, b, relu=False)
    ref = np.zeros((1, 5, 6, 3))
    for h in range(5):
        for ww in range(6):
            for co in range(3):
                s = b[co]
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        if 0 <= h + dy < 5 and 0 <= ww + dx < 6:
                            s += x[0, h + dy, ww + dx] @ w[co, 3 * (dy + 1) + dx + 1]
                ref[0, h, ww, co] = s
    assert np.allclose(y, ref, atol=1e-12)


@pytest.mark.parametrize("levels", [1, 2, 3])
def test_backward_matches_finite_differences(levels):
    size, base = 8, 8
    rows, n = layer_table(size, levels, base)
    net = UNetRef(rows, size, levels, base, bf16=False)
    rs = np.random.default_rng(levels)
    flat = rs.standard_normal(n) * 0.3
    x = rs.uniform(0, 1, (2, size, size))
    tgt = np.stack([rs.uniform(0, 1, (2, size, size)), rs.uniform(-3, 3, (2, size, size)),
                    (rs.uniform(size=(2, size, size)) > 0.5).astype(float)], axis=-1)

    def L(f):
        y, _ = net.forward(x, net.unpack(f))
        return net.loss(y, tgt)[0]

    y, cache = net.forward(x, net.unpack(flat))
    _, dy = net.loss(y, tgt)
    g = net.pack_grads(net.backward(dy, cache), n)
    eps = 1e-6
    for cin, cout, taps, kp, wo, bo in rows:
        for idx in [wo + rs.integers(cout) * kp + rs.integers(taps * cin), bo + rs.integers(cout)]:
            fp, fm = flat.copy(), flat.copy()
            fp[idx] += eps
            fm[idx] -= eps
            fd = (L(fp) - L(fm)) / (2 * eps)
            assert abs(fd - g[idx]) <= 1e-6 + 1e-4 * abs(fd), (cin, cout, idx, fd, g[idx])
