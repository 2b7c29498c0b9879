"""The level schedule of the ptychographic sweep (cdx_ptycho_levels, host
only).  The GPU sweep runs the positions of one group concurrently, so the
object it leaves equals the sequential loop's (util/ptycho.cpp:514) exactly
when every position appears once, the windows of a group are pairwise
disjoint inside the object, and a position comes after every earlier
position (in the caller's order) whose window overlaps its own.  Checked on
random scans with windows partly and wholly outside the object, against
these properties and a brute-force restatement of the levels."""
import numpy as np
import pytest


def _clip(n, shape, a, b):
    return max(a, 0), min(a + n, shape[0]), max(b, 0), min(b + n, shape[1])


def _overlap(r, s):
    return r[0] < r[1] and r[2] < r[3] and s[0] < s[1] and s[2] < s[3] and \
        r[0] < s[1] and s[0] < r[1] and r[2] < s[3] and s[2] < r[3]


def _brute(n, shape, sx, sy, order, cap):
    rects = [_clip(n, shape, sx[i], sy[i]) for i in order]
    lvl = []
    for j in range(len(order)):
        lvl.append(max([lvl[i] + 1 for i in range(j) if _overlap(rects[i], rects[j])], default=0))
    groups = []
    for L in range(max(lvl) + 1 if lvl else 0):
        members = [order[j] for j in range(len(order)) if lvl[j] == L]
        groups += [members[a:a + cap] for a in range(0, len(members), cap)]
    return groups


@pytest.mark.parametrize("seed", range(6))
def test_schedule_is_sequentially_equivalent(seed):
    import cdixx_amd as cdx
    rs = np.random.default_rng(seed)
    n = int(rs.choice([64, 128, 256]))
    shape = (int(rs.integers(n // 2, 6 * n)), int(rs.integers(n // 2, 6 * n)))
    m = int(rs.integers(1, 400))
    sx = rs.integers(-n - 10, shape[0] + 10, m)
    sy = rs.integers(-n - 10, shape[1] + 10, m)
    order = rs.permutation(m)[: int(rs.integers(1, m + 1))]
    cap = int(rs.choice([0, 1, 3, 64]))
    groups = cdx.ptycho_levels(n, shape, sx, sy, order=order, max_concurrent=cap)
    flat = [i for g in groups for i in g]
    assert sorted(flat) == sorted(order.tolist())
    assert all(1 <= len(g) <= (cap or 64) for g in groups)
    where = {s: k for k, g in enumerate(groups) for s in g}
    rects = {i: _clip(n, shape, sx[i], sy[i]) for i in order}
    for g in groups:
        for a in range(len(g)):
            for b in range(a + 1, len(g)):
                assert not _overlap(rects[g[a]], rects[g[b]])
    pos = {s: j for j, s in enumerate(order)}
    for i in order:
        for j in order:
            if pos[i] < pos[j] and _overlap(rects[i], rects[j]):
                assert where[i] < where[j]
    assert groups == _brute(n, shape, sx, sy, order.tolist(), cap or 64)


def test_raster_schedule_and_errors():
    import cdixx_amd as cdx
    n, shape = 64, (420, 400)
    xs = np.arange(-12, 420 - 64 + 13, 16)
    ys = np.arange(-12, 400 - 64 + 13, 16)
    sx, sy = np.repeat(xs, len(ys)), np.tile(ys, len(xs))
    groups = cdx.ptycho_levels(n, shape, sx, sy)
    assert len(groups) < len(sx) // 4  # several disjoint windows per level
    assert len(cdx.ptycho_levels(n, shape, sx, sy, max_concurrent=1)) == len(sx)
    with pytest.raises(cdx.CdxError):
        cdx.ptycho_levels(n, shape, sx, sy, order=[0, len(sx)])
