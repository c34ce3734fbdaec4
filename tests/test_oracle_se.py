"""Pins the oracle's 18- and 26-neighbour structuring elements (the metric
option of the dilation) against scipy.ndimage.binary_dilation, known-answer
point pairs and closed forms.

The 18 element is the one of the reference's second engine with CHEBYSHEV 1
(forBench/HausdorffDistance.cuh:7-8): its offsets are listed below as data,
one line of forBench/HausdorffDistance.cu each, so the scipy check does not
share the oracle's rule for building them."""
from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle

# (dx, dy, dz) offsets written by forBench/HausdorffDistance.cu dilate()
FORBENCH_18 = [
    (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),   # :37-44
    (1, -1, 0), (-1, -1, 0), (1, 1, 0), (-1, 1, 0),                          # :48-51
    (1, -1, 1), (-1, -1, 1), (1, 1, 1), (-1, 1, 1),                          # :53-56
    (1, -1, -1), (-1, -1, -1), (1, 1, -1), (-1, 1, -1),                      # :57-60
]


def structure(conn):
    """3x3x3 [dz, dy, dx] structuring element for scipy."""
    st = np.zeros((3, 3, 3), bool)
    st[1, 1, 1] = True
    if conn == 6:
        offs = FORBENCH_18[:6]
    elif conn == 18:
        offs = FORBENCH_18
    else:
        offs = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]
    for dx, dy, dz in offs:
        st[dz + 1, dy + 1, dx + 1] = True
    return st


def scipy_expect(g, s, conn, r=1.0):
    """First-hit d(v) by iterated scipy dilation (border voxels clipped, as
    the reference's bounds checks do); the result fields the oracle reports."""
    st = structure(conn)
    out = {}
    maps = np.zeros(g.shape, np.uint16)
    for tag, src, tgt in (("fp", g, s & ~g), ("fn", s, g & ~s)):
        n = int(tgt.sum())
        cut = int(np.ceil(r * n)) if n else 0
        cur, t, cov, hd, tot = src.copy(), 0, 0, 0, 0
        while cov < cut:
            t += 1
            nxt = ndimage.binary_dilation(cur, structure=st)
            new = nxt & ~cur & tgt
            k = int(new.sum())
            if k:
                hd = t
            cov += k
            tot += k * t
            maps[new] = t
            cur = nxt
            assert t < 4 * sum(g.shape)
        out["hd_" + tag], out["sum_" + tag], out["cnt_" + tag] = hd, tot, cov
    out["hd"] = max(out["hd_fp"], out["hd_fn"])
    return out, maps


KEYS = ("hd", "hd_fp", "hd_fn", "sum_fp", "cnt_fp", "sum_fn", "cnt_fn")
METHODS = [("naive", True), ("tiled", True), ("tiled", False)]


def run_all(g, s, conn, r=1.0, want_map=False):
    outs = []
    for method, crop in METHODS:
        rc, res, _, _, mp = oracle.run(g.astype(np.uint8), s.astype(np.uint8), robustness=r,
                                       method=method, crop=crop, conn=conn, want_map=want_map)
        assert rc == 0, (method, crop, rc)
        outs.append((method, crop, res, mp))
    return outs


def points(shape_zyx, a_xyz, b_xyz):
    g = np.zeros(shape_zyx, bool)
    s = np.zeros(shape_zyx, bool)
    g[a_xyz[2], a_xyz[1], a_xyz[0]] = True
    s[b_xyz[2], b_xyz[1], b_xyz[0]] = True
    return g, s


@pytest.mark.parametrize("conn", [6, 18, 26])
@pytest.mark.parametrize("seed", range(12))
def test_se_random_volumes_match_scipy(conn, seed):
    rng = np.random.default_rng(1000 * conn + seed)
    nz, ny, nx = (int(v) for v in rng.integers(3, 20, 3))
    if seed % 3 == 0:
        nx = int(rng.integers(60, 140))  # several 64-bit words per row
    p = [0.002, 0.01, 0.05][seed % 3]
    g = rng.random((nz, ny, nx)) < p
    s = rng.random((nz, ny, nx)) < p
    g[0, 0, 0] = True
    s[-1, -1, -1] = True
    want, wmap = scipy_expect(g, s, conn)
    for method, crop, res, mp in run_all(g, s, conn, want_map=True):
        assert {k: res[k] for k in KEYS} == {k: want[k] for k in KEYS}, (method, crop)
        np.testing.assert_array_equal(mp, wmap)


@pytest.mark.parametrize("conn", [18, 26])
@pytest.mark.parametrize("r", [0.5, 0.9])
def test_se_robust_cut_matches_scipy(conn, r):
    rng = np.random.default_rng(77 + conn)
    g = np.zeros((30, 40, 70), bool)
    s = np.zeros_like(g)
    g[10:20, 10:25, 20:50] = True
    s[12:22, 14:30, 18:45] = True
    s |= rng.random(g.shape) < 0.003
    want, _ = scipy_expect(g, s, conn, r)
    for method, crop, res, _ in run_all(g, s, conn, r):
        assert {k: res[k] for k in KEYS} == {k: want[k] for k in KEYS}, (method, crop)


# offset (dx, dy, dz) from a gold voxel to a segm voxel, with room around both
# -> HD.  The forBench element has the xy diagonals and the corners but no
# xz / yz diagonals, so (1,0,1) and (0,1,1) take two steps, (1,1,1) one; the
# textbook 18-neighbourhood would give the opposite.
KAT_18 = [((1, 0, 0), 1), ((1, 1, 0), 1), ((1, 1, 1), 1), ((1, 0, 1), 2), ((0, 1, 1), 2),
          ((2, 0, 2), 2), ((0, 2, 2), 2), ((3, 1, 0), 3), ((0, 0, 3), 3), ((2, 2, 2), 2),
          ((1, 0, 2), 3), ((4, 0, 4), 4)]


@pytest.mark.parametrize("off,hd", KAT_18)
def test_forbench_element_known_answers(off, hd):
    a = (10, 10, 10)
    b = tuple(a[k] + off[k] for k in range(3))
    g, s = points((24, 24, 24), a, b)
    assert scipy_expect(g, s, 18)[0]["hd"] == hd
    for method, crop, res, _ in run_all(g, s, 18):
        assert res["hd"] == hd and res["sum_fp"] == hd, (method, crop)


@pytest.mark.parametrize("seed", range(6))
def test_cube_element_is_chebyshev_distance(seed):
    """26 element: HD of two single voxels is their L-infinity distance."""
    rng = np.random.default_rng(seed)
    shape = (20, 24, 90)
    a = tuple(int(v) for v in rng.integers(0, [90, 24, 20]))
    b = tuple(int(v) for v in rng.integers(0, [90, 24, 20]))
    if a == b:
        return
    g, s = points(shape, a, b)
    cheb = max(abs(a[k] - b[k]) for k in range(3))
    for method, crop, res, _ in run_all(g, s, 26):
        assert res["hd"] == cheb and res["sum_fp"] == cheb and res["sum_fn"] == cheb


@pytest.mark.parametrize("name,shape,a,b", [
    # both voxels on the only row of the tail y tile (ny = 9): the 18 element's
    # two-step path (1,1,1) + (1,-1,1) needs row y = 7 of the tile below
    ("y_tail_tile", (16, 9, 16), (5, 8, 5), (7, 8, 7)),
    # the same across the tail x word (nx = 65): (1,1,1) + (-1,1,1) needs x = 63
    ("x_tail_word", (16, 16, 65), (64, 5, 5), (64, 7, 7)),
])
def test_forbench_paths_leave_the_box(name, shape, a, b):
    """The 18 element's shortest path leaves the voxels' box by one voxel; the
    tile box must be widened (the unwidened box gives 4 here)."""
    g, s = points(shape, a, b)
    assert scipy_expect(g, s, 18)[0]["hd"] == 2
    sub = (slice(None), slice(8, 9), slice(None)) if name == "y_tail_tile" else \
          (slice(None), slice(None), slice(64, 65))
    assert scipy_expect(g[sub], s[sub], 18)[0]["hd"] == 4
    for method, crop, res, _ in run_all(g, s, 18):
        assert res["hd"] == 2, (method, crop)


def test_conn_six_is_the_default_path():
    rng = np.random.default_rng(5)
    g = rng.random((12, 14, 70)) < 0.02
    s = rng.random((12, 14, 70)) < 0.02
    g[0, 0, 0] = s[-1, -1, -1] = True
    rc, a, *_ = oracle.run(g.astype(np.uint8), s.astype(np.uint8))
    rc2, b, *_ = oracle.run(g.astype(np.uint8), s.astype(np.uint8), conn=6)
    assert rc == rc2 == 0 and a == b


def test_unknown_conn_is_rejected():
    g = np.ones((4, 4, 4), np.uint8)
    for method in ("naive", "tiled"):
        rc, *_ = oracle.run(g, g, method=method, conn=10)
        assert rc == -1
