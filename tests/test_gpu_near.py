"""The near pass of the single-pair launch (64x8x8 tiles, 6 element): after
the pack, every region of 1 x 2 x 2 tiles that holds voxels to be covered
dilates the other mask eight steps in its own window, and a pair whose cut is
reached within those steps finishes in that one launch; a pair with HD > 8
(or a robust cut reached later) is handed to the region launch.  Every case
is compared field by field, histograms included, with the CPU oracle, on both
sides of the eight-step boundary, at volume borders and partial tiles, for
robust cuts and for empty masks."""
from __future__ import annotations

import numpy as np
import pytest

import housdorffhdf5cmake_amd as hdm
from oracle import oracle

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

KEYS = ("hd", "hd_fp", "hd_fn", "iterations", "sum", "count", "sum_fp", "cnt_fp", "sum_fn",
        "cnt_fn", "n_gold", "n_segm", "n_fp", "n_fn")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def check(g, s, robustness=1.0, what=""):
    """GPU against the oracle on every integer output; returns the GPU result
    and the launches the fused path used."""
    hl = sum(g.shape) + 2
    res = hdm.hausdorff(dev(g), dev(s), robustness=robustness, hist_len=hl, timing=True)
    rc, want, hf, hn, _ = oracle.run(g, s, robustness=robustness)
    assert rc == 0, what
    bad = {k: (res[k], want[k]) for k in KEYS if res[k] != want[k]}
    assert not bad, f"{what}: {bad}"
    if robustness == 1.0:
        assert np.array_equal(res["hist_fp"], hf), what  # both nx + ny + nz + 2 long
        assert np.array_equal(res["hist_fn"], hn), what
    return res, hdm.last_timings()


def blob(shape, centre, r):
    z, y, x = np.ogrid[:shape[0], :shape[1], :shape[2]]
    cz, cy, cx = centre
    return ((z - cz) ** 2 + (y - cy) ** 2 + (x - cx) ** 2 <= r * r).astype(np.uint8)


@pytest.mark.parametrize("dist", [1, 7, 8, 9, 12, 17])
@pytest.mark.parametrize("axis", ["x", "yz"])
def test_single_far_voxel_around_the_eight_step_boundary(dist, axis):
    """One FP voxel at L1 distance dist from a ball: up to 8 the near pass
    finishes alone (one launch), beyond it the pair goes on to the region
    launch (two launches); the result equals the oracle's either way."""
    shape = (40, 48, 150)
    g = blob(shape, (20, 24, 60), 5)
    s = g.copy()
    if axis == "x":
        s[20, 24, 65 + dist] = 1  # the ball's +x pole is x = 65
    else:  # off the ball's +y pole (y = 29) in y and z
        dy = dist // 2
        s[20 + (dist - dy), 29 + dy, 60] = 1
    res, lt = check(g, s, what=f"{axis} {dist}")
    if axis == "x":
        assert res["hd"] == dist
    assert lt["launches"] == (1 if res["hd"] <= 8 else 2), (res["hd"], lt)


def test_robust_cut_within_eight_steps_of_a_far_pair():
    """Many FP voxels one step away and one 40 steps away: with robustness
    0.9 the cut is reached at step 1 and the near pass finishes the pair;
    with 1.0 the far voxel sends it to the region launch."""
    shape = (32, 40, 200)
    g = np.zeros(shape, np.uint8)
    g[8:24, 10:30, 40:100] = 1
    s = np.zeros_like(g)
    s[8:24, 10:30, 41:101] = 1
    s[16, 20, 99 + 40] = 1  # gold ends at x = 99
    r9, lt9 = check(g, s, robustness=0.9, what="r=0.9")
    assert r9["hd"] == 1 and lt9["launches"] == 1
    r1, lt1 = check(g, s, robustness=1.0, what="r=1.0")
    assert r1["hd"] == 40 and lt1["launches"] == 2


@pytest.mark.parametrize("shape", [(13, 21, 70), (9, 9, 65), (17, 40, 130), (8, 8, 64)])
@pytest.mark.parametrize("seed", [0, 1])
def test_partial_tiles_and_borders(shape, seed):
    """Random sparse masks on volumes that end inside a tile in every axis
    (and one that is exactly one tile): windows that reach past the volume
    read nothing there."""
    rng = np.random.default_rng(seed * 100 + shape[2])
    g = (rng.random(shape) < 0.04).astype(np.uint8)
    s = (rng.random(shape) < 0.04).astype(np.uint8)
    check(g, s, what=f"{shape} seed {seed}")


def test_masks_touching_every_face_of_the_volume():
    shape = (24, 40, 192)
    g = np.zeros(shape, np.uint8)
    s = np.zeros(shape, np.uint8)
    g[0, :, :] = 1
    g[-1, :, 5] = 1
    s[:, 0, :] = 1
    s[:, -1, -1] = 1
    s[3, 7, 0] = 1
    res, _ = check(g, s, what="faces")
    assert res["hd"] > 0


def test_empty_and_identical_masks():
    shape = (16, 16, 64)
    g = blob(shape, (8, 8, 30), 4)
    z = np.zeros_like(g)
    assert hdm.hausdorff(dev(g), dev(g))["hd"] == 0
    assert hdm.hausdorff(dev(z), dev(z))["hd"] == 0
    r = hdm.hausdorff(dev(z), dev(g), raise_on_error=False)
    assert r["status"] == -2  # HDM_E_EMPTY_MASK
    r = hdm.hausdorff(dev(g), dev(z), raise_on_error=False)
    assert r["status"] == -2  # HDM_E_EMPTY_MASK


def test_calls_of_changing_size_in_a_row():
    """Region marks are epochs of the launch that wrote them: a call after a
    larger or a smaller pair (other region grid, other marks) sees only its
    own."""
    pairs = []
    for shape, seed in [((40, 64, 256), 3), ((12, 20, 70), 4), ((64, 64, 128), 5),
                        ((12, 20, 70), 6)]:
        rng = np.random.default_rng(seed)
        g = blob(shape, tuple(v // 2 for v in shape), min(shape) // 3)
        s = np.roll(g, 2, axis=2)
        s[rng.integers(0, shape[0]), rng.integers(0, shape[1]), rng.integers(0, shape[2])] = 1
        pairs.append((g, s))
    for i, (g, s) in enumerate(pairs * 2):
        check(g, s, what=f"call {i}")


@pytest.mark.parametrize("grid", [1, 2, 3, 5, 8, 13])
def test_small_grids_take_every_pack_step(grid):
    """The fused launches' pack hands its last quarter of steps out from one
    counter per group of workgroups; with fewer workgroups than groups every
    step must still be packed (a group without workgroups would leave its
    steps out): a short pair and a pair beyond eight steps, each at a grid of
    a few workgroups, equal the oracle."""
    short = blob((40, 48, 150), (20, 24, 60), 6)
    s1 = short.copy()
    s1[20, 24, 60 + 6 + 3] = 1  # HD 3
    lg, ls = hdm.synth_device(hdm.synth_spec("long", 128, 96, 64, 16, 48, 32, 12, 10, 8, 40, 3))
    torch.cuda.synchronize()
    for g, s in ((short, s1), (lg.cpu().numpy(), ls.cpu().numpy())):
        hl = sum(g.shape) + 2
        res = hdm.hausdorff(dev(g), dev(s), hist_len=hl, grid_ctas=grid)
        rc, want, hf, hn, _ = oracle.run(g, s)
        assert rc == 0
        bad = {k: (res[k], want[k]) for k in KEYS if res[k] != want[k]}
        assert not bad, (grid, g.shape, bad)
        assert np.array_equal(res["hist_fp"], hf) and np.array_equal(res["hist_fn"], hn)


@pytest.mark.parametrize("far", [False, True])
def test_empty_tiles_after_a_dense_pair_read_as_empty(far):
    """The near launch's pack writes no rows of a mask that is empty in a
    tile, so after a dense pair of the same size the packed buffers still
    hold that pair's rows there; the near pass and the region launch must
    take those tiles as empty from their flags.  A dense pair (most tiles
    full) runs first, then a sparse pair of the same dims (a small ball and
    a shifted copy, plus with far=True a voxel 30 steps away, which hands the
    pair to the region launch); the sparse pair equals the oracle."""
    shape = (48, 64, 256)
    rng = np.random.default_rng(7)
    dense_g = (rng.random(shape) < 0.6).astype(np.uint8)
    dense_s = (rng.random(shape) < 0.6).astype(np.uint8)
    check(dense_g, dense_s, what="dense pair")
    g = blob(shape, (24, 32, 128), 9)
    s = np.roll(g, 1, axis=2)
    if far:
        s[24, 32, 137 + 30] = 1  # the ball's +x pole is x = 137
    res, lt = check(g, s, what=f"sparse pair after dense, far={far}")
    assert lt["launches"] == (2 if far else 1), (res["hd"], lt)
    if far:
        assert res["hd"] == 30
