"""Parity of the HIP engine's 18- and 26-neighbour structuring elements
(hdm_opts.connectivity) with the CPU oracle, bit-exact on HD, sums, counts,
histograms and first-hit maps, through every launch path: the fused
single-pair launch, the separate-kernel single-pair path (a batch of one),
the batched multi-pair launch and the serving engine."""
from __future__ import annotations

import numpy as np
import pytest

import housdorffhdf5cmake_amd as hdm
from housdorffhdf5cmake_amd import serving
from oracle import oracle
from tests.golden_cases import case, manifest, volumes

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CONNS = [18, 26]
KEYS = ("hd", "hd_fp", "hd_fn", "sum", "count", "sum_fp", "cnt_fp", "sum_fn", "cnt_fn",
        "n_gold", "n_segm", "n_fp", "n_fn")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def expect(g, s, conn, label=1, robustness=1.0):
    rc, res, hf, hn, _ = oracle.run(g, s, label=label, robustness=robustness, conn=conn)
    assert rc == 0
    return res, hf, hn


def same(res, want, where):
    got = {k: int(res[k]) for k in KEYS}
    exp = {k: int(want[k]) for k in KEYS}
    assert got == exp, where


@pytest.mark.parametrize("conn", CONNS)
@pytest.mark.parametrize("name", [c["name"] for c in manifest()])
def test_golden_volumes(conn, name):
    c = case(name)
    label = c.get("label", 1)
    g, s = volumes(c, label)
    G, S = dev(g), dev(s)
    for r in sorted({e["robustness"] for e in c["expected"]}):
        want, hf, hn = expect(g, s, conn, label, r)
        L = sum(G.shape) + 2
        res = hdm.hausdorff(G, S, label=label, robustness=r, connectivity=conn, hist_len=L)
        same(res, want, f"{name}/conn{conn}/r={r}")
        assert np.array_equal(res["hist_fp"], hf) and np.array_equal(res["hist_fn"], hn)
        if r == 1.0:
            assert res["iterations"] == want["hd"]


@pytest.mark.parametrize("conn", CONNS)
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_random_ragged_volumes(conn, seed):
    rng = np.random.default_rng(100 * conn + seed)
    nz, ny, nx = (int(v) for v in rng.integers(5, 90, size=3))
    if seed == 4:
        nx = 129  # a 1-voxel tail word
    g = (rng.random((nz, ny, nx)) < 0.005).astype(np.uint8)
    s = (rng.random((nz, ny, nx)) < 0.005).astype(np.uint8)
    g[0, 0, 0] = s[-1, -1, -1] = 1
    want, hf, hn = expect(g, s, conn)
    res = hdm.hausdorff(dev(g), dev(s), connectivity=conn, hist_len=nx + ny + nz + 2)
    same(res, want, f"seed{seed}/conn{conn}")
    assert np.array_equal(res["hist_fp"], hf) and np.array_equal(res["hist_fn"], hn)
    same(hdm.hausdorff(dev(g), dev(s), connectivity=conn, crop=False), want, "nocrop")
    same(hdm.hausdorff(g, s, connectivity=conn), want, "host pointers")


@pytest.mark.parametrize("shape,a,b", [((16, 9, 16), (5, 8, 5), (7, 8, 7)),
                                       ((16, 16, 65), (64, 5, 5), (64, 7, 7))])
def test_forbench_path_leaves_the_box(shape, a, b):
    """Both voxels on the one row of a tail tile: the 18 element's two-step
    path needs the tile beside the box (HD 2, the unwidened box gives 4)."""
    g = np.zeros(shape, np.uint8)
    s = np.zeros(shape, np.uint8)
    g[a[2], a[1], a[0]] = 1
    s[b[2], b[1], b[0]] = 1
    res = hdm.hausdorff(dev(g), dev(s), connectivity=18)
    assert res["hd"] == 2 and res["sum"] == 4 and res["count"] == 2
    assert hdm.hausdorff(dev(g), dev(s), connectivity=26)["hd"] == 2
    assert hdm.hausdorff(dev(g), dev(s), connectivity=6)["hd"] == 4


@pytest.mark.parametrize("conn", CONNS)
def test_first_hit_map(conn):
    for name in ("sphere64", "random3_67x13x9", "long_200x48x40"):
        g, s = volumes(case(name))
        _, _, _, _, want = oracle.run(g, s, want_map=True, conn=conn)
        _, mp = hdm.first_hit_map(dev(g), dev(s), connectivity=conn)
        assert np.array_equal(mp.cpu().numpy().view(np.uint16), want), name


@pytest.mark.parametrize("conn", CONNS)
def test_batch_and_single_kernel_paths(conn):
    """Several pairs through the multi-pair kernel; one pair through the
    separate-kernel single-pair path (a batch of one)."""
    vols, wants = [], []
    for i, dims in enumerate([(64, 64, 64), (96, 80, 48), (130, 70, 45), (40, 33, 29)]):
        sp = hdm.synth_spec("ellipsoids", *dims, 6, 4, 20, 2, 0.005, 300 + i)
        G, S = hdm.synth_device(sp)
        vols.append((G, S))
    torch.cuda.synchronize()
    for G, S in vols:
        wants.append(expect(G.cpu().numpy(), S.cpu().numpy(), conn)[0])
    b = hdm.Batch(vols, connectivity=conn, graph=True)
    for _ in range(2):  # a replayed graph gives the same results
        b.run()
        out = b.sync()
        for i, (o, w) in enumerate(zip(out, wants)):
            same(o, w, f"batch pair {i}")
    b.close()
    one = hdm.Batch(vols[:1], connectivity=conn, graph=False)
    one.run()
    same(one.sync()[0], wants[0], "batch of one")
    one.close()


@pytest.mark.parametrize("conn", CONNS)
def test_serving_engine(conn):
    srv = serving.Server(hbm_budget=4 << 30, max_slots=8, slice_iters=4,
                         max_dims=(160, 128, 96), connectivity=conn)
    pairs, ids = {}, []
    for i, dims in enumerate([(64, 64, 64), (160, 128, 96), (96, 80, 48)]):
        sp = hdm.synth_spec("ellipsoids", *dims, 6, 4, 20, 2, 0.005, 500 + i)
        G, S = hdm.synth_device(sp)
        torch.cuda.synchronize()
        rid = srv.submit(G, S)
        pairs[rid] = (G, S)
    done = srv.drain(60.0)
    srv.close()
    assert sorted(d["id"] for d in done) == sorted(pairs)
    for d in done:
        G, S = pairs[d["id"]]
        assert d["status"] == 0
        same(d, expect(G.cpu().numpy(), S.cpu().numpy(), conn)[0], f"request {d['id']}")


@pytest.mark.parametrize("conn", CONNS)
def test_ct_pair_full_size(conn):
    """512x512x256 union of 8 ellipsoids: bit-exact against the oracle."""
    sp = hdm.synth_spec("ellipsoids", 512, 512, 256, 8, 10, 60, 2, 0.005, 2024)
    G, S = hdm.synth_device(sp)
    res = hdm.hausdorff(G, S, connectivity=conn, hist_len=512 + 512 + 256 + 2)
    want, hf, hn = expect(G.cpu().numpy(), S.cpu().numpy(), conn)
    same(res, want, f"ct512/conn{conn}")
    assert np.array_equal(res["hist_fp"], hf) and np.array_equal(res["hist_fn"], hn)
