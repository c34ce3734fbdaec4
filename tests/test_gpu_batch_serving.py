"""Batched launches (hipGraph replays), the autotuner and the serving engine:
every result must equal the same pair computed alone, bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

import housdorffhdf5cmake_amd as hdm
from housdorffhdf5cmake_amd import serving
from oracle import oracle

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

KEYS = ("hd", "sum", "count", "sum_fp", "cnt_fp", "sum_fn", "cnt_fn", "n_fp", "n_fn", "n_gold",
        "n_segm", "hd_fp", "hd_fn")


def same(a, b):
    return all(a[k] == b[k] for k in KEYS)


def mixed_pairs(n, seed=3):
    rng = np.random.default_rng(seed)
    dims = [(64, 64, 64), (128, 96, 64), (67, 45, 31), (256, 128, 64), (100, 100, 100)]
    vols = []
    for i in range(n):
        nx, ny, nz = dims[int(rng.integers(0, len(dims)))]
        sp = hdm.synth_spec("ellipsoids", nx, ny, nz, 6, 4, 20, 2, 0.01, 40 + i)
        vols.append(hdm.synth_device(sp))
    torch.cuda.synchronize()
    return vols


def test_batch_equals_single_and_oracle():
    vols = mixed_pairs(12)
    alone = [hdm.hausdorff(g, s) for g, s in vols]
    for tile in (hdm.TILE_64x8x8, hdm.TILE_64x4x4, hdm.TILE_32x8x8):
        b = hdm.Batch(vols, tile=tile, graph=False)
        b.run()
        out = b.sync()
        b.close()
        for o, a in zip(out, alone):
            assert same(o, a)
    for (g, s), a in zip(vols[:4], alone[:4]):
        rc, want, _, _, _ = oracle.run(g.cpu().numpy(), s.cpu().numpy())
        assert rc == 0 and all(int(want[k]) == int(a[k]) for k in KEYS)


@pytest.mark.parametrize("robust", [1.0, 0.9])
def test_batch_near_pass_with_short_and_long_pairs(robust):
    """The batch's near pass finishes the pairs whose cut is reached within
    eight steps and leaves the others to the seed and the dilation: a mix of
    both equals every pair alone on two graph replays, and the entries the
    near pass stepped are reported."""
    short = mixed_pairs(4, seed=51)
    lg = hdm.synth_device(hdm.synth_spec("long", 128, 96, 128, 30, 64, 48, 20, 15, 12, 60, 4))
    torch.cuda.synchronize()
    vols = [short[0], lg, short[1], short[2], (lg[1], lg[0]), short[3]]
    alone = [hdm.hausdorff(g, s, robustness=robust) for g, s in vols]
    assert max(a["hd"] for a in alone) > 8 and min(a["hd"] for a in alone) <= 8
    b = hdm.Batch(vols, robustness=robust, graph=True)
    for _ in range(2):
        b.run()
        out = b.sync()
        for o, a in zip(out, alone):
            assert same(o, a)
        assert b.near_entries() > 0
    b.close()


@pytest.mark.parametrize("robust", [0.5, 0.95])
def test_batch_robust_cut_equals_single_and_oracle(robust):
    """A batch runs two dilation steps per grid barrier (64x8x8 tiles, 6
    element): each step's first hits are their own histogram bin, and a pass
    stops at the bin whose covered count reaches ceil(r * n)
    (hello.cu:1748-1759) -- also when that bin is the first of the two, whose
    second step's hits then count for nothing.  One-pair batches run the
    single-pair engine loop."""
    vols = mixed_pairs(6, seed=29)
    alone = [hdm.hausdorff(g, s, robustness=robust) for g, s in vols]
    b = hdm.Batch(vols, robustness=robust, graph=False)
    b.run()
    out = b.sync()
    b.close()
    for o, a in zip(out, alone):
        assert same(o, a)
    one = hdm.Batch(vols[2:3], robustness=robust, graph=False)
    one.run()
    assert same(one.sync()[0], alone[2])
    one.close()
    for (g, s), a in zip(vols[:3], alone[:3]):
        rc, want, _, _, _ = oracle.run(g.cpu().numpy(), s.cpu().numpy(), robustness=robust)
        assert rc == 0 and all(int(want[k]) == int(a[k]) for k in KEYS)


def test_batch_of_small_and_degenerate_volumes_equals_single():
    """Lines, single voxels, volumes smaller than one tile, odd dims and
    random sparse volumes in one batch (two dilation steps per barrier over
    pairs of every size): each result equals the pair computed alone."""
    rng = np.random.default_rng(31)
    host = []
    for shape in ((1, 1, 200), (200, 1, 1), (1, 150, 1), (3, 1, 129), (5, 7, 9), (1, 1, 1)):
        g = np.zeros(shape, np.uint8)
        s = np.zeros(shape, np.uint8)
        g.reshape(-1)[0] = 1
        s.reshape(-1)[-1] = 1
        host.append((g, s))
    for _ in range(6):
        nz, ny, nx = (int(v) for v in rng.integers(3, 70, size=3))
        g = (rng.random((nz, ny, nx)) < 0.01).astype(np.uint8)
        s = (rng.random((nz, ny, nx)) < 0.01).astype(np.uint8)
        g[0, 0, 0] = s[-1, -1, -1] = 1
        host.append((g, s))
    vols = [(torch.from_numpy(g).cuda(), torch.from_numpy(s).cuda()) for g, s in host]
    alone = [hdm.hausdorff(g, s) for g, s in vols]
    for robust in (1.0, 0.75):
        want = alone if robust == 1.0 else [hdm.hausdorff(g, s, robustness=robust) for g, s in vols]
        b = hdm.Batch(vols, robustness=robust, graph=False)
        b.run()
        out = b.sync()
        b.close()
        for i, (o, a) in enumerate(zip(out, want)):
            assert same(o, a), (robust, i, host[i][0].shape)


def test_graph_replays_are_identical_with_eager_work_between():
    vols = mixed_pairs(8, seed=5)
    alone = [hdm.hausdorff(g, s) for g, s in vols]
    b = hdm.Batch(vols, graph=True)
    for rep in range(4):
        b.run()
        out = b.sync()
        for o, a in zip(out, alone):
            assert same(o, a), rep
        # eager engine work and synchronizations between the replays
        g, s = vols[rep]
        assert same(hdm.hausdorff(g, s), alone[rep])
        torch.cuda.synchronize()
    assert b.last_ms() > 0
    b.close()


def test_tiny_batch_packs_every_step():
    """A batch small enough that its pack runs on fewer workgroups than the
    pack's step counters: every step is still packed (results equal alone)."""
    vols = []
    for i, shape in enumerate([(16, 16, 64), (8, 24, 64), (16, 8, 128)]):
        sp = hdm.synth_spec("ellipsoids", shape[2], shape[1], shape[0], 3, 2, 6, 1, 0.02, 60 + i)
        vols.append(hdm.synth_device(sp))
    torch.cuda.synchronize()
    alone = [hdm.hausdorff(g, s) for g, s in vols]
    b = hdm.Batch(vols, graph=False)
    b.run()
    out = b.sync()
    b.close()
    for o, a in zip(out, alone):
        assert same(o, a)


def test_batch_with_empty_mask_pair_reports_error_for_it():
    vols = mixed_pairs(3, seed=9)
    g0 = torch.zeros((20, 30, 40), dtype=torch.uint8, device="cuda")
    s0 = g0.clone()
    s0[5, 5, 5] = 1
    b = hdm.Batch(vols + [(g0, s0)], graph=False)
    b.run()
    with pytest.raises(hdm.HDMError) as e:
        b.sync()
    assert e.value.code == -2
    b.close()


def test_autotune_table_consistent():
    vols = mixed_pairs(6, seed=7)
    table = hdm.autotune(vols, tiles=(hdm.TILE_64x8x8, hdm.TILE_64x8x4), chunks=(1, 2),
                         grids=(0, 256), reps=2)
    assert len(table) == 8
    assert all(e["status"] == 0 for e in table)
    ms = [e["ms"] for e in table]
    assert ms == sorted(ms) and ms[0] > 0


def test_server_mix_equals_alone():
    vols = mixed_pairs(10, seed=13)
    alone = [hdm.hausdorff(g, s) for g, s in vols]
    srv = serving.Server(hbm_budget=4 << 30, max_slots=4, slice_iters=2, max_dims=(256, 128, 100))
    ids = {srv.submit(g, s): i for i, (g, s) in enumerate(vols)}
    got = {}
    for c in srv.drain(60):
        got[ids[c["id"]]] = c
    assert len(got) == len(vols)
    for i, c in got.items():
        assert c["status"] == 0 and same(c, alone[i]), i
    # waves: new requests arrive while earlier ones are in flight
    ids = {}
    done = []
    for i, (g, s) in enumerate(vols):
        ids[srv.submit(g, s)] = i
        done += srv.step()
    done += srv.drain(60)
    assert len(done) == len(vols)
    for c in done:
        assert c["status"] == 0 and same(c, alone[ids[c["id"]]])
    srv.close()


def test_servers_in_a_row_each_one_slice():
    """Each server here runs exactly one slice (two requests, admitted
    together), so the one created after it may get the same pinned block back
    with the first slice's ready marker still in it; every server must still
    wait for its own slice."""
    vols = mixed_pairs(2, seed=37)
    alone = [hdm.hausdorff(g, s) for g, s in vols]
    for k in range(6):
        srv = serving.Server(hbm_budget=1 << 30, max_slots=4, slice_iters=64,
                             max_dims=(256, 128, 100))
        order = [k % 2, 1 - k % 2]
        rids = [srv.submit(*vols[i]) for i in order]
        done = srv.drain(60)
        srv.close()
        assert sorted(c["id"] for c in done) == sorted(rids)
        for c in done:
            i = order[rids.index(c["id"])]
            assert c["status"] == 0 and same(c, alone[i]), k


@pytest.mark.parametrize("robust,conn", [(1.0, 6), (0.9, 6), (1.0, 18)])
def test_lone_requests_equal_alone(robust, conn):
    """A request that finds nothing in flight and nothing else waiting runs as
    the one-pair call on the server's stream and completes in the same step:
    its result equals the pair computed alone (device and host pointers,
    int32 volumes, a robust cut, the 18 element), and an empty mask completes
    it with status -2."""
    vols = mixed_pairs(3, seed=41)
    alone = [hdm.hausdorff(g, s, robustness=robust, connectivity=conn) for g, s in vols]
    srv = serving.Server(hbm_budget=1 << 30, max_slots=4, robustness=robust, connectivity=conn,
                         max_dims=(256, 128, 100))
    for i, (g, s) in enumerate(vols):
        for kind in ("device", "host", "i32"):
            if kind == "host":
                gg, ss = g.cpu().numpy(), s.cpu().numpy()
            elif kind == "i32":
                gg, ss = g.int(), s.int()
                torch.cuda.synchronize()  # the server's stream does not wait for torch's
            else:
                gg, ss = g, s
            rid = srv.submit(gg, ss)
            done = srv.step()
            assert [c["id"] for c in done] == [rid], (i, kind)
            assert done[0]["status"] == 0 and same(done[0], alone[i]), (i, kind)
    g0 = torch.zeros((20, 30, 40), dtype=torch.uint8, device="cuda")
    s0 = g0.clone()
    s0[5, 5, 5] = 1
    torch.cuda.synchronize()
    rid = srv.submit(g0, s0)
    done = srv.step()
    assert [(c["id"], c["status"]) for c in done] == [(rid, -2)]
    srv.close()


def test_server_errors_budget_and_capacity():
    vols = mixed_pairs(3, seed=17)
    g0 = torch.zeros((16, 16, 64), dtype=torch.uint8, device="cuda")
    s0 = g0.clone()
    s0[3, 3, 3] = 1
    torch.cuda.synchronize()  # the server's stream does not wait for torch's
    srv = serving.Server(hbm_budget=64 << 20, max_slots=2, slice_iters=4, max_dims=(256, 128, 100))
    ids = {srv.submit(g, s): i for i, (g, s) in enumerate(vols)}
    bad = srv.submit(g0, s0)
    res = {c["id"]: c for c in srv.drain(60)}
    assert res[bad]["status"] == -2
    alone = [hdm.hausdorff(g, s) for g, s in vols]
    for rid, i in ids.items():
        assert res[rid]["status"] == 0 and same(res[rid], alone[i])
    big = torch.zeros((100, 128, 300), dtype=torch.uint8, device="cuda")
    with pytest.raises(hdm.HDMError) as e:
        srv.submit(big, big)
    assert e.value.code == -9  # dims beyond the server's max
    srv.close()
    tiny = serving.Server(hbm_budget=2 << 20, max_slots=2, max_dims=(256, 256, 256))
    g, s = mixed_pairs(1, seed=19)[0]
    big = torch.zeros((200, 200, 200), dtype=torch.uint8, device="cuda")
    with pytest.raises(hdm.HDMError) as e:
        tiny.submit(big, big)
    assert e.value.code == -6  # larger than the whole budget
    tiny.close()


def test_server_host_pointer_requests():
    vols = mixed_pairs(4, seed=23)
    alone = [hdm.hausdorff(g, s) for g, s in vols]
    host = [(g.cpu().numpy(), s.cpu().numpy()) for g, s in vols]
    srv = serving.Server(hbm_budget=1 << 30, max_slots=3, slice_iters=3, max_dims=(256, 128, 100))
    ids = {srv.submit(g, s): i for i, (g, s) in enumerate(host)}
    res = srv.drain(60)
    assert len(res) == 4
    for c in res:
        assert c["status"] == 0 and same(c, alone[ids[c["id"]]])
    srv.close()


def test_poisson_serving_short_run():
    def verify(g, s, a):
        rc, want, _, _, _ = oracle.run(g.cpu().numpy(), s.cpu().numpy())
        return rc == 0 and all(int(want[k]) == int(a[k]) for k in KEYS)

    r = serving.poisson_run(seconds=1.5, rate=100, budget_bytes=8 << 30, verify=verify,
                            per_class=2, max_slots=16)
    assert r["completed"] == r["requests"] > 0
    assert r["identical_alone_vs_mix"] and r["bit_exact_vs_cpu"]
    assert r["p99_ms"] >= r["p50_ms"] > 0


@pytest.mark.parametrize("robust,conn", [(1.0, 6), (0.9, 6), (1.0, 18)])
def test_registered_masks_equal_fresh_submit(robust, conn):
    """Volumes registered once (packed on the device, no volume read at
    admission) give the same result as the same pair submitted as volumes and
    as the pair computed alone: each registered gold mask is used again with
    several segm masks, in one wave and request by request; int32 and host
    volumes register too.  A mask unregistered while requests name it stays
    until they complete; a later request naming it is refused."""
    base = mixed_pairs(4, seed=53)
    vols = []
    for g, s in base:  # a second segm per gold, same dims
        s2 = s.clone()
        s2[: s2.shape[0] // 3] = 0
        vols += [(g, s), (g, s2)]
    torch.cuda.synchronize()
    alone = [hdm.hausdorff(g, s, robustness=robust, connectivity=conn) for g, s in vols]
    srv = serving.Server(hbm_budget=2 << 30, max_slots=4, slice_iters=2, robustness=robust,
                         connectivity=conn, max_dims=(256, 128, 100))
    hg = {}
    hs = []
    for k, (g, s) in enumerate(vols):
        if k % 2 == 0:
            src = g.cpu().numpy() if k % 4 == 0 else g.int()  # host uint8, device int32
            torch.cuda.synchronize()
            hg[k // 2] = srv.register(src)
        hs.append(srv.register(s))
    fresh = {}
    for rep in range(2):  # every registered mask named again in the second round
        ids = {srv.submit_masks(hg[k // 2], hs[k]): k for k in range(len(vols))}
        for c in srv.drain(60):
            k = ids[c["id"]]
            assert c["status"] == 0 and same(c, alone[k]), (rep, k)
            fresh[k] = c
    ids = {srv.submit(g, s): k for k, (g, s) in enumerate(vols)}
    for c in srv.drain(60):
        k = ids[c["id"]]
        assert same(c, fresh[k]) and c["iterations"] == fresh[k]["iterations"], k
    for k in range(len(vols)):  # one by one: the first request of a fresh engine
        rid = srv.submit_masks(hg[k // 2], hs[k])
        done = srv.drain(60)
        assert [c["id"] for c in done] == [rid] and same(done[0], alone[k]), k
    st = srv.stats()
    assert st["registered_bytes"] > 0
    assert st["admitted"] + st["near_finished"] + st["lone"] >= 3 * len(vols)
    rid = srv.submit_masks(hg[0], hs[0])
    srv.unregister(hs[0])
    done = srv.drain(60)
    assert [c["id"] for c in done] == [rid] and same(done[0], alone[0])
    with pytest.raises(hdm.HDMError) as e:
        srv.submit_masks(hg[0], hs[0])
    assert e.value.code == -1
    small = torch.ones((8, 8, 8), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    hsmall = srv.register(small)
    with pytest.raises(hdm.HDMError) as e:
        srv.submit_masks(hg[0], hsmall)  # unequal dims
    assert e.value.code == -1
    srv.close()


@pytest.mark.parametrize("resident", [False, True])
def test_loaded_serving_batches_and_equals_alone(resident):
    """Under load -- Poisson arrivals at 6000 requests/s, and a closed loop
    that keeps 24 requests in the server -- every result equals the pair
    computed alone, with slices of eight steps and of two (pairs then stay in
    flight across slices and new ones join between them); in the closed loop
    several requests are admitted together, and the admission group's near
    pass finishes them (this mix has HD <= 8 in every pair).  The slices'
    own batching under load runs with the near pass off, below."""
    for si in (8, 2):
        r = serving.poisson_run(seconds=0.5, rate=6000, budget_bytes=8 << 30, per_class=2,
                                max_slots=32, slice_iters=si, resident=resident)
        assert r["completed"] == r["requests"] > 1000, si
        assert r["identical_alone_vs_mix"], si
        c = serving.closed_loop_run(seconds=0.5, outstanding=24, budget_bytes=8 << 30,
                                    per_class=2, max_slots=32, slice_iters=si,
                                    resident=resident)
        assert c["completed"] > 500 and c["identical_alone_vs_mix"], si
        assert c["group_mean"] > 4 and c["near_finished"] > 500, (si, c)


CHILD_LOADED = r"""
import sys, json
sys.path.insert(0, sys.argv[1])
from housdorffhdf5cmake_amd import serving
out = []
for si in (8, 2):
    c = serving.closed_loop_run(seconds=0.5, outstanding=24, budget_bytes=8 << 30, per_class=2,
                                max_slots=32, slice_iters=si, resident=True)
    out.append(c)
print(json.dumps(out))
"""


def test_loaded_serving_through_the_slices():
    """With the near pass off (HDM_NEAR=0) every admitted request goes
    through the slots and the slices: under a closed loop of 24 requests the
    engine runs several pairs per slice, and every result equals the pair
    computed alone."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, HDM_NEAR="0")
    r = subprocess.run([sys.executable, "-c", CHILD_LOADED, root], capture_output=True, text=True,
                       timeout=240, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    for c in json.loads(r.stdout.strip().splitlines()[-1]):
        assert c["completed"] > 500 and c["identical_alone_vs_mix"], c
        assert c["batch_mean"] > 4 and c["max_in_flight"] > 8 and c["near_finished"] == 0, c
