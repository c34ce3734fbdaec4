"""Parity of the HIP engine with the golden fixtures and the CPU oracle,
bit-exact on every integer output (HD, sums, counts, histograms, maps)."""
from __future__ import annotations

import numpy as np
import pytest

import housdorffhdf5cmake_amd as hdm
from oracle import oracle
from tests.golden_cases import case, check_hist, check_result, manifest, volumes

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TILES = [hdm.TILE_64x8x8, hdm.TILE_64x4x4, hdm.TILE_64x8x4, hdm.TILE_32x8x8]
CASES = [c["name"] for c in manifest()]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def oracle_expect(g, s, label=1, robustness=1.0):
    rc, res, hf, hn, _ = oracle.run(g, s, label=label, robustness=robustness)
    assert rc == 0
    return res, hf, hn


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("tile", TILES)
def test_golden_cases_all_tiles(name, tile):
    c = case(name)
    label = c.get("label", 1)
    g, s = volumes(c, label)
    G, S = dev(g), dev(s)
    for exp in c["expected"]:
        r = exp["robustness"]
        res = hdm.hausdorff(G, S, label=label, robustness=r, tile=tile,
                            hist_len=sum(G.shape) + 2)
        check_result(res, exp, f"{name}/tile{tile}/r={r}")
        if r == 1.0:
            check_hist(res["hist_fp"], res["hist_fn"], exp, f"{name}/tile{tile}")
            assert res["iterations"] == exp["hd"]


@pytest.mark.parametrize("name", ["sphere64", "ell_130x70x45", "random3_67x13x9", "long_200x48x40"])
def test_host_pointers_and_crop_off(name):
    c = case(name)
    g, s = volumes(c)
    exp = c["expected"][0]
    check_result(hdm.hausdorff(g, s), exp, f"{name}/host")
    check_result(hdm.hausdorff(dev(g), dev(s), crop=False), exp, f"{name}/nocrop")


@pytest.mark.parametrize("name", ["sphere64", "random3_67x13x9", "long_200x48x40", "kat_testdata"])
def test_first_hit_map_matches_oracle(name):
    c = case(name)
    g, s = volumes(c)
    _, _, _, _, want = oracle.run(g, s, want_map=True)
    res, mp = hdm.first_hit_map(dev(g), dev(s))
    got = mp.cpu().numpy().view(np.uint16)
    assert np.array_equal(got, want)
    res_h, mp_h = hdm.first_hit_map(g, s)
    assert np.array_equal(mp_h, want)


def test_kat_points_first_and_second_dilation():
    """old/testData.cu:61-132 pairs at L1 1 and 2 are covered at those iterations."""
    c = case("kat_testdata")
    g, s = volumes(c)
    _, mp = hdm.first_hit_map(dev(g), dev(s))
    m = mp.cpu().numpy().view(np.uint16)
    for x, y, z, where, group in c["points"]:
        if group == "first":
            assert m[z, y, x] == 1
        if group == "second":
            assert m[z, y, x] == 2


def test_label_dtype_and_bool():
    c = case("ell_96x80x64")
    g, s = volumes(c)
    exp = c["expected"][0]
    rng = np.random.default_rng(3)
    g7 = np.where(g == 1, 7, rng.choice([0, 1, 2], size=g.shape)).astype(np.uint8)
    s7 = np.where(s == 1, 7, rng.choice([0, 1, 2], size=s.shape)).astype(np.uint8)
    check_result(hdm.hausdorff(dev(g7), dev(s7), label=7), exp, "label7")
    check_result(hdm.hausdorff(dev(g7.astype(np.int32)), dev(s7.astype(np.int32)), label=7), exp,
                 "i32")
    check_result(hdm.hausdorff(dev(g.astype(bool)), dev(s.astype(bool))), exp, "bool")
    assert hdm.getHausdorffDistance_CUDA_Generic(dev(g7), dev(s7), 96, 80, 64, 1.0, False, 7,
                                                 False) == exp["hd"]


def test_edge_cases():
    g = np.zeros((9, 10, 70), np.uint8)
    s = g.copy()
    s[4, 5, 66] = 1
    with pytest.raises(hdm.HDMError) as e:
        hdm.hausdorff(dev(g), dev(s))
    assert e.value.code == -2
    with pytest.raises(hdm.HDMError) as e:
        hdm.hausdorff(dev(s), dev(g))
    assert e.value.code == -2
    r = hdm.hausdorff(dev(g), dev(g))
    assert r["hd"] == 0 and r["count"] == 0 and r["avg_hd"] == 0.0
    r = hdm.hausdorff(dev(s), dev(s))
    assert r["hd"] == 0 and r["iterations"] == 0
    # 1x1xN and Nx1x1 lines
    for shape in ((1, 1, 200), (200, 1, 1), (1, 150, 1), (3, 1, 129)):
        g = np.zeros(shape, np.uint8)
        s = np.zeros(shape, np.uint8)
        g.reshape(-1)[0] = 1
        s.reshape(-1)[-1] = 1
        want, _, _ = oracle_expect(g, s)
        check_result(hdm.hausdorff(dev(g), dev(s)), want, f"line{shape}")


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_random_ragged_volumes(seed):
    rng = np.random.default_rng(seed)
    nz, ny, nx = (int(v) for v in rng.integers(5, 90, size=3))
    g = (rng.random((nz, ny, nx)) < 0.01).astype(np.uint8)
    s = (rng.random((nz, ny, nx)) < 0.01).astype(np.uint8)
    g[0, 0, 0] = s[-1, -1, -1] = 1
    want, hf, hn = oracle_expect(g, s)
    for tile in TILES:
        res = hdm.hausdorff(dev(g), dev(s), tile=tile, hist_len=nx + ny + nz + 2)
        check_result(res, want, f"seed{seed}/tile{tile}")
        assert np.array_equal(res["hist_fp"], hf) and np.array_equal(res["hist_fn"], hn)


def test_device_generator_matches_host():
    for kind, args in (("sphere", (64, 64, 64, 32, 32, 32, 20, 2, 1, 0, 0.01, 1)),
                       ("ellipsoids", (160, 128, 96, 8, 6, 24, 2, 0.005, 3)),
                       ("long", (200, 48, 40, 30, 24, 20, 20, 12, 10, 120, 3))):
        sp = hdm.synth_spec(kind, *args)
        g, s = hdm.synth_host(sp)
        G, S = hdm.synth_device(sp)
        torch.cuda.synchronize()
        assert np.array_equal(G.cpu().numpy(), g) and np.array_equal(S.cpu().numpy(), s), kind


def test_workload_ct_pair_full_size_vs_oracle():
    """512x512x256 union of 8 ellipsoids (semi-axes 10-60), segm +-0..2 voxels
    per ellipsoid and 0.5% boundary flips: bit-exact against the oracle."""
    sp = hdm.synth_spec("ellipsoids", 512, 512, 256, 8, 10, 60, 2, 0.005, 2024)
    G, S = hdm.synth_device(sp)
    res = hdm.hausdorff(G, S, hist_len=512 + 512 + 256 + 2)
    g, s = G.cpu().numpy(), S.cpu().numpy()
    want, hf, hn = oracle_expect(g, s)
    check_result(res, want, "ct512")
    assert np.array_equal(res["hist_fp"], hf) and np.array_equal(res["hist_fn"], hn)
    for tile in TILES[1:]:
        check_result(hdm.hausdorff(G, S, tile=tile), want, f"ct512/tile{tile}")


def test_workload_long_iteration_full_size_vs_oracle():
    """512^3, ellipsoid (80,60,50) plus a 5^3 cube 300 voxels away."""
    sp = hdm.synth_spec("long", 512, 512, 512, 100, 256, 256, 80, 60, 50, 300, 5)
    G, S = hdm.synth_device(sp)
    res = hdm.hausdorff(G, S)
    want, _, _ = oracle_expect(G.cpu().numpy(), S.cpu().numpy())
    check_result(res, want, "long512")
    assert res["hd"] >= 300 and res["iterations"] == res["hd"]
    assert res["queue_max"] > 0 and res["queue_total"] >= res["queue_max"]


def test_repeat_calls_are_identical():
    sp = hdm.synth_spec("ellipsoids", 256, 256, 128, 8, 10, 60, 2, 0.005, 77)
    G, S = hdm.synth_device(sp)
    a = hdm.hausdorff(G, S)
    for _ in range(5):
        b = hdm.hausdorff(G, S)
        for k in ("hd", "sum", "count", "sum_fp", "sum_fn", "n_fp", "n_fn"):
            assert a[k] == b[k]


def test_calls_across_streams_and_workspace_growth():
    """A call returns as soon as the kernel has stored its result; the next
    call on another stream, or one that grows the workspace, waits for the
    previous launch first.  Inputs may be overwritten once a call returns."""
    small = hdm.synth_spec("ellipsoids", 96, 80, 64, 6, 4, 20, 2, 0.005, 5)
    big = hdm.synth_spec("ellipsoids", 256, 256, 160, 8, 10, 40, 2, 0.005, 6)
    gs, ss = hdm.synth_device(small)
    gb, sb = hdm.synth_device(big)
    torch.cuda.synchronize()
    want_s = oracle_expect(gs.cpu().numpy(), ss.cpu().numpy())[0]
    want_b = oracle_expect(gb.cpu().numpy(), sb.cpu().numpy())[0]
    streams = [torch.cuda.Stream(), torch.cuda.Stream(), torch.cuda.default_stream()]
    for i in range(9):
        st = streams[i % 3]
        st.wait_stream(torch.cuda.default_stream())
        with torch.cuda.stream(st):
            if i % 4 == 3:
                check_result(hdm.hausdorff(gb, sb), want_b, f"big/{i}")
            else:
                check_result(hdm.hausdorff(gs, ss), want_s, f"small/{i}")
    # overwrite the inputs right after a call returns, then use them again
    g2, s2 = gs.clone(), ss.clone()
    check_result(hdm.hausdorff(g2, s2), want_s, "before overwrite")
    g2.zero_()
    s2.zero_()
    r = hdm.hausdorff(g2, s2)
    assert r["hd"] == 0 and r["count"] == 0
    check_result(hdm.hausdorff(gs, ss), want_s, "after overwrite")
