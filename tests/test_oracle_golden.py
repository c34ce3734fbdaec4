"""Pins the CPU oracle: golden fixtures (scipy taxicab distance transform and
brute force), closed forms, invariants, the reference's known-answer point
sets, and planted mistakes that the checks must catch."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from oracle import oracle
from tests.golden_cases import case, check_hist, check_result, manifest, volumes

CASES = manifest()
NAIVE_MAX = 2_500_000  # voxels; the byte-per-voxel oracle is O(N * HD)


def _cases(kinds=None):
    return [c["name"] for c in CASES if kinds is None or c["kind"] in kinds]


@pytest.mark.parametrize("name", _cases())
@pytest.mark.parametrize("method", ["naive", "tiled", "tiled_nocrop", "tiled_1thread"])
def test_oracle_matches_golden(name, method):
    c = case(name)
    nx, ny, nz = c["dims"]
    if method == "naive" and nx * ny * nz > NAIVE_MAX:
        pytest.skip("naive oracle limited to small volumes")
    label = c.get("label", 1)
    g, s = volumes(c, label)
    for exp in c["expected"]:
        r = exp["robustness"]
        kw = {}
        if method.startswith("tiled"):
            kw = {"method": "tiled", "crop": method != "tiled_nocrop",
                  "threads": 1 if method == "tiled_1thread" else 0}
        else:
            kw = {"method": "naive"}
        rc, res, hf, hn, _ = oracle.run(g, s, label=label, robustness=r, **kw)
        assert rc == 0, (name, method, rc)
        check_result(res, exp, f"{name}/{method}/r={r}")
        if r == 1.0:
            check_hist(hf, hn, exp, f"{name}/{method}")
            assert res["iterations"] == res["hd"]


def test_closed_forms():
    for name in ("sphere64_noflip", "single_voxels", "nested_boxes"):
        c = case(name)
        g, s = volumes(c)
        rc, res, *_ = oracle.run(g, s)
        assert rc == 0
        assert res["hd"] == c["closed_form_hd"], name
    # one voxel each: avg HD = HD = L1 distance
    rc, res, *_ = oracle.run(*volumes(case("single_voxels")))
    assert res["avg_hd"] == 99.0 and res["count"] == 2
    # identical masks: HD 0 (the reference reports 1, hello.cu:1893)
    rc, res, *_ = oracle.run(*volumes(case("identical")))
    assert rc == 0 and res["hd"] == 0 and res["count"] == 0 and res["avg_hd"] == 0.0


def test_shift_closed_form_any_direction():
    # HD(K, K + v) = |v|_1 for a convex K away from the border
    nz, ny, nx = 40, 40, 90
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    for v in ((3, 0, 0), (-2, 1, 0), (1, -1, 2), (0, 0, -4), (5, 2, -1)):
        g = (((xx - 45) ** 2 + (yy - 20) ** 2 + (zz - 20) ** 2) <= 100).astype(np.uint8)
        s = (((xx - 45 - v[0]) ** 2 + (yy - 20 - v[1]) ** 2 + (zz - 20 - v[2]) ** 2)
             <= 100).astype(np.uint8)
        for m in ("naive", "tiled"):
            rc, res, *_ = oracle.run(g, s, method=m)
            assert rc == 0 and res["hd"] == sum(abs(a) for a in v), (v, m)


def test_swap_symmetry():
    for name in ("sphere64", "ell_130x70x45", "random3_67x13x9"):
        g, s = volumes(case(name))
        _, a, af, an, _ = oracle.run(g, s)
        _, b, bf, bn, _ = oracle.run(s, g)
        assert a["hd"] == b["hd"] and a["sum"] == b["sum"] and a["count"] == b["count"]
        assert a["sum_fp"] == b["sum_fn"] and a["cnt_fn"] == b["cnt_fp"]
        assert np.array_equal(af, bn) and np.array_equal(an, bf)


def test_empty_mask_is_an_error_not_a_hang():
    g = np.zeros((8, 8, 70), np.uint8)
    s = g.copy()
    s[3, 3, 3] = 1
    for m in ("naive", "tiled"):
        assert oracle.run(g, s, method=m)[0] == -2
        assert oracle.run(s, g, method=m)[0] == -2
    rc, res, *_ = oracle.run(g, g.copy())
    assert rc == 0 and res["hd"] == 0


def test_label_and_dtype():
    c = case("ell_96x80x64")
    g, s = volumes(c)
    exp = c["expected"][0]
    rng = np.random.default_rng(7)
    # other labels around: only == label counts (hello.cu:453-454)
    g3 = np.where(g == 1, 3, rng.choice([0, 1, 2], size=g.shape)).astype(np.uint8)
    s3 = np.where(s == 1, 3, rng.choice([0, 1, 2], size=s.shape)).astype(np.uint8)
    rc, res, *_ = oracle.run(g3, s3, label=3)
    assert rc == 0
    check_result(res, exp, "label3")
    rc, res, *_ = oracle.run(g3.astype(np.int32), s3.astype(np.int32), label=3)
    assert rc == 0
    check_result(res, exp, "i32")


def test_first_hit_map_matches_scipy():
    from scipy import ndimage
    for name in ("sphere64", "random3_67x13x9", "long_200x48x40"):
        g, s = volumes(case(name))
        gb, sb = g.astype(bool), s.astype(bool)
        dg = ndimage.distance_transform_cdt(~gb, metric="taxicab")
        ds = ndimage.distance_transform_cdt(~sb, metric="taxicab")
        want = np.zeros(g.shape, np.uint16)
        want[sb & ~gb] = dg[sb & ~gb]
        want[gb & ~sb] = ds[gb & ~sb]
        for m in ("naive", "tiled"):
            _, _, _, _, mp = oracle.run(g, s, method=m, want_map=True)
            assert np.array_equal(mp, want), (name, m)


def test_kat_reference_points():
    """old/testData.cu: 'covered by first dilatation' pairs have d = 1,
    'covered by second dilatation' pairs d = 2, overlapping voxels are neither
    FP nor FN."""
    c = case("kat_testdata")
    g, s = volumes(c)
    rc, res, _, _, mp = oracle.run(g, s, want_map=True)
    assert rc == 0
    check_result(res, c["expected"][0], "kat_testdata")
    seen = {"first": 0, "second": 0, "overlap": 0}
    for x, y, z, where, group in c["points"]:
        if group == "first":
            assert mp[z, y, x] == 1, (x, y, z)
        elif group == "second":
            assert mp[z, y, x] == 2, (x, y, z)
        elif group == "overlap":
            assert mp[z, y, x] == 0 and g[z, y, x] and s[z, y, x]
        if group in seen:
            seen[group] += 1
    assert seen == {"first": 8, "second": 4, "overlap": 2}
    # old/FirstMetaDataPassTest.cu:137-155: 7 gold and 3 segm voxels of label 2
    c = case("kat_firstmeta")
    g, s = volumes(c, label=2)
    rc, res, *_ = oracle.run(g, s, label=2)
    assert rc == 0
    assert res["n_fn"] == 7 and res["n_fp"] == 3
    check_result(res, c["expected"][0], "kat_firstmeta")


MUTANTS = {1: "first-hit index off by one", 2: "+z neighbour dropped",
           3: "wrong reference mask", 4: "one-sided x shift"}


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_planted_mistakes_are_caught(mutant, tmp_path):
    """Builds the oracle with a planted mistake and checks the golden suite
    rejects it, so the fixtures pin the oracle rather than restate it."""
    import ctypes as C
    so = tmp_path / f"liboracle_m{mutant}.so"
    src = os.path.join(os.path.dirname(oracle.__file__), "hdm_oracle.c")
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-shared", f"-DORACLE_MUTANT={mutant}",
                    "-o", str(so), src, "-lm"], check=True)
    L = C.CDLL(str(so))
    L.oracle_tiled.restype = C.c_int
    caught = False
    for name in ("sphere64", "ell_96x80x64", "long_200x48x40", "random3_67x13x9"):
        c = case(name)
        g, s = volumes(c)
        nz, ny, nx = g.shape
        hl = nx + ny + nz + 2
        hf = np.zeros(hl, np.uint32)
        hn = np.zeros(hl, np.uint32)
        res = oracle.OracleResult()
        rc = L.oracle_tiled(C.c_void_p(g.ctypes.data), C.c_void_p(s.ctypes.data), 0, 1, nx, ny,
                            nz, C.c_double(1.0), 0, 1, C.c_void_p(hf.ctypes.data),
                            C.c_void_p(hn.ctypes.data), hl, None, C.byref(res))
        exp = c["expected"][0]
        d = res.as_dict()
        if rc != 0 or any(int(d[k]) != int(exp[k]) for k in ("hd", "sum", "count")):
            caught = True
            break
    assert caught, f"mutant {mutant} ({MUTANTS[mutant]}) passed the golden checks"
