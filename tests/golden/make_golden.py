"""Writes the golden fixtures under tests/golden/ (run once, in a container
with numpy and scipy; the outputs are committed).

Expected values come from tools independent of the engine and of the oracle:

* first-hit d(v) of an FP voxel equals its L1 (taxicab) distance to the gold
  mask, of an FN voxel its distance to segm (6-connectivity dilation inside a
  box; an L1 geodesic between two voxels stays inside their bounding box).
  ``scipy.ndimage.distance_transform_cdt(..., metric='taxicab')`` gives that
  distance exactly; for tiny volumes brute force O(N^2) confirms it.
* the synthetic generator is restated here in numpy (integer arithmetic, same
  RNG stream) so the fixture masks also pin the library's generator.
* the known-answer point sets restate the reference's own fixtures:
  old/testData.cu:4-336 (block-relative points with the behaviour stated in
  their comments) and old/FirstMetaDataPassTest.cu:137-155 (label 2).

The reference is never imported or executed (it is CUDA C++).
"""
from __future__ import annotations

import json
import math
import os

import numpy as np
from scipy import ndimage

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


# ---------------------------------------------------------------- generator
def mix64(z: int) -> int:
    z &= M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


class Rng:
    def __init__(self, seed: int):
        self.s = seed & M64

    def next(self) -> int:
        self.s = (self.s + GOLDEN) & M64
        return mix64(self.s)

    def u01(self) -> float:
        return float(self.next() >> 11) * (1.0 / 9007199254740992.0)

    def range(self, lo: int, hi: int) -> int:
        return lo + self.next() % (hi - lo + 1)


def ell(c, s, r):
    return {"c": list(c), "s": list(s), "r": list(r)}


EYE = [1024, 0, 0, 0, 1024, 0, 0, 0, 1024]


def spec_sphere(dims, c, r, d, p_flip, seed):
    g = ell(c, (r, r, r), EYE)
    s = ell((c[0] + d[0], c[1] + d[1], c[2] + d[2]), (r, r, r), EYE)
    return {"dims": list(dims), "gold": [g], "segm": [s], "flip_mode": 1 if p_flip > 0 else 0,
            "flip_seed": mix64(seed), "flip_thresh": int(p_flip * 9007199254740992.0),
            "cube_lo": [0, 0, 0], "cube_hi": [0, 0, 0]}


def spec_ellipsoids(dims, n, amin, amax, dmax, p_flip, seed):
    g = Rng(seed)
    gold, segm = [], []
    for _ in range(n):
        s = [g.range(amin, amax) for _ in range(3)]
        c = []
        for k in range(3):
            m = dims[k] // 8
            c.append(g.range(m, max(m, dims[k] - 1 - m)))
        while True:
            w = 2.0 * g.u01() - 1.0
            x = 2.0 * g.u01() - 1.0
            y = 2.0 * g.u01() - 1.0
            z = 2.0 * g.u01() - 1.0
            n2 = w * w + x * x + y * y + z * z
            if not (n2 > 1.0 or n2 < 1e-4):
                break
        nr = math.sqrt(n2)
        w, x, y, z = w / nr, x / nr, y / nr, z / nr
        R = [1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - z * w), 2.0 * (x * z + y * w),
             2.0 * (x * y + z * w), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - x * w),
             2.0 * (x * z - y * w), 2.0 * (y * z + x * w), 1.0 - 2.0 * (x * x + y * y)]
        rq = [int(math.floor(v * 1024.0 + 0.5)) for v in R]
        delta = g.range(-dmax, dmax)
        gold.append(ell(c, s, rq))
        segm.append(ell(c, [max(1, v + delta) for v in s], rq))
    return {"dims": list(dims), "gold": gold, "segm": segm, "flip_mode": 2 if p_flip > 0 else 0,
            "flip_seed": mix64(seed), "flip_thresh": int(p_flip * 9007199254740992.0),
            "cube_lo": [0, 0, 0], "cube_hi": [0, 0, 0]}


def spec_long(dims, c, a, gap, edge):
    g = ell(c, a, EYE)
    lo = [c[0] + a[0] + gap, c[1] - edge // 2, c[2] - edge // 2]
    return {"dims": list(dims), "gold": [g], "segm": [dict(g)], "flip_mode": 0,
            "flip_seed": 0, "flip_thresh": 0, "cube_lo": lo, "cube_hi": [v + edge for v in lo]}


def _ell_mask(e, X, Y, Z):
    d0 = X - e["c"][0]
    d1 = Y - e["c"][1]
    d2 = Z - e["c"][2]
    r = e["r"]
    q = [r[0] * d0 + r[1] * d1 + r[2] * d2, r[3] * d0 + r[4] * d1 + r[5] * d2,
         r[6] * d0 + r[7] * d1 + r[8] * d2]
    a = [np.abs(v).astype(np.uint64) for v in q]
    s = [np.uint64(v) for v in e["s"]]
    ok = (a[0] <= np.uint64(1024) * s[0]) & (a[1] <= np.uint64(1024) * s[1]) & \
         (a[2] <= np.uint64(1024) * s[2])
    a = [np.where(ok, v, np.uint64(0)) for v in a]
    s2 = [v * v for v in s]
    lhs = a[0] * a[0] * (s2[1] * s2[2]) + a[1] * a[1] * (s2[0] * s2[2]) + \
        a[2] * a[2] * (s2[0] * s2[1])
    rhs = (s2[0] * s2[1] * s2[2]) << np.uint64(20)
    return ok & (lhs <= rhs)


def synth_eval(spec):
    """Returns (gold, segm) as uint8 arrays indexed [z, y, x]."""
    nx, ny, nz = spec["dims"]
    Z, Y, X = np.meshgrid(np.arange(nz, dtype=np.int64), np.arange(ny, dtype=np.int64),
                          np.arange(nx, dtype=np.int64), indexing="ij")
    gold = np.zeros((nz, ny, nx), bool)
    for e in spec["gold"]:
        gold |= _ell_mask(e, X, Y, Z)
    base = np.zeros((nz, ny, nx), bool)
    lo, hi = spec["cube_lo"], spec["cube_hi"]
    base |= (X >= lo[0]) & (X < hi[0]) & (Y >= lo[1]) & (Y < hi[1]) & (Z >= lo[2]) & (Z < hi[2])
    for e in spec["segm"]:
        base |= _ell_mask(e, X, Y, Z)
    segm = base.copy()
    mode = spec["flip_mode"]
    if mode:
        v = (X + nx * (Y + ny * Z)).astype(np.uint64)
        with np.errstate(over="ignore"):
            h = np.uint64(spec["flip_seed"]) + (v + np.uint64(1)) * np.uint64(GOLDEN)
            h = (h ^ (h >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            h = (h ^ (h >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            h = h ^ (h >> np.uint64(31))
        cand = (h >> np.uint64(11)) < np.uint64(spec["flip_thresh"])
        pad = np.pad(base, 1, constant_values=False)
        nbs = [pad[1:-1, 1:-1, :-2], pad[1:-1, 1:-1, 2:], pad[1:-1, :-2, 1:-1],
               pad[1:-1, 2:, 1:-1], pad[:-2, 1:-1, 1:-1], pad[2:, 1:-1, 1:-1]]
        if mode == 1:
            any_off = np.zeros_like(base)
            for nb in nbs:  # out-of-volume counts as off
                any_off |= ~nb
            segm = np.where(cand & base & any_off, False, base)
        else:
            inside = np.ones((nz + 2, ny + 2, nx + 2), bool)
            inside[0, :, :] = inside[-1, :, :] = False
            inside[:, 0, :] = inside[:, -1, :] = False
            inside[:, :, 0] = inside[:, :, -1] = False
            ins = [inside[1:-1, 1:-1, :-2], inside[1:-1, 1:-1, 2:], inside[1:-1, :-2, 1:-1],
                   inside[1:-1, 2:, 1:-1], inside[:-2, 1:-1, 1:-1], inside[2:, 1:-1, 1:-1]]
            differs = np.zeros_like(base)
            for nb, ok in zip(nbs, ins):
                differs |= ok & (nb != base)
            segm = np.where(cand & differs, ~base, base)
    return gold.astype(np.uint8), segm.astype(np.uint8)


# ---------------------------------------------------------------- expectations
def first_hits(gold: np.ndarray, segm: np.ndarray):
    """d(v) on FP (distance to gold) and FN (distance to segm) via scipy."""
    g = gold.astype(bool)
    s = segm.astype(bool)
    fp = s & ~g
    fn = g & ~s
    d_fp = np.zeros(g.shape, np.int64)
    d_fn = np.zeros(g.shape, np.int64)
    if fp.any():
        d_fp = ndimage.distance_transform_cdt(~g, metric="taxicab").astype(np.int64)
    if fn.any():
        d_fn = ndimage.distance_transform_cdt(~s, metric="taxicab").astype(np.int64)
    return fp, fn, d_fp, d_fn


def brute_first_hits(gold, segm):
    g = np.argwhere(gold.astype(bool))
    s = np.argwhere(segm.astype(bool))
    fp = np.argwhere(segm.astype(bool) & ~gold.astype(bool))
    fn = np.argwhere(gold.astype(bool) & ~segm.astype(bool))
    out_fp = [int(np.abs(g - p).sum(1).min()) for p in fp] if len(g) else []
    out_fn = [int(np.abs(s - p).sum(1).min()) for p in fn] if len(s) else []
    return sorted(out_fp), sorted(out_fn)


def expected(gold, segm, robust=1.0):
    fp, fn, d_fp, d_fn = first_hits(gold, segm)
    vf = d_fp[fp]
    vn = d_fn[fn]
    L = int(max(vf.max(initial=0), vn.max(initial=0))) + 1
    hf = np.bincount(vf, minlength=L).astype(np.int64)
    hn = np.bincount(vn, minlength=L).astype(np.int64)
    out = {"n_gold": int(gold.astype(bool).sum()), "n_segm": int(segm.astype(bool).sum()),
           "n_fp": int(fp.sum()), "n_fn": int(fn.sum()),
           "hist_fp": hf.tolist(), "hist_fn": hn.tolist()}
    res = {}
    for name, h, tot in (("fp", hf, out["n_fp"]), ("fn", hn, out["n_fn"])):
        tgt = min(tot, math.ceil(robust * tot)) if tot else 0
        cnt = s_ = hd = 0
        if tgt:
            for t in range(1, L):
                cnt += int(h[t])
                s_ += t * int(h[t])
                if h[t]:
                    hd = t
                if cnt >= tgt:
                    break
        res[name] = (hd, s_, cnt)
    out.update({"hd_fp": res["fp"][0], "sum_fp": res["fp"][1], "cnt_fp": res["fp"][2],
                "hd_fn": res["fn"][0], "sum_fn": res["fn"][1], "cnt_fn": res["fn"][2]})
    out["hd"] = max(out["hd_fp"], out["hd_fn"])
    out["sum"] = out["sum_fp"] + out["sum_fn"]
    out["count"] = out["cnt_fp"] + out["cnt_fn"]
    out["robustness"] = robust
    return out


def save_masks(name, gold, segm):
    path = os.path.join(HERE, f"masks_{name}.npz")
    np.savez_compressed(path, shape=np.array(gold.shape, np.int64),
                        gold=np.packbits(gold.astype(bool).ravel()),
                        segm=np.packbits(segm.astype(bool).ravel()))
    return os.path.basename(path)


# ---------------------------------------------------------------- KAT points
def kat_testdata_points(db=(32, 8, 32), meta=(10, 6, 12)):
    """Points of old/testData.cu:4-336 as (x, y, z, where, group): where is
    'gold', 'segm' or 'both' (the 'overlapping' points, :46-57); group names
    the behaviour the reference's comments state for the point."""
    mx, my, mz = meta
    dx, dy, dz = db

    def P(x, y, z, is_gold, xm, ym, zm, group="", both=False):
        return (xm * dx + x, ym * dy + y, zm * dz + z, "both" if both else
                ("gold" if is_gold else "segm"), group)

    pts = [
        P(2, 5, 8, True, 2, 2, 2), P(3, 3, 9, True, 2, 2, 2), P(1, 5, 3, True, 2, 2, 2),
        P(0, 0, 1, True, 0, 0, 0), P(dx - 2, dy - 2, dz - 2, True, mx - 1, my - 1, mz - 1),
        P(0, 0, 0, False, 0, 0, 0), P(dx - 2, dy - 2, dz - 2, False, mx - 1, my - 1, mz - 1),
        P(5, 6, 7, False, 3, 4, 6, "overlap", both=True),
        P(9, 11, 7, False, 3, 4, 6, "overlap", both=True),
        # "covered by first dilatation" (:61-106)
        P(9, 11, 7, False, 7, 4, 6, "first"), P(9, 11, 8, True, 7, 4, 6, "first"),
        P(9, 3, 7, False, 7, 4, 6, "first"), P(9, 2, 7, True, 7, 4, 6, "first"),
        P(9, 5, 7, False, 7, 4, 6, "first"), P(9, 6, 7, True, 7, 4, 6, "first"),
        P(2, 3, 7, False, 7, 4, 6, "first"), P(3, 3, 7, True, 7, 4, 6, "first"),
        # "covered by second dilatation" (:110-132)
        P(9, 11, 7, False, 7, 2, 6, "second"), P(9, 11, 9, True, 7, 2, 6, "second"),
        P(9, 3, 7, False, 7, 2, 6, "second"), P(9, 1, 7, True, 7, 2, 6, "second"),
        # face-crossing probes (:136-210)
        P(2, 2, 0, False, 0, 0, 2, "face"), P(2, 2, dz - 1, False, 0, 0, 4, "face"),
        P(0, 2, 2, False, 8, 0, 6, "face"), P(dx - 1, 3, 7, False, 0, 0, 8, "face"),
        P(9, dy - 1, 7, False, 0, 0, 10, "face"), P(9, 0, 7, False, 2, 2, 4, "face"),
        # corner probes (:266-326)
        P(0, dy - 1, 0, False, 2, 2, 6, "corner"), P(dx - 1, dy - 1, 0, False, 2, 2, 8, "corner"),
        P(dx - 1, 0, dz - 1, False, 4, 4, 4, "corner"),
        # "activated after two dilatations" (:328-333)
        P(1, 1, 1, False, 4, 4, 8, "late"),
    ]
    dims = (mx * dx, my * dy, mz * dz)
    return dims, pts


def kat_firstmeta_points():
    """old/FirstMetaDataPassTest.cu:137-155: label-2 voxels in a 448x112x544
    volume of 32x8x32 blocks."""
    dx, dy, dz = 32, 8, 32
    gold = [(dx + 2, dy + 1, dz * 3 + 1), (dx + 2, dy + 2, dz * 3 + 1),
            (dx + 3, dy + 2, dz * 3 + 1), (dx * 2 + 2, dy * 2 + 3, dz * 2 + 4),
            (dx * 2 + 3, dy * 2 + 3, dz * 2 + 4), (dx * 3 + 2, dy + 2, dz * 2 + 5),
            (dx * 4 + 9, dy * 2, dz * 2 + 1)]
    segm = [(dx * 6 + 1, dy * 2, dz * 2 + 1), (dx * 6 + 5, dy * 2, dz * 2 + 1),
            (dx * 7 + 4, dy * 2, dz * 2 + 1)]
    return (14 * dx, 14 * dy, 17 * dz), gold, segm


def points_volume(dims, pts):
    nx, ny, nz = dims
    g = np.zeros((nz, ny, nx), np.uint8)
    s = np.zeros((nz, ny, nx), np.uint8)
    for x, y, z, where, _ in pts:
        if where in ("gold", "both"):
            g[z, y, x] = 1
        if where in ("segm", "both"):
            s[z, y, x] = 1
    return g, s


# ---------------------------------------------------------------- main
def main():
    cases = []

    def add(name, gold, segm, kind, extra=None, robust_list=(1.0,), brute=False, store=True):
        exp = [expected(gold, segm, r) for r in robust_list]
        if brute:
            bf, bn = brute_first_hits(gold, segm)
            fp, fn, d_fp, d_fn = first_hits(gold, segm)
            assert sorted(d_fp[fp].tolist()) == bf, name
            assert sorted(d_fn[fn].tolist()) == bn, name
        c = {"name": name, "kind": kind, "dims": [int(gold.shape[2]), int(gold.shape[1]),
                                                  int(gold.shape[0])], "expected": exp}
        if store:
            c["masks"] = save_masks(name, gold, segm)
        if extra:
            c.update(extra)
        cases.append(c)
        print(name, "hd", exp[0]["hd"], "count", exp[0]["count"], "sum", exp[0]["sum"])

    # workload 1: 64^3 sphere r=20, segm shifted (+2,+1,0), 1% surface flips
    sp = spec_sphere((64, 64, 64), (32, 32, 32), 20, (2, 1, 0), 0.01, 1)
    g, s = synth_eval(sp)
    add("sphere64", g, s, "synth", {"spec": sp, "builder": ["sphere", [64, 64, 64, 32, 32, 32,
        20, 2, 1, 0, 0.01, 1]]}, robust_list=(1.0, 0.95, 0.5))
    sp = spec_sphere((64, 64, 64), (32, 32, 32), 20, (2, 1, 0), 0.0, 1)
    g, s = synth_eval(sp)
    add("sphere64_noflip", g, s, "synth", {"spec": sp, "builder": ["sphere", [64, 64, 64, 32,
        32, 32, 20, 2, 1, 0, 0.0, 1]], "closed_form_hd": 3})
    # ellipsoid unions (workload 2 generator at reduced size)
    for name, dims, amin, amax, seed in (("ell_96x80x64", (96, 80, 64), 4, 16, 2),
                                          ("ell_160x128x96", (160, 128, 96), 6, 24, 3),
                                          ("ell_130x70x45", (130, 70, 45), 5, 20, 4)):
        sp = spec_ellipsoids(dims, 8, amin, amax, 2, 0.005, seed)
        g, s = synth_eval(sp)
        add(name, g, s, "synth", {"spec": sp, "builder": ["ellipsoids", [*dims, 8, amin, amax,
            2, 0.005, seed]]}, robust_list=(1.0, 0.9))
    # long-iteration shape (workload 3 generator at reduced size)
    sp = spec_long((200, 48, 40), (30, 24, 20), (20, 12, 10), 120, 3)
    g, s = synth_eval(sp)
    add("long_200x48x40", g, s, "synth", {"spec": sp, "builder": ["long", [200, 48, 40, 30, 24,
        20, 20, 12, 10, 120, 3]]})
    # random blobs, brute-force confirmed
    rs = np.random.default_rng(12345)
    for i, dims in enumerate(((12, 10, 9), (16, 16, 16), (7, 13, 5), (67, 13, 9))):
        nx, ny, nz = dims
        g = (rs.random((nz, ny, nx)) < 0.06).astype(np.uint8)
        s = (rs.random((nz, ny, nx)) < 0.06).astype(np.uint8)
        add(f"random{i}_{nx}x{ny}x{nz}", g, s, "random", robust_list=(1.0, 0.75),
            brute=nx * ny * nz <= 6000)
    # identical masks: HD 0
    g = np.zeros((20, 24, 40), np.uint8)
    g[5:12, 3:20, 7:30] = 1
    add("identical", g, g.copy(), "edge")
    # single voxels: HD = avg = L1
    g = np.zeros((30, 31, 70), np.uint8)
    s = np.zeros_like(g)
    g[3, 4, 5] = 1
    s[25, 20, 66] = 1
    add("single_voxels", g, s, "edge", {"closed_form_hd": 22 + 16 + 61})
    # both empty
    add("both_empty", np.zeros((8, 8, 64), np.uint8), np.zeros((8, 8, 64), np.uint8), "edge")
    # nested boxes (closed form: HD = max face gap)
    g = np.zeros((40, 40, 130), np.uint8)
    s = np.zeros_like(g)
    g[10:30, 10:30, 30:100] = 1
    s[13:25, 12:30, 34:99] = 1
    add("nested_boxes", g, s, "edge", {"closed_form_hd": 4 + 2 + 5})
    # KATs
    dims, pts = kat_testdata_points()
    g, s = points_volume(dims, pts)
    add("kat_testdata", g, s, "kat", {"points": [list(p) for p in pts]}, store=False)
    dims, gp, spts = kat_firstmeta_points()
    pts = [(x, y, z, "gold", "") for x, y, z in gp] + [(x, y, z, "segm", "") for x, y, z in spts]
    g, s = points_volume(dims, pts)
    add("kat_firstmeta", g, s, "kat", {"points": [list(p) for p in pts], "label": 2},
        store=False)
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump({"generated_by": "tests/golden/make_golden.py", "cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
