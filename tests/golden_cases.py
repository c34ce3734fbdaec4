"""Loader for the committed golden fixtures (tests/golden/)."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def manifest():
    with open(os.path.join(GOLDEN, "cases.json")) as f:
        return json.load(f)["cases"]


def case(name):
    for c in manifest():
        if c["name"] == name:
            return c
    raise KeyError(name)


def volumes(c, label=1):
    """[z, y, x] uint8 volumes of a case; voxels on are set to `label`."""
    nx, ny, nz = c["dims"]
    if "masks" in c:
        d = np.load(os.path.join(GOLDEN, c["masks"]))
        n = nx * ny * nz
        g = np.unpackbits(d["gold"])[:n].reshape(nz, ny, nx)
        s = np.unpackbits(d["segm"])[:n].reshape(nz, ny, nx)
    else:
        g = np.zeros((nz, ny, nx), np.uint8)
        s = np.zeros((nz, ny, nx), np.uint8)
        for x, y, z, where, _ in c["points"]:
            if where in ("gold", "both"):
                g[z, y, x] = 1
            if where in ("segm", "both"):
                s[z, y, x] = 1
    return (g * np.uint8(label)).astype(np.uint8), (s * np.uint8(label)).astype(np.uint8)


RESULT_KEYS = ("hd", "hd_fp", "hd_fn", "sum", "count", "sum_fp", "cnt_fp", "sum_fn", "cnt_fn",
               "n_gold", "n_segm", "n_fp", "n_fn")


def check_result(res, exp, where=""):
    for k in RESULT_KEYS:
        assert int(res[k]) == int(exp[k]), f"{where}: {k} {res[k]} != {exp[k]}"
    avg = exp["sum"] / exp["count"] if exp["count"] else 0.0
    assert res["avg_hd"] == avg, f"{where}: avg {res['avg_hd']} != {avg}"


def check_hist(hf, hn, exp, where=""):
    ef = np.array(exp["hist_fp"], np.int64)
    en = np.array(exp["hist_fn"], np.int64)
    L = max(len(ef), len(en))
    got_f = np.zeros(max(L, len(hf)), np.int64)
    got_f[:len(hf)] = hf
    got_n = np.zeros(max(L, len(hn)), np.int64)
    got_n[:len(hn)] = hn
    want_f = np.zeros_like(got_f)
    want_f[:len(ef)] = ef
    want_n = np.zeros_like(got_n)
    want_n[:len(en)] = en
    assert np.array_equal(got_f, want_f), f"{where}: FP histogram differs"
    assert np.array_equal(got_n, want_n), f"{where}: FN histogram differs"
