"""hdm_hausdorff_h5 (the reference's HDF5 benchmark path,
forBench/hdf5Manag.cu:196-220) on the GPU: two flat boolean datasets of one
file -> HD, bit-exact against the oracle and against the in-memory entry
point on the same volumes."""
from __future__ import annotations

import os

import numpy as np
import pytest

import housdorffhdf5cmake_amd as hdm
from housdorffhdf5cmake_amd import h5
from oracle import oracle

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

H5 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "h5")
KEYS = ("hd", "sum", "count", "sum_fp", "cnt_fp", "sum_fn", "cnt_fn", "n_fp", "n_fn")


def spheres():
    z, y, x = np.meshgrid(np.arange(20), np.arange(30), np.arange(40), indexing="ij")
    g = (x - 20) ** 2 + (y - 15) ** 2 + (z - 10) ** 2 <= 49
    s = (x - 22) ** 2 + (y - 16) ** 2 + (z - 10) ** 2 <= 49
    s[5, 4, 3] = True
    return g.astype(np.uint8), s.astype(np.uint8)


@pytest.mark.parametrize("conn", [6, 18, 26])
def test_reference_layout_file(conn):
    f = os.path.join(H5, "bool_flat.h5")
    res = hdm.hausdorff_h5(f, "onlyBladderBoolFlat", f, "onlyLungsBoolFlat", 40, 30, 20,
                           connectivity=conn)
    g, s = spheres()
    rc, want, *_ = oracle.run(g, s, conn=conn)
    assert rc == 0
    assert {k: res[k] for k in KEYS} == {k: want[k] for k in KEYS}
    assert res["load_ms"] >= 0


def test_written_pair_matches_in_memory(tmp_path):
    sp = hdm.synth_spec("ellipsoids", 160, 128, 96, 8, 6, 24, 2, 0.005, 41)
    g, s = hdm.synth_host(sp)
    f = str(tmp_path / "pair.h5")
    h5.write(f, {"gold": g.reshape(-1).astype(bool), "segm": s.reshape(-1).astype(bool)})
    a = hdm.hausdorff_h5(f, "gold", f, "segm", 160, 128, 96)
    b = hdm.hausdorff(torch.from_numpy(g).cuda(), torch.from_numpy(s).cuda())
    assert {k: a[k] for k in KEYS} == {k: b[k] for k in KEYS}
