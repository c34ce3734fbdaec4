"""The 256-pair batched workload on the GPU: volumes drawn uniformly from
{64^3, 128^3, 256x256x128, 384x384x256}, each an 8-ellipsoid union pair from
its own seed (1000 + i; the draw from numpy's default_rng(7)), through one
persistent launch captured in a hipGraph with the tile and queue chunk the
autotuner picked.  Every result must equal the pair computed alone (one
plain call each, the path that replaces the reference's per-pair
cooperative launch, hello.cu:1999) and the CPU oracle, bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

import housdorffhdf5cmake_amd as hdm
from oracle import oracle

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

VOLUMES = [(64, 64, 64), (128, 128, 128), (256, 256, 128), (384, 384, 256)]
KEYS = ("hd", "sum", "count", "sum_fp", "cnt_fp", "sum_fn", "cnt_fn", "n_fp", "n_fn", "n_gold",
        "n_segm", "hd_fp", "hd_fn")


def mix(n_pairs=256):
    rng = np.random.default_rng(7)
    dims, vols = [], []
    for i in range(n_pairs):
        nx, ny, nz = VOLUMES[int(rng.integers(0, 4))]
        sp = hdm.synth_spec("ellipsoids", nx, ny, nz, 8, 10, 60, 2, 0.005, 1000 + i)
        vols.append(hdm.synth_device(sp))
        dims.append((nx, ny, nz))
    torch.cuda.synchronize()
    return dims, vols


def test_config4_batch_graph_autotuned_equals_alone_and_oracle():
    dims, vols = mix()
    counts = {d: dims.count(d) for d in VOLUMES}
    assert min(counts.values()) >= 4, counts  # every class is represented
    print("classes", counts, flush=True)
    table = hdm.autotune(vols, reps=1)
    ok = [e for e in table if e["status"] == 0]
    assert ok and len(ok) == len(table), table  # every candidate agreed bit for bit
    best = ok[0]
    print("autotuned", best, flush=True)
    b = hdm.Batch(vols, tile=best["tile_id"], chunk=best["chunk"], graph=True)
    runs = []
    for _ in range(2):  # two replays of the captured graph
        b.run()
        runs.append(b.sync())
    b.close()
    assert runs[0] == runs[1]
    alone = [hdm.hausdorff(g, s) for g, s in vols]
    for i, (o, a) in enumerate(zip(runs[1], alone)):
        assert all(o[k] == a[k] for k in KEYS), (i, dims[i])
    bad = []
    for i, ((g, s), a) in enumerate(zip(vols, alone)):
        rc, want, _, _, _ = oracle.run(g.cpu().numpy(), s.cpu().numpy())
        if rc != 0 or any(int(want[k]) != int(a[k]) for k in KEYS):
            bad.append((i, dims[i]))
    assert not bad, bad
