"""Benchmark of the Hausdorff dilation engine on one MI355X.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]

Headline (value): mask-voxels/s of exact HD + average HD on the CT-sized pair
(512x512x256 uint8, union of 8 seeded ellipsoids with semi-axes 10-60, segm
with +-0..2 voxels per ellipsoid and 0.5% boundary flips), inputs resident in
HBM, one step = one call of the C-ABI entry point (pack, metadata, persistent
dilation, finalize, result copy).  Extra legs on the same JSON line:
  batched  256 pairs of uneven volumes through one persistent launch captured
           in a hipGraph (tile and chunk picked by the autotuner) versus 256
           plain per-pair calls, pairs/s and equality of every result;
  serving  Poisson arrivals against an HBM budget, p50/p99 latency, pairs/s;
  roofline achieved HBM GB/s of the dominant kernel from HIP-event phase times;
  cpu_baseline  the CPU oracle (same tiled algorithm, OpenMP) on the host cores,
           bit-exact against the GPU result.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
CT = dict(dims=(512, 512, 256), n=8, amin=10, amax=60, dmax=2, p=0.005, seed=2024)
BATCH_VOLUMES = [(64, 64, 64), (128, 128, 128), (256, 256, 128), (384, 384, 256)]
TILE_BYTES = {  # per queued (tile, mask): read source + write target + read the
    # other mask's packed original (3 V/8) plus the six halo faces
    1: 3 * 64 * 8 * 8 // 8 + (2 * 8 * 8 + 4 * 64 * 8) // 8,
    2: 3 * 64 * 4 * 4 // 8 + (2 * 4 * 4 + 4 * 64 * 4) // 8,
    3: 3 * 64 * 8 * 4 // 8 + (2 * 8 * 4 + 2 * 64 * 4 + 2 * 64 * 8) // 8,
    4: 3 * 32 * 8 * 8 // 8 + (2 * 8 * 8 + 4 * 32 * 8) // 8,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ct_pair(hdm, seed=CT["seed"]):
    nx, ny, nz = CT["dims"]
    sp = hdm.synth_spec("ellipsoids", nx, ny, nz, CT["n"], CT["amin"], CT["amax"], CT["dmax"],
                        CT["p"], seed)
    return hdm.synth_device(sp)


def leg_single(hdm, torch, steps, warmup):
    G, S = ct_pair(hdm)
    torch.cuda.synchronize()
    nvox = G.numel()
    for _ in range(warmup):
        res = hdm.hausdorff(G, S, timing=True)
    torch.cuda.synchronize()
    phases = {"pack": [], "meta": [], "dilate": [], "finalize": [], "total": []}
    t0 = time.perf_counter()
    for _ in range(steps):
        res = hdm.hausdorff(G, S, timing=True)
        for k, v in hdm.last_timings().items():
            phases[k].append(v)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    avg = {k: sum(v) / len(v) for k, v in phases.items()}
    return G, S, res, dt, nvox, avg


def roofline(res, avg, nvox, tile=1):
    """Algorithmic bytes per launch of each phase over its mean event time."""
    algo = {
        "pack": 2.25 * nvox,  # read 2 uint8 masks, write 2 bits per voxel
        "meta": 0.25 * nvox,  # read both packed masks
        "dilate": float(res["queue_total"]) * TILE_BYTES[tile],
    }
    rows = {}
    for k, b in algo.items():
        ms = avg[k]
        rows[k] = {"ms": round(ms, 4), "algo_bytes": int(b),
                   "gbs": round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 else None}
    dom = max(("pack", "meta", "dilate"), key=lambda k: avg[k])
    r = rows[dom]
    out = {"bound": "hbm", "kernel": dom, "achieved": r["gbs"], "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(r["gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
           "per_phase": rows}
    tf = os.path.join(ROOT, "profiles", f"traffic_ct512_{dom}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            out["traffic"] = json.load(f).get("hbm_bytes_per_launch")
    return out


def leg_cpu(hdm, G, S, res, seconds):
    """The CPU oracle (same tiled algorithm, OpenMP over tiles) on this host."""
    from oracle import oracle
    g, s = G.cpu().numpy(), S.cpu().numpy()
    threads = min(os.cpu_count() or 1, 16)
    times = []
    t_end = time.perf_counter() + seconds
    want = None
    while time.perf_counter() < t_end or len(times) < 3:
        t0 = time.perf_counter()
        rc, want, _, _, _ = oracle.run(g, s, threads=threads)
        times.append(time.perf_counter() - t0)
        if len(times) >= 200:
            break
    exact = rc == 0 and all(int(want[k]) == int(res[k]) for k in
                            ("hd", "sum", "count", "sum_fp", "cnt_fp", "sum_fn", "cnt_fn"))
    med = statistics.median(times)
    return {"value": round(2 * g.size / med, 1), "unit": "mask-voxels/s", "cores": threads,
            "kind": "port", "ms_median": round(med * 1e3, 3),
            "sample": f"the same 512x512x256 pair, {len(times)} runs of the tiled oracle over "
                      f"{sum(times):.1f} s, median", "bit_exact_vs_gpu": bool(exact)}


def leg_batched(hdm, torch, n_pairs, reps, tune):
    import numpy as np
    rng = np.random.default_rng(7)
    vols = []
    for i in range(n_pairs):
        nx, ny, nz = BATCH_VOLUMES[int(rng.integers(0, 4))]
        sp = hdm.synth_spec("ellipsoids", nx, ny, nz, 8, 10, 60, 2, 0.005, 1000 + i)
        vols.append(hdm.synth_device(sp))
    torch.cuda.synchronize()
    nvox = sum(int(g.numel()) for g, _ in vols)
    # plain: one call per pair
    plain = [hdm.hausdorff(g, s) for g, s in vols[:8]]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain = [hdm.hausdorff(g, s) for g, s in vols]
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    best = {"tile_id": hdm.TILE_64x8x8, "chunk": 1, "tile": "64x8x8"}
    table = None
    if tune:
        table = hdm.autotune(vols, reps=2)
        ok = [e for e in table if e["status"] == 0]
        if ok:
            best = ok[0]
    b = hdm.Batch(vols, tile=best["tile_id"], chunk=best["chunk"], graph=True)
    b.run()
    out = b.sync()
    ms = []
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run()
        b.sync(results=False)
        ms.append(b.last_ms())
    t_batch = (time.perf_counter() - t0) / reps
    out = b.sync()
    b.close()
    keys = ("hd", "sum", "count", "sum_fp", "sum_fn", "n_fp", "n_fn")
    same = all(all(o[k] == p[k] for k in keys) for o, p in zip(out, plain))
    return {"pairs": n_pairs, "mask_voxels": 2 * nvox,
            "plain_pairs_per_s": round(n_pairs / t_plain, 1),
            "batched_pairs_per_s": round(n_pairs / t_batch, 1),
            "speedup": round(t_plain / t_batch, 2),
            "batched_mask_voxels_per_s": round(2 * nvox / t_batch, 1),
            "batched_device_ms_median": round(statistics.median(ms), 3),
            "tile": best["tile"], "chunk": best["chunk"], "graph": True,
            "identical_to_plain": bool(same),
            "autotune": [{k: e[k] for k in ("tile", "chunk", "ms", "status")}
                         for e in (table or [])][:12]}


def leg_serving(hdm, torch, seconds, rate, budget_gb):
    from housdorffhdf5cmake_amd import serving
    return serving.poisson_run(seconds=seconds, rate=rate, budget_bytes=int(budget_gb * (1 << 30)),
                               verify=None)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch-pairs", type=int, default=256)
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run a plain device-to-device copy of MIB MiB three times "
                         "(the known-size kernel that scales HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine is single-device (a pair or a batch per "
              "GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    import housdorffhdf5cmake_amd as hdm
    log("single pair 512x512x256 ...")
    G, S, res, dt, nvox, avg = leg_single(hdm, torch, args.steps, args.warmup)
    value = 2 * nvox * args.steps / dt
    line = {
        "metric": "mask-voxels/s + pairs/s for exact HD & avg HD; batched vs plain; p99 lat",
        "value": round(value, 1), "unit": "mask-voxels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "ct_pair_512x512x256_8_ellipsoids", "dims": list(CT["dims"]),
                   "seed": CT["seed"], "connectivity": 6, "tile": "64x8x8"},
        "result": {k: res[k] for k in ("hd", "avg_hd", "count", "sum", "iterations",
                                       "queue_total", "queue_max")},
        "pairs_per_s": round(args.steps / dt, 2),
        "phase_ms": {k: round(v, 4) for k, v in avg.items()},
    }
    line["roofline"] = roofline(res, avg, nvox)
    if args.copy_probe > 0:
        a = torch.ones(args.copy_probe << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if not args.no_batch:
        log("batched vs plain ...")
        line["batched"] = leg_batched(hdm, torch, args.batch_pairs, args.batch_reps,
                                      not args.no_tune)
    if not args.no_serve:
        log("serving ...")
        line["serving"] = leg_serving(hdm, torch, args.serve_seconds, args.serve_rate, 64.0)
    log("cpu baseline ...")
    line["cpu_baseline"] = leg_cpu(hdm, G, S, res, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
