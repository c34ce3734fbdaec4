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
  serving_loaded  the same mix with resident packed masks at 200 and 5000
           requests/s and in a closed loop of 64 requests (sustained pairs/s,
           pairs per slice), and the closed loop with volumes;
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
# the same generator on a 512^3 volume, twice the ellipsoids for the same
# foreground share: a short-HD 512^3 pair (the long-iteration 512^3 pair
# is the other kind)
CT_CUBE = dict(dims=(512, 512, 512), n=16, amin=10, amax=60, dmax=2, p=0.005, seed=2025)
BATCH_VOLUMES = [(64, 64, 64), (128, 128, 128), (256, 256, 128), (384, 384, 256)]
TILE_BYTES = {  # per queued (tile, mask): read source + write target + read the
    # other mask's packed original (3 V/8) plus the six halo faces.  An entry
    # moves these bytes whether it runs one dilation step or two.
    1: 3 * 64 * 8 * 8 // 8 + (2 * 8 * 8 + 4 * 64 * 8) // 8,
    2: 3 * 64 * 4 * 4 // 8 + (2 * 4 * 4 + 4 * 64 * 4) // 8,
    3: 3 * 64 * 8 * 4 // 8 + (2 * 8 * 4 + 2 * 64 * 4 + 2 * 64 * 8) // 8,
    4: 3 * 32 * 8 * 8 // 8 + (2 * 8 * 8 + 4 * 32 * 8) // 8,
}
# per queued (region, mask) of the region launch (64x16x16 voxels, eight steps):
# read the 32x32-row window of the source (the region's rows and an 8-row ring,
# 8 B each, plus 8 columns on each x side, 2 B per row), read the other mask's
# 256 core rows and write the 256 core rows of the result
REGION_BYTES = 32 * 32 * (8 + 2) + 2 * 16 * 16 * 8
# per (region, mask) entry of the near pass (the same window, eight steps at
# most, after the pack in the same launch): read the window of the source and
# the other mask's 256 core rows; nothing is written back
NEAR_BYTES = 32 * 32 * (8 + 2) + 16 * 16 * 8


def entry_bytes(steps, tile=1, near=False):
    """Algorithmic bytes one queued entry moves: a near-pass entry, a region
    entry (eight steps per barrier) or a tile entry (one or two)."""
    if near:
        return NEAR_BYTES
    return REGION_BYTES if steps == 8 else TILE_BYTES[tile]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ct_pair(hdm, seed=None, spec=CT):
    nx, ny, nz = spec["dims"]
    sp = hdm.synth_spec("ellipsoids", nx, ny, nz, spec["n"], spec["amin"], spec["amax"],
                        spec["dmax"], spec["p"], spec["seed"] if seed is None else seed)
    return hdm.synth_device(sp)


COPIES = 4  # 4 x 134 MB of input, more than the 256 MB Infinity Cache


def leg_single(hdm, torch, steps, warmup, spec=CT, copies=COPIES):
    """Timed steps cycle over COPIES device copies of the pair, so no step
    reads its input from the Infinity Cache where the previous step left it,
    and bracket the dominant launch with HIP events (timing 3: the fused
    launch, or the dilation kernel of a separate-kernel call; the events are
    read after the timed region); the phases are timed by as many extra calls
    with timing 1."""
    G, S = ct_pair(hdm, spec=spec)
    pairs = [(G, S)] + [(G.clone(), S.clone()) for _ in range(copies - 1)]
    torch.cuda.synchronize()
    nvox = G.numel()
    for i in range(warmup):
        res = hdm.hausdorff(*pairs[i % copies], timing=3)
    torch.cuda.synchronize()
    hdm.timing_log()
    t0 = time.perf_counter()
    for i in range(steps):
        res = hdm.hausdorff(*pairs[i % copies], timing=3)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    dom = hdm.timing_log()  # the events of the timed steps, read after them
    lt = hdm.last_timings()
    fused, steps_per_barrier, launches = lt["fused"], lt["steps"], lt.get("launches", 0)
    assert len(dom) == steps and min(dom) > 0, dom
    phases = {"pack": [], "meta": [], "dilate": [], "finalize": [], "total": []}
    for i in range(steps):
        hdm.hausdorff(*pairs[i % copies], timing=True)
        for k, v in hdm.last_timings().items():
            if k in phases:
                phases[k].append(v)
    avg = {k: sum(v) / len(v) for k, v in phases.items()}
    avg["dominant_timed_steps"] = sum(dom) / len(dom)
    avg["fused"] = fused
    avg["steps_per_barrier"] = steps_per_barrier
    avg["launches"] = launches
    del pairs
    return G, S, res, dt, nvox, avg


def copy_bandwidth(torch, mib=1024, reps=5):
    """Attainable HBM bandwidth on this box: a device-to-device copy of mib
    MiB (read + write bytes over the event time, torch's current stream, where
    the copy runs), mean of reps copies after one warm-up."""
    a = torch.empty(mib << 20, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return round(2 * (mib << 20) / (ms * 1e-3) / 1e9, 1)


def roofline_obj(kernel, algo_bytes, ms, traffic_name, **extra):
    """The roofline object of one workload: algorithmic bytes per launch over
    the launch's mean HIP-event time, against MI355X's 8 TB/s; traffic is the
    counted HBM bytes per launch from profiles/traffic_<traffic_name>.json
    (two rocprofv3 counter passes, scaled by a copy of known size), or None."""
    gbs = algo_bytes / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "kernel": kernel, "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
           "algo_bytes": int(algo_bytes), "kernel_ms": round(ms, 4)}
    tf = os.path.join(ROOT, "profiles", f"traffic_{traffic_name}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            out["traffic"] = json.load(f).get("hbm_bytes_per_launch")
    if out["traffic"]:
        # the same launch time over the counted HBM bytes: below the
        # algorithmic fraction where the launch moves fewer bytes than the
        # algorithm's count (the near launch writes no rows of a mask that is
        # empty in a tile), above it where it re-reads
        cg = out["traffic"] / (ms * 1e-3) / 1e9
        out["achieved_counted"] = round(cg, 1)
        out["frac_counted"] = round(cg / HBM_PEAK_GBS, 4)
    out.update(extra)
    return out


def roofline(res, avg, nvox, tile=1, traffic_name="ct512"):
    """Algorithmic bytes per call of each phase over its mean event time; the
    dominant kernel is the fused launch (and the region launch after it when
    the pair was handed over)."""
    steps = avg.get("steps_per_barrier", 1)
    # the near pass finished the pair in the fused launch itself
    near = bool(avg.get("fused")) and steps == 8 and avg.get("launches") == 1
    moved = float(res["queue_total"]) * entry_bytes(steps, tile, near)
    algo = {
        "pack": 2.25 * nvox,  # read 2 uint8 masks, write 2 bits per voxel
        # the seed reads both packed masks; the near pass's count reads only
        # the region bitmap
        "meta": 0.0 if near else 0.25 * nvox,
        "dilate": moved,      # counted as moved: one entry's bytes per entry
    }
    rows = {}
    for k, b in algo.items():
        ms = avg[k]
        rows[k] = {"ms": round(ms, 4), "algo_bytes": int(b),
                   "gbs": round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 else None}
    if avg.get("fused"):
        # one launch does it all: its algorithmic bytes are the sum of the
        # phases' (the seed reads 2 flag bytes per tile and is left out).  The
        # second count charges a two-step tile entry one tile's bytes per step.
        extra = {"per_phase": rows, "dilation_steps_per_entry": steps,
                 "body": "near pass" if near else "two-step" if steps == 2 else
                 "region launch" if steps == 8 else "one step"}
        if steps in (1, 2):
            extra["algo_bytes_counting_each_step"] = int(algo["pack"] + moved * steps)
        return roofline_obj("fused", algo["pack"] + algo["dilate"], avg["dominant_timed_steps"],
                            traffic_name, **extra)
    dom = max(("pack", "meta", "dilate"), key=lambda k: avg[k])
    ms = avg["dominant_timed_steps"] if dom == "dilate" else avg[dom]
    return roofline_obj(dom, algo[dom], ms, f"ct512_{dom}", per_phase=rows)


def _oracle_times(oracle, g, s, threads, seconds, min_runs=5):
    """One warm-up run, then runs until `seconds` have passed and at least
    min_runs were timed; returns (median seconds, runs, total seconds, result)."""
    rc, want, _, _, _ = oracle.run(g, s, threads=threads)
    times = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end or len(times) < min_runs:
        t0 = time.perf_counter()
        rc, want, _, _, _ = oracle.run(g, s, threads=threads)
        times.append(time.perf_counter() - t0)
        if len(times) >= 2000:
            break
    return statistics.median(times), len(times), sum(times), (rc, want)


EXACT_KEYS = ("hd", "sum", "count", "sum_fp", "cnt_fp", "sum_fn", "cnt_fn", "n_fp", "n_fn")


def _same(rcw, res):
    rc, want = rcw
    return rc == 0 and all(int(want[k]) == int(res[k]) for k in EXACT_KEYS)


def leg_cpu(hdm, G, S, res, seconds, extra=()):
    """The CPU oracle (the same tiled dilation, OpenMP over tiles) on this
    host's cores, single-core and all-core, bit-exact against the GPU; `extra`
    is a list of (name, gold, segm, gpu result) also checked bit for bit."""
    from oracle import oracle
    g, s = G.cpu().numpy(), S.cpu().numpy()
    threads = min(os.cpu_count() or 1, 16)
    med, runs, tot, rcw = _oracle_times(oracle, g, s, threads, seconds)
    med1, runs1, tot1, rcw1 = _oracle_times(oracle, g, s, 1, seconds / 2, min_runs=3)
    checks = {}
    for name, eg, es, eres in extra:
        rc, want, _, _, _ = oracle.run(eg.cpu().numpy(), es.cpu().numpy(), threads=threads)
        checks[name] = checks.get(name, True) and _same((rc, want), eres)
    return {"value": round(2 * g.size / med, 1), "unit": "mask-voxels/s", "cores": threads,
            "kind": "port", "ms_median": round(med * 1e3, 3),
            "single_core_ms_median": round(med1 * 1e3, 3),
            "single_core_value": round(2 * g.size / med1, 1),
            "sample": f"the same 512x512x256 pair: {runs} runs of the tiled oracle on {threads} "
                      f"threads over {tot:.1f} s and {runs1} runs on 1 thread over {tot1:.1f} s, "
                      f"medians after one warm-up each",
            "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "bit_exact_vs_gpu": bool(_same(rcw, res) and _same(rcw1, res)),
            "also_bit_exact": checks}


def leg_small(hdm, torch, steps):
    """64^3 sphere r 20, segm shifted (2,1,0) with 1% surface flips."""
    sp = hdm.synth_spec("sphere", 64, 64, 64, 32, 32, 32, 20, 2, 1, 0, 0.01, 1)
    G, S = hdm.synth_device(sp)
    torch.cuda.synchronize()
    for _ in range(3):
        res = hdm.hausdorff(G, S)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        res = hdm.hausdorff(G, S)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return G, S, res, {"workload": "sphere_64x64x64_r20_shift_2_1_0_flip1pct", "hd": res["hd"],
                       "avg_hd": res["avg_hd"], "ms_per_pair": round(dt * 1e3, 4),
                       "pairs_per_s": round(1 / dt, 1), "mask_voxels_per_s": round(2 * G.numel() / dt, 1)}


def leg_ct_cube(hdm, torch, steps, warmup):
    """The config-2 generator on 512^3 (16 ellipsoids): a short-HD 512^3 pair,
    timed like the headline (input copies cycled past the Infinity Cache,
    HIP events around the fused launch, phases from extra calls)."""
    G, S, res, dt, nvox, avg = leg_single(hdm, torch, steps, warmup, spec=CT_CUBE, copies=2)
    out = {"workload": "ct_pair_512x512x512_16_ellipsoids", "dims": list(CT_CUBE["dims"]),
           "seed": CT_CUBE["seed"], "hd": res["hd"], "avg_hd": res["avg_hd"],
           "iterations": res["iterations"], "queue_total": res["queue_total"],
           "ms_per_pair": round(dt / steps * 1e3, 4), "pairs_per_s": round(steps / dt, 2),
           "mask_voxels_per_s": round(2 * nvox * steps / dt, 1),
           "phase_ms": {k: round(v, 4) for k, v in avg.items()
                        if k not in ("fused", "steps_per_barrier", "launches")},
           "roofline": roofline(res, avg, nvox, traffic_name="ct512cube")}
    return G, S, res, out


def leg_long(hdm, torch, reps):
    """512^3: one ellipsoid (80,60,50) in both masks plus a 5^3 cube 300 voxels
    beyond it in segm only, about 300 iterations with a sparse queue."""
    sp = hdm.synth_spec("long", 512, 512, 512, 100, 256, 256, 80, 60, 50, 300, 5)
    G, S = hdm.synth_device(sp)
    torch.cuda.synchronize()
    nvox = G.numel()
    res = hdm.hausdorff(G, S, timing=3)
    torch.cuda.synchronize()
    hdm.timing_log()
    t0 = time.perf_counter()
    for _ in range(reps):
        res = hdm.hausdorff(G, S, timing=3)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    dom = hdm.timing_log()  # HIP events around the call's launches
    phases = {"pack": [], "meta": [], "dilate": [], "finalize": [], "total": []}
    for _ in range(reps):
        hdm.hausdorff(G, S, timing=True)
        for k, v in hdm.last_timings().items():
            if k in phases:
                phases[k].append(v)
    avg = {k: sum(v) / len(v) for k, v in phases.items()}
    hdm.hausdorff(G, S, timing=2)
    tr = hdm.last_trace()
    steps = hdm.last_timings()["steps"]  # dilation steps per grid barrier
    occ = [r["queued"] for r in tr]
    it_us = [r["next_start_us"] for r in tr if r["next_start_us"]]
    dil_bytes = float(res["queue_total"]) * entry_bytes(steps)
    dom_ms = sum(dom) / len(dom)
    return G, S, res, {
        "workload": "long_512x512x512_ellipsoid_80_60_50_cube5_gap300", "hd": res["hd"],
        "avg_hd": res["avg_hd"], "iterations": res["iterations"],
        "dilation_steps_per_barrier": steps,
        "ms_per_pair": round(dt * 1e3, 3), "mask_voxels_per_s": round(2 * nvox / dt, 1),
        "device_ms_per_pair": round(dom_ms, 4),
        "phase_ms": {k: round(v, 4) for k, v in avg.items()},
        "dilate_us_per_step": round(avg["dilate"] * 1e3 / max(1, res["iterations"]), 2),
        "queue_total": res["queue_total"], "queue_max": res["queue_max"],
        "queue_per_barrier_period": {"min": min(occ), "median": statistics.median(occ),
                                     "max": max(occ), "first8": occ[:8], "last8": occ[-8:]},
        "us_per_barrier_period_median_traced": round(statistics.median(it_us), 2) if it_us else None,
        "dilate_algo_bytes": int(dil_bytes),
        "dilate_gbs": round(dil_bytes / (avg["dilate"] * 1e-3) / 1e9, 1),
        # the call's launches (two-step launch: reset, pack, hand-off; region
        # launch: seed and dilation) as one: pack bytes plus the moved bytes
        # of every queued entry
        "roofline": roofline_obj("fused+regions", 2.25 * nvox + dil_bytes, dom_ms, "long512",
                                 dilation_steps_per_entry=steps)}


def leg_h5(hdm, torch, reps):
    """The reference's HDF5 benchmark path at its size (forBench/hdf5Manag.cu:
    196-220): two flat boolean datasets of 512x512x826 in one file, loaded by
    the library's reader, then HD with the reference engine's 6 element and
    its forBench engine's 18 element.  The reference's lung / bladder file is
    not shipped: here a union of 8 ellipsoids of that shape, written with the
    package's writer.  Load times are from the page cache (the file was just
    written); the HD call copies both volumes from host memory."""
    import tempfile
    from housdorffhdf5cmake_amd import h5
    nx, ny, nz = 512, 512, 826
    sp = hdm.synth_spec("ellipsoids", nx, ny, nz, 8, 10, 60, 2, 0.005, 826)
    G, S = hdm.synth_device(sp)
    torch.cuda.synchronize()
    path = os.path.join(tempfile.gettempdir(), f"hdm_bench_pair_{os.getpid()}.h5")
    h5.write(path, {"onlyBladderBoolFlat": G.cpu().numpy().reshape(-1).view(bool),
                    "onlyLungsBoolFlat": S.cpu().numpy().reshape(-1).view(bool)})
    out = {"workload": "hdf5_flat_bool_pair_512x512x826_8_ellipsoids",
           "file_bytes": os.path.getsize(path)}
    try:
        for conn in (6, 18):
            want = hdm.hausdorff(G, S, connectivity=conn)
            args = (path, "onlyBladderBoolFlat", path, "onlyLungsBoolFlat", nx, ny, nz)
            hdm.hausdorff_h5(*args, connectivity=conn)
            tot, load = [], []
            for _ in range(reps):
                t0 = time.perf_counter()
                r = hdm.hausdorff_h5(*args, connectivity=conn)
                tot.append(time.perf_counter() - t0)
                load.append(r["load_ms"])
            lm = statistics.median(load)
            out[f"conn{conn}"] = {
                "hd": r["hd"], "avg_hd": r["avg_hd"],
                "ms_load_and_hd_median": round(statistics.median(tot) * 1e3, 2),
                "ms_load_median": round(lm, 2),
                "load_gbs": round(2 * nx * ny * nz / (lm * 1e-3) / 1e9, 2),
                "identical_to_in_memory": all(r[k] == want[k] for k in EXACT_KEYS)}
    finally:
        os.remove(path)
    return out


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
    near_entries = b.near_entries()
    b.close()
    keys = ("hd", "sum", "count", "sum_fp", "sum_fn", "n_fp", "n_fn")
    same = all(all(o[k] == p[k] for k in keys) for o, p in zip(out, plain))
    # the batch's launches (reset, pack, totals, near pass, seed, dilation) per
    # replay: pack bytes of every pair plus the moved bytes of every near-pass
    # entry and every queued entry (queue_total is the batch's, reported with
    # each pair)
    moved_tiles = float(out[0]["queue_total"]) * TILE_BYTES[best["tile_id"]]
    moved = moved_tiles + float(near_entries) * NEAR_BYTES
    dev_ms = statistics.median(ms)
    return {"pairs": n_pairs, "mask_voxels": 2 * nvox,
            "plain_pairs_per_s": round(n_pairs / t_plain, 1),
            "batched_pairs_per_s": round(n_pairs / t_batch, 1),
            "speedup": round(t_plain / t_batch, 2),
            "batched_mask_voxels_per_s": round(2 * nvox / t_batch, 1),
            "batched_device_ms_median": round(statistics.median(ms), 3),
            "queue_total": int(out[0]["queue_total"]), "queue_max": int(out[0]["queue_max"]),
            "near_entries": near_entries,
            "tile": best["tile"], "chunk": best["chunk"], "graph": True,
            "identical_to_plain": bool(same),
            "roofline": roofline_obj("batch graph replay", 2.25 * nvox + moved, dev_ms,
                                     "batch256", dilation_algo_bytes=int(moved),
                                     pack_algo_bytes=int(2.25 * nvox),
                                     # 64x8x8 tiles run two steps per entry
                                     algo_bytes_counting_each_step=int(
                                         2.25 * nvox + moved - moved_tiles + moved_tiles * (
                                             2 if best["tile_id"] == hdm.TILE_64x8x8 else 1))),
            "autotune": [{k: e[k] for k in ("tile", "chunk", "ms", "status")}
                         for e in (table or [])][:12]}


def leg_serving(hdm, torch, seconds, rate, budget_gb):
    """Returns the summary and the pool pairs with their results computed
    alone (every served result equalled those; the CPU leg checks them)."""
    from housdorffhdf5cmake_amd import serving
    return serving.poisson_run(seconds=seconds, rate=rate, budget_bytes=int(budget_gb * (1 << 30)),
                               verify=None, return_pool=True)


def leg_serving_loaded(hdm, torch, seconds, budget_gb):
    """The serving engine under load: config 5's volume mix with its packed
    masks resident (each pool volume registered once) at 200 requests/s (config
    5's rate), at 5000 requests/s, and in a closed loop keeping 64 requests in
    the server (the rate it sustains); the closed loop also with volumes
    packed at every admission.  Every result is compared with the same pair
    computed alone."""
    from housdorffhdf5cmake_amd import serving
    b = int(budget_gb * (1 << 30))
    return {
        "resident_poisson_200": serving.poisson_run(seconds=seconds, rate=200, budget_bytes=b,
                                                    resident=True),
        "resident_poisson_5000": serving.poisson_run(seconds=seconds, rate=5000, budget_bytes=b,
                                                     resident=True),
        "resident_closed_loop_64": serving.closed_loop_run(seconds=seconds, outstanding=64,
                                                           budget_bytes=b, resident=True),
        "volumes_closed_loop_64": serving.closed_loop_run(seconds=seconds, outstanding=64,
                                                          budget_bytes=b, resident=False),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch-pairs", type=int, default=256)
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-loaded-seconds", type=float, default=5.0,
                    help="seconds of each loaded serving leg (0 skips them)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the 64^3 and the long-iteration 512^3 lines")
    ap.add_argument("--no-small", action="store_true", help="skip the 64^3 line")
    ap.add_argument("--no-h5", action="store_true", help="skip the HDF5 512x512x826 line")
    ap.add_argument("--long-reps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="size of the device-to-device copy timed for the attainable "
                         "bandwidth (default 1024); a run's copies all have this size, the "
                         "known-size kernel that scales HBM byte counters")
    ap.add_argument("--no-cube", action="store_true", help="skip the short-HD 512^3 line")
    ap.add_argument("--only", choices=("small", "long", "cube", "batched", "serving_loaded"),
                    help="run this one leg (and the copy probe) and print its object alone: "
                         "a profile of one workload")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine is single-device (a pair or a batch per "
              "GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    import housdorffhdf5cmake_amd as hdm
    if args.only:
        one = {"only": args.only}
        if args.only == "small":
            one["small_64"] = leg_small(hdm, torch, 200)[3]
        elif args.only == "long":
            one["long_512"] = leg_long(hdm, torch, args.long_reps)[3]
        elif args.only == "cube":
            one["ct_512cube"] = leg_ct_cube(hdm, torch, args.steps, args.warmup)[3]
        elif args.only == "serving_loaded":
            one["serving_loaded"] = leg_serving_loaded(hdm, torch, args.serve_loaded_seconds, 64.0)
        else:
            one["batched"] = leg_batched(hdm, torch, args.batch_pairs, args.batch_reps,
                                         not args.no_tune)
        one["attainable_copy_gbs"] = copy_bandwidth(
            torch, mib=args.copy_probe if args.copy_probe > 0 else 1024)
        print(json.dumps(one), flush=True)
        return
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
        "phase_ms": {k: round(v, 4) for k, v in avg.items()
                     if k not in ("fused", "steps_per_barrier", "launches")},
        "fused_launch": bool(avg["fused"]),
        "dilation_steps_per_barrier": int(avg["steps_per_barrier"]),
    }
    line["roofline"] = roofline(res, avg, nvox)
    # every copy of the run has this size, so it is also the known-size
    # kernel that scales the HBM byte counters
    copy_gbs = copy_bandwidth(torch, mib=args.copy_probe if args.copy_probe > 0 else 1024)
    line["roofline"]["attainable_copy_gbs"] = copy_gbs
    line["roofline"]["frac_of_attainable"] = round(line["roofline"]["achieved"] / copy_gbs, 4)
    extra = []
    if not args.no_extra and not args.no_small:
        log("small 64^3 ...")
        g1, s1, r1, line["small_64"] = leg_small(hdm, torch, 200)
        extra.append(("small_64", g1, s1, r1))
    if not args.no_extra and not args.no_cube:
        log("short-HD 512^3 ...")
        g4, s4, r4, line["ct_512cube"] = leg_ct_cube(hdm, torch, args.steps, args.warmup)
        line["ct_512cube"]["roofline"]["attainable_copy_gbs"] = copy_gbs
        extra.append(("ct_512cube", g4, s4, r4))
    if not args.no_extra:
        log("long-iteration 512^3 ...")
        g3, s3, r3, line["long_512"] = leg_long(hdm, torch, args.long_reps)
        extra.append(("long_512", g3, s3, r3))
    if not args.no_extra and not args.no_h5:
        log("HDF5 pair 512x512x826 ...")
        line["hdf5_512x512x826"] = leg_h5(hdm, torch, 3)
    if not args.no_batch:
        log("batched vs plain ...")
        line["batched"] = leg_batched(hdm, torch, args.batch_pairs, args.batch_reps,
                                      not args.no_tune)
    if not args.no_serve:
        log("serving ...")
        line["serving"], pool = leg_serving(hdm, torch, args.serve_seconds, args.serve_rate,
                                            64.0)
        extra += [("serving_pool", g, s, a) for g, s, a in pool]
        if args.serve_loaded_seconds > 0:
            log("serving under load ...")
            line["serving_loaded"] = leg_serving_loaded(hdm, torch, args.serve_loaded_seconds, 64.0)
    log("cpu baseline ...")
    line["cpu_baseline"] = leg_cpu(hdm, G, S, res, args.cpu_seconds, extra)
    if "serving" in line:
        line["serving"]["bit_exact_vs_cpu"] = line["cpu_baseline"]["also_bit_exact"].get(
            "serving_pool")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
