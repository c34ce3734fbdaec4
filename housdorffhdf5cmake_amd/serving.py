"""Serving loop over the engine's continuous batching (hdm_server_*).

Server wraps the C ABI; poisson_run drives it with Poisson arrivals of
requests drawn from a pool of device-resident mask pairs (volumes 64^3,
128^3, 256x256x128, 384x384x256 weighted 40/30/20/10 %), admits them against
an HBM budget, and reports p50 / p99 latency (scheduled arrival to result on
the host), pairs/s, and whether every result equals the same pair computed
alone (and, with check_exact, the CPU path).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import (DTYPE_I32, DTYPE_U8, PTR_DEVICE, PTR_HOST, TILE_64x8x8, Completion, HDMError,
               ServerOpts, ServerStats, _volume_args, check, hausdorff, lib, synth_device,
               synth_spec)

VOLUMES = [(64, 64, 64), (128, 128, 128), (256, 256, 128), (384, 384, 256)]
WEIGHTS = [0.4, 0.3, 0.2, 0.1]
RESULT_KEYS = ("hd", "sum", "count", "sum_fp", "cnt_fp", "sum_fn", "cnt_fn", "n_fp", "n_fn")


class Server:
    def __init__(self, hbm_budget=64 << 30, max_slots=64, slice_iters=8, tile=TILE_64x8x8,
                 grid_ctas=0, robustness=1.0, max_dims=(512, 512, 512), connectivity=6):
        o = ServerOpts()
        lib().hdm_server_opts_init(C.byref(o))
        o.hbm_budget = int(hbm_budget)
        o.max_slots = int(max_slots)
        o.slice_iters = int(slice_iters)
        o.tile = int(tile)
        o.grid_ctas = int(grid_ctas)
        o.robustness = float(robustness)
        o.max_nx, o.max_ny, o.max_nz = (int(v) for v in max_dims)
        o.connectivity = int(connectivity)
        self.h = C.c_void_p()
        check(lib().hdm_server_create(C.byref(o), C.byref(self.h)), "hdm_server_create")
        self.max_slots = max_slots
        self._keep = {}
        self._done = (Completion * 256)()

    def submit(self, gold, segm, label=1) -> int:
        """Queue a pair; device tensors must be complete (synchronize after
        writing them): the server's stream does not wait for torch's."""
        gp, sp, nx, ny, nz, dt, pk, keep = _volume_args(gold, segm)
        rid = C.c_uint64()
        check(lib().hdm_server_submit(self.h, gp, sp, nx, ny, nz, dt, label, pk, C.byref(rid)),
              "hdm_server_submit")
        self._keep[rid.value] = keep
        return rid.value

    def register(self, vol, label=1) -> int:
        """Pack one volume (the voxels equal to label) once and keep it on the
        device; returns a handle for submit_masks.  Device tensors must be
        complete; the volume may be freed when this returns."""
        vp, _, nx, ny, nz, dt, pk, keep = _volume_args(vol, vol)
        h = C.c_uint64()
        check(lib().hdm_server_register(self.h, vp, nx, ny, nz, dt, label, pk, C.byref(h)),
              "hdm_server_register")
        del keep
        return h.value

    def submit_masks(self, gold_handle, segm_handle) -> int:
        """Queue a pair of registered masks (no pack at admission)."""
        rid = C.c_uint64()
        check(lib().hdm_server_submit_masks(self.h, int(gold_handle), int(segm_handle),
                                            C.byref(rid)), "hdm_server_submit_masks")
        return rid.value

    def unregister(self, handle):
        check(lib().hdm_server_unregister(self.h, int(handle)), "hdm_server_unregister")

    def stats(self) -> dict:
        """Slices run, pairs in flight summed over them (batch_mean = the
        ratio), the most in one slice, lone one-pair runs, admissions through
        the slots, the arena bytes of registered masks, and (64x8x8 tiles, 6
        element) the admission groups, their requests (group_mean = the
        ratio) and the requests their near pass finished."""
        st = ServerStats()
        check(lib().hdm_server_stats_get(self.h, C.byref(st)), "hdm_server_stats_get")
        d = st.as_dict()
        d["batch_mean"] = round(d["pair_slices"] / d["slices"], 2) if d["slices"] else 0.0
        d["group_mean"] = round(d["group_pairs"] / d["groups"], 2) if d["groups"] else 0.0
        return d

    def step(self):
        n = C.c_int()
        check(lib().hdm_server_step(self.h, self._done, len(self._done), C.byref(n)),
              "hdm_server_step")
        out = []
        for c in self._done[:n.value]:
            d = c.result.as_dict()
            d.update(id=c.id, status=c.status, t_submit=c.t_submit, t_admit=c.t_admit,
                     t_done=c.t_done)
            self._keep.pop(c.id, None)
            out.append(d)
        return out

    def load(self):
        q, f, b = C.c_int(), C.c_int(), C.c_uint64()
        check(lib().hdm_server_load(self.h, C.byref(q), C.byref(f), C.byref(b)))
        return q.value, f.value, b.value

    def drain(self, timeout_s=60.0):
        out = []
        t_end = time.perf_counter() + timeout_s
        while time.perf_counter() < t_end:
            out += self.step()
            q, f, _ = self.load()
            if q == 0 and f == 0:
                break
        return out

    def close(self):
        if self.h:
            lib().hdm_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def now() -> float:
    return float(lib().hdm_server_now())


def make_pool(per_class=4, seed=500):
    """Device-resident request pool: per_class seeded ellipsoid-union pairs per
    volume class; returns [(class, gold, segm)]."""
    pool = []
    for k, (nx, ny, nz) in enumerate(VOLUMES):
        for j in range(per_class):
            sp = synth_spec("ellipsoids", nx, ny, nz, 8, 10, 60, 2, 0.005, seed + 97 * k + j)
            g, s = synth_device(sp)
            pool.append((k, g, s))
    return pool


def poisson_run(seconds=10.0, rate=200.0, budget_bytes=64 << 30, verify=None,
                per_class=4, max_slots=64, slice_iters=8, seed=11, return_pool=False,
                grid_ctas=0, resident=False):
    """verify (optional): callable(gold, segm, result) -> bool, an independent
    check of each pool pair's result computed alone (the caller's own check).
    resident: every pool volume is registered once (hdm_server_register) and
    requests name the packed masks, so no admission re-packs a volume.
    With return_pool, returns (summary, [(gold, segm, result alone)])."""
    import torch
    pool = make_pool(per_class)
    torch.cuda.synchronize()
    # the same pair computed alone, and optionally checked by the caller
    alone = [hausdorff(g, s) for _, g, s in pool]
    exact_cpu = None
    if verify is not None:
        exact_cpu = all(bool(verify(g, s, a)) for (_, g, s), a in zip(pool, alone))
    rng = np.random.default_rng(seed)
    n = max(1, int(rng.poisson(rate * seconds)))
    arrivals = np.sort(rng.uniform(0.0, seconds, size=n))
    classes = rng.choice(len(VOLUMES), size=n, p=WEIGHTS)
    members = rng.integers(0, per_class, size=n)
    picks = [int(c) * per_class + int(m) for c, m in zip(classes, members)]
    srv = Server(hbm_budget=budget_bytes, max_slots=max_slots, slice_iters=slice_iters,
                 max_dims=max(VOLUMES), grid_ctas=grid_ctas)
    handles = None
    if resident:  # each pool volume packed once, before the clock starts
        handles = [(srv.register(g), srv.register(s)) for _, g, s in pool]
    lat, waits, ok = [], [], True
    id2req = {}
    done_n, nvox_done, i = 0, 0, 0
    t0 = now()
    t_last = t0
    while True:
        t = now() - t0
        while i < n and arrivals[i] <= t:
            if handles:
                id2req[srv.submit_masks(*handles[picks[i]])] = i
            else:
                _, g, s = pool[picks[i]]
                id2req[srv.submit(g, s)] = i
            i += 1
        for c in srv.step():
            r = id2req.pop(c["id"])
            a = alone[picks[r]]
            ok &= c["status"] == 0 and all(c[k] == a[k] for k in RESULT_KEYS)
            lat.append(c["t_done"] - (t0 + arrivals[r]))
            waits.append(c["t_admit"] - c["t_submit"])
            nvox_done += 2 * int(pool[picks[r]][1].numel())
            done_n += 1
            t_last = c["t_done"]
        q, f, _ = srv.load()
        if i >= n and q == 0 and f == 0:
            break
        if q == 0 and f == 0 and i < n:  # idle until the next arrival:
            # sleep to 0.2 ms before it, then spin (a sleep overshoots by the
            # kernel's timer slack, which would count in the latency)
            gap = arrivals[i] - (now() - t0)
            if gap > 3e-4:
                time.sleep(min(0.001, gap - 2e-4))
    st = srv.stats()
    srv.close()
    span = max(t_last - t0, 1e-9)
    lat_ms = np.array(lat) * 1e3
    summary = {"requests": n, "completed": done_n, "seconds": seconds, "rate_per_s": rate,
               "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
               "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
               "mean_ms": round(float(lat_ms.mean()), 3),
               "admit_wait_p99_ms": round(float(np.percentile(np.array(waits) * 1e3, 99)), 3),
               "pairs_per_s": round(done_n / span, 1),
               "mask_voxels_per_s": round(nvox_done / span, 1),
               # pairs in flight per slice of the several-pair engine, and the
               # requests that ran as one-pair calls on an idle server
               "batch_mean": st["batch_mean"], "max_in_flight": st["max_in_flight"],
               "slices": st["slices"], "lone_runs": st["lone"],
               "admitted_through_slots": st["admitted"],
               # requests per admission group, and those its near pass finished
               "group_mean": st["group_mean"], "near_finished": st["near_finished"],
               "max_slots": max_slots, "hbm_budget_gb": budget_bytes / (1 << 30),
               "identical_alone_vs_mix": bool(ok), "bit_exact_vs_cpu": exact_cpu,
               "pool": f"{per_class} pairs per volume class, device resident" + (
                   ", each volume registered (packed once)" if resident else "")}
    if return_pool:
        return summary, [(g, s, a) for (_, g, s), a in zip(pool, alone)]
    return summary


def closed_loop_run(seconds=5.0, outstanding=64, budget_bytes=64 << 30, per_class=4,
                    max_slots=64, slice_iters=8, seed=13, resident=False, grid_ctas=0):
    """Saturating load: `outstanding` requests are kept in the server at all
    times (a new one, drawn from the config-5 mix, is submitted as each one
    completes), for `seconds`.  Reports the pairs/s the engine sustains, the
    latency from submission to result, pairs in flight per slice, and whether
    every result equals the same pair computed alone."""
    import torch
    pool = make_pool(per_class)
    torch.cuda.synchronize()
    alone = [hausdorff(g, s) for _, g, s in pool]
    rng = np.random.default_rng(seed)
    srv = Server(hbm_budget=budget_bytes, max_slots=max_slots, slice_iters=slice_iters,
                 max_dims=max(VOLUMES), grid_ctas=grid_ctas)
    handles = [(srv.register(g), srv.register(s)) for _, g, s in pool] if resident else None

    def submit():
        k = int(rng.choice(len(VOLUMES), p=WEIGHTS)) * per_class + int(rng.integers(0, per_class))
        rid = srv.submit_masks(*handles[k]) if handles else srv.submit(pool[k][1], pool[k][2])
        return rid, k

    id2k = dict(submit() for _ in range(outstanding))
    lat, ok, done_n, nvox_done = [], True, 0, 0
    t0 = now()
    t_end = t0 + seconds
    while id2k:
        for c in srv.step():
            k = id2k.pop(c["id"])
            a = alone[k]
            ok &= c["status"] == 0 and all(c[x] == a[x] for x in RESULT_KEYS)
            lat.append(c["t_done"] - c["t_submit"])
            nvox_done += 2 * int(pool[k][1].numel())
            done_n += 1
            if c["t_done"] < t_end:
                rid, kk = submit()
                id2k[rid] = kk
    span = now() - t0
    st = srv.stats()
    srv.close()
    lat_ms = np.array(lat) * 1e3
    return {"mode": "closed loop", "outstanding": outstanding, "seconds": round(span, 3),
            "completed": done_n, "pairs_per_s": round(done_n / span, 1),
            "mask_voxels_per_s": round(nvox_done / span, 1),
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
            "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
            "batch_mean": st["batch_mean"], "max_in_flight": st["max_in_flight"],
            "slices": st["slices"], "lone_runs": st["lone"],
            "group_mean": st["group_mean"], "near_finished": st["near_finished"],
            "max_slots": max_slots, "slice_iters": slice_iters,
            "identical_alone_vs_mix": bool(ok),
            "pool": f"{per_class} pairs per volume class, device resident" + (
                ", each volume registered (packed once)" if resident else "")}
