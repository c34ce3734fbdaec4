"""The comparison switches: HDM_DS=0 (one dilation step per grid barrier) and
HDM_FUSED=0 (a single pair as four kernels) give the same results as the
default fused launch (the region launch), bit for bit, for single pairs, batches and
robust cuts; HDM_SERVER_LONE=0 (a lone server request through a slot) gives
the same results as the lone one-pair launch.  The switches are read once per
process (or server), so the switched runs happen in a child process on the
same GPU."""
from __future__ import annotations

import json
import os
import subprocess
import sys

import pytest

import housdorffhdf5cmake_amd as hdm

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KEYS = ("hd", "sum", "count", "sum_fp", "cnt_fp", "sum_fn", "cnt_fn", "n_fp", "n_fn", "hd_fp",
        "hd_fn", "iterations")
SPECS = [("ellipsoids", 160, 128, 96, 8, 6, 24, 2, 0.005, 41),
         ("ellipsoids", 67, 45, 31, 6, 4, 20, 2, 0.01, 7),
         ("long", 128, 96, 128, 30, 64, 48, 20, 15, 12, 60, 4)]

CHILD = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
import torch
import housdorffhdf5cmake_amd as hdm
specs = json.loads(sys.argv[2])
keys = json.loads(sys.argv[3])
vols = [hdm.synth_device(hdm.synth_spec(s[0], *s[1:])) for s in specs]
torch.cuda.synchronize()
out = {"single": [], "robust": [], "batch": []}
for g, s in vols:
    out["single"].append({k: hdm.hausdorff(g, s)[k] for k in keys})
    out["robust"].append({k: hdm.hausdorff(g, s, robustness=0.9)[k] for k in keys})
b = hdm.Batch(vols, graph=False)
b.run()
out["batch"] = [{k: r[k] for k in keys if k != "iterations"} for r in b.sync()]
b.close()
print(json.dumps(out))
"""


def run_here():
    vols = [hdm.synth_device(hdm.synth_spec(s[0], *s[1:])) for s in SPECS]
    torch.cuda.synchronize()
    out = {"single": [], "robust": [], "batch": []}
    for g, s in vols:
        out["single"].append({k: hdm.hausdorff(g, s)[k] for k in KEYS})
        out["robust"].append({k: hdm.hausdorff(g, s, robustness=0.9)[k] for k in KEYS})
    b = hdm.Batch(vols, graph=False)
    b.run()
    out["batch"] = [{k: r[k] for k in KEYS if k != "iterations"} for r in b.sync()]
    b.close()
    return out


@pytest.mark.parametrize("switch,value", [("HDM_DS", "0"), ("HDM_FUSED", "0"), ("HDM_COOP", "1"),
                                          ("HDM_TB", "0"), ("HDM_TB", "2"), ("HDM_NEAR", "0"),
                                          ("HDM_PACK_POOLS", "0"), ("HDM_SKIP_ROWS", "0")])
def test_switched_off_paths_give_the_same_results(switch, value):
    """HDM_COOP=1 (the fused launch as a cooperative launch), HDM_TB=0 (the
    two-step fused kernel only: no near pass, no region launch), HDM_TB=2 (the
    region body for every pair, right after the pack), HDM_NEAR=0 (the
    two-step fused kernel first, handing pairs whose HD bound reaches 16 steps
    to the region launch) and HDM_PACK_POOLS=0 (the fused pack in even static
    shares, without the per-XCD counters for its last quarter) and
    HDM_SKIP_ROWS=0 (a batch writes every row in its pack, where by default it
    skips the rows of masks empty in a tile and zero-fills them afterwards for
    the pairs its near pass leaves -- here the long pair) are compared the
    same way."""
    here = run_here()
    env = dict(os.environ, **{switch: value})
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT, json.dumps(SPECS), json.dumps(KEYS)],
                       capture_output=True, text=True, timeout=240, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    there = json.loads(r.stdout.strip().splitlines()[-1])
    assert there == here


CHILD_TIMEOUT = r"""
import json, os, sys
sys.path.insert(0, sys.argv[1])
import torch
import housdorffhdf5cmake_amd as hdm
keys = json.loads(sys.argv[2])
g, s = hdm.synth_device(hdm.synth_spec("ellipsoids", 160, 128, 96, 8, 6, 24, 2, 0.005, 41))
# a pair beyond the near pass's eight steps, so the batch runs the barriers too
lg, ls = hdm.synth_device(hdm.synth_spec("long", 128, 96, 128, 30, 64, 48, 20, 15, 12, 60, 4))
torch.cuda.synchronize()
def pick(r):
    return {k: r[k] for k in keys}
def batch_run():
    b = hdm.Batch([(g, s), (s, g), (lg, ls)], graph=False)
    b.run()
    try:
        return [pick(r) for r in b.sync()]
    except hdm.HDMError as e:
        return e.code
    finally:
        b.close()
out = {"before": pick(hdm.hausdorff(g, s)), "batch_before": batch_run()}
os.environ["HDM_SPIN_TICKS"] = "0"  # every barrier wait expires at once
out["timeout"] = hdm.hausdorff(g, s, raise_on_error=False)["status"]
out["batch_timeout"] = batch_run()
del os.environ["HDM_SPIN_TICKS"]
out["after"] = pick(hdm.hausdorff(g, s))
out["batch_after"] = batch_run()
print(json.dumps(out))
"""


def test_barrier_timeout_is_reported_and_the_next_calls_are_exact():
    """With the barrier wait bound forced to zero (HDM_SPIN_TICKS=0), a single
    pair and a batch end with HDM_E_TIMEOUT (-5) instead of waiting, every
    workgroup leaving its loop; the calls after it, with the normal bound,
    return exactly the results from before (engine state, epoch tags and the
    alternating pack partials recover)."""
    keys = [k for k in KEYS if k != "iterations"]
    r = subprocess.run([sys.executable, "-c", CHILD_TIMEOUT, ROOT, json.dumps(keys)],
                       capture_output=True, text=True, timeout=240,
                       env={k: v for k, v in os.environ.items() if k != "HDM_SPIN_TICKS"})
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["timeout"] == -5
    assert out["batch_timeout"] == -5
    assert out["after"] == out["before"]
    assert out["batch_after"] == out["batch_before"]


CHILD_SERVE = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
import torch
import housdorffhdf5cmake_amd as hdm
from housdorffhdf5cmake_amd import serving
specs = json.loads(sys.argv[2])
keys = json.loads(sys.argv[3])
vols = [hdm.synth_device(hdm.synth_spec(s[0], *s[1:])) for s in specs]
torch.cuda.synchronize()
srv = serving.Server(hbm_budget=1 << 30, max_slots=4, max_dims=(160, 128, 128))
out = []
for g, s in vols:
    rid = srv.submit(g, s)
    done = srv.drain(60)
    out.append([done[0]["id"] == rid, len(done), done[0]["status"]] + [done[0][k] for k in keys])
srv.close()
print(json.dumps(out))
"""


def test_server_lone_switch_gives_the_same_results():
    """Requests submitted one at a time, each finding the server idle: by
    default each runs as a one-pair launch; with HDM_SERVER_LONE=0 each goes
    through a slot and the several-pair kernel.  Both equal the pair computed
    alone."""
    keys = [k for k in KEYS if k != "iterations"]
    vols = [hdm.synth_device(hdm.synth_spec(s[0], *s[1:])) for s in SPECS]
    torch.cuda.synchronize()
    want = [[True, 1, 0] + [hdm.hausdorff(g, s)[k] for k in keys] for g, s in vols]
    for lone in ("1", "0"):
        env = dict(os.environ, HDM_SERVER_LONE=lone)
        r = subprocess.run([sys.executable, "-c", CHILD_SERVE, ROOT, json.dumps(SPECS),
                            json.dumps(keys)], capture_output=True, text=True, timeout=240, env=env)
        assert r.returncode == 0, r.stderr[-2000:]
        assert json.loads(r.stdout.strip().splitlines()[-1]) == want, lone


CHILD_SERVER_TIMEOUT = r"""
import json, os, sys
sys.path.insert(0, sys.argv[1])
import torch
import housdorffhdf5cmake_amd as hdm
from housdorffhdf5cmake_amd import serving
keys = json.loads(sys.argv[2])
# pairs with HD > 8: the admission group's near pass hands them to the slots
lg, ls = hdm.synth_device(hdm.synth_spec("long", 128, 96, 128, 30, 64, 48, 20, 15, 12, 60, 4))
vols = [(lg, ls), (ls, lg)]
torch.cuda.synchronize()
alone = [{k: hdm.hausdorff(g, s)[k] for k in keys} for g, s in vols]
srv = serving.Server(hbm_budget=1 << 30, max_slots=4, max_dims=(160, 128, 128))
def round_trip():
    ids = [srv.submit(g, s) for g, s in vols]  # two at once: the slot path
    done = srv.drain(60)
    by = {d["id"]: d for d in done}
    return [[by[i]["status"]] + [by[i][k] for k in keys] for i in ids]
os.environ["HDM_SPIN_TICKS"] = "0"  # every barrier wait of the slice expires at once
failed = round_trip()
del os.environ["HDM_SPIN_TICKS"]
after = [round_trip() for _ in range(2)]
srv.close()
print(json.dumps({"failed": [f[0] for f in failed], "after": after,
                  "want": [[0] + [a[k] for k in keys] for a in alone]}))
"""


def test_server_recovers_after_a_barrier_timeout():
    """A slice whose grid-barrier wait expires (HDM_SPIN_TICKS=0, read per
    slice) fails every request in flight with HDM_E_TIMEOUT (-5); the engine
    then restarts (iteration 0, empty queue, error flag cleared), so the
    requests submitted after it, with the normal bound, give exactly the
    results of the pairs computed alone."""
    keys = [k for k in KEYS if k != "iterations"]
    r = subprocess.run([sys.executable, "-c", CHILD_SERVER_TIMEOUT, ROOT, json.dumps(keys)],
                       capture_output=True, text=True, timeout=240,
                       env={k: v for k, v in os.environ.items() if k != "HDM_SPIN_TICKS"})
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["failed"] == [-5, -5]
    assert out["after"] == [out["want"], out["want"]]


def test_lone_request_then_server_destroyed_then_single_call():
    """A lone request runs as a one-pair call on the server's stream; after
    the server (and that stream) is destroyed, a one-pair call on torch's
    stream succeeds and is exact (the library no longer waits for the
    destroyed stream)."""
    from housdorffhdf5cmake_amd import serving
    g, s = hdm.synth_device(hdm.synth_spec("ellipsoids", 96, 80, 64, 6, 4, 20, 2, 0.01, 5))
    torch.cuda.synchronize()
    want = hdm.hausdorff(g, s)
    for _ in range(3):
        srv = serving.Server(hbm_budget=1 << 30, max_slots=4, max_dims=(128, 128, 128))
        rid = srv.submit(g, s)
        done = srv.drain(60)
        srv.close()
        assert len(done) == 1 and done[0]["id"] == rid and done[0]["status"] == 0
        got = hdm.hausdorff(g, s)  # on torch's current stream
        assert all(got[k] == want[k] for k in KEYS)
