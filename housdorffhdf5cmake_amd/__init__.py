"""MI355X Hausdorff-distance engine: Python host mirror of the C ABI (libhdm.so).

Volumes are 3-D arrays indexed [z, y, x] (x fastest, the reference's layout
x + y*Nx + z*Nx*Ny, hello.cu:453).  torch tensors on the GPU are passed as
device pointers on torch's current stream; numpy arrays are copied by the
library.  The compute always runs in the HIP library; importing this package
without libhdm.so built raises.

Reference-shaped entry point (hello.cu:2025-2027):
    getHausdorffDistance_CUDA_Generic(gold, segm, WIDTH, HEIGHT, DEPTH,
        robustnessPercent, resIterNeeded, numberToLookFor, res3DNeeded) -> int
"""
from __future__ import annotations

import ctypes as C
import struct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhdm.so")

HDM_OK = 0
ERRORS = {-1: "HDM_E_INVALID", -2: "HDM_E_EMPTY_MASK", -3: "HDM_E_HIP", -4: "HDM_E_NOMEM",
          -5: "HDM_E_TIMEOUT", -6: "HDM_E_BUDGET", -7: "HDM_E_NODEVICE", -8: "HDM_E_STUCK",
          -9: "HDM_E_CAPACITY", -10: "HDM_E_IO", -11: "HDM_E_UNSUPPORTED", -12: "HDM_E_FORMAT",
          -13: "HDM_E_NOTFOUND"}
PTR_DEVICE, PTR_HOST = 0, 1
DTYPE_U8, DTYPE_I32 = 0, 1
TILE_AUTO, TILE_64x8x8, TILE_64x4x4, TILE_64x8x4, TILE_32x8x8 = 0, 1, 2, 3, 4
TILE_NAMES = {TILE_64x8x8: "64x8x8", TILE_64x4x4: "64x4x4", TILE_64x8x4: "64x8x4",
              TILE_32x8x8: "32x8x8"}


class HDMError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)} ({code}) {what}".strip())


class Opts(C.Structure):
    _fields_ = [("robustness", C.c_double), ("ptr_kind", C.c_int32), ("tile", C.c_int32),
                ("grid_ctas", C.c_int32), ("chunk", C.c_int32), ("crop", C.c_int32),
                ("timing", C.c_int32), ("stream", C.c_void_p), ("map_out", C.c_void_p),
                ("hist_out", C.c_void_p), ("hist_len", C.c_int32), ("connectivity", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("hd", C.c_int32), ("hd_fp", C.c_int32), ("hd_fn", C.c_int32),
                ("iterations", C.c_int32), ("sum", C.c_uint64), ("count", C.c_uint64),
                ("sum_fp", C.c_uint64), ("cnt_fp", C.c_uint64), ("sum_fn", C.c_uint64),
                ("cnt_fn", C.c_uint64), ("n_gold", C.c_uint64), ("n_segm", C.c_uint64),
                ("n_fp", C.c_uint64), ("n_fn", C.c_uint64), ("avg_hd", C.c_double),
                ("queue_total", C.c_uint64), ("queue_max", C.c_uint32), ("tiles", C.c_uint32)]

    def as_dict(self) -> dict:
        return dict(zip(_RESULT_NAMES, _RESULT_LAYOUT.unpack_from(self)))


# hdm_result as one unpack: 4 int32, 10 uint64, a double, a uint64, 2 uint32
_RESULT_NAMES = tuple(k for k, _ in Result._fields_)
_RESULT_LAYOUT = struct.Struct("<4i10QdQ2I")
assert _RESULT_LAYOUT.size == C.sizeof(Result)


class Pair(C.Structure):
    _fields_ = [("gold", C.c_void_p), ("segm", C.c_void_p), ("nx", C.c_int32), ("ny", C.c_int32),
                ("nz", C.c_int32), ("dtype", C.c_int32), ("label", C.c_int32),
                ("reserved", C.c_int32)]


class TuneEntry(C.Structure):
    _fields_ = [("tile", C.c_int32), ("chunk", C.c_int32), ("grid_ctas", C.c_int32),
                ("status", C.c_int32), ("ms", C.c_float)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_slots", C.c_int32), ("slice_iters", C.c_int32),
                ("tile", C.c_int32), ("grid_ctas", C.c_int32), ("robustness", C.c_double),
                ("max_nx", C.c_int32), ("max_ny", C.c_int32), ("max_nz", C.c_int32),
                ("connectivity", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("slices", C.c_uint64), ("pair_slices", C.c_uint64), ("max_in_flight", C.c_int32),
                ("reserved", C.c_int32), ("lone", C.c_uint64), ("admitted", C.c_uint64),
                ("registered_bytes", C.c_uint64), ("near_finished", C.c_uint64),
                ("groups", C.c_uint64), ("group_pairs", C.c_uint64)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_ if k != "reserved"}


class H5Info(C.Structure):
    _fields_ = [("rank", C.c_int32), ("type_class", C.c_int32), ("elem_size", C.c_int32),
                ("layout", C.c_int32), ("dims", C.c_uint64 * 8), ("n_elems", C.c_uint64),
                ("chunk", C.c_uint64 * 8), ("n_filters", C.c_int32), ("big_endian", C.c_int32)]

    def as_dict(self) -> dict:
        r = self.rank
        return {"shape": tuple(int(v) for v in self.dims[:r]),
                "chunks": tuple(int(v) for v in self.chunk[:r]) if self.layout == 2 else None,
                "type_class": ("int", "uint", "float", "enum")[self.type_class],
                "elem_size": self.elem_size, "layout": ("compact", "contiguous", "chunked")[self.layout],
                "n_filters": self.n_filters, "big_endian": bool(self.big_endian)}


class Completion(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int32), ("reserved", C.c_int32),
                ("result", Result), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double)]


class Ell(C.Structure):
    _fields_ = [("c", C.c_int32 * 3), ("s", C.c_int32 * 3), ("r", C.c_int32 * 9)]


class SynthSpec(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("n_gold", C.c_int32),
                ("n_segm", C.c_int32), ("flip_mode", C.c_int32), ("flip_seed", C.c_uint64),
                ("flip_thresh", C.c_uint64), ("cube_lo", C.c_int32 * 3),
                ("cube_hi", C.c_int32 * 3), ("gold", Ell * 16), ("segm", Ell * 16)]


_lib = None


def lib() -> C.CDLL:
    """The loaded libhdm.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with "
                          "`python -m housdorffhdf5cmake_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i32, u64 = C.c_void_p, C.c_int, C.c_uint64
    sig = {
        "hdm_abi_version": ([], i32), "hdm_strerror": ([i32], C.c_char_p),
        "hdm_opts_init": ([C.POINTER(Opts)], None), "hdm_device_count": ([], i32),
        "hdm_workspace_bytes": ([i32, i32, i32, i32], u64),
        "hdm_last_timings": ([C.POINTER(C.c_float), i32], i32),
        "hdm_timing_log": ([C.POINTER(C.c_float), i32], i32),
        "hdm_last_trace": ([C.POINTER(C.c_uint64), i32, C.POINTER(C.c_double)], i32),
        "hdm_hausdorff_u8": ([vp, vp, i32, i32, i32, C.c_uint8, C.POINTER(Opts),
                              C.POINTER(Result)], i32),
        "hdm_hausdorff_i32": ([vp, vp, i32, i32, i32, C.c_int32, C.POINTER(Opts),
                               C.POINTER(Result)], i32),
        "hdm_hausdorff_bool": ([vp, vp, i32, i32, i32, C.POINTER(Opts), C.POINTER(Result)], i32),
        "hdm_first_hit_map_u8": ([vp, vp, i32, i32, i32, C.c_uint8, C.POINTER(Opts), vp,
                                  C.POINTER(Result)], i32),
        "hdm_hausdorff_u8_noalloc": ([vp, vp, i32, i32, i32, C.c_uint8, C.POINTER(Opts),
                                      C.POINTER(Result)], i32),
        "hdm_noalloc_max_bytes": ([], u64),
        "hdm_batch_create": ([C.POINTER(Pair), i32, C.POINTER(Opts), i32, C.POINTER(vp)], i32),
        "hdm_batch_run": ([vp], i32), "hdm_batch_sync": ([vp, C.POINTER(Result)], i32),
        "hdm_batch_last_ms": ([vp], C.c_float), "hdm_batch_destroy": ([vp], None),
        "hdm_batch_near_entries": ([vp], C.c_uint32),
        "hdm_hausdorff_batch": ([C.POINTER(Pair), i32, C.POINTER(Opts), C.POINTER(Result)], i32),
        "hdm_autotune": ([C.POINTER(Pair), i32, vp, i32, vp, i32, vp, i32, i32,
                          C.POINTER(TuneEntry), i32, C.POINTER(C.c_int)], i32),
        "hdm_server_opts_init": ([C.POINTER(ServerOpts)], None),
        "hdm_server_create": ([C.POINTER(ServerOpts), C.POINTER(vp)], i32),
        "hdm_server_submit": ([vp, vp, vp, i32, i32, i32, i32, i32, i32,
                               C.POINTER(C.c_uint64)], i32),
        "hdm_server_step": ([vp, C.POINTER(Completion), i32, C.POINTER(C.c_int)], i32),
        "hdm_server_load": ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                             C.POINTER(C.c_uint64)], i32),
        "hdm_server_now": ([], C.c_double), "hdm_server_destroy": ([vp], None),
        "hdm_server_register": ([vp, vp, i32, i32, i32, i32, i32, i32, C.POINTER(C.c_uint64)],
                                i32),
        "hdm_server_submit_masks": ([vp, u64, u64, C.POINTER(C.c_uint64)], i32),
        "hdm_server_unregister": ([vp, u64], i32),
        "hdm_server_stats_get": ([vp, C.POINTER(ServerStats)], i32),
        "hdm_synth_spec_sphere": ([C.POINTER(SynthSpec), i32, i32, i32, i32, i32, i32, i32, i32,
                                   i32, i32, C.c_double, u64], i32),
        "hdm_synth_spec_ellipsoids": ([C.POINTER(SynthSpec), i32, i32, i32, i32, i32, i32, i32,
                                       C.c_double, u64], i32),
        "hdm_synth_spec_long": ([C.POINTER(SynthSpec), i32, i32, i32, i32, i32, i32, i32, i32,
                                 i32, i32, i32], i32),
        "hdm_synth_host": ([C.POINTER(SynthSpec), vp, vp], i32),
        "hdm_h5_info_get": ([C.c_char_p, C.c_char_p, C.POINTER(H5Info)], i32),
        "hdm_h5_read": ([C.c_char_p, C.c_char_p, i32, vp, u64, i32], i32),
        "hdm_hausdorff_h5": ([C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, i32, i32, i32,
                              C.POINTER(Opts), C.POINTER(Result), C.POINTER(C.c_double)], i32),
        "hdm_synth_device": ([C.POINTER(SynthSpec), vp, vp, vp], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != HDM_OK:
        raise HDMError(rc, what)


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


_OPTS_DEFAULT = None


_RAW_STREAM = None


def _current_stream_ptr():
    """torch's current stream on the current device (a call's default)."""
    global _RAW_STREAM
    if _RAW_STREAM is None:
        import torch
        raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
        if raw is not None:  # the handle of current_stream().cuda_stream, without the object
            dev = torch.cuda.current_device
            _RAW_STREAM = lambda: raw(dev())
        else:
            _RAW_STREAM = lambda: torch.cuda.current_stream().cuda_stream
    return C.c_void_p(_RAW_STREAM())


_TORCH_DTYPES = None


def _volume_args(gold, segm):
    """(gold_ptr, segm_ptr, nx, ny, nz, dtype, ptr_kind, keepalive)."""
    global _TORCH_DTYPES
    shape = gold.shape
    if shape != segm.shape or len(shape) != 3:
        raise ValueError(f"need two equal 3-D volumes, got {tuple(gold.shape)} {tuple(segm.shape)}")
    nz, ny, nx = int(shape[0]), int(shape[1]), int(shape[2])
    if _is_torch(gold):
        if not (gold.is_cuda and segm.is_cuda):
            gold, segm = gold.numpy(), segm.numpy()
        else:
            if _TORCH_DTYPES is None:
                import torch
                _TORCH_DTYPES = {torch.uint8: DTYPE_U8, torch.bool: DTYPE_U8,
                                 torch.int32: DTYPE_I32}
            dt = _TORCH_DTYPES.get(gold.dtype)
            if dt is None or segm.dtype != gold.dtype:
                raise TypeError(f"unsupported or mismatched dtypes {gold.dtype} {segm.dtype}")
            if not (gold.is_contiguous() and segm.is_contiguous()):
                gold, segm = gold.contiguous(), segm.contiguous()
            return gold.data_ptr(), segm.data_ptr(), nx, ny, nz, dt, PTR_DEVICE, (gold, segm)
    g = np.ascontiguousarray(gold)
    s = np.ascontiguousarray(segm)
    if g.dtype == np.bool_:
        g, s = g.view(np.uint8), s.view(np.uint8)
    if g.dtype != s.dtype:
        raise ValueError("dtype mismatch")
    if g.dtype == np.uint8:
        dt = DTYPE_U8
    elif g.dtype == np.int32:
        dt = DTYPE_I32
    else:
        raise TypeError(f"unsupported dtype {g.dtype}")
    return (C.c_void_p(g.ctypes.data), C.c_void_p(s.ctypes.data), nx, ny, nz, dt, PTR_HOST,
            (g, s))


def make_opts(robustness=1.0, tile=TILE_AUTO, grid_ctas=0, chunk=0, crop=True, stream=None,
              ptr_kind=PTR_DEVICE, timing=False, connectivity=6) -> Opts:
    global _OPTS_DEFAULT
    if _OPTS_DEFAULT is None:
        _OPTS_DEFAULT = Opts()
        lib().hdm_opts_init(C.byref(_OPTS_DEFAULT))
    o = Opts.from_buffer_copy(_OPTS_DEFAULT)
    o.robustness = float(robustness)
    o.tile = int(tile)
    o.grid_ctas = int(grid_ctas)
    o.chunk = int(chunk)
    o.crop = 1 if crop else 0
    o.ptr_kind = ptr_kind
    o.timing = int(timing)
    o.connectivity = int(connectivity)
    if stream is not None:
        o.stream = stream
    return o


def last_timings() -> dict:
    """Device ms of the phases of the last call made with timing=True; "fused"
    is True when the call ran the fused single-pair path (phases then come
    from the kernel's own clock, and timing=3 fills only "total"), "launches"
    counts its launches (1, or 2 when the pair was handed to the region
    launch) and "steps" the dilation steps per grid barrier of the body that
    finished (8 for the near pass and the region launch)."""
    ms = (C.c_float * 7)()
    n = lib().hdm_last_timings(ms, 7)
    names = ("pack", "meta", "dilate", "finalize", "total")
    out = {names[i]: float(ms[i]) for i in range(min(n, 5))}
    out["fused"] = n > 5 and ms[5] >= 1.0
    # launches of the fused path: 2 when the first handed the pair to the region launch
    out["launches"] = int(ms[5]) if n > 5 else 0
    out["steps"] = int(ms[6]) if n > 6 else 1  # dilation steps per grid barrier
    return out


def timing_log(max_calls: int = 1024) -> list:
    """Device ms of the dominant launch of every call made with timing=3
    since the last read, oldest first (waits for their events)."""
    ms = (C.c_float * max_calls)()
    n = lib().hdm_timing_log(ms, max_calls)
    return [float(ms[i]) for i in range(n)]


def last_trace(max_iters: int = 4096):
    """Per-iteration trace of the last call made with timing=2: a list of dicts
    with the phase times in microseconds from the iteration start and the
    queue length."""
    buf = (C.c_uint64 * (8 * max_iters))()
    tick = C.c_double()
    n = lib().hdm_last_trace(buf, max_iters, C.byref(tick))
    out = []
    for i in range(n):
        r = buf[8 * i: 8 * i + 8]
        us = lambda k: round((r[k] - r[0]) * tick.value / 1e3, 2) if r[k] else None
        nxt = buf[8 * (i + 1)] if i + 1 < n else 0
        out.append({"t": i, "queued": int(r[5]), "all_done_us": us(1), "wg0_done_us": us(2),
                    "arrived_us": us(3), "folded_us": us(4),
                    "next_start_us": round((nxt - r[0]) * tick.value / 1e3, 2) if nxt else None,
                    # region kernel: entries stepped, mean step phase of one
                    "entries": int(r[6]),
                    "entry_steps_us_mean": round(r[7] / r[6] * tick.value / 1e3, 2) if r[6] else None})
    return out


def hausdorff(gold, segm, label: int = 1, robustness: float = 1.0, tile: int = TILE_AUTO,
              grid_ctas: int = 0, chunk: int = 0, crop: bool = True, raise_on_error: bool = True,
              hist_len: int = 0, timing: bool = False, connectivity: int = 6):
    """HD and average HD of two volumes.  Returns a dict of hdm_result fields
    (plus 'status', and 'hist_fp'/'hist_fn' when hist_len > 0).  connectivity:
    the structuring element, 6 (the reference's cross), 18 (its forBench
    engine's CHEBYSHEV element) or 26 (the cube)."""
    gp, sp, nx, ny, nz, dt, pk, keep = _volume_args(gold, segm)
    o = make_opts(robustness, tile, grid_ctas, chunk, crop,
                  _current_stream_ptr() if pk == PTR_DEVICE else None, pk, timing, connectivity)
    hist = None
    if hist_len > 0:
        hist = np.zeros(2 * hist_len, np.uint32)
        o.hist_out = hist.ctypes.data
        o.hist_len = hist_len
    res = Result()
    if dt == DTYPE_U8:
        rc = lib().hdm_hausdorff_u8(gp, sp, nx, ny, nz, label, C.byref(o), C.byref(res))
    else:
        rc = lib().hdm_hausdorff_i32(gp, sp, nx, ny, nz, label, C.byref(o), C.byref(res))
    del keep
    if raise_on_error:
        check(rc, "hdm_hausdorff")
    d = res.as_dict()
    d["status"] = rc
    if hist is not None:
        d["hist_fp"] = hist[:hist_len].copy()
        d["hist_fn"] = hist[hist_len:].copy()
    return d


def first_hit_map(gold, segm, label: int = 1, robustness: float = 1.0, tile: int = TILE_AUTO,
                  connectivity: int = 6):
    """(result dict, uint16 map [z,y,x]) with d(v) on FP and FN voxels."""
    gp, sp, nx, ny, nz, dt, pk, keep = _volume_args(gold, segm)
    if dt != DTYPE_U8:
        raise TypeError("first_hit_map takes uint8 volumes")
    if pk == PTR_DEVICE:
        import torch
        mp = torch.empty((nz, ny, nx), dtype=torch.int16, device=gold.device)
        mptr = C.c_void_p(mp.data_ptr())
        o = make_opts(robustness, tile, stream=_current_stream_ptr(), ptr_kind=pk,
                      connectivity=connectivity)
    else:
        mp = np.zeros((nz, ny, nx), np.uint16)
        mptr = C.c_void_p(mp.ctypes.data)
        o = make_opts(robustness, tile, ptr_kind=pk, connectivity=connectivity)
    res = Result()
    rc = lib().hdm_first_hit_map_u8(gp, sp, nx, ny, nz, label, C.byref(o), mptr, C.byref(res))
    del keep
    check(rc, "hdm_first_hit_map_u8")
    return res.as_dict(), mp


# ---------------------------------------------------------------- HDF5
H5_AS_U8, H5_AS_I32, H5_AS_NONZERO = 0, 1, 2


def h5_info(path: str, dataset: str) -> dict:
    """Shape, element type and layout of an HDF5 dataset (the library's own
    reader; no libhdf5)."""
    info = H5Info()
    check(lib().hdm_h5_info_get(os.fsencode(path), dataset.encode(), C.byref(info)),
          f"hdm_h5_info_get {path}:{dataset}")
    return info.as_dict()


def h5_read(path: str, dataset: str, conversion: int = H5_AS_U8, threads: int = 0) -> np.ndarray:
    """The whole dataset as a numpy array of its shape: uint8 (H5_AS_U8,
    values must fit), int32 (H5_AS_I32) or uint8 0/1 (H5_AS_NONZERO, the
    reference's bool read, forBench/hdf5Manag.cu:76)."""
    info = h5_info(path, dataset)
    out = np.empty(info["shape"], np.int32 if conversion == H5_AS_I32 else np.uint8)
    check(lib().hdm_h5_read(os.fsencode(path), dataset.encode(), conversion, out.ctypes.data,
                            out.size, threads), f"hdm_h5_read {path}:{dataset}")
    return out


def hausdorff_h5(gold_path: str, gold_dataset: str, segm_path: str, segm_dataset: str,
                 nx: int, ny: int, nz: int, robustness: float = 1.0, connectivity: int = 6):
    """The reference's loadHDF (forBench/hdf5Manag.cu:196-220): two datasets
    of nx*ny*nz booleans (x fastest) -> result dict with 'load_ms'."""
    o = make_opts(robustness, ptr_kind=PTR_HOST, connectivity=connectivity)
    res = Result()
    ms = C.c_double()
    check(lib().hdm_hausdorff_h5(os.fsencode(gold_path), gold_dataset.encode(),
                                 os.fsencode(segm_path), segm_dataset.encode(), nx, ny, nz,
                                 C.byref(o), C.byref(res), C.byref(ms)), "hdm_hausdorff_h5")
    d = res.as_dict()
    d["load_ms"] = ms.value
    return d


def getHausdorffDistance_CUDA_Generic(goldStandard, algoOutput, WIDTH: int, HEIGHT: int,
                                      DEPTH: int, robustnessPercent: float = 1.0,
                                      resIterNeeded: bool = False, numberToLookFor: int = 1,
                                      res3DNeeded: bool = False) -> int:
    """The reference's entry shape (hello.cu:2025-2062): returns the HD.
    Raises HDMError where the reference would hang (empty mask) or return
    garbage; identical masks give 0 (the reference returns 1)."""
    for v in (goldStandard, algoOutput):
        if tuple(int(s) for s in v.shape) != (DEPTH, HEIGHT, WIDTH):
            raise ValueError("volume shape must be (DEPTH, HEIGHT, WIDTH)")
    return int(hausdorff(goldStandard, algoOutput, label=numberToLookFor,
                         robustness=robustnessPercent)["hd"])


# ---------------------------------------------------------------- synthetic
def synth_spec(kind: str, *args) -> SynthSpec:
    sp = SynthSpec()
    f = {"sphere": lib().hdm_synth_spec_sphere, "ellipsoids": lib().hdm_synth_spec_ellipsoids,
         "long": lib().hdm_synth_spec_long}[kind]
    check(f(C.byref(sp), *args), f"synth_spec {kind}")
    return sp


def synth_host(sp: SynthSpec):
    g = np.empty((sp.nz, sp.ny, sp.nx), np.uint8)
    s = np.empty_like(g)
    check(lib().hdm_synth_host(C.byref(sp), C.c_void_p(g.ctypes.data), C.c_void_p(s.ctypes.data)))
    return g, s


def synth_device(sp: SynthSpec, device="cuda"):
    import torch
    g = torch.empty((sp.nz, sp.ny, sp.nx), dtype=torch.uint8, device=device)
    s = torch.empty_like(g)
    check(lib().hdm_synth_device(C.byref(sp), C.c_void_p(g.data_ptr()), C.c_void_p(s.data_ptr()),
                                 _current_stream_ptr()))
    return g, s


# ---------------------------------------------------------------- batch
def _pairs(vols, label=1):
    arr = (Pair * len(vols))()
    keep = []
    for i, (g, s) in enumerate(vols):
        gp, sp, nx, ny, nz, dt, pk, k = _volume_args(g, s)
        if pk != PTR_DEVICE:
            raise ValueError("batched pairs must be device tensors")
        arr[i].gold, arr[i].segm = gp, sp
        arr[i].nx, arr[i].ny, arr[i].nz = nx, ny, nz
        arr[i].dtype, arr[i].label = dt, label
        keep.append(k)
    return arr, keep


class Batch:
    """Many pairs through one persistent launch; with graph=True the launch
    sequence is captured once and replayed."""

    def __init__(self, vols, label=1, robustness=1.0, tile=TILE_AUTO, chunk=0, grid_ctas=0,
                 graph=True, crop=True, connectivity=6):
        self.pairs, self._keep = _pairs(vols, label)
        self.n = len(vols)
        o = make_opts(robustness, tile, grid_ctas, chunk, crop, connectivity=connectivity)
        self.h = C.c_void_p()
        check(lib().hdm_batch_create(self.pairs, self.n, C.byref(o), 1 if graph else 0,
                                     C.byref(self.h)), "hdm_batch_create")

    def run(self):
        check(lib().hdm_batch_run(self.h), "hdm_batch_run")

    def sync(self, results=True):
        if not results:
            check(lib().hdm_batch_sync(self.h, None), "hdm_batch_sync")
            return None
        out = (Result * self.n)()
        rc = lib().hdm_batch_sync(self.h, out)
        res = [r.as_dict() for r in out]
        check(rc, "hdm_batch_sync")
        return res

    def last_ms(self) -> float:
        return float(lib().hdm_batch_last_ms(self.h))

    def near_entries(self) -> int:
        """(region, mask) entries the near pass of the last synced run stepped."""
        return int(lib().hdm_batch_near_entries(self.h))

    def close(self):
        if self.h:
            lib().hdm_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(vols, tiles=(TILE_64x8x8, TILE_64x4x4, TILE_64x8x4, TILE_32x8x8), chunks=(1, 2, 4),
             grids=(0,), reps=3, label=1):
    """Times the batched launch for each (tile, chunk, grid); fastest first."""
    pairs, keep = _pairs(vols, label)
    t = (C.c_int32 * len(tiles))(*tiles)
    c = (C.c_int32 * len(chunks))(*chunks)
    g = (C.c_int32 * len(grids))(*grids)
    cap = len(tiles) * len(chunks) * len(grids)
    out = (TuneEntry * cap)()
    n = C.c_int()
    check(lib().hdm_autotune(pairs, len(vols), t, len(tiles), c, len(chunks), g, len(grids), reps,
                             out, cap, C.byref(n)), "hdm_autotune")
    del keep
    return [{"tile": TILE_NAMES.get(e.tile, e.tile), "tile_id": e.tile, "chunk": e.chunk,
             "grid_ctas": e.grid_ctas, "status": e.status, "ms": e.ms} for e in out[:n.value]]
