"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
benchmark's cpu_baseline leg, to compare results against.  The product package never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OracleResult(C.Structure):
    _fields_ = [("hd", C.c_int32), ("hd_fp", C.c_int32), ("hd_fn", C.c_int32),
                ("iterations", C.c_int32), ("sum", C.c_uint64), ("count", C.c_uint64),
                ("sum_fp", C.c_uint64), ("cnt_fp", C.c_uint64), ("sum_fn", C.c_uint64),
                ("cnt_fn", C.c_uint64), ("n_gold", C.c_uint64), ("n_segm", C.c_uint64),
                ("n_fp", C.c_uint64), ("n_fn", C.c_uint64), ("avg_hd", C.c_double),
                ("queue_total", C.c_uint64), ("queue_max", C.c_uint32), ("tiles", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        args = [C.c_void_p, C.c_void_p, C.c_int, C.c_int32, C.c_int, C.c_int, C.c_int,
                C.c_double]
        L.oracle_naive.argtypes = args + [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                          C.POINTER(OracleResult)]
        L.oracle_naive.restype = C.c_int
        L.oracle_tiled.argtypes = args + [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                          C.c_void_p, C.POINTER(OracleResult)]
        L.oracle_tiled.restype = C.c_int
        L.oracle_naive_se.argtypes = args + [C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                             C.c_void_p, C.POINTER(OracleResult)]
        L.oracle_naive_se.restype = C.c_int
        L.oracle_tiled_se.argtypes = args + [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                             C.c_int, C.c_void_p, C.POINTER(OracleResult)]
        L.oracle_tiled_se.restype = C.c_int
        L.oracle_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _prep(gold: np.ndarray, segm: np.ndarray):
    assert gold.shape == segm.shape and gold.ndim == 3, (gold.shape, segm.shape)
    assert gold.dtype == segm.dtype
    if gold.dtype == np.uint8:
        dt = 0
    elif gold.dtype == np.int32:
        dt = 1
    else:
        raise TypeError(gold.dtype)
    g = np.ascontiguousarray(gold)
    s = np.ascontiguousarray(segm)
    nz, ny, nx = g.shape
    return g, s, dt, nx, ny, nz


def run(gold: np.ndarray, segm: np.ndarray, label: int = 1, robustness: float = 1.0,
        method: str = "tiled", threads: int = 0, crop: bool = True, want_map: bool = False,
        conn: int = 6):
    """Runs the oracle on [z, y, x] volumes with structuring element conn (6,
    18 or 26).  Returns (rc, result dict, hist_fp, hist_fn, map or None)."""
    g, s, dt, nx, ny, nz = _prep(gold, segm)
    L = nx + ny + nz + 2
    hf = np.zeros(L, np.uint32)
    hn = np.zeros(L, np.uint32)
    mp = np.zeros(g.shape, np.uint16) if want_map else None
    res = OracleResult()
    mptr = mp.ctypes.data if mp is not None else None
    if method == "naive":
        rc = lib().oracle_naive_se(g.ctypes.data, s.ctypes.data, dt, label, nx, ny, nz,
                                   robustness, conn, hf.ctypes.data, hn.ctypes.data, L, mptr,
                                   C.byref(res))
    else:
        rc = lib().oracle_tiled_se(g.ctypes.data, s.ctypes.data, dt, label, nx, ny, nz,
                                   robustness, conn, threads, 1 if crop else 0, hf.ctypes.data,
                                   hn.ctypes.data, L, mptr, C.byref(res))
    return rc, res.as_dict(), hf, hn, mp


def max_threads() -> int:
    return int(lib().oracle_max_threads())
