"""ctypes binding of the CPU oracle (oracle/liboracle.so). Used by tests only, as the checker."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

_lib = None
_P = ctypes.POINTER


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all"], check=True,
                           capture_output=True)
        so = ctypes.CDLL(ORACLE_SO)
        i64p, i32p = _P(ctypes.c_int64), _P(ctypes.c_int32)
        so.oracle_static_mochy.argtypes = [ctypes.c_int32, i64p, i32p, i64p, ctypes.c_int]
        so.oracle_static_brute.argtypes = [ctypes.c_int32, i64p, i32p, i64p]
        so.oracle_thyme_temporal.argtypes = [ctypes.c_int32, i64p, i32p, i64p, ctypes.c_int64, i64p]
        so.oracle_temporal_brute.argtypes = [ctypes.c_int32, i64p, i32p, i64p, ctypes.c_int64, i64p]
        so.oracle_thyme_temporal_range.argtypes = [ctypes.c_int32, i64p, i32p, i64p, ctypes.c_int64,
                                                   ctypes.c_int32, ctypes.c_int32, i64p]
        so.oracle_projection_degrees.argtypes = [ctypes.c_int32, i64p, i32p, i64p]
        so.oracle_mochy_table.restype = _P(ctypes.c_int)
        so.oracle_thyme_temporal_table.restype = _P(ctypes.c_int)
        for fn in ("oracle_static_mochy", "oracle_static_brute", "oracle_thyme_temporal",
                   "oracle_thyme_temporal_range",
                   "oracle_temporal_brute", "oracle_projection_degrees"):
            getattr(so, fn).restype = ctypes.c_int
        _lib = so
    return _lib


def _arr(off, nodes):
    off = np.ascontiguousarray(off, dtype=np.int64)
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    if nodes.size == 0:
        nodes = np.zeros(1, dtype=np.int32)
    return off, nodes, off.ctypes.data_as(_P(ctypes.c_int64)), nodes.ctypes.data_as(_P(ctypes.c_int32))


def static_mochy(off, nodes, threads: int = 8) -> np.ndarray:
    off, nodes, po, pn = _arr(off, nodes)
    out = np.zeros(30, dtype=np.int64)
    assert lib().oracle_static_mochy(off.size - 1, po, pn, out.ctypes.data_as(_P(ctypes.c_int64)),
                                     threads) == 0
    return out


def static_brute(off, nodes) -> np.ndarray:
    off, nodes, po, pn = _arr(off, nodes)
    out = np.zeros(30, dtype=np.int64)
    assert lib().oracle_static_brute(off.size - 1, po, pn, out.ctypes.data_as(_P(ctypes.c_int64))) == 0
    return out


def thyme_temporal(off, nodes, times, delta: int) -> np.ndarray:
    off, nodes, po, pn = _arr(off, nodes)
    times = np.ascontiguousarray(times, dtype=np.int64)
    out = np.zeros(96, dtype=np.int64)
    assert lib().oracle_thyme_temporal(off.size - 1, po, pn, times.ctypes.data_as(_P(ctypes.c_int64)),
                                       delta, out.ctypes.data_as(_P(ctypes.c_int64))) == 0
    return out


def thyme_temporal_range(off, nodes, times, delta: int, lo: int, hi: int) -> np.ndarray:
    """THyMe+ counts of the instances whose newest edge has time-order position in [lo, hi)."""
    off, nodes, po, pn = _arr(off, nodes)
    times = np.ascontiguousarray(times, dtype=np.int64)
    out = np.zeros(96, dtype=np.int64)
    assert lib().oracle_thyme_temporal_range(off.size - 1, po, pn, times.ctypes.data_as(_P(ctypes.c_int64)),
                                             delta, lo, hi, out.ctypes.data_as(_P(ctypes.c_int64))) == 0
    return out


def temporal_brute(off, nodes, times, delta: int) -> np.ndarray:
    off, nodes, po, pn = _arr(off, nodes)
    times = np.ascontiguousarray(times, dtype=np.int64)
    out = np.zeros(96, dtype=np.int64)
    assert lib().oracle_temporal_brute(off.size - 1, po, pn, times.ctypes.data_as(_P(ctypes.c_int64)),
                                       delta, out.ctypes.data_as(_P(ctypes.c_int64))) == 0
    return out


def projection_degrees(off, nodes) -> np.ndarray:
    off, nodes, po, pn = _arr(off, nodes)
    deg = np.zeros(max(off.size - 1, 1), dtype=np.int64)
    assert lib().oracle_projection_degrees(off.size - 1, po, pn, deg.ctypes.data_as(_P(ctypes.c_int64))) == 0
    return deg[: off.size - 1]


def mochy_table() -> list[int]:
    t = lib().oracle_mochy_table()
    return [t[i] for i in range(128)]


def thyme_temporal_table() -> list[int]:
    t = lib().oracle_thyme_temporal_table()
    return [t[i] for i in range(128)]


def _layers(layers):
    offs = [np.ascontiguousarray(o, dtype=np.int64) for o, _ in layers]
    nods = [np.ascontiguousarray(n if len(n) else np.zeros(1), dtype=np.int32) for _, n in layers]
    po = (ctypes.c_void_p * 3)(*[o.ctypes.data for o in offs])
    pn = (ctypes.c_void_p * 3)(*[n.ctypes.data for n in nods])
    nr = (ctypes.c_int32 * 3)(*[o.size - 1 for o in offs])
    return offs, nods, po, pn, nr


def strict_inc(older, middle, newest) -> np.ndarray:
    """The reference's three-snapshot strict-increasing count (restatement of its kernel)."""
    keep = _layers([older, middle, newest])
    out = np.zeros(30, dtype=np.int64)
    assert lib().oracle_strict_inc(keep[2], keep[3], keep[4], out.ctypes.data_as(_P(ctypes.c_int64))) == 0
    return out


def strict_inc_brute(older, middle, newest) -> np.ndarray:
    keep = _layers([older, middle, newest])
    out = np.zeros(30, dtype=np.int64)
    assert lib().oracle_strict_inc_brute(keep[2], keep[3], keep[4], out.ctypes.data_as(_P(ctypes.c_int64))) == 0
    return out
