"""ipdps26_amd — Python host side of libhmotif, an MI355X engine for exact hypergraph motifs.

Everything here is a thin ctypes layer over the C ABI declared in ``include/hmotif.h``; the
counting itself runs in the hand-written HIP kernels of ``libhmotif.so`` (built in-tree by
``__graft_entry__.build()`` or ``make -C ipdps26_amd/csrc``). If the library is missing, every
entry point raises: there is no CPU fallback.

Output numbering follows the reference: static counts are the 30 MoCHy bins
(MoCHy-master/motif_id.cpp:5-34; ``printed_motifs`` gives MoCHy's 26 printed ones), temporal
counts the 96 THyMe bins (THyMe-main/code/motif_id.cpp:25-42).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhmotif.so")

STATIC_BINS = 30
TEMPORAL_BINS = 96
HIDDEN_STATIC_BINS = (0, 1, 4, 6)
PRINTED_STATIC_BINS = tuple(b for b in range(STATIC_BINS) if b not in HIDDEN_STATIC_BINS)

_STATUS = {
    0: "ok",
    1: "out of memory",
    2: "over budget",
    3: "bad input",
    4: "device error",
    5: "unsupported input",
    6: "buffer too small",
}


class HMotifError(RuntimeError):
    """A libhmotif call returned a status other than HM_OK."""

    def __init__(self, status: int, message: str):
        self.status = status
        super().__init__(f"{_STATUS.get(status, status)}: {message}")


class _GenParams(ctypes.Structure):
    _fields_ = [
        ("num_nodes", ctypes.c_int64),
        ("num_edges", ctypes.c_int64),
        ("alpha", ctypes.c_double),
        ("kmin", ctypes.c_int32),
        ("kmax", ctypes.c_int32),
        ("zipf_s", ctypes.c_double),
        ("seed", ctypes.c_uint64),
    ]


class _GraphStats(ctypes.Structure):
    _fields_ = [
        ("num_edges", ctypes.c_int64),
        ("num_nodes", ctypes.c_int64),
        ("num_incidences", ctypes.c_int64),
        ("proj_pairs", ctypes.c_int64),
        ("max_proj_degree", ctypes.c_int64),
        ("oriented_row_sum", ctypes.c_int64),
        ("wedges", ctypes.c_int64),
        ("triangles", ctypes.c_int64),
        ("max_node_degree", ctypes.c_int64),
    ]


class _HubStats(ctypes.Structure):
    _fields_ = [
        ("hubs", ctypes.c_int32),
        ("classes", ctypes.c_int64),
        ("light_pairs", ctypes.c_int64),
        ("pair_types", ctypes.c_int64),
        ("light_triangles", ctypes.c_int64),
        ("ms_classes", ctypes.c_double),
        ("ms_wedges", ctypes.c_double),
        ("ms_types", ctypes.c_double),
        ("ms_triangles", ctypes.c_double),
        ("ms_tri_kernel", ctypes.c_double),
        ("big_centres", ctypes.c_int64),
    ]


_P = ctypes.POINTER
_i32p = _P(ctypes.c_int32)
_i64p = _P(ctypes.c_int64)
_u64p = _P(ctypes.c_uint64)

# (name, restype, argtypes) of every C entry point this module binds.
_SIGNATURES = [
    ("hm_status_string", ctypes.c_char_p, [ctypes.c_int]),
    ("hm_last_error", ctypes.c_char_p, []),
    ("hm_bin_table", ctypes.c_int, [ctypes.c_int32, _P(ctypes.c_int8)]),
    ("hm_gen_static_size", ctypes.c_int, [_P(_GenParams), _i64p]),
    ("hm_gen_static", ctypes.c_int, [_P(_GenParams), _i64p, _i32p, ctypes.c_int64]),
    ("hm_gen_timestamps", ctypes.c_int,
     [ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_int32, ctypes.c_double,
      ctypes.c_uint64, _i64p]),
    ("hm_gen_pick", ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64, _i64p]),
    ("hm_init", ctypes.c_int, [ctypes.c_int32]),
    ("hm_graph_create", ctypes.c_int,
     [_i64p, _i32p, ctypes.c_int32, ctypes.c_int32, _P(ctypes.c_void_p)]),
    ("hm_graph_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("hm_count_static", ctypes.c_int, [ctypes.c_void_p, _u64p]),
    ("hm_graph_get_sizes", ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int32), _P(ctypes.c_int32), _P(ctypes.c_int64)]),
    ("hm_graph_read_csr", ctypes.c_int, [ctypes.c_void_p, _i64p, _i32p, _i64p, _i32p]),
    ("hm_count_static_noalloc", ctypes.c_int, [ctypes.c_void_p, _u64p]),
    ("hm_graph_get_stats", ctypes.c_int, [ctypes.c_void_p, _P(_GraphStats)]),
    ("hm_graph_phase_ms", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, _P(ctypes.c_double)]),
    ("hm_graph_set_tile", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("hm_graph_set_hubs", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("hm_graph_get_hub_stats", ctypes.c_int, [ctypes.c_void_p, _P(_HubStats)]),
    ("hm_graph_count_hub_part", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64,
                                               ctypes.c_int32, ctypes.c_double, _P(ctypes.c_uint64),
                                               _P(ctypes.c_int64), _P(ctypes.c_int64)]),
    ("hm_graph_set_replay", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("hm_graph_set_tri_form", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    ("hm_graph_get_tri_form", ctypes.c_int, [ctypes.c_void_p, _i32p, _i32p, _i32p]),
    ("hm_graph_replay_state", ctypes.c_int, [ctypes.c_void_p, _i32p, _i32p]),
    ("hm_graph_set_update_capture", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("hm_graph_update_capture_state", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("hm_tgraph_set_tile", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("hm_graph_update_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("hm_count_strict_inc", ctypes.c_int,
     [ctypes.c_void_p * 3, ctypes.c_void_p * 3, ctypes.c_int32 * 3, _P(ctypes.c_uint64)]),
    ("hm_graph_update", ctypes.c_int,
     [ctypes.c_void_p, _i32p, ctypes.c_int32, _i64p, _i32p, ctypes.c_int32, _i32p, _i64p]),
    ("hm_count_delta", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, _i32p, ctypes.c_int32, _i32p, ctypes.c_int32, _i64p]),
    ("hm_batch_create", ctypes.c_int, [_P(ctypes.c_void_p)]),
    ("hm_batch_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("hm_batch_count_static", ctypes.c_int,
     [ctypes.c_void_p, _P(ctypes.c_void_p), ctypes.c_int32, _u64p]),
    ("hm_batch_set_form", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    ("hm_batch_get_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("hm_server_create", ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_void_p)]),
    ("hm_server_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("hm_server_add_graph", ctypes.c_int, [ctypes.c_void_p, _i64p, _i32p, ctypes.c_int32, _i32p]),
    ("hm_server_remove_graph", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("hm_server_submit_count", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _u64p]),
    ("hm_server_submit_update", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int32, _i32p, ctypes.c_int32, _i64p, _i32p, ctypes.c_int32, _u64p]),
    ("hm_server_wait", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]),
    ("hm_server_poll", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint64, _P(ctypes.c_int32), ctypes.c_void_p]),
    ("hm_server_get_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("hm_probe_valu", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]),
    ("hm_tgraph_create", ctypes.c_int,
     [_i64p, _i32p, _i64p, ctypes.c_int32, _P(ctypes.c_void_p)]),
    ("hm_tgraph_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("hm_count_temporal", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, _u64p]),
    ("hm_tgraph_get_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("hm_tgraph_phase_ms", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, _P(ctypes.c_double)]),
]

_lib = None


def lib() -> ctypes.CDLL:
    """Loads libhmotif.so from the package directory (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() or "
                "`make -C ipdps26_amd/csrc`")
        so = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGNATURES:
            fn = getattr(so, name)
            fn.restype = res
            fn.argtypes = args
        _lib = so
    return _lib


def _check(status: int) -> None:
    if status != 0:
        raise HMotifError(status, lib().hm_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(_P(ctype))


def probe_valu(iters: int, lanes: int = 64) -> float:
    """Duration (ms) of the VALU calibration kernel (hm_probe_valu) on the current device."""
    ms = ctypes.c_double(0.0)
    _check(lib().hm_probe_valu(int(iters), int(lanes), ctypes.byref(ms)))
    return ms.value


def bin_table(which: str = "static") -> np.ndarray:
    """The kernels' Venn-id -> 0-based bin table (int8[128], -1 = not a connected triple)."""
    out = np.zeros(128, dtype=np.int8)
    _check(lib().hm_bin_table(0 if which == "static" else 1, _ptr(out, ctypes.c_int8)))
    return out


# ------------------------------------------------------------------------------------------
# Generator
# ------------------------------------------------------------------------------------------
@dataclass(frozen=True)
class GenParams:
    """Synthetic hypergraph parameters: power-law sizes on [kmin, kmax], Zipf node popularity."""

    num_nodes: int
    num_edges: int
    alpha: float
    kmin: int
    kmax: int
    zipf_s: float
    seed: int

    def _c(self) -> _GenParams:
        return _GenParams(self.num_nodes, self.num_edges, self.alpha, self.kmin, self.kmax,
                          self.zipf_s, self.seed)


def generate_static(p: GenParams) -> tuple[np.ndarray, np.ndarray]:
    """Returns (offsets int64[E+1], nodes int32[I]) of the seeded synthetic hypergraph."""
    cp = p._c()
    total = ctypes.c_int64(0)
    _check(lib().hm_gen_static_size(ctypes.byref(cp), ctypes.byref(total)))
    offsets = np.zeros(p.num_edges + 1, dtype=np.int64)
    nodes = np.zeros(max(total.value, 1), dtype=np.int32)
    _check(lib().hm_gen_static(ctypes.byref(cp), _ptr(offsets, ctypes.c_int64),
                               _ptr(nodes, ctypes.c_int32), nodes.size))
    return offsets, nodes[: total.value]


def generate_timestamps(n: int, seed: int, mean_gap_us: float = 1e6, burst_fraction: float = 0.3,
                        burst_len: int = 50, burst_gap_us: float = 1e4) -> np.ndarray:
    """Strictly increasing int64 microsecond timestamps (exponential gaps plus bursts)."""
    out = np.zeros(max(n, 1), dtype=np.int64)
    _check(lib().hm_gen_timestamps(n, mean_gap_us, burst_fraction, burst_len, burst_gap_us, seed,
                                   _ptr(out, ctypes.c_int64)))
    return out[:n]


def pick(n: int, k: int, seed: int) -> np.ndarray:
    """k distinct indices drawn uniformly from [0, n)."""
    out = np.zeros(max(k, 1), dtype=np.int64)
    _check(lib().hm_gen_pick(n, k, seed, _ptr(out, ctypes.c_int64)))
    return out[:k]


# ------------------------------------------------------------------------------------------
# Static graph handle
# ------------------------------------------------------------------------------------------
def init(device: int = 0) -> None:
    """Selects the HIP device for handles created afterwards (must be gfx950)."""
    _check(lib().hm_init(device))


def as_csr(edges) -> tuple[np.ndarray, np.ndarray]:
    """Converts a list of node lists into (offsets int64, nodes int32)."""
    offsets = np.zeros(len(edges) + 1, dtype=np.int64)
    flat = []
    for i, e in enumerate(edges):
        flat.extend(int(v) for v in e)
        offsets[i + 1] = len(flat)
    return offsets, np.asarray(flat, dtype=np.int32)


class Graph:
    """A hypergraph resident in HBM (flat sorted H2V and V2H arrays)."""

    def __init__(self, offsets: np.ndarray, nodes: np.ndarray, num_nodes: int = 0):
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        if nodes.size == 0:
            nodes = np.zeros(1, dtype=np.int32)
        self._h = ctypes.c_void_p()
        _check(lib().hm_graph_create(_ptr(offsets, ctypes.c_int64), _ptr(nodes, ctypes.c_int32),
                                     offsets.size - 1, num_nodes, ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().hm_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def count_static(self) -> np.ndarray:
        """Exact static h-motif counts: uint64[30] in MoCHy bin order."""
        out = np.zeros(STATIC_BINS, dtype=np.uint64)
        _check(lib().hm_count_static(self._h, _ptr(out, ctypes.c_uint64)))
        return out

    def count_static_noalloc(self) -> np.ndarray:
        """The same counts without any device allocation (hm_count_static_noalloc: the library's
        static per-GPU scratch; one-table graphs of at most ~64 MiB of count scratch)."""
        out = np.zeros(STATIC_BINS, dtype=np.uint64)
        _check(lib().hm_count_static_noalloc(self._h, _ptr(out, ctypes.c_uint64)))
        return out

    def read_csr(self):
        """The store as it is now: (offsets int64[E+1], nodes int32[I], voffsets int64[V+1],
        vedges int32[I]); a deleted hyperedge is an empty row."""
        E, V, I = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int64(0)
        _check(lib().hm_graph_get_sizes(self._h, ctypes.byref(E), ctypes.byref(V), ctypes.byref(I)))
        off = np.zeros(E.value + 1, dtype=np.int64)
        nodes = np.zeros(max(I.value, 1), dtype=np.int32)
        voff = np.zeros(V.value + 1, dtype=np.int64)
        vedges = np.zeros(max(I.value, 1), dtype=np.int32)
        _check(lib().hm_graph_read_csr(self._h, _ptr(off, ctypes.c_int64), _ptr(nodes, ctypes.c_int32),
                                       _ptr(voff, ctypes.c_int64), _ptr(vedges, ctypes.c_int32)))
        return off, nodes[: I.value], voff, vedges[: I.value]

    def stats(self) -> dict:
        st = _GraphStats()
        _check(lib().hm_graph_get_stats(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in _GraphStats._fields_}

    def phase_ms(self, phase: str) -> float:
        ms = ctypes.c_double(0.0)
        _check(lib().hm_graph_phase_ms(self._h, phase.encode(), ctypes.byref(ms)))
        return ms.value

    def set_tile(self, tile_ids: int) -> None:
        """Key tile of count_static: 0 = automatic, > 0 = forced tiles of that many ids."""
        _check(lib().hm_graph_set_tile(self._h, int(tile_ids)))

    def set_hubs(self, hubs: int) -> None:
        """Hub-aware count_static: 0 = off (triangles of the full projection), 1..23 = that many
        hubs, -1 = hub count from the cost model (hm_graph_set_hubs)."""
        _check(lib().hm_graph_set_hubs(self._h, int(hubs)))

    def hub_stats(self) -> dict:
        st = _HubStats()
        _check(lib().hm_graph_get_hub_stats(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in _HubStats._fields_}

    def count_hub_part(self, hubs: int, centre_begin: int = 0, centre_end: int = -1,
                       other_parts: bool = True, seconds: float = 0.0, with_done: bool = False):
        """One part of a hub-aware count (hm_graph_count_hub_part): the light triangles centred at
        light ranks [centre_begin, centre_end), plus the other parts if other_parts; seconds > 0
        stops starting launches after that long. Returns (uint64[30] bins, light triangles of the
        range), plus the end of the ranks enumerated if with_done."""
        out = np.zeros(STATIC_BINS, dtype=np.uint64)
        lt, done = ctypes.c_int64(0), ctypes.c_int64(0)
        _check(lib().hm_graph_count_hub_part(self._h, int(hubs), int(centre_begin), int(centre_end),
                                             int(bool(other_parts)), float(seconds),
                                             out.ctypes.data_as(_P(ctypes.c_uint64)), ctypes.byref(lt),
                                             ctypes.byref(done)))
        return (out, int(lt.value), int(done.value)) if with_done else (out, int(lt.value))

    def update(self, deleted, ins_offsets=None, ins_nodes=None, delta: bool = True):
        """Deletes the ids in `deleted`, inserts the CSR (ins_offsets, ins_nodes); returns
        (assigned ids int32[n_ins], delta int64[30] or None)."""
        dels = np.ascontiguousarray(np.asarray(deleted, dtype=np.int32).reshape(-1))
        if ins_offsets is None:
            ins_offsets, ins_nodes = np.zeros(1, dtype=np.int64), np.zeros(0, dtype=np.int32)
        ioff = np.ascontiguousarray(ins_offsets, dtype=np.int64)
        inodes = np.ascontiguousarray(ins_nodes, dtype=np.int32)
        n_ins = ioff.size - 1
        assigned = np.zeros(max(n_ins, 1), dtype=np.int32)
        out = np.zeros(STATIC_BINS, dtype=np.int64)
        _check(lib().hm_graph_update(
            self._h, _ptr(dels if dels.size else np.zeros(1, np.int32), ctypes.c_int32), dels.size,
            _ptr(ioff, ctypes.c_int64), _ptr(inodes if inodes.size else np.zeros(1, np.int32), ctypes.c_int32),
            n_ins, _ptr(assigned, ctypes.c_int32), _ptr(out, ctypes.c_int64) if delta else None))
        return assigned[:n_ins], (out if delta else None)

    def set_tri_form(self, waves: int = 0, copies: int = 16) -> None:
        """waves=0: autotuned per shape; -1: tune again on the next count; else 4 or 8 waves with
        8 or 16 histogram copies."""
        _check(lib().hm_graph_set_tri_form(self._h, int(waves), int(copies)))

    def tri_form(self) -> tuple[int, int, bool]:
        """(waves, copies, tuned) of the last one-table count."""
        w, c, t = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
        _check(lib().hm_graph_get_tri_form(self._h, ctypes.byref(w), ctypes.byref(c), ctypes.byref(t)))
        return w.value, c.value, bool(t.value)

    def set_replay(self, on: bool = True) -> None:
        """Capture the count of the unchanged graph as a hipGraph and replay it on later counts."""
        _check(lib().hm_graph_set_replay(self._h, int(bool(on))))

    def replay_state(self) -> tuple[bool, bool]:
        """(a capture of the current graph is held, the last count replayed it)."""
        cap, last = ctypes.c_int32(0), ctypes.c_int32(0)
        _check(lib().hm_graph_replay_state(self._h, ctypes.byref(cap), ctypes.byref(last)))
        return bool(cap.value), bool(last.value)

    def set_update_capture(self, on: bool = True) -> None:
        """Run every later update with a delta as one captured hipGraph (one launch, one wait),
        its buffers sized to a plan with headroom that grows when a batch does not fit."""
        _check(lib().hm_graph_set_update_capture(self._h, int(bool(on))))

    def update_capture_state(self) -> dict:
        """Plan of the captured update: on / disabled / captured, captures made, whether the
        last update ran captured (and overflowed its rows), and the plan's capacities."""
        st = _UpdateCaptureState()
        _check(lib().hm_graph_update_capture_state(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _UpdateCaptureState._fields_}

    def update_stats(self) -> dict:
        """Of the last update with a delta: removed / added triples (held a deleted hyperedge
        before, hold an inserted one after), row entries the delta kernel scanned, the
        milliseconds of the row build, the delta kernel and the store update, and whether it
        ran as the captured graph (then ms_delta holds the whole launch)."""
        st = _UpdateStats()
        _check(lib().hm_graph_update_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _UpdateStats._fields_}


class _UpdateStats(ctypes.Structure):
    _fields_ = [("removed", ctypes.c_int64), ("added", ctypes.c_int64),
                ("row_entries", ctypes.c_int64), ("ms_rows", ctypes.c_double),
                ("ms_delta", ctypes.c_double), ("ms_store", ctypes.c_double),
                ("captured", ctypes.c_int32), ("rows_overflow", ctypes.c_int32)]


class _UpdateCaptureState(ctypes.Structure):
    _fields_ = [("on", ctypes.c_int32), ("disabled", ctypes.c_int32),
                ("captured", ctypes.c_int32), ("captures", ctypes.c_int32),
                ("last_captured", ctypes.c_int32), ("last_overflow", ctypes.c_int32),
                ("id_capacity", ctypes.c_int64), ("node_capacity", ctypes.c_int64),
                ("incidence_capacity", ctypes.c_int64), ("row_capacity", ctypes.c_int64),
                ("batch_deletes", ctypes.c_int32), ("batch_inserts", ctypes.c_int32),
                ("batch_incidences", ctypes.c_int64)]


def count_delta(old: Graph, new: Graph, deleted, inserted) -> np.ndarray:
    """counts(new) - counts(old) (int64[30]) from the deleted / inserted ids of two snapshots."""
    d = np.ascontiguousarray(np.asarray(deleted, dtype=np.int32).reshape(-1))
    i = np.ascontiguousarray(np.asarray(inserted, dtype=np.int32).reshape(-1))
    out = np.zeros(STATIC_BINS, dtype=np.int64)
    _check(lib().hm_count_delta(old._h, new._h, _ptr(d if d.size else np.zeros(1, np.int32), ctypes.c_int32),
                                d.size, _ptr(i if i.size else np.zeros(1, np.int32), ctypes.c_int32),
                                i.size, _ptr(out, ctypes.c_int64)))
    return out


def count_strict_inc(older, middle, newest) -> np.ndarray:
    """The reference's three-snapshot strict-increasing count (30 static bins, uint64): triples
    (i, j, k), i < j < k, of older row i, middle row j, newest row k, pairwise overlapping. Each
    snapshot is (offsets int64[n+1], nodes int32[...]) indexed by row id."""
    offs = [np.ascontiguousarray(l[0], dtype=np.int64) for l in (older, middle, newest)]
    nods = [np.ascontiguousarray(l[1], dtype=np.int32) for l in (older, middle, newest)]
    nods = [n if n.size else np.zeros(1, dtype=np.int32) for n in nods]
    po = (ctypes.c_void_p * 3)(*[o.ctypes.data for o in offs])
    pn = (ctypes.c_void_p * 3)(*[n.ctypes.data for n in nods])
    nr = (ctypes.c_int32 * 3)(*[o.size - 1 for o in offs])
    out = np.zeros(STATIC_BINS, dtype=np.uint64)
    _check(lib().hm_count_strict_inc(po, pn, nr, _ptr(out, ctypes.c_uint64)))
    return out


class _TemporalStats(ctypes.Structure):
    _fields_ = [
        ("num_edges", ctypes.c_int64),
        ("num_incidences", ctypes.c_int64),
        ("proj_pairs", ctypes.c_int64),
        ("max_window", ctypes.c_int64),
        ("neighbour_pairs", ctypes.c_int64),
        ("instances", ctypes.c_int64),
        ("oriented_row_sum", ctypes.c_int64),
        ("num_nodes", ctypes.c_int64),
    ]


class TemporalGraph:
    """Temporal hyperedges (node sets with int64 timestamps) resident in HBM, in time order."""

    def __init__(self, offsets: np.ndarray, nodes: np.ndarray, times: np.ndarray):
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        times = np.ascontiguousarray(times, dtype=np.int64)
        if nodes.size == 0:
            nodes = np.zeros(1, dtype=np.int32)
        if times.size == 0:
            times = np.zeros(1, dtype=np.int64)
        self._h = ctypes.c_void_p()
        _check(lib().hm_tgraph_create(_ptr(offsets, ctypes.c_int64), _ptr(nodes, ctypes.c_int32),
                                      _ptr(times, ctypes.c_int64), offsets.size - 1,
                                      ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().hm_tgraph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def count(self, delta: int) -> np.ndarray:
        """Exact temporal h-motif counts within delta: uint64[96] in THyMe order (motif i+1)."""
        out = np.zeros(TEMPORAL_BINS, dtype=np.uint64)
        _check(lib().hm_count_temporal(self._h, int(delta), _ptr(out, ctypes.c_uint64)))
        return out

    def stats(self) -> dict:
        st = _TemporalStats()
        _check(lib().hm_tgraph_get_stats(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in _TemporalStats._fields_}

    def set_tile(self, tile_positions: int) -> None:
        """Position tile of count: 0 = automatic, > 0 = forced tiles of that many positions."""
        _check(lib().hm_tgraph_set_tile(self._h, int(tile_positions)))

    def phase_ms(self, phase: str) -> float:
        """Milliseconds of the last count's phase: "project" (window bounds and temporal
        projection rows), "count" (the counting kernel) or "wedge" (k_twedge)."""
        ms = ctypes.c_double(0.0)
        _check(lib().hm_tgraph_phase_ms(self._h, phase.encode(), ctypes.byref(ms)))
        return ms.value


class _BatchStats(ctypes.Structure):
    _fields_ = [("graphs_batched", ctypes.c_int32), ("graphs_alone", ctypes.c_int32),
                ("waves", ctypes.c_int32), ("copies", ctypes.c_int32),
                ("persistent", ctypes.c_int32), ("order", ctypes.c_int32), ("tuned", ctypes.c_int32),
                ("hyperedges", ctypes.c_int64), ("ms_prepare", ctypes.c_double),
                ("ms_triangles", ctypes.c_double)]


class Batch:
    """Static counts of many graphs in one launch sequence (hm_batch_* in hmotif.h). The
    triangle kernel's form (waves per centre, LDS histogram copies, persistent work queue or one
    workgroup per centre) and centre order are autotuned per batch shape unless set_form fixes
    them."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        _check(lib().hm_batch_create(ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().hm_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def count_static(self, graphs) -> np.ndarray:
        """uint64[len(graphs), 30]: the exact static counts of every graph."""
        n = len(graphs)
        arr = (ctypes.c_void_p * max(n, 1))(*[g._h.value for g in graphs])
        out = np.zeros((max(n, 1), STATIC_BINS), dtype=np.uint64)
        _check(lib().hm_batch_count_static(self._h, arr, n, _ptr(out, ctypes.c_uint64)))
        return out[:n]

    def set_form(self, waves: int = 0, copies: int = 16, persistent: bool = False,
                 order: int = 2) -> None:
        """waves=0: autotuned; else waves 4 or 8, copies 8 or 16, persistent work queue or not,
        centre order 0 (graph by graph), 1 (chunks of 64 MB of rows) or 2 (whole batch by work)."""
        _check(lib().hm_batch_set_form(self._h, int(waves), int(copies), int(bool(persistent)),
                                       int(order)))

    def stats(self) -> dict:
        st = _BatchStats()
        _check(lib().hm_batch_get_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _BatchStats._fields_}


class _ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("slots", ctypes.c_int32),
                ("maintain_counts", ctypes.c_int32), ("batch_graphs", ctypes.c_int32),
                ("capture_updates", ctypes.c_int32), ("hub_min_edges", ctypes.c_int32)]


class _Result(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("kind", ctypes.c_int32),
        ("counts", ctypes.c_int64 * STATIC_BINS),
        ("queued_ms", ctypes.c_double),
        ("run_ms", ctypes.c_double),
        ("latency_ms", ctypes.c_double),
        ("first_new_id", ctypes.c_int32),
        ("n_new", ctypes.c_int32),
    ]


class _ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in
                ("graphs", "reserved_bytes", "budget_bytes", "completed", "coalesced",
                 "maintained", "batched_counts", "batch_launches")]


class Server:
    """Request loop over many device-resident hypergraphs (see hm_server_* in hmotif.h).
    maintain_counts=True keeps each graph's counts current from the exact update deltas after
    its first count and answers count queries from them without a launch; batch_graphs=n > 1
    lets a worker count up to n ready graphs in one batched launch sequence; capture_updates=True
    runs every graph's updates as one captured hipGraph each; hub_min_edges=n > 0 counts and
    updates every graph of at least n hyperedges hub-aware (Graph.set_hubs(-1))."""

    COUNT, UPDATE = 0, 1

    def __init__(self, hbm_budget_bytes: int, slots: int = 4, maintain_counts: bool = False,
                 batch_graphs: int = 0, capture_updates: bool = False, hub_min_edges: int = 0):
        cfg = _ServerConfig(int(hbm_budget_bytes), int(slots), int(bool(maintain_counts)),
                            int(batch_graphs), int(bool(capture_updates)), int(hub_min_edges))
        self._h = ctypes.c_void_p()
        _check(lib().hm_server_create(ctypes.byref(cfg), ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().hm_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_graph(self, offsets: np.ndarray, nodes: np.ndarray) -> int:
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        nodes = np.ascontiguousarray(nodes if nodes.size else np.zeros(1), dtype=np.int32)
        gid = ctypes.c_int32(-1)
        _check(lib().hm_server_add_graph(self._h, _ptr(offsets, ctypes.c_int64),
                                         _ptr(nodes, ctypes.c_int32), offsets.size - 1,
                                         ctypes.byref(gid)))
        return gid.value

    def remove_graph(self, gid: int) -> None:
        _check(lib().hm_server_remove_graph(self._h, gid))

    def submit_count(self, gid: int) -> int:
        t = ctypes.c_uint64(0)
        _check(lib().hm_server_submit_count(self._h, gid, ctypes.byref(t)))
        return t.value

    def submit_update(self, gid: int, deleted, ins_offsets, ins_nodes) -> int:
        d = np.ascontiguousarray(np.asarray(deleted, dtype=np.int32).reshape(-1))
        o = np.ascontiguousarray(ins_offsets, dtype=np.int64)
        n = np.ascontiguousarray(ins_nodes if np.size(ins_nodes) else np.zeros(1), dtype=np.int32)
        t = ctypes.c_uint64(0)
        _check(lib().hm_server_submit_update(
            self._h, gid, _ptr(d if d.size else np.zeros(1, np.int32), ctypes.c_int32), d.size,
            _ptr(o, ctypes.c_int64), _ptr(n, ctypes.c_int32), o.size - 1, ctypes.byref(t)))
        return t.value

    @staticmethod
    def _result(r) -> dict:
        return {"kind": r.kind, "counts": np.array(r.counts[:], dtype=np.int64),
                "queued_ms": r.queued_ms, "run_ms": r.run_ms, "latency_ms": r.latency_ms,
                "first_new_id": r.first_new_id, "n_new": r.n_new}

    def wait(self, ticket: int) -> dict:
        r = _Result()
        _check(lib().hm_server_wait(self._h, ticket, ctypes.byref(r)))
        return self._result(r)

    def poll(self, ticket: int):
        """None while the request runs, else its result (the ticket is then consumed)."""
        r, done = _Result(), ctypes.c_int32(0)
        _check(lib().hm_server_poll(self._h, ticket, ctypes.byref(done), ctypes.byref(r)))
        return self._result(r) if done.value else None

    def stats(self) -> dict:
        s = _ServerStats()
        _check(lib().hm_server_get_stats(self._h, ctypes.byref(s)))
        return {n: getattr(s, n) for n, _ in _ServerStats._fields_}


def printed_motifs(counts30) -> list[int]:
    """MoCHy's 26 printed motifs (motif 1..26) from the 30 bins (MoCHy-master/main_exact.cpp:133-137)."""
    return [int(counts30[b]) for b in PRINTED_STATIC_BINS]


def algorithmic_bytes_temporal(st: dict) -> int:
    """The static byte model over the delta-window projection, plus 8 B per timestamp and the
    96-bin histogram: 4(2I + E + V + 2) + 16 P + 8 * sum_{x < y overlapping} d_y + 8 E + 768,
    with P = 2 * proj_pairs directed pairs."""
    return (4 * (2 * st["num_incidences"] + st["num_edges"] + st["num_nodes"] + 2)
            + 16 * 2 * st["proj_pairs"] + 8 * st["oriented_row_sum"] + 8 * st["num_edges"] + 768)


def algorithmic_bytes_static(st: dict) -> int:
    """Bytes a static count must move at minimum under the engine's cost model:
    4(2I + E + V + 2) (CSR in) + 16 P (projection written and read at 8 B per pair)
    + 8 * sum_a sum_{b in N(a), b > a} d_b (neighbour rows read for closed triples) + 240 (histogram)."""
    return (4 * (2 * st["num_incidences"] + st["num_edges"] + st["num_nodes"] + 2)
            + 16 * st["proj_pairs"] + 8 * st["oriented_row_sum"] + 240)
