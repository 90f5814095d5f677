"""scout_amd: Python binding of libscout_hip.so, the MI355X SCOUT repair engine.

The product path is the C ABI declared in include/scout_hip.h; this module loads the
in-tree library with ctypes and exposes it as an ``Engine`` object whose methods mirror the
reference's per-batch flow (src/*/scout_*.cu processBatch): stage a batch, classify its
lines, apply them to the scratchpad, repair, read the solution back.

There is no CPU fallback: if the library is missing the import of ``Engine`` fails loudly.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ROOT, "libscout_hip.so")
BIN_DIR = os.path.join(ROOT, "bin")

OK = 0
STATE_MIS, STATE_COLOR, STATE_LEVEL, STATE_LABEL = 0, 1, 2, 3
CLASSIFY_CURRENT, CLASSIFY_BASE = 0, 1
STRATEGY_LOCAL, STRATEGY_RECOMPUTE, STRATEGY_LEARNED = 0, 1, 2
MIS_RULES, MIS_COMPLETE = 0, 1
CC_AUTO, CC_WALK, CC_SCAN = 0, 1, 2
BFS_ASYNC, BFS_LEVEL_SYNC, BFS_ASYNC_INVALIDATION = 0, 1, 2
GRAPH_INFO_COUNT = 8  # entries scout_graph_info writes


class ScoutError(RuntimeError):
    pass


class Stats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "lines", "insertions", "deletions", "noop", "seeds", "touched", "rounds", "fallback",
        "splits", "conflicts", "sub_batches", "edges_scanned")] + [
        ("repair_ms", ctypes.c_double), ("apply_ms", ctypes.c_double), ("readbacks", ctypes.c_int64),
        ("inval_rounds", ctypes.c_int64), ("relax_rounds", ctypes.c_int64), ("completed", ctypes.c_int64),
        ("cc_mode", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64

# name -> argument types (all functions return int status except scout_last_error)
_SIGNATURES = {
    "scout_ctx_create": [ctypes.c_int, _P, ctypes.POINTER(_P)],
    "scout_ctx_destroy": [_P],
    "scout_device_name": [_P, ctypes.c_char_p, ctypes.c_int],
    "scout_mem_info": [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64)],
    "scout_graph_load": [_P, _I32, _I64, _P, ctypes.c_int, _P],
    "scout_graph_info": [_P, _P],
    "scout_graph_download": [_P, _P, _P],
    "scout_graph_compact": [_P],
    "scout_graph_degrees": [_P, _P],
    "scout_state_upload": [_P, ctypes.c_int, _P, _I64],
    "scout_state_download": [_P, ctypes.c_int, _P, _I64],
    "scout_set_source": [_P, _I32],
    "scout_stage_batch": [_P, ctypes.c_int, _P, _I64],
    "scout_stage_batch_device": [_P, ctypes.c_int, _P, _I64],
    "scout_classify": [_P, ctypes.c_int, ctypes.c_int, _P],
    "scout_apply": [_P, ctypes.c_int, ctypes.c_int],
    "scout_repair_bfs": [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Stats)],
    "scout_repair_cc": [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Stats)],
    "scout_commit": [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Stats), ctypes.POINTER(Stats)],
    "scout_repair_color": [_P, ctypes.c_int, ctypes.POINTER(Stats)],
    "scout_repair_mis": [_P, ctypes.c_int, _I64, ctypes.POINTER(Stats)],
    "scout_repair_mis_multi": [_P, ctypes.c_int, _I32, _P, _I64, ctypes.POINTER(Stats)],
    "scout_mis_config": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "scout_mis_mode": [_P, ctypes.c_int],
    "scout_cc_mode": [_P, ctypes.c_int],
    "scout_mis_limits": [_P, _I64, _I64],
    "scout_mis_get_config": [_P, _P],
    "scout_mis_autotune": [_P, ctypes.c_int, _I32, _P, _I64, ctypes.POINTER(ctypes.c_double)],
    "scout_color_config": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "scout_color_get_config": [_P, _P],
    "scout_color_autotune": [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
    "scout_bfs_full": [_P, _I32],
    "scout_cc_full": [_P],
    "scout_cc_full_noalloc": [_P, _P],
    "scout_mis_greedy": [_P],
    "scout_color_greedy": [_P],
    "scout_sync": [_P],
    "scout_gen_rmat": [_P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                       ctypes.c_uint64],
    "scout_gen_batch": [_P, ctypes.c_int, _I64, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _I64,
                        ctypes.POINTER(_I64)],
    "scout_batch_download": [_P, ctypes.c_int, _P, _I64],
    "scout_gen_mode": [_P, ctypes.c_int],
    "scout_batch_store": [_P, ctypes.c_int, ctypes.c_int],
    "scout_batch_load": [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_I64)],
    "scout_snapshot": [_P],
    "scout_restore": [_P],
    "scout_prof_enable": [_P, ctypes.c_int],
    "scout_prof_read": [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I64),
                        ctypes.POINTER(_I64)],
    "scout_batch_append": [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_I64)],
    "scout_path_counts": [_P, _P, ctypes.c_int],
    "scout_stage_info": [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(_I64)],
    "scout_bfs_bound": [_P, _I64],
    "scout_bfs_schedule": [_P, ctypes.c_int, ctypes.c_int],
    "scout_bfs_mode": [_P, ctypes.c_int],
    "scout_touched_enable": [_P, ctypes.c_int],
    "scout_touched_download": [_P, _P, _I64, ctypes.POINTER(_I64)],
    "scout_apply_segments": [_P, ctypes.c_int, _I32, _P, ctypes.c_int, ctypes.c_int, _P],
    "scout_mlp_init": [_P, ctypes.c_uint64],
    "scout_mlp_set_params": [_P, _P],
    "scout_mlp_get_params": [_P, _P, ctypes.c_int],
    "scout_mlp_forward": [_P, _P, _I64, _P],
    "scout_mlp_grad": [_P, _P, _P, _I64, _P, ctypes.POINTER(ctypes.c_float)],
    "scout_mlp_log": [_P, ctypes.c_int],
    "scout_mlp_samples": [_P, _P, _I64, ctypes.POINTER(_I64)],
    "scout_mlp_train": [_P, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, ctypes.POINTER(ctypes.c_float),
                        ctypes.POINTER(ctypes.c_float)],
}

MLP_FEATURES, MLP_OUTPUTS, MLP_PARAMS = 12, 3, 5187
GEN_UNIFORM, GEN_DEGREE = 0, 1

PROF_BFS_RELAX, PROF_BFS_INVALIDATE, PROF_CC_EXPAND, PROF_APPLY, PROF_CLASSIFY = 0, 1, 2, 3, 4
PROF_COLOR, PROF_MIS_EVAL, PROF_CC_HOOK, PROF_BFS_INVALIDATE_KERNEL = 5, 6, 7, 8
PROF_NAMES = ["bfs_relax", "bfs_invalidate", "cc_expand", "apply", "classify", "color", "mis_eval", "cc_hook",
              "bfs_invalidate_kernel", "k_relax_short", "k_relax_rows", "k_invalidate_short", "k_inval_push",
              "k_cc_explore", "k_cc_collect"]
PATH_NAMES = ["inval_tasks", "inval_push_tasks", "inval_kids_overflow", "cc_tasks", "relax_tasks",
              "tentative_tasks", "cc_walk_rows", "inval_deferred"]
LONG_ROW_PATHS = PATH_NAMES[:7]

_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libscout_hip.so and sets every exported signature. Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ScoutError(f"{path} is not built; run __graft_entry__.build() or make -C scout_amd")
    lib = ctypes.CDLL(path)
    for name, argtypes in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    lib.scout_last_error.argtypes = []
    lib.scout_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def declared_symbols(header: str) -> list[str]:
    """Function names declared in a C header of this repository (for export checks)."""
    import re
    text = open(header).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(scout_\w+)\(", text, re.M)))


@dataclass
class Graph:
    n: int
    off: np.ndarray   # int64, n+1
    col: np.ndarray   # int32

    @staticmethod
    def read_egr(path: str) -> "Graph":
        raw = np.fromfile(path, dtype=np.uint8)
        n, e32 = np.frombuffer(raw[:8].tobytes(), dtype=np.int32)
        if e32 == -1:
            e = int(np.frombuffer(raw[8:16].tobytes(), dtype=np.uint64)[0])
            off = np.frombuffer(raw[16:16 + 8 * (n + 1)].tobytes(), dtype=np.int64)
            base = 16 + 8 * (n + 1)
        else:
            e = int(e32)
            off = np.frombuffer(raw[8:8 + 4 * (n + 1)].tobytes(), dtype=np.int32).astype(np.int64)
            base = 8 + 4 * (n + 1)
        col = np.frombuffer(raw[base:base + 4 * e].tobytes(), dtype=np.int32)
        return Graph(int(n), np.ascontiguousarray(off), np.ascontiguousarray(col))


class Engine:
    """One device context (scout_ctx) holding a graph and its maintained solutions."""

    def __init__(self, gpu: int = 0, hbm_bytes: int = 0, pinned_bytes: int = 0):
        self.lib = load_library()
        budget = (ctypes.c_int64 * 2)(hbm_bytes, pinned_bytes)
        h = ctypes.c_void_p()
        self._check(self.lib.scout_ctx_create(gpu, ctypes.cast(budget, ctypes.c_void_p), ctypes.byref(h)),
                    "ctx_create")
        self.h = h
        self.n = 0
        self._staged = {0: 0, 1: 0}  # lines staged per channel (sizes classify's output)

    def _check(self, rc: int, what: str):
        if rc != OK:
            msg = self.lib.scout_last_error().decode(errors="replace")
            raise ScoutError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.scout_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # graph
    def load_graph(self, g: Graph):
        off = np.ascontiguousarray(g.off, dtype=np.int64)
        col = np.ascontiguousarray(g.col, dtype=np.int32)
        self._check(self.lib.scout_graph_load(self.h, g.n, int(off[-1]), _ptr(off), 1, _ptr(col)), "graph_load")
        self.n = g.n

    def gen_rmat(self, scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1):
        self._check(self.lib.scout_gen_rmat(self.h, scale, edge_factor, a, b, c, seed), "gen_rmat")
        self.n = 1 << scale

    def mem_info(self):
        """(free, total) device memory in bytes."""
        free, total = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.scout_mem_info(self.h, ctypes.byref(free), ctypes.byref(total)), "mem_info")
        return free.value, total.value

    def graph_info(self) -> dict:
        out = np.zeros(GRAPH_INFO_COUNT, dtype=np.int64)
        self._check(self.lib.scout_graph_info(self.h, _ptr(out)), "graph_info")
        return {"n": int(out[0]), "base_entries": int(out[1]), "inserted_entries": int(out[2]),
                "logical_entries": int(out[3]), "sorted": bool(out[4]), "host_resident": bool(out[5]),
                "fetch_bytes": int(out[6]), "fetch_us": int(out[7])}

    def download_graph(self) -> Graph:
        off = np.zeros(self.n + 1, dtype=np.int64)
        self._check(self.lib.scout_graph_download(self.h, _ptr(off), None), "graph_download")
        col = np.zeros(int(off[-1]), dtype=np.int32)
        self._check(self.lib.scout_graph_download(self.h, _ptr(off), _ptr(col)), "graph_download")
        return Graph(self.n, off, col)

    def compact(self):
        self._check(self.lib.scout_graph_compact(self.h), "graph_compact")

    def degrees(self) -> np.ndarray:
        d = np.zeros(self.n, dtype=np.int64)
        self._check(self.lib.scout_graph_degrees(self.h, _ptr(d)), "graph_degrees")
        return d

    # state
    def upload(self, kind: int, arr: np.ndarray):
        dt = np.uint8 if kind == STATE_MIS else np.int32
        a = np.ascontiguousarray(arr, dtype=dt)
        self._check(self.lib.scout_state_upload(self.h, kind, _ptr(a), self.n), "state_upload")

    def download(self, kind: int) -> np.ndarray:
        a = np.zeros(self.n, dtype=np.uint8 if kind == STATE_MIS else np.int32)
        self._check(self.lib.scout_state_download(self.h, kind, _ptr(a), self.n), "state_download")
        return a

    # batches
    def stage(self, uv: np.ndarray, channel: int = 0):
        a = np.ascontiguousarray(uv, dtype=np.int32).reshape(-1)
        self._check(self.lib.scout_stage_batch(self.h, channel, _ptr(a), a.size // 2), "stage_batch")
        self._staged[channel] = a.size // 2

    def classify(self, rule: int = CLASSIFY_CURRENT, channel: int = 0, download: bool = True):
        """Labels the staged lines (1 = insertion). Returns the labels, or None with
        download=False (they stay on the device for apply and the repairs)."""
        if not download:
            self._check(self.lib.scout_classify(self.h, channel, rule, None), "classify")
            return None
        m = self._staged[channel]
        out = np.zeros(max(m, 1), dtype=np.uint8)
        self._check(self.lib.scout_classify(self.h, channel, rule, _ptr(out)), "classify")
        return out[:m]

    def apply(self, conditional: bool = False, channel: int = 0):
        self._check(self.lib.scout_apply(self.h, channel, 1 if conditional else 0), "apply")

    def repair_bfs(self, strategy: int = STRATEGY_LOCAL, channel: int = 0) -> Stats:
        st = Stats()
        self._check(self.lib.scout_repair_bfs(self.h, channel, strategy, ctypes.byref(st)), "repair_bfs")
        return st

    def repair_cc(self, strategy: int = STRATEGY_LOCAL, channel: int = 0) -> Stats:
        st = Stats()
        self._check(self.lib.scout_repair_cc(self.h, channel, strategy, ctypes.byref(st)), "repair_cc")
        return st

    def commit(self, strategy: int = STRATEGY_LOCAL, channel: int = 0):
        """classify (labels kept on the device), apply, repair_bfs and repair_cc of the staged
        lines in one call; returns the BFS and the CC stats."""
        sb, sc = Stats(), Stats()
        self._check(self.lib.scout_commit(self.h, channel, strategy, ctypes.byref(sb), ctypes.byref(sc)), "commit")
        return sb, sc

    def repair_color(self, channel: int = 0) -> Stats:
        st = Stats()
        self._check(self.lib.scout_repair_color(self.h, channel, ctypes.byref(st)), "repair_color")
        return st

    def repair_mis(self, max_csr: int, channel: int = 0) -> Stats:
        st = Stats()
        self._check(self.lib.scout_repair_mis(self.h, channel, max_csr, ctypes.byref(st)), "repair_mis")
        return st

    def repair_mis_multi(self, line_off, max_csr: int, channel: int = 0) -> Stats:
        """MIS repair of a staged batch whose lines are grouped by instance: lines
        [line_off[i], line_off[i+1]) belong to instance i (disjoint vertex sets)."""
        off = np.ascontiguousarray(line_off, dtype=np.int64)
        st = Stats()
        self._check(self.lib.scout_repair_mis_multi(self.h, channel, len(off) - 1, _ptr(off), max_csr,
                                                    ctypes.byref(st)), "repair_mis_multi")
        return st

    def mis_mode(self, mode: int):
        """MIS_RULES (the reference's rules) or MIS_COMPLETE (rules, then greedy completion over
        the batch's 2-hop region, which keeps the set maximal)."""
        self._check(self.lib.scout_mis_mode(self.h, mode), "mis_mode")

    def cc_mode(self, mode: int):
        """CC_AUTO (by batch size), CC_WALK (walk from the batch endpoints) or CC_SCAN (one pass
        over every vertex): how the local CC repair finds the vertices to relabel."""
        self._check(self.lib.scout_cc_mode(self.h, mode), "cc_mode")

    def mis_config(self, grid: int = 1024, group_rounds: int = 4, use_graph: bool = True):
        self._check(self.lib.scout_mis_config(self.h, grid, group_rounds, 1 if use_graph else 0), "mis_config")

    def mis_get_config(self):
        out = np.zeros(3, dtype=np.int32)
        self._check(self.lib.scout_mis_get_config(self.h, _ptr(out)), "mis_get_config")
        return {"grid": int(out[0]), "group_rounds": int(out[1]), "use_graph": bool(out[2])}

    def mis_autotune(self, line_off, max_csr: int, channel: int = 0) -> float:
        """Times the MIS repair of the staged, applied batch (lines grouped by instance) under
        each grid and round-group setting and keeps the fastest; the membership is left as it
        was. Returns the best time in ms."""
        off = np.ascontiguousarray(line_off, dtype=np.int64)
        best = ctypes.c_double()
        self._check(self.lib.scout_mis_autotune(self.h, channel, len(off) - 1, _ptr(off), max_csr,
                                                ctypes.byref(best)), "mis_autotune")
        return best.value

    def mis_limits(self, hop2_requests: int = 65536, hop2_tasks: int = 1 << 22):
        self._check(self.lib.scout_mis_limits(self.h, hop2_requests, hop2_tasks), "mis_limits")

    def color_config(self, thread_max: int = 8, block_min: int = 4096, grid: int = 1024, use_graph: bool = True):
        self._check(self.lib.scout_color_config(self.h, thread_max, block_min, grid, 1 if use_graph else 0),
                    "color_config")

    def color_get_config(self):
        out = np.zeros(4, dtype=np.int32)
        self._check(self.lib.scout_color_get_config(self.h, _ptr(out)), "color_get_config")
        return {"thread_max": int(out[0]), "block_min": int(out[1]), "grid": int(out[2]),
                "use_graph": bool(out[3])}

    def color_autotune(self, channel: int = 0) -> float:
        """Times the colouring repair of the staged, applied batch under each candidate setting
        and keeps the fastest; the colours are left unchanged. Returns the best time in ms."""
        best = ctypes.c_double()
        self._check(self.lib.scout_color_autotune(self.h, channel, ctypes.byref(best)), "color_autotune")
        return best.value

    def bfs_full(self, source: int = 0):
        self._check(self.lib.scout_bfs_full(self.h, source), "bfs_full")

    def cc_full(self):
        self._check(self.lib.scout_cc_full(self.h), "cc_full")

    def cc_full_noalloc(self, d_label: int):
        """Full CC recompute into a caller-owned device array (address d_label, n int32) with
        no device allocation; the context's own labels are left unchanged."""
        self._check(self.lib.scout_cc_full_noalloc(self.h, ctypes.c_void_p(d_label)), "cc_full_noalloc")

    def mis_greedy(self):
        self._check(self.lib.scout_mis_greedy(self.h), "mis_greedy")

    def color_greedy(self):
        self._check(self.lib.scout_color_greedy(self.h), "color_greedy")

    def sync(self):
        self._check(self.lib.scout_sync(self.h), "sync")

    # synthetic streams, snapshots, kernel timing
    def gen_batch(self, size, delete_frac=0.5, tree_frac=0.0, seed=1, expected_total=0, channel=0) -> int:
        m = ctypes.c_int64()
        self._check(self.lib.scout_gen_batch(self.h, channel, size, delete_frac, tree_frac, seed, expected_total,
                                             ctypes.byref(m)), "gen_batch")
        self._staged[channel] = m.value
        return m.value

    def gen_mode(self, mode: int):
        """GEN_UNIFORM (default) or GEN_DEGREE: endpoints proportional to degree + 1."""
        self._check(self.lib.scout_gen_mode(self.h, mode), "gen_mode")

    def batch_download(self, m, channel=0) -> np.ndarray:
        out = np.zeros((max(m, 1), 2), dtype=np.int32)
        self._check(self.lib.scout_batch_download(self.h, channel, _ptr(out), m), "batch_download")
        return out[:m]

    def batch_append(self, slot, channel=0):
        m = ctypes.c_int64()
        self._check(self.lib.scout_batch_append(self.h, channel, slot, ctypes.byref(m)), "batch_append")
        self._staged[channel] = m.value
        return m.value

    def apply_segments(self, seg_off, rule: int = CLASSIFY_CURRENT, conditional: bool = False, channel: int = 0):
        """Classifies and applies the staged lines as consecutive segments, each against the
        graph the earlier ones left; returns every line's label."""
        off = np.ascontiguousarray(seg_off, dtype=np.int64)
        m = self._staged[channel]
        out = np.zeros(max(m, 1), dtype=np.uint8)
        self._check(self.lib.scout_apply_segments(self.h, channel, len(off) - 1, _ptr(off), rule,
                                                  1 if conditional else 0, _ptr(out)), "apply_segments")
        return out[:m]

    def batch_store(self, slot, channel=0):
        self._check(self.lib.scout_batch_store(self.h, channel, slot), "batch_store")

    def batch_load(self, slot, channel=0):
        m = ctypes.c_int64()
        self._check(self.lib.scout_batch_load(self.h, channel, slot, ctypes.byref(m)), "batch_load")
        self._staged[channel] = m.value
        return m.value

    def snapshot(self):
        self._check(self.lib.scout_snapshot(self.h), "snapshot")

    def restore(self):
        self._check(self.lib.scout_restore(self.h), "restore")

    def prof_enable(self, on=True):
        self._check(self.lib.scout_prof_enable(self.h, 1 if on else 0), "prof_enable")

    def prof_read(self, kid: int) -> dict:
        ms, cnt, by = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.scout_prof_read(self.h, kid, ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(by)),
                    "prof_read")
        return {"ms": ms.value, "launches": cnt.value, "bytes": by.value}

    def stage_info(self, channel: int = 0) -> dict:
        """The channel's last upload: host pack time, copy-stream H2D time (ms) and bytes."""
        pack, h2d, by = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        self._check(self.lib.scout_stage_info(self.h, channel, ctypes.byref(pack), ctypes.byref(h2d),
                                              ctypes.byref(by)), "stage_info")
        return {"pack_ms": pack.value, "h2d_ms": h2d.value, "bytes": by.value}

    def bfs_schedule(self, inval_rounds: int, relax_rounds: int):
        """Round groups of the next BFS repair only (a perfect predictor's schedule)."""
        self._check(self.lib.scout_bfs_schedule(self.h, inval_rounds, relax_rounds), "bfs_schedule")

    def bfs_mode(self, mode: int):
        """BFS_ASYNC (every level's candidates and seeds in the first round) or BFS_LEVEL_SYNC (one
        level per round); the levels are the same."""
        self._check(self.lib.scout_bfs_mode(self.h, mode), "bfs_mode")

    def bfs_bound(self, bound: int):
        self._check(self.lib.scout_bfs_bound(self.h, bound), "bfs_bound")

    def touched_enable(self, on: bool = True):
        self._check(self.lib.scout_touched_enable(self.h, 1 if on else 0), "touched_enable")

    def touched(self) -> np.ndarray:
        cnt = ctypes.c_int64()
        self._check(self.lib.scout_touched_download(self.h, None, 0, ctypes.byref(cnt)), "touched_download")
        out = np.zeros(max(cnt.value, 1), dtype=np.int32)
        self._check(self.lib.scout_touched_download(self.h, _ptr(out), cnt.value, ctypes.byref(cnt)),
                    "touched_download")
        return out[:cnt.value]

    def path_counts(self) -> dict:
        """Cumulative counts of the long-row code paths (chunk tasks per kernel family) and of
        the vertices the asynchronous BFS invalidation made wait a round."""
        out = np.zeros(len(PATH_NAMES), dtype=np.int64)
        self._check(self.lib.scout_path_counts(self.h, _ptr(out), len(PATH_NAMES)), "path_counts")
        return {k: int(v) for k, v in zip(PATH_NAMES, out)}


    # --- learned scheduler -------------------------------------------------------------
    def mlp_init(self, seed: int = 1):
        self._check(self.lib.scout_mlp_init(self.h, seed), "mlp_init")

    def mlp_set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == MLP_PARAMS
        self._check(self.lib.scout_mlp_set_params(self.h, _ptr(p)), "mlp_set_params")

    def mlp_get_params(self, rounded: bool = False) -> np.ndarray:
        p = np.zeros(MLP_PARAMS, dtype=np.float32)
        self._check(self.lib.scout_mlp_get_params(self.h, _ptr(p), 1 if rounded else 0), "mlp_get_params")
        return p

    def mlp_forward(self, feats) -> np.ndarray:
        f = np.ascontiguousarray(feats, dtype=np.float32).reshape(-1, MLP_FEATURES)
        out = np.zeros((len(f), MLP_OUTPUTS), dtype=np.float32)
        self._check(self.lib.scout_mlp_forward(self.h, _ptr(f), len(f), _ptr(out)), "mlp_forward")
        return out

    def mlp_grad(self, feats, targets):
        f = np.ascontiguousarray(feats, dtype=np.float32).reshape(-1, MLP_FEATURES)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, MLP_OUTPUTS)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        self._check(self.lib.scout_mlp_grad(self.h, _ptr(f), _ptr(t), len(f), _ptr(g), ctypes.byref(loss)),
                    "mlp_grad")
        return g, loss.value

    def mlp_log(self, on: bool = True):
        self._check(self.lib.scout_mlp_log(self.h, 1 if on else 0), "mlp_log")

    def mlp_samples(self, download: bool = False):
        cnt = ctypes.c_int64()
        self._check(self.lib.scout_mlp_samples(self.h, None, 0, ctypes.byref(cnt)), "mlp_samples")
        if not download:
            return cnt.value
        rows = np.zeros((max(cnt.value, 1), MLP_FEATURES + MLP_OUTPUTS), dtype=np.float32)
        self._check(self.lib.scout_mlp_samples(self.h, _ptr(rows), cnt.value, ctypes.byref(cnt)), "mlp_samples")
        return rows[: cnt.value]

    def mlp_train(self, epochs: int = 4, lr: float = 1e-3, seed: int = 1):
        a, b = ctypes.c_float(), ctypes.c_float()
        self._check(self.lib.scout_mlp_train(self.h, epochs, lr, seed, ctypes.byref(a), ctypes.byref(b)), "mlp_train")
        return a.value, b.value


def mis_max_csr(avg_degree: float, free_bytes: float) -> int:
    """MAX_CSR_ELEMENTS of src/mis/scout_mis.cu:800-815."""
    est = int(1048576.0 * 2.0 * avg_degree * avg_degree)
    mx = min(800000000, est) if est > 0 else 50000000
    mx = max(mx, 1000000)
    fit = int(free_bytes / 4.0 / 8.0)
    if fit > 1000000 and mx > fit:
        mx = fit
    return mx
