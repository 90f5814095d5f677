"""ctypes binding of oracle/liboracle.so (the CPU restatement used as the test oracle)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")

_P = ctypes.c_void_p
_I32, _I64 = ctypes.c_int32, ctypes.c_int64

_SIG = {
    "orc_graph_open": ([ctypes.c_char_p], _P),
    "orc_graph_from_csr": ([_I32, _P, _P], _P),
    "orc_graph_free": ([_P], None),
    "orc_n": ([_P], _I32),
    "orc_base_edges": ([_P], _I64),
    "orc_classify": ([_P, _P, _I64, _P], None),
    "orc_apply": ([_P, _P, _P, _I64], None),
    "orc_apply_conditional": ([_P, _P, _P, _I64], None),
    "orc_is_adjacent": ([_P, _I32, _I32], ctypes.c_int),
    "orc_degree": ([_P, _I32], _I64),
    "orc_row": ([_P, _I32, _P, _I64], _I64),
    "orc_logical_csr": ([_P, _P, _P], _I64),
    "orc_compact_to": ([_P, ctypes.c_char_p], ctypes.c_int),
    "orc_bfs": ([_P, _I32, _P], None),
    "orc_cc": ([_P, _P], _I32),
    "orc_color_batch": ([_P, _P, _P, _I64, _P, _P], _I64),
    "orc_mis_max_csr": ([ctypes.c_double, ctypes.c_double], _I64),
    "orc_mis_batch": ([_P, _P, _P, _I64, _I64, _P, _P], ctypes.c_int),
    "orc_check_bfs": ([_P, _I32, _P], _I64),
    "orc_check_cc": ([_P, _P], _I64),
    "orc_check_coloring": ([_P, _P], _I64),
    "orc_check_mis_independent": ([_P, _P], _I64),
    "orc_mis_maximality_violations": ([_P, _P], _I64),
    "orc_mis_complete": ([_P, _P, _I64, _P], _I64),
    "orc_greedy_mis": ([_P, _P], None),
    "orc_greedy_color": ([_P, _P], None),
    "orc_refbfs_create": ([_P, _P, _I32, _I64, ctypes.c_int], _P),
    "orc_refbfs_free": ([_P], None),
    "orc_refbfs_batch": ([_P, _P, _P, _I64, _P], None),
    "orc_refbfs_finish": ([_P, _P], None),
    "orc_refcc_create": ([_P, _P, _I64, ctypes.c_int], _P),
    "orc_refcc_free": ([_P], None),
    "orc_refcc_batch": ([_P, _P, _P, _I64, _P], None),
    "orc_refcc_finish": ([_P, _P], None),
    "orc_mlp_loss_grad": ([_P, _P, _P, _I64, _P], ctypes.c_double),
    "orc_mlp_forward": ([_P, _P, _I64, _P], None),
    "orc_bf16_round": ([ctypes.c_float], ctypes.c_float),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(ORACLE_SO)
        for name, (args, res) in _SIG.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleGraph:
    """Base graph + scratchpad overlay with the reference's DeltaOverlay semantics."""

    def __init__(self, path=None, n=None, off=None, col=None):
        L = lib()
        if path is not None:
            self.h = L.orc_graph_open(path.encode())
        else:
            off = np.ascontiguousarray(off, dtype=np.int64)
            col = np.ascontiguousarray(col, dtype=np.int32)
            self.h = L.orc_graph_from_csr(int(n), _p(off), _p(col))
        if not self.h:
            raise RuntimeError("oracle could not open the graph")
        self.n = L.orc_n(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_graph_free(self.h)
            self.h = None

    def classify(self, uv):
        uv = np.ascontiguousarray(uv, dtype=np.int32).reshape(-1)
        ins = np.zeros(max(uv.size // 2, 1), dtype=np.uint8)
        lib().orc_classify(self.h, _p(uv), uv.size // 2, _p(ins))
        return ins[: uv.size // 2]

    def apply(self, uv, ins, conditional=False):
        uv = np.ascontiguousarray(uv, dtype=np.int32).reshape(-1)
        ins = np.ascontiguousarray(ins, dtype=np.uint8)
        f = lib().orc_apply_conditional if conditional else lib().orc_apply
        f(self.h, _p(uv), _p(ins), uv.size // 2)

    def is_adjacent(self, u, v):
        return bool(lib().orc_is_adjacent(self.h, u, v))

    def degree(self, v):
        return int(lib().orc_degree(self.h, v))

    def row(self, v):
        cap = 1 << 16
        out = np.zeros(cap, dtype=np.int32)
        k = lib().orc_row(self.h, v, _p(out), cap)
        if k > cap:
            out = np.zeros(k, dtype=np.int32)
            lib().orc_row(self.h, v, _p(out), k)
        return out[:k].copy()

    def logical_csr(self):
        off = np.zeros(self.n + 1, dtype=np.int64)
        e = lib().orc_logical_csr(self.h, _p(off), None)
        assert e >= 0, "inconsistent delta"
        col = np.zeros(max(e, 1), dtype=np.int32)
        assert lib().orc_logical_csr(self.h, _p(off), _p(col)) >= 0, "degree mismatch (compactor would exit)"
        return off, col[:e]

    def bfs(self, src=0):
        out = np.zeros(self.n, dtype=np.int32)
        lib().orc_bfs(self.h, src, _p(out))
        return out

    def cc(self):
        out = np.zeros(self.n, dtype=np.int32)
        lib().orc_cc(self.h, _p(out))
        return out

    def color_batch(self, uv, ins, color):
        uv = np.ascontiguousarray(uv, dtype=np.int32).reshape(-1)
        ins = np.ascontiguousarray(ins, dtype=np.uint8)
        color = np.ascontiguousarray(color, dtype=np.int32).copy()
        st = np.zeros(2, dtype=np.int64)
        lib().orc_color_batch(self.h, _p(uv), _p(ins), uv.size // 2, _p(color), _p(st))
        return color, st

    def mis_batch(self, uv, ins, mis, max_csr):
        uv = np.ascontiguousarray(uv, dtype=np.int32).reshape(-1)
        ins = np.ascontiguousarray(ins, dtype=np.uint8)
        mis = np.ascontiguousarray(mis, dtype=np.uint8).copy()
        st = np.zeros(3, dtype=np.int64)
        rc = lib().orc_mis_batch(self.h, _p(uv), _p(ins), uv.size // 2, int(max_csr), _p(mis), _p(st))
        assert rc == 0, "MIS sub-batch exceeds the reference's row pool"
        return mis, st

    def check_bfs(self, level, src=0):
        return int(lib().orc_check_bfs(self.h, src, _p(np.ascontiguousarray(level, dtype=np.int32))))

    def check_cc(self, label):
        return int(lib().orc_check_cc(self.h, _p(np.ascontiguousarray(label, dtype=np.int32))))

    def check_coloring(self, color):
        return int(lib().orc_check_coloring(self.h, _p(np.ascontiguousarray(color, dtype=np.int32))))

    def check_mis_independent(self, mis):
        return int(lib().orc_check_mis_independent(self.h, _p(np.ascontiguousarray(mis, dtype=np.uint8))))

    def mis_complete(self, uv, mis):
        """MIS complete mode: greedy ascending completion over the batch's 2-hop region."""
        uv = np.ascontiguousarray(uv, dtype=np.int32).reshape(-1)
        mis = np.ascontiguousarray(mis, dtype=np.uint8).copy()
        added = lib().orc_mis_complete(self.h, _p(uv), uv.size // 2, _p(mis))
        return mis, int(added)

    def mis_maximality_violations(self, mis):
        return int(lib().orc_mis_maximality_violations(self.h, _p(np.ascontiguousarray(mis, dtype=np.uint8))))

    def greedy_mis(self):
        out = np.zeros(self.n, dtype=np.uint8)
        lib().orc_greedy_mis(self.h, _p(out))
        return out

    def greedy_color(self):
        out = np.zeros(self.n, dtype=np.int32)
        lib().orc_greedy_color(self.h, _p(out))
        return out


def mis_max_csr(avg_degree, free_bytes):
    return int(lib().orc_mis_max_csr(float(avg_degree), float(free_bytes)))


def read_batch(path, n):
    """Batch file -> (m, 2) int32 0-indexed pairs with the reference's line rules."""
    out = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) < 2:
                continue
            try:
                u, v = int(parts[0]) - 1, int(parts[1]) - 1
            except ValueError:
                continue
            if 0 <= u < n and 0 <= v < n and u != v:
                out.append((u, v))
    return np.array(out, dtype=np.int32).reshape(-1, 2)


def mlp_loss_grad(params, feats, targets, with_grad=True):
    """Mean loss (double) and its gradient from the oracle's double-precision MLP."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    f = np.ascontiguousarray(feats, dtype=np.float32)
    t = np.ascontiguousarray(targets, dtype=np.float32)
    g = np.zeros(p.size, dtype=np.float64) if with_grad else None
    loss = lib().orc_mlp_loss_grad(_p(p), _p(f), _p(t), len(f), _p(g) if with_grad else None)
    return loss, g


def mlp_forward(params, feats):
    p = np.ascontiguousarray(params, dtype=np.float32)
    f = np.ascontiguousarray(feats, dtype=np.float32)
    out = np.zeros((len(f), 3), dtype=np.float64)
    lib().orc_mlp_forward(_p(p), _p(f), len(f), _p(out))
    return out


def bf16_round(x):
    return np.array([lib().orc_bf16_round(float(v)) for v in np.ravel(x)], dtype=np.float32).reshape(np.shape(x))

