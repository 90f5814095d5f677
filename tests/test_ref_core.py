"""The oracle's scratchpad restatement, the compactor and the report lines checked against
the reference's own host code (core/FlashGraph.h, core/flash_compactor.h, core/reporting.h),
compiled in place by oracle/ref_build.sh into oracle/_ref/libref_core.so.

Skipped where oracle/_ref is not built (it needs the reference checkout at build time).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle_lib
from oracle_lib import OracleGraph

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SO = os.path.join(REPO, "oracle", "_ref", "libref_core.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built (no reference checkout)")

_P, _I32, _I64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64


def _ref():
    L = ctypes.CDLL(REF_SO)
    sig = {
        "ref_open": ([ctypes.c_char_p], _P),
        "ref_close": ([_P], None),
        "ref_n": ([_P], _I32),
        "ref_is_adjacent": ([_P, _I32, _I32], ctypes.c_int),
        "ref_insert": ([_P, _I32, _I32], None),
        "ref_remove": ([_P, _I32, _I32], None),
        "ref_degree": ([_P, _I32], _I64),
        "ref_row": ([_P, _I32, _P, _I64], _I64),
        "ref_classify_apply": ([_P, _P, _I64, _P], None),
        "ref_compact": ([_P, ctypes.c_char_p], None),
        "ref_report_sample": ([_P, _I64], _I64),
    }
    for k, (a, r) in sig.items():
        getattr(L, k).argtypes = a
        getattr(L, k).restype = r
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class RefGraph:
    def __init__(self, path):
        self.L = _ref()
        self.h = self.L.ref_open(path.encode())
        self.n = self.L.ref_n(self.h)

    def close(self):
        self.L.ref_close(self.h)

    def classify_apply(self, uv):
        uv = np.ascontiguousarray(uv, dtype=np.int32)
        ins = np.zeros(max(len(uv), 1), dtype=np.uint8)
        self.L.ref_classify_apply(self.h, _p(uv), len(uv), _p(ins))
        return ins[: len(uv)]

    def row(self, v):
        out = np.zeros(1 << 14, dtype=np.int32)
        k = self.L.ref_row(self.h, v, _p(out), out.size)
        assert k <= out.size
        return out[:k].copy()


def _mixed_batches(og, rng, nb, m):
    """Batches with deletions of live edges, insertions of non-edges, re-insertions of deleted
    base edges, deletions of inserted edges, and pairs repeated in both orientations."""
    n = og.n
    history = []
    out = []
    for _ in range(nb):
        lines = []
        for _ in range(m):
            r = rng.random()
            if r < 0.35:
                u = int(rng.integers(n))
                row = og.row(u)
                if len(row):
                    lines.append((u, int(row[rng.integers(len(row))])))
            elif r < 0.65:
                u, v = map(int, rng.integers(n, size=2))
                if u != v:
                    lines.append((u, v))
            elif r < 0.85 and history:
                u, v = history[rng.integers(len(history))]
                lines.append((v, u) if rng.random() < 0.5 else (u, v))
            elif lines:
                lines.append(lines[rng.integers(len(lines))])
        uv = np.array(lines, dtype=np.int32).reshape(-1, 2)
        history.extend(map(tuple, uv.tolist()))
        out.append(uv)
        ins = og.classify(uv)
        og.apply(uv, ins)
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_scratchpad_restatement_equals_reference_hybrid_graph(sample_dir, seed):
    egr = os.path.join(sample_dir, "sample.egr")
    rng = np.random.default_rng(seed)
    gen = OracleGraph(egr)
    batches = _mixed_batches(gen, rng, nb=4, m=600)
    og, ref = OracleGraph(egr), RefGraph(egr)
    try:
        for uv in batches:
            ins_o = og.classify(uv)
            og.apply(uv, ins_o)
            ins_r = ref.classify_apply(uv)
            assert np.array_equal(ins_o, ins_r)
            for v in range(og.n):
                assert np.array_equal(og.row(v), ref.row(v)), f"row {v}"
                assert og.degree(v) == ref.L.ref_degree(ref.h, v)
            pairs = rng.integers(og.n, size=(500, 2))
            for u, v in pairs:
                assert og.is_adjacent(int(u), int(v)) == bool(ref.L.ref_is_adjacent(ref.h, int(u), int(v)))
    finally:
        ref.close()


def _distinct_pairs(uv):
    key = np.minimum(uv[:, 0], uv[:, 1]).astype(np.int64) << 32 | np.maximum(uv[:, 0], uv[:, 1])
    _, first = np.unique(key, return_index=True)
    return uv[np.sort(first)]


def test_compaction_is_byte_identical_to_reference(tmp_path, sample_dir):
    """Batches of distinct pairs keep the delta consistent; both compactors write the same file."""
    egr = os.path.join(sample_dir, "sample.egr")
    rng = np.random.default_rng(7)
    gen = OracleGraph(egr)
    batches = [_distinct_pairs(b) for b in _mixed_batches(gen, rng, nb=3, m=500)]
    og, ref = OracleGraph(egr), RefGraph(egr)
    try:
        for uv in batches:
            og.apply(uv, og.classify(uv))
            ref.classify_apply(uv)
        a, b = tmp_path / "ours.egr", tmp_path / "ref.egr"
        assert oracle_lib.lib().orc_compact_to(og.h, str(a).encode()) == 0
        ref.L.ref_compact(ref.h, str(b).encode())
        assert a.read_bytes() == b.read_bytes()
    finally:
        ref.close()


_CHILD = """
import ctypes, sys, numpy as np
L = ctypes.CDLL(sys.argv[1])
L.ref_open.restype = ctypes.c_void_p
h = ctypes.c_void_p(L.ref_open(sys.argv[2].encode()))
for path in sys.argv[4:]:
    uv = np.load(path)
    ins = np.zeros(max(len(uv), 1), dtype=np.uint8)
    L.ref_classify_apply(h, uv.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(len(uv)),
                         ins.ctypes.data_as(ctypes.c_void_p))
L.ref_compact(h, sys.argv[3].encode())
"""


def test_inconsistent_delta_fails_compaction_like_reference(tmp_path, sample_dir):
    """An inserted pair deleted twice leaves a tombstone the base row does not hold: the
    reference compactor exits with an error (core/flash_compactor.h:75-80); the oracle's
    compaction reports failure too."""
    egr = os.path.join(sample_dir, "sample.egr")
    og = OracleGraph(egr)
    u, v = 0, 1
    while og.is_adjacent(u, v):
        v += 1
    batches = [np.array([[u, v]], dtype=np.int32), np.array([[u, v], [v, u]], dtype=np.int32)]
    paths = []
    for i, b in enumerate(batches):
        og.apply(b, og.classify(b))
        paths.append(str(tmp_path / f"b{i}.npy"))
        np.save(paths[-1], b)
    assert oracle_lib.lib().orc_compact_to(og.h, str(tmp_path / "ours.egr").encode()) != 0
    r = subprocess.run(["python3", "-c", _CHILD, REF_SO, egr, str(tmp_path / "ref.egr"), *paths],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "flash_compactor" in r.stderr


def test_report_lines_are_byte_identical_to_reference(tmp_path):
    exe = tmp_path / "report_sample"
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", str(exe), os.path.join(REPO, "tests", "report_sample.cpp")],
                   check=True)
    ours = subprocess.run([str(exe)], capture_output=True, check=True).stdout
    buf = ctypes.create_string_buffer(1 << 14)
    k = _ref().ref_report_sample(ctypes.cast(buf, ctypes.c_void_p), len(buf))
    assert ours == buf.raw[:k]
