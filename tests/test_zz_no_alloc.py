"""The allocation-free CC recompute (scout_cc_full_noalloc) against scout_cc_full.

The variant writes labels into a caller-owned device array and keeps its chunk tasks and
counters in the library's fixed per-GPU scratch. Two calls in a row, the second on a smaller
graph in a new context on the same GPU, must each give the labels scout_cc_full gives (and
the oracle's min-id labels), whatever the scratch held from the call before.
"""
import ctypes

import numpy as np
import pytest

from oracle_lib import OracleGraph
import scout_amd as sa

pytestmark = pytest.mark.gpu


class DeviceArray:
    """A caller-owned device array, allocated through the HIP runtime the library itself links
    (libamdhip64.so.7), so the test owns the output memory as a real caller would."""

    def __init__(self, count, fill_byte):
        sa.load_library()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.bytes = 4 * count
        self.p = ctypes.c_void_p()
        assert self.hip.hipSetDevice(0) == 0
        assert self.hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(self.bytes)) == 0
        assert self.hip.hipMemset(self.p, fill_byte, ctypes.c_size_t(self.bytes)) == 0
        assert self.hip.hipDeviceSynchronize() == 0

    def download(self):
        out = np.zeros(self.bytes // 4, dtype=np.int32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(self.bytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def labels_both_ways(scale, seed, lines):
    with sa.Engine(0) as eng:
        eng.gen_rmat(scale, 16, seed=seed)
        # a batch of toggles, so the graph has inserted rows and deleted base entries
        eng.gen_batch(lines, 0.5, 0.0, seed=seed + 1, expected_total=lines)
        eng.classify(download=False)
        eng.apply()
        eng.cc_full()
        want = eng.download(sa.STATE_LABEL)
        out = DeviceArray(eng.n, 0xF9)
        eng.cc_full_noalloc(out.p.value)
        got = out.download()
        out.free()
        # the context's own labels are untouched by the variant
        assert np.array_equal(eng.download(sa.STATE_LABEL), want)
        g = eng.download_graph()
        paths = eng.path_counts()
    oracle = OracleGraph(n=g.n, off=g.off, col=g.col).cc()
    return got, want, oracle, int(np.diff(g.off).max()), paths


def test_cc_full_noalloc_equals_cc_full_twice():
    print("[noalloc] call 1: R-MAT scale 15 (hub rows split into chunk tasks)", flush=True)
    got, want, oracle, longest, paths = labels_both_ways(15, 41, 4000)
    assert longest > 4096 and paths["cc_tasks"] > 0
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} labels differ from scout_cc_full"
    assert np.array_equal(want, oracle)
    print("[noalloc] call 2: R-MAT scale 11 (smaller)", flush=True)
    got, want, oracle, _, _ = labels_both_ways(11, 42, 500)
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} labels differ from scout_cc_full"
    assert np.array_equal(want, oracle)


def test_cc_full_noalloc_rejects_null_output():
    with sa.Engine(0) as eng:
        eng.gen_rmat(10, 8, seed=1)
        with pytest.raises(sa.ScoutError, match="null label array"):
            eng.cc_full_noalloc(0)
