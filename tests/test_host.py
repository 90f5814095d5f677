"""CPU tests of the host side: generators, .egr I/O, batch-line parsing, the binaries'
argument contract and the C ABI exports of libscout_hip.so (no device calls)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import GEN, REPO, run_gen
from oracle_lib import OracleGraph
from scout_amd import BIN_DIR, LIB_PATH, Graph, declared_symbols


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB_PATH)
    names = declared_symbols(os.path.join(REPO, "include", "scout_hip.h"))
    assert len(names) >= 27
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_graph_info_callers_size_their_buffer_from_the_header():
    """scout_graph_info writes SCOUT_GRAPH_INFO_COUNT entries: the binding and every binary that
    calls it size their buffer from that count (a shorter array is overwritten past its end)."""
    import re
    import scout_amd
    header = open(os.path.join(REPO, "include", "scout_hip.h")).read()
    count = int(re.search(r"SCOUT_GRAPH_INFO_COUNT = (\d+)", header).group(1))
    assert scout_amd.GRAPH_INFO_COUNT == count
    bins = os.path.join(REPO, "scout_amd", "csrc", "bin")
    for name in sorted(os.listdir(bins)):
        text = open(os.path.join(bins, name)).read()
        for m in re.finditer(r"scout_graph_info\(ctx, (\w+)\)", text):
            decl = re.search(r"int64_t " + m.group(1) + r"\[([^\]]+)\]", text)
            assert decl and decl.group(1) == "SCOUT_GRAPH_INFO_COUNT", f"{name}: {m.group(0)}"


@pytest.mark.parametrize("prog", ["scout_bfs", "scout_cc", "scout_coloring", "scout_mis"])
def test_binaries_print_usage_with_too_few_arguments(prog):
    r = subprocess.run([os.path.join(BIN_DIR, prog), "a", "b"], capture_output=True, text=True)
    assert r.returncode == 1
    algo = prog.split("_", 1)[1]
    assert f"<initial_{algo}_context.txt>" in r.stderr and r.stderr.startswith("Usage: ")


def _check_simple(g: Graph):
    deg = np.diff(g.off)
    src = np.repeat(np.arange(g.n), deg)
    assert (g.col >= 0).all() and (g.col < g.n).all()
    assert (src != g.col).all(), "self-loop"
    for v in range(0, g.n, max(1, g.n // 64)):
        row = g.col[g.off[v]:g.off[v + 1]]
        assert (np.diff(row) > 0).all(), "row not strictly ascending"
    a = set(zip(src.tolist(), g.col.tolist()))
    assert all((w, u) in a for (u, w) in list(a)[:5000]), "not symmetric"


def test_rmat_generator_is_deterministic_and_simple(tmp_path):
    p1, p2 = tmp_path / "a.egr", tmp_path / "b.egr"
    run_gen("rmat", 10, 16, 7, p1)
    run_gen("rmat", 10, 16, 7, p2)
    assert p1.read_bytes() == p2.read_bytes()
    g = Graph.read_egr(str(p1))
    assert g.n == 1024
    _check_simple(g)
    deg = np.diff(g.off)
    assert deg[0] == deg.max(), "vertex 0 is the hub of unpermuted R-MAT"


def test_er_generator_has_exact_edge_count(tmp_path):
    p = tmp_path / "er.egr"
    run_gen("er", 1000, 5000, 0, p)
    g = Graph.read_egr(str(p))
    assert g.off[-1] == 10000
    _check_simple(g)


def test_ba_generator_degrees(tmp_path):
    p = tmp_path / "ba.egr"
    run_gen("ba", 2000, 8, 1, p)
    g = Graph.read_egr(str(p))
    _check_simple(g)
    assert g.off[-1] == 2 * 8 * (2000 - 8)


def test_egr_reader_matches_oracle_and_compactor_roundtrip(tmp_path, sample_dir):
    src = os.path.join(sample_dir, "sample.egr")
    g = Graph.read_egr(src)
    og = OracleGraph(src)
    off, col = og.logical_csr()
    assert np.array_equal(off, g.off) and np.array_equal(col, g.col)
    # the sample is already wide: compacting an unchanged graph reproduces the file
    out = tmp_path / "c.egr"
    from oracle_lib import lib
    assert lib().orc_compact_to(og.h, str(out).encode()) == 0
    assert out.read_bytes() == open(src, "rb").read()


def test_init_files_have_reference_formats(tmp_path, sample_dir):
    src = os.path.join(sample_dir, "sample.egr")
    for kind, fixture in [("mis", "sample_mis.txt"), ("bfs", "sample_bfs.txt"), ("cc", "sample_components.txt")]:
        out = tmp_path / f"{kind}.txt"
        run_gen("init", kind, src, out)
        assert out.read_text() == open(os.path.join(sample_dir, fixture)).read(), kind


def test_stream_generator_no_repeats_and_toggles(tmp_path):
    gp = tmp_path / "g.egr"
    run_gen("er", 500, 2000, 3, gp)
    sd = tmp_path / "batches"
    run_gen("stream", "mixed", gp, sd, 4, 100, 9)
    g = OracleGraph(str(gp))
    seen = set()
    for k in range(1, 5):
        lines = [tuple(map(int, l.split())) for l in open(sd / f"{k}.mtx")]
        assert len(lines) == 100
        keys = {(min(a, b), max(a, b)) for a, b in lines}
        assert len(keys) == 100 and not (keys & seen)
        seen |= keys
        uv = np.array(lines, dtype=np.int32) - 1
        ins = g.classify(uv)
        assert ins.sum() == 50
        g.apply(uv, ins)


def test_tree_stream_deletes_consecutive_level_edges(tmp_path):
    gp = tmp_path / "r.egr"
    run_gen("rmat", 9, 8, 2, gp)
    sd = tmp_path / "b"
    run_gen("stream", "tree", gp, sd, 2, 200, 5, 0.7, 0.5)
    g = OracleGraph(str(gp))
    lv = g.bfs(0)
    uv = np.array([tuple(map(int, l.split())) for l in open(sd / "1.mtx")], dtype=np.int32) - 1
    ins = g.classify(uv)
    dels = uv[ins == 0]
    assert len(dels) == 140
    tree = np.abs(lv[dels[:, 0]] - lv[dels[:, 1]]) == 1
    assert tree.sum() >= 70


def test_batch_line_parser_rules(tmp_path):
    p = tmp_path / "b.mtx"
    p.write_text("%%MatrixMarket comment\n1 2\n3 3\n0 5\n4 5 extra tokens\n7 2abc\n12abc 3\n+2 -1\n 6\t1\r\n9 10\n")
    out = run_gen("parse", p, 8)
    pairs = [tuple(map(int, l.split())) for l in out.strip().splitlines()]
    # 1-indexed -> 0-indexed; self-loop, out-of-range and malformed lines dropped;
    # text after the second integer ignored
    assert pairs == [(0, 1), (3, 4), (6, 1), (5, 0)]


def _write_mtx(path, n, pairs, comment=True):
    with open(path, "w") as f:
        if comment:
            f.write("%%MatrixMarket matrix coordinate pattern general\n% written by a test\n")
        f.write(f"{n} {n} {len(pairs)}\n")
        for u, v in pairs:
            f.write(f"{u + 1} {v + 1}\n")


def test_mtx_converter_reproduces_the_reference_sample_egr(tmp_path, sample_dir):
    """The reference's sample graph (wide, sorted, symmetric) listed row by row as a .mtx
    converts back to the identical file (tools/efficient_converter.py writes wide output with
    rows in input order)."""
    egr = os.path.join(sample_dir, "sample.egr")
    g = Graph.read_egr(egr)
    src = np.repeat(np.arange(g.n), np.diff(g.off))
    mtx, out = tmp_path / "s.mtx", tmp_path / "s.egr"
    _write_mtx(mtx, g.n, list(zip(src.tolist(), g.col.tolist())))
    run_gen("convert", mtx, out)
    assert out.read_bytes() == open(egr, "rb").read()


def test_mtx_converter_keeps_input_order_and_skips_rows_past_the_count(tmp_path):
    rng = np.random.default_rng(4)
    n = 300
    pairs = [(int(u), int(v)) for u, v in rng.integers(0, n, (2000, 2))]
    mtx, out = tmp_path / "r.mtx", tmp_path / "r.egr"
    # a line whose row is past the vertex count is skipped, one with a single field too
    _write_mtx(mtx, n, pairs)
    with open(mtx, "a") as f:
        f.write(f"{n + 5} 1\n7\n")
    run_gen("convert", mtx, out)
    raw = out.read_bytes()
    assert np.frombuffer(raw[4:8], np.int32)[0] == -1, "always wide"
    g = Graph.read_egr(str(out))
    assert g.n == n and g.off[-1] == len(pairs)
    for v in range(n):
        assert g.col[g.off[v]:g.off[v + 1]].tolist() == [b for a, b in pairs if a == v], f"row {v}"


def test_mtx_converter_rejects_malformed_entries(tmp_path):
    mtx = tmp_path / "bad.mtx"
    mtx.write_text("4 4 2\n1 2\nx 3\n")
    r = subprocess.run([GEN, "convert", str(mtx), str(tmp_path / "bad.egr")], capture_output=True, text=True)
    assert r.returncode != 0 and "line 3" in r.stderr


def test_narrow_egr_reads_like_the_wide_file(tmp_path, sample_dir):
    """The narrow layout (int32 entry count and int32 offsets, core/FlashGraph.h:130-192) is read
    to the same rows as the wide one, by the host reader the binaries and the oracle share and
    by the Python reader."""
    egr = os.path.join(sample_dir, "sample.egr")
    g = Graph.read_egr(egr)
    narrow = tmp_path / "narrow.egr"
    with open(narrow, "wb") as f:
        f.write(np.array([g.n, g.off[-1]], np.int32).tobytes())
        f.write(g.off.astype(np.int32).tobytes())
        f.write(g.col.astype(np.int32).tobytes())
    gn = Graph.read_egr(str(narrow))
    assert np.array_equal(gn.off, g.off) and np.array_equal(gn.col, g.col)
    ow, on = OracleGraph(egr), OracleGraph(str(narrow))
    assert on.n == ow.n
    for v in range(0, g.n, 7):
        assert np.array_equal(on.row(v), ow.row(v))
