"""GPU parity of the storage-resident mode: a graph whose base entries do not fit the context's
HBM budget keeps them in pinned host memory (kernels read the rows over PCIe, full CC
recomputes stream them through two device buffers), and every repair gives the same answers
as with the graph in HBM -- the oracle's, batch by batch."""
import os
import subprocess

import numpy as np
import pytest

from conftest import run_gen
from oracle_lib import OracleGraph, mis_max_csr, read_batch
import scout_amd as sa

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def storage_engine(g, slack_words=12):
    """A context whose HBM budget holds the offsets and per-vertex state but not the base
    entries."""
    small = 16 * (g.n + 1) + 12 * g.n
    budget = small + 4 * slack_words * (g.n + 1)
    assert budget < small + 4 * int(g.off[-1]), "the budget must not hold the base entries"
    eng = sa.Engine(0, hbm_bytes=budget)
    eng.load_graph(g)
    assert eng.graph_info()["host_resident"]
    return eng


@pytest.fixture(scope="module")
def rmat(tmp_path_factory):
    root = tmp_path_factory.mktemp("storage")
    egr = root / "r14.egr"
    run_gen("rmat", 14, 16, 6, egr)
    run_gen("stream", "tree", egr, root / "b", 5, 3000, 7, 0.7, 0.5)
    run_gen("stream", "mixed", egr, root / "m", 4, 1500, 8)
    return str(egr), str(root / "b"), str(root / "m")


def test_storage_resident_bfs_cc_exact_every_batch(rmat):
    egr, bdir, _ = rmat
    say("[storage] BFS + CC, R-MAT scale 14 with its base entries in host memory")
    g = sa.Graph.read_egr(egr)
    eng = storage_engine(g)
    og = OracleGraph(egr)
    eng.upload(sa.STATE_LEVEL, og.bfs(0))
    eng.upload(sa.STATE_LABEL, og.cc())
    for k in range(1, 6):
        uv = read_batch(os.path.join(bdir, f"{k}.mtx"), g.n)
        eng.stage(uv)
        ins = eng.classify()
        assert np.array_equal(ins, og.classify(uv))
        eng.apply()
        og.apply(uv, ins)
        eng.repair_bfs()
        eng.repair_cc()
        assert np.array_equal(eng.download(sa.STATE_LEVEL), og.bfs(0)), f"batch {k}: levels"
        assert np.array_equal(eng.download(sa.STATE_LABEL), og.cc()), f"batch {k}: labels"
    # full recomputes: BFS reads the rows in host memory, CC streams them in chunks
    info0 = eng.graph_info()
    eng.bfs_full(0)
    eng.cc_full()
    assert np.array_equal(eng.download(sa.STATE_LEVEL), og.bfs(0))
    assert np.array_equal(eng.download(sa.STATE_LABEL), og.cc())
    info = eng.graph_info()
    streamed = info["fetch_bytes"] - info0["fetch_bytes"]
    say(f"[storage] streamed {streamed} B in {info['fetch_us'] - info0['fetch_us']} us")
    assert streamed == 4 * info["base_entries"]
    off, col = og.logical_csr()
    dg = eng.download_graph()
    assert np.array_equal(dg.off, off) and np.array_equal(dg.col, col)
    with pytest.raises(sa.ScoutError, match="host-resident"):
        eng.compact()
    eng.close()


def test_storage_resident_coloring_and_mis_exact_every_batch(rmat):
    egr, _, mdir = rmat
    say("[storage] colouring + MIS")
    g = sa.Graph.read_egr(egr)
    og, base = OracleGraph(egr), OracleGraph(egr)
    color, mis = og.greedy_color(), og.greedy_mis()
    max_csr = mis_max_csr(g.off[-1] / g.n, 288e9)
    ec, em = storage_engine(g), storage_engine(g)
    ec.upload(sa.STATE_COLOR, color)
    em.upload(sa.STATE_MIS, mis)
    og2 = OracleGraph(egr)
    for k in range(1, 5):
        uv = read_batch(os.path.join(mdir, f"{k}.mtx"), g.n)
        ec.stage(uv)
        ins = ec.classify()
        ec.apply()
        og.apply(uv, ins)
        ec.repair_color()
        color, _ = og.color_batch(uv, ins, color)
        assert np.array_equal(ec.download(sa.STATE_COLOR), color), f"batch {k}: colours"
        em.stage(uv)
        insb = em.classify(sa.CLASSIFY_BASE)
        assert np.array_equal(insb, base.classify(uv))
        em.apply(True)
        og2.apply(uv, insb, conditional=True)
        em.repair_mis(max_csr)
        mis, _ = og2.mis_batch(uv, insb, mis, max_csr)
        assert np.array_equal(em.download(sa.STATE_MIS), mis), f"batch {k}: MIS"
    ec.close()
    em.close()


def test_budget_errors():
    """A budget that cannot even hold the offsets and per-vertex state is refused with
    SCOUT_ERR_BUDGET (4); so is a state array past the budget."""
    g = sa.Graph.read_egr(os.path.join(os.path.dirname(__file__), "golden", "sample", "sample.egr"))
    with pytest.raises(sa.ScoutError, match=r"\(4\)"):
        sa.Engine(0, hbm_bytes=1000).load_graph(g)
    # offsets (16 B per vertex), three scratch arrays (12 B) and room for one more int32 array
    eng = sa.Engine(0, hbm_bytes=32 * (g.n + 1) + 100)
    eng.load_graph(g)
    assert eng.graph_info()["host_resident"]
    with pytest.raises(sa.ScoutError, match=r"\(4\)"):
        eng.upload(sa.STATE_LEVEL, np.zeros(g.n, np.int32))  # levels and parents: 8 B per vertex
    eng.close()


def test_binary_with_storage_resident_graph(tmp_path, sample_dir):
    """scout_bfs and scout_cc under SCOUT_HBM_BUDGET keep the sample's base entries in host
    memory and dump the same states as with the graph in HBM."""
    egr = os.path.join(sample_dir, "sample.egr")
    bdir = os.path.join(sample_dir, "batches")
    say("[storage] binaries under SCOUT_HBM_BUDGET")
    for prog, init in (("bfs", "sample_bfs.txt"), ("cc", "sample_components.txt")):
        for name, budget in (("hbm", None), ("host", 110000)):
            env = dict(os.environ, SCOUT_DUMP_STATE=str(tmp_path / name), SCOUT_SKIP_COMPACTION="1")
            if budget:
                env["SCOUT_HBM_BUDGET"] = str(budget)
            r = subprocess.run([os.path.join(sa.BIN_DIR, f"scout_{prog}"), egr, os.path.join(sample_dir, init), bdir,
                                "5", "0"], capture_output=True, text=True, env=env, cwd=tmp_path, timeout=120)
            assert r.returncode == 0, r.stderr
        for k in range(1, 6):
            a = (tmp_path / "hbm" / f"{prog}_{k}.bin").read_bytes()
            assert a == (tmp_path / "host" / f"{prog}_{k}.bin").read_bytes(), f"{prog} batch {k}"
