"""GPU tests of the stream generators. Degree-biased streams (GEN_DEGREE): every line is a
canonical u < v pair, no pair repeats within or across batches, deletions are current edges
and insertions current non-edges (labelled by the oracle on the same graph), deletions and
insertions are interleaved, and endpoints lean towards high degree. Both modes: the same
graph and seed give the same batches, line for line."""
import numpy as np
import pytest

import scout_amd as sa
from oracle_lib import OracleGraph

pytestmark = pytest.mark.gpu


def test_degree_stream_is_valid_and_biased():
    print("stream: degree-biased batches on R-MAT scale 12", flush=True)
    with sa.Engine(0) as eng:
        eng.gen_rmat(12, 16, seed=3)
        g = eng.download_graph()
        og = OracleGraph(n=g.n, off=g.off, col=g.col)
        deg = np.diff(g.off)
        eng.gen_mode(sa.GEN_DEGREE)
        seen = set()
        ins_deg, uni_deg = [], float(deg.mean())
        for b in range(3):
            m = eng.gen_batch(2000, 0.5, 0.0, seed=5, expected_total=6000)
            uv = eng.batch_download(m)
            assert m == 2000
            assert np.all(uv[:, 0] < uv[:, 1])
            pairs = set(map(tuple, uv.tolist()))
            assert len(pairs) == m and not (pairs & seen)
            seen |= pairs
            ins = og.classify(uv)
            assert int(ins.sum()) == 1000  # the 1000 insertions are exactly the non-edges
            # interleaved: insertions are not one block at the end
            first_ins = int(np.argmax(ins))
            assert first_ins < 100 and ins[-100:].sum() < 100
            ins_deg.append(deg[uv[ins == 1]].mean())
            eng.classify()
            eng.apply()
            og.apply(uv, ins)
        # endpoints proportional to degree + 1 have a mean degree far above the vertex mean
        assert min(ins_deg) > 3 * uni_deg


@pytest.mark.parametrize("mode", ["uniform", "degree"])
def test_streams_are_reproducible(mode):
    """Two contexts with the same graph and seed generate the same batches, line for line
    (the generator's draws run in parallel, but a pair goes to its earliest draw and accepted
    pairs keep draw order)."""
    print(f"stream: {mode} batches generated twice on R-MAT scale 12", flush=True)
    runs = []
    for _ in range(2):
        with sa.Engine(0) as eng:
            eng.gen_rmat(12, 16, seed=3)
            eng.bfs_full(0)
            if mode == "degree":
                eng.gen_mode(sa.GEN_DEGREE)
            out = []
            for b in range(3):
                m = eng.gen_batch(3000, 0.7, 0.5 if mode == "uniform" else 0.0, seed=9, expected_total=9000)
                out.append(eng.batch_download(m).copy())
                eng.classify()
                eng.apply()
                if mode == "uniform":
                    eng.repair_bfs(sa.STRATEGY_LOCAL)
            runs.append(out)
    for a, b in zip(*runs):
        assert a.shape == b.shape and np.array_equal(a, b)


def test_toggle_set_grows_and_keeps_the_stream():
    """A context whose toggle set was sized for a short stream grows it as batches accumulate:
    its batches equal those of a context sized for the whole stream up front, line for line,
    and no pair repeats across batches."""
    print("stream: toggle set sized for 1 line vs 1M lines, 8 batches of 3000", flush=True)
    runs = []
    for expected in (1, 1_000_000):
        with sa.Engine(0) as eng:
            eng.gen_rmat(12, 16, seed=3)
            eng.bfs_full(0)
            out = []
            for b in range(8):
                m = eng.gen_batch(3000, 0.7, 0.0, seed=9, expected_total=expected)
                assert m == 3000, f"batch {b}: {m} lines"
                out.append(eng.batch_download(m).copy())
                eng.classify()
                eng.apply()
                eng.repair_bfs(sa.STRATEGY_LOCAL)
            runs.append(out)
    seen = set()
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
        pairs = set(map(tuple, np.sort(a, axis=1).tolist()))
        assert len(pairs) == len(a) and not (pairs & seen)
        seen |= pairs
