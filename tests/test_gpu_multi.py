"""GPU tests of the multi-instance path: several Barabasi-Albert graphs held as one disjoint
union, every batch carrying toggles for all of them, MIS repaired by scout_repair_mis_multi
and colours by scout_repair_color on the union.

Pinned two ways after every batch: each instance's MIS bits and colours equal (1) the oracle
run on that instance's own graph with its own lines (local ids) and (2) the GPU run that
serves each instance alone in its own steps. Exact equality (integer state).

Also: the colouring result does not depend on the launch settings (thread/wave/block
thresholds, grid, captured graph on or off), and the autotuner leaves the colours unchanged.
"""
import numpy as np
import pytest

import scout_amd as sa
from conftest import run_gen
from oracle_lib import OracleGraph

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _instances(tmp_path, sizes, m, seed):
    graphs = []
    for i, n in enumerate(sizes):
        p = tmp_path / f"ba{i}.egr"
        run_gen("ba", n, m, seed + i, p)
        graphs.append(sa.Graph.read_egr(str(p)))
    return graphs


def _union(graphs):
    base = np.cumsum([0] + [g.n for g in graphs])
    eoff = np.cumsum([0] + [int(g.off[-1]) for g in graphs])
    off = np.concatenate([g.off[:-1] + eoff[i] for i, g in enumerate(graphs)] + [np.array([eoff[-1]])])
    col = np.concatenate([g.col + base[i] for i, g in enumerate(graphs)]).astype(np.int32)
    return sa.Graph(int(base[-1]), off.astype(np.int64), col), base


def _edge_sets(graphs):
    sets = []
    for g in graphs:
        s = set()
        for u in range(g.n):
            for v in g.col[g.off[u]:g.off[u + 1]]:
                if u < v:
                    s.add((u, int(v)))
        sets.append(s)
    return sets


def _toggles(rng, n, edges, size):
    """size lines on local ids: half deletions of current edges, half insertions of non-edges;
    the edge set is updated."""
    ndel = size // 2
    cur = sorted(edges)
    pick = rng.choice(len(cur), size=min(ndel, len(cur)), replace=False)
    lines = [cur[k] for k in pick]
    seen = set(lines)
    while len(lines) < size:
        u, v = (int(x) for x in rng.integers(0, n, 2))
        if u == v:
            continue
        e = (min(u, v), max(u, v))
        if e in edges or e in seen:
            continue
        seen.add(e)
        lines.append(e)
    order = rng.permutation(len(lines))
    lines = [lines[k] for k in order]
    for e in lines:
        if e in edges:
            edges.remove(e)
        else:
            edges.add(e)
    return np.array(lines, dtype=np.int32).reshape(-1, 2)


@pytest.mark.parametrize("max_csr", [150000, 10 ** 9])
def test_multi_instance_equals_oracle_and_alone(tmp_path, max_csr):
    sizes, nb = [300, 1200, 2500, 700, 4000], 4
    say(f"multi: {len(sizes)} BA instances {sizes}, {nb} batches, max_csr {max_csr}")
    graphs = _instances(tmp_path, sizes, 4, 7)
    union, base = _union(graphs)
    oracles = [OracleGraph(n=g.n, off=g.off, col=g.col) for g in graphs]
    mis_o = [og.greedy_mis() for og in oracles]
    col_o = [og.greedy_color() for og in oracles]
    rng = np.random.default_rng(3)
    edges = _edge_sets(graphs)
    batches = [[_toggles(rng, g.n, edges[i], int(rng.integers(20, 400))) for i, g in enumerate(graphs)]
               for _ in range(nb)]
    with sa.Engine(0) as mix, sa.Engine(0) as alone:
        for eng in (mix, alone):
            eng.load_graph(union)
            eng.mis_greedy()
            eng.color_greedy()
        assert np.array_equal(mix.download(sa.STATE_MIS), np.concatenate(mis_o))
        assert np.array_equal(mix.download(sa.STATE_COLOR), np.concatenate(col_o))
        for b, per in enumerate(batches):
            uv = np.concatenate([l + base[i] for i, l in enumerate(per)])
            off = np.cumsum([0] + [len(l) for l in per])
            mix.stage(uv)
            mix.classify()
            mix.apply()
            st = mix.repair_mis_multi(off, max_csr)
            mix.repair_color()
            for i, l in enumerate(per):
                alone.stage(l + base[i])
                alone.classify()
                alone.apply()
                alone.repair_mis(max_csr)
                alone.repair_color()
                # no pair is toggled twice, so labels against the current graph equal the
                # reference MIS labels against the initial graph; repairs run on the applied graph
                ins = oracles[i].classify(l)
                oracles[i].apply(l, ins)
                mis_o[i], _ = oracles[i].mis_batch(l, ins, mis_o[i], max_csr)
                col_o[i], _ = oracles[i].color_batch(l, ins, col_o[i])
            mm, cm = mix.download(sa.STATE_MIS), mix.download(sa.STATE_COLOR)
            say(f"multi: batch {b}: {len(uv)} lines, {st.sub_batches} sub-batches, {st.rounds} rounds")
            assert np.array_equal(mm, np.concatenate(mis_o)), f"batch {b}: MIS differs from the oracle"
            assert np.array_equal(cm, np.concatenate(col_o)), f"batch {b}: colours differ from the oracle"
            assert np.array_equal(mm, alone.download(sa.STATE_MIS))
            assert np.array_equal(cm, alone.download(sa.STATE_COLOR))
        for i, og in enumerate(oracles):
            assert og.check_coloring(col_o[i]) == 0
            assert og.check_mis_independent(mis_o[i]) == 0


def test_multi_rejects_bad_offsets():
    with sa.Engine(0) as eng:
        eng.gen_rmat(8, 8, seed=1)
        eng.mis_greedy()
        eng.stage(np.array([[1, 2], [3, 4]], dtype=np.int32))
        eng.classify()
        eng.apply()
        with pytest.raises(sa.ScoutError):
            eng.repair_mis_multi([0, 1], 10 ** 6)  # does not end at the staged count
        with pytest.raises(sa.ScoutError):
            eng.repair_mis_multi([0, 2, 1, 2], 10 ** 6)  # decreasing


CONFIGS = [
    dict(thread_max=0, block_min=0, grid=64, use_graph=True),       # every row per block
    dict(thread_max=0, block_min=10 ** 9, grid=256, use_graph=False),  # every row per wave
    dict(thread_max=10 ** 6, block_min=10 ** 6, grid=1024, use_graph=True),  # every row per thread
    dict(thread_max=4, block_min=16, grid=128, use_graph=True),     # all three tiers
    dict(thread_max=4, block_min=16, grid=128, use_graph=False),
]


def test_color_settings_do_not_change_results(tmp_path):
    """Colours after each of 5 batches on an R-MAT graph (hub rows of hundreds of entries)
    are identical under every launch setting, graph replays included, and equal the oracle."""
    egr = tmp_path / "g.egr"
    run_gen("rmat", 12, 16, 5, egr)
    run_gen("stream", "mixed", egr, tmp_path / "b", 5, 3000, 8, 0.4)
    g = sa.Graph.read_egr(str(egr))
    from oracle_lib import read_batch
    import os
    batches = [read_batch(os.path.join(tmp_path / "b", f"{k}.mtx"), g.n) for k in range(1, 6)]
    og = OracleGraph(str(egr))
    col = og.greedy_color()
    want = []
    for uv in batches:
        ins = og.classify(uv)
        og.apply(uv, ins)
        col, _ = og.color_batch(uv, ins, col)
        want.append(col.copy())
    for cfg in CONFIGS:
        say(f"color settings {cfg}")
        with sa.Engine(0) as eng:
            eng.load_graph(g)
            eng.color_greedy()
            eng.color_config(**cfg)
            assert eng.color_get_config()["thread_max"] == cfg["thread_max"]
            for k, uv in enumerate(batches):
                eng.stage(uv)
                eng.classify()
                eng.apply()
                st = eng.repair_color()
                assert np.array_equal(eng.download(sa.STATE_COLOR), want[k]), (cfg, k, st.rounds)


def test_autotune_keeps_colours_and_picks_a_setting(tmp_path):
    egr = tmp_path / "g.egr"
    run_gen("rmat", 12, 16, 6, egr)
    run_gen("stream", "mixed", egr, tmp_path / "b", 1, 4000, 9, 0.4)
    g = sa.Graph.read_egr(str(egr))
    from oracle_lib import read_batch
    import os
    uv = read_batch(os.path.join(tmp_path / "b", "1.mtx"), g.n)
    og = OracleGraph(str(egr))
    ins = og.classify(uv)
    col0 = og.greedy_color()
    og.apply(uv, ins)
    want, _ = og.color_batch(uv, ins, col0)
    with sa.Engine(0) as eng:
        eng.load_graph(g)
        eng.color_greedy()
        before = eng.download(sa.STATE_COLOR)
        eng.stage(uv)
        eng.classify()
        eng.apply()
        best = eng.color_autotune()
        cfg = eng.color_get_config()
        say(f"autotune: {cfg}, {best:.3f} ms")
        assert best > 0
        assert cfg["thread_max"] in (0, 4, 8, 16, 32)
        assert np.array_equal(eng.download(sa.STATE_COLOR), before)
        eng.repair_color()
        assert np.array_equal(eng.download(sa.STATE_COLOR), want)


def test_mis_autotune_keeps_membership_and_picks_a_setting(tmp_path):
    """The MIS autotuner times the staged multi-instance batch under each launch setting, leaves
    the membership as it was, and the repair that follows (with the kept setting) equals the
    oracle's, on this batch and the next."""
    sizes = [300, 1200, 2500, 700]
    say(f"mis autotune: {len(sizes)} BA instances {sizes}")
    graphs = _instances(tmp_path, sizes, 4, 5)
    union, base = _union(graphs)
    oracles = [OracleGraph(n=g.n, off=g.off, col=g.col) for g in graphs]
    mis_o = [og.greedy_mis() for og in oracles]
    rng = np.random.default_rng(8)
    edges = _edge_sets(graphs)
    max_csr = 150000
    with sa.Engine(0) as eng:
        eng.load_graph(union)
        eng.mis_greedy()
        for b in range(2):
            per = [_toggles(rng, g.n, edges[i], int(rng.integers(20, 400))) for i, g in enumerate(graphs)]
            uv = np.concatenate([l + base[i] for i, l in enumerate(per)])
            off = np.cumsum([0] + [len(l) for l in per])
            eng.stage(uv)
            eng.classify()
            eng.apply()
            if b == 0:
                before = eng.download(sa.STATE_MIS)
                best = eng.mis_autotune(off, max_csr)
                cfg = eng.mis_get_config()
                say(f"mis autotune: {cfg}, {best:.3f} ms")
                assert best > 0 and cfg["grid"] in (512, 1024, 2048, 4096) and cfg["group_rounds"] in (4, 8, 16, 32)
                assert np.array_equal(eng.download(sa.STATE_MIS), before)
            eng.repair_mis_multi(off, max_csr)
            for i, l in enumerate(per):
                ins = oracles[i].classify(l)
                oracles[i].apply(l, ins)
                mis_o[i], _ = oracles[i].mis_batch(l, ins, mis_o[i], max_csr)
            assert np.array_equal(eng.download(sa.STATE_MIS), np.concatenate(mis_o)), f"batch {b}"


@pytest.mark.parametrize("max_csr,limits", [(300000, None), (10 ** 9, None), (10 ** 9, (2, 16))])
def test_mis_hop2_searches_equal_oracle(max_csr, limits):
    """Degree-biased toggles on an R-MAT graph make many insertions between two members whose
    candidates in [r] + N(r) are all rejected: their hop-2 searches run as chunk tasks over
    many waves, or serially when the request or task list is full. MIS and colours equal the
    oracle after every batch, and the hop-2 path ran."""
    say(f"mis hop-2: R-MAT scale 12, degree-biased stream, max_csr {max_csr}, list limits {limits}")
    with sa.Engine(0) as eng:
        if limits:
            # tiny request and task lists: most searches take the serial fallback
            eng.mis_limits(*limits)
        eng.gen_rmat(12, 16, seed=4)
        g = eng.download_graph()
        og = OracleGraph(n=g.n, off=g.off, col=g.col)
        eng.mis_greedy()
        eng.color_greedy()
        mis, col = og.greedy_mis(), og.greedy_color()
        assert np.array_equal(eng.download(sa.STATE_MIS), mis)
        eng.gen_mode(sa.GEN_DEGREE)
        hop2 = 0
        for b in range(6):
            m = eng.gen_batch(3000, 0.5, 0.0, seed=7, expected_total=18000)
            uv = eng.batch_download(m)
            eng.classify()
            eng.apply()
            st = eng.repair_mis(max_csr)
            eng.repair_color()
            hop2 += st.seeds
            ins = og.classify(uv)
            og.apply(uv, ins)
            mis, _ = og.mis_batch(uv, ins, mis, max_csr)
            col, _ = og.color_batch(uv, ins, col)
            assert np.array_equal(eng.download(sa.STATE_MIS), mis), f"batch {b}: MIS differs"
            assert np.array_equal(eng.download(sa.STATE_COLOR), col), f"batch {b}: colours differ"
        say(f"mis hop-2: {hop2} hop-2 searches")
        assert hop2 > 0
