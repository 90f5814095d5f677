"""GPU parity: libscout_hip (through its C ABI) against the CPU oracle, batch by batch.

Bit-exact on every batch: classification labels, BFS levels, CC min-id labels, colours, MIS
membership, and the logical graph (compaction order). Graphs and streams come from the
host generator (scout_gen) with fixed seeds; sizes stay small enough for the oracle.
"""
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


@pytest.fixture(scope="module")
def data(tmp_path_factory):
    root = tmp_path_factory.mktemp("gpu_data")
    d = {}
    # small ER with a mixed stream
    er = root / "er.egr"
    run_gen("er", 3000, 20000, 0, er)
    run_gen("stream", "mixed", er, root / "er_b", 5, 800, 11)
    d["er"] = (str(er), str(root / "er_b"), 5)
    # R-MAT with deletions drawn from BFS-tree edges (70% deletions)
    rm = root / "rmat.egr"
    run_gen("rmat", 12, 16, 3, rm)
    run_gen("stream", "tree", rm, root / "rmat_b", 5, 3000, 12, 0.7, 0.5)
    d["rmat"] = (str(rm), str(root / "rmat_b"), 5)
    # degree-biased stream (hub endpoints)
    run_gen("stream", "degree", rm, root / "deg_b", 4, 2000, 13)
    d["deg"] = (str(rm), str(root / "deg_b"), 4)
    # heavy deletions: many pieces split off vertex 0's component, later batches delete inside
    # and insert between the unreached pieces
    run_gen("stream", "tree", rm, root / "heavy_b", 8, 6000, 14, 0.9, 0.6)
    d["heavy"] = (str(rm), str(root / "heavy_b"), 8)
    # the reference's bundled sample (deletion-only batches)
    sample = os.path.join(os.path.dirname(__file__), "golden", "sample")
    d["sample"] = (os.path.join(sample, "sample.egr"), os.path.join(sample, "batches"), 5)
    return d


def batches(bdir, count, n):
    for k in range(1, count + 1):
        yield k, read_batch(os.path.join(bdir, f"{k}.mtx"), n)


def engine_for(egr):
    g = sa.Graph.read_egr(egr)
    eng = sa.Engine(0)
    eng.load_graph(g)
    return eng, g


CC_MODES = {"walk": [sa.CC_WALK], "scan": [sa.CC_SCAN], "mixed": [sa.CC_WALK, sa.CC_SCAN, sa.CC_AUTO]}


def components(lab):
    return int(np.count_nonzero(lab == np.arange(len(lab))))


@pytest.mark.parametrize("mode", list(CC_MODES))
@pytest.mark.parametrize("name", ["sample", "er", "rmat", "deg", "heavy"])
def test_bfs_cc_exact_every_batch(data, name, mode):
    """Levels and labels equal the oracle's after every batch, in each CC mode (walk from the
    endpoints, scan of every vertex, or the modes in turn), and the split count equals the rise
    in the oracle's component count."""
    egr, bdir, nb = data[name]
    say(f"[bfs+cc] {name} {mode}")
    eng, g = engine_for(egr)
    og = OracleGraph(egr)
    eng.upload(sa.STATE_LEVEL, og.bfs(0))
    eng.upload(sa.STATE_LABEL, og.cc())
    comps = components(og.cc())
    for k, uv in batches(bdir, nb, g.n):
        eng.cc_mode(CC_MODES[mode][k % len(CC_MODES[mode])])
        eng.stage(uv)
        ins = eng.classify(sa.CLASSIFY_CURRENT)
        ins_o = og.classify(uv)
        assert np.array_equal(ins, ins_o), f"classification differs in batch {k}"
        eng.apply(False)
        og.apply(uv, ins_o)
        eng.repair_bfs()
        st = eng.repair_cc()
        lv, lab = eng.download(sa.STATE_LEVEL), eng.download(sa.STATE_LABEL)
        want_lv, want_lab = og.bfs(0), og.cc()
        assert np.array_equal(lv, want_lv), f"{name} batch {k}: {np.count_nonzero(lv != want_lv)} levels differ"
        assert np.array_equal(lab, want_lab), f"{name} batch {k}: {np.count_nonzero(lab != want_lab)} labels differ"
        now = components(want_lab)
        if k >= 2:  # the first repair after an upload counts the components without a delta
            assert st.splits == max(0, now - comps), f"{name} batch {k}: splits {st.splits}, want {now - comps}"
        comps = now
    off, col = og.logical_csr()
    dg = eng.download_graph()
    assert np.array_equal(dg.off, off) and np.array_equal(dg.col, col), "logical graph differs"
    eng.close()


# statistics fixed by the batch; seeds, rounds, scanned entries and read-backs depend on the
# order in which concurrent lines and rounds lower the same vertex
STAT_FIELDS = ("lines", "insertions", "deletions", "noop", "touched", "fallback", "splits", "cc_mode")


@pytest.mark.parametrize("strategy", [sa.STRATEGY_LOCAL, sa.STRATEGY_RECOMPUTE])
@pytest.mark.parametrize("name", ["sample", "rmat", "deg", "heavy"])
def test_commit_equals_the_separate_calls(data, name, strategy):
    """scout_commit (classify, apply, BFS and CC repairs in one call) leaves the levels, labels
    and logical graph of classify + apply + repair_bfs + repair_cc one by one, with the same
    batch-determined statistics, and both equal the oracle after every batch."""
    egr, bdir, nb = data[name]
    say(f"[commit] {name} strategy {strategy}")
    one, g = engine_for(egr)
    many, _ = engine_for(egr)
    og = OracleGraph(egr)
    for e in (one, many):
        e.upload(sa.STATE_LEVEL, og.bfs(0))
        e.upload(sa.STATE_LABEL, og.cc())
    for k, uv in batches(bdir, nb, g.n):
        one.stage(uv)
        sb1, sc1 = one.commit(strategy)
        many.stage(uv)
        many.classify(download=False)
        many.apply(False)
        sb2 = many.repair_bfs(strategy)
        sc2 = many.repair_cc(strategy)
        og.apply(uv, og.classify(uv))
        want_lv, want_lab = og.bfs(0), og.cc()
        for e in (one, many):
            assert np.array_equal(e.download(sa.STATE_LEVEL), want_lv), f"{name} batch {k}: levels differ"
            assert np.array_equal(e.download(sa.STATE_LABEL), want_lab), f"{name} batch {k}: labels differ"
        for f in STAT_FIELDS:
            assert getattr(sb1, f) == getattr(sb2, f), f"{name} batch {k}: BFS stat {f}"
            assert getattr(sc1, f) == getattr(sc2, f), f"{name} batch {k}: CC stat {f}"
    d1, d2 = one.download_graph(), many.download_graph()
    assert np.array_equal(d1.off, d2.off) and np.array_equal(d1.col, d2.col)
    one.close()
    many.close()


@pytest.mark.parametrize("bfs_mode", [sa.BFS_LEVEL_SYNC, sa.BFS_ASYNC_INVALIDATION])
@pytest.mark.parametrize("name", ["sample", "er", "rmat", "deg", "heavy"])
def test_bfs_level_sync_exact_every_batch(data, name, bfs_mode):
    """The level-synchronous BFS round order (one level per round), and the asynchronous
    invalidation with a level-synchronous relax, give the oracle's levels and labels after every
    batch too; the tests above run the default asynchronous order."""
    egr, bdir, nb = data[name]
    say(f"[bfs mode {bfs_mode}] {name}")
    eng, g = engine_for(egr)
    eng.bfs_mode(bfs_mode)
    og = OracleGraph(egr)
    eng.upload(sa.STATE_LEVEL, og.bfs(0))
    eng.upload(sa.STATE_LABEL, og.cc())
    for k, uv in batches(bdir, nb, g.n):
        eng.stage(uv)
        ins = eng.classify(sa.CLASSIFY_CURRENT)
        og.apply(uv, og.classify(uv))
        eng.apply(False)
        eng.repair_bfs()
        eng.repair_cc()
        lv, want_lv = eng.download(sa.STATE_LEVEL), og.bfs(0)
        assert np.array_equal(lv, want_lv), f"{name} batch {k}: {np.count_nonzero(lv != want_lv)} levels differ"
        assert np.array_equal(eng.download(sa.STATE_LABEL), og.cc()), f"{name} batch {k}: labels differ"
        assert ins is not None
    eng.close()


def test_inserted_rows_keep_line_order_past_the_block_sort():
    """New inserted-row entries are placed in line order per row. A batch of 30,000 insertions
    over 1,500 vertices gives almost every source several new entries (more than one block
    groups, so the library sort runs) and rows of more than 64 entries move to fresh spans; a
    mixed batch then appends to, tombstones and forgets those entries (grouped by one block);
    a last batch gives one source 400 new entries (a run longer than the block ranks: the
    library sort again). The logical rows equal the oracle's after each batch."""
    rng = np.random.default_rng(21)
    n = 1500
    pairs = rng.integers(0, n, size=(3000, 2))
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    key = np.unique(np.minimum(pairs[:, 0], pairs[:, 1]) * n + np.maximum(pairs[:, 0], pairs[:, 1]))
    u, v = key // n, key % n
    both = np.concatenate([np.stack([u, v], 1), np.stack([v, u], 1)])
    both = both[np.lexsort((both[:, 1], both[:, 0]))]
    off = np.zeros(n + 1, dtype=np.int64)
    np.add.at(off, both[:, 0] + 1, 1)
    off = np.cumsum(off)
    col = both[:, 1].astype(np.int32)
    og = OracleGraph(n=n, off=off, col=col)
    eng = sa.Engine(0)
    eng.load_graph(sa.Graph(n, off, col))
    present = set(key.tolist())

    def fresh(count):
        out, seen = [], set()
        while len(out) < count:
            a, b = (int(x) for x in rng.integers(0, n, 2))
            k = min(a, b) * n + max(a, b)
            if a == b or k in seen or k in present:
                continue
            seen.add(k)
            out.append((a, b))
        return np.array(out, dtype=np.int32)

    first = fresh(30000)
    for k in (1, 2, 3):
        if k == 1:
            uv = first
        elif k == 2:
            # new pairs (not present after batch 1) and deletions of batch 1's: all distinct;
            # ~8,000 new entries over 1,500 sources, grouped by one block
            second = np.concatenate([fresh(4000), first[rng.choice(len(first), 3000, replace=False)]])
            uv = np.ascontiguousarray(second[rng.permutation(len(second))])
        else:
            # one source gains 400 new entries (a run longer than one block ranks), among others
            hub = 7
            new_nb = [w for w in range(n) if w != hub and min(hub, w) * n + max(hub, w) not in present][:400]
            others = fresh(600)
            others = others[(others[:, 0] != hub) & (others[:, 1] != hub)]
            uv = np.ascontiguousarray(np.concatenate([np.array([[hub, w] for w in new_nb], dtype=np.int32),
                                                      others])[rng.permutation(len(new_nb) + len(others))])
        eng.stage(uv)
        ins = eng.classify(sa.CLASSIFY_CURRENT)
        ins_o = og.classify(uv)
        assert np.array_equal(ins, ins_o), f"classification differs in batch {k}"
        eng.apply(False)
        og.apply(uv, ins_o)
        want_off, want_col = og.logical_csr()
        dg = eng.download_graph()
        assert np.array_equal(dg.off, want_off), f"batch {k}: row lengths differ"
        assert np.array_equal(dg.col, want_col), f"batch {k}: row entries or their order differ"
        for a, b in uv.tolist():
            present ^= {min(a, b) * n + max(a, b)}
    eng.close()


def test_segments_match_sequential_batches(data):
    """Three batches staged back to back and applied as segments (one serving step) give the
    labels, the logical graph and the BFS/CC solutions of applying them one by one, also when
    later batches toggle pairs of earlier ones."""
    egr, bdir, _ = data["rmat"]
    say("[segments] rmat")
    eng, g = engine_for(egr)
    og = OracleGraph(egr)
    eng.upload(sa.STATE_LEVEL, og.bfs(0))
    eng.upload(sa.STATE_LABEL, og.cc())
    parts = [uv for _, uv in batches(bdir, 3, g.n)]
    rng = np.random.default_rng(3)
    # later segments toggle some pairs of earlier ones again
    parts[1] = np.concatenate([parts[1], parts[0][rng.choice(len(parts[0]), 200, replace=False)][:, ::-1]])
    parts[2] = np.concatenate([parts[2], parts[1][rng.choice(len(parts[1]), 200, replace=False)]])
    allv = np.ascontiguousarray(np.concatenate(parts).astype(np.int32))
    off = np.cumsum([0] + [len(p) for p in parts])
    eng.stage(allv)
    ins = eng.apply_segments(off)
    want = []
    for p in parts:
        i = og.classify(p)
        og.apply(p, i)
        want.append(i)
    assert np.array_equal(ins, np.concatenate(want))
    eng.repair_bfs()
    eng.repair_cc()
    assert np.array_equal(eng.download(sa.STATE_LEVEL), og.bfs(0))
    assert np.array_equal(eng.download(sa.STATE_LABEL), og.cc())
    off_o, col_o = og.logical_csr()
    dg = eng.download_graph()
    assert np.array_equal(dg.off, off_o) and np.array_equal(dg.col, col_o)
    eng.close()


def test_serving_loop_steps_equal_the_oracle(tmp_path):
    """The serving loop (Poisson arrivals, log-uniform request sizes, 60% deletions over
    adjacency entries) at a rate that queues requests: every committed step's levels and labels
    equal the oracle after the step's requests are applied in order, each against the graph the
    earlier ones left; the final state equals serving every request alone and a full
    recompute; steps are staged on one channel while the device repairs the other."""
    import json
    say("[serve] R-MAT scale 12, 60 requests")
    exe = os.path.join(sa.BIN_DIR, "scout_serve")
    dump = tmp_path / "serve"
    r = subprocess.run([exe, "--scale", "12", "--rate", "3e6", "--min", "50", "--max", "4000", "--requests", "60",
                        "--pinned-mb", "1", "--hbm-mb", "8", "--seed", "3", "--dump", str(dump)],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    say(f"[serve] {out['steps']} steps, at most {out['max_requests_per_step']} requests per step, "
        f"{out['steps_staged_while_device_busy']} staged while busy, PCIe {out['pcie_h2d_GBps']} GB/s")
    assert out["same_as_alone"] and out["verified_exact"]
    assert out["max_requests_per_step"] > 1 and out["steps"] > 2
    egr = tmp_path / "g.egr"
    run_gen("rmat", 12, 16, 3, egr)
    og = OracleGraph(str(egr))
    steps = [tuple(map(int, ln.split())) for ln in (dump / "steps.txt").read_text().split("\n") if ln]
    assert steps[0][0] == 0 and steps[-1][1] == 60 and all(a[1] == b[0] for a, b in zip(steps, steps[1:]))
    for k, (first, last) in enumerate(steps):
        for q in range(first, last):
            uv = np.fromfile(dump / f"req_{q}.bin", dtype=np.int32).reshape(-1, 2)
            og.apply(uv, og.classify(uv))
        lv = np.fromfile(dump / f"lv_{k}.bin", dtype=np.int32)
        lab = np.fromfile(dump / f"lab_{k}.bin", dtype=np.int32)
        assert np.array_equal(lv, og.bfs(0)), f"step {k}: levels differ"
        assert np.array_equal(lab, og.cc()), f"step {k}: labels differ"


@pytest.mark.parametrize("size", [2000, 20000])
def test_cc_listings_equal_the_oracle_at_scale_18(size):
    """R-MAT scale 18 (262k vertices, about 8M entries) with device-generated batches of `size`
    toggles (70% deletions, half of them BFS-tree edges): the walk listing and the scan listing
    give the oracle's levels and labels after every batch."""
    gen = sa.Engine(0)
    gen.gen_rmat(18, 16, seed=21)
    gen.bfs_full(0)
    gen.cc_full()
    g = gen.download_graph()
    batches = []
    for _ in range(4):
        m = gen.gen_batch(size, 0.7, 0.5, seed=22, expected_total=4 * size)
        batches.append(gen.batch_download(m).copy())
        gen.classify(download=False)
        gen.apply()
        gen.repair_bfs()
        gen.repair_cc()
    gen.close()
    say(f"[cc scale 18] n={g.n} entries={len(g.col)} batches of {size}")
    og = OracleGraph(n=g.n, off=g.off, col=g.col)
    lv0, lab0 = og.bfs(0), og.cc()
    engines = {}
    for mode in ("walk", "scan"):
        e = sa.Engine(0)
        e.load_graph(g)
        e.cc_mode(CC_MODES[mode][0])
        e.upload(sa.STATE_LEVEL, lv0)
        e.upload(sa.STATE_LABEL, lab0)
        engines[mode] = e
    try:
        for k, uv in enumerate(batches):
            ins = og.classify(uv)
            og.apply(uv, ins)
            want_lv, want_lab = og.bfs(0), og.cc()
            for mode, e in engines.items():
                e.stage(uv)
                e.classify(download=False)
                e.apply()
                e.repair_bfs()
                st = e.repair_cc()
                assert st.cc_mode == CC_MODES[mode][0]
                assert np.array_equal(e.download(sa.STATE_LEVEL), want_lv), f"{mode} batch {k}: levels"
                assert np.array_equal(e.download(sa.STATE_LABEL), want_lab), f"{mode} batch {k}: labels"
    finally:
        for e in engines.values():
            e.close()


@pytest.mark.parametrize("mode", ["walk", "scan"])
@pytest.mark.parametrize("name", ["rmat", "heavy"])
def test_cc_alone_exact_every_batch(data, name, mode):
    """CC repaired without BFS maintained on the context: the CC state keeps its own levels
    from vertex 0 (and their parents) and repairs them batch by batch."""
    egr, bdir, nb = data[name]
    say(f"[cc alone] {name} {mode}")
    eng, g = engine_for(egr)
    eng.cc_mode(CC_MODES[mode][0])
    og = OracleGraph(egr)
    eng.upload(sa.STATE_LABEL, og.cc())
    for k, uv in batches(bdir, nb, g.n):
        eng.stage(uv)
        ins = eng.classify(sa.CLASSIFY_CURRENT)
        eng.apply(False)
        og.apply(uv, og.classify(uv))
        st = eng.repair_cc()
        assert np.array_equal(eng.download(sa.STATE_LABEL), og.cc()), f"{name} batch {k}"
        assert st.lines == len(ins)
    eng.close()


@pytest.mark.parametrize("name", ["er", "rmat"])
def test_recompute_strategy_equals_local(data, name):
    egr, bdir, nb = data[name]
    say(f"[strategies] {name}")
    og = OracleGraph(egr)
    lv0, lab0 = og.bfs(0), og.cc()
    results = []
    for strat in (sa.STRATEGY_LOCAL, sa.STRATEGY_RECOMPUTE):
        eng, g = engine_for(egr)
        eng.upload(sa.STATE_LEVEL, lv0)
        eng.upload(sa.STATE_LABEL, lab0)
        out = []
        for k, uv in batches(bdir, nb, g.n):
            eng.stage(uv)
            eng.classify()
            eng.apply()
            eng.repair_bfs(strat)
            eng.repair_cc(strat)
            out.append((eng.download(sa.STATE_LEVEL), eng.download(sa.STATE_LABEL)))
        eng.close()
        results.append(out)
    for (a1, b1), (a2, b2) in zip(*results):
        assert np.array_equal(a1, a2) and np.array_equal(b1, b2)


@pytest.mark.parametrize("name", ["sample", "er", "rmat", "deg"])
def test_coloring_exact_every_batch(data, name):
    egr, bdir, nb = data[name]
    say(f"[coloring] {name}")
    eng, g = engine_for(egr)
    og = OracleGraph(egr)
    color = og.greedy_color()
    eng.upload(sa.STATE_COLOR, color)
    for k, uv in batches(bdir, nb, g.n):
        eng.stage(uv)
        ins = eng.classify()
        eng.apply()
        og.apply(uv, ins)
        st = eng.repair_color()
        color, ost = og.color_batch(uv, ins, color)
        got = eng.download(sa.STATE_COLOR)
        assert np.array_equal(got, color), f"{name} batch {k}: {np.count_nonzero(got != color)} colours differ"
        assert st.conflicts == ost[0] and st.rounds == ost[1]
        assert og.check_coloring(got) == 0


@pytest.mark.parametrize("name", ["sample", "er", "rmat", "deg"])
def test_mis_exact_every_batch(data, name):
    egr, bdir, nb = data[name]
    say(f"[mis] {name}")
    eng, g = engine_for(egr)
    og = OracleGraph(egr)
    base = OracleGraph(egr)       # the initial graph: MIS labels every batch against it
    mis = og.greedy_mis()
    eng.upload(sa.STATE_MIS, mis)
    avg = g.off[-1] / g.n
    max_csr = mis_max_csr(avg, 288e9)
    for k, uv in batches(bdir, nb, g.n):
        eng.stage(uv)
        ins = eng.classify(sa.CLASSIFY_BASE)
        assert np.array_equal(ins, base.classify(uv))
        eng.apply(True)
        og.apply(uv, ins, conditional=True)
        st = eng.repair_mis(max_csr)
        mis, ost = og.mis_batch(uv, ins, mis, max_csr)
        got = eng.download(sa.STATE_MIS)
        assert np.array_equal(got, mis), f"{name} batch {k}: {np.count_nonzero(got != mis)} memberships differ"
        assert st.sub_batches == ost[0]
        assert og.check_mis_independent(got) == 0


@pytest.mark.parametrize("name", ["sample", "er", "rmat", "deg"])
def test_mis_complete_mode_exact_and_maximal_every_batch(data, name):
    """Complete mode: the reference's rules, then the greedy completion over the batch's 2-hop
    region. Equal to the oracle's sequential completion after every batch, independent, and
    maximal on the whole graph (the greedy start is maximal)."""
    egr, bdir, nb = data[name]
    say(f"[mis complete] {name}")
    eng, g = engine_for(egr)
    eng.mis_mode(sa.MIS_COMPLETE)
    og, base = OracleGraph(egr), OracleGraph(egr)
    mis = og.greedy_mis()
    eng.upload(sa.STATE_MIS, mis)
    max_csr = mis_max_csr(g.off[-1] / g.n, 288e9)
    for k, uv in batches(bdir, nb, g.n):
        eng.stage(uv)
        ins = eng.classify(sa.CLASSIFY_BASE)
        eng.apply(True)
        og.apply(uv, ins, conditional=True)
        st = eng.repair_mis(max_csr)
        mis, _ = og.mis_batch(uv, ins, mis, max_csr)
        mis, added = og.mis_complete(uv, mis)
        got = eng.download(sa.STATE_MIS)
        assert np.array_equal(got, mis), f"{name} batch {k}: {np.count_nonzero(got != mis)} memberships differ"
        assert st.completed == added
        assert og.check_mis_independent(got) == 0 and og.mis_maximality_violations(got) == 0
    eng.close()


def test_mis_small_subbatches_match(data):
    # a tiny MAX_CSR_ELEMENTS cuts every batch into many sub-batches with different rows
    egr, bdir, nb = data["er"]
    say("[mis] sub-batches")
    eng, g = engine_for(egr)
    og = OracleGraph(egr)
    mis = og.greedy_mis()
    eng.upload(sa.STATE_MIS, mis)
    for k, uv in batches(bdir, 2, g.n):
        eng.stage(uv)
        ins = eng.classify(sa.CLASSIFY_BASE)
        eng.apply(True)
        og.apply(uv, ins, conditional=True)
        st = eng.repair_mis(200000)
        mis, ost = og.mis_batch(uv, ins, mis, 200000)
        assert st.sub_batches == ost[0] and st.sub_batches > 1
        assert np.array_equal(eng.download(sa.STATE_MIS), mis)


def test_repeated_pairs_in_one_batch(data):
    egr, _, _ = data["er"]
    say("[graph] repeated pairs")
    eng, g = engine_for(egr)
    og = OracleGraph(egr)
    eng.upload(sa.STATE_LEVEL, og.bfs(0))
    eng.upload(sa.STATE_LABEL, og.cc())
    rng = np.random.default_rng(5)
    edges = np.stack([np.repeat(np.arange(g.n), np.diff(g.off)), g.col], 1)
    pick = edges[rng.integers(0, len(edges), 300)]
    rand = rng.integers(0, g.n, (300, 2))
    uv = np.concatenate([pick, rand, pick[:100, ::-1], rand[:100]]).astype(np.int32)
    uv = uv[uv[:, 0] != uv[:, 1]]
    for rep in range(3):
        eng.stage(uv)
        ins = eng.classify()
        assert np.array_equal(ins, og.classify(uv))
        eng.apply()
        og.apply(uv, ins)
        eng.repair_bfs()
        eng.repair_cc()
        assert np.array_equal(eng.download(sa.STATE_LEVEL), og.bfs(0))
        assert np.array_equal(eng.download(sa.STATE_LABEL), og.cc())
        dg = eng.download_graph()
        rows_o = [og.row(v) for v in range(g.n)]
        rows_g = [dg.col[dg.off[v]:dg.off[v + 1]] for v in range(g.n)]
        assert all(np.array_equal(a, b) for a, b in zip(rows_o, rows_g)), "rows differ"


def test_commit_settles_repeated_pairs_and_the_library_sort(data):
    """scout_commit queues the BFS repair's first kernels before reading the graph update's
    counters back; rounds that read rows are shut by a device gate while the update is not
    complete. Batches that leave it incomplete: repeated pairs (the device applies nothing, the
    host applies runs of distinct pairs) and more multi entries than one block groups (the
    library sort places them), one batch with a 400-entry run. Levels, labels and the logical
    rows equal the oracle's after every batch, and the stream ends with batches that settle on
    the device."""
    egr, _, _ = data["er"]
    say("[commit] incomplete updates")
    eng, g = engine_for(egr)
    og = OracleGraph(egr)
    eng.upload(sa.STATE_LEVEL, og.bfs(0))
    eng.upload(sa.STATE_LABEL, og.cc())
    rng = np.random.default_rng(9)
    n = g.n
    edges = np.stack([np.repeat(np.arange(n), np.diff(g.off)), g.col], 1)

    def absent(count, avoid=()):
        out, seen = [], set()
        while len(out) < count:
            a, b = (int(x) for x in rng.integers(0, n, 2))
            k = (min(a, b), max(a, b))
            if a == b or k in seen or og.is_adjacent(a, b) or a in avoid or b in avoid:
                continue
            seen.add(k)
            out.append((a, b))
        return np.array(out, dtype=np.int32)

    def present(count):
        return edges[rng.integers(0, len(edges), count)].astype(np.int32)

    hub = 11
    for k in range(1, 7):
        if k == 1:  # repeated pairs
            pick, rand = present(300), absent(300)
            uv = np.concatenate([pick, rand, pick[:100, ::-1], rand[:100]])
        elif k == 2:  # ~15,000 multi entries (most sources gain several): the library sort
            uv = absent(15000)
        elif k == 3:  # one source gains 400 entries among others
            nb = [w for w in range(n) if w != hub and not og.is_adjacent(hub, w)][:400]
            uv = np.concatenate([np.array([[hub, w] for w in nb], dtype=np.int32), absent(500, {hub})])
            uv = uv[rng.permutation(len(uv))]
        else:  # settled on the device
            uv = np.concatenate([present(400), absent(400)])
        uv = np.ascontiguousarray(uv[uv[:, 0] != uv[:, 1]])
        eng.stage(uv)
        eng.commit(sa.STRATEGY_LOCAL)
        og.apply(uv, og.classify(uv))
        assert np.array_equal(eng.download(sa.STATE_LEVEL), og.bfs(0)), f"batch {k}: levels differ"
        assert np.array_equal(eng.download(sa.STATE_LABEL), og.cc()), f"batch {k}: labels differ"
        # row by row (repeated pairs can leave the delta in a state whose compaction the
        # reference refuses, so the oracle's logical CSR is not taken)
        dg = eng.download_graph()
        rows_g = [dg.col[dg.off[v]:dg.off[v + 1]] for v in range(n)]
        assert all(np.array_equal(og.row(v), rows_g[v]) for v in range(n)), f"batch {k}: rows differ"
        edges = np.stack([np.repeat(np.arange(n), np.diff(dg.off)), dg.col], 1)
    eng.close()


def test_device_generator_matches_host(tmp_path):
    say("[gen] rmat scale 12")
    p = tmp_path / "h.egr"
    run_gen("rmat", 12, 16, 21, p)
    host = sa.Graph.read_egr(str(p))
    eng = sa.Engine(0)
    eng.gen_rmat(12, 16, seed=21)
    dev = eng.download_graph()
    assert np.array_equal(dev.off, host.off) and np.array_equal(dev.col, host.col)
    # greedy initial states and full recomputes on the device
    og = OracleGraph(str(p))
    eng.mis_greedy()
    eng.color_greedy()
    eng.bfs_full(0)
    eng.cc_full()
    assert np.array_equal(eng.download(sa.STATE_MIS), og.greedy_mis())
    assert np.array_equal(eng.download(sa.STATE_COLOR), og.greedy_color())
    assert np.array_equal(eng.download(sa.STATE_LEVEL), og.bfs(0))
    assert np.array_equal(eng.download(sa.STATE_LABEL), og.cc())
    eng.close()


def test_compaction_on_device_keeps_logical_graph(data):
    egr, bdir, nb = data["rmat"]
    say("[graph] compaction")
    eng, g = engine_for(egr)
    og = OracleGraph(egr)
    for k, uv in batches(bdir, 2, g.n):
        eng.stage(uv)
        ins = eng.classify()
        eng.apply()
        og.apply(uv, ins)
    before = eng.download_graph()
    eng.compact()
    after = eng.download_graph()
    off, col = og.logical_csr()
    assert np.array_equal(after.off, off) and np.array_equal(after.col, col)
    assert np.array_equal(before.col, after.col)
    info = eng.graph_info()
    assert info["inserted_entries"] == 0 and info["base_entries"] == len(col)


@pytest.mark.parametrize("prog,init_name", [("bfs", "sample_bfs.txt"), ("cc", "sample_components.txt"),
                                             ("coloring", "sample_colors.txt"), ("mis", "sample_mis.txt")])
def test_binaries_on_sample_dump_exact_states(tmp_path, data, prog, init_name):
    egr, bdir, nb = data["sample"]
    say(f"[binary] scout_{prog}")
    work = tmp_path / "w"
    work.mkdir()
    local = work / "sample.egr"
    local.write_bytes(open(egr, "rb").read())
    dump = work / "dump"
    env = dict(os.environ, SCOUT_DUMP_STATE=str(dump), SCOUT_SKIP_COMPACTION="1")
    init = os.path.join(os.path.dirname(egr), init_name)
    r = subprocess.run([os.path.join(sa.BIN_DIR, f"scout_{prog}"), str(local), init, bdir, "5", "0"],
                       capture_output=True, text=True, env=env, cwd=work, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("Batch ") >= 5 and "=== FINAL SUMMARY ===" in r.stdout
    og = OracleGraph(egr)
    base = OracleGraph(egr)
    n = og.n
    vals = np.loadtxt(init, dtype=np.int64)
    if prog == "bfs":
        state = np.full(n, -1, np.int32)
        state[vals[:, 0] - 1] = vals[:, 1]
    elif prog == "cc":
        state = vals.astype(np.int32)
    elif prog == "coloring":
        state = vals.astype(np.int32)
    else:
        state = np.zeros(n, np.uint8)
        state[vals - 1] = 1
    for k, uv in batches(bdir, 5, n):
        ins = (base if prog == "mis" else og).classify(uv)
        og.apply(uv, ins, conditional=(prog == "mis"))
        if prog == "bfs":
            state = og.bfs(0)
        elif prog == "cc":
            state = og.cc()
        elif prog == "coloring":
            state, _ = og.color_batch(uv, ins, state)
        else:
            state, _ = og.mis_batch(uv, ins, state, mis_max_csr(8.0, 288e9))
        got = np.fromfile(dump / f"{prog}_{k}.bin", dtype=np.uint8 if prog == "mis" else np.int32)
        assert np.array_equal(got, state), f"scout_{prog} batch {k}"


def test_mis_er_100k_full_size_exact_every_batch(tmp_path):
    """The MIS configuration at its full size: ER n=100,000 with 800,000 edges (seed 0),
    greedy-ascending initial MIS, 10 batches of 10,000 toggles (half insertions of non-edges,
    half deletions of edges). Every batch's MIS equals the oracle's and is independent; the
    maximality violations (which the reference's rules allow) are reported."""
    say("[mis] ER 100k / 800k edges, 10 x 10k toggles")
    egr = tmp_path / "er100k.egr"
    run_gen("er", 100000, 800000, 0, egr)
    run_gen("stream", "mixed", egr, tmp_path / "b", 10, 10000, 1)
    eng, g = engine_for(str(egr))
    og = OracleGraph(str(egr))
    base = OracleGraph(str(egr))
    eng.mis_greedy()
    mis = og.greedy_mis()
    assert np.array_equal(eng.download(sa.STATE_MIS), mis)
    max_csr = mis_max_csr(g.off[-1] / g.n, 288e9)
    for k, uv in batches(str(tmp_path / "b"), 10, g.n):
        eng.stage(uv)
        ins = eng.classify(sa.CLASSIFY_BASE)
        assert np.array_equal(ins, base.classify(uv))
        eng.apply(True)
        og.apply(uv, ins, conditional=True)
        st = eng.repair_mis(max_csr)
        mis, ost = og.mis_batch(uv, ins, mis, max_csr)
        got = eng.download(sa.STATE_MIS)
        assert np.array_equal(got, mis), f"batch {k}: {np.count_nonzero(got != mis)} memberships differ"
        assert st.sub_batches == ost[0]
        assert og.check_mis_independent(got) == 0
        say(f"[mis] batch {k}: {int(got.sum())} members, {og.mis_maximality_violations(got)} maximality "
            f"violations, {st.rounds} rounds, {st.repair_ms:.2f} ms")
    eng.close()
    # complete mode on the same stream: equal to the oracle's completion, independent and maximal
    say("[mis] ER 100k, complete mode")
    eng, g = engine_for(str(egr))
    eng.mis_mode(sa.MIS_COMPLETE)
    og, base = OracleGraph(str(egr)), OracleGraph(str(egr))
    mis = og.greedy_mis()
    eng.upload(sa.STATE_MIS, mis)
    for k, uv in batches(str(tmp_path / "b"), 10, g.n):
        eng.stage(uv)
        ins = eng.classify(sa.CLASSIFY_BASE)
        eng.apply(True)
        og.apply(uv, ins, conditional=True)
        st = eng.repair_mis(max_csr)
        mis, _ = og.mis_batch(uv, ins, mis, max_csr)
        mis, added = og.mis_complete(uv, mis)
        got = eng.download(sa.STATE_MIS)
        assert np.array_equal(got, mis), f"complete batch {k}: {np.count_nonzero(got != mis)} memberships differ"
        assert og.check_mis_independent(got) == 0 and og.mis_maximality_violations(got) == 0
        say(f"[mis] complete batch {k}: {int(got.sum())} members, {added} added by the completion, "
            f"{st.repair_ms:.2f} ms")
    eng.close()


def run_binary(prog, egr, init, bdir, nb, cwd, **env):
    e = dict(os.environ, **{k: str(v) for k, v in env.items()})
    r = subprocess.run([os.path.join(sa.BIN_DIR, f"scout_{prog}"), str(egr), init, bdir, str(nb), "0"],
                       capture_output=True, text=True, env=e, cwd=cwd, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_binaries_read_the_narrow_egr(tmp_path, data):
    """scout_bfs and scout_cc on the sample written in the narrow layout dump the same states
    as on the wide file."""
    egr, bdir, nb = data["sample"]
    say("[binary] narrow .egr")
    g = sa.Graph.read_egr(egr)
    narrow = tmp_path / "narrow.egr"
    with open(narrow, "wb") as f:
        f.write(np.array([g.n, g.off[-1]], np.int32).tobytes())
        f.write(g.off.astype(np.int32).tobytes())
        f.write(g.col.astype(np.int32).tobytes())
    sample = os.path.dirname(egr)
    for prog, init in (("bfs", "sample_bfs.txt"), ("cc", "sample_components.txt")):
        for name, path in (("wide", egr), ("narrow", narrow)):
            run_binary(prog, path, os.path.join(sample, init), bdir, nb, tmp_path, SCOUT_DUMP_STATE=tmp_path / name)
        for k in range(1, nb + 1):
            a = (tmp_path / "wide" / f"{prog}_{k}.bin").read_bytes()
            assert a == (tmp_path / "narrow" / f"{prog}_{k}.bin").read_bytes(), f"{prog} batch {k}"


def test_mis_binary_compacts_the_egr_in_place(tmp_path, data):
    """Without SCOUT_SKIP_COMPACTION, scout_mis rewrites its .egr at the end of the run
    (src/mis/scout_mis.cu:949-965, core/flash_compactor.h:23-119): the file must equal the
    oracle's compaction of the logical graph after the batches, applied with MIS's rules
    (classification against the initial graph, conditional apply)."""
    egr, bdir, nb = data["sample"]
    say("[binary] scout_mis compaction")
    local = tmp_path / "sample.egr"
    local.write_bytes(open(egr, "rb").read())
    out = run_binary("mis", local, os.path.join(os.path.dirname(egr), "sample_mis.txt"), bdir, nb, tmp_path)
    assert "End-of-cycle compaction time" in out and "Peak program GPU memory:" in out
    og, base = OracleGraph(egr), OracleGraph(egr)
    for k, uv in batches(bdir, nb, og.n):
        og.apply(uv, base.classify(uv), conditional=True)
    want = tmp_path / "oracle.egr"
    from oracle_lib import lib
    assert lib().orc_compact_to(og.h, str(want).encode()) == 0
    assert local.read_bytes() == want.read_bytes()
    # the run wrote its final MIS where the reference does
    assert (tmp_path / "datasets" / "Final_MISs" / "sample_flash_final_mis.txt").exists()


def test_binary_switches_and_stage_lines(tmp_path, data):
    """SCOUT_BFS_BOUND is a scheduling switch: with b = 0 no batch falls back, with b = 1 the
    batches whose cascades pass one vertex per deleted parent edge are settled by a full BFS
    (fallback=1); the dumped levels are identical either way. [STAGE] lines carry the measured
    pack and copy times, and SCOUT_DUMP_TOUCHED writes the ascending ids of the rows read."""
    egr, bdir, nb = data["heavy"]
    init = str(tmp_path / "bfs_init.txt")
    run_gen("init", "bfs", egr, init)
    say("[binary] bound, stage lines, touched rows")
    outs = {}
    for b in (0, 1):
        outs[b] = run_binary("bfs", egr, init, bdir, nb, tmp_path, SCOUT_BFS_BOUND=b, SCOUT_DUMP_STATE=tmp_path / f"b{b}",
                             SCOUT_DUMP_TOUCHED=tmp_path / f"t{b}.bin", SCOUT_NO_MEMO=1)
        assert "[Deferred fallback: ON (default)]" in outs[b]
    fb = {b: sum(int(t.split("=")[1]) for t in o.split() if t.startswith("fallback=")) for b, o in outs.items()}
    assert fb[0] == 0 and fb[1] > 0, fb
    for k in range(1, nb + 1):
        assert (tmp_path / "b0" / f"bfs_{k}.bin").read_bytes() == (tmp_path / "b1" / f"bfs_{k}.bin").read_bytes()
    stages = [ln for ln in outs[0].splitlines() if ln.startswith("[STAGE]")]
    assert len(stages) == nb
    for ln in stages:
        f = dict(kv.split("=") for kv in ln.split()[1:])
        assert float(f["h2d_ms"]) > 0 and float(f["seed_pack_ms"]) >= 0 and float(f["repair_ms"]) > 0
    ids = np.fromfile(tmp_path / "t0.bin", dtype=np.int32)
    assert len(ids) > 0 and np.all(np.diff(ids) > 0) and ids[0] >= 0 and ids[-1] < sa.Graph.read_egr(egr).n
