"""Pins the CPU oracle (oracle/oracle.cpp) before any GPU result is compared with it.

Every oracle function is checked against something it does not compute itself: the
reference's golden fixtures (datasets/sample), textbook library routines (scipy.sparse.csgraph),
hand-worked cases of the reference's rules, or invariants that a plausible mistake breaks.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components, shortest_path

from oracle_lib import OracleGraph, mis_max_csr, read_batch


def csr_from_edges(n, edges):
    """Symmetric, sorted, duplicate-free CSR from undirected edges."""
    keys = set()
    for u, v in edges:
        if u != v:
            keys.add((u, v))
            keys.add((v, u))
    keys = sorted(keys)
    off = np.zeros(n + 1, dtype=np.int64)
    for u, _ in keys:
        off[u + 1] += 1
    off = np.cumsum(off)
    col = np.array([v for _, v in keys], dtype=np.int32)
    return off, col


def random_graph(n, m, rng):
    edges = set()
    while len(edges) < m:
        u, v = rng.integers(0, n, 2)
        if u != v:
            edges.add((min(u, v), max(u, v)))
    return sorted(edges)


def scipy_graph(g: OracleGraph):
    """Adjacency of the logical rows (a batch that repeats a pair can leave a tombstone on a
    pair the base never held, and then the reference's degree formula, used by its
    compactor, no longer matches the row; the rows themselves are the graph)."""
    rows = [g.row(v) for v in range(g.n)]
    off = np.concatenate([[0], np.cumsum([r.size for r in rows])]).astype(np.int64)
    col = np.concatenate(rows).astype(np.int32) if off[-1] else np.zeros(0, np.int32)
    return sp.csr_matrix((np.ones(col.size), col, off), shape=(g.n, g.n))


# --- golden fixtures (datasets/sample) ------------------------------------------------

def test_sample_fixtures_are_closed_forms(sample_dir):
    g = OracleGraph(os.path.join(sample_dir, "sample.egr"))
    assert g.n == 2000
    ref = np.loadtxt(os.path.join(sample_dir, "sample_bfs.txt"), dtype=np.int64)
    want = np.full(g.n, -1, dtype=np.int32)
    want[ref[:, 0] - 1] = ref[:, 1]
    assert np.array_equal(g.bfs(0), want)
    assert np.array_equal(g.cc(), np.loadtxt(os.path.join(sample_dir, "sample_components.txt"), dtype=np.int32))
    mis = np.zeros(g.n, dtype=np.uint8)
    mis[np.loadtxt(os.path.join(sample_dir, "sample_mis.txt"), dtype=np.int64) - 1] = 1
    assert np.array_equal(g.greedy_mis(), mis)
    colors = np.loadtxt(os.path.join(sample_dir, "sample_colors.txt"), dtype=np.int32)
    assert g.check_coloring(colors) == 0  # ECL-GC output: proper, 6 colours
    assert colors.max() + 1 == 6


def test_sample_batches_are_deletions_of_base_edges(sample_dir):
    g = OracleGraph(os.path.join(sample_dir, "sample.egr"))
    for k in range(1, 6):
        uv = read_batch(os.path.join(sample_dir, "batches", f"{k}.mtx"), g.n)
        assert uv.shape == (50, 2)
        ins = g.classify(uv)
        assert not ins.any()
        g.apply(uv, ins)
        for u, v in uv:
            assert not g.is_adjacent(u, v) and not g.is_adjacent(v, u)


# --- closed forms vs library routines ---------------------------------------------------

@pytest.mark.parametrize("seed", range(6))
def test_bfs_and_cc_match_scipy_after_random_updates(seed):
    rng = np.random.default_rng(seed)
    n = 60
    off, col = csr_from_edges(n, random_graph(n, 70, rng))
    g = OracleGraph(n=n, off=off, col=col)
    for _ in range(4):
        uv = rng.integers(0, n, (25, 2)).astype(np.int32)
        uv = uv[uv[:, 0] != uv[:, 1]]
        g.apply(uv, g.classify(uv))
        a = scipy_graph(g)
        d = shortest_path(a, unweighted=True, indices=0)
        want = np.where(np.isinf(d), -1, d).astype(np.int32)
        assert np.array_equal(g.bfs(0), want)
        ncomp, lab = connected_components(a, directed=False)
        mins = np.full(ncomp, n, dtype=np.int64)
        np.minimum.at(mins, lab, np.arange(n))
        assert np.array_equal(g.cc(), mins[lab].astype(np.int32))
        assert g.check_bfs(g.bfs(0)) == 0


def test_validity_checks_catch_planted_mistakes():
    rng = np.random.default_rng(7)
    n = 80
    off, col = csr_from_edges(n, random_graph(n, 160, rng))
    g = OracleGraph(n=n, off=off, col=col)
    lv = g.bfs(0)
    bad = lv.copy()
    k = int(np.flatnonzero(lv > 1)[0])
    bad[k] += 1                      # an off-by-one level
    assert g.check_bfs(bad) > 0
    bad = lv.copy()
    bad[k] -= 1
    assert g.check_bfs(bad) > 0
    lab = g.cc()
    bad = lab.copy()
    bad[n - 1] = n - 1 if lab[n - 1] != n - 1 else 0
    assert g.check_cc(bad) > 0
    colors = g.greedy_color()
    assert g.check_coloring(colors) == 0
    u = int(np.flatnonzero(off[1:] - off[:-1])[0])
    w = int(col[off[u]])
    bad = colors.copy()
    bad[w] = bad[u]
    assert g.check_coloring(bad) > 0
    mis = g.greedy_mis()
    assert g.check_mis_independent(mis) == 0 and g.mis_maximality_violations(mis) == 0
    bad = mis.copy()
    bad[w] = bad[u] = 1
    assert g.check_mis_independent(bad) > 0


def test_greedy_states_are_first_fit():
    rng = np.random.default_rng(3)
    n = 50
    off, col = csr_from_edges(n, random_graph(n, 120, rng))
    g = OracleGraph(n=n, off=off, col=col)
    c = g.greedy_color()
    for v in range(n):
        lower = {int(c[w]) for w in col[off[v]:off[v + 1]] if w < v}
        assert c[v] == min(set(range(len(lower) + 1)) - lower)
    m = g.greedy_mis()
    for v in range(n):
        lower_in = any(m[w] for w in col[off[v]:off[v + 1]] if w < v)
        assert m[v] == (0 if lower_in else 1)


# --- scratchpad semantics (core/FlashGraph.h:200-310) -------------------------------------

def test_scratchpad_row_order_and_toggles():
    # base: 0-1, 0-2, 0-3 ; rows sorted
    off, col = csr_from_edges(5, [(0, 1), (0, 2), (0, 3)])
    g = OracleGraph(n=5, off=off, col=col)
    uv = np.array([[0, 2], [0, 4], [1, 2], [0, 3]], dtype=np.int32)
    ins = g.classify(uv)
    assert ins.tolist() == [0, 1, 1, 0]
    g.apply(uv, ins)
    # base row minus deleted entries, then inserted entries in insertion order
    assert g.row(0).tolist() == [1, 4]
    assert g.row(2).tolist() == [1]
    assert g.row(1).tolist() == [0, 2]
    # re-inserting a deleted base edge restores it in its base position
    uv2 = np.array([[2, 0], [4, 0]], dtype=np.int32)
    ins2 = g.classify(uv2)
    assert ins2.tolist() == [1, 0]
    g.apply(uv2, ins2)
    assert g.row(0).tolist() == [1, 2]
    assert g.row(4).tolist() == []
    assert g.degree(0) == 2


@pytest.mark.parametrize("seed", range(4))
def test_scratchpad_toggles_are_symmetric_and_consistent(seed):
    rng = np.random.default_rng(100 + seed)
    n = 30
    off, col = csr_from_edges(n, random_graph(n, 40, rng))
    g = OracleGraph(n=n, off=off, col=col)
    adj = np.zeros((n, n), dtype=bool)
    adj[np.repeat(np.arange(n), np.diff(off)), col] = True
    for _ in range(20):
        uv = rng.integers(0, n, (12, 2)).astype(np.int32)
        uv = uv[uv[:, 0] != uv[:, 1]]
        # distinct unordered pairs per batch, as the generators emit
        _, keep = np.unique(np.sort(uv, axis=1), axis=0, return_index=True)
        uv = uv[np.sort(keep)]
        ins = g.classify(uv)
        assert np.array_equal(ins.astype(bool), ~adj[uv[:, 0], uv[:, 1]])
        g.apply(uv, ins)
        adj[uv[:, 0], uv[:, 1]] ^= True
        adj[uv[:, 1], uv[:, 0]] ^= True
        for v in range(n):
            r = g.row(v)
            assert sorted(r.tolist()) == np.flatnonzero(adj[v]).tolist()
            assert g.degree(v) == r.size


# --- colouring rounds (src/coloring/scout_coloring.cu:46-100,243-277) ----------------------

def test_coloring_rounds_hand_worked():
    g = OracleGraph(n=4, off=np.zeros(5, dtype=np.int64), col=np.zeros(0, dtype=np.int32))
    uv = np.array([[0, 1], [1, 2], [2, 3]], dtype=np.int32)
    ins = g.classify(uv)
    g.apply(uv, ins)
    colors, st = g.color_batch(uv, ins, np.zeros(4, dtype=np.int32))
    # marks {1,2,3}; round 1 proposes 1,1,1 (2 and 3 lose); round 2 proposes 0,0 (3 loses);
    # round 3 gives 3 colour 1
    assert colors.tolist() == [0, 1, 0, 1]
    assert st.tolist() == [3, 3]


def test_coloring_path_insert():
    off, col = csr_from_edges(3, [(0, 1), (1, 2)])
    g = OracleGraph(n=3, off=off, col=col)
    uv = np.array([[0, 2]], dtype=np.int32)
    ins = g.classify(uv)
    g.apply(uv, ins)
    colors, st = g.color_batch(uv, ins, np.array([0, 1, 0], dtype=np.int32))
    assert colors.tolist() == [0, 1, 2] and st.tolist() == [1, 1]


@pytest.mark.parametrize("seed", range(5))
def test_coloring_stays_proper(seed):
    rng = np.random.default_rng(seed)
    n = 200
    off, col = csr_from_edges(n, random_graph(n, 600, rng))
    g = OracleGraph(n=n, off=off, col=col)
    colors = g.greedy_color()
    for _ in range(5):
        uv = rng.integers(0, n, (80, 2)).astype(np.int32)
        uv = uv[uv[:, 0] != uv[:, 1]]
        ins = g.classify(uv)
        g.apply(uv, ins)
        colors, _ = g.color_batch(uv, ins, colors)
        assert g.check_coloring(colors) == 0


# --- MIS rules (src/mis/scout_mis.cu:290-362) -----------------------------------------------

BIG = mis_max_csr(16.0, 1e12)


def _mis_case(n, edges, members, batch):
    off, col = csr_from_edges(n, edges)
    g = OracleGraph(n=n, off=off, col=col)
    uv = np.array(batch, dtype=np.int32).reshape(-1, 2)
    ins = g.classify(uv)
    g.apply(uv, ins, conditional=True)
    mis = np.zeros(n, dtype=np.uint8)
    mis[list(members)] = 1
    out, st = g.mis_batch(uv, ins, mis, BIG)
    return set(np.flatnonzero(out).tolist()), st


def test_mis_deletion_rejected_by_two_hop_member():
    # path 0-1-2-3, M = {0, 3}; deleting (0,1) leaves 1 undominated but 3 is in its 2-hop
    # ball, so 1 stays out (the reference does not restore maximality here)
    out, _ = _mis_case(4, [(0, 1), (1, 2), (2, 3)], {0, 3}, [0, 1])
    assert out == {0, 3}


def test_mis_deletion_accepted():
    out, _ = _mis_case(2, [(0, 1)], {0}, [0, 1])
    assert out == {0, 1}


def test_mis_insertion_adds_first_addable_in_ball_order():
    # 1-2 edge, 3-4 edge, M = {0, 1, 3}; insert (0,1): 1 leaves; ball(1) = [1, 2, 0];
    # 2's hop-1 row [1] holds no member, so 2 joins
    out, _ = _mis_case(5, [(1, 2), (3, 4)], {0, 1, 3}, [0, 1])
    assert out == {0, 2, 3}


def test_mis_insertion_skips_vertices_without_row():
    # edges 1-2, 2-3, 2-4; M = {0, 1, 4}; insert (0,1): ball(1) = [1, 2, 0, 3, 4];
    # 2's row holds 4; 3 is two hops from the endpoints and has no row, so it is skipped
    out, _ = _mis_case(5, [(1, 2), (2, 3), (2, 4)], {0, 1, 4}, [0, 1])
    assert out == {0, 4}


def _mis_case_csr(n, edges, members, batch, max_csr):
    off, col = csr_from_edges(n, edges)
    g = OracleGraph(n=n, off=off, col=col)
    uv = np.array(batch, dtype=np.int32).reshape(-1, 2)
    ins = g.classify(uv)
    g.apply(uv, ins, conditional=True)
    mis = np.zeros(n, dtype=np.uint8)
    mis[list(members)] = 1
    out, st = g.mis_batch(uv, ins, mis, max_csr)
    return set(np.flatnonzero(out).tolist()), int(st[0])


def test_mis_sub_batch_cut_decides_the_row_kinds():
    """The deg*1000 estimate against MAX_CSR_ELEMENTS (src/mis/scout_mis.cu:519-542) cuts the
    batch, and the cut changes the answer, because a vertex's row is its 2-hop ball only when it
    is an endpoint of a line in the same sub-batch. Path 1-2-3-4 plus vertices 0 and 5,
    M = {0, 1, 4}; lines: insert (0,1), insert (2,5). Degrees after the batch: 0:1, 1:2, 2:3,
    5:1, so the first line estimates 3000 and the second 4000.
      * max_csr 7001 (3000 + 4000 < 7001): one sub-batch. 1 leaves; ball(1) = [1, 2, 0, 3, 5];
        2 is an endpoint, its ball holds members 0 and 4; 3's adjacency row holds 4; 5's ball
        [5, 2, 1, 3] holds none, so 5 joins: {0, 4, 5}.
      * max_csr 7000 (3000 + 4000 >= 7000 cuts): line 1 alone; 2 is only next to an endpoint,
        its adjacency row [1, 3, 5] holds no member, so 2 joins; line 2 then has one member
        endpoint on an insertion and changes nothing: {0, 2, 4}.
    Degrees taken before the batch (0:0, 1:1, 2:2, 5:0) would keep one sub-batch at 7000."""
    edges = [(1, 2), (2, 3), (3, 4)]
    batch = [0, 1, 2, 5]
    out, subs = _mis_case_csr(6, edges, {0, 1, 4}, batch, 7001)
    assert subs == 1 and out == {0, 4, 5}
    out, subs = _mis_case_csr(6, edges, {0, 1, 4}, batch, 7000)
    assert subs == 2 and out == {0, 2, 4}


def test_mis_complete_restores_maximality_on_a_hand_case():
    # path 0-1-2-3, M = {0, 3}; deleting (0,1) leaves 1 undominated (the reference's rule keeps
    # it out: 3 is in its 2-hop ball); the completion scans the region {0, 1, 2, 3} in id
    # order and adds 1 (no member neighbour left), after which 2 stays dominated
    n = 4
    off, col = csr_from_edges(n, [(0, 1), (1, 2), (2, 3)])
    g = OracleGraph(n=n, off=off, col=col)
    uv = np.array([[0, 1]], dtype=np.int32)
    ins = g.classify(uv)
    g.apply(uv, ins, conditional=True)
    mis = np.array([1, 0, 0, 1], dtype=np.uint8)
    mis, _ = g.mis_batch(uv, ins, mis, BIG)
    assert g.mis_maximality_violations(mis) == 1
    done, added = g.mis_complete(uv, mis)
    assert added == 1 and np.flatnonzero(done).tolist() == [0, 1, 3]
    assert g.mis_maximality_violations(done) == 0


@pytest.mark.parametrize("seed", range(4))
def test_mis_complete_keeps_the_set_maximal_and_touches_only_the_region(seed):
    """Rules, then completion, batch after batch from the greedy set: independent and maximal
    on the whole graph after every batch; members are never removed by the completion, and
    vertices more than two hops from every endpoint keep their membership."""
    rng = np.random.default_rng(seed)
    n = 300
    off, col = csr_from_edges(n, random_graph(n, 700, rng))
    g = OracleGraph(n=n, off=off, col=col)
    mis = g.greedy_mis()
    for _ in range(6):
        uv = rng.integers(0, n, (40, 2)).astype(np.int32)
        uv = uv[uv[:, 0] != uv[:, 1]]
        _, keep = np.unique(np.sort(uv, axis=1), axis=0, return_index=True)
        uv = uv[np.sort(keep)]
        ins = g.classify(uv)
        g.apply(uv, ins, conditional=True)
        mis, _ = g.mis_batch(uv, ins, mis, BIG)
        done, added = g.mis_complete(uv, mis)
        assert added == int(done.sum()) - int(mis.sum()) and np.all(done >= mis)
        # the 2-hop region on the graph after the batch
        region = set(uv.ravel().tolist())
        for _hop in range(2):
            region |= {int(w) for v in list(region) for w in g.row(v)}
        changed = set(np.flatnonzero(done != mis).tolist())
        assert changed <= region
        assert g.check_mis_independent(done) == 0
        assert g.mis_maximality_violations(done) == 0
        mis = done


def test_mis_max_csr_formula():
    assert mis_max_csr(16.0, 288e9) == min(800000000, int(1048576 * 2 * 256))
    assert mis_max_csr(0.0, 288e9) == 50000000
    assert mis_max_csr(100.0, 8e9) == int(8e9 / 32)


@pytest.mark.parametrize("seed", range(4))
def test_mis_rules_keep_independence(seed):
    rng = np.random.default_rng(seed)
    n = 150
    off, col = csr_from_edges(n, random_graph(n, 400, rng))
    g = OracleGraph(n=n, off=off, col=col)
    mis = g.greedy_mis()
    for _ in range(5):
        uv = rng.integers(0, n, (60, 2)).astype(np.int32)
        uv = uv[uv[:, 0] != uv[:, 1]]
        _, keep = np.unique(np.sort(uv, axis=1), axis=0, return_index=True)
        uv = uv[np.sort(keep)]
        ins = g.classify(uv)
        g.apply(uv, ins, conditional=True)
        mis, _ = g.mis_batch(uv, ins, mis, BIG)
        assert g.check_mis_independent(mis) == 0


def _distinct_stream(og, rng, m, del_frac):
    uv = []
    for _ in range(m):
        u = int(rng.integers(og.n))
        row = og.row(u)
        if rng.random() < del_frac and len(row):
            uv.append((u, int(row[rng.integers(len(row))])))
        else:
            v = int(rng.integers(og.n))
            if u != v:
                uv.append((u, v))
    uv = np.array(uv, dtype=np.int32)
    key = np.minimum(uv[:, 0], uv[:, 1]).astype(np.int64) << 32 | np.maximum(uv[:, 0], uv[:, 1])
    _, first = np.unique(key, return_index=True)
    return np.ascontiguousarray(uv[np.sort(first)])


@pytest.mark.parametrize("seed", [0, 1])
def test_per_edit_reference_cc_is_exact_every_batch(sample_dir, seed):
    """The per-edit CC baseline (bidirectional search, src/cc/scout_cc.cu:300-441) with no
    bound settles every deletion, so its canonicalised labels equal the closed form after
    every batch."""
    from oracle_lib import _p, lib
    egr = os.path.join(sample_dir, "sample.egr")
    rng = np.random.default_rng(seed)
    og, g2 = OracleGraph(egr), OracleGraph(egr)
    L = lib()
    rc = L.orc_refcc_create(g2.h, _p(og.cc()), 1 << 40, 0)
    try:
        for _ in range(5):
            uv = _distinct_stream(og, rng, 300, 0.7)
            ins = og.classify(uv)
            og.apply(uv, ins)
            st = np.zeros(8, dtype=np.int64)
            L.orc_refcc_batch(rc, _p(uv), _p(ins), len(uv), _p(st))
            lab = np.zeros(og.n, dtype=np.int32)
            L.orc_refcc_finish(rc, _p(lab))
            assert np.array_equal(lab, og.cc())
            assert st[7] == 0  # nothing inconclusive without a bound
    finally:
        L.orc_refcc_free(rc)


@pytest.mark.parametrize("seed", [0, 1])
def test_per_edit_reference_bfs_insertions_are_exact(sample_dir, seed):
    """The per-edit BFS baseline's insertion relax (src/bfs/scout_bfs.cu:460-480) is exact on
    insertion-only streams; its deletion worklist (min neighbour level + 1) is not exact in
    general, which is why the baseline is timed and not used as a test oracle."""
    from oracle_lib import _p, lib
    egr = os.path.join(sample_dir, "sample.egr")
    rng = np.random.default_rng(seed)
    og, g2 = OracleGraph(egr), OracleGraph(egr)
    L = lib()
    rb = L.orc_refbfs_create(g2.h, _p(og.bfs(0)), 0, 1 << 40, 0)
    try:
        for _ in range(4):
            uv = _distinct_stream(og, rng, 200, 0.0)
            uv = np.ascontiguousarray(uv[og.classify(uv) == 1])  # random pairs that are not edges
            ins = og.classify(uv)
            assert ins.all()
            og.apply(uv, ins)
            st = np.zeros(8, dtype=np.int64)
            L.orc_refbfs_batch(rb, _p(uv), _p(ins), len(uv), _p(st))
            lv = np.zeros(og.n, dtype=np.int32)
            L.orc_refbfs_finish(rb, _p(lv))
            assert np.array_equal(lv, og.bfs(0))
    finally:
        L.orc_refbfs_free(rb)


# --- learned scheduler's MLP (oracle side) ---------------------------------------------

def _mlp_case(seed, n=40):
    rng = np.random.default_rng(seed)
    params = (rng.standard_normal(5187) * 0.2).astype(np.float32)
    feats = rng.standard_normal((n, 12)).astype(np.float32)
    targets = np.stack([rng.integers(0, 2, n), rng.random(n) * 3, rng.random(n) * 4], 1).astype(np.float32)
    return params, feats, targets


def _numpy_forward(p, x):
    """Independent forward pass (float64 numpy) from the documented parameter layout."""
    p = p.astype(np.float64)
    w1 = p[0:768].reshape(12, 64); b1 = p[768:832]
    w2 = p[832:4928].reshape(64, 64); b2 = p[4928:4992]
    w3 = p[4992:5184].reshape(64, 3); b3 = p[5184:5187]
    h1 = np.maximum(x.astype(np.float64) @ w1 + b1, 0)
    h2 = np.maximum(h1 @ w2 + b2, 0)
    return h2 @ w3 + b3


def test_mlp_oracle_forward_matches_numpy():
    from oracle_lib import mlp_forward
    p, x, _ = _mlp_case(3)
    assert np.allclose(mlp_forward(p, x), _numpy_forward(p, x), rtol=1e-12, atol=1e-12)


def test_mlp_oracle_loss_matches_numpy_closed_form():
    from oracle_lib import mlp_loss_grad
    p, x, t = _mlp_case(4)
    out = _numpy_forward(p, x)
    z = out[:, 0]
    bce = np.maximum(z, 0) - z * t[:, 0] + np.log1p(np.exp(-np.abs(z)))
    expect = np.mean(bce + 0.5 * ((out[:, 1] - t[:, 1]) ** 2 + (out[:, 2] - t[:, 2]) ** 2))
    loss, _ = mlp_loss_grad(p, x, t, with_grad=False)
    assert abs(loss - expect) < 1e-10


@pytest.mark.parametrize("seed", [5, 6])
def test_mlp_oracle_gradient_matches_finite_differences(seed):
    """Central differences of the oracle's own loss in float64 pin its backward pass; every
    parameter block (W1, b1, W2, b2, W3, b3) is probed."""
    from oracle_lib import mlp_loss_grad
    p, x, t = _mlp_case(seed)
    _, g = mlp_loss_grad(p, x, t)
    rng = np.random.default_rng(seed)
    blocks = [(0, 768), (768, 832), (832, 4928), (4928, 4992), (4992, 5184), (5184, 5187)]
    idx = np.concatenate([rng.integers(lo, hi, size=min(8, hi - lo)) for lo, hi in blocks])
    # the parameters are float32: step in float64 via a scaled copy of the loss
    pd = p.astype(np.float64)
    for i in idx:
        h = 1e-6
        up, dn = pd.copy(), pd.copy()
        up[i] += h
        dn[i] -= h
        # evaluate with the numpy closed form in float64 so the step is exact
        def loss_of(q):
            out = _numpy_forward(q, x)
            z = out[:, 0]
            bce = np.maximum(z, 0) - z * t[:, 0] + np.log1p(np.exp(-np.abs(z)))
            return np.mean(bce + 0.5 * ((out[:, 1] - t[:, 1]) ** 2 + (out[:, 2] - t[:, 2]) ** 2))
        fd = (loss_of(up) - loss_of(dn)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_mlp_oracle_gradient_catches_a_planted_mistake():
    """A backward pass that drops the layer-1 ReLU mask gives a W1 gradient far from the
    oracle's (which the finite differences above pin), so that mistake cannot pass."""
    from oracle_lib import mlp_loss_grad
    p, x, t = _mlp_case(7)
    _, g = mlp_loss_grad(p, x, t)
    q = p.astype(np.float64)
    w1, b1 = q[0:768].reshape(12, 64), q[768:832]
    w2, b2 = q[832:4928].reshape(64, 64), q[4928:4992]
    w3, b3 = q[4992:5184].reshape(64, 3), q[5184:5187]
    xd = x.astype(np.float64)
    z1 = xd @ w1 + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w2 + b2
    out = np.maximum(z2, 0) @ w3 + b3
    d3 = np.stack([1 / (1 + np.exp(-out[:, 0])) - t[:, 0], out[:, 1] - t[:, 1], out[:, 2] - t[:, 2]], 1) / len(x)
    d2 = (d3 @ w3.T) * (z2 > 0)
    right = xd.T @ ((d2 @ w2.T) * (z1 > 0))
    planted = xd.T @ (d2 @ w2.T)
    assert np.allclose(right.ravel(), g[0:768], rtol=1e-9, atol=1e-12)
    assert np.abs(planted.ravel() - g[0:768]).max() > 1e-2 * np.abs(g[0:768]).max()


def test_bf16_round_is_nearest_even():
    from oracle_lib import bf16_round
    def rne(v):
        u = np.array([v], dtype=np.float32).view(np.uint32)[0].item()
        lsb = (u >> 16) & 1
        r = ((u + 0x7FFF + lsb) & 0xFFFF0000) & 0xFFFFFFFF
        return np.array([r], dtype=np.uint32).view(np.float32)[0]
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # bf16 spacing at 1.0
    cases = [one, one + ulp / 2, one + ulp + ulp / 2, one + ulp / 4, -(one + 3 * ulp / 2), np.float32(3.14159),
             np.float32(1e-30), np.float32(-7.5e5)]
    for v in cases:
        assert bf16_round(v) == rne(v)
    assert bf16_round(one + ulp / 2) == one                     # tie to even (down)
    assert bf16_round(one + ulp + ulp / 2) == one + 2 * ulp     # tie to even (up)
