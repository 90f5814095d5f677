"""GPU parity on hub rows: the long-row code paths of the BFS and CC repairs against the oracle.

Rows longer than one wave's share are split into chunk tasks spread over many waves (BFS
invalidation support search and child pushes, tentative levels, relax; CC union-find), and an
invalid row whose one-level-down neighbours overflow the wave's LDS list re-reads its row.
The reference gives every row to one thread (src/bfs/scout_bfs.cu:73-105,
src/cc/scout_cc.cu:38-61), so these paths have no counterpart there; they must still give the
textbook answers. The graph is an R-MAT of scale 16 (hub rows up to ~9,700 entries) with
planted structures that force each path:
  * H1: a level-1 vertex with 2,000 leaves one level down (a medium row whose children do not
    fit the LDS list once its edge to vertex 0 is deleted);
  * H2: a level-1 vertex with 6,000 leaves (a long row that loses its only support);
  * H3: a level-2 vertex with 5,002 entries that keeps support after its parent edge goes;
and batches delete the edges of vertex 0 to its largest hubs, then every edge of vertex 0
(the whole graph leaves vertex 0's component), then re-insert hub edges (relax over hubs).
Levels and labels are compared with the oracle after every batch, and the device's path
counters must show that every long-row path ran.
"""
import numpy as np
import pytest

from conftest import run_gen
from oracle_lib import OracleGraph
import scout_amd as sa

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def csr_from_pairs(n, pairs):
    """Symmetric CSR with sorted rows, no duplicates, no self-loops."""
    p = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    p = p[p[:, 0] != p[:, 1]]
    both = np.concatenate([p, p[:, ::-1]])
    key = np.unique(both[:, 0] * n + both[:, 1])
    src, dst = key // n, key % n
    off = np.zeros(n + 1, dtype=np.int64)
    np.add.at(off, src + 1, 1)
    return np.cumsum(off), dst.astype(np.int32)


def hub_graph(tmp_path):
    p = tmp_path / "r16.egr"
    run_gen("rmat", 16, 16, 5, p)
    g = sa.Graph.read_egr(str(p))
    lv = OracleGraph(str(p)).bfs(0)
    rng = np.random.default_rng(7)
    src = np.repeat(np.arange(g.n), np.diff(g.off))
    pairs = [np.stack([src, g.col], 1)]
    deep = np.where(lv >= 2)[0]
    l1 = np.where(lv == 1)[0]
    nid = g.n
    plant = {}
    for name, leaves in (("H1", 2000), ("H2", 6000)):
        h = nid
        lf = np.arange(nid + 1, nid + 1 + leaves)
        nid += 1 + leaves
        pairs += [np.array([[0, h]]), np.stack([np.full(leaves, h), lf], 1),
                  np.stack([lf, rng.choice(deep, leaves)], 1)]
        plant[name] = h
    h3 = nid
    nid += 1
    nb = rng.choice(deep, 5000, replace=False)
    pairs += [np.array([[h3, l1.min()], [h3, l1.max()]]), np.stack([np.full(len(nb), h3), nb], 1)]
    plant["H3"] = h3
    plant["A"] = int(l1.min())
    n = nid
    off, col = csr_from_pairs(n, np.concatenate(pairs))
    return n, off, col, plant


def random_mix(og, n, rng, dels, inss, off, col):
    """dels existing edges (uniform over the loaded adjacency, kept if still present) and inss
    uniform non-edges."""
    out = []
    src = np.repeat(np.arange(n), np.diff(off))
    for k in rng.integers(0, len(col), 3 * dels):
        u, v = int(src[k]), int(col[k])
        if og.is_adjacent(u, v):
            out.append((u, v))
        if len(out) >= dels:
            break
    while len(out) < dels + inss:
        u, v = (int(x) for x in rng.integers(0, n, 2))
        if u != v and not og.is_adjacent(u, v):
            out.append((u, v))
    return out


def dedup(lines):
    seen, out = set(), []
    for u, v in lines:
        k = (min(u, v), max(u, v))
        if k not in seen and u != v:
            seen.add(k)
            out.append((u, v))
    return np.array(out, dtype=np.int32).reshape(-1, 2)


def test_hub_rows_bfs_cc_exact_every_batch(tmp_path):
    n, off, col, plant = hub_graph(tmp_path)
    deg = np.diff(off)
    say(f"[hubs] n={n} entries={len(col)} largest row {deg.max()} rows>4096 {(deg > 4096).sum()}")
    assert deg.max() > 4096 and deg[plant["H1"]] < 4096 < deg[plant["H2"]]
    og = OracleGraph(n=n, off=off, col=col)
    g = sa.Graph(n, off, col)
    eng = sa.Engine(0)
    eng.load_graph(g)
    eng.cc_mode(sa.CC_WALK)
    cc_only = sa.Engine(0)  # CC without BFS maintained: its own levels from vertex 0
    cc_only.load_graph(g)
    cc_only.cc_mode(sa.CC_SCAN)
    lv0 = og.bfs(0)
    eng.upload(sa.STATE_LEVEL, lv0)
    eng.upload(sa.STATE_LABEL, og.cc())
    cc_only.upload(sa.STATE_LABEL, og.cc())
    rng = np.random.default_rng(11)
    hubs = [int(v) for v in np.argsort(-deg) if lv0[v] == 1][:6]

    def batch(k):
        lv = og.bfs(0)
        if k == 1:
            lines = [(0, plant["H1"]), (0, plant["H2"]), (plant["H3"], plant["A"])] + [(0, h) for h in hubs]
            return lines + random_mix(og, n, rng, 400, 400, off, col)
        if k == 2:
            lines = []
            for h in hubs + [plant["H3"]]:
                r = og.row(h)
                up = np.sort(r[lv[r] == lv[h] - 1])[:3] if lv[h] > 0 else []
                lines += [(h, int(x)) for x in up]
            return lines + random_mix(og, n, rng, 300, 300, off, col)
        if k == 3:
            return [(0, int(x)) for x in og.row(0)]
        if k == 4:
            return [(0, h) for h in hubs + [plant["H1"], plant["H2"]]] + random_mix(og, n, rng, 200, 600, off, col)
        return random_mix(og, n, rng, 1500, 1500, off, col)

    for k in range(1, 6):
        uv = dedup(batch(k))
        for e in (eng, cc_only):
            e.stage(uv)
            ins = e.classify(sa.CLASSIFY_CURRENT)
        ins_o = og.classify(uv)
        assert np.array_equal(ins, ins_o), f"classification differs in batch {k}"
        og.apply(uv, ins_o)
        eng.apply(False)
        cc_only.apply(False)
        sb = eng.repair_bfs()
        eng.repair_cc()
        cc_only.repair_cc()
        want_lv, want_lab = og.bfs(0), og.cc()
        lv, lab = eng.download(sa.STATE_LEVEL), eng.download(sa.STATE_LABEL)
        assert np.array_equal(lv, want_lv), f"batch {k}: {np.count_nonzero(lv != want_lv)} levels differ"
        assert np.array_equal(lab, want_lab), f"batch {k}: {np.count_nonzero(lab != want_lab)} labels differ"
        assert np.array_equal(cc_only.download(sa.STATE_LABEL), want_lab), f"batch {k}: CC-only labels differ"
        say(f"[hubs] batch {k}: {len(uv)} lines, {sb.touched} invalidated, {sb.rounds} rounds, "
            f"reached {int((want_lv >= 0).sum())}, paths {eng.path_counts()}")
    # full recomputes over the hub rows
    eng.bfs_full(0)
    eng.cc_full()
    assert np.array_equal(eng.download(sa.STATE_LEVEL), og.bfs(0))
    assert np.array_equal(eng.download(sa.STATE_LABEL), og.cc())
    paths = eng.path_counts()
    say(f"[hubs] path counts {paths}")
    for name in sa.LONG_ROW_PATHS:
        assert paths[name] > 0, f"the {name} path never ran"
    assert cc_only.path_counts()["cc_tasks"] > 0 and cc_only.path_counts()["cc_walk_rows"] == 0
    eng.close()
    cc_only.close()
