"""GPU parity of the two BFS round orders: asynchronous (every level's deletion candidates in the
first invalidation round, every insertion seed in the first relax round) and level-synchronous.

The asynchronous invalidation has one path the level-synchronous one does not: a vertex whose
only possible support one level up is itself being decided in the same round waits for the next
round. The graph below plants many small structures that reach it:

    0 - a, 0 - b            level 1
    a - x, a - w, b - w     level 2 (x and w take a as parent: rows are sorted, the smallest id
                            one level up is the parent)
    x - y, w - y            level 3 (y takes x as parent)

and the batch deletes (a, w) and (x, y), listed next to each other: w and y are both candidates
of the first round, y's only other neighbour one level up is w, and w is still pending when y's
wave reads it (both are in one wave, or both in the wave kernel's list when their rows are
longer than a thread's share), so y waits a round and then takes w. Levels equal the oracle's
after every batch in both orders, and the device counter of waiting vertices is non-zero.
"""
import numpy as np
import pytest

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
    return n, np.cumsum(off), dst.astype(np.int32)


def gadget_graph(short, medium, leaves=20):
    """`short` gadgets with rows of at most 3 entries and `medium` gadgets whose w and y carry
    `leaves` extra leaves each (rows longer than 16 entries). Returns the CSR and the (a, w),
    (x, y) pairs of every gadget in that order."""
    pairs, dels = [], []
    nid = 1
    for j in range(short + medium):
        a, b, x, w, y = range(nid, nid + 5)
        nid += 5
        pairs += [(0, a), (0, b), (a, x), (a, w), (b, w), (x, y), (w, y)]
        if j >= short:
            for hub in (w, y):
                pairs += [(hub, nid + k) for k in range(leaves)]
                nid += leaves
        dels += [(a, w), (x, y)]
    n, off, col = csr_from_pairs(nid, pairs)
    return n, off, col, np.array(dels, dtype=np.int32)


def run_stream(mode, n, off, col, batches, schedule=None):
    og = OracleGraph(n=n, off=off, col=col)
    eng = sa.Engine(0)
    eng.load_graph(sa.Graph(n, off, col))
    eng.bfs_mode(mode)
    eng.upload(sa.STATE_LEVEL, og.bfs(0))
    rounds = []
    for k, uv in enumerate(batches, 1):
        eng.stage(uv)
        ins = eng.classify(sa.CLASSIFY_CURRENT)
        ins_o = og.classify(uv)
        assert np.array_equal(ins, ins_o), f"classification differs in batch {k}"
        eng.apply(False)
        og.apply(uv, ins_o)
        if schedule:
            eng.bfs_schedule(*schedule)
        st = eng.repair_bfs()
        want, got = og.bfs(0), eng.download(sa.STATE_LEVEL)
        assert np.array_equal(got, want), f"mode {mode} batch {k}: {np.count_nonzero(got != want)} levels differ"
        rounds.append((st.inval_rounds, st.relax_rounds))
    paths = eng.path_counts()
    eng.close()
    return rounds, paths


def test_async_invalidation_waits_for_pending_support():
    n, off, col, dels = gadget_graph(short=3000, medium=1500)
    rng = np.random.default_rng(5)
    # batch 1: the planted deletions; batch 2: re-insert them; batch 3: cut every w off
    # both its level-1 neighbours (w and its leaves cascade; y keeps x); batch 4: a random
    # mix of re-insertions
    a_w = dels[0::2]
    b_w = np.stack([a_w[:, 0] + 1, a_w[:, 1]], 1)
    batches = [dels, dels[:, ::-1].copy(), np.concatenate([a_w, b_w]).astype(np.int32),
               np.ascontiguousarray(np.concatenate([a_w, b_w])[rng.permutation(2 * len(a_w))[:2000]])]
    say(f"[bfs modes] gadget graph n={n} entries={len(col)}")
    r_async, p_async = run_stream(sa.BFS_ASYNC, n, off, col, batches)
    r_sync, p_sync = run_stream(sa.BFS_LEVEL_SYNC, n, off, col, batches)
    say(f"[bfs modes] async rounds {r_async}, deferred {p_async['inval_deferred']}; level-sync rounds {r_sync}")
    assert p_async["inval_deferred"] > 0, "no vertex waited for a pending support"
    assert p_sync["inval_deferred"] == 0


@pytest.mark.parametrize("mode", [sa.BFS_ASYNC, sa.BFS_LEVEL_SYNC])
def test_bfs_modes_exact_on_deep_cascades(mode):
    """A path of 3,000 vertices from vertex 0, also reached through a 6-edge detour to its
    middle: deleting the first edge invalidates the ~750 vertices the path still served, one
    cascade that both orders walk one vertex per round, and the relax then runs back up the
    path. Shortcuts inserted next are seeds at several depths, so the asynchronous relax
    lowers some vertices in more than one round. Levels equal the oracle's after every batch."""
    L = 3000
    path = [(i, i + 1) for i in range(L)]
    # a detour that keeps the tail reachable once (0, 1) goes: 0 - D1 - ... - D5 - 1500
    det = list(range(L + 1, L + 6))
    extra = [(0, det[0])] + [(det[i], det[i + 1]) for i in range(4)] + [(det[-1], 1500)]
    n, off, col = csr_from_pairs(L + 6, path + extra)
    shortcuts = np.array([[0, 2000], [0, 2500], [det[0], 700], [5, 2900], [0, 1200]], dtype=np.int32)
    batches = [np.array([[0, 1]], dtype=np.int32), shortcuts, np.array([[0, 1], [det[-1], 1500]], dtype=np.int32),
               shortcuts[:, ::-1].copy()]
    rounds, _ = run_stream(mode, n, off, col, batches)
    say(f"[bfs modes] path mode {mode}: rounds {rounds}")


@pytest.mark.parametrize("schedule", [(1, 1), (40, 40)])
def test_relax_queued_behind_the_invalidation_read_back_exact(schedule):
    """A first invalidation group sized in advance (here handed in with bfs_schedule; under the
    learned strategy by the prediction) is expected to end the invalidation, so the asynchronous
    repair queues the tentative levels, the seeds and the relax's first group behind that
    group's read-back, each kernel shut by a device gate unless the invalidation has ended. The
    asynchronous order queues two rounds whatever the schedule, so under both schedules the
    gadget graph's two-round cascades open the gate, and the 753-round path keeps it shut (the
    repair goes on with more invalidation rounds and queues the relax again). Levels equal the
    oracle's after every batch either way."""
    n, off, col, dels = gadget_graph(short=300, medium=150)
    a_w = dels[0::2]
    batches = [dels, dels[:, ::-1].copy(), a_w]
    rounds, _ = run_stream(sa.BFS_ASYNC, n, off, col, batches, schedule=schedule)
    say(f"[bfs modes] gadgets, schedule {schedule}: rounds {rounds}")
    L = 3000
    path = [(i, i + 1) for i in range(L)]
    det = list(range(L + 1, L + 6))
    extra = [(0, det[0])] + [(det[i], det[i + 1]) for i in range(4)] + [(det[-1], 1500)]
    n, off, col = csr_from_pairs(L + 6, path + extra)
    batches = [np.array([[0, 1]], dtype=np.int32), np.array([[0, 2000], [5, 2900]], dtype=np.int32)]
    rounds, _ = run_stream(sa.BFS_ASYNC, n, off, col, batches, schedule=schedule)
    say(f"[bfs modes] path, schedule {schedule}: rounds {rounds}")
