"""GPU tests of the learned scheduler: the HIP forward and backward kernels against the CPU
oracle's double-precision MLP, training on samples logged by the plain repair, and learned
repair == plain repair bit for bit, with what the network changes in the work measured.

Tolerances (fp32 device arithmetic against the float64 oracle, same bf16 weights): forward
outputs within 1e-4 absolute + 1e-4 relative; gradients within 2e-3 of the largest gradient
entry; losses within 1e-4 relative.
"""
import numpy as np
import pytest

import scout_amd as sa
from oracle_lib import OracleGraph, bf16_round, mlp_forward, mlp_loss_grad

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _case(seed, n):
    rng = np.random.default_rng(seed)
    feats = rng.standard_normal((n, 12)).astype(np.float32)
    targets = np.stack([rng.integers(0, 2, n), rng.random(n) * 3, rng.random(n) * 4], 1).astype(np.float32)
    return feats, targets


def test_bf16_weights_are_round_to_nearest_even():
    say("mlp: bf16 rounding of the master weights")
    with sa.Engine(0) as eng:
        eng.mlp_init(11)
        master = eng.mlp_get_params(rounded=False)
        w16 = eng.mlp_get_params(rounded=True)
    assert np.array_equal(w16, bf16_round(master))
    assert not np.array_equal(w16, master)


def test_forward_matches_oracle():
    say("mlp: forward vs oracle")
    feats, _ = _case(1, 777)
    with sa.Engine(0) as eng:
        eng.mlp_init(12)
        w = eng.mlp_get_params(rounded=True)
        out = eng.mlp_forward(feats)
    ref = mlp_forward(w, feats)
    assert np.allclose(out, ref, rtol=1e-4, atol=1e-4), np.abs(out - ref).max()


@pytest.mark.parametrize("n", [1, 100, 4096])
def test_gradient_matches_oracle(n):
    say(f"mlp: gradient vs oracle, n={n}")
    feats, targets = _case(2 + n, n)
    with sa.Engine(0) as eng:
        eng.mlp_init(13)
        w = eng.mlp_get_params(rounded=True)
        g, loss = eng.mlp_grad(feats, targets)
    ref_loss, ref_g = mlp_loss_grad(w, feats, targets)
    assert abs(loss - ref_loss) <= 1e-4 * abs(ref_loss)
    scale = np.abs(ref_g).max()
    assert np.abs(g - ref_g).max() <= 2e-3 * scale, (np.abs(g - ref_g).max(), scale)


def _stream_engine(scale, seed):
    eng = sa.Engine(0)
    eng.gen_rmat(scale, 16, 0.57, 0.19, 0.19, seed)
    eng.bfs_full(0)
    eng.cc_full()
    return eng


def test_training_on_logged_repairs_and_learned_equals_plain():
    """Logs 20 plain batches, trains on the device, then replays 6 more batches with the plain
    strategy, the learned one and the schedule a perfect predictor would choose (each batch's
    first invalidation and relax groups sized to the rounds the plain repair needed). Levels and
    labels equal the oracle after every learned batch and are identical across the three; the
    learned schedule needs fewer host read-backs than the plain one, and the perfect schedule
    bounds what any learned schedule can save."""
    scale, bsize, ntrain, ntest = 15, 20000, 20, 6
    say(f"mlp: logged training, R-MAT scale {scale}, {ntrain} + {ntest} batches of {bsize}")
    eng = _stream_engine(scale, 5)
    try:
        eng.mlp_log(True)
        total = ntrain + ntest
        eng.snapshot()
        for b in range(total):
            eng.gen_batch(bsize, 0.7, 0.5, 9, expected_total=bsize * total)
            eng.batch_store(b)
            eng.classify()
            eng.apply()
            eng.repair_bfs(sa.STRATEGY_LOCAL)
            eng.repair_cc(sa.STRATEGY_LOCAL)
            if b == ntrain - 1:
                eng.mlp_log(False)
        nsamp = eng.mlp_samples()
        assert nsamp >= 4096
        rows = eng.mlp_samples(download=True)
        assert np.isfinite(rows).all()
        assert set(np.unique(rows[:, 12])) <= {0.0, 1.0}
        eng.mlp_init(3)
        first, last = eng.mlp_train(epochs=8, lr=2e-3, seed=1)
        say(f"mlp: {nsamp} samples, loss {first:.4f} -> {last:.4f}")
        assert last < first

        def replay(strategy, schedule=None, oracle=False):
            eng.restore()
            for b in range(ntrain):  # bring the graph to the state after the training batches
                eng.batch_load(b)
                eng.classify()
                eng.apply()
                eng.repair_bfs(sa.STRATEGY_LOCAL)
                eng.repair_cc(sa.STRATEGY_LOCAL)
            og = None
            if oracle:
                g = eng.download_graph()
                og = OracleGraph(n=g.n, off=g.off, col=g.col)
            states, readbacks, rounds = [], 0, []
            for i, b in enumerate(range(ntrain, total)):
                m = eng.batch_load(b)
                eng.classify()
                eng.apply()
                if schedule:
                    eng.bfs_schedule(*schedule[i])
                st = eng.repair_bfs(strategy)
                eng.repair_cc(strategy)
                readbacks += st.readbacks
                rounds.append((st.inval_rounds, st.relax_rounds))
                lv, lab = eng.download(sa.STATE_LEVEL), eng.download(sa.STATE_LABEL)
                if og is not None:
                    uv = eng.batch_download(m)
                    og.apply(uv, og.classify(uv))
                    assert np.array_equal(lv, og.bfs(0)), f"learned batch {b}: levels differ from the oracle"
                    assert np.array_equal(lab, og.cc()), f"learned batch {b}: labels differ from the oracle"
                states.append((lv, lab))
            return states, readbacks, rounds

        plain, rb_plain, rounds = replay(sa.STRATEGY_LOCAL)
        learned, rb_learned, _ = replay(sa.STRATEGY_LEARNED, oracle=True)
        perfect, rb_perfect, _ = replay(sa.STRATEGY_LOCAL, schedule=[(max(a, 1), max(b, 1)) for a, b in rounds])
        say(f"mlp: BFS read-backs over {ntest} batches: plain {rb_plain}, learned {rb_learned}, "
            f"perfect schedule {rb_perfect}; rounds per batch {rounds}")
        for (lp, cp), (ll, cl), (lq, cq) in zip(plain, learned, perfect):
            assert np.array_equal(lp, ll) and np.array_equal(lp, lq)
            assert np.array_equal(cp, cl) and np.array_equal(cp, cq)
        assert rb_perfect <= rb_learned < rb_plain
        # and all equal a full recompute
        eng.bfs_full(0)
        eng.cc_full()
        assert np.array_equal(eng.download(sa.STATE_LEVEL), learned[-1][0])
        assert np.array_equal(eng.download(sa.STATE_LABEL), learned[-1][1])
    finally:
        eng.close()


def test_learned_recompute_switch_keeps_results_exact(tmp_path):
    """A network whose size head predicts huge cascades switches every batch to a full
    recompute; the answers stay those of the oracle."""
    from conftest import run_gen
    egr = tmp_path / "g.egr"
    run_gen("rmat", 11, 16, 4, egr)
    run_gen("stream", "tree", egr, tmp_path / "b", 3, 1500, 5, 0.7, 0.5)
    g = sa.Graph.read_egr(str(egr))
    og = OracleGraph(str(egr))
    with sa.Engine(0) as eng:
        eng.load_graph(g)
        eng.upload(sa.STATE_LEVEL, og.bfs(0))
        eng.upload(sa.STATE_LABEL, og.cc())
        p = np.zeros(sa.MLP_PARAMS, dtype=np.float32)
        p[5184 + 1] = 20.0  # b3[1]: log1p(size) = 20 for every candidate
        eng.mlp_set_params(p)
        from oracle_lib import read_batch
        import os
        for k in range(1, 4):
            uv = read_batch(os.path.join(tmp_path / "b", f"{k}.mtx"), g.n)
            eng.stage(uv)
            ins = eng.classify()
            eng.apply()
            og.apply(uv, og.classify(uv))
            st = eng.repair_bfs(sa.STRATEGY_LEARNED)
            eng.repair_cc(sa.STRATEGY_LEARNED)
            assert st.fallback == 1
            assert np.array_equal(eng.download(sa.STATE_LEVEL), og.bfs(0))
            assert np.array_equal(eng.download(sa.STATE_LABEL), og.cc())
            assert ins is not None


@pytest.mark.parametrize("logit,want", [(-20.0, "walk"), (20.0, "scan")])
def test_split_head_chooses_the_cc_listing(tmp_path, logit, want):
    """The split head's expected cut-off decides how the learned CC repair lists the vertices to
    relabel: a network sure that no deletion splits vertex 0's component leaves small batches to
    the walk, one sure that every deletion cuts off its predicted cascade moves them to the scan.
    Each batch deletes 10 edges to a vertex's only neighbour one level up (a parent edge in every
    BFS tree, so the network sees candidates). Labels equal the oracle's either way."""
    from conftest import run_gen
    egr = tmp_path / "g.egr"
    run_gen("rmat", 14, 16, 4, egr)
    g = sa.Graph.read_egr(str(egr))
    og = OracleGraph(str(egr))
    rng = np.random.default_rng(17)
    print(f"mlp: split head logit {logit}, n={g.n}, 10-line batches of parent-edge deletions", flush=True)
    with sa.Engine(0) as eng:
        eng.load_graph(g)
        eng.upload(sa.STATE_LEVEL, og.bfs(0))
        eng.upload(sa.STATE_LABEL, og.cc())
        p = np.zeros(sa.MLP_PARAMS, dtype=np.float32)
        p[5184 + 0] = logit  # b3[0]: logit of "ends unreached"
        p[5184 + 1] = 5.0    # b3[1]: log1p(cascade size) = 5, far below the recompute switch
        eng.mlp_set_params(p)
        for k in range(1, 4):
            lv = og.bfs(0)
            lines = []
            for v in rng.permutation(g.n):
                if lv[v] <= 0:
                    continue
                r = og.row(int(v))
                up = r[lv[r] == lv[v] - 1]
                if len(up) == 1:
                    lines.append((int(up[0]), int(v)))
                if len(lines) == 10:
                    break
            uv = np.array(lines, dtype=np.int32)
            eng.stage(uv)
            eng.classify()
            eng.apply()
            og.apply(uv, og.classify(uv))
            sb = eng.repair_bfs(sa.STRATEGY_LEARNED)
            sc = eng.repair_cc(sa.STRATEGY_LEARNED)
            assert sb.fallback == 0
            assert 2 * len(uv) * 256 <= g.n  # by size alone the batch is walked
            assert sc.cc_mode == (sa.CC_WALK if want == "walk" else sa.CC_SCAN), f"batch {k}"
            assert np.array_equal(eng.download(sa.STATE_LEVEL), og.bfs(0))
            assert np.array_equal(eng.download(sa.STATE_LABEL), og.cc())
