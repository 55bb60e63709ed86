"""GPU: the learned proposal against the oracle.

- single learned moves (features, network, Gaussian step, Hastings term,
  energy) on the same walkers as the oracle, within float tolerance (the
  device's exp/log/sin differ from the host's in the last bits, so learned
  trajectories are compared move by move, not bit for bit);
- probe records from the same walkers and streams;
- the training gradient and loss on the same records and weights;
- Adam's first step and a falling training loss;
- detailed balance on the device: at infinite temperature two atoms must
  end uniform in the sphere (closed-form pair-distance law);
- what the trained network gains over the fixed step on held-out records.
"""
import numpy as np
import pytest
from scipy import stats

import oracle as O
import sadmc_amd as S
from test_oracle_proposal import _hand_weights, _pair_distance_cdf, init_weights, synthetic_records

pytestmark = pytest.mark.gpu


def lj13(lanes=4, **kw):
    base = dict(system=S.LJ, n=13, lanes=lanes, method=S.SAD, method_param=0.05, move_param=0.1, energy_bin=0.01,
                has_max_e=1, max_e=0.0)
    base.update(kw)
    return S.make_params(**base)


def test_learned_moves_match_oracle(gpu_ctx_factory):
    print("case learned plan vs oracle", flush=True)
    w = init_weights(3)
    ctx = gpu_ctx_factory()
    ctx.prop_set_weights(w)
    sp = lj13(move_kind=S.MOVE_LEARNED)
    seeds = np.arange(4, dtype=np.uint64) + 50
    sid = ctx.add_set(sp, seeds, geom=2.5)
    ctx.init()
    worst = 0.0
    for k, seed in enumerate(seeds):
        orc = O.Walker(O.make_params(system=O.LJ, n=13, radius=2.5, method=O.SAD, method_param=0.05, move_param=0.1,
                                     energy_bin=0.01, has_max_e=1, max_e=0.0, order_g=4, seed=int(seed), mlp=w))
        np.testing.assert_array_equal(ctx.coords(sid, k), orc.coords())
        for step in range(30):
            g = ctx.plan_learned(sid, k)
            o = orc.plan_learned()
            assert (g is None) == (o is None), f"walker {k} step {step}: wall decision differs"
            if g is None:
                continue
            # the float step puts the two proposals ~1e-7 apart, and the
            # steep pair terms amplify that in the energy
            np.testing.assert_allclose(g[2], o[2], atol=1e-6)
            assert abs(g[0] - o[0]) <= 1e-5 * (1.0 + abs(o[0])), (step, g[0], o[0])
            assert abs(g[1] - o[1]) <= 1e-4 * (1.0 + abs(o[1])), (step, g[1], o[1])
            # the device energy of its own proposal against an extended-
            # precision recompute of that configuration
            ctx.confirm(sid, k)
            orc.confirm()
            e_ld = O.lj_energy_ld(ctx.coords(sid, k))
            assert abs(ctx.state(sid, k)["E"] - e_ld) <= 1e-9 * max(1.0, abs(e_ld)), (step, e_ld)
            worst = max(worst, abs(ctx.state(sid, k)["E"] - e_ld) / max(1.0, abs(e_ld)))
    print("max relative energy error against the recompute", worst)


def test_probe_records_match_oracle(gpu_ctx_factory):
    print("case probe vs oracle", flush=True)
    ctx = gpu_ctx_factory()
    sp = lj13()
    seeds = np.arange(6, dtype=np.uint64) + 3
    sid = ctx.add_set(sp, seeds, geom=2.5)
    ctx.init()
    ctx.run(3000)
    ctx.prop_collect(sid, 77, 4)
    rec = ctx.prop_records().reshape(len(seeds), 4, 24)
    mismatched = 0
    for k, seed in enumerate(seeds):
        orc = O.Walker(O.make_params(system=O.LJ, n=13, radius=2.5, method=O.SAD, method_param=0.05,
                                     move_param=0.1, energy_bin=0.01, has_max_e=1, max_e=0.0, order_g=4,
                                     seed=int(seed)))
        orc.run(3000)
        np.testing.assert_array_equal(ctx.coords(sid, k), orc.coords())
        want = orc.probe(77 + k, 4)
        np.testing.assert_allclose(rec[k, :, :8], want[:, :8], rtol=1e-5, atol=1e-6)
        close = np.isclose(rec[k, :, 8:], want[:, 8:], rtol=1e-3, atol=1e-7)
        mismatched += int((~close).sum())
    # a step that lands within rounding of the wall or of a bin edge may flip
    assert mismatched <= 0.02 * rec[:, :, 8:].size


def test_gradient_matches_oracle(gpu_ctx_factory):
    print("case gradient vs oracle", flush=True)
    ctx = gpu_ctx_factory()
    w = init_weights(1)
    rec = synthetic_records(2, 1000)
    ctx.prop_set_weights(w)
    ctx.prop_load_records(rec)
    loss, g = ctx.prop_gradient(0, 1000)
    lo, go = O.prop_loss_grad(w, rec)
    assert abs(loss - lo) < 1e-5 * abs(lo)
    assert np.abs(go).max() > 0
    np.testing.assert_allclose(g, go, atol=2e-5 * np.abs(go).max())
    # a cyclic window past the end wraps to the start
    loss2, g2 = ctx.prop_gradient(900, 200)
    lo2, go2 = O.prop_loss_grad(w, np.concatenate([rec[900:], rec[:100]]))
    np.testing.assert_allclose(g2, go2, atol=2e-5 * np.abs(go2).max())


def test_adam_first_step_and_training_reduces_loss(gpu_ctx_factory):
    print("case training", flush=True)
    ctx = gpu_ctx_factory()
    w = init_weights(5)
    rec = synthetic_records(6, 8192)
    ctx.prop_set_weights(w)
    ctx.prop_load_records(rec)
    _, g = ctx.prop_gradient(0, 1024)
    w0, _ = ctx.prop_get_weights()
    lr = 1e-3
    ctx.prop_train(1, 1024, lr)
    w1, b1 = ctx.prop_get_weights()
    big = np.abs(g) > 1e-4
    assert big.sum() > 100
    # Adam's first step moves every weight by lr against its gradient's sign
    np.testing.assert_allclose((w1 - w0)[big], -lr * np.sign(g[big]), rtol=1e-3)
    assert np.all(b1 == O.bf16_round(w1))
    losses = ctx.prop_train(400, 1024, 3e-3)
    assert losses[-50:].mean() < losses[:10].mean() - 0.02


def test_learned_walkers_sample_uniform_sphere(gpu_ctx_factory):
    print("case learned detailed balance", flush=True)
    R = 1.5
    ctx = gpu_ctx_factory()
    ctx.prop_set_weights(_hand_weights())
    sp = S.make_params(system=S.LJ, n=2, lanes=2, method=S.CANONICAL, method_param=1e300, energy_bin=1e9,
                       has_max_e=1, max_e=1e12, move_kind=S.MOVE_LEARNED)
    W = 4096
    start = np.tile(np.array([0.5, -0.5, 0.0, 0.0, 0.0, 0.0]), W)
    sid = ctx.add_set(sp, np.arange(W, dtype=np.uint64), geom=R, coords=start)
    ctx.init()
    ctx.run(6000)
    c = np.array([ctx.coords(sid, k) for k in range(W)]).reshape(W, 3, 2)
    d = np.linalg.norm(c[:, :, 0] - c[:, :, 1], axis=1)
    assert np.all(np.linalg.norm(c, axis=1) <= R + 1e-12)
    assert stats.kstest(d, lambda r: _pair_distance_cdf(r, R)).pvalue > 1e-3


def test_trained_network_gains_over_fixed_step(gpu_ctx_factory):
    """Train on records from plain SAD walkers at three stages of their run
    (hot, cooling, cold), then score held-out records of the same stages by
    the mean relative expected squared jump: the trained network against the
    Gaussian with the cube move's second moment (the method without it) and
    against the best single scale chosen on the training records."""
    print("case learned gain", flush=True)
    ctx = gpu_ctx_factory()
    sp = lj13()
    W = 2048
    sid = ctx.add_set(sp, np.arange(W, dtype=np.uint64), geom=2.5)
    ctx.init()
    done = 0
    train, held = [], []
    for k, stage in enumerate((20000, 80000, 200000)):
        ctx.run(stage - done)
        done = stage
        ctx.prop_load_records(np.zeros((0, 24), np.float32))
        ctx.prop_collect(sid, 1000 + k, 8)
        train.append(ctx.prop_records())
        ctx.prop_load_records(np.zeros((0, 24), np.float32))
        ctx.prop_collect(sid, 9000 + k, 2)
        held.append(ctx.prop_records())
    train, held = np.concatenate(train), np.concatenate(held)
    ctx.prop_load_records(train[np.random.default_rng(0).permutation(len(train))])
    ctx.prop_init(0)
    ctx.prop_train(1500, 2048, 3e-3)
    _, wb = ctx.prop_get_weights()
    learned = -O.prop_loss_grad(wb, held)[0]

    def constant(s, rec):
        w = np.zeros(O.PROP_NP, np.float32)
        w[1376] = s
        return -O.prop_loss_grad(w, rec)[0]

    cube_equivalent = constant(np.log(0.1 / np.sqrt(3.0)), held)  # same second moment as the 0.1 cube step
    grid = np.linspace(O.PROP_SMIN, O.PROP_SMAX, 61)
    s_best = grid[int(np.argmax([constant(s, train) for s in grid]))]
    best_constant = constant(s_best, held)
    print("held-out relative ESJD: learned %.4f, best single scale (sigma %.3f) %.4f, cube-equivalent %.4f"
          % (learned, np.exp(s_best), best_constant, cube_equivalent))
    assert learned > 1.5 * cube_equivalent
    assert learned > 1.03 * best_constant
