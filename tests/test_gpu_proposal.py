"""GPU: the learned proposal against the oracle.

- single learned moves (features, network, Gaussian step, Hastings term,
  energy) on the same walkers as the oracle, within float tolerance (the
  device's exp/log/sin differ from the host's in the last bits, so learned
  trajectories are compared move by move, not bit for bit), for LJ13 and
  for the learned-proposal workload (LJ38, sphere 3.2, min_T 0.03) at 8 and
  16 lanes per walker;
- the persistent learned walker over 2000 SAD moves against the oracle's,
  move by move (its waves full, so the network runs on the matrix cores;
  single moves take the vector-unit form), and the workload's full
  16384-walker launch;
- probe records from the same walkers and streams;
- the training gradient and loss on the same records and weights;
- Adam's first step and a falling training loss;
- detailed balance on the device, with either form of the network: at
  infinite temperature two atoms must end uniform in the sphere
  (closed-form pair-distance law);
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


# (atoms, sphere radius, SAD min_T, lanes per walker): LJ13 at 4 lanes, and
# the learned-proposal workload (LJ38 in a 3.2 sigma sphere, min_T 0.03) at 8
# lanes (5 atoms per lane: the KMAX = 8 kernels) and at 16 lanes (KMAX = 4).
CONFIGS = {"lj13_g4": (13, 2.5, 0.05, 4), "lj38_g8": (38, 3.2, 0.03, 8), "lj38_g16": (38, 3.2, 0.03, 16)}


def learned_params(case):
    n, R, min_t, lanes = CONFIGS[case]
    return S.make_params(system=S.LJ, n=n, lanes=lanes, method=S.SAD, method_param=min_t, move_param=0.1,
                         energy_bin=0.01, has_max_e=1, max_e=0.0, move_kind=S.MOVE_LEARNED), R


def oracle_learned(case, seed, w):
    n, R, min_t, lanes = CONFIGS[case]
    return O.Walker(O.make_params(system=O.LJ, n=n, radius=R, method=O.SAD, method_param=min_t, move_param=0.1,
                                  energy_bin=0.01, has_max_e=1, max_e=0.0, order_g=lanes, seed=int(seed), mlp=w))


@pytest.mark.parametrize("case", list(CONFIGS))
def test_learned_moves_match_oracle(gpu_ctx_factory, case):
    print(f"case learned plan vs oracle {case}", flush=True)
    w = init_weights(3)
    ctx = gpu_ctx_factory()
    ctx.prop_set_weights(w)
    sp, R = learned_params(case)
    seeds = np.arange(4, dtype=np.uint64) + 50
    sid = ctx.add_set(sp, seeds, geom=R)
    ctx.init()
    worst = 0.0
    for k, seed in enumerate(seeds):
        orc = oracle_learned(case, seed, w)
        np.testing.assert_array_equal(ctx.coords(sid, k), orc.coords())
        assert ctx.state(sid, k)["E"] == orc.scalars()["E"]
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


# Decisions a float-tolerance trajectory may take differently from the
# oracle: the device's exp/log/sin/cos differ from the host's in the last
# bits, which moves a learned proposal by ~1e-8 sigma; the pair forces of a
# strained configuration (up to ~1e4 epsilon/sigma) carry that into the
# energy, and ln w's extension below too_lo (slope 1/min_T) and ln H carry it
# into the acceptance exponent.  A decision whose input lies within that gap
# of its threshold can go either way; `gap` is the largest energy difference
# between the two walkers seen so far.
def _explained(lm, st_prev, width, gap, min_t):
    e2, lr, u, wall = lm["e2"], lm["lr"], lm["u"], lm["wall"]
    tol_e = 10.0 * gap + 1e-6
    why = []
    if np.isfinite(wall) and abs(wall) < 1e-4:
        why.append(f"wall |to|^2-R^2={wall:.2e}")
    if np.isfinite(lr) and np.isfinite(u) and abs(lr - np.log(u)) < 10.0 * tol_e / min_t + 1e-4:
        why.append(f"acceptance lr-ln(u)={lr - np.log(u):.2e}")
    if np.isfinite(e2):
        x = (e2 - st_prev["bins_min"]) / width
        if abs(x - np.round(x)) * width < tol_e:
            why.append(f"bin edge {abs(x - np.round(x)) * width:.2e}")
        for name in ("too_lo", "too_hi"):
            if abs(e2 - st_prev[name]) < tol_e:
                why.append(f"{name} {e2 - st_prev[name]:.2e}")
        if abs(e2) < tol_e:
            why.append("max_allowed_energy")
    return why


def _oracle_view(g):
    """A device walker-state record as the oracle's scalars."""
    from conftest import SCALAR_KEYS
    d = {ok: g[gk] for gk, ok in SCALAR_KEYS if gk in g.dtype.names}
    d["rng"] = [int(x) for x in g["rng"]]
    d["min_T"] = g["min_T"]
    return d


@pytest.mark.parametrize("case", ["lj38_g8", "lj38_g16"])
def test_learned_trajectory_matches_oracle(gpu_ctx_factory, case):
    """The persistent learned walker (k_run_learned: learned move, ln H in
    the acceptance, SAD bookkeeping) of the learned-proposal workload against
    the oracle's learned walker, move by move for 2000 moves of 8 walkers.
    After every move: every counter, the generator, every SAD index and the
    current bin identical; every 50 moves the positions within 1e-5 and each
    side's energy within 1e-9 of an extended-precision recompute of its own
    configuration.  When a walker's discrete state parts from the oracle's,
    the decision must have been taken within the walkers' float-level gap of
    its threshold; the oracle is then reloaded with the device walker's
    state and the comparison goes on.  At the end every bin's histogram,
    t_found and round-trip count identical, ln w within 1e-6 relative, and
    each bin's energy sums within its count times the largest gap.
    The one-move launches are then checked against a single 2000-move
    launch of the same walkers, bit for bit."""
    print(f"case learned trajectory {case}", flush=True)
    w = init_weights(3)
    sp, R = learned_params(case)
    min_t = CONFIGS[case][2]
    W, M = 8, 2000
    seeds = np.arange(W, dtype=np.uint64) + 200
    ctx = gpu_ctx_factory()
    ctx.prop_set_weights(w)
    sid = ctx.add_set(sp, seeds, geom=R)
    ctx.init()
    orcs = [oracle_learned(case, s, w) for s in seeds]
    reports = []
    gap = [0.0] * W
    worst_gap, worst_x = 0.0, 0.0
    prev, _ = ctx.download_raw(sid, with_bins=False)
    for m in range(1, M + 1):
        ctx.run(1)
        st, _ = ctx.download_raw(sid, with_bins=False)
        for k in range(W):
            o = orcs[k]
            o.run(1)
            s = o.scalars()
            g = st[k]
            same = all(int(g[f]) == int(s[o_f]) for f, o_f in (
                ("moves", "moves"), ("accepted_moves", "accepted"), ("len", "len"), ("num_states", "num_states"),
                ("tL", "tL"), ("tF", "tF"), ("highest_hist", "highest_hist"), ("max_S_index", "max_S_index"),
                ("status", "status")))
            same = same and list(map(int, g["rng"])) == list(map(int, s["rng"]))
            same = same and all(float(g[f]) == float(s[f]) for f in ("bins_min", "too_lo", "too_hi"))
            same = same and int((g["E"] - g["bins_min"]) / 0.01) == int((s["E"] - s["bins_min"]) / 0.01)
            if not same:
                why = _explained(o.last_move(), prev[k], 0.01, gap[k], min_t)
                reports.append((k, m, why))
                print(f"walker {k} parts from the oracle at move {m} (gap {gap[k]:.1e}): {why}", flush=True)
                if not why:
                    diff = {f: (g[f], s[o_f]) for f, o_f in (
                        ("moves", "moves"), ("accepted_moves", "accepted"), ("len", "len"), ("num_states", "num_states"),
                        ("tL", "tL"), ("tF", "tF"), ("highest_hist", "highest_hist"),
                        ("max_S_index", "max_S_index"), ("status", "status"), ("bins_min", "bins_min"),
                        ("too_lo", "too_lo"), ("too_hi", "too_hi"), ("E", "E")) if g[f] != s[o_f]}
                    print("differing:", diff, "rng", list(map(int, g["rng"])), s["rng"], "last move",
                          o.last_move(), "previous device state E", prev[k]["E"], flush=True)
                assert why, f"walker {k} move {m}: discrete state differs with no decision near its threshold"
                o.load_state(_oracle_view(g), ctx.bins(sid, k), ctx.coords(sid, k))
                gap[k] = 0.0
                continue
            gap[k] = max(gap[k], abs(g["E"] - s["E"]))
            worst_gap = max(worst_gap, gap[k])
            assert gap[k] <= 1e-2 * max(1.0, abs(s["E"])), (k, m, g["E"], s["E"])
            if m % 50 == 0:
                c, oc = ctx.coords(sid, k), o.coords()
                worst_x = max(worst_x, float(np.abs(c - oc).max()))
                np.testing.assert_allclose(c, oc, atol=1e-5, err_msg=f"walker {k} move {m}")
                for e, cc in ((g["E"], c), (s["E"], oc)):
                    e_ld = O.lj_energy_ld(cc)
                    assert abs(e - e_ld) <= 1e-9 * max(1.0, abs(e_ld)), (k, m, e, e_ld)
        prev = st
    # partings are rare: a decision lands within ~1e-4 of its threshold
    assert len(reports) <= 0.01 * W * M, reports
    for k in range(W):
        gb, ob = ctx.bins(sid, k), orcs[k].bins()
        for name in ("histogram", "t_found", "round_trips", "visited"):
            np.testing.assert_array_equal(gb[name], ob[name], err_msg=f"walker {k} {name}")
        # lambda = (E - too_lo) / min_T when the range grows upward carries
        # the energy's float-level difference into gamma
        np.testing.assert_allclose(gb["lnw"], ob["lnw"], rtol=1e-6, atol=1e-9, err_msg=f"walker {k} lnw")
        # each of a bin's hist energies is within the float gap of the oracle's
        h = gb["histogram"].astype(np.float64)
        g = worst_gap + 1e-12
        assert np.all(np.abs(gb["energy_total"] - ob["energy_total"]) <= h * g), f"walker {k} energy_total"
        emax = np.abs(gb["energy_total"]) / np.maximum(h, 1.0) + np.abs(gb["energy_squared_total"] / np.maximum(
            h, 1.0)) ** 0.5
        assert np.all(np.abs(gb["energy_squared_total"] - ob["energy_squared_total"]) <= h * g * (2 * emax + g)), \
            f"walker {k} energy_squared_total"
    print(f"{W} walkers x {M} moves compared; {len(reports)} partings at decisions within the float gap "
          f"(largest energy gap {worst_gap:.1e}, largest position difference {worst_x:.1e}): {reports}")
    # the same walkers in one persistent launch
    one = gpu_ctx_factory()
    one.prop_set_weights(w)
    sid1 = one.add_set(sp, seeds, geom=R)
    one.init()
    one.run(M)
    a, ab = ctx.download_raw(sid)
    b, bb = one.download_raw(sid1)
    assert a.tobytes() == b.tobytes() and ab.tobytes() == bb.tobytes()
    for k in range(W):
        np.testing.assert_array_equal(ctx.coords(sid, k), one.coords(sid1, k))


def test_learned_workload_full_set(gpu_ctx_factory):
    """The learned-proposal workload at its size: 16384 LJ38 walkers (sphere
    3.2, min_T 0.03) at 8 lanes in one k_run_learned launch of 2000 moves;
    every walker runs, and a sample's energies agree with an extended-
    precision recompute of its configuration to 1e-9."""
    print("case learned workload 16384 walkers", flush=True)
    w = init_weights(3)
    sp, R = learned_params("lj38_g8")
    W = 16384
    ctx = gpu_ctx_factory()
    ctx.prop_set_weights(w)
    sid = ctx.add_set(sp, np.arange(W, dtype=np.uint64), geom=R)
    ctx.init()
    ctx.run(2000)
    st, _ = ctx.download_raw(sid, with_bins=False)
    assert np.all(st["status"] == 0) and np.all(st["moves"] == 2000)
    assert st["accepted_moves"].mean() > 100
    worst = 0.0
    for k in range(0, W, 257):
        e_ld = O.lj_energy_ld(ctx.coords(sid, k))
        worst = max(worst, abs(st[k]["E"] - e_ld) / max(1.0, abs(e_ld)))
    assert worst <= 1e-9, worst
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


@pytest.mark.parametrize("lanes", [2, 4])
def test_learned_walkers_sample_uniform_sphere(gpu_ctx_factory, lanes):
    """At 2 lanes the network runs on the vector units, at 4 (every wave
    full: 4096 walkers) on the matrix cores."""
    print(f"case learned detailed balance, {lanes} lanes", flush=True)
    R = 1.5
    ctx = gpu_ctx_factory()
    ctx.prop_set_weights(_hand_weights())
    sp = S.make_params(system=S.LJ, n=2, lanes=lanes, method=S.CANONICAL, method_param=1e300, energy_bin=1e9,
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
