"""Pins of the oracle's learned proposal (no reference counterpart):

- the analytic gradient of the training loss against central finite
  differences (double precision), and the float-forward gradient against it;
- Gaussian radial features against their closed form on two atoms;
- the Box-Muller draws against N(0, 1) (KS, moments);
- the two-sided Hastings term: ln H(x->x') + ln H(x'->x) = 0 exactly;
- detailed balance: at infinite temperature the learned move must leave
  two atoms uniform in the sphere, whose pair-distance law is known in
  closed form; the same run with the Hastings term dropped (a mistake
  planted in the oracle) must fail that test;
- probe records: the RNG draw order and non-negative finite targets.
"""
import numpy as np
import pytest
from scipy import stats

import oracle as O


def init_weights(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    w = np.zeros(O.PROP_NP, np.float32)
    w[:256] = rng.normal(0, scale * np.sqrt(2 / 8) / 4, 256)
    w[288:1312] = rng.normal(0, scale * np.sqrt(2 / 32), 1024)
    w[256:288] = 0.1
    w[1312:1344] = 0.1
    w[1344:1376] = rng.normal(0, 0.1, 32)
    w[1376] = np.log(0.1)
    return O.bf16_round(w)


def synthetic_records(seed, n):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, O.PROP_REC), np.float32)
    r[:, :8] = rng.uniform(0, 4, (n, 8))
    grid = np.linspace(O.PROP_SMIN, O.PROP_SMAX, O.PROP_NS)
    peak = rng.uniform(-3.5, -1.0, n)
    r[:, 8:] = np.exp(2 * grid)[None, :] * np.exp(-np.exp(2 * (grid[None, :] - peak[:, None])))
    return r


def test_loss_gradient_matches_finite_differences():
    w = init_weights(1).astype(np.float64)
    rec = synthetic_records(2, 64)
    loss, g = O.prop_loss_grad(w, rec, f64=True)
    assert np.abs(g).max() > 0
    rng = np.random.default_rng(3)
    idx = np.concatenate([[1376], rng.choice(np.arange(1344, 1376), 8, replace=False),
                          rng.choice(np.arange(288, 1312), 16, replace=False),
                          rng.choice(np.arange(0, 256), 16, replace=False), [260, 1320]])
    h = 1e-6
    bad = 0
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (O.prop_loss_grad(wp, rec, f64=True)[0] - O.prop_loss_grad(wm, rec, f64=True)[0]) / (2 * h)
        if abs(fd - g[i]) > 1e-6 * max(1.0, abs(g).max()) + 1e-4 * abs(g[i]):
            bad += 1
    # a kink of ReLU or of the interpolated target inside +-h can spoil one
    assert bad <= 1


def test_float_gradient_matches_double():
    w = init_weights(4)
    rec = synthetic_records(5, 256)
    l32, g32 = O.prop_loss_grad(w, rec)
    l64, g64 = O.prop_loss_grad(w.astype(np.float64), rec, f64=True)
    assert abs(l32 - l64) < 1e-5 * abs(l64)
    np.testing.assert_allclose(g32, g64, atol=1e-4 * np.abs(g64).max())


def test_features_closed_form_two_atoms():
    c0 = 0.9 + 3 * 1.6 / 7
    coords = np.array([0.0, c0, 0.0, 0.0, 0.0, 0.0])  # SoA: x = (0, c0), y = z = 0
    f = O.prop_features(coords, 0, (0.0, 0.0, 0.0))
    assert abs(f[3] - 1.0) < 1e-6
    d = 1.6 / 7
    for k in range(8):
        assert abs(f[k] - np.exp(-12.5 * ((k - 3) * d) ** 2)) < 2e-6
    # the skipped atom does not count, and features add over atoms
    # skipping atom 1 leaves atom 0 at distance 0
    f0 = O.prop_features(coords, 1, (0.0, 0.0, 0.0))
    np.testing.assert_allclose(f0, np.exp(-12.5 * (0.9 + np.arange(8) * d) ** 2), rtol=1e-5)


def test_gaussian_draws_are_standard_normal():
    rng = O.RNG(11)
    xi = np.array([O.prop_gaussian(rng) for _ in range(20000)])
    for k in range(3):
        assert stats.kstest(xi[:, k], "norm").pvalue > 1e-3
    assert abs(xi.mean()) < 0.02 and abs(xi.var() - 1) < 0.03
    assert abs(np.corrcoef(xi[:, 0], xi[:, 1])[0, 1]) < 0.03


def test_hastings_term_is_antisymmetric():
    rng = np.random.default_rng(6)
    for _ in range(200):
        s, s2 = rng.uniform(O.PROP_SMIN, 0, 2).astype(np.float32)
        xi2 = np.float32(rng.chisquare(3))
        back = np.float32(xi2 * np.exp(2 * (float(s) - float(s2))))  # |xi_rev|^2 = |d|^2 / sigma'^2
        a = O.prop_lnh(s, s2, xi2)
        assert abs(a + O.prop_lnh(s2, s, back)) < 1e-6 * (1 + abs(a))


def _hand_weights():
    """sigma = 0.5 exp(-2 (f0 + f1 + f2)): small steps near a neighbour."""
    w = np.zeros(O.PROP_NP, np.float32)
    w[0:3] = 1.0          # W1[0][0..2]
    w[288] = 1.0          # W2[0][0]
    w[1344] = -2.0        # w3[0]
    w[1376] = np.log(0.5)
    return w


def _pair_distance_cdf(r, R):
    x = r / R
    return x ** 3 - 9 / 16 * x ** 4 + x ** 6 / 32


@pytest.mark.parametrize("drop", [0, 1])
def test_learned_move_detailed_balance_uniform_sphere(drop):
    R = 1.5
    p = O.make_params(system=O.LJ, n=2, radius=R, method=O.CANONICAL, method_param=1e300, energy_bin=1e9,
                      has_max_e=1, max_e=1e12, seed=21, mlp=_hand_weights(), drop_hastings=drop)
    w = O.Walker(p, coords=np.array([0.5, -0.5, 0.0, 0.0, 0.0, 0.0]))
    d = []
    # samples 400 moves apart: the small steps near a neighbour decorrelate slowly
    for _ in range(8000):
        w.run(400)
        c = w.coords().reshape(3, 2)
        d.append(np.linalg.norm(c[:, 0] - c[:, 1]))
    pval = stats.kstest(np.array(d), lambda r: _pair_distance_cdf(r, R)).pvalue
    if drop:
        assert pval < 1e-6
    else:
        assert pval > 1e-3


def test_probe_records_draw_order_and_targets():
    w = init_weights(7)
    p = O.make_params(system=O.LJ, n=13, radius=2.5, method=O.SAD, method_param=0.05, energy_bin=0.01,
                      has_max_e=1, max_e=0.0, order_g=4, seed=3, mlp=w)
    wk = O.Walker(p)
    wk.run(5000)
    rec = wk.probe(99, 8)
    rng = O.RNG(99)
    coords = wk.coords()
    for r in rec:
        atom = rng.uniform_index(13)
        pos = coords.reshape(3, 13)[:, atom]
        np.testing.assert_array_equal(r[:8], O.prop_features(coords, atom, pos))
        for _ in range(O.PROP_NS * O.PROP_R):
            O.prop_gaussian(rng)
    J = rec[:, 8:]
    assert np.all(np.isfinite(J)) and np.all(J >= 0)
    # the smallest scale is almost always accepted: J_0 ~ sigma_0^2 E|xi|^2 = 3e-4
    assert 0.5e-4 < np.median(J[:, 0]) < 1e-3


def test_load_state_resynchronises_a_walker():
    """The oracle's state loader (used to resynchronise on a device walker
    after a float-level parting): a walker loaded with another's state
    follows it exactly afterwards."""
    w = init_weights(3)

    def mk(seed):
        return O.Walker(O.make_params(system=O.LJ, n=38, radius=3.2, method=O.SAD, method_param=0.03,
                                      move_param=0.1, energy_bin=0.01, has_max_e=1, max_e=0.0, order_g=8,
                                      seed=seed, mlp=w))

    a, b = mk(1), mk(2)
    a.run(3000)
    b.run(1234)
    b.load_state(a.scalars(), a.bins(), a.coords())
    a.run(2000)
    b.run(2000)
    assert a.scalars() == b.scalars()
    for k, v in a.bins().items():
        np.testing.assert_array_equal(v, b.bins()[k], err_msg=k)
    np.testing.assert_array_equal(a.coords(), b.coords())
