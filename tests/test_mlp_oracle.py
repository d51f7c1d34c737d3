"""Pins the learned-damping oracle on the CPU: finite differences for the
back-propagation, closed forms for features and damping targets."""
import numpy as np

import mlp_oracle as mo
import oracle as o


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n_in, h, B = 8, 16, 11
    _, n = mo.param_layout(n_in, h)
    p = 0.3 * rng.standard_normal(n)
    X = rng.uniform(0.5, 1.5, (B, n_in))
    t = rng.uniform(0.2, 1.2, B)
    g = mo.grad(p, X, t, n_in, h)
    for i in rng.choice(n, 40, replace=False):
        e = np.zeros(n)
        e[i] = 1e-6
        fd = (mo.loss(p + e, X, t, n_in, h) - mo.loss(p - e, X, t, n_in, h)) / 2e-6
        assert abs(fd - g[i]) < 1e-8 + 1e-6 * abs(fd), (i, fd, g[i])


def test_output_range_and_bf16_rounding():
    rng = np.random.default_rng(1)
    _, n = mo.param_layout(24, 64)
    p = rng.standard_normal(n)
    w, _ = mo.forward(p, rng.uniform(0, 2, (50, 24)), 24, 64)
    assert np.all((w >= 0.2) & (w <= 1.2))
    # spacing at 1 is 2^-7: 1 + 2^-8 is a tie (to even: 1), 1 + 3*2^-8 rounds up
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -3.14159, 1e-3])
    r = mo.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6
    assert abs(r[3] + 3.140625) < 1e-12


def test_features_cartesian_are_ones():
    sm = o.Smoother(3, 1, 4)
    f = mo.patch_features(3, 4, None, sm.patch_vertices())
    np.testing.assert_allclose(f, 1.0, rtol=1e-14)
    # stretched mesh: singular values are the stretch factors (sorted)
    V = o.uniform_vertices(3, 4) * np.array([1.0, 2.0, 0.5])
    f = mo.patch_features(3, 4, V, sm.patch_vertices())
    np.testing.assert_allclose(f.reshape(-1, 8, 3), np.broadcast_to([2.0, 1.0, 0.5], (27, 8, 3)),
                               rtol=1e-13)


def test_optimal_damping_is_one_for_exact_local_solvers():
    """On Cartesian meshes the FD local solver is the exact patch inverse,
    so the A-norm line search along every local correction gives 1."""
    sm = o.Smoother(2, 2, 5, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK)
    w = mo.optimal_damping(sm, 2, 5, lo=0.0, hi=10.0)
    np.testing.assert_allclose(w, 1.0, rtol=1e-11)
    V = o.distorted_vertices(2, 5, 0.25, seed=3)
    smd = o.Smoother(2, 2, 5, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK, vertices=V)
    wd = mo.optimal_damping(smd, 2, 5, lo=0.0, hi=10.0)
    assert np.abs(wd - 1.0).max() > 1e-3  # surrogate local solvers are inexact
