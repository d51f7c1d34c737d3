"""Learned per-patch damping on the device vs the FP64 numpy oracle (needs an MI355X).

Features and damping targets are FP32 outputs of FP64 computations (rel. 1e-6);
the network computes in FP32 with BF16 weights: forward and gradient are compared
with the oracle evaluated on the same BF16-rounded weights (rel. 1e-4)."""
import numpy as np
import pytest

import mlp_oracle as mo
import oracle as o

pytestmark = pytest.mark.gpu


def _tpss():
    import tpss_amd
    if tpss_amd.device_count() < 1:
        pytest.skip("no HIP device")
    return tpss_amd


@pytest.mark.parametrize("dim,n", [(3, 4), (2, 6)])
def test_patch_features(dim, n):
    t = _tpss()
    V = o.distorted_vertices(dim, n, 0.2, seed=3)
    lv = t.Level(dim, 2, n, vertices=V)
    sm = t.SchwarzSmoother(lv, variant=t.MULTIPLICATIVE)
    f = sm.patch_features()
    ref = mo.patch_features(dim, n, V, sm.patch_vertices())
    np.testing.assert_allclose(f, ref, rtol=2e-6)


@pytest.mark.parametrize("distorted", [False, True])
def test_optimal_damping(distorted):
    t = _tpss()
    dim, p, n = 3, 2, 4
    V = o.distorted_vertices(dim, n, 0.2, seed=3) if distorted else None
    lv = t.Level(dim, p, n, ip_factor=4.0, vertices=V)
    sm = t.SchwarzSmoother(lv, variant=t.MULTIPLICATIVE)
    w = sm.optimal_damping(n_samples=2, seed=5, omega_min=0.0, omega_max=10.0)
    ref_sm = o.Smoother(dim, p, n, ip=4.0, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK,
                        vertices=V)
    ref = mo.optimal_damping(ref_sm, 2, 5, lo=0.0, hi=10.0)
    np.testing.assert_allclose(w, ref, rtol=2e-6)
    if not distorted:
        np.testing.assert_allclose(w, 1.0, rtol=1e-6)


def test_optimal_damping_elasticity():
    """Block-diagonal elasticity local solvers (3D Q2, mu 1, lambda 10): the line-search
    damping over all component blocks of a patch, against the oracle's brute force."""
    t = _tpss()
    dim, p, n = 3, 2, 4
    lv = t.Level(dim, p, n, ip_factor=10.0, elasticity=(1.0, 10.0))
    sm = t.SchwarzSmoother(lv, variant=t.MULTIPLICATIVE)
    w = sm.optimal_damping(n_samples=2, seed=5, omega_min=0.0, omega_max=10.0)
    ref_sm = o.Smoother(dim, p, n, ip=10.0, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK,
                        elasticity=(1.0, 10.0))
    ref = mo.optimal_damping(ref_sm, 2, 5, lo=0.0, hi=10.0)
    np.testing.assert_allclose(w, ref, rtol=2e-6)
    # the block-diagonal solver drops the lambda coupling of the components: not exact
    assert np.abs(w - 1.0).max() > 1e-3
    f = sm.patch_features()
    assert np.abs(f - f[0]).max() <= 1e-6 * np.abs(f).max()  # Cartesian: one feature vector


def test_mlp_forward_and_gradient():
    t = _tpss()
    mlp = t.DampingMLP(24, 64, 0.2, 1.2, seed=4)
    params = mlp.get_params().astype(np.float64)
    pb = mo.bf16_round(params)
    rng = np.random.default_rng(2)
    X = rng.uniform(0.6, 1.4, (300, 24)).astype(np.float32)
    target = rng.uniform(0.3, 1.1, 300).astype(np.float32)
    w = mlp.forward(X)
    ref, _ = mo.forward(pb, X.astype(np.float64), 24, 64)
    np.testing.assert_allclose(w, ref, rtol=1e-4, atol=1e-5)
    loss, g = mlp.loss_grad(X, target)
    gref = mo.grad(pb, X.astype(np.float64), target.astype(np.float64), 24, 64)
    assert abs(loss - mo.loss(pb, X.astype(np.float64), target.astype(np.float64), 24, 64)) < 1e-4 * loss
    np.testing.assert_allclose(g, gref, rtol=1e-3, atol=1e-4 * np.abs(gref).max())


def test_training_reduces_loss_and_sets_damping():
    t = _tpss()
    dim, p, n = 3, 3, 8
    V = o.distorted_vertices(dim, n, 0.2, seed=11)
    lv = t.Level(dim, p, n, ip_factor=4.0, vertices=V)
    sm = t.SchwarzSmoother(lv, variant=t.MULTIPLICATIVE)
    X = sm.patch_features()
    y = sm.optimal_damping(n_samples=2, seed=5)
    mlp = t.DampingMLP(24, 64, 0.2, 1.2, seed=4)
    l0, _ = mlp.loss_grad(X, y)
    l1 = mlp.train(X, y, epochs=200, batch=128, lr=3e-3)
    assert l1 < 0.5 * l0, (l0, l1)
    sm.set_damping_mlp(mlp)
    src = o.splitmix_uniform(3, lv.n_dofs)
    d = t.DeviceVector.zeros(lv.n_dofs)
    sm.precond_vmult(d, t.DeviceVector.from_numpy(src))
    ref = o.Smoother(dim, p, n, ip=4.0, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK,
                     vertices=V)
    omega = mlp.forward(X).astype(np.float64)
    ref.set_patch_damping(omega)
    out = d.numpy()
    expect = ref.precond_vmult(src)
    assert np.abs(out - expect).max() < 1e-12 * np.abs(expect).max()


@pytest.mark.parametrize("pattern", ["half", "uniform", "graded"])
def test_shipped_model_beats_the_best_constant_damping(pattern):
    """The trained network (tpss_amd/models/damping_dgq3_3d.npy) as smoother damping,
    omega_j = 1 - g (1 - w_j) with the gain g calibrated on other meshes, against every
    constant damping 1.0, 0.9, ..., 0.6 swept on the evaluation mesh itself: 3D DG Q3 on
    64^3 cells, interior vertices moved by 0.2 h (seed 3, neither a training nor a
    calibration mesh) everywhere, growing with x, or only where x > 1/2 (the other
    half Cartesian); mixed-precision GMG-CG to 1e-8.  The network's V-cycle contracts
    faster than the best constant's on every pattern and, on the half-distorted mesh,
    needs one CG iteration less."""
    t = _tpss()
    from tpss_amd import learned_damping as ld
    n = 64
    V = ld.distorted_vertices(3, n, 0.2, seed=3, pattern=pattern)
    consts = {c: ld.gmg_solve(V, n, relaxation=c)[:2]
              for c in (1.0, 0.9, 0.8, 0.75, 0.7, 0.65, 0.6)}
    best_it = min(it for it, _ in consts.values() if it > 0)
    best_rate = min(rate for _, rate in consts.values())
    it, rate = ld.gmg_solve(V, n, mlp=ld.load())[:2]
    assert 0 < it <= best_it and rate < 0.95 * best_rate, (it, rate, consts)
    if pattern == "half":
        assert it < best_it, (it, rate, consts)
    # the reference's fixed damping 1 (TPSS.h:120-163) is far behind
    assert consts[1.0][0] < 0 or consts[1.0][0] >= it + 3, consts[1.0]


def test_shipped_elasticity_model_beats_the_reference_damping():
    """The elasticity network (tpss_amd/models/damping_elasticity_dgq2_3d.npy, trained on
    the 32^3 and 48^3 levels' patches) as damping of the block-diagonal multiplicative
    smoother: 3D SIPG DG Q2 elasticity (mu 1, lambda 10) on 64^3 cells, mixed-precision
    GMG-CG to 1e-8 needs fewer iterations than with the reference's omega = 1, and its
    V-cycle contracts at least as fast as the best constant damping of a sweep."""
    t = _tpss()
    from tpss_amd import learned_damping as ld
    n = 64
    mlp = ld.load_elasticity()
    w = float(mlp.forward(np.ones((1, 24), dtype=np.float32))[0])
    assert 0.9 < w < 1.0, w
    it, rate = ld.gmg_solve_elasticity(n, mlp=mlp)[:2]
    consts = {c: ld.gmg_solve_elasticity(n, relaxation=c)[:2] for c in (1.0, 0.95, 0.9, 0.85)}
    assert 0 < it < consts[1.0][0], (it, rate, consts)
    assert rate <= 1.01 * min(r for _, r in consts.values()), (it, rate, consts)


def test_local_lambda_max_matches_dense_eigenproblem():
    """largest eigenvalue of M_j^{-1} A_j per patch (surrogate fast-diagonalization
    solver M_j against the exact local matrix A_j of a moved-vertex patch) against
    scipy's generalized symmetric eigensolver on the oracle's dense matrices;
    exactly 1 on a Cartesian mesh, where the local solver is exact"""
    import scipy.linalg

    t = _tpss()
    dim, p, n = 3, 2, 4
    V = o.distorted_vertices(dim, n, 0.2, seed=3)
    lv = t.Level(dim, p, n, ip_factor=4.0, vertices=V)
    sm = t.SchwarzSmoother(lv, variant=t.MULTIPLICATIVE)
    lam = sm.local_lambda_max(n_iter=300)
    ref = o.Smoother(dim, p, n, ip=4.0, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK,
                     vertices=V)
    ndp = ref.ndp
    for j in range(ref.P):
        A = ref.patch_local_matrix(j)
        Minv = np.stack([ref.local_apply(j, e) for e in np.eye(ndp)], axis=1)
        M = np.linalg.inv(0.5 * (Minv + Minv.T))
        ev = np.sort(scipy.linalg.eigh(A, 0.5 * (M + M.T), eigvals_only=True))
        # the Rayleigh quotient approaches the largest eigenvalue from below, at a
        # rate set by the gap to the next one
        assert ev[-1] * (1 - 2e-3) < lam[j] < ev[-1] * (1 + 1e-9), (j, lam[j], ev[-2:])
        if ev[-1] - ev[-2] > 0.05 * ev[-1]:
            assert abs(lam[j] - ev[-1]) < 1e-8 * ev[-1], (j, lam[j], ev[-2:])
    lvc = t.Level(dim, p, n, ip_factor=4.0)
    smc = t.SchwarzSmoother(lvc, variant=t.MULTIPLICATIVE)
    np.testing.assert_allclose(smc.local_lambda_max(n_iter=20), 1.0, rtol=1e-10)
