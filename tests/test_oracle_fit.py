"""Pins the oracle's separable fit of exact local matrices (the local solvers of
distorted patches built from a Kronecker fit, kroneckersvd.h:215-585).

What fixes the fit without the reference: the exact local matrix R_j A R_j^T
agrees with an independent dense assembly; on Cartesian patches the local
matrix is exactly separable (fast diagonalization is exact there,
tests/apps_laplace_integrator.cc:56-167), so the fit must recover it from a
wrong initial guess; alternating least squares never increases the objective;
on distorted patches the fitted solver is spectrally closer to A_j than the
surrogate-length solver it starts from.
"""
import numpy as np
import pytest

import dense_sipg as ds
import oracle as o


def test_local_matrix_matches_dense_assembly():
    dim, p, n, ip = 2, 2, 4, 1.5
    A = ds.assemble(dim, p, n, ip)
    sm = o.Smoother(dim, p, n, ip=ip)
    idx = sm.patch_dof_indices().astype(np.int64)
    for j in (0, 4, sm.P - 1):
        Aj = sm.patch_local_matrix(j)
        assert np.abs(Aj - A[np.ix_(idx[j], idx[j])]).max() <= 1e-12 * np.abs(Aj).max()


def kron_sum(M, L):
    dim = len(M)
    tot = 0.0
    for d in range(dim):
        term = np.ones((1, 1))
        for e in reversed(range(dim)):
            term = np.kron(term, L[e] if e == d else M[e])
        tot = tot + term
    return tot


@pytest.mark.parametrize("dim,p", [(2, 3), (3, 1), (3, 2)])
def test_fit_recovers_separable_matrix(dim, p):
    n = 4
    sm = o.Smoother(dim, p, n, ip=2.0)
    j = sm.P // 2
    Aj = sm.patch_local_matrix(j)
    v = sm.patch_vertices()[j]
    # wrong initial guess: lengths scaled by 1.3 / 0.8
    M0, L0 = [], []
    for d in range(dim):
        M, L = o.patch_matrices_1d(p, [1.3 / n, 0.8 / n], v[d] == 1, v[d] == n - 1, 2.0)
        M0.append(M)
        L0.append(L)
    err0 = np.linalg.norm(Aj - kron_sum(M0, L0)) / np.linalg.norm(Aj)
    M, L, err = o.separable_fit(dim, Aj, np.array(M0), np.array(L0), 30)
    assert err0 > 1e-2 and err < 1e-9, (err0, err)
    assert np.linalg.norm(Aj - kron_sum(M, L)) / np.linalg.norm(Aj) == pytest.approx(err, abs=1e-12)


def test_fit_objective_is_monotone_and_beats_surrogate():
    dim, p, n = 3, 2, 4
    V = o.distorted_vertices(dim, n, 0.15, seed=3)
    sm = o.Smoother(dim, p, n, ip=4.0, vertices=V)
    h = sm.patch_h()
    vt = sm.patch_vertices()
    for j in (0, 13):
        Aj = sm.patch_local_matrix(j)
        M0, L0 = [], []
        for d in range(dim):
            M, L = o.patch_matrices_1d(p, h[j, d], vt[j, d] == 1, vt[j, d] == n - 1, 4.0)
            M0.append(M)
            L0.append(L)
        errs = [np.linalg.norm(Aj - kron_sum(M0, L0)) / np.linalg.norm(Aj)]
        for s in range(1, 6):
            errs.append(o.separable_fit(dim, Aj, np.array(M0), np.array(L0), s)[2])
        assert all(b <= a * (1 + 1e-12) for a, b in zip(errs, errs[1:])), errs
        assert errs[-1] < 0.95 * errs[0]


def test_fitted_local_solver_spectrum():
    """Eigenvalues of B_j^{-1} A_j (B_j: local solver) over all patches of a
    distorted 3D Q1 mesh: the fitted solvers are s.p.d. and lower the largest
    eigenvalue (the one that limits the additive damping)."""
    dim, p, n = 3, 1, 4
    V = o.distorted_vertices(dim, n, 0.2, seed=5)
    sur = o.Smoother(dim, p, n, ip=4.0, vertices=V)
    fit = o.Smoother(dim, p, n, ip=4.0, vertices=V, local_solver=o.LOCAL_FIT, fit_sweeps=4)
    assert fit.fit_fallbacks() == 0
    ranges = {}
    for name, s in (("surrogate", sur), ("fit", fit)):
        lo, hi = np.inf, -np.inf
        for j in range(s.P):
            Aj = s.patch_local_matrix(j)
            Binv = np.stack([s.local_apply(j, c) for c in np.eye(s.ndp)], 1)
            w = np.linalg.eigvals(Binv @ Aj).real
            lo, hi = min(lo, w.min()), max(hi, w.max())
        ranges[name] = (lo, hi)
    assert ranges["fit"][0] > 0.0
    assert ranges["fit"][1] < ranges["surrogate"][1], ranges
    np.testing.assert_array_less(fit.fit_errors(), 0.3)
