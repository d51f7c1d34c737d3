"""Pins the oracle's linear elasticity path (operator, block-diagonal local
solvers, smoother, GMG) without a GPU.

The reference elasticity operator is assembled by deal.II's MeshWorker
integrators (apps/linelasticity_integrator.h:31-124), which are not in the
reference tree; it is pinned here by an independently written dense assembly
(tests/dense_elasticity.py), by what the mathematics fixes (rigid motions are
in the kernel away from the boundary; for a continuous quadratic field the
operator reproduces -div sigma(u) tested with the basis) and by the reference's
own iteration bounds (tests/apps_linelasticity.cc:50-105).
"""
import itertools

import numpy as np
import pytest

import dense_elasticity as de
import dense_sipg as ds
import oracle as o

MU, LAM, IP = 1.3, 7.0, 2.5


def nodal_coordinates(dim, p, n):
    nodes = ds.gll_nodes(p)
    nb, h = p + 1, 1.0 / n
    ndc = nb ** dim
    X = np.zeros((n ** dim * ndc, dim))
    inner = np.zeros(n ** dim * ndc, dtype=bool)
    for cell in range(n ** dim):
        cc = [(cell // n ** e) % n for e in range(dim)]
        for i in range(ndc):
            li = [(i // nb ** e) % nb for e in range(dim)]
            X[cell * ndc + i] = [(cc[e] + nodes[li[e]]) * h for e in range(dim)]
            inner[cell * ndc + i] = all(0 < cc[e] < n - 1 for e in range(dim))
    return X, inner


@pytest.mark.parametrize("dim,p,n", [(2, 1, 2), (2, 2, 3), (2, 3, 2), (3, 1, 2), (3, 2, 2)])
def test_operator_matches_dense_assembly(dim, p, n):
    A = de.assemble(dim, p, n, IP, MU, LAM)
    sm = o.Smoother(dim, p, n, ip=IP, elasticity=(MU, LAM))
    assert sm.N == A.shape[0]
    x = o.splitmix_uniform(1, sm.N)
    y = sm.operator(x)
    assert np.abs(y - A @ x).max() <= 1e-13 * np.abs(A @ x).max()
    np.testing.assert_allclose(A, A.T, atol=1e-12 * np.abs(A).max())
    assert np.linalg.eigvalsh(A).min() > 0.0


@pytest.mark.parametrize("dim,p,n", [(2, 2, 4), (3, 2, 3)])
def test_operator_rigid_motions_and_consistency(dim, p, n):
    """Rigid motions (eps(u) = 0, div u = 0) give A u = 0 on cells away from the
    boundary; for u = (x^2, x y, [x z]) the rows of those cells equal the mass
    matrix applied to -div sigma(u) (exact for continuous polynomials of degree p)."""
    A, M = de.assemble(dim, p, n, IP, MU, LAM, with_mass=True)
    X, inner = nodal_coordinates(dim, p, n)
    rows = np.concatenate([inner] * dim)
    Ns = X.shape[0]
    x, y = X[:, 0], X[:, 1]
    rigid = [np.concatenate([np.ones(Ns)] + [np.zeros(Ns)] * (dim - 1)),
             np.concatenate([-y, x] + [np.zeros(Ns)] * (dim - 2))]
    if dim == 3:
        z = X[:, 2]
        rigid.append(np.concatenate([np.zeros(Ns), -z, y]))
    for u in rigid:
        assert np.abs((A @ u)[rows]).max() < 1e-11 * np.abs(A).max()
    if dim == 2:
        u = np.concatenate([x ** 2, x * y])
        f = np.concatenate([-(5 * MU + 3 * LAM) * np.ones(Ns), np.zeros(Ns)])
    else:
        z = X[:, 2]
        u = np.concatenate([x ** 2, x * y, x * z])
        # grad u: [[2x,0,0],[y,x,0],[z,0,x]], div u = 4x
        # -div sigma = -(mu lap u + (mu + lam) grad div u) = -((2 mu) + 4 (mu + lam), 0, 0)
        f = np.concatenate([-(2 * MU + 4 * (MU + LAM)) * np.ones(Ns), np.zeros(Ns), np.zeros(Ns)])
    assert np.abs((A @ u - M @ f)[rows]).max() < 1e-10 * np.abs(A @ u).max()


@pytest.mark.parametrize("dim,p,n", [(2, 2, 4), (2, 3, 3), (3, 2, 3)])
def test_block_diagonal_local_solvers(dim, p, n):
    """FD::MatrixIntegrator::assemble_subspace_inverses (linelasticity_integrator.h:604-647):
    block c of the local solver is the (c, c) block of R_j A R_j^T (Cartesian), and
    the fast-diagonalization inverse applies its inverse."""
    A = de.assemble(dim, p, n, IP, MU, LAM)
    sm = o.Smoother(dim, p, n, ip=IP, elasticity=(MU, LAM))
    idx = sm.patch_dof_indices().reshape(sm.P, -1).astype(np.int64)
    kd = sm.ndp // dim
    V = sm.patch_vertices()
    for j in range(sm.P):
        Aj = A[np.ix_(idx[j], idx[j])]
        D = np.zeros_like(Aj)
        for c in range(dim):
            blk = slice(c * kd, (c + 1) * kd)
            D[blk, blk] = Aj[blk, blk]
            # the Kronecker sum of the 1D matrices with the component's coefficients
            mats = []
            for d in range(dim):
                a_cons = 2 * MU * (1.0 if c == d else 0.5) + (LAM if c == d else 0.0)
                a_pen = 2 * MU + (LAM if c == d else 0.0)
                mats.append(o.patch_matrices_1d_coef(p, [1.0 / n] * 2, V[j, d] == 1,
                                                     V[j, d] == n - 1, IP, a_cons, a_pen))
            K = 0.0
            for d in range(dim):
                term = np.ones((1, 1))
                for e in reversed(range(dim)):
                    term = np.kron(term, mats[e][1] if e == d else mats[e][0])
                K = K + term
            assert np.abs(K - Aj[blk, blk]).max() <= 1e-12 * np.abs(Aj).max()
        r = o.splitmix_uniform(j + 3, sm.ndp)
        u = sm.local_apply(j, r)
        assert np.abs(u - np.linalg.solve(D, r)).max() <= 1e-12 * np.abs(u).max()


def test_patch_dof_indices_are_component_blocks():
    dim, p, n = 3, 2, 3
    sm = o.Smoother(dim, p, n, elasticity=(1.0, 1.0))
    ps = o.Smoother(dim, p, n)
    idx = sm.patch_dof_indices().reshape(sm.P, dim, -1)
    base = ps.patch_dof_indices()
    Ns = n ** dim * (p + 1) ** dim
    for c in range(dim):
        np.testing.assert_array_equal(idx[:, c, :], base + c * Ns)


def _brute_force_precond(A, sm, src, variant, omega):
    idx = sm.patch_dof_indices().reshape(sm.P, -1).astype(np.int64)
    cp = sm.color_ptr()
    kd = sm.ndp // sm.dim
    dst = np.zeros_like(src)

    def local(j, r):
        Aj = A[np.ix_(idx[j], idx[j])]
        D = np.zeros_like(Aj)
        for c in range(sm.dim):
            blk = slice(c * kd, (c + 1) * kd)
            D[blk, blk] = Aj[blk, blk]
        return np.linalg.solve(D, r)

    if variant == o.ADDITIVE:
        for j in range(sm.P):
            dst[idx[j]] += omega * local(j, src[idx[j]])
        return dst
    for k, c in enumerate(sm.color_sequence()):
        res = src - A @ dst
        upd = np.zeros_like(dst)
        for j in range(cp[c], cp[c + 1]):
            upd[idx[j]] += omega * local(j, res[idx[j]])
        dst += upd
    return dst


@pytest.mark.parametrize("dim,p,n,variant", [(2, 2, 4, o.ADDITIVE), (2, 2, 4, o.MULTIPLICATIVE),
                                             (3, 1, 3, o.MULTIPLICATIVE)])
def test_precond_brute_force(dim, p, n, variant):
    A = de.assemble(dim, p, n, IP, MU, LAM)
    omega = 0.25 if variant == o.ADDITIVE else 1.0
    col = o.COLOR_NONE if variant == o.ADDITIVE else o.COLOR_REDBLACK
    sm = o.Smoother(dim, p, n, ip=IP, variant=variant, coloring=col, relaxation=omega,
                    elasticity=(MU, LAM))
    src = o.splitmix_uniform(5, sm.N)
    ref = _brute_force_precond(A, sm, src, variant, omega)
    got = sm.precond_vmult(src)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # Richardson step x += P (b - A x)
    x0 = o.splitmix_uniform(6, sm.N)
    x = sm.step(x0.copy(), src)
    np.testing.assert_allclose(x, x0 + sm.precond_vmult(src - A @ x0), rtol=0,
                               atol=1e-12 * np.abs(x).max())


def mass_load(dim, p, n, ncomp):
    """b_i = int phi_i (a constant unit body force in every component)."""
    nodes = o.gauss_lobatto(p + 1)
    qx, qw = o.gauss_legendre(p + 1)
    w1 = np.array([sum(qw[q] * o.lagrange(nodes, qx[q])[0][i] for q in range(p + 1))
                   for i in range(p + 1)]) / n
    wc = w1
    for _ in range(1, dim):
        wc = np.kron(w1, wc)
    return np.tile(wc, ncomp * n ** dim)


@pytest.mark.slow
@pytest.mark.parametrize("p,variant,bound", [(1, o.MULTIPLICATIVE, 9.0), (3, o.MULTIPLICATIVE, 7.0),
                                             (1, o.ADDITIVE, 36.0), (3, o.ADDITIVE, 29.0)])
def test_gmg_iteration_bounds(p, variant, bound):
    """TestLinElasticityDiagOnly (tests/apps_linelasticity.cc:50-105,
    apps_linelasticity.h:17-104): 2D, mu = lambda = ip = 1, vertex patches with
    the block-diagonal local solvers, root mesh 2x2, coarse CG 1e-12, one pre-
    and one post-smoothing step, CG to 1e-8; bounds MVP 9 (Q1), 7 (Q3), AVP
    (damping 0.25) 36 (Q1), 29 (Q3) on the fractional iteration count."""
    omega = 1.0 if variant == o.MULTIPLICATIVE else 0.25
    col = o.COLOR_REDBLACK if variant == o.MULTIPLICATIVE else o.COLOR_NONE
    for levels in (3, 4, 5):
        g = o.GMG(2, p, 2, levels, ip=1.0, variant=variant, coloring=col, relaxation=omega,
                  elasticity=(1.0, 1.0))
        b = mass_load(2, p, 2 << (levels - 1), 2)
        x, it, hist = g.cg(b, rel_tol=1e-8, max_it=100)
        n_frac = np.log(1e-8) / np.log((hist[-1] / hist[0]) ** (1.0 / (len(hist) - 1)))
        assert it > 0 and n_frac < bound, (levels, it, n_frac)
