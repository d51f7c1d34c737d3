"""Pins the CPU oracle against closed forms, invariants, textbook routines and
the reference's own known-answer tests (no GPU)."""
import itertools

import numpy as np
import pytest
import scipy.linalg

import oracle as o
from dense_sipg import assemble, gll_nodes


# ---------------------------------------------------------------- 1D basics

@pytest.mark.parametrize("nq", [1, 2, 3, 4, 5, 6, 8])
def test_gauss_exactness(nq):
    x, w = o.gauss_legendre(nq)
    assert abs(w.sum() - 1.0) < 1e-15
    for m in range(2 * nq):
        assert abs(np.dot(w, x ** m) - 1.0 / (m + 1)) < 1e-14
    # and not exact beyond (degree 2nq)
    assert abs(np.dot(w, x ** (2 * nq)) - 1.0 / (2 * nq + 1)) > 1e-13


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 7])
def test_gauss_lobatto_nodes(p):
    np.testing.assert_allclose(o.gauss_lobatto(p + 1), gll_nodes(p), atol=1e-15)


@pytest.mark.parametrize("p", [1, 3, 4])
def test_lagrange_cardinal(p):
    nodes = o.gauss_lobatto(p + 1)
    for i, xi in enumerate(nodes):
        v, _ = o.lagrange(nodes, xi)
        np.testing.assert_allclose(v, np.eye(p + 1)[i], atol=1e-13)
    # derivative reproduces d/dx x^m
    for m in range(p + 1):
        x = 0.3141
        _, d = o.lagrange(nodes, x)
        assert abs(np.dot(d, nodes ** m) - (m * x ** (m - 1) if m else 0.0)) < 1e-12


# ---------------------------------------------------------------- operator

CASES = [(2, 1, 3), (2, 2, 2), (2, 3, 3), (2, 4, 2), (3, 1, 2), (3, 2, 3), (3, 3, 2)]


@pytest.mark.parametrize("dim,p,n", CASES)
def test_operator_matches_dense_assembly(dim, p, n):
    ip = 12.34
    A = assemble(dim, p, n, ip)
    sm = o.Smoother(dim, p, n, ip=ip)
    x = o.splitmix_uniform(7, sm.N)
    np.testing.assert_allclose(sm.operator(x), A @ x, rtol=0, atol=1e-11 * np.abs(A @ x).max())
    # and the general-geometry code on the same (uniform) vertices
    smv = o.Smoother(dim, p, n, ip=ip, vertices=o.uniform_vertices(dim, n))
    np.testing.assert_allclose(smv.operator(x), A @ x, rtol=0, atol=1e-11 * np.abs(A @ x).max())
    assert np.allclose(A, A.T, atol=1e-10 * np.abs(A).max())


@pytest.mark.parametrize("dim,p,n", [(2, 1, 4), (2, 3, 3), (3, 2, 2), (3, 3, 3)])
def test_operator_energy_closed_form(dim, p, n):
    """a(u,u) for u = x_0 on [0,1]^dim: |grad u|^2 = 1, no interior jumps,
    boundary x_0=1 gives sigma_b - 2, the 2(dim-1) side faces sigma_b/3 each."""
    ip = 1.7
    sm = o.Smoother(dim, p, n, ip=ip)
    nodes = o.gauss_lobatto(p + 1)
    ndc = (p + 1) ** dim
    u = np.zeros(sm.N)
    for c in range(n ** dim):
        c0 = c % n
        for i in range(ndc):
            u[c * ndc + i] = (c0 + nodes[i % (p + 1)]) / n
    sig_b = ip * 2 * p * (p + 1) * n
    expected = 1.0 + sig_b - 2.0 + 2 * (dim - 1) * sig_b / 3.0
    assert abs(u @ sm.operator(u) - expected) < 1e-10 * expected


def test_distorted_operator_spd():
    dim, p, n = 2, 2, 4
    V = o.distorted_vertices(dim, n, 0.25, seed=3)
    sm = o.Smoother(dim, p, n, vertices=V)
    A = np.stack([sm.operator(e) for e in np.eye(sm.N)], axis=1)
    assert np.allclose(A, A.T, atol=1e-10 * np.abs(A).max())
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0


# ---------------------------------------------------------------- patches

def brute_patch_dofs(dim, p, n, v):
    k, nb = 2 * (p + 1), p + 1
    out = []
    for i in range(k ** dim):
        mi = [(i // k ** e) % k for e in range(dim)]
        cell = [v[e] - 1 + mi[e] // nb for e in range(dim)]
        loc = [mi[e] % nb for e in range(dim)]
        cidx = sum(cell[e] * n ** e for e in range(dim))
        lidx = sum(loc[e] * nb ** e for e in range(dim))
        out.append(cidx * nb ** dim + lidx)
    return np.array(out, dtype=np.uint32)


@pytest.mark.parametrize("dim,p,n", [(2, 3, 4), (3, 1, 3), (3, 3, 4)])
def test_patch_dof_indices_and_order(dim, p, n):
    sm = o.Smoother(dim, p, n, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK)
    assert sm.P == (n - 1) ** dim
    assert sm.n_colors == 2 * 2 ** dim
    V = sm.patch_vertices()
    idx = sm.patch_dof_indices()
    for j in range(sm.P):
        np.testing.assert_array_equal(idx[j], brute_patch_dofs(dim, p, n, V[j]))
    # each interior vertex exactly once
    assert len({tuple(v) for v in V}) == sm.P
    # colors: same-color patches share no DoF and are not face-coupled
    cp = sm.color_ptr()
    for c in range(sm.n_colors):
        cells = {}
        for j in range(cp[c], cp[c + 1]):
            for off in itertools.product((0, 1), repeat=dim):
                cell = tuple(V[j][e] - 1 + off[e] for e in range(dim))
                assert cell not in cells
                cells[cell] = j
        for cell, j in cells.items():
            for e in range(dim):
                for s in (-1, 1):
                    nb = list(cell)
                    nb[e] += s
                    assert cells.get(tuple(nb), j) == j
        # interior patches precede boundary patches inside the color
        bd = [int(any(V[j][e] in (1, n - 1) for e in range(dim))) for j in range(cp[c], cp[c + 1])]
        assert bd == sorted(bd)


def test_redblack_color_formula():
    # apps/coloring.h:286-324 written out for a 2D example
    sm = o.Smoother(2, 1, 6, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK)
    V, cp = sm.patch_vertices(), sm.color_ptr()
    for c in range(sm.n_colors):
        for j in range(cp[c], cp[c + 1]):
            x = V[j] - 1
            shift = x % 2
            color = 2 * (shift[0] + 2 * shift[1]) + ((x[0] + shift[0]) // 2 + (x[1] + shift[1]) // 2) % 2
            assert color == c


@pytest.mark.parametrize("dim,p,n", [(2, 2, 3), (3, 1, 3)])
def test_restrict_prolongate_counts(dim, p, n):
    """sum_j R_j^T R_j x = (#patches containing the DoF) x
    (tests/patch_transfer.cc:172-248 invariant)."""
    sm = o.Smoother(dim, p, n)
    idx = sm.patch_dof_indices()
    counts = np.zeros(sm.N)
    for j in range(sm.P):
        counts[idx[j]] += 1
    ndc = (p + 1) ** dim
    for c in range(n ** dim):
        cc = [(c // n ** e) % n for e in range(dim)]
        expect = np.prod([min(ci, n - 2) - max(ci - 1, 0) + 1 for ci in cc])
        assert np.all(counts[c * ndc:(c + 1) * ndc] == expect)


# ---------------------------------------------------------------- local solvers

@pytest.mark.parametrize("dim,p,n", [(2, 1, 2), (2, 2, 4), (2, 5, 2), (3, 1, 2), (3, 2, 4)])
def test_patch_matrix_is_submatrix(dim, p, n):
    """FD patch matrix == R_j A R_j^T of the assembled level matrix, and its
    FD inverse inverts it (tests/apps_laplace_integrator.cc:56-167, 406-412)."""
    ip = 12.34
    A = assemble(dim, p, n, ip)
    sm = o.Smoother(dim, p, n, ip=ip)
    idx = sm.patch_dof_indices()
    V = sm.patch_vertices()
    k = 2 * (p + 1)
    for j in range(sm.P):
        Ms, Ls = [], []
        for d in range(dim):
            M, L = o.patch_matrices_1d(p, [1.0 / n, 1.0 / n], V[j][d] == 1, V[j][d] == n - 1, ip)
            Ms.append(M)
            Ls.append(L)
        Aj = np.zeros((k ** dim, k ** dim))
        for d in range(dim):
            term = np.ones((1, 1))
            for e in reversed(range(dim)):
                term = np.kron(term, Ls[e] if e == d else Ms[e])
            Aj += term
        Asub = A[np.ix_(idx[j], idx[j])]
        np.testing.assert_allclose(Aj, Asub, rtol=0, atol=1e-10 * np.abs(Asub).max())
        r = o.splitmix_uniform(11 + j, k ** dim)
        u = sm.local_apply(j, r)
        np.testing.assert_allclose(Asub @ u, r, rtol=0, atol=1e-10)


@pytest.mark.parametrize("p", [1, 3, 4])
def test_gen_eig_vs_lapack(p):
    M, L = o.patch_matrices_1d(p, [0.3, 0.45], True, False, 2.5)
    S, lam = o.gen_eig(L, M)
    ref = scipy.linalg.eigh(L, M, eigvals_only=True)
    np.testing.assert_allclose(lam, ref, rtol=1e-12)
    np.testing.assert_allclose(S.T @ M @ S, np.eye(len(lam)), atol=1e-12)
    np.testing.assert_allclose(S.T @ L @ S, np.diag(lam), atol=1e-10 * lam.max())


def test_fd_apply_equals_dense_inverse():
    dim, p = 3, 1
    k = 2 * (p + 1)
    Ms, Ls = [], []
    for d in range(dim):
        M, L = o.patch_matrices_1d(p, [0.2 + 0.1 * d, 0.25], d == 0, d == 2, 3.0)
        Ms.append(M)
        Ls.append(L)
    S = np.stack([o.gen_eig(Ls[d], Ms[d])[0] for d in range(dim)])
    lam = np.stack([o.gen_eig(Ls[d], Ms[d])[1] for d in range(dim)])
    A = np.zeros((k ** dim,) * 2)
    for d in range(dim):
        term = np.ones((1, 1))
        for e in reversed(range(dim)):
            term = np.kron(term, Ls[e] if e == d else Ms[e])
        A += term
    r = o.splitmix_uniform(5, k ** dim)
    np.testing.assert_allclose(o.fd_apply(dim, S, lam, r), np.linalg.solve(A, r), rtol=1e-11, atol=1e-12)


# ---------------------------------------------------------------- smoothers

def dense_level(dim, p, n, ip):
    A = assemble(dim, p, n, ip)
    return A


def brute_avp(A, idx, omega, src):
    dst = np.zeros_like(src)
    for j in range(idx.shape[0]):
        Aj = A[np.ix_(idx[j], idx[j])]
        dst[idx[j]] += omega * np.linalg.solve(Aj, src[idx[j]])
    return dst


def brute_mvp(A, idx, cp, seq, src):
    x = np.zeros_like(src)
    for s, c in enumerate(seq):
        r = src - A @ x
        upd = np.zeros_like(src)
        for j in range(cp[c], cp[c + 1]):
            Aj = A[np.ix_(idx[j], idx[j])]
            upd[idx[j]] += np.linalg.solve(Aj, r[idx[j]])
        x += upd
    return x


@pytest.mark.parametrize("dim,p,n", [(2, 3, 4), (3, 1, 3)])
def test_avp_bruteforce(dim, p, n):
    ip = 1.0
    A = dense_level(dim, p, n, ip)
    sm = o.Smoother(dim, p, n, ip=ip, variant=o.ADDITIVE, relaxation=0.25)
    src = o.splitmix_uniform(1, sm.N)
    ref = brute_avp(A, sm.patch_dof_indices(), 0.25, src)
    np.testing.assert_allclose(sm.precond_vmult(src), ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_avp_single_patch_is_exact_solve():
    dim, p, n = 2, 3, 2
    A = dense_level(dim, p, n, 1.0)
    sm = o.Smoother(dim, p, n, relaxation=1.0)
    b = o.splitmix_uniform(2, sm.N)
    np.testing.assert_allclose(sm.precond_vmult(b), np.linalg.solve(A, b), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("dim,p,n,rev,sym", [(2, 3, 5, False, False), (2, 2, 4, True, False),
                                             (2, 1, 4, False, True), (3, 1, 3, False, False)])
def test_mvp_bruteforce(dim, p, n, rev, sym):
    A = dense_level(dim, p, n, 1.0)
    sm = o.Smoother(dim, p, n, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK,
                    reverse=rev, symmetrized=sym)
    seq = sm.color_sequence()
    nc = sm.n_colors
    expect = list(range(nc))[::-1] if rev else list(range(nc))
    if sym:
        expect = expect + expect[-2::-1]
    assert list(seq) == expect
    src = o.splitmix_uniform(1, sm.N)
    ref = brute_mvp(A, sm.patch_dof_indices(), sm.color_ptr(), seq, src)
    np.testing.assert_allclose(sm.precond_vmult(src), ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_symmetrized_mvp_is_symmetric():
    dim, p, n = 2, 1, 4
    sm = o.Smoother(dim, p, n, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK, symmetrized=True)
    B = np.stack([sm.precond_vmult(e) for e in np.eye(sm.N)], axis=1)
    np.testing.assert_allclose(B, B.T, atol=1e-11 * np.abs(B).max())


def test_richardson_step():
    dim, p, n = 2, 2, 4
    sm = o.Smoother(dim, p, n, relaxation=0.25, damping=0.8, n_steps=2)
    A = dense_level(dim, p, n, 1.0)
    b = o.splitmix_uniform(3, sm.N)
    x0 = o.splitmix_uniform(4, sm.N)
    x = x0.copy()
    for _ in range(2):
        x = x + 0.8 * sm.precond_vmult(b - A @ x)
    np.testing.assert_allclose(sm.step(x0.copy(), b), x, rtol=0, atol=1e-12 * np.abs(x).max())
    np.testing.assert_allclose(sm.smoother_vmult(b), sm.step(np.zeros(sm.N), b), atol=0)


# ---------------------------------------------------------------- transfer / GMG

@pytest.mark.parametrize("dim,p", [(2, 3), (3, 2)])
def test_prolongation_exact_on_polynomials(dim, p):
    nc = 2
    nodes = o.gauss_lobatto(p + 1)
    ndc = (p + 1) ** dim

    def interp(n):
        u = np.zeros(n ** dim * ndc)
        for c in range(n ** dim):
            cc = [(c // n ** e) % n for e in range(dim)]
            for i in range(ndc):
                li = [(i // (p + 1) ** e) % (p + 1) for e in range(dim)]
                x = [(cc[e] + nodes[li[e]]) / n for e in range(dim)]
                u[c * ndc + i] = x[0] ** p + 0.5 * x[-1] ** (p - 1) * x[0]
        return u

    uf = np.zeros((2 * nc) ** dim * ndc)
    o.prolongate_add(dim, p, nc, interp(nc), uf)
    np.testing.assert_allclose(uf, interp(2 * nc), atol=1e-13)
    # restriction is the transpose
    a = o.splitmix_uniform(1, nc ** dim * ndc)
    b = o.splitmix_uniform(2, (2 * nc) ** dim * ndc)
    pa = np.zeros_like(b)
    o.prolongate_add(dim, p, nc, a, pa)
    assert abs(pa @ b - a @ o.restrict(dim, p, nc, b)) < 1e-12 * abs(pa @ b)


@pytest.mark.slow
@pytest.mark.parametrize("dim,p,levels,variant,bound", [
    (2, 4, 3, o.MULTIPLICATIVE, 3), (2, 4, 3, o.ADDITIVE, 15),
    (3, 2, 2, o.MULTIPLICATIVE, 3), (3, 2, 2, o.ADDITIVE, 21)])
def test_gmg_iteration_bounds(dim, p, levels, variant, bound):
    """tests/apps_poisson.cc:87-135: rel. tol 1e-6, root 2^dim cells, MVP < 3, AVP < 15 (2D p4), < 21 (3D p2)."""
    omega = 1.0 if variant == o.MULTIPLICATIVE else (0.25 if dim == 2 else 0.15)
    col = o.COLOR_REDBLACK if variant == o.MULTIPLICATIVE else o.COLOR_NONE
    g = o.GMG(dim, p, 2, levels, variant=variant, coloring=col, relaxation=omega)
    # a smooth load (the reference solves with its manufactured load)
    b = np.ones(g.N)
    x, it, hist = g.cg(b, rel_tol=1e-6, max_it=100)
    assert it > 0
    assert fractional_steps(hist, 1e-6) < bound, (it, hist)


def test_gmg_threads_do_not_change_the_result():
    """The threaded oracle GMG (the CPU baseline's solve) computes bit for bit what the
    serial one does: the threads split the patches of a color and the cells of an
    operator application, whose results do not depend on the split."""
    b = o.splitmix_uniform(2, o.GMG(3, 2, 2, 3).N)
    x1, it1, h1 = o.GMG(3, 2, 2, 3).cg(b, rel_tol=1e-8)
    x4, it4, h4 = o.GMG(3, 2, 2, 3, threads=4).cg(b, rel_tol=1e-8)
    assert it1 == it4 > 0
    assert np.array_equal(x1, x4) and np.array_equal(h1, h4)


def fractional_steps(hist, tol):
    """compute_fractional_steps, apps/postprocess.h:40-60."""
    n = len(hist) - 1
    rho = (hist[-1] / hist[0]) ** (1.0 / n)
    return np.log(tol) / np.log(rho)


# ---------------------------------------------------------------- Kronecker SVD

def test_ksvd_diagonal_golden_values():
    """tests/kroneckersvd.cc:221-284: rank-1 KSVD of diag(1..12) with factors
    4x4 (x) 3x3, and of the second SIMD lane with entries 3 and 8 replaced by
    -12 and -33 (values "generated by working python program", 1e-6)."""
    d = np.arange(1.0, 13.0)
    B1, B0, s = o.ksvd_diag(d, 4, 3, 2)
    np.testing.assert_allclose(B1[0], [0.710868, 1.735563, 2.760258, 3.784953], atol=1e-6)
    np.testing.assert_allclose(B0[0], [2.545889, 2.899023, 3.252157], atol=1e-6)
    d2 = d.copy()
    d2[3], d2[8] = -12.0, -33.0
    B1, B0, s = o.ksvd_diag(d2, 4, 3, 2)
    np.testing.assert_allclose(B1[0], [-0.422376, -1.250524, 5.699786, -1.457304], atol=1e-6)
    np.testing.assert_allclose(B0[0], [1.097814, 0.618112, -5.896291], atol=1e-6)


def test_ksvd_exact_rank2():
    rng = np.random.default_rng(0)
    A1, A0 = rng.standard_normal((2, 3, 4)), rng.standard_normal((2, 2, 5))
    A = np.kron(A1[0], A0[0]) + np.kron(A1[1], A0[1])
    B1, B0, s = o.ksvd(A, 3, 4, 2, 5, 2)
    approx = np.kron(B1[0], B0[0]) + np.kron(B1[1], B0[1])
    np.testing.assert_allclose(approx, A, atol=1e-12)


# ---------------------------------------------------------------- surrogate patch lengths

def test_surrogate_h_affine_and_cartesian():
    """mapping_info.templates.h:42-240: on Cartesian meshes h is the cell size;
    on an affinely stretched mesh the stretched cell size; on a sheared mesh
    the distance between opposite faces along reference lines (edge length)."""
    dim, p, n = 3, 2, 4
    sm = o.Smoother(dim, p, n, vertices=o.uniform_vertices(dim, n))
    np.testing.assert_allclose(sm.patch_h(), 1.0 / n, rtol=1e-14)
    V = o.uniform_vertices(dim, n)
    Vs = V * np.array([1.0, 2.0, 0.5])
    sm = o.Smoother(dim, p, n, vertices=Vs)
    np.testing.assert_allclose(sm.patch_h()[:, 1, :], 2.0 / n, rtol=1e-14)
    np.testing.assert_allclose(sm.patch_h()[:, 2, :], 0.5 / n, rtol=1e-14)
    Vsh = V.copy()
    Vsh[:, 0] += 0.3 * V[:, 1]  # shear: cell edges along y have length h*sqrt(1+0.09)
    sm = o.Smoother(dim, p, n, vertices=Vsh)
    np.testing.assert_allclose(sm.patch_h()[:, 1, :], np.sqrt(1.09) / n, rtol=1e-13)
    np.testing.assert_allclose(sm.patch_h()[:, 0, :], 1.0 / n, rtol=1e-13)
