"""Pins the oracle's moved-vertex SIPG operator (or_* operator_distorted,
restating apps/laplace_integrator.h:613-743 on multilinear cells) with checks
that a transposed inverse Jacobian, a wrong face normal or surface element, or
a penalty length taken from the wrong row of J^{-1} fail:

* an independent dense assembly from the geometry (tests/dense_sipg_mapped.py)
  on affine sheared/stretched meshes and on randomly moved vertices;
* linear consistency: A (a.x + c) vanishes on the rows of cells without a
  boundary face (QGauss(p+1) integrates the discrete integration by parts
  exactly on multilinear cells);
* the closed-form energy a(x_0, x_0) on a sheared parallelepiped, whose
  penalty lengths are the cell heights, not the side lengths.
"""
import itertools

import numpy as np
import pytest

import oracle as o
from dense_sipg_mapped import _corners, _map, affine_vertices, assemble_mapped


def _shear(dim):
    M = np.eye(dim)
    M[0, 1], M[1, 1] = 0.4, 1.3
    if dim == 3:
        M[2, 0], M[1, 2], M[2, 2] = -0.3, 0.2, 0.8
    return M


def _node_coordinates(dim, p, n, V):
    """physical coordinates of every DoF node (cell-major, local lexicographic)"""
    nodes = o.gauss_lobatto(p + 1)
    nb = p + 1
    out = []
    for c in range(n ** dim):
        cc = [(c // n ** e) % n for e in range(dim)]
        X = _corners(V, dim, n, cc)
        for i in range(nb ** dim):
            xi = [nodes[(i // nb ** e) % nb] for e in range(dim)]
            out.append(_map(X, dim, xi)[0])
    return np.array(out)


MESHES = ["shear", "stretch", "moved"]


def _vertices(kind, dim, n):
    if kind == "shear":
        return affine_vertices(dim, n, _shear(dim), shift=[0.1] * dim)
    if kind == "stretch":
        return affine_vertices(dim, n, np.diag([0.5, 2.0, 1.5][:dim]))
    return o.distorted_vertices(dim, n, 0.25, seed=5)


@pytest.mark.parametrize("kind", MESHES)
@pytest.mark.parametrize("dim,p,n", [(2, 1, 3), (2, 2, 3), (2, 3, 2), (3, 1, 2), (3, 2, 2)])
def test_moved_vertex_operator_matches_dense_assembly(kind, dim, p, n):
    ip = 12.34
    V = _vertices(kind, dim, n)
    A = assemble_mapped(dim, p, n, ip, V)
    sm = o.Smoother(dim, p, n, ip=ip, vertices=V)
    for seed in (7, 8):
        x = o.splitmix_uniform(seed, sm.N)
        z = A @ x
        np.testing.assert_allclose(sm.operator(x), z, rtol=0, atol=1e-12 * np.abs(z).max())
    assert np.allclose(A, A.T, atol=1e-11 * np.abs(A).max())


def test_dense_assembly_is_the_cartesian_one_on_uniform_vertices():
    from dense_sipg import assemble
    for dim, p, n in [(2, 2, 3), (3, 1, 2)]:
        A0 = assemble(dim, p, n, 3.0)
        A1 = assemble_mapped(dim, p, n, 3.0, o.uniform_vertices(dim, n))
        np.testing.assert_allclose(A1, A0, rtol=0, atol=1e-12 * np.abs(A0).max())


@pytest.mark.parametrize("kind", ["moved", "shear"])
@pytest.mark.parametrize("dim,p,n", [(2, 1, 4), (2, 2, 4), (2, 3, 4), (3, 1, 3), (3, 2, 3),
                                     (3, 3, 3)])
def test_linear_functions_are_in_the_kernel_away_from_the_boundary(kind, dim, p, n):
    V = _vertices(kind, dim, n)
    sm = o.Smoother(dim, p, n, ip=4.0, vertices=V)
    X = _node_coordinates(dim, p, n, V)
    ndc = (p + 1) ** dim
    a = np.array([0.7, -1.3, 0.4][:dim])
    u = X @ a + 0.25
    y = sm.operator(u)
    inner = [c for c in range(n ** dim)
             if all(0 < (c // n ** e) % n < n - 1 for e in range(dim))]
    assert inner
    rows = np.concatenate([np.arange(c * ndc, (c + 1) * ndc) for c in inner])
    scale = np.abs(y).max()
    assert scale > 1.0  # the boundary rows do not vanish
    assert np.abs(y[rows]).max() < 1e-12 * scale


def _sheared_energy_closed_form(dim, p, n, ip, M):
    """a(u, u) for u = x_0 on the parallelepiped M [0,1]^dim (no shift)."""
    gamma = p * (p + 1)
    vol = abs(np.linalg.det(M))
    energy = vol  # |grad u|^2 = 1
    for d in range(dim):
        tang = [M[:, e] for e in range(dim) if e != d]
        if dim == 2:
            area = np.linalg.norm(tang[0])
            nrm = np.array([tang[0][1], -tang[0][0]]) / area
        else:
            cr = np.cross(tang[0], tang[1])
            area = np.linalg.norm(cr)
            nrm = cr / area
        h = vol / (n * area)  # cell height over this face
        sigma = ip * 2.0 * gamma / h
        for side in (0, 1):
            nv = nrm if nrm @ M[:, d] * (2 * side - 1) > 0 else -nrm
            # u on the face: c + sum_t s_t a_t, s uniform on [0,1]^(dim-1)
            c = side * M[0, d]
            at = np.array([M[0, e] for e in range(dim) if e != d])
            mean_u = c + at.sum() / 2.0
            mean_u2 = mean_u ** 2 + (at ** 2).sum() / 12.0
            energy += area * (sigma * mean_u2 - 2.0 * nv[0] * mean_u)
    return energy


@pytest.mark.parametrize("dim,p,n", [(2, 1, 3), (2, 3, 2), (3, 1, 2), (3, 2, 3)])
def test_sheared_energy_closed_form(dim, p, n):
    ip = 1.7
    M = _shear(dim)
    V = affine_vertices(dim, n, M)
    sm = o.Smoother(dim, p, n, ip=ip, vertices=V)
    u = _node_coordinates(dim, p, n, V)[:, 0]
    expected = _sheared_energy_closed_form(dim, p, n, ip, M)
    assert abs(u @ sm.operator(u) - expected) < 1e-11 * abs(expected)
    # with the Cartesian side lengths instead of the heights the energy differs
    side = np.linalg.norm(M[:, 0]) / n
    assert abs(side - abs(np.linalg.det(M)) / (n * np.linalg.norm(M[:, 1] if dim == 2 else
                                                                   np.cross(M[:, 1], M[:, 2])))) > 1e-3


@pytest.mark.parametrize("dim,p,n", [(2, 2, 3), (3, 1, 3), (3, 2, 2)])
def test_local_matrix_is_the_submatrix_of_the_dense_assembly(dim, p, n):
    """or_patch_local_matrix (unit vectors applied cell by cell on moved-vertex
    meshes) == R_j A R_j^T of the independent dense assembly"""
    V = o.distorted_vertices(dim, n, 0.25, seed=6)
    A = assemble_mapped(dim, p, n, 3.0, V)
    sm = o.Smoother(dim, p, n, ip=3.0, vertices=V)
    idx = sm.patch_dof_indices().astype(np.int64)
    for j in range(sm.P):
        Aj = sm.patch_local_matrix(j)
        ref = A[np.ix_(idx[j], idx[j])]
        np.testing.assert_allclose(Aj, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
