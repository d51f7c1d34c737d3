"""Pins the oracle's restricted additive Schwarz (RAS) data and operator:
ras weights (compute_ras_weights, schwarz_preconditioner.templates.h:126-165),
restricted-DoF flags (compute_restricted_dofs_impl, dof_info.templates.h:192-313)
and the weighted / restricted scatter of apply_local_solvers (:585-662).

Checks: brute force from the patch index tables with the reference's sort
order written out; the partition-of-unity invariant of the reference's own
test (tests/patch_transfer.cc:1360-1510: rscatter_add of ones from every patch
gives the vector of ones); dense brute force of the preconditioner."""
import numpy as np
import pytest

import oracle as o
from dense_sipg import assemble


def _brute_force_flags(dim, p, n, idx):
    """sort (distance, local index, patch) per DoF as the reference does and
    keep the first; distance = unsigned(max_d |k/2 - 0.5 - m_d|)"""
    k = 2 * (p + 1)
    mid = k / 2.0 - 0.5
    entries = {}
    for j in range(idx.shape[0]):
        for i in range(idx.shape[1]):
            m = [(i // k ** d) % k for d in range(dim)]
            dist = int(max(abs(mid - md) for md in m))
            entries.setdefault(int(idx[j, i]), []).append((dist, i, j))
    flags = np.ones(idx.shape, dtype=np.uint8)
    for dof, lst in entries.items():
        dist, i, j = sorted(lst)[0]
        flags[j, i] = 0
    return flags, entries


CASES = [(2, 1, 4), (2, 2, 4), (2, 3, 3), (3, 1, 3), (3, 2, 3)]


@pytest.mark.parametrize("dim,p,n", CASES)
def test_ras_weights_and_flags_brute_force(dim, p, n):
    for ras in (o.RAS_WEIGHTS, o.RAS_BOOLEAN):
        sm = o.Smoother(dim, p, n, variant=o.ADDITIVE, relaxation=0.7, ras=ras)
        idx = sm.patch_dof_indices()
        counts = np.bincount(idx.ravel().astype(np.int64), minlength=sm.N)
        np.testing.assert_array_equal(sm.ras_weights(), counts)
        flags, entries = _brute_force_flags(dim, p, n, idx)
        np.testing.assert_array_equal(sm.restricted_flags(), flags)
        # every DoF lies in a patch (interior-vertex patches cover every cell)
        assert len(entries) == sm.N and counts.min() >= 1


@pytest.mark.parametrize("dim,p,n", CASES)
def test_restricted_prolongation_is_a_partition_of_unity(dim, p, n):
    """tests/patch_transfer.cc:1360-1510: rscatter_add of a local vector of ones
    from every patch gives exactly one on every DoF; with the RAS weights the
    weighted scatter of ones does as well."""
    sm = o.Smoother(dim, p, n, variant=o.ADDITIVE, ras=o.RAS_BOOLEAN)
    idx = sm.patch_dof_indices()
    keep = sm.restricted_flags() == 0
    v = np.zeros(sm.N)
    np.add.at(v, idx[keep].astype(np.int64), 1.0)
    np.testing.assert_array_equal(v, np.ones(sm.N))
    # and sum_j R~_j^T R_j x = x for any x
    x = o.splitmix_uniform(5, sm.N)
    y = np.zeros(sm.N)
    for j in range(sm.P):
        loc = x[idx[j]]
        np.add.at(y, idx[j][keep[j]].astype(np.int64), loc[keep[j]])
    np.testing.assert_array_equal(y, x)
    smw = o.Smoother(dim, p, n, variant=o.ADDITIVE, ras=o.RAS_WEIGHTS)
    w = smw.ras_weights()
    z = np.zeros(sm.N)
    for j in range(sm.P):
        np.add.at(z, idx[j].astype(np.int64), x[idx[j]] / w[idx[j]])
    np.testing.assert_allclose(z, x, rtol=1e-15, atol=0)


@pytest.mark.parametrize("dim,p,n", [(2, 2, 4), (2, 3, 3), (3, 1, 3), (3, 2, 3)])
@pytest.mark.parametrize("ras", [1, 2])
def test_ras_preconditioner_dense_brute_force(dim, p, n, ras):
    ip, omega = 2.0, 0.8
    A = assemble(dim, p, n, ip)
    sm = o.Smoother(dim, p, n, ip=ip, variant=o.ADDITIVE, relaxation=omega, ras=ras)
    idx = sm.patch_dof_indices().astype(np.int64)
    flags = sm.restricted_flags()
    w = sm.ras_weights()
    src = o.splitmix_uniform(3, sm.N)
    ref = np.zeros(sm.N)
    for j in range(sm.P):
        Aj = A[np.ix_(idx[j], idx[j])]
        u = omega * np.linalg.solve(Aj, src[idx[j]])
        if ras == 1:
            np.add.at(ref, idx[j], u / w[idx[j]])
        else:
            keep = flags[j] == 0
            np.add.at(ref, idx[j][keep], u[keep])
    got = sm.precond_vmult(src)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_ras_on_a_single_patch_is_the_exact_solve():
    """n = 2: one interior vertex, one patch covering the whole mesh: every DoF
    is kept by it and has weight one, so RAS with omega = 1 solves exactly."""
    for dim, p in ((2, 3), (3, 2)):
        A = assemble(dim, p, 2, 1.0)
        b = o.splitmix_uniform(9, A.shape[0])
        for ras in (1, 2):
            sm = o.Smoother(dim, p, 2, variant=o.ADDITIVE, relaxation=1.0, ras=ras)
            np.testing.assert_allclose(A @ sm.precond_vmult(b), b, rtol=0, atol=1e-10)


def test_ras_needs_the_additive_scalar_variant():
    with pytest.raises(ValueError):
        o.Smoother(2, 2, 4, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK, ras=o.RAS_BOOLEAN)
    with pytest.raises(ValueError):
        o.Smoother(2, 1, 4, variant=o.ADDITIVE, elasticity=(1.0, 1.0), ras=o.RAS_WEIGHTS)


def test_restricted_flags_prefer_the_patch_nearest_the_dof():
    """2D Q3 (k = 8, patch centre 3.5): a DoF at local node l of its cell is
    kept by the patch where it is nearest the vertex -- nodes 0, 1 (cell is the
    upper cell of the patch: distance l) and 2, 3 (lower cell: distance 3 - l) --
    except at the boundary where that patch does not exist."""
    sm = o.Smoother(2, 3, 4, variant=o.ADDITIVE, ras=o.RAS_BOOLEAN)
    idx = sm.patch_dof_indices()
    keep = sm.restricted_flags() == 0
    V = sm.patch_vertices()
    owner = {}
    for j in range(sm.P):
        for i in np.nonzero(keep[j])[0]:
            owner[int(idx[j, i])] = tuple(V[j])
    # DoF in cell (1, 1) at local (0, 0): patch at vertex (1, 1) (upper cell in both)
    cell, ndc = 1 + 4 * 1, 16
    assert owner[cell * ndc + 0] == (1, 1)
    # local (3, 3): patch at vertex (2, 2) (lower cell in both)
    assert owner[cell * ndc + 15] == (2, 2)
    # local (1, 2) -> x: node 1 upper (vertex 1); y: node 2 lower (vertex 2)
    assert owner[cell * ndc + 1 + 4 * 2] == (1, 2)
    # cell (0, 0) at local (0, 0): only the patch at vertex (1, 1) exists
    assert owner[0] == (1, 1)
