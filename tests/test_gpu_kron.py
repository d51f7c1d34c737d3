"""Generic Kronecker-sum matrices, their inverse, the Kronecker SVD of a sum
of Kronecker products and CP-ALS on the device vs the oracle (C ABI) and the
reference's known-answer tests (tests/tensor_product_matrix.cc:379-693,
tests/kroneckersvd.cc:67-416).  FP64: within 1e-12 of the oracle; the
decompositions reconstruct their input to 1e-10."""
import numpy as np
import pytest

import kron_oracle as ko
import oracle as o
from test_gpu_parity import _tpss

pytestmark = pytest.mark.gpu

M23 = np.array([[1.0, 3.0, 5.0], [2.0, 4.0, 6.0]])


def _dense(t, terms, transpose=False):
    return t.kron_table(terms, transpose=transpose)[0]


@pytest.mark.parametrize("case", ["id_id", "id_2x2", "id_2x3", "id3x1_2x3", "2x3_2x3",
                                  "sum_2x3_2x3", "aniso_one", "aniso_sum", "aniso_rand",
                                  "aniso_rand_sum", "order1", "order3"])
def test_kron_vmult_tvmult_table(case):
    t = _tpss()
    rng = np.random.default_rng(7)
    Id22 = np.eye(2)
    Id23 = np.zeros((2, 3))
    Id23[0, 0] = Id23[1, 1] = 1
    Id31 = np.zeros((3, 1))
    Id31[0, 0] = 1
    R = lambda *s: rng.uniform(-1, 1, s)
    terms = {
        "id_id": [[np.eye(3), np.eye(3)]],
        "id_2x2": [[np.array([[1.0, 3.0], [2.0, 4.0]]), Id22]],
        "id_2x3": [[M23, Id23]],
        "id3x1_2x3": [[M23, Id31]],
        "2x3_2x3": [[M23, M23]],
        "sum_2x3_2x3": [[R(2, 3), R(2, 3)], [R(2, 3), R(2, 3)]],
        "aniso_one": [[M23, np.ones((4, 2))]],
        "aniso_sum": [[M23, np.ones((4, 2))], [M23, np.ones((4, 2))]],
        "aniso_rand": [[R(2, 3), R(4, 2)]],
        "aniso_rand_sum": [[R(2, 3), R(4, 2)], [R(2, 3), R(4, 2)]],
        "order1": [[R(5, 3)], [R(5, 3)]],
        "order3": [[R(3, 2), R(4, 4), R(2, 5)], [R(3, 2), R(4, 4), R(2, 5)],
                   [R(3, 2), R(4, 4), R(2, 5)]],
    }[case]
    ref = ko.kron_matrix(terms)
    np.testing.assert_allclose(_dense(t, terms), ref, rtol=0, atol=1e-14)
    np.testing.assert_allclose(_dense(t, terms, transpose=True), ref.T, rtol=0, atol=1e-14)
    x = rng.uniform(-1, 1, ref.shape[1])
    y = rng.uniform(-1, 1, ref.shape[0])
    np.testing.assert_allclose(t.kron_vmult(terms, x), ref @ x, rtol=0, atol=1e-13)
    np.testing.assert_allclose(t.kron_vmult(terms, y, transpose=True), ref.T @ y, rtol=0,
                               atol=1e-13)
    d0 = rng.uniform(-1, 1, ref.shape[0])
    np.testing.assert_allclose(t.kron_vmult(terms, x, add_to=d0), d0 + ref @ x, rtol=0, atol=1e-13)


def test_kron_batched_patch_sized():
    """a batch of 3D separable patch matrices (k = 8): M (x) M (x) L + ... against the
    dense oracle, and the inverse of one of them against numpy"""
    t = _tpss()
    rng = np.random.default_rng(3)
    batch, k = 5, 8
    all_terms = []
    for _ in range(batch):
        Ms = [np.eye(k) + 0.1 * rng.normal(size=(k, k)) for _ in range(3)]
        Ls = [rng.normal(size=(k, k)) for _ in range(3)]
        all_terms.append([[Ls[0], Ms[1], Ms[2]], [Ms[0], Ls[1], Ms[2]], [Ms[0], Ms[1], Ls[2]]])
    X = rng.uniform(-1, 1, (batch, k ** 3))
    Y = t.kron_vmult_batched(all_terms, X)
    for b in range(batch):
        np.testing.assert_allclose(Y[b], ko.kron_matrix(all_terms[b]) @ X[b], rtol=0,
                                   atol=1e-12 * np.abs(Y[b]).max())


@pytest.mark.parametrize("n,threshold,kernel", [(9, 0.0, -1), (16, 1e-6, -1), (12, 0.0, 2),
                                                 (64, 0.0, -1), (1, 0.0, -1)])
def test_pinv_matches_oracle(n, threshold, kernel):
    t = _tpss()
    rng = np.random.default_rng(n)
    A = rng.normal(size=(3, n, n))
    if kernel > 0 or threshold > 0:  # make it numerically rank-deficient
        for b in range(3):
            U, s, Vt = np.linalg.svd(A[b])
            s[-2:] = [1e-9, 1e-12]
            A[b] = U @ np.diag(s) @ Vt
    P = t.pinv_batched(A, threshold=threshold, kernel_size=kernel)
    for b in range(3):
        ref = ko.pinv(A[b], threshold=threshold, kernel_size=kernel)
        np.testing.assert_allclose(P[b], ref, rtol=0, atol=1e-9 * np.abs(ref).max())
        if kernel < 0 and threshold == 0:
            np.testing.assert_allclose(P[b] @ A[b], np.eye(n), atol=1e-10)


T1 = np.array([[1, 0, 1], [0, 1, 0.1], [1, 0.1, -7], [0, -1, 0]], dtype=float)
T2 = np.array([[1, 2], [2, 4]], dtype=float)
T3 = np.array([[0.5, 3, 2], [3.4, 1, -0.4], [2, -0.5, 0], [0, 0, 0]], dtype=float)
T4 = np.array([[1, 1], [1, 0]], dtype=float)


def test_ksvd_of_a_sum_of_kronecker_products():
    """tests/kroneckersvd.cc:67-133, 175-180: rank-two KSVD of t1 (x) t2 + t3 (x) t4
    given as its terms reproduces the matrix; the same factors as the dense input"""
    t = _tpss()
    Mat = np.kron(T1, T2) + np.kron(T3, T4)
    B1, B0, s = t.ksvd_sum([T1, T3], [T2, T4], 2)
    rec = sum(np.kron(B1[r], B0[r]) for r in range(2))
    np.testing.assert_allclose(rec, Mat, rtol=0, atol=1e-10 * np.abs(Mat).max())
    D1, D0, ds = t.ksvd(Mat, 4, 3, 2, 2, 2)
    np.testing.assert_allclose(s, ds, rtol=1e-12)
    np.testing.assert_allclose(B1, D1, atol=1e-10)
    np.testing.assert_allclose(B0, D0, atol=1e-10)
    # and the oracle's dense KSVD
    O1, O0, os_ = o.ksvd(Mat, 4, 3, 2, 2, 2)
    np.testing.assert_allclose(s, os_, rtol=1e-12)
    # a batch of random sums, rank 3 of 4 terms
    rng = np.random.default_rng(4)
    big = rng.normal(size=(6, 4, 5, 5))
    small = rng.normal(size=(6, 4, 3, 4))
    B1, B0, s = t.ksvd_sum(big, small, 3)
    for b in range(6):
        M = sum(np.kron(big[b, i], small[b, i]) for i in range(4))
        R1, R0, rs = o.ksvd(M, 5, 5, 3, 4, 3)
        np.testing.assert_allclose(s[b], rs, rtol=1e-10)


def test_kcp_matches_oracle_and_reconstructs():
    """tests/kroneckersvd.cc:287-340 (three-dimensional CP): exact reconstruction of a
    Kronecker-rank-2 matrix given with three terms; a random batch against the
    oracle's literal CP-ALS"""
    t = _tpss()
    T5 = 0.8 * T3
    terms = [[T1, T2, T2], [T3, T4, T4], [T5, T4, T4]]
    out = t.kcp([terms], 2)[0]
    full = lambda tt: np.kron(tt[0], np.kron(tt[1], tt[2]))
    ref = sum(full(tt) for tt in terms)
    np.testing.assert_allclose(sum(full(tt) for tt in out), ref, rtol=0,
                               atol=1e-10 * np.abs(ref).max())
    oref = ko.kcp(terms, 2)
    for r in range(2):
        for k in range(3):
            np.testing.assert_allclose(out[r][k], oref[r][k], rtol=0,
                                       atol=1e-9 * np.abs(oref[r][k]).max())
    rng = np.random.default_rng(5)
    batch = [[[rng.normal(size=(3, 3)), rng.normal(size=(2, 4)), rng.normal(size=(4, 2))]
              for _ in range(4)] for _ in range(5)]
    outs = t.kcp(batch, 3)
    for b in range(5):
        oref = ko.kcp(batch[b], 3)
        got = sum(full(tt) for tt in outs[b])
        exp = sum(full(tt) for tt in oref)
        np.testing.assert_allclose(got, exp, rtol=0, atol=1e-8 * np.abs(exp).max())
