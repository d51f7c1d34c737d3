"""Pins the Kronecker oracles (oracle/kron_oracle.py and the C oracle's KSVD)
with the reference's own known-answer tests: tests/tensor_product_matrix.cc:
379-693 (explicit Kronecker-sum matrices, vmult and Tvmult) and
tests/kroneckersvd.cc:287-416 (exact reconstruction by the Kronecker SVD of a
sum of Kronecker products and by CP-ALS)."""
import numpy as np
import pytest

import kron_oracle as ko
import oracle as o

M23 = np.array([[1.0, 3.0, 5.0], [2.0, 4.0, 6.0]])  # columns (1,2), (3,4), (5,6)


def test_id_times_id():
    I3 = np.eye(3)
    np.testing.assert_array_equal(ko.kron_matrix([[I3, I3]]), np.eye(9))


def test_id_times_2x2_is_block_diagonal():
    M = np.array([[1.0, 3.0], [2.0, 4.0]])
    A = ko.kron_matrix([[M, np.eye(2)]])
    ref = np.zeros((4, 4))
    ref[:2, :2] = M
    ref[2:, 2:] = M
    np.testing.assert_array_equal(A, ref)


def test_id_times_2x3():
    Id = np.zeros((2, 3))
    Id[0, 0] = Id[1, 1] = 1.0
    A = ko.kron_matrix([[M23, Id]])
    ref = np.zeros((4, 9))
    ref[:2, :3] = M23
    ref[2:, 3:6] = M23
    np.testing.assert_array_equal(A, ref)


def test_id3x1_times_2x3():
    Id = np.zeros((3, 1))
    Id[0, 0] = 1.0
    A = ko.kron_matrix([[M23, Id]])
    ref = np.zeros((6, 3))
    ref[:2, :] = M23
    np.testing.assert_array_equal(A, ref)


def test_anisotropic_sums_and_transpose():
    rng = np.random.default_rng(0)
    for random in (False, True):
        Ms = [rng.uniform(-1, 1, (2, 3)) if random else M23 for _ in range(2)]
        Ns = [rng.uniform(-1, 1, (4, 2)) if random else np.ones((4, 2)) for _ in range(2)]
        for rank1 in (True, False):
            terms = [[Ms[0], Ns[0]]] + ([] if rank1 else [[Ms[1], Ns[1]]])
            ref = np.kron(Ns[0], Ms[0]) + (0 if rank1 else np.kron(Ns[1], Ms[1]))
            np.testing.assert_allclose(ko.kron_matrix(terms), ref, rtol=0, atol=1e-15)
            np.testing.assert_allclose(ko.kron_matrix(terms).T, ref.T, rtol=0, atol=1e-15)


def test_pinv_threshold_and_kernel():
    rng = np.random.default_rng(1)
    Q, _ = np.linalg.qr(rng.normal(size=(6, 6)))
    s = np.array([5.0, 3.0, 1.0, 0.5, 1e-3, 1e-9])
    A = Q @ np.diag(s) @ Q.T
    np.testing.assert_allclose(ko.pinv(A) @ A, np.eye(6), atol=1e-6)
    P = ko.pinv(A, threshold=1e-3)
    np.testing.assert_allclose(P, Q[:, :4] @ np.diag(1 / s[:4]) @ Q[:, :4].T, atol=1e-12)
    P2 = ko.pinv(A, kernel_size=2)
    np.testing.assert_allclose(P2, P, atol=1e-12)


T1 = np.array([[1, 0, 1], [0, 1, 0.1], [1, 0.1, -7], [0, -1, 0]], dtype=float)
T2 = np.array([[1, 2], [2, 4]], dtype=float)
T3 = np.array([[0.5, 3, 2], [3.4, 1, -0.4], [2, -0.5, 0], [0, 0, 0]], dtype=float)
T4 = np.array([[1, 1], [1, 0]], dtype=float)


def test_cp_exact_reconstruction():
    """tests/kroneckersvd.cc:287-340: t1 x t2 x t2 + t3 x t4 x t4 + 0.8 t3 x t4 x t4
    (Kronecker rank 2 written with 3 terms) is reproduced exactly by rank-2 CP-ALS"""
    T5 = 0.8 * T3
    terms = [[T1, T2, T2], [T3, T4, T4], [T5, T4, T4]]
    approx = ko.kcp(terms, 2)
    full = lambda t: np.kron(t[0], np.kron(t[1], t[2]))
    ref = sum(full(t) for t in terms)
    got = sum(full(t) for t in approx)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


def test_cp_fixed_point_and_rank_one():
    rng = np.random.default_rng(2)
    terms = [[rng.normal(size=(3, 2)), rng.normal(size=(2, 2)), rng.normal(size=(2, 3))]]
    out = ko.kcp(terms, 1)
    full = lambda t: np.kron(t[0], np.kron(t[1], t[2]))
    np.testing.assert_allclose(full(out[0]), full(terms[0]), atol=1e-12)


def test_ksvd_of_a_sum_exact_reconstruction():
    """tests/kroneckersvd.cc:67-133: the rank-two KSVD of t1 (x) t2 + t3 (x) t4
    reproduces it, from the dense matrix (C oracle)"""
    Mat = np.kron(T1, T2) + np.kron(T3, T4)
    B1, B0, sig = o.ksvd(Mat, 4, 3, 2, 2, 2)
    rec = sum(np.kron(B1[r], B0[r]) for r in range(2))
    np.testing.assert_allclose(rec, Mat, rtol=0, atol=1e-10 * np.abs(Mat).max())
    # rank one: kron(t2, t1)
    Mat1 = np.kron(T4, T1)
    B1, B0, sig = o.ksvd(Mat1, 2, 2, 4, 3, 1)
    np.testing.assert_allclose(np.kron(B1[0], B0[0]), Mat1, atol=1e-12 * np.abs(Mat1).max())
