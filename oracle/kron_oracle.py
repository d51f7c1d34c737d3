"""CPU restatement (numpy, FP64) of the reference's generic Kronecker-sum
matrix and Kronecker decompositions -- test infrastructure only.

* kron_matrix: Tensors::TensorProductMatrix in the basic state,
  sum_r A_{r,order-1} (x) ... (x) A_{r,0}, dimension 0 rightmost
  (tensor_product_matrix.h:136-304; vmult/Tvmult tensor_product_matrix.templates.h:200-291
  applied to unit vectors give this dense matrix, as_table).
* pinv: InverseTable (matrix_helper.h:21-68) -- LAPACK SVD inverse with the
  singular values s_i / s_max < threshold, or the kernel_size smallest, dropped.
* kcp: compute_kcp (kroneckersvd.h:501-585) written out as the reference does
  it -- explicit unfoldings (unfold_rank1 / unfold_rankk, :158-204), Khatri-Rao
  products (alignedlinalg.h:214-223), Hadamard products of Gram matrices and
  the SVD pseudo-inverse U diag(invert_safe(s)) V^T (alignedlinalg.h:103-109,
  471-485), R^2 + 10 sweeps, initial factors = the first R input terms.
"""
import numpy as np


def kron_matrix(terms):
    """terms: list over rank of [A_0, A_1, ..., A_{order-1}] (A_d: m_d x n_d)"""
    out = None
    for t in terms:
        k = np.array([[1.0]])
        for A in t:  # A_0 rightmost: kron(A_d, previous)
            k = np.kron(np.asarray(A, dtype=float), k)
        out = k if out is None else out + k
    return out


def pinv(A, threshold=0.0, kernel_size=-1):
    U, s, Vt = np.linalg.svd(A)
    inv = np.zeros_like(s)
    keep = len(s) if kernel_size < 0 else len(s) - kernel_size
    for i in range(keep):
        if s[i] > 0 and (kernel_size >= 0 or s[i] >= threshold * s[0]):
            inv[i] = 1.0 / s[i]
    return Vt.T @ np.diag(inv) @ U.T


def _invert_safe(x):
    return 0.0 if abs(x) <= np.finfo(float).eps else 1.0 / x


def _pseudo_inverse(V):
    # alignedlinalg.h:471-485: U diag(invert_safe(s)) V^T (V symmetric here)
    U, s, Vt = np.linalg.svd(V)
    return U @ np.diag([_invert_safe(x) for x in s]) @ Vt


def _unfold_rank1(f, direction):
    first, second, third = {0: (f[0], f[2], f[1]), 1: (f[1], f[2], f[0]),
                            2: (f[2], f[1], f[0])}[direction]
    return np.einsum("i,j,k->ijk", first, second, third).reshape(len(first), -1)


def _khatri_rao(a, b):
    return np.einsum("ik,jk->ijk", a, b).reshape(a.shape[0] * b.shape[0], a.shape[1])


def kcp(terms, rank, iterations=None):
    """terms: list over input rank of [C_i, B_i, A_i] (= in[i][0], in[i][1],
    in[i][2]); returns the rank-R list in the same order."""
    vec = [[np.asarray(t[2], float).ravel(), np.asarray(t[1], float).ravel(),
            np.asarray(t[0], float).ravel()] for t in terms]
    A = [np.stack([vec[i][k] for i in range(rank)], axis=1) for k in range(3)]
    iters = rank * rank + 10 if iterations is None else iterations

    def unfold(direction):
        return sum(_unfold_rank1(v, direction) for v in vec)

    for _ in range(iters):
        V = (A[1].T @ A[1]) * (A[2].T @ A[2])
        A[0] = unfold(0) @ (_khatri_rao(A[2], A[1]) @ _pseudo_inverse(V))
        V = (A[0].T @ A[0]) * (A[2].T @ A[2])
        A[1] = unfold(1) @ (_khatri_rao(A[2], A[0]) @ _pseudo_inverse(V))
        V = (A[0].T @ A[0]) * (A[1].T @ A[1])
        A[2] = unfold(2) @ (_khatri_rao(A[1], A[0]) @ _pseudo_inverse(V))
    shapes = [np.shape(t) for t in terms[0]]
    return [[A[2][:, r].reshape(shapes[0]), A[1][:, r].reshape(shapes[1]),
             A[0][:, r].reshape(shapes[2])] for r in range(rank)]
