"""Device Kronecker SVD vs the reference's golden values and the FP64 oracle
(needs an MI355X)."""
import numpy as np
import pytest

import oracle as o

pytestmark = pytest.mark.gpu


def _tpss():
    import tpss_amd
    if tpss_amd.device_count() < 1:
        pytest.skip("no HIP device")
    return tpss_amd


def test_golden_values_diagonal():
    """tests/kroneckersvd.cc:221-284 (both SIMD lanes)."""
    t = _tpss()
    for d, g1, g0 in [
        (np.arange(1.0, 13.0), [0.710868, 1.735563, 2.760258, 3.784953], [2.545889, 2.899023, 3.252157]),
        (np.array([1, 2, 3, -12, 5, 6, 7, 8, -33, 10, 11, 12.0]),
         [-0.422376, -1.250524, 5.699786, -1.457304], [1.097814, 0.618112, -5.896291])]:
        B1, B0, s = t.ksvd(np.diag(d), 4, 4, 3, 3, 2)
        np.testing.assert_allclose(np.diag(B1[0]), g1, atol=1e-6)
        np.testing.assert_allclose(np.diag(B0[0]), g0, atol=1e-6)


def test_exact_rank_and_oracle_parity():
    t = _tpss()
    rng = np.random.default_rng(3)
    batch, m1, n1, m0, n0 = 5, 4, 3, 5, 6
    A = np.zeros((batch, m1 * m0, n1 * n0))
    for b in range(batch):
        for r in range(2):
            A[b] += np.kron(rng.standard_normal((m1, n1)), rng.standard_normal((m0, n0)))
    B1, B0, s = t.ksvd(A, m1, n1, m0, n0, 2)
    for b in range(batch):
        approx = np.kron(B1[b, 0], B0[b, 0]) + np.kron(B1[b, 1], B0[b, 1])
        np.testing.assert_allclose(approx, A[b], atol=1e-11 * np.abs(A[b]).max())
        r1, r0, rs = o.ksvd(A[b], m1, n1, m0, n0, 2)
        np.testing.assert_allclose(s[b], rs, rtol=1e-12)
        np.testing.assert_allclose(B1[b], r1, atol=1e-10)
        np.testing.assert_allclose(B0[b], r0, atol=1e-10)


def test_patch_matrix_kronecker_sum_is_rank_two():
    """A 2D FD patch matrix L (x) M + M (x) L has Kronecker rank 2 (config-3
    separable structure); its rank-2 KSVD reproduces it exactly."""
    t = _tpss()
    M1, L1 = o.patch_matrices_1d(3, [0.2, 0.3], True, False, 4.0)
    M0, L0 = o.patch_matrices_1d(3, [0.25, 0.25], False, False, 4.0)
    A = np.kron(L1, M0) + np.kron(M1, L0)
    B1, B0, s = t.ksvd(A, 8, 8, 8, 8, 2)
    approx = np.kron(B1[0], B0[0]) + np.kron(B1[1], B0[1])
    np.testing.assert_allclose(approx, A, atol=1e-10 * np.abs(A).max())
