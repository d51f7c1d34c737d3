"""The rest of the drop-in boundary on the device: vmult_add / Tvmult /
Tvmult_add (schwarz_preconditioner.h:185-194), the multiplicative sweep with
the tiled 2^dim colors, and the error paths of the C ABI (non-finite output,
collapsed cells, CG without convergence)."""
import numpy as np
import pytest

import oracle as o
from test_gpu_parity import _tpss, dev, host, make, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,p,n,variant", [(2, 3, 7, 0), (2, 2, 6, 1), (3, 2, 4, 1), (3, 3, 4, 0)])
def test_vmult_add_and_transpose(dim, p, n, variant):
    t = _tpss()
    lv, sm, ref = make(dim, p, n, variant)
    src = o.splitmix_uniform(11, ref.N)
    d0 = o.splitmix_uniform(12, ref.N)
    Psrc = ref.precond_vmult(src)
    # vmult_add: dst += P src
    dst = dev(d0)
    sm.vmult_add(dst, dev(src))
    assert rel_err(host(dst), d0 + Psrc) < 1e-12
    # Tvmult: the reversed color sequence (the additive operator is symmetric)
    _, _, ref_t = make(dim, p, n, variant, reverse=True)
    dst = dev(np.zeros(ref.N))
    sm.Tvmult(dst, dev(src))
    PTsrc = host(dst)
    assert rel_err(PTsrc, ref_t.precond_vmult(src)) < 1e-12
    if variant == t.ADDITIVE:
        assert rel_err(PTsrc, Psrc) < 1e-12
    # adjoint identity <P^T x, y> = <x, P y>
    y = o.splitmix_uniform(13, ref.N)
    dy = dev(np.zeros(ref.N))
    sm.vmult(dy, dev(y))
    lhs, rhs = PTsrc @ y, src @ host(dy)
    assert abs(lhs - rhs) < 1e-12 * (np.abs(PTsrc).max() * np.abs(y).sum())
    # Tvmult_add
    dst = dev(d0)
    sm.Tvmult_add(dst, dev(src))
    assert rel_err(host(dst), d0 + PTsrc) < 1e-12


@pytest.mark.parametrize("dim,p,n,fused", [(2, 3, 8, True), (3, 3, 5, True), (3, 2, 6, False)])
def test_tiled_coloring_mvp_parity(dim, p, n, fused):
    """The multiplicative smoother with the 2^dim tiled colors (8 in 3D, the
    color count of the 64^3 Q3 configuration) against the oracle."""
    t = _tpss()
    lv, sm, ref = make(dim, p, n, t.MULTIPLICATIVE, coloring=t.COLOR_TILED)
    assert sm.n_colors == 2 ** dim == ref.n_colors
    np.testing.assert_array_equal(sm.color_ptr(), ref.color_ptr())
    sm.set_fused_residual(fused)
    src = o.splitmix_uniform(3, ref.N)
    dst = dev(np.zeros(ref.N))
    sm.precond_vmult(dst, dev(src))
    assert rel_err(host(dst), ref.precond_vmult(src)) < 1e-12
    x = dev(o.splitmix_uniform(4, ref.N))
    sm.vmult(x, dev(src))
    assert rel_err(host(x), ref.smoother_vmult(src)) < 1e-12


def test_nonfinite_output_is_reported():
    t = _tpss()
    lv, sm, ref = make(3, 2, 4, t.MULTIPLICATIVE)
    src = o.splitmix_uniform(3, ref.N)
    bad = src.copy()
    bad[17] = np.nan
    dst = dev(np.zeros(ref.N))
    # default mode: no synchronisation in the call, reported afterwards
    sm.vmult(dst, dev(bad))
    with pytest.raises(t.TPSSError) as e:
        sm.finite_status()
    assert e.value.code == t.E_NONFINITE
    sm.finite_status()  # the report clears the flag
    # reported by the next apply call once the check has run
    sm.vmult(dst, dev(bad))
    t.synchronize()
    with pytest.raises(t.TPSSError) as e:
        sm.vmult(dst, dev(src))
    assert e.value.code == t.E_NONFINITE
    sm.vmult(dst, dev(src))
    assert rel_err(host(dst), ref.smoother_vmult(src)) < 1e-12
    sm.finite_status()
    # synchronous mode reports in the call itself
    sm.set_check_finite(2)
    with pytest.raises(t.TPSSError) as e:
        sm.precond_vmult(dst, dev(bad))
    assert e.value.code == t.E_NONFINITE
    # off: nothing is reported
    sm.set_check_finite(0)
    sm.precond_vmult(dst, dev(bad))
    sm.finite_status()
    # Inf in an FP32 level vector
    lv32 = t.Level(2, 3, 5, precision=t.F32)
    sm32 = t.SchwarzSmoother(lv32, variant=t.ADDITIVE)
    b32 = np.ones(lv32.n_dofs, dtype=np.float32)
    b32[3] = np.inf
    d32 = dev(np.zeros(lv32.n_dofs, dtype=np.float32))
    sm32.precond_vmult(d32, dev(b32))
    with pytest.raises(t.TPSSError) as e:
        sm32.finite_status()
    assert e.value.code == t.E_NONFINITE


@pytest.mark.parametrize("dim,p,n", [(3, 3, 6), (3, 2, 4), (3, 3, 16)])
def test_nonfinite_check_of_the_residual_free_passes(dim, p, n):
    """The residual-free color passes check the values they store themselves
    (no separate pass over the output); the additive path and the accumulating
    entry points keep the pass over the output.  (3, 3, 16): colors of at most
    512 patches run the four-wave pass."""
    t = _tpss()
    lv, sm, ref = make(dim, p, n, t.MULTIPLICATIVE)
    N = ref.N
    src = o.splitmix_uniform(3, N)
    # a NaN in the right-hand side of a Richardson step (read by the patches
    # holding that DoF) reaches the output
    b = src.copy()
    b[N // 2] = np.nan
    sm.step(dev(o.splitmix_uniform(4, N)), dev(b))
    with pytest.raises(t.TPSSError) as e:
        sm.finite_status()
    assert e.value.code == t.E_NONFINITE
    # Tvmult with an Inf in the source
    bad = src.copy()
    bad[N - 1] = np.inf
    dst = dev(np.zeros(N))
    sm.Tvmult(dst, dev(bad))
    with pytest.raises(t.TPSSError) as e:
        sm.finite_status()
    assert e.value.code == t.E_NONFINITE
    # vmult_add: the passes write a temporary; a NaN already in dst is found by
    # the pass over the output
    d = np.zeros(N)
    d[7] = np.nan
    sm.vmult_add(dev(d), dev(src))
    with pytest.raises(t.TPSSError) as e:
        sm.finite_status()
    assert e.value.code == t.E_NONFINITE
    # finite data: nothing reported, and the results are unchanged
    dst = dev(np.zeros(N))
    sm.precond_vmult(dst, dev(src))
    sm.finite_status()
    assert rel_err(host(dst), ref.precond_vmult(src)) < 1e-12


def test_gmg_solve_reports_nonfinite_and_no_convergence():
    t = _tpss()
    g = t.GMGSolver(2, 2, 2, 3)
    b = o.splitmix_uniform(2, g.n_dofs)
    x = dev(np.zeros(g.n_dofs))
    with pytest.raises(t.TPSSError) as e:
        g.solve(x, dev(b), rel_tol=1e-14, max_iter=1)
    assert e.value.code == t.E_NOT_CONVERGED
    info, hist = g.solve(x, dev(b), rel_tol=1e-14, max_iter=1, raise_on_fail=False)
    assert len(hist) == 2 and hist[1] > 1e-14 * hist[0]
    bad = b.copy()
    bad[5] = np.nan
    with pytest.raises(t.TPSSError) as e:
        g.solve(x, dev(bad), rel_tol=1e-8, max_iter=10)
    assert e.value.code == t.E_NONFINITE


def test_collapsed_cell_is_singular():
    t = _tpss()
    V = o.uniform_vertices(2, 4).reshape(5, 5, 2).copy()
    V[:, 1, 0] = V[:, 2, 0]  # the column of cells between x-vertices 1 and 2 has zero width
    with pytest.raises(t.TPSSError) as e:
        t.Level(2, 2, 4, vertices=V.reshape(-1, 2))
    assert e.value.code == t.E_SINGULAR
    with pytest.raises(t.TPSSError) as e:
        t.GMGSolver(2, 2, 2, 2, vertices=V.reshape(-1, 2))
    assert e.value.code == t.E_SINGULAR
