"""Linear elasticity on the device vs the CPU oracle through the C ABI: the SIPG
elasticity operator, the block-diagonal vertex-patch smoother (additive and
multiplicative) and GMG-CG.  Bar: FP64 within 1e-12 relative of the oracle,
index tables bit-exact, FP32 patch solves within 2e-5."""
import numpy as np
import pytest

import oracle as o
from test_gpu_parity import _tpss, dev, host, rel_err
from test_oracle_elasticity import mass_load

pytestmark = pytest.mark.gpu

MU, LAM = 1.0, 10.0  # Lame coefficients of the 3D elasticity workload


def make(dim, p, n, variant, ip=10.0, precision=0, local_precision=None, relaxation=None,
         reverse=False, damping=1.0, n_steps=1):
    t = _tpss()
    coloring = t.COLOR_NONE if variant == t.ADDITIVE else t.COLOR_REDBLACK
    if relaxation is None:
        relaxation = t.DEFAULT_DAMPING[(variant, dim)]
    lv = t.Level(dim, p, n, ip_factor=ip, precision=precision, elasticity=(MU, LAM))
    sm = t.SchwarzSmoother(lv, variant=variant, coloring=coloring, relaxation=relaxation,
                           reverse=reverse, damping_factor=damping, n_steps=n_steps,
                           local_precision=local_precision)
    ref = o.Smoother(dim, p, n, ip=ip, variant=variant, coloring=coloring, relaxation=relaxation,
                     reverse=reverse, damping=damping, n_steps=n_steps, elasticity=(MU, LAM))
    return lv, sm, ref


@pytest.mark.parametrize("dim,p,n", [(2, 1, 5), (2, 3, 6), (2, 4, 4), (3, 1, 4), (3, 2, 5),
                                     (3, 3, 3), (3, 4, 3)])
def test_elasticity_operator_parity(dim, p, n):
    t = _tpss()
    lv = t.Level(dim, p, n, ip_factor=10.0, elasticity=(MU, LAM))
    ref = o.Smoother(dim, p, n, ip=10.0, elasticity=(MU, LAM))
    assert lv.n_dofs == ref.N
    x = o.splitmix_uniform(1, ref.N)
    y = dev(np.zeros_like(x))
    lv.vmult(y, dev(x))
    assert rel_err(host(y), ref.operator(x)) < 1e-12
    b = o.splitmix_uniform(2, ref.N)
    lv.residual(y, dev(b), dev(x))
    assert rel_err(host(y), b - ref.operator(x)) < 1e-12
    # FP32 vectors and FP32 arithmetic: within 1e-6 of the FP64 oracle (max norm, relative)
    lv32 = t.Level(dim, p, n, ip_factor=10.0, elasticity=(MU, LAM), precision=t.F32)
    y32 = dev(np.zeros(ref.N, np.float32))
    lv32.vmult(y32, dev(x.astype(np.float32)))
    assert rel_err(host(y32), ref.operator(x.astype(np.float32).astype(np.float64))) < 1e-6


@pytest.mark.parametrize("dim,p,n", [(2, 2, 6), (3, 2, 4)])
def test_elasticity_index_tables_bit_exact(dim, p, n):
    t = _tpss()
    for variant in (t.ADDITIVE, t.MULTIPLICATIVE):
        lv, sm, ref = make(dim, p, n, variant)
        assert sm.n_dofs_per_patch == dim * (2 * (p + 1)) ** dim
        np.testing.assert_array_equal(sm.color_ptr(), ref.color_ptr())
        np.testing.assert_array_equal(sm.patch_dof_indices(),
                                      ref.patch_dof_indices().reshape(sm.n_patches, -1))


@pytest.mark.parametrize("dim,p,n,variant,rev", [(2, 3, 8, 0, False), (2, 2, 7, 1, False),
                                                 (2, 2, 7, 1, True), (3, 2, 5, 0, False),
                                                 (3, 2, 5, 1, False), (3, 1, 6, 1, True)])
def test_elasticity_precond_parity(dim, p, n, variant, rev):
    lv, sm, ref = make(dim, p, n, variant, reverse=rev)
    src = o.splitmix_uniform(3, ref.N)
    d = dev(np.zeros(ref.N))
    sm.precond_vmult(d, dev(src))
    assert rel_err(host(d), ref.precond_vmult(src)) < 1e-12


@pytest.mark.parametrize("variant", [0, 1])
def test_elasticity_richardson_step_parity(variant):
    lv, sm, ref = make(3, 2, 5, variant, damping=0.9, n_steps=2)
    b = o.splitmix_uniform(4, ref.N)
    x0 = o.splitmix_uniform(5, ref.N)
    d = dev(x0)
    sm.step(d, dev(b))
    assert rel_err(host(d), ref.step(x0.copy(), b)) < 1e-12


@pytest.mark.parametrize("variant", [0, 1])
def test_elasticity_fp32_patch_solves(variant):
    t = _tpss()
    lv, sm, ref = make(3, 2, 5, variant, precision=t.F32, local_precision=t.F32)
    src = o.splitmix_uniform(7, ref.N)
    d = dev(np.zeros(ref.N, np.float32))
    sm.precond_vmult(d, dev(src.astype(np.float32)))
    assert rel_err(host(d), ref.precond_vmult(src)) < 2e-5


@pytest.mark.parametrize("vec_prec,rev", [(1, False), (0, True)])
def test_elasticity_bf16_mfma_patch_solves(vec_prec, rev):
    """3D Q2 elasticity, block-diagonal multiplicative smoother with BF16-MFMA patch
    solves (each 6^3 component block zero-padded into the 8^3 layout of the matrix-core
    solve, FP32 accumulation): one sweep within 3e-2 (max norm) of the FP64 oracle, and
    really computed in reduced precision."""
    t = _tpss()
    lv, sm, ref = make(3, 2, 5, t.MULTIPLICATIVE, precision=vec_prec, local_precision=t.BF16,
                       reverse=rev)
    src = o.splitmix_uniform(7, ref.N)
    d = dev(np.zeros(ref.N), vec_prec)
    sm.precond_vmult(d, dev(src, vec_prec))
    err = rel_err(host(d), ref.precond_vmult(src))
    assert 1e-7 < err < 3e-2, err
    # the Richardson step runs the same sweep on the residual
    x0 = o.splitmix_uniform(8, ref.N)
    x = dev(x0, vec_prec)
    sm.step(x, dev(src, vec_prec))
    err = rel_err(host(x), ref.step(x0.copy(), src))
    assert err < 3e-2, err


def test_elasticity_bf16_outside_its_scope_is_rejected():
    """BF16-MFMA solves of elasticity blocks: 3D Q2, multiplicative, red-black only."""
    t = _tpss()
    lv = t.Level(3, 2, 4, ip_factor=10.0, elasticity=(MU, LAM), precision=t.F32)
    with pytest.raises(t.TPSSError):
        t.SchwarzSmoother(lv, variant=t.ADDITIVE, local_precision=t.BF16)
    lv3 = t.Level(3, 3, 3, ip_factor=10.0, elasticity=(MU, LAM), precision=t.F32)
    with pytest.raises(t.TPSSError):
        t.SchwarzSmoother(lv3, variant=t.MULTIPLICATIVE, local_precision=t.BF16)


def test_elasticity_bf16_gmg_reaches_tolerance():
    """The BF16-MFMA smoother inside the FP32 V-cycle of the FP64 CG (3D Q2, 16^3)."""
    t = _tpss()
    g = t.GMGSolver(3, 2, 2, 4, ip_factor=10.0, mg_precision=t.F32, local_precision=t.BF16,
                    elasticity=(MU, LAM))
    b = o.splitmix_uniform(2, g.n_dofs)
    x = dev(np.zeros(g.n_dofs))
    info, hist = g.solve(x, dev(b), rel_tol=1e-8, max_iter=40)
    assert hist[-1] <= 1e-8 * hist[0] and info.iterations <= 20, (info.iterations, hist[-1] / hist[0])


@pytest.mark.parametrize("dim,p,levels,variant", [(2, 2, 3, 1), (2, 3, 3, 0), (3, 2, 2, 1)])
def test_elasticity_gmg_matches_oracle(dim, p, levels, variant):
    t = _tpss()
    relax = t.DEFAULT_DAMPING[(variant, dim)]
    col = t.COLOR_REDBLACK if variant == 1 else t.COLOR_NONE
    g = t.GMGSolver(dim, p, 2, levels, ip_factor=10.0, variant=variant, coloring=col,
                    relaxation=relax, elasticity=(MU, LAM))
    ref = o.GMG(dim, p, 2, levels, ip=10.0, variant=variant, coloring=col, relaxation=relax,
                elasticity=(MU, LAM))
    assert g.n_dofs == ref.N
    b = o.splitmix_uniform(2, g.n_dofs)
    z = dev(np.zeros(g.n_dofs))
    g.vcycle(z, dev(b))
    assert rel_err(host(z), ref.vcycle(b)) < 1e-11
    x = dev(np.zeros(g.n_dofs))
    info, hist = g.solve(x, dev(b), rel_tol=1e-8)
    xr, itr, histr = ref.cg(b, rel_tol=1e-8)
    assert info.iterations == itr
    assert rel_err(host(x), xr) < 1e-9


def test_elasticity_gmg_reference_iteration_bounds():
    """tests/apps_linelasticity.cc:50-105 (2D, mu = lambda = ip = 1, block-diagonal
    vertex patches): MVP < 7 (Q3), AVP < 29 (Q3), fractional iterations to 1e-8."""
    t = _tpss()
    for variant, bound in ((t.MULTIPLICATIVE, 7.0), (t.ADDITIVE, 29.0)):
        col = t.COLOR_REDBLACK if variant == 1 else t.COLOR_NONE
        for levels in (4, 6):
            g = t.GMGSolver(2, 3, 2, levels, ip_factor=1.0, variant=variant, coloring=col,
                            relaxation=1.0 if variant == 1 else 0.25, elasticity=(1.0, 1.0))
            b = mass_load(2, 3, 2 << (levels - 1), 2)
            x = dev(np.zeros(g.n_dofs))
            info, hist = g.solve(x, dev(b), rel_tol=1e-8)
            assert info.n_frac < bound, (variant, levels, info.n_frac)


def test_elasticity_mixed_precision_gmg_reaches_tolerance():
    """3D DG Q2 elasticity (mu 1, lambda 10, ip 10): FP32 V-cycle inside FP64 CG."""
    t = _tpss()
    g = t.GMGSolver(3, 2, 2, 4, ip_factor=10.0, mg_precision=t.F32, elasticity=(MU, LAM))
    b = o.splitmix_uniform(2, g.n_dofs)
    x = dev(np.zeros(g.n_dofs))
    info, hist = g.solve(x, dev(b), rel_tol=1e-8)
    assert hist[-1] <= 1e-8 * hist[0]
    ref = o.Smoother(3, 2, 16, ip=10.0, elasticity=(MU, LAM))
    assert np.linalg.norm(b - ref.operator(host(x))) <= 1.01e-8 * np.linalg.norm(b)
