"""HIP path vs the CPU oracle through the C ABI (needs an MI355X).

Bar: FP64 results within 1e-12 relative (max-norm) of the oracle, index
tables bit-exact; FP32 patch solves within 1e-5 relative."""
import numpy as np
import pytest

import oracle as o

pytestmark = pytest.mark.gpu


def _tpss():
    import tpss_amd
    try:
        n = tpss_amd.device_count()
    except tpss_amd.TPSSError:
        n = 0
    if n < 1:
        pytest.skip("no HIP device")
    return tpss_amd


def dev(a, precision=None):
    t = _tpss()
    a = np.ascontiguousarray(a)
    if precision is None:
        precision = t.F32 if a.dtype == np.float32 else t.F64
    return t.DeviceVector.from_numpy(a, precision)


def host(v):
    _tpss().synchronize()
    return v.numpy().astype(np.float64)


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def make(dim, p, n, variant, coloring=None, relaxation=None, reverse=False, symmetrized=False,
         damping=1.0, n_steps=1, ip=1.0, precision=0, local_precision=None, vertices=None):
    import tpss_amd as t
    if coloring is None:
        coloring = t.COLOR_NONE if variant == t.ADDITIVE else t.COLOR_REDBLACK
    if relaxation is None:
        relaxation = t.DEFAULT_DAMPING[(variant, dim)]
    lv = t.Level(dim, p, n, ip_factor=ip, precision=precision, vertices=vertices)
    sm = t.SchwarzSmoother(lv, variant=variant, coloring=coloring, relaxation=relaxation,
                           reverse=reverse, symmetrized=symmetrized, damping_factor=damping,
                           n_steps=n_steps, local_precision=local_precision)
    ref = o.Smoother(dim, p, n, ip=ip, variant=variant, coloring=coloring, relaxation=relaxation,
                     reverse=reverse, symmetrized=symmetrized, damping=damping, n_steps=n_steps,
                     vertices=vertices)
    return lv, sm, ref


OPERATOR_CASES = [(2, 1, 5), (2, 3, 8), (2, 4, 6), (3, 1, 4), (3, 2, 5), (3, 3, 6), (3, 4, 3)]


@pytest.mark.parametrize("dim,p,n", OPERATOR_CASES)
def test_operator_parity(dim, p, n):
    t = _tpss()
    lv = t.Level(dim, p, n, ip_factor=2.5)
    ref = o.Smoother(dim, p, n, ip=2.5)
    x = o.splitmix_uniform(1, ref.N)
    y = dev(np.zeros_like(x))
    lv.vmult(y, dev(x))
    assert rel_err(host(y), ref.operator(x)) < 1e-13
    b = o.splitmix_uniform(2, ref.N)
    lv.residual(y, dev(b), dev(x))
    assert rel_err(host(y), b - ref.operator(x)) < 1e-13


@pytest.mark.parametrize("dim,p,n", [(2, 3, 8), (3, 3, 5), (3, 2, 4), (2, 1, 6)])
def test_patch_index_tables_bit_exact(dim, p, n):
    import tpss_amd as t
    for variant in (t.ADDITIVE, t.MULTIPLICATIVE):
        lv, sm, ref = make(dim, p, n, variant)
        np.testing.assert_array_equal(sm.color_ptr(), ref.color_ptr())
        np.testing.assert_array_equal(sm.patch_vertices(), ref.patch_vertices())
        np.testing.assert_array_equal(sm.patch_dof_indices(), ref.patch_dof_indices())


def test_c1_avp_fp64_one_step():
    """2D SIPG DG Q3 on 128x128, one AVP smoothing step (omega 0.25) on U[-1,1] (seed 1)."""
    import tpss_amd as t
    lv, sm, ref = make(2, 3, 128, t.ADDITIVE)
    src = o.splitmix_uniform(1, ref.N)
    d = dev(np.zeros(ref.N))
    sm.precond_vmult(d, dev(src))
    assert rel_err(host(d), ref.precond_vmult(src)) < 1e-12
    # smoothing step from zero = smoother vmult
    sm.vmult(d, dev(src))
    assert rel_err(host(d), ref.smoother_vmult(src)) < 1e-12


@pytest.mark.parametrize("dim,p,n,variant,rev,sym", [
    (2, 3, 16, 0, False, False), (2, 2, 9, 1, False, False), (2, 3, 10, 1, True, False),
    (2, 1, 7, 1, False, True), (3, 3, 8, 0, False, False), (3, 3, 8, 1, False, False),
    (3, 2, 7, 1, True, False), (3, 4, 5, 0, False, False), (3, 1, 6, 1, False, True)])
def test_precond_parity(dim, p, n, variant, rev, sym):
    lv, sm, ref = make(dim, p, n, variant, reverse=rev, symmetrized=sym)
    src = o.splitmix_uniform(3, ref.N)
    d = dev(np.zeros(ref.N))
    sm.precond_vmult(d, dev(src))
    assert rel_err(host(d), ref.precond_vmult(src)) < 1e-12


@pytest.mark.parametrize("variant", [0, 1])
def test_richardson_step_parity(variant):
    lv, sm, ref = make(3, 2, 6, variant, damping=0.9, n_steps=2)
    b = o.splitmix_uniform(4, ref.N)
    x0 = o.splitmix_uniform(5, ref.N)
    d = dev(x0)
    sm.step(d, dev(b))
    assert rel_err(host(d), ref.step(x0.copy(), b)) < 1e-12


def test_tiled_coloring_avp_matches_single_color():
    import tpss_amd as t
    lv, sm, ref = make(3, 3, 6, t.ADDITIVE, coloring=t.COLOR_TILED)
    src = o.splitmix_uniform(6, ref.N)
    d = dev(np.zeros(ref.N))
    sm.precond_vmult(d, dev(src))
    ref1 = o.Smoother(3, 3, 6, variant=t.ADDITIVE, relaxation=0.15)
    assert rel_err(host(d), ref1.precond_vmult(src)) < 1e-12


@pytest.mark.parametrize("variant", [0, 1])
def test_fp32_patch_solves(variant):
    import tpss_amd as t
    lv, sm, ref = make(3, 3, 6, variant, precision=t.F32, local_precision=t.F32)
    src = o.splitmix_uniform(7, ref.N)
    d = dev(np.zeros(ref.N), t.F32)
    sm.precond_vmult(d, dev(src, t.F32))
    assert rel_err(host(d), ref.precond_vmult(src)) < 2e-5
    # FP32 solves on FP64 vectors
    lv64, sm64, _ = make(3, 3, 6, variant, local_precision=t.F32)
    d64 = dev(np.zeros(ref.N))
    sm64.precond_vmult(d64, dev(src))
    assert rel_err(host(d64), ref.precond_vmult(src)) < 2e-5


@pytest.mark.parametrize("dim,p,n,mult", [(2, 3, 12, False), (3, 3, 5, True)])
def test_per_patch_damping(dim, p, n, mult):
    import tpss_amd as t
    lv, sm, ref = make(dim, p, n, t.MULTIPLICATIVE if mult else t.ADDITIVE)
    omega = 0.2 + 0.5 * (o.splitmix_uniform(9, sm.n_patches) + 1.0) / 2.0
    omega = omega.astype(np.float32).astype(np.float64)  # stored as FP32 on the device
    sm.set_patch_damping(omega)
    ref.set_patch_damping(omega)
    src = o.splitmix_uniform(8, ref.N)
    d = dev(np.zeros(ref.N))
    sm.precond_vmult(d, dev(src))
    assert rel_err(host(d), ref.precond_vmult(src)) < 1e-12


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 3), (3, 1), (3, 2), (3, 3), (3, 4)])
def test_transfer_parity(dim, p):
    """Sum-factorised prolongation / restriction per coarse cell vs the oracle's
    dense 1D-product loops: FP64 within 1e-13, FP32 within 1e-6 (relative)."""
    import tpss_amd as t
    xc = o.splitmix_uniform(1, t.Level(dim, p, 3).n_dofs)
    xf = o.splitmix_uniform(2, t.Level(dim, p, 6).n_dofs)
    ref_p = xf.copy()
    o.prolongate_add(dim, p, 3, xc, ref_p)
    ref_r = o.restrict(dim, p, 3, xf)
    for prec, tol in ((t.F64, 1e-13), (t.F32, 1e-6)):
        lc, lf = t.Level(dim, p, 3, precision=prec), t.Level(dim, p, 6, precision=prec)
        d = dev(xf.copy(), prec)
        t.lib.tpss_transfer_prolongate_add(lc.handle, lf.handle, t._ptr(d), t._ptr(dev(xc, prec)),
                                           None)
        assert rel_err(host(d), ref_p) < tol
        dc = dev(np.zeros(lc.n_dofs), prec)
        t.lib.tpss_transfer_restrict(lc.handle, lf.handle, t._ptr(dc), t._ptr(dev(xf, prec)), None)
        assert rel_err(host(dc), ref_r) < tol


def test_random_fill_matches_oracle():
    t = _tpss()
    x = t.DeviceVector(10001)
    t.fill_random(x, seed=1)
    np.testing.assert_array_equal(host(x), o.splitmix_uniform(1, 10001))


@pytest.mark.parametrize("dim,p,levels,variant", [(2, 3, 4, 1), (2, 3, 4, 0), (3, 2, 3, 1),
                                                   (3, 3, 3, 0)])
def test_gmg_cg_matches_oracle(dim, p, levels, variant):
    import tpss_amd as t
    relax = t.DEFAULT_DAMPING[(variant, dim)]
    col = t.COLOR_REDBLACK if variant == 1 else t.COLOR_NONE
    g = t.GMGSolver(dim, p, 2, levels, variant=variant, coloring=col, relaxation=relax)
    ref = o.GMG(dim, p, 2, levels, variant=variant, coloring=col, relaxation=relax)
    b = o.splitmix_uniform(2, g.n_dofs)
    # one V-cycle (multiplicative: residual-update path first, then fused passes)
    z = dev(np.zeros(g.n_dofs))
    if variant == 1:
        for l in range(1, levels):
            for post in (False, True):
                g.smoother(l, post).set_fused_residual(False)
        g.vcycle(z, dev(b))
        assert rel_err(host(z), ref.vcycle(b)) < 1e-11, "residual-update path"
        for l in range(1, levels):
            for post in (False, True):
                g.smoother(l, post).set_fused_residual(True)
    g.vcycle(z, dev(b))
    zr = ref.vcycle(b)
    assert rel_err(host(z), zr) < 1e-11
    # in place (dst == src): the finest level then keeps its own vectors
    w = dev(b)
    g.vcycle(w, w)
    assert rel_err(host(w), zr) < 1e-11
    x = dev(np.zeros(g.n_dofs))
    info, hist = g.solve(x, dev(b), rel_tol=1e-8)
    xr, itr, histr = ref.cg(b, rel_tol=1e-8)
    assert info.iterations == itr
    assert rel_err(host(x), xr) < 1e-9
    np.testing.assert_allclose(hist, histr, rtol=1e-7)


def test_gmg_cg_at_48_cubed_matches_the_threaded_oracle():
    """3D Q3 on 48^3 cells (3^3 coarse, 5 levels): the outer CG's operator pass forms
    p.Ap in 6,912 block partials, enough for every branch of the wide final sum; the
    iterations, the residual history and the solution follow the oracle's FP64 GMG-CG
    (run on up to 16 host threads, bit for bit the serial oracle)."""
    import os
    import tpss_amd as t
    g = t.GMGSolver(3, 3, 3, 5, variant=t.MULTIPLICATIVE, coloring=t.COLOR_REDBLACK,
                    relaxation=1.0)
    b = o.splitmix_uniform(5, g.n_dofs)
    x = dev(np.zeros(g.n_dofs))
    info, hist = g.solve(x, dev(b), rel_tol=1e-8)
    ref = o.GMG(3, 3, 3, 5, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK, relaxation=1.0,
                threads=max(1, min(16, os.cpu_count() or 1)))
    xr, itr, histr = ref.cg(b, rel_tol=1e-8)
    assert info.iterations == itr
    np.testing.assert_allclose(hist, histr, rtol=1e-7)
    assert rel_err(host(x), xr) < 1e-9


def test_gmg_mixed_precision_reaches_tolerance():
    import tpss_amd as t
    g = t.GMGSolver(3, 3, 2, 3, mg_precision=t.F32)
    b = o.splitmix_uniform(2, g.n_dofs)
    x = dev(np.zeros(g.n_dofs))
    info, hist = g.solve(x, dev(b), rel_tol=1e-8)
    assert hist[-1] <= 1e-8 * hist[0]
    lv = o.Smoother(3, 3, 8)
    assert np.linalg.norm(b - lv.operator(host(x))) <= 1.01e-8 * np.linalg.norm(b)


# ---------------------------------------------------------------- distorted meshes

@pytest.mark.parametrize("dim,p,n", [(2, 2, 5), (2, 3, 4), (3, 1, 3), (3, 2, 3), (3, 3, 4),
                                     (3, 4, 3)])
def test_distorted_operator_parity(dim, p, n):
    t = _tpss()
    V = o.distorted_vertices(dim, n, 0.2, seed=3)
    lv = t.Level(dim, p, n, ip_factor=4.0, vertices=V)
    ref = o.Smoother(dim, p, n, ip=4.0, vertices=V)
    x = o.splitmix_uniform(1, ref.N)
    y = dev(np.zeros_like(x))
    lv.vmult(y, dev(x))
    assert rel_err(host(y), ref.operator(x)) < 1e-12
    b = o.splitmix_uniform(2, ref.N)
    lv.residual(y, dev(b), dev(x))
    assert rel_err(host(y), b - ref.operator(x)) < 1e-12


@pytest.mark.parametrize("dim,p,n", [(2, 3, 4), (3, 2, 3), (3, 3, 4), (3, 4, 3)])
def test_distorted_operator_parity_f32(dim, p, n):
    """FP32 levels evaluate the moved-vertex operator in FP32 arithmetic (the
    Jacobians too, from map coefficients formed in FP64): max-norm error within
    1e-5 of the FP64 oracle, relative, for the operator and the residual."""
    t = _tpss()
    V = o.distorted_vertices(dim, n, 0.2, seed=3)
    lv = t.Level(dim, p, n, ip_factor=4.0, vertices=V, precision=t.F32)
    ref = o.Smoother(dim, p, n, ip=4.0, vertices=V)
    x = o.splitmix_uniform(1, ref.N).astype(np.float32).astype(np.float64)
    y = dev(np.zeros(ref.N), t.F32)
    lv.vmult(y, dev(x, t.F32))
    assert rel_err(host(y), ref.operator(x)) < 1e-5
    b = o.splitmix_uniform(2, ref.N).astype(np.float32).astype(np.float64)
    lv.residual(y, dev(b, t.F32), dev(x, t.F32))
    assert rel_err(host(y), b - ref.operator(x)) < 1e-5


@pytest.mark.parametrize("dim,p,n,variant", [(2, 3, 6, 0), (3, 3, 4, 0), (3, 2, 5, 1), (3, 4, 3, 0)])
def test_distorted_smoother_parity(dim, p, n, variant):
    V = o.distorted_vertices(dim, n, 0.15, seed=3)
    lv, sm, ref = make(dim, p, n, variant, ip=4.0, vertices=V)
    np.testing.assert_allclose(sm.patch_h(), ref.patch_h(), rtol=1e-13, atol=0)
    src = o.splitmix_uniform(4, ref.N)
    d = dev(np.zeros(ref.N))
    sm.vmult(d, dev(src))
    assert rel_err(host(d), ref.smoother_vmult(src)) < 1e-12


def test_distorted_gmg_matches_oracle():
    import tpss_amd as t
    dim, p, levels = 3, 2, 3
    V = o.distorted_vertices(dim, 2 << (levels - 1), 0.2, seed=3)
    # AVP: one V-cycle (the surrogate AVP with the fixed reference damping does
    # not make a convergent CG preconditioner on this mesh, in the oracle too)
    g = t.GMGSolver(dim, p, 2, levels, ip_factor=4.0, vertices=V, variant=t.ADDITIVE,
                    relaxation=0.15)
    ref = o.GMG(dim, p, 2, levels, ip=4.0, variant=t.ADDITIVE, coloring=t.COLOR_NONE,
                relaxation=0.15, vertices=V)
    b = o.splitmix_uniform(2, g.n_dofs)
    z = dev(np.zeros(g.n_dofs))
    g.vcycle(z, dev(b))
    assert rel_err(host(z), ref.vcycle(b)) < 1e-11
    # MVP: full GMG-CG
    g = t.GMGSolver(dim, p, 2, levels, ip_factor=4.0, vertices=V)
    ref = o.GMG(dim, p, 2, levels, ip=4.0, vertices=V)
    x = dev(np.zeros(g.n_dofs))
    info, hist = g.solve(x, dev(b), rel_tol=1e-8)
    xr, itr, histr = ref.cg(b, rel_tol=1e-8)
    assert info.iterations == itr
    assert rel_err(host(x), xr) < 1e-9


@pytest.mark.parametrize("dim,p,n,relax,rev", [(2, 3, 9, 1.0, False), (3, 3, 7, 1.0, False),
                                               (3, 2, 6, 0.7, True), (2, 2, 8, 0.6, False),
                                               (2, 3, 4, 1.0, False), (2, 3, 8, 1.0, True),
                                               (2, 3, 16, 1.0, False), (3, 2, 4, 1.0, False),
                                               (3, 3, 6, 0.8, True), (3, 3, 2, 1.0, False),
                                               (3, 3, 3, 1.0, True)])
def test_mvp_fused_and_residual_paths(dim, p, n, relax, rev):
    """Residual-free color passes and the residual-update path both match the
    oracle (which recomputes the full residual before every color)."""
    import tpss_amd as t
    lv, sm, ref = make(dim, p, n, t.MULTIPLICATIVE, relaxation=relax, reverse=rev)
    src = o.splitmix_uniform(12, ref.N)
    expect = ref.precond_vmult(src)
    for fused in (True, False):
        sm.set_fused_residual(fused)
        d = dev(o.splitmix_uniform(13, ref.N))  # vmult overwrites whatever dst holds
        sm.precond_vmult(d, dev(src))
        assert rel_err(host(d), expect) < 1e-12, fused
        # twice in a row (scratch buffers keep the previous call's contents)
        sm.precond_vmult(d, dev(src))
        assert rel_err(host(d), expect) < 1e-12, fused


@pytest.mark.parametrize("relax,rev,prec", [(1.0, False, 0), (0.8, True, 0), (1.0, False, 1)])
def test_fine_level_color_passes_against_the_oracle(relax, rev, prec):
    """3D Q3 on 23^3 cells: every color holds 665 or 666 patches, so the step runs the
    one-wave color pass (the batched FP64 form, the FP32 form) in the XCD launch
    order, as on the 64^3 level (smaller levels run the four-wave pass); against the
    oracle's step and sweep, which recompute the full residual before every color."""
    t = _tpss()
    n = 23
    lv = t.Level(3, 3, n, precision=prec)
    sm = t.SchwarzSmoother(lv, variant=t.MULTIPLICATIVE, relaxation=relax, reverse=rev,
                           local_precision=prec)
    ref = o.Smoother(3, 3, n, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK,
                     relaxation=relax, reverse=rev, threads=8)
    assert np.diff(ref.color_ptr()).min() > 512
    b = o.splitmix_uniform(21, ref.N)
    x0 = o.splitmix_uniform(22, ref.N)
    tol = 1e-12 if prec == t.F64 else 2e-5
    x = dev(x0, prec)
    sm.step(x, dev(b, prec))
    assert rel_err(host(x), ref.step(x0.copy(), b)) < tol
    d = dev(np.zeros(ref.N), prec)
    sm.precond_vmult(d, dev(b, prec))
    assert rel_err(host(d), ref.precond_vmult(b)) < tol


@pytest.mark.parametrize("vec_prec", [0, 1])
def test_bf16_mfma_patch_solves(vec_prec):
    """BF16 products with FP32 accumulation in the FD contractions: one MVP
    sweep within 3e-2 (max norm) of the FP64 oracle."""
    import tpss_amd as t
    lv, sm, ref = make(3, 3, 6, t.MULTIPLICATIVE, precision=vec_prec, local_precision=t.BF16)
    src = o.splitmix_uniform(7, ref.N)
    d = dev(np.zeros(ref.N), vec_prec)
    sm.precond_vmult(d, dev(src, vec_prec))
    err = rel_err(host(d), ref.precond_vmult(src))
    assert err < 3e-2, err
    assert err > 1e-7  # really computed in reduced precision


def test_gmg_bf16_reaches_tolerance():
    import tpss_amd as t
    g = t.GMGSolver(3, 3, 2, 4, mg_precision=t.F32, local_precision=t.BF16)
    b = o.splitmix_uniform(2, g.n_dofs)
    x = dev(np.zeros(g.n_dofs))
    info, hist = g.solve(x, dev(b), rel_tol=1e-8)
    assert hist[-1] <= 1e-8 * hist[0]
    assert info.iterations <= 12


@pytest.mark.parametrize("dim,p,n,rev,n_steps", [(2, 3, 9, False, 1), (3, 2, 6, True, 2),
                                                 (3, 3, 5, False, 1), (3, 3, 24, True, 1)])
def test_residual_free_richardson_step(dim, p, n, rev, n_steps):
    """With omega_R = 1 the multiplicative step runs as color passes on (x, b)
    without a residual update; it must match the reference's residual + sweep.
    3D Q3: colors of up to 512 patches (n = 5) run the four-wave pass, larger
    ones (n = 24: about 760 patches per color) one wave per patch."""
    import tpss_amd as t
    lv, sm, ref = make(dim, p, n, t.MULTIPLICATIVE, reverse=rev, n_steps=n_steps)
    b = o.splitmix_uniform(4, ref.N)
    x0 = o.splitmix_uniform(5, ref.N)
    expect = ref.step(x0.copy(), b)
    d = dev(x0)
    sm.step(d, dev(b))
    assert rel_err(host(d), expect) < 1e-12
    sm.set_fused_residual(False)
    d = dev(x0)
    sm.step(d, dev(b))
    assert rel_err(host(d), expect) < 1e-12
    sm.set_fused_residual(True)
    sm.vmult(d, dev(b))  # from zero
    assert rel_err(host(d), ref.smoother_vmult(b)) < 1e-12


def test_gmg_iterative_coarse_solve():
    """Coarse level above the dense-inverse size (3D Q2 on 5^3 coarse cells,
    3375 DoFs): the coarse CG runs inside the outer CG's preconditioner."""
    import tpss_amd as t
    g = t.GMGSolver(3, 2, 5, 2, variant=t.MULTIPLICATIVE, coloring=t.COLOR_REDBLACK)
    ref = o.GMG(3, 2, 5, 2, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK)
    b = o.splitmix_uniform(2, g.n_dofs)
    z = dev(np.zeros(g.n_dofs))
    g.vcycle(z, dev(b))
    assert rel_err(host(z), ref.vcycle(b)) < 1e-9
    x = dev(np.zeros(g.n_dofs))
    info, hist = g.solve(x, dev(b), rel_tol=1e-8)
    xr, itr, histr = ref.cg(b, rel_tol=1e-8)
    assert info.iterations == itr
    assert rel_err(host(x), xr) < 1e-7
