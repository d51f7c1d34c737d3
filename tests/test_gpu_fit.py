"""Fitted local solvers (separable Kronecker fit of the exact local matrices,
alternating least squares from the surrogate-length matrices) on the device vs
the oracle through the C ABI.  FP64 setup: fit errors within 1e-9 relative of
the oracle's, patch solves within 1e-10."""
import numpy as np
import pytest

import oracle as o
from test_gpu_parity import _tpss, dev, host, rel_err

pytestmark = pytest.mark.gpu


def make_fit(dim, p, n, variant, amp, seed=3, sweeps=4, ip=4.0, precision=0):
    t = _tpss()
    V = None if amp == 0 else o.distorted_vertices(dim, n, amp, seed=seed)
    coloring = t.COLOR_NONE if variant == t.ADDITIVE else t.COLOR_REDBLACK
    relax = 0.7 if variant == t.ADDITIVE else 1.0
    lv = t.Level(dim, p, n, ip_factor=ip, vertices=V, precision=precision)
    sm = t.SchwarzSmoother(lv, variant=variant, coloring=coloring, relaxation=relax,
                           local_solver=t.LOCAL_FIT, fit_sweeps=sweeps)
    ref = o.Smoother(dim, p, n, ip=ip, vertices=V, variant=variant, coloring=coloring,
                     relaxation=relax, local_solver=o.LOCAL_FIT, fit_sweeps=sweeps, threads=8)
    return lv, sm, ref


@pytest.mark.parametrize("n,variant", [(3, 0), (3, 1), (4, 0)])
def test_fit_parity_q4_vertices_moved_by_0_15h(n, variant):
    """The fitted local solver of the 48^3 Q4 workload: 3D DG Q4 (k = 10, 1000 DoFs per
    patch), interior vertices moved by 0.15 h, rank-2 separable fit of the exact local
    matrices; the
    additive smoother with damping 0.7 and the multiplicative one, FP64 and FP32 patch
    solves (the FP32 smoother keeps the FP64 setup)."""
    t = _tpss()
    lv, sm, ref = make_fit(3, 4, n, variant, 0.15)
    assert ref.fit_fallbacks() == 0
    np.testing.assert_allclose(sm.fit_errors(), ref.fit_errors(), rtol=1e-9, atol=1e-14)
    src = o.splitmix_uniform(3, ref.N)
    d = dev(np.zeros(ref.N))
    sm.precond_vmult(d, dev(src))
    expect = ref.precond_vmult(src)
    assert rel_err(host(d), expect) < 1e-10
    lv32, sm32, _ = make_fit(3, 4, n, variant, 0.15, precision=t.F32)
    d32 = dev(np.zeros(ref.N, dtype=np.float32))
    sm32.precond_vmult(d32, dev(src.astype(np.float32)))
    assert rel_err(host(d32), expect) < 2e-5


@pytest.mark.parametrize("dim,p,n,variant,amp", [(3, 1, 4, 0, 0.2), (3, 2, 3, 0, 0.15),
                                                 (2, 3, 5, 1, 0.2), (3, 1, 5, 1, 0.2)])
def test_fit_parity(dim, p, n, variant, amp):
    lv, sm, ref = make_fit(dim, p, n, variant, amp)
    assert ref.fit_fallbacks() == 0
    np.testing.assert_allclose(sm.fit_errors(), ref.fit_errors(), rtol=1e-9, atol=1e-14)
    src = o.splitmix_uniform(3, ref.N)
    d = dev(np.zeros(ref.N))
    sm.precond_vmult(d, dev(src))
    assert rel_err(host(d), ref.precond_vmult(src)) < 1e-10


def test_fit_is_exact_on_cartesian_patches():
    t = _tpss()
    lv, sm, ref = make_fit(3, 2, 4, t.ADDITIVE, 0.0)
    assert np.abs(sm.fit_errors()).max() < 1e-12
    sur = t.SchwarzSmoother(lv, variant=t.ADDITIVE, relaxation=0.7)
    src = o.splitmix_uniform(4, ref.N)
    a, b = dev(np.zeros(ref.N)), dev(np.zeros(ref.N))
    sm.precond_vmult(a, dev(src))
    sur.precond_vmult(b, dev(src))
    assert rel_err(host(a), host(b)) < 1e-11


def test_fit_gmg_converges():
    """3D DG Q2, 16^3 cells, interior vertices moved by 0.15 h, multiplicative
    vertex patches: GMG-CG with fitted local solvers reaches 1e-8.  (With
    fine-level perturbations the level meshes are not nested, so the DG
    embedding is only approximate; the additive smoother with the reference's
    damping 0.15 stalls on such hierarchies for either local solver.)"""
    t = _tpss()
    n = 16
    V = o.distorted_vertices(3, n, 0.15, seed=3)
    res = {}
    for name, ls in (("surrogate", t.LOCAL_SURROGATE), ("fit", t.LOCAL_FIT)):
        g = t.GMGSolver(3, 2, 2, 4, ip_factor=4.0, vertices=V, variant=t.MULTIPLICATIVE,
                        relaxation=1.0, local_solver=ls)
        b = o.splitmix_uniform(2, g.n_dofs)
        x = dev(np.zeros(g.n_dofs))
        info, hist = g.solve(x, dev(b), rel_tol=1e-8, max_iter=200, raise_on_fail=False)
        res[name] = (info.iterations, info.reduction_rate, info.final_residual / info.initial_residual)
        if ls == t.LOCAL_FIT:
            err = g.smoother(3).fit_errors()
            res["fit_err"] = (float(err.max()), int((err < 0).sum()))
    assert res["fit"][0] > 0 and res["fit"][2] <= 1e-8, res


def test_nested_distorted_gmg_matches_oracle():
    """The reference's distorted meshes (root mesh distorted, then refined;
    apps/mesh.h:153-204) give nested GMG levels: one AVP V-cycle and the MVP
    GMG-CG on 3D DG Q2, root 2^3 cells distorted by 0.15 h_root, 8^3 fine cells."""
    t = _tpss()
    from tpss_amd import meshes
    V = meshes.nested_distorted_vertices(3, 2, 8, 0.15, seed=3)
    b = o.splitmix_uniform(2, 8 ** 3 * 27)
    g = t.GMGSolver(3, 2, 2, 3, ip_factor=4.0, vertices=V, variant=t.ADDITIVE,
                    relaxation=0.15)
    ref = o.GMG(3, 2, 2, 3, ip=4.0, variant=t.ADDITIVE, coloring=t.COLOR_NONE,
                relaxation=0.15, vertices=V)
    z = dev(np.zeros(g.n_dofs))
    g.vcycle(z, dev(b))
    assert rel_err(host(z), ref.vcycle(b)) < 1e-11
    g = t.GMGSolver(3, 2, 2, 3, ip_factor=4.0, vertices=V)
    ref = o.GMG(3, 2, 2, 3, ip=4.0, vertices=V)
    x = dev(np.zeros(g.n_dofs))
    info, hist = g.solve(x, dev(b), rel_tol=1e-8)
    xr, itr, histr = ref.cg(b, rel_tol=1e-8)
    assert info.iterations == itr
    assert rel_err(host(x), xr) < 1e-9


def test_fit_improves_additive_smoother_on_nested_meshes():
    """On a nested distorted hierarchy (3D DG Q3, root 3^3 cells distorted by
    0.15 h_root, 12^3 fine cells) the additive vertex-patch smoother with the
    reference's 3D damping 0.15 converges, and the fitted local solvers need
    fewer GMG-CG iterations than the surrogate-length ones."""
    t = _tpss()
    from tpss_amd import meshes
    V = meshes.nested_distorted_vertices(3, 3, 12, 0.15, seed=3)
    res = {}
    for name, ls in (("surrogate", t.LOCAL_SURROGATE), ("fit", t.LOCAL_FIT)):
        g = t.GMGSolver(3, 3, 3, 3, ip_factor=4.0, vertices=V, variant=t.ADDITIVE,
                        relaxation=0.15, local_solver=ls)
        b = o.splitmix_uniform(2, g.n_dofs)
        x = dev(np.zeros(g.n_dofs))
        info, hist = g.solve(x, dev(b), rel_tol=1e-8, max_iter=200, raise_on_fail=False)
        res[name] = (info.iterations, info.reduction_rate)
    assert res["surrogate"][0] > 0 and res["fit"][0] > 0, res
    assert res["fit"][0] < res["surrogate"][0], res
