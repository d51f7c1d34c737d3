"""Learned per-patch damping on the device vs the FP64 numpy oracle (needs an MI355X).

Features and damping targets are FP32 outputs of FP64 computations (rel. 1e-6);
the network computes in FP32 with BF16 weights: forward and gradient are compared
with the oracle evaluated on the same BF16-rounded weights (rel. 1e-4)."""
import numpy as np
import pytest

import mlp_oracle as mo
import oracle as o

pytestmark = pytest.mark.gpu


def _tpss():
    import tpss_amd
    if tpss_amd.device_count() < 1:
        pytest.skip("no HIP device")
    return tpss_amd


@pytest.mark.parametrize("dim,n", [(3, 4), (2, 6)])
def test_patch_features(dim, n):
    t = _tpss()
    V = o.distorted_vertices(dim, n, 0.2, seed=3)
    lv = t.Level(dim, 2, n, vertices=V)
    sm = t.SchwarzSmoother(lv, variant=t.MULTIPLICATIVE)
    f = sm.patch_features()
    ref = mo.patch_features(dim, n, V, sm.patch_vertices())
    np.testing.assert_allclose(f, ref, rtol=2e-6)


@pytest.mark.parametrize("distorted", [False, True])
def test_optimal_damping(distorted):
    t = _tpss()
    dim, p, n = 3, 2, 4
    V = o.distorted_vertices(dim, n, 0.2, seed=3) if distorted else None
    lv = t.Level(dim, p, n, ip_factor=4.0, vertices=V)
    sm = t.SchwarzSmoother(lv, variant=t.MULTIPLICATIVE)
    w = sm.optimal_damping(n_samples=2, seed=5, omega_min=0.0, omega_max=10.0)
    ref_sm = o.Smoother(dim, p, n, ip=4.0, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK,
                        vertices=V)
    ref = mo.optimal_damping(ref_sm, 2, 5, lo=0.0, hi=10.0)
    np.testing.assert_allclose(w, ref, rtol=2e-6)
    if not distorted:
        np.testing.assert_allclose(w, 1.0, rtol=1e-6)


def test_mlp_forward_and_gradient():
    t = _tpss()
    mlp = t.DampingMLP(24, 64, 0.2, 1.2, seed=4)
    params = mlp.get_params().astype(np.float64)
    pb = mo.bf16_round(params)
    rng = np.random.default_rng(2)
    X = rng.uniform(0.6, 1.4, (300, 24)).astype(np.float32)
    target = rng.uniform(0.3, 1.1, 300).astype(np.float32)
    w = mlp.forward(X)
    ref, _ = mo.forward(pb, X.astype(np.float64), 24, 64)
    np.testing.assert_allclose(w, ref, rtol=1e-4, atol=1e-5)
    loss, g = mlp.loss_grad(X, target)
    gref = mo.grad(pb, X.astype(np.float64), target.astype(np.float64), 24, 64)
    assert abs(loss - mo.loss(pb, X.astype(np.float64), target.astype(np.float64), 24, 64)) < 1e-4 * loss
    np.testing.assert_allclose(g, gref, rtol=1e-3, atol=1e-4 * np.abs(gref).max())


def test_training_reduces_loss_and_sets_damping():
    t = _tpss()
    dim, p, n = 3, 3, 8
    V = o.distorted_vertices(dim, n, 0.2, seed=11)
    lv = t.Level(dim, p, n, ip_factor=4.0, vertices=V)
    sm = t.SchwarzSmoother(lv, variant=t.MULTIPLICATIVE)
    X = sm.patch_features()
    y = sm.optimal_damping(n_samples=2, seed=5)
    mlp = t.DampingMLP(24, 64, 0.2, 1.2, seed=4)
    l0, _ = mlp.loss_grad(X, y)
    l1 = mlp.train(X, y, epochs=200, batch=128, lr=3e-3)
    assert l1 < 0.5 * l0, (l0, l1)
    sm.set_damping_mlp(mlp)
    src = o.splitmix_uniform(3, lv.n_dofs)
    d = t.DeviceVector.zeros(lv.n_dofs)
    sm.precond_vmult(d, t.DeviceVector.from_numpy(src))
    ref = o.Smoother(dim, p, n, ip=4.0, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK,
                     vertices=V)
    omega = mlp.forward(X).astype(np.float64)
    ref.set_patch_damping(omega)
    out = d.numpy()
    expect = ref.precond_vmult(src)
    assert np.abs(out - expect).max() < 1e-12 * np.abs(expect).max()


@pytest.mark.parametrize("calibrated,n", [(False, 16), (True, 32)])
def test_shipped_model_beats_fixed_damping(calibrated, n):
    """The trained network (tpss_amd/models/damping_dgq3_3d.npy), with its own
    prediction (16^3 cells) and with the output scale calibrated on a 64^3
    training mesh (32^3 cells), against the reference's fixed multiplicative
    damping 1 (TPSS.h:120-163): GMG-CG on a distorted DG Q3 mesh (interior
    vertices moved by 0.2 h, seed 3) contracts faster and needs fewer iterations
    to 1e-8."""
    t = _tpss()
    from tpss_amd import learned_damping as ld
    mlp = ld.load(scale=ld.OUTPUT_SCALE if calibrated else 1.0)
    V = ld.distorted_vertices(3, n, 0.2, seed=3)
    levels = int(np.log2(n // 2)) + 1
    res = {}
    for use in (False, True):
        g = t.GMGSolver(3, 3, 2, levels, ip_factor=4.0, vertices=V, variant=t.MULTIPLICATIVE,
                        relaxation=1.0)
        if use:
            ld.apply_to_gmg(g, mlp)
        b = o.splitmix_uniform(2, g.n_dofs)
        x = t.DeviceVector.zeros(g.n_dofs)
        info, hist = g.solve(x, t.DeviceVector.from_numpy(b), rel_tol=1e-8, max_iter=100)
        res[use] = (info.iterations, info.reduction_rate)
    assert res[True][1] < 0.8 * res[False][1] and res[True][0] < res[False][0], res
