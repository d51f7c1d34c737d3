"""Restricted additive Schwarz on the device vs the oracle (C ABI): restricted
flags bit-exact, the weighted and the restricted additive preconditioner in
FP64 within 1e-12 (FP32 patch solves within 2e-5), on Cartesian and
moved-vertex meshes."""
import numpy as np
import pytest

import oracle as o
from test_gpu_parity import _tpss, dev, host, rel_err

pytestmark = pytest.mark.gpu


def _make(dim, p, n, ras, relax=0.7, vertices=None, precision=0, local_precision=None,
          coloring=None):
    t = _tpss()
    coloring = t.COLOR_NONE if coloring is None else coloring
    lv = t.Level(dim, p, n, ip_factor=1.5, vertices=vertices, precision=precision)
    sm = t.SchwarzSmoother(lv, variant=t.ADDITIVE, coloring=coloring, relaxation=relax, ras=ras,
                           local_precision=local_precision)
    ref = o.Smoother(dim, p, n, ip=1.5, variant=o.ADDITIVE, coloring=coloring, relaxation=relax,
                     vertices=vertices, ras=ras)
    return lv, sm, ref


@pytest.mark.parametrize("dim,p,n", [(2, 1, 5), (2, 3, 6), (2, 4, 4), (3, 1, 4), (3, 2, 4),
                                     (3, 3, 4), (3, 4, 3)])
@pytest.mark.parametrize("ras", [1, 2])
def test_ras_parity(dim, p, n, ras):
    lv, sm, ref = _make(dim, p, n, ras)
    np.testing.assert_array_equal(sm.restricted_dof_flags(), ref.restricted_flags()
                                  if ras == 2 else _flags_of(ref, dim, p, n))
    src = o.splitmix_uniform(21, ref.N)
    d = dev(np.zeros(ref.N))
    sm.precond_vmult(d, dev(src))
    assert rel_err(host(d), ref.precond_vmult(src)) < 1e-12
    # Richardson step through the same preconditioner
    x0 = o.splitmix_uniform(22, ref.N)
    x = dev(x0)
    sm.step(x, dev(src))
    assert rel_err(host(x), ref.step(x0.copy(), src)) < 1e-12


def _flags_of(ref, dim, p, n):
    b = o.Smoother(dim, p, n, variant=o.ADDITIVE, ras=o.RAS_BOOLEAN)
    return b.restricted_flags()


def test_ras_tiled_colors_fp32_and_moved_vertices():
    t = _tpss()
    V = o.distorted_vertices(3, 4, 0.2, seed=4)
    for ras in (1, 2):
        lv, sm, ref = _make(3, 2, 4, ras, vertices=V, coloring=t.COLOR_TILED)
        src = o.splitmix_uniform(23, ref.N)
        d = dev(np.zeros(ref.N))
        sm.precond_vmult(d, dev(src))
        assert rel_err(host(d), ref.precond_vmult(src)) < 1e-12
        lv32, sm32, _ = _make(3, 2, 4, ras, vertices=V, precision=t.F32,
                              coloring=t.COLOR_TILED)
        d32 = dev(np.zeros(ref.N, dtype=np.float32))
        sm32.precond_vmult(d32, dev(src.astype(np.float32)))
        assert rel_err(host(d32), ref.precond_vmult(src)) < 2e-5


def test_ras_rejected_where_the_reference_rejects_it():
    t = _tpss()
    lv = t.Level(2, 2, 4)
    with pytest.raises(t.TPSSError) as e:
        t.SchwarzSmoother(lv, variant=t.MULTIPLICATIVE, ras=t.RAS_BOOLEAN)
    assert e.value.code == t.E_INVALID_ARG
    le = t.Level(2, 2, 4, elasticity=(1.0, 1.0))
    with pytest.raises(t.TPSSError) as e:
        t.SchwarzSmoother(le, variant=t.ADDITIVE, ras=t.RAS_WEIGHTS)
    assert e.value.code == t.E_UNSUPPORTED
