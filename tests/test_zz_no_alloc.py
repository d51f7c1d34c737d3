"""The allocation-free preconditioner entry point (tpss_precond_vmult_noalloc)
against tpss_precond_vmult through the C ABI: bit-identical outputs on two calls
in a row, the second on a smaller level, for the additive variant (slot buffer
in the scratch array) and the multiplicative one with localized residuals
(residual in the scratch array); oversized smoothers are rejected."""
import ctypes

import numpy as np
import pytest

import oracle as o
from test_gpu_parity import _tpss, dev, host

pytestmark = pytest.mark.gpu


def _noalloc(t, sm, dst, src):
    return t.lib.tpss_precond_vmult_noalloc(sm.handle, ctypes.c_void_p(dst.data_ptr()),
                                           ctypes.c_void_p(src.data_ptr()), None)


def _case(t, dim, p, n, variant, **kw):
    lv = t.Level(dim, p, n, precision=kw.pop("precision", t.F64), vertices=kw.pop("vertices", None))
    sm = t.SchwarzSmoother(lv, variant=variant, **kw)
    return lv, sm


@pytest.mark.parametrize("variant", ["additive", "multiplicative"])
def test_noalloc_matches_the_allocating_entry_point(variant):
    t = _tpss()
    assert t.lib.tpss_noalloc_init() == t.OK
    if variant == "additive":
        big = _case(t, 2, 3, 64, t.ADDITIVE, relaxation=0.25)
        small = _case(t, 3, 2, 6, t.ADDITIVE, relaxation=0.15, ras=t.RAS_BOOLEAN)
    else:
        V = o.distorted_vertices(3, 8, 0.2, seed=3)
        big = _case(t, 3, 3, 8, t.MULTIPLICATIVE, vertices=V)
        small = _case(t, 2, 2, 7, t.MULTIPLICATIVE, coloring=t.COLOR_TILED, precision=t.F32)
    for lv, sm in (big, small):  # two calls in a row, the second smaller
        dt = np.float32 if lv.precision == t.F32 else np.float64
        src = dev(o.splitmix_uniform(31, lv.n_dofs).astype(dt))
        ref, got = dev(np.zeros(lv.n_dofs, dt)), dev(np.full(lv.n_dofs, 7.0, dt))
        # the allocation-free call first, on a smoother that has allocated no scratch
        assert _noalloc(t, sm, got, src) == t.OK, t.lib.tpss_last_error()
        sm.precond_vmult(ref, src)
        np.testing.assert_array_equal(host(got), host(ref))
    # and once more on the first smoother after the second used the array
    lv, sm = big
    src = dev(o.splitmix_uniform(32, lv.n_dofs))
    ref, got = dev(np.zeros(lv.n_dofs)), dev(np.zeros(lv.n_dofs))
    sm.precond_vmult(ref, src)
    assert _noalloc(t, sm, got, src) == t.OK
    np.testing.assert_array_equal(host(got), host(ref))


def test_noalloc_rejects_smoothers_beyond_the_scratch_size():
    t = _tpss()
    lv = t.Level(3, 3, 24)  # additive slot buffer 24^3 * 64 * 8 * 8 B = 56.6 MiB: fits
    sm = t.SchwarzSmoother(lv, variant=t.ADDITIVE)
    src, dst = dev(np.zeros(lv.n_dofs)), dev(np.zeros(lv.n_dofs))
    assert _noalloc(t, sm, dst, src) == t.OK
    lv2 = t.Level(3, 3, 26)  # 72.0 MiB: rejected
    sm2 = t.SchwarzSmoother(lv2, variant=t.ADDITIVE)
    src2, dst2 = dev(np.zeros(lv2.n_dofs)), dev(np.zeros(lv2.n_dofs))
    assert _noalloc(t, sm2, dst2, src2) == t.E_INVALID_ARG
    assert b"scratch" in t.lib.tpss_last_error()
