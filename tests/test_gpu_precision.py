"""Errors of the reduced-precision paths against the FP64 oracle, for one
smoothing step and for the final GMG-CG solution.

Workload: 3D SIPG DG Q3, 16^3 Cartesian cells (262,144 DoFs), multiplicative
vertex-patch smoother with red-black colors, omega = 1, right-hand side
U[-1,1] (seed 2).  Paths: FP64 (vectors and solves), FP32 (FP32 vectors and
patch solves; the residual-free color passes), BF16-MFMA patch solves with FP32
accumulation on FP32 vectors.  The GMG-CG runs 4 levels (2^3 .. 16^3 cells)
with the outer CG, coarse solve and setup in FP64.

Measured errors (max norm, relative to the oracle's max norm) are printed
(`pytest -s`; on one MI355X: step FP64 5.0e-15, FP32 3.2e-7, BF16 1.2e-2; final
solution FP64 6.1e-16, FP32 2.1e-12, BF16 1.3e-9, all in the oracle's 4
iterations) and bounded here:
  one smoothing step:  FP64 < 1e-12, FP32 < 1e-5, BF16 < 3e-2
  final solution:      FP64 < 1e-12, mixed FP32 < 1e-10, mixed BF16 < 1e-7, and
                       every path's true FP64 residual <= 1.01e-8 of the initial one
"""
import numpy as np
import pytest

import oracle as o
from test_gpu_parity import _tpss, dev, host, rel_err

pytestmark = pytest.mark.gpu

N_CELLS = 16


def paths(t):
    return (("fp64", t.F64, t.F64), ("fp32", t.F32, t.F32), ("bf16", t.F32, t.BF16))


def test_smoothing_step_error_by_precision():
    t = _tpss()
    ref = o.Smoother(3, 3, N_CELLS, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK,
                     relaxation=1.0)
    b = o.splitmix_uniform(2, ref.N)
    expect = ref.step(np.zeros(ref.N), b)
    errs = {}
    for name, vp, lp in paths(t):
        lv = t.Level(3, 3, N_CELLS, precision=vp)
        sm = t.SchwarzSmoother(lv, variant=t.MULTIPLICATIVE, coloring=t.COLOR_REDBLACK,
                               relaxation=1.0, local_precision=lp)
        x = dev(np.zeros(ref.N), vp)
        sm.step(x, dev(b, vp))
        errs[name] = rel_err(host(x), expect)
    print("one smoothing step, error vs FP64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()})
    assert errs["fp64"] < 1e-12
    assert errs["fp32"] < 1e-5
    assert errs["bf16"] < 3e-2


def test_final_solution_error_by_precision():
    t = _tpss()
    levels = 4
    ref = o.GMG(3, 3, 2, levels, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK,
                relaxation=1.0)
    b = o.splitmix_uniform(2, ref.N)
    xr, itr, histr = ref.cg(b, rel_tol=1e-8)
    assert itr > 0
    lv64 = t.Level(3, 3, 2 << (levels - 1))
    res = {}
    for name, mp, lp in paths(t):
        g = t.GMGSolver(3, 3, 2, levels, mg_precision=mp, variant=t.MULTIPLICATIVE,
                        coloring=t.COLOR_REDBLACK, relaxation=1.0, local_precision=lp)
        x = dev(np.zeros(g.n_dofs))
        info, hist = g.solve(x, dev(b), rel_tol=1e-8)
        r = dev(np.zeros(g.n_dofs))
        lv64.residual(r, dev(b), x)
        true_res = np.linalg.norm(host(r)) / np.linalg.norm(b)
        res[name] = (rel_err(host(x), xr), info.iterations, true_res)
    print("GMG-CG to 1e-8, solution error vs FP64 oracle (error, iterations, true residual):",
          {k: (f"{e:.2e}", it, f"{r:.2e}") for k, (e, it, r) in res.items()}, "oracle its", itr)
    assert res["fp64"][0] < 1e-12 and res["fp64"][1] == itr
    assert res["fp32"][0] < 1e-10 and res["bf16"][0] < 1e-7
    for name, (e, it, r) in res.items():
        assert r <= 1.01e-8, (name, r)
