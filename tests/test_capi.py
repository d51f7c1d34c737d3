"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/tpss.h, and rejects invalid arguments the way the
reference's AssertThrow checks do (no device work is started)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "tpss.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tpss_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import tpss_amd
    lib = ctypes.CDLL(tpss_amd.LIB_PATH)
    names = declared_entry_points()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(tpss_amd.EXPORTS) == names


def test_version_string():
    import tpss_amd
    assert "gfx950" in tpss_amd.version()


@pytest.mark.parametrize("kw,msg", [
    (dict(dim=4, degree=3, n_cells_1d=4), "dim"),
    (dict(dim=2, degree=0, n_cells_1d=4), "degree"),
    (dict(dim=2, degree=3, n_cells_1d=1), "n_cells_1d"),
    (dict(dim=2, degree=3, n_cells_1d=4, ip_factor=-1.0), "ip_factor"),
])
def test_level_argument_checks(kw, msg):
    import tpss_amd
    with pytest.raises(tpss_amd.TPSSError) as e:
        tpss_amd.Level(**kw)
    assert e.value.code == tpss_amd.E_INVALID_ARG
    assert msg in str(e.value)


def test_level_n_dofs_without_device_work():
    import tpss_amd
    lv = tpss_amd.Level(3, 3, 8)
    assert lv.n_dofs == 8 ** 3 * 64


def test_smoother_rejects_nonpositive_relaxation():
    import tpss_amd
    lv = tpss_amd.Level(2, 3, 4)
    with pytest.raises(tpss_amd.TPSSError) as e:
        tpss_amd.SchwarzPreconditioner(lv, relaxation=0.0)
    assert e.value.code == tpss_amd.E_INVALID_ARG
    with pytest.raises(tpss_amd.TPSSError):
        tpss_amd.SchwarzPreconditioner(lv, variant=tpss_amd.MULTIPLICATIVE,
                                       coloring=tpss_amd.COLOR_NONE)


def test_null_pointers_rejected():
    import tpss_amd
    lv = tpss_amd.Level(2, 3, 4)
    st = tpss_amd.lib.tpss_operator_vmult(lv.handle, None, None, None)
    assert st == tpss_amd.E_INVALID_ARG
    assert b"null" in tpss_amd.lib.tpss_last_error()


def test_level_size_limits():
    """one block of at most 256 threads per cell keeps every launch below 2^32
    work-items: at most 2^24 cells (3D: n <= 255) and 2^31 DoFs"""
    import tpss_amd
    assert tpss_amd.Level(3, 1, 255).n_dofs == 255 ** 3 * 8
    assert tpss_amd.Level(2, 4, 1000).n_dofs == 1000 ** 2 * 25
    assert tpss_amd.Level(3, 4, 255).n_dofs == 255 ** 3 * 125
    for kw in (dict(dim=3, degree=1, n_cells_1d=256),
               dict(dim=3, degree=4, n_cells_1d=200, elasticity=(1.0, 10.0))):
        with pytest.raises(tpss_amd.TPSSError) as e:
            tpss_amd.Level(**kw)
        assert e.value.code == tpss_amd.E_INVALID_ARG
        assert "too large" in str(e.value)


@pytest.mark.parametrize("dim", [2, 3])
def test_collapsed_or_inverted_cells_rejected_before_device_work(dim):
    import numpy as np

    import oracle as o
    import tpss_amd
    n = 4
    V = o.uniform_vertices(dim, n).reshape((n + 1,) * dim + (dim,)).copy()
    sel = (slice(None),) * (dim - 1) + (1, 0)
    V[sel] = V[(slice(None),) * (dim - 1) + (2, 0)]  # zero-width column of cells
    with pytest.raises(tpss_amd.TPSSError) as e:
        tpss_amd.Level(dim, 2, n, vertices=V.reshape(-1, dim))
    assert e.value.code == tpss_amd.E_SINGULAR
    assert "collapsed or inverted" in str(e.value)
    # one interior vertex pushed past its x-neighbour inverts two cells
    W = o.uniform_vertices(dim, n).copy()
    W[(n + 1) + 1 if dim == 2 else (n + 1) ** 2 + (n + 1) + 1, 0] += 1.5 / n
    with pytest.raises(tpss_amd.TPSSError) as e:
        tpss_amd.Level(dim, 1, n, vertices=W)
    assert e.value.code == tpss_amd.E_SINGULAR


def test_aliased_operator_arguments_rejected():
    import tpss_amd
    lv = tpss_amd.Level(2, 3, 4)
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    assert tpss_amd.lib.tpss_operator_vmult(lv.handle, a, a, None) == tpss_amd.E_INVALID_ARG
    assert b"alias" in tpss_amd.lib.tpss_last_error()
    for args in ((a, b, a), (a, a, b)):
        st = tpss_amd.lib.tpss_operator_residual(lv.handle, *args, None)
        assert st == tpss_amd.E_INVALID_ARG


def test_borrowed_wrappers_keep_their_solver_alive(monkeypatch):
    """GMGSolver.level()/smoother() wrap handles owned by the C solver: the
    solver must not be destroyed while such a wrapper is referenced (checked
    with a stand-in library, no device work)."""
    import gc

    import tpss_amd as t

    destroyed = []

    class FakeLib:
        def __getattr__(self, name):
            def f(*args):
                if name == "tpss_gmg_destroy":
                    destroyed.append(args[0])
                elif name in ("tpss_gmg_level", "tpss_gmg_smoother"):
                    args[-1]._obj.value = 0x1234
                elif name == "tpss_level_n_dofs":
                    args[-1]._obj.value = 64
                return t.OK
            return f

    monkeypatch.setattr(t, "lib", FakeLib())
    g = t.GMGSolver.__new__(t.GMGSolver)
    g.handle = ctypes.c_void_p(0x99)
    g.dim, g.degree, g.n_levels, g.mg_precision = 2, 1, 2, t.F64
    g.desc = t.GMGDesc(2, 1, 2, 2, 1.0, None, t.F64,
                       t._smoother_desc(1, 2, 1.0, 0, 0, 1.0, 1, t.F64), 1e-12, 0, 0.0, 0.0)
    sm = g.smoother(1)
    lv = g.level(0)
    del g
    gc.collect()
    assert destroyed == []
    del sm
    gc.collect()
    assert destroyed == []
    del lv
    gc.collect()
    assert len(destroyed) == 1
