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
    return sorted(set(re.findall(r"\b(tpss_[a-z0-9_]+)\s*\(", text)))


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
