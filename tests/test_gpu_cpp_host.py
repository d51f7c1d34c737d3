"""The C++ host mirror (include/tpss_schwarz.hpp) drives the library like the
reference's smoother is driven: built with g++ against the C ABI, run, and
compared with the oracle (needs an MI355X)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as o

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_mirror(tmp_path):
    import tpss_amd
    if tpss_amd.device_count() < 1:
        pytest.skip("no HIP device")
    exe = tmp_path / "schwarz_cpp"
    libdir = os.path.join(ROOT, "tpss_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "schwarz_cpp.cpp"), "-o", str(exe),
                    "-L", libdir, "-ltpss_amd", f"-Wl,-rpath,{libdir}"], check=True)
    out = tmp_path / "out.bin"
    res = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    N = 6 ** 3 * 27
    data = np.fromfile(out, dtype=np.float64)
    src = o.splitmix_uniform(21, N)
    ref = o.Smoother(3, 2, 6, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK)
    expect = ref.smoother_vmult(src)
    assert np.abs(data[:N] - expect).max() < 1e-12 * np.abs(expect).max()
    refr = o.Smoother(3, 2, 6, variant=o.MULTIPLICATIVE, coloring=o.COLOR_REDBLACK, reverse=True)
    expect2 = refr.precond_vmult(src)
    assert np.abs(data[N:] - expect2).max() < 1e-12 * np.abs(expect2).max()
