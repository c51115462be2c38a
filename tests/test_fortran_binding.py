"""The Fortran iso_c_binding module (chapsim2_amd/fortran/chapsim_hip.F90)
compiles with flang, links against the library, and its host-only calls
agree with the Python binding (no GPU needed)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("amdflang") or shutil.which("flang") or "/opt/rocm/bin/amdflang"


@pytest.mark.skipif(not os.path.exists(FLANG), reason="no Fortran compiler")
def test_fortran_binding(tmp_path):
    from chapsim2_amd import _lib
    from chapsim2_amd import operations as ops
    from chapsim2_amd.domain import make_desc

    _lib.load()
    libdir = os.path.join(ROOT, "chapsim2_amd", "lib")
    exe = str(tmp_path / "binding_check")
    cmd = [FLANG, f"-J{tmp_path}", os.path.join(ROOT, "chapsim2_amd", "fortran", "chapsim_hip.F90"),
           os.path.join(ROOT, "tests", "fortran", "binding_check.f90"), f"-L{libdir}", "-lchapsim_hip",
           f"-Wl,-rpath,{libdir}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    assert lines[0].split()[1] == str(_lib.load().chs_abi_version())
    assert float(lines[1].split()[1]) == _lib.load().chs_inspect_coeff(2, 1, 3, 1, 0, 4)
    got = np.array([float(v) for v in lines[2:] if v.strip()])
    d = make_desc((16, 24, 8), 1.0, -1.0, 1.0, 1.0, istret=2, rstret=0.075)
    j = np.arange(1, 25)
    fi = np.sin(0.3 * j) + 0.01 * j * j
    ref = ops.inspect_op1d(d, 1, 4, 4, [5, 5], fi, 25, np.array([0.5, -0.25, 0.0, 0.0]))
    assert np.abs(got - ref).max() < 1e-15 * np.abs(ref).max() * 10
