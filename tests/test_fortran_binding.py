"""The Fortran interface module (wrf_amd/fortran/wrfmi_bindc.f90) compiles, links against
libwrfmi.so, agrees with the C struct layout and reaches the library's option check."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"


def test_fortran_host_links_and_checks_layout():
    if not os.path.exists(FLANG):
        pytest.skip("no Fortran compiler")
    lib = os.path.join(ROOT, "wrf_amd")
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "check_binding")
        subprocess.run([FLANG, "-module-dir", tmp, os.path.join(ROOT, "wrf_amd", "fortran", "wrfmi_bindc.f90"),
                        os.path.join(ROOT, "tests", "fortran", "check_binding.f90"), "-L" + lib, "-lwrfmi",
                        "-Wl,-rpath," + lib, "-o", exe], check=True, capture_output=True, cwd=tmp)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK" in r.stdout and "rk_ord" in r.stdout


def test_fortran_interface_matches_generator():
    """wrfmi_bindc.f90 is what gen_wrfmi_bindc.py writes for the current field lists: a
    field or config member added to the C structs without regenerating it would shift
    every later member of the Fortran derived types."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen_wrfmi_bindc",
                                                  os.path.join(ROOT, "wrf_amd", "fortran", "gen_wrfmi_bindc.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    with open(os.path.join(ROOT, "wrf_amd", "fortran", "wrfmi_bindc.f90")) as f:
        assert f.read() == gen.generate(), "wrfmi_bindc.f90 is stale: run wrf_amd/fortran/gen_wrfmi_bindc.py"
