"""The entry points that keep the reference's argument lists (include/wrfmi_routines.h) and the
Fortran solve_em-shaped host that calls them (wrf_amd/fortran/solve_em_host.f90, built into
wrf_amd/libwrfmi_fortran.so), on the GPU, against the oracle (needs an MI355X).

Bars: each of advect_u/v/w/scalar, calc_p_rho, calc_coef_w, advance_uv, advance_mu_t,
advance_w, sumflux and rk_update_scalar bitwise equal to the oracle routine of the same name
from identical inputs (every array compared); two RK3 steps of the Fortran host bitwise equal
to the oracle's two steps (the oracle whose EOS rounds LOG and ** once from double, as the
GPU does), every array compared.
"""
import ctypes
import os

import numpy as np
import pytest

from routine_parity import routine_parity
from wrf_amd import grid, runtime

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FORTRAN_LIB = os.path.join(ROOT, "wrf_amd", "libwrfmi_fortran.so")

CASES = {
    "periodic": dict(nx=24, ny=20, nz=16, bc="periodic", n_passive=2, n_bubbles=2, bubble_amp=2.0, seed=11),
    "open": dict(nx=22, ny=19, nz=14, bc="open", n_bubbles=1, seed=12),
    "terrain": dict(nx=26, ny=20, nz=16, dx=1000.0, dy=1000.0, dt=3.0, bc_x="open", bc_y="periodic", hybrid_opt=2,
                    sounding="stable", n_ridges=3, ridge_a=(3000.0, 8000.0), n_passive=2, n_bubbles=0, noise_amp=0.0,
                    seed=21),
}


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    runtime.load_library()
    return torch


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("rk_step", [1, 2])
def test_reference_argument_entry_points(gpu, case, rk_step):
    from pyoracle import Oracle
    print(f"reference-argument entry points, case {case}, rk_step {rk_step}", flush=True)
    hs = grid.build_initial_state(grid.DomainSpec(**CASES[case]))
    dev = runtime.DeviceDomain(hs)
    fails = routine_parity(dev, Oracle("f32cr").session(hs), hs, rk_step=rk_step)
    dev.release()
    assert not fails, "\n".join(fails)


@pytest.mark.parametrize("case", sorted(CASES))
def test_fortran_host_two_steps(gpu, case):
    from pyoracle import Oracle
    print(f"Fortran solve_em host, case {case}", flush=True)
    if not os.path.exists(FORTRAN_LIB):
        pytest.fail(f"{FORTRAN_LIB} is not built (make -C wrf_amd/fortran)")
    flib = ctypes.CDLL(FORTRAN_LIB)
    flib.wrfmi_fortran_solve_em_step.argtypes = [ctypes.c_void_p] * 3
    flib.wrfmi_fortran_solve_em_step.restype = ctypes.c_int
    hs = grid.build_initial_state(grid.DomainSpec(**CASES[case]))
    dev = runtime.DeviceDomain(hs)
    for _ in range(2):
        runtime.check(flib.wrfmi_fortran_solve_em_step(ctypes.byref(dev.state), ctypes.byref(dev.cfg),
                                                       ctypes.byref(dev.dims)))
    got = dev.download()
    dev.release()
    exp = Oracle("f32cr").run(hs, 2)
    bad = [k for k in exp if not np.array_equal(got[k], exp[k])]
    assert not bad, f"the Fortran host's two steps differ from the oracle in {bad}"
