"""The allocation-free whole step (wrfmi_solve_em_step_noalloc, work arrays in the library's
static device array) against the allocating whole step (wrfmi_solve_em_step), both called
through the C ABI: bit for bit on every array, on two calls in a row, the second on a
smaller domain, so that the second call runs on work-array contents left by the first."""
import ctypes

import numpy as np
import pytest

from wrf_amd import grid, runtime

pytestmark = pytest.mark.gpu

VP = ctypes.c_void_p


def _entry(lib, name):
    fn = getattr(lib, name)
    fn.argtypes = [VP, VP, VP, VP, ctypes.c_int]
    fn.restype = ctypes.c_int
    return fn


def _step(hs, name, nsteps):
    dev = runtime.DeviceDomain(hs)
    fn = _entry(dev.lib, name)
    for _ in range(nsteps):
        runtime.check(fn(ctypes.byref(dev.state), ctypes.byref(dev.cfg), ctypes.byref(dev.dims),
                         VP(dev.stream_handle()), 1))
    out = dev.download()
    dev.release()
    return out


# the larger domain first (aligned rows: four-point kernels; several species), then a
# smaller one with open boundaries and a single species
CASES = [
    dict(nx=96, ny=72, nz=30, bc="periodic", n_passive=2, n_bubbles=2, bubble_amp=2.0, seed=31),
    dict(nx=40, ny=36, nz=18, bc="open", n_bubbles=1, seed=32),
]


def test_noalloc_step_equals_allocating_step():
    for n, kw in enumerate(CASES):
        print(f"no-alloc case {n}: {kw}", flush=True)
        hs = grid.build_initial_state(grid.DomainSpec(**kw))
        want = _step(hs, "wrfmi_solve_em_step", 2)
        got = _step(hs, "wrfmi_solve_em_step_noalloc", 2)
        bad = [k for k in want if not np.array_equal(got[k], want[k], equal_nan=True)]
        assert not bad, f"case {n}: arrays differ from the allocating step: {bad}"


def test_noalloc_refuses_large_domains():
    lib = runtime.load_library()
    fn = _entry(lib, "wrfmi_solve_em_step_noalloc")
    spec = grid.DomainSpec(nx=400, ny=400, nz=40)
    cfg, dims = grid.make_config(spec), grid.make_dims(spec)
    st = runtime.State32()
    rc = fn(ctypes.byref(st), ctypes.byref(cfg), ctypes.byref(dims), None, 1)
    assert rc == runtime.WRFMI_E_DIMS
    assert "allocation-free" in lib.wrfmi_last_error().decode()
