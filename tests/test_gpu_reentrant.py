"""Re-entrancy of the library (needs an MI355X): every domain has its own work arrays, so
two domains of the same size stepped concurrently on two streams each equal their own
single-stream run bit for bit; and a graph cached for registered host arrays is not
replayed on a device mirror that was replaced."""
import ctypes

import numpy as np
import pytest

from wrf_amd import fields, grid, runtime

pytestmark = pytest.mark.gpu

SPEC = dict(nx=48, ny=40, nz=20, bc="periodic", n_passive=3, n_bubbles=2, bubble_amp=2.0)


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    runtime.load_library()
    return torch


def test_two_domains_on_two_streams(gpu):
    torch = gpu
    hs_a = grid.build_initial_state(grid.DomainSpec(seed=31, **SPEC))
    hs_b = grid.build_initial_state(grid.DomainSpec(seed=32, **SPEC))
    print("two same-size domains, two streams, 3 steps", flush=True)
    ref = []
    for hs in (hs_a, hs_b):
        dev = runtime.DeviceDomain(hs)
        dev.step(3)
        ref.append(dev.download())
        dev.release()
        del dev
    da, db = runtime.DeviceDomain(hs_a), runtime.DeviceDomain(hs_b)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    for _ in range(3):
        da.step(1, stream=sa, sync=False)
        db.step(1, stream=sb, sync=False)
    torch.cuda.synchronize()
    for name, dev, exp in (("a", da, ref[0]), ("b", db, ref[1])):
        got = dev.download()
        bad = [k for k in exp if not np.array_equal(got[k], exp[k])]
        assert not bad, f"domain {name}: {bad} differ from its single-stream run"


def test_graph_after_mirror_replaced(gpu):
    """Registered host arrays: step_graph, then re-register one array with a larger size
    (its device mirror is replaced), step_graph again; both graph steps equal eager steps."""
    hs = grid.build_initial_state(grid.DomainSpec(nx=20, ny=16, nz=12, bc="periodic", n_passive=1, seed=33))
    lib = runtime.load_library()
    runtime.check(lib.wrfmi_init())
    print("registered host arrays, graph replay across a re-registration", flush=True)

    def host_domain():
        arrays = {k: np.array(v, copy=True) for k, v in hs.arrays().items()}
        st = fields.State32()
        P = ctypes.POINTER(ctypes.c_float)
        for n in fields.FIELDS_3D + fields.FIELDS_2D + fields.FIELDS_1D:
            setattr(st, n, arrays[n].ctypes.data_as(P))
        for n in fields.SCALARS_R:
            setattr(st, n, hs.scal[n])
        st.a, st.alpha, st.gamma = (arrays[k].ctypes.data for k in ("a", "alpha", "gamma"))
        st.n_moist, st.n_species = hs.n_moist, hs.n_species
        for m in range(hs.n_species):
            st.q[m] = arrays[f"q{m}"].ctypes.data_as(P)
            st.q_old[m] = arrays[f"q_old{m}"].ctypes.data_as(P)
            st.q_tend[m] = arrays[f"q_tend{m}"].ctypes.data_as(P)
        return arrays, st

    def step(st, graph):
        fn = lib.wrfmi_solve_em_step_graph if graph else lib.wrfmi_solve_em_step
        runtime.check(fn(ctypes.byref(st), ctypes.byref(hs.cfg), ctypes.byref(hs.dims), None, 1))

    results = {}
    for graph in (False, True):
        arrays, st = host_domain()
        # t_tendf lives in a larger host buffer, so that it can be registered with two sizes
        big = np.zeros(arrays["t_tendf"].size * 2, np.float32)
        big[: arrays["t_tendf"].size] = arrays["t_tendf"].ravel()
        st.t_tendf = big.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        regs = [v for k, v in arrays.items() if k != "t_tendf"]
        for v in regs:
            runtime.check(lib.wrfmi_register_field(ctypes.c_void_p(v.ctypes.data), v.nbytes))
        runtime.check(lib.wrfmi_register_field(ctypes.c_void_p(big.ctypes.data), arrays["t_tendf"].nbytes))
        step(st, graph)
        # replace the mirror: same host address, new size
        runtime.check(lib.wrfmi_register_field(ctypes.c_void_p(big.ctypes.data), big.nbytes))
        step(st, graph)
        for v in regs:
            runtime.check(lib.wrfmi_sync_to_host(ctypes.c_void_p(v.ctypes.data)))
        results[graph] = {k: np.array(v, copy=True) for k, v in arrays.items() if k != "t_tendf"}
        for v in regs + [big]:
            runtime.check(lib.wrfmi_unregister_field(ctypes.c_void_p(v.ctypes.data)))
    bad = [k for k in results[False] if not np.array_equal(results[False][k], results[True][k])]
    assert not bad, f"graph steps differ from eager steps in {bad}"
