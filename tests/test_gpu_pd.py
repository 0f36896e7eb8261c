"""Positive-definite scalar transport on the gfx950 path (needs an MI355X): advect_scalar_pd
and rk_update_scalar_pd (moist_adv_opt / scalar_adv_opt 1) against the reference's own
advect_scalar_pd (fixture) and the FP32 oracle, bit for bit; through the whole step, its
hipGraph replay, and at 400x400x80, where the species stay non-negative.
"""
import numpy as np
import pytest

from parity import interior, phase_parity
from pd_parity import fixture_failures, pd_state, update_pd_failures
from wrf_amd import grid, runtime

pytestmark = pytest.mark.gpu

CASES = {
    "pd_periodic": (dict(nx=24, ny=20, nz=16, bc="periodic", n_passive=2, n_bubbles=2, bubble_amp=2.0, seed=31), (1, 1)),
    "pd_open": (dict(nx=22, ny=19, nz=14, bc="open", n_passive=1, n_bubbles=1, seed=32), (1, 1)),
    # moist species unlimited, passive scalars positive definite; rows of 25 floats
    "pd_mixed_unaligned": (dict(nx=21, ny=17, nz=12, bc_x="open", bc_y="periodic", n_passive=1, n_bubbles=2, seed=33,
                                pad_i=1), (0, 1)),
}


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    runtime.load_library()
    return torch


def _state(case):
    kw, opts = CASES[case]
    return pd_state(kw, *opts, seed=kw["seed"])


def test_pd_routines_match_reference(gpu):
    from pyoracle import Oracle
    print("advect_scalar_pd / rk_update_scalar_pd entry points", flush=True)
    lib = runtime.load_library()
    fails = fixture_failures(lib, Oracle("f32")) + update_pd_failures(lib, Oracle("f32"))
    assert not fails, "\n".join(fails)


@pytest.mark.parametrize("case", sorted(CASES))
def test_pd_per_routine_bitwise(gpu, case):
    from pyoracle import Oracle
    hs = _state(case)
    print(f"per-routine parity, case {case}", flush=True)
    dev = runtime.DeviceDomain(hs)
    sess = Oracle("f32cr").session(hs)
    phases = runtime.step_phases(hs.cfg, hs.scal, hs.n_species, hs.n_moist)
    assert "rk_scalar_pd_prep" in [n for n, _ in phases]
    fails = phase_parity(dev, sess, phases, nsteps=2, eos_ulps=1)
    assert not fails, "\n".join(fails)


@pytest.mark.parametrize("case", sorted(CASES))
def test_pd_steps_bitwise(gpu, case):
    from pyoracle import Oracle
    hs = _state(case)
    print(f"3-step parity, case {case}", flush=True)
    dev = runtime.DeviceDomain(hs)
    dev.step(3)
    got = dev.download()
    exp = Oracle("f32cr").run(hs, 3)
    bad = [k for k in exp if k not in ("a", "alpha") and not np.array_equal(got[k], exp[k])]
    assert not bad, f"arrays differ from the oracle after three steps: {bad}"


def test_pd_graph_replay_matches_eager(gpu):
    hs = _state("pd_periodic")
    print("graph replay vs eager, positive definite", flush=True)
    eager, graphed = runtime.DeviceDomain(hs), runtime.DeviceDomain(hs)
    torch = gpu
    for n in range(3):
        eager.step(1)
        graphed.step(1, graph=True)
        x = torch.randn(1 << 20, device="cuda")
        _ = (x * 2).sum().item()
        a, b = eager.download(), graphed.download()
        bad = [k for k in a if k not in ("a", "alpha") and not np.array_equal(a[k], b[k])]
        assert not bad, f"replay {n + 1}: {bad} differ from the eager step"


def test_pd_full_size_bitwise_and_non_negative(gpu):
    """400x400x80 with every species positive definite and half-empty passive scalars: one
    step bitwise equal to the oracle; after three more the passive scalars are still
    non-negative (to rounding) everywhere."""
    from pyoracle import Oracle
    hs = pd_state(grid.large_c1(), 1, 1, seed=7)
    print("400x400x80 positive definite, one step against the oracle", flush=True)
    dev = runtime.DeviceDomain(hs)
    dev.step(1)
    got = dev.download()
    exp = Oracle("f32cr").run(hs, 1)
    bad = [k for k in exp if k not in ("a", "alpha") and not np.array_equal(got[k], exp[k])]
    assert not bad, f"arrays differ from the oracle after one full-size step: {bad}"
    dev.step(3)
    d = hs.dims
    out = dev.to_host([f"q{m}" for m in range(hs.n_moist, hs.n_species)])
    for k, q in out.items():
        qi = interior(q, d)
        print(f"  {k}: min {qi.min():.3e} max {qi.max():.3e} empty cells {int(np.sum(qi == 0))}", flush=True)
        assert qi.min() >= -1e-6 * qi.max(), k
    assert dev.check_finite() == 0
