"""Specified lateral boundaries on the gfx950 path (needs an MI355X): relaxation and spec
zones of the dry fields (relax_bdy_dry, spec_bdy_dry, the acoustic spec_bdyupdate,
spec_bdyupdate_ph and zero_grad_bdy), the species' relaxation or flow-dependent
boundaries, against the FP32 oracle bit for bit: per routine, over three steps, and at
400x400x80; the host's dtbc advances by dt per step.
"""
import numpy as np
import pytest

from parity import phase_parity
from pd_parity import pd_state
from wrf_amd import grid, runtime

pytestmark = pytest.mark.gpu

# (DomainSpec keywords, config options, positive-definite options or None, boundary seed)
CASES = {
    "spec_all": (dict(nx=24, ny=20, nz=16, bc="specified", n_passive=1, n_bubbles=2, bubble_amp=2.0, seed=61), {},
                 None, 71),
    "spec_channel": (dict(nx=22, ny=19, nz=14, bc_x="periodic", bc_y="specified", n_passive=1, n_bubbles=1, seed=62),
                     {}, None, 72),
    # every species with boundary arrays, positive-definite transport, mixing and damping
    "spec_mixed": (dict(nx=23, ny=18, nz=14, bc="specified", n_passive=2, n_bubbles=1, seed=63),
                   dict(have_bcs_scalar=1, diff_6th_opt=2, damp_opt=3, zdamp=5000.0), (1, 1), 73),
}


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    runtime.load_library()
    return torch


def _state(case):
    kw, opts, pd, seed = CASES[case]
    hs = pd_state(kw, *pd, seed=kw["seed"]) if pd else grid.build_initial_state(grid.DomainSpec(**kw))
    for k, v in opts.items():
        setattr(hs.cfg, k, v)
    grid.build_lbc(hs, seed=seed)
    return hs


@pytest.mark.parametrize("case", sorted(CASES))
def test_specified_per_routine_bitwise(gpu, case):
    from pyoracle import Oracle
    hs = _state(case)
    print(f"per-routine parity, case {case}", flush=True)
    dev = runtime.DeviceDomain(hs)
    sess = Oracle("f32cr").session(hs)
    phases = runtime.step_phases(hs.cfg, hs.scal, hs.n_species, hs.n_moist)
    fails = phase_parity(dev, sess, phases, nsteps=2, eos_ulps=1)
    assert not fails, "\n".join(fails)


@pytest.mark.parametrize("case", sorted(CASES))
def test_specified_steps_bitwise(gpu, case):
    from pyoracle import Oracle
    hs = _state(case)
    print(f"3-step parity, case {case}", flush=True)
    dev = runtime.DeviceDomain(hs)
    dev.step(3)
    got = dev.download()
    exp = Oracle("f32cr").run(hs, 3)
    bad = [k for k in exp if k not in ("a", "alpha") and not np.array_equal(got[k], exp[k])]
    assert not bad, f"arrays differ from the oracle after three steps: {bad}"
    assert dev.lbc.dtbc == np.float32(3 * np.float32(hs.scal["dt"])), "dtbc advances by dt per step"


def test_specified_boundaries_act(gpu):
    """The boundary arrays drive the domain: stepping with their tendencies doubled changes
    the spec zone's values."""
    hs = _state("spec_all")
    other = hs.copy()
    for n in grid.LBC_FIELDS:
        for a in other.lbc[n + "_bt"]:
            a *= 2.0
    a_dev, b_dev = runtime.DeviceDomain(hs), runtime.DeviceDomain(other)
    a_dev.step(1)
    b_dev.step(1)
    a, b = a_dev.download(), b_dev.download()
    d = hs.dims
    J0, I0 = d.jds - d.jms, d.ids - d.ims
    assert not np.array_equal(a["t_2"][J0, :, I0:I0 + 10], b["t_2"][J0, :, I0:I0 + 10])
    assert not np.array_equal(a["mu_2"][J0, I0:I0 + 10], b["mu_2"][J0, I0:I0 + 10])


def test_specified_full_size_bitwise(gpu):
    """400x400x80 with specified boundaries on all four sides: one step bitwise equal to the
    oracle's."""
    from pyoracle import Oracle
    hs = grid.build_initial_state(grid.large_c1(bc="specified"))
    grid.build_lbc(hs, seed=74)
    print("400x400x80 specified boundaries, one step against the oracle", flush=True)
    dev = runtime.DeviceDomain(hs)
    dev.step(1)
    got = dev.download()
    exp = Oracle("f32cr").run(hs, 1)
    bad = [k for k in exp if k not in ("a", "alpha") and not np.array_equal(got[k], exp[k])]
    assert not bad, f"arrays differ from the oracle after one full-size step: {bad}"
    assert dev.check_finite() == 0
