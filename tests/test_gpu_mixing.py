"""Mixing and damping of the first RK substep on the gfx950 path (needs an MI355X):
second-order diffusion with constant eddy coefficients (diff_opt 1, km_opt 1), sixth-order
diffusion (diff_6th_opt 1 and 2, with and without slope tapering), the Rayleigh damper toward
the base profiles (damp_opt 2) and advance_w's implicit upper damping layer (damp_opt 3),
against the FP32 oracle bit for bit: per routine, over three steps, under a hipGraph replay
and at 400x400x80.
"""
import numpy as np
import pytest

from parity import phase_parity
from pd_parity import pd_state
from wrf_amd import grid, runtime

pytestmark = pytest.mark.gpu

# (DomainSpec keywords, config options, positive-definite transport options or None)
CASES = {
    "mix_periodic": (dict(nx=24, ny=20, nz=16, bc="periodic", n_passive=1, n_bubbles=2, bubble_amp=2.0, seed=41),
                     dict(diff_6th_opt=2, diff_6th_slopeopt=1, damp_opt=2, zdamp=6000.0), None),
    "mix_open": (dict(nx=22, ny=19, nz=14, bc="open", n_passive=1, n_bubbles=1, seed=42),
                 dict(diff_6th_opt=1, damp_opt=3, zdamp=5000.0, dampcoef=0.3), None),
    "mix_terrain": (dict(nx=24, ny=16, nz=14, dx=2000.0, dy=2000.0, dt=6.0, bc_x="open", bc_y="periodic",
                         hybrid_opt=2, sounding="stable", n_ridges=3, ridge_a=(3000.0, 8000.0), n_passive=1,
                         n_bubbles=0, noise_amp=0.0, seed=43),
                    dict(diff_6th_opt=2, diff_6th_slopeopt=1, diff_6th_thresh=0.02, damp_opt=2, zdamp=6000.0), None),
    "diff_periodic": (dict(nx=24, ny=20, nz=16, bc="periodic", n_passive=1, n_bubbles=2, bubble_amp=2.0, seed=45),
                      dict(diff_opt=1, km_opt=1, khdif=300.0, kvdif=50.0), None),
    "diff_terrain": (dict(nx=24, ny=16, nz=14, dx=2000.0, dy=2000.0, dt=6.0, bc_x="open", bc_y="periodic",
                          hybrid_opt=2, sounding="stable", n_ridges=3, ridge_a=(3000.0, 8000.0), n_passive=1,
                          n_bubbles=0, noise_amp=0.0, seed=46),
                     dict(diff_opt=1, km_opt=1, khdif=500.0, kvdif=30.0, diff_6th_opt=1, damp_opt=2, zdamp=6000.0),
                     None),
    # mixing together with positive-definite transport of every species
    "mix_pd": (dict(nx=24, ny=20, nz=16, bc="periodic", n_passive=2, n_bubbles=2, bubble_amp=2.0, seed=44),
               dict(diff_opt=1, km_opt=1, khdif=300.0, kvdif=50.0, diff_6th_opt=2, damp_opt=3, zdamp=6000.0,
                    dampcoef=0.2), (1, 1)),
}


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    runtime.load_library()
    return torch


def _state(kw, opts, pd):
    hs = pd_state(kw, *pd, seed=kw["seed"]) if pd else grid.build_initial_state(grid.DomainSpec(**kw))
    for k, v in opts.items():
        setattr(hs.cfg, k, v)
    return hs


def _case(case):
    return _state(*CASES[case])


@pytest.mark.parametrize("case", sorted(CASES))
def test_mixing_per_routine_bitwise(gpu, case):
    from pyoracle import Oracle
    hs = _case(case)
    print(f"per-routine parity, case {case}", flush=True)
    dev = runtime.DeviceDomain(hs)
    sess = Oracle("f32cr").session(hs)
    phases = runtime.step_phases(hs.cfg, hs.scal, hs.n_species, hs.n_moist)
    fails = phase_parity(dev, sess, phases, nsteps=2, eos_ulps=1)
    assert not fails, "\n".join(fails)


@pytest.mark.parametrize("case", sorted(CASES))
def test_mixing_steps_bitwise(gpu, case):
    from pyoracle import Oracle
    hs = _case(case)
    print(f"3-step parity, case {case}", flush=True)
    dev = runtime.DeviceDomain(hs)
    dev.step(3)
    got = dev.download()
    exp = Oracle("f32cr").run(hs, 3)
    bad = [k for k in exp if k not in ("a", "alpha") and not np.array_equal(got[k], exp[k])]
    assert not bad, f"arrays differ from the oracle after three steps: {bad}"


@pytest.mark.parametrize("case", ["mix_open", "diff_periodic"])
def test_mixing_changes_the_step(gpu, case):
    """The options act: the same state stepped with mixing and damping off differs from the
    one stepped with them on (so that the bitwise tests above compare the mixing itself)."""
    kw, opts, _ = CASES[case]
    on, off = runtime.DeviceDomain(_state(kw, opts, None)), runtime.DeviceDomain(_state(kw, {}, None))
    on.step(1)
    off.step(1)
    a, b = on.download(), off.download()
    for k in ("u_2", "v_2", "w_2", "t_2", "q0"):
        assert not np.array_equal(a[k], b[k]), k


def test_mixing_graph_replay_matches_eager(gpu):
    hs = _case("mix_pd")
    print("graph replay vs eager, mixing and damping", flush=True)
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


@pytest.mark.parametrize("opts", [dict(diff_6th_opt=2, diff_6th_slopeopt=1, damp_opt=2, zdamp=5000.0),
                                  dict(diff_opt=1, km_opt=1, khdif=300.0, kvdif=50.0, diff_6th_opt=1, damp_opt=3,
                                       zdamp=5000.0, dampcoef=0.2)])
def test_mixing_full_size_bitwise(gpu, opts):
    """400x400x80 with the mixing and damping options on: one step bitwise equal to the
    oracle's."""
    from pyoracle import Oracle
    hs = grid.build_initial_state(grid.large_c1())
    for k, v in opts.items():
        setattr(hs.cfg, k, v)
    print(f"400x400x80 one step against the oracle, {opts}", flush=True)
    dev = runtime.DeviceDomain(hs)
    dev.step(1)
    got = dev.download()
    exp = Oracle("f32cr").run(hs, 1)
    bad = [k for k in exp if k not in ("a", "alpha") and not np.array_equal(got[k], exp[k])]
    assert not bad, f"arrays differ from the oracle after one full-size step: {bad}"
    assert dev.check_finite() == 0
