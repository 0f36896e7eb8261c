"""Batched ensembles on the GPU (wrfmi_ensemble_*): members of uneven sizes advanced by one
launch per kernel are bit for bit the members advanced one at a time by
wrfmi_solve_em_step, eagerly and from the captured graph (needs an MI355X)."""
import numpy as np
import pytest

from wrf_amd import grid, runtime

pytestmark = pytest.mark.gpu
KEYS = ("u_2", "v_2", "w_2", "t_2", "mu_2", "ph_2", "p", "al", "ww", "q0")


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    runtime.load_library()
    return torch


def _singles(hss, nsteps):
    out = []
    for hs in hss:
        d = runtime.DeviceDomain(hs)
        d.step(nsteps)
        out.append(d.to_host(KEYS))
        del d
    return out


def _check(ens, expected, tag):
    for m, (mem, exp) in enumerate(zip(ens.members, expected)):
        got = mem.to_host(KEYS)
        for k in KEYS:
            assert np.array_equal(got[k], exp[k]), f"{tag}: member {m} field {k} differs from its single-domain step"


def test_uneven_ensemble_matches_single_domains(gpu):
    specs = [grid.DomainSpec(nx=nx, ny=ny, nz=nz, bc="open", n_bubbles=1, bubble_amp=a, noise_amp=0.2, seed=4100 + m)
             for m, (nx, ny, nz, a) in enumerate([(16, 24, 12, 1.0), (33, 17, 12, 2.5), (40, 40, 14, 4.0),
                                                   (24, 16, 12, 3.0), (21, 35, 13, 1.5)])]
    hss = [grid.build_initial_state(s) for s in specs]
    print("ensemble: 5 uneven open-boundary members, eager", flush=True)
    exp2 = _singles(hss, 2)
    ens = runtime.DeviceEnsemble([runtime.DeviceDomain(hs) for hs in hss])
    ens.step(2)
    _check(ens, exp2, "eager")
    ens.close()
    print("ensemble: graph replays", flush=True)
    exp = _singles(hss, 3)
    ens = runtime.DeviceEnsemble([runtime.DeviceDomain(hs) for hs in hss])
    torch = gpu
    for n in range(3):
        ens.step(1, graph=True)
        x = torch.randn(1 << 20, device="cuda")
        _ = (x * 2).sum().item()
    _check(ens, exp, "graph")
    ens.close()


def test_full_size_ensemble_members(gpu):
    """The first four members of the 64-member uneven ensemble (nz = 50, nx, ny from
    {64 ... 256}) at full size: one batched step equals the members' own steps."""
    specs = grid.ensemble_c3()[:4]
    print("ensemble: 4 full-size members", [(s.nx, s.ny) for s in specs], flush=True)
    hss = [grid.build_initial_state(s) for s in specs]
    exp = _singles(hss, 1)
    ens = runtime.DeviceEnsemble([runtime.DeviceDomain(hs) for hs in hss])
    ens.step(1, graph=True)
    _check(ens, exp, "full size")
    for mem in ens.members:
        assert mem.check_finite() == 0
    ens.close()


def test_mismatched_members_refused(gpu):
    a = grid.build_initial_state(grid.DomainSpec(nx=12, ny=12, nz=10, bc="open", seed=1))
    b = grid.build_initial_state(grid.DomainSpec(nx=12, ny=12, nz=10, bc="open", seed=2, time_step_sound=8))
    with pytest.raises(runtime.WrfmiError) as ei:
        runtime.DeviceEnsemble([runtime.DeviceDomain(a), runtime.DeviceDomain(b)])
    assert ei.value.code == runtime.WRFMI_E_BADOPT
