"""The 1200x1200x80 terrain case at full size on the gfx950 path (needs an MI355X):
hybrid levels, 12 Witch-of-Agnesi ridges, periodic in y and open in x, qv + 10 scalars.

The oracle cannot step this domain in test time, so the full size is checked through
properties that do not depend on it (the small terrain cases of test_gpu_parity are
bitwise against the oracle).

Translation in the periodic direction: every point runs the same operations whatever
its j, so a state shifted by s rows and stepped is, bit for bit, the stepped state
shifted by s rows -- with one exception that the reference itself makes.  Its advect_v
open-x boundary terms clamp the j-neighbours of ru to jds..jde-1 even when y is periodic
(dyn_em/module_advect_em.F:2749-2790), so the v tendency at the two open-x boundary
columns of row jds differs from a periodic neighbour's.  One RK3 step reaches at most 45
points horizontally (3 stages of 5th-order advection of a mass-weighted flux, reach 5, and
1 + 3 + 6 acoustic steps of reach 3), so after one step the rows farther than 64 from the
wrap row in both frames are compared bit for bit; the doubly periodic form of the same
domain has no such row and is compared everywhere after 20 steps, with its dry-air mass
conserved to rounding.
"""
import numpy as np
import pytest

from wrf_amd import grid, runtime

pytestmark = pytest.mark.gpu

NSTEPS = 20
BAND = 64
FIELDS = dict(u_2="u", v_2="v", w_2="w", t_2="mass", mu_2="mass", p="mass", ph_2="w", al="mass", q0="mass", q10="mass")


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    runtime.load_library()
    return torch


def _roll_j(hs, shift):
    """Roll every 2-D and 3-D array of `hs` in place: interior rows by `shift` in j, then
    refill the halos (the domain is periodic in y)."""
    d = hs.dims
    j0, ny = d.jds - d.jms, hs.spec.ny
    per = (hs.spec.bcx == "periodic", hs.spec.bcy == "periodic")
    for name, a in hs.arrays().items():
        if a.ndim == 1 or not np.any(a):
            continue
        a[j0:j0 + ny] = np.roll(a[j0:j0 + ny], shift, axis=0)
        grid._halo_fill_np(a, d, per, name in ("u_1", "u_2"), name in ("v_1", "v_2"))


def _cols(a, d, kind):
    """i and k extent of the domain for rows already cut to j = jds..jde-1."""
    i0 = d.ids - d.ims
    ni = d.ide - d.ids + (1 if kind == "u" else 0)
    if a.ndim == 3:
        return a[:, : (d.kde if kind == "w" else d.kde - 1), i0:i0 + ni]
    return a[:, i0:i0 + ni]


def _dry_mass(a, d):
    i0, j0 = d.ids - d.ims, d.jds - d.jms
    nx, ny = d.ide - d.ids, d.jde - d.jds
    return float(np.sum(a["mu_2"][j0:j0 + ny, i0:i0 + nx].astype(np.float64)))


def _two_runs(torch, hs, shift, snap_after):
    """Step `hs` and its copy shifted by `shift` rows; host copies of FIELDS after
    `snap_after` steps and after NSTEPS, and the finiteness after NSTEPS."""
    out = []
    for label in ("unshifted", f"shifted by {shift} rows"):
        if out:
            _roll_j(hs, shift)
        print(f"  {label}: upload and {NSTEPS} steps", flush=True)
        dev = runtime.DeviceDomain(hs)
        m0 = _dry_mass(dev.to_host(["mu_2"]), hs.dims)
        dev.step(snap_after)
        early = dev.to_host(list(FIELDS))
        dev.step(NSTEPS - snap_after)
        late = dev.to_host(list(FIELDS))
        nbad = dev.check_finite()
        print(f"    non-finite values after {NSTEPS} steps: {nbad}", flush=True)
        out.append((early, late, nbad, m0))
        del dev
        torch.cuda.empty_cache()
    return out


def _compare(a_run, b_run, d, ny, shift, keep_rows):
    fails = []
    j0 = d.jds - d.jms
    for k, kind in FIELDS.items():
        a = np.roll(a_run[k][j0:j0 + ny], shift, axis=0)[keep_rows]
        b = b_run[k][j0:j0 + ny][keep_rows]
        nbad = int(np.sum(_cols(a, d, kind) != _cols(b, d, kind)))
        print(f"    {k:5s} max|x| {float(np.abs(_cols(b, d, kind)).max()):.4g}  points differing {nbad}", flush=True)
        if nbad:
            fails.append(f"{k}: {nbad} points differ")
    return fails


@pytest.mark.timeout(900)
def test_terrain_full_size(torch_gpu):
    spec = grid.terrain_c4()
    print(f"terrain {spec.nx}x{spec.ny}x{spec.nz}, open x / periodic y: building the state", flush=True)
    hs = grid.build_initial_state(spec)
    shift = spec.ny // 3 + 7
    (a1, _, bad_a, _), (b1, _, bad_b, _) = _two_runs(torch_gpu, hs, shift, 1)
    assert bad_a == 0 and bad_b == 0
    # rows of the shifted frame farther than BAND from both frames' wrap rows
    r = np.arange(spec.ny)
    dist = lambda c: np.minimum((r - c) % spec.ny, (c - r) % spec.ny)
    keep = (dist(0) > BAND) & (dist(shift) > BAND)
    print(f"  after one step, {int(keep.sum())} of {spec.ny} rows compared", flush=True)
    fails = _compare(a1, b1, hs.dims, spec.ny, shift, keep)
    assert not fails, "; ".join(fails)


@pytest.mark.timeout(900)
def test_terrain_full_size_doubly_periodic(torch_gpu):
    spec = grid.terrain_c4(bc_x="periodic")
    print(f"terrain {spec.nx}x{spec.ny}x{spec.nz}, doubly periodic: building the state", flush=True)
    hs = grid.build_initial_state(spec)
    shift = spec.ny // 3 + 7
    mub = _dry_mass({"mu_2": hs.a2["mub"]}, hs.dims)
    (_, a, bad_a, m0), (_, b, bad_b, _) = _two_runs(torch_gpu, hs, shift, 1)
    assert bad_a == 0 and bad_b == 0
    drift = abs(_dry_mass(a, hs.dims) - m0) / mub
    print(f"  dry-air mass drift over {NSTEPS} steps: {drift:.2e}", flush=True)
    assert drift < 1e-6
    fails = _compare(a, b, hs.dims, spec.ny, shift, np.ones(spec.ny, bool))
    assert not fails, "; ".join(fails)
