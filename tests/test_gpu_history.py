"""History output on the gfx950 path (needs an MI355X): wrfmi_history_* writes the
reference's wrfout layout from the device state; three records, with steps between them,
read back with scipy's netCDF reader and compared bit for bit with the downloaded state."""
import os

import pytest

from history_check import history_failures
from wrf_amd import grid, runtime

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    runtime.load_library()
    return torch


@pytest.mark.parametrize("kw", [dict(nx=160, ny=128, nz=40, bc="periodic", n_passive=3, n_bubbles=3, seed=51),
                                dict(nx=96, ny=64, nz=30, dx=2000.0, dy=2000.0, dt=6.0, bc_x="open", bc_y="periodic",
                                     hybrid_opt=2, sounding="stable", n_ridges=3, ridge_a=(3000.0, 8000.0),
                                     n_passive=1, n_bubbles=0, noise_amp=0.0, seed=52)])
def test_history_records_match_state(gpu, kw, tmp_path):
    hs = grid.build_initial_state(grid.DomainSpec(**kw))
    print(f"history file, {kw['nx']}x{kw['ny']}x{kw['nz']}", flush=True)
    dev = runtime.DeviceDomain(hs)
    path = os.path.join(tmp_path, "wrfout_d01_2026-06-01_12:00:00")
    start = "2026-06-01_12:00:00"
    recs = []
    with runtime.History(dev.lib, dev.state, dev.cfg, dev.dims, path, start) as h:
        for r in range(3):
            date = f"2026-06-01_12:{r:02d}:00"
            h.write(date, float(r), stream=dev.stream_handle())
            recs.append((date, float(r), dev.download()))
            dev.step(10)
        assert h.records == 3
    fails = history_failures(path, recs, hs.dims, hs.cfg, hs.scal, hs.n_moist, hs.n_species, start)
    assert not fails, "\n".join(fails)
