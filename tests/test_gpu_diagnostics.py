"""FP64 domain totals on the GPU (wrfmi_domain_totals, needs an MI355X): against the same sums
in float64 numpy, and the conservation of dry mass and of the theta budget over ten steps of a
doubly periodic domain with no diabatic heating."""
import numpy as np
import pytest

from parity import domain_totals_numpy
from wrf_amd import grid, runtime

pytestmark = pytest.mark.gpu


def test_domain_totals_and_conservation():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    hs = grid.build_initial_state(grid.DomainSpec(nx=96, ny=80, nz=30, bc="periodic", n_passive=1, n_bubbles=3,
                                                  bubble_amp=2.0, seed=44))
    print("domain totals: 96x80x30 periodic", flush=True)
    dev = runtime.DeviceDomain(hs)
    tol = np.array([1e-12, 1e-6, 1e-12, 1e-12, 1e-12, 1e-12])
    res = []
    for stage in range(2):
        got = np.array(list(dev.domain_totals().values()))
        exp = domain_totals_numpy(dev.download(), hs)
        rel = np.abs(got - exp) / np.abs(exp)
        assert np.all(rel <= tol), f"stage {stage}: relative differences {rel}"
        res.append(got)
        dev.step(10)
    dev.release()
    assert abs(res[1][0] - res[0][0]) / res[0][0] < 1e-6, "dry mass not conserved"
    assert abs(res[1][1] - res[0][1]) / res[0][1] < 1e-5, "theta budget not conserved"
