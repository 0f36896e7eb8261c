"""Launch shapes on the GPU (needs an MI355X): the autotuner (wrf_amd/autotune.py) measures
and writes a table for a small domain, the library loads it, and a step with the tuned shapes
(and with each offered shape set) is bit for bit the step with the default shapes."""
import ctypes

import numpy as np
import pytest

from wrf_amd import autotune, grid, runtime

pytestmark = pytest.mark.gpu

SPEC = dict(nx=64, ny=48, nz=20, bc="periodic", n_passive=2, n_bubbles=2, bubble_amp=2.0, seed=51)


def _step(hs):
    dev = runtime.DeviceDomain(hs)
    dev.step(1)
    out = dev.download()
    dev.release()
    return out


def test_tuned_shapes_keep_results(tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    lib = runtime.load_library()
    autotune._bind(lib)
    spec = grid.DomainSpec(**SPEC)
    hs = grid.build_initial_state(spec)
    print("launch shapes: default step", flush=True)
    runtime.check(lib.wrfmi_tuning_clear())
    want = _step(hs)
    print("launch shapes: autotuning a 64x48x20 domain", flush=True)
    best = autotune.tune(spec, steps=1, log=lambda m: None)
    path = str(tmp_path / "tuning.txt")
    autotune.write_table(path, spec, best)
    runtime.check(lib.wrfmi_tuning_clear())
    runtime.check(lib.wrfmi_tuning_load(path.encode()))
    for k, shape in best.items():
        assert autotune.current_shape(lib, k, spec.nx, spec.ny, spec.nz) == shape
    # (kernel, rows, levels or None, prefetch or None); None keeps the kernel's current value
    sets = [("tuned", None), ("rows 4", [(k, 4, None, None) for k in autotune.KERNELS]),
            ("levels 8 / prefetch 8", [("k_advance_uv", 2, 8, None), ("k_ac_div", 1, 8, None),
                                       ("k_ac_wsolve", 1, None, 8)])]
    for label, entries in sets:
        print(f"launch shapes: {label}", flush=True)
        if entries:
            for k, r, lv, pf in entries:
                _, lv0, pf0 = autotune.current_shape(lib, k, 0, 0, 0)
                runtime.check(lib.wrfmi_tuning_set(k.encode(), 0, 0, 0, r, lv0 if lv is None else lv,
                                                   pf0 if pf is None else pf))
        got = _step(hs)
        bad = [k for k in want if not np.array_equal(got[k], want[k])]
        assert not bad, f"{label}: the step differs from the default shapes' in {bad}"
    runtime.check(lib.wrfmi_tuning_clear())
    if autotune.os.path.exists(autotune.DEFAULT_TABLE):
        runtime.check(lib.wrfmi_tuning_load(autotune.DEFAULT_TABLE.encode()))
