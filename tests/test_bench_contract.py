"""bench.py's contract, checked without a GPU: the algorithmic bytes per mass-point step
of the benchmark workloads, and the refusal of more than one GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.mark.parametrize("n_sound,n_species,expected", [
    (6, 7, 3684),   # 400x400x80, qv + 6 scalars
    (6, 1, 3396),   # 41x41x40 and the 64-member ensemble, qv only
    (6, 11, 3876),  # 1200x1200x80 terrain case, qv + 10 scalars
])
def test_algorithmic_bytes_per_point_step(n_sound, n_species, expected):
    # 4 * (327 + 12 n_s + 51 N_ac) with N_ac = 1 + n_sound/2 + n_sound acoustic substeps
    assert bench.b_alg(n_sound, n_species) == expected


def test_dominant_kernel_field_counts_cover_the_step_kernels():
    assert {"k_ac_wsolve", "k_ac_theta", "k_advance_uv", "k_tend_u", "k_adv_scalar"} <= set(bench.KERNEL_FIELDS)
    assert all(isinstance(v, int) and v > 0 for v in bench.KERNEL_FIELDS.values())


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_cpu_baseline_reports_threads_tiles_and_median():
    from wrf_amd import grid
    hs = grid.build_initial_state(grid.large_c1(nx=24, ny=20, nz=10, n_bubbles=1))
    cb = bench.cpu_baseline(hs, runs=3)
    assert cb["kind"] == "port" and cb["unit"] == "Mpoint-RK3-steps/s"
    assert cb["cores"] == cb["omp_threads"] >= 1 and cb["tiles"] >= 1
    assert len(cb["seconds_per_step"]) == 3 and cb["value"] > 0
    assert "24x20x10" in cb["sample"] and cb["cpu_model"]
