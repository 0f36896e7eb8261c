"""Launch-shape autotuner of the MI355X dynamics step.

For one domain shape it measures, kernel by kernel, the candidate launch shapes the library
offers (rows per workgroup of the four-point and column kernels -- up to 8 for the 5th-order
stencil kernels, 4 for the others --, levels per thread of the
four-point advance_uv and column divergence, prefetch depth of the column solve; see
wrfmi_tuning_set in include/wrfmi.h) with the library's own per-kernel HIP-event timing, keeps
the fastest for each kernel in turn (coordinate descent, starting from the current table), and
writes the table that wrfmi_init loads (wrf_amd/wrfmi_tuning.txt beside the library, or
$WRFMI_TUNING_FILE).  Entries of other domain shapes already in the file are kept.  Every shape
gives the same results bit for bit, so tuning changes only speed.

    python -m wrf_amd.autotune --preset large_c1 [--steps 3] [--out wrf_amd/wrfmi_tuning.txt]
    python -m wrf_amd.autotune --nx 256 --ny 192 --nz 50 --bc open

This is the counterpart of the reference's tile decomposition (frame/module_tiles.F:200-460),
which splits a patch into OpenMP tiles by a namelist count; here the shapes are measured.
"""
from __future__ import annotations

import argparse
import ctypes
import os
import sys

from . import grid, runtime

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "wrfmi_tuning.txt")

ROWS = (1, 2, 4)
TALL_ROWS = (1, 2, 4, 8)  # the four-point stencil kernels take workgroups of up to 8 rows
TALL = {"k_tend_u", "k_tend_v", "k_tend_w", "k_tend_t", "k_adv_scalar"}
LEVELS = (2, 4, 8)
PREFETCH = (2, 4, 8)
# the tunable kernels: which parameter besides rows each has
KERNELS = {
    "k_step_prep": None, "k_ww_scan": None, "k_tend_u": None, "k_tend_v": None, "k_tend_w": None,
    "k_tend_t": None, "k_small_prep": None, "k_p_rho0": None, "k_coef_w": None,
    "k_advance_uv": "levels", "k_ac_div": "levels", "k_ac_mu_ww": None, "k_ac_theta": None,
    "k_ac_wsolve": "prefetch", "k_small_finish": None, "k_adv_scalar": None, "k_update_scalar": None,
    "k_p_rho_phi": None,
}


def _bind(lib):
    cp, ci = ctypes.c_char_p, ctypes.c_int
    lib.wrfmi_tuning_set.argtypes = [cp, ci, ci, ci, ci, ci, ci]
    lib.wrfmi_tuning_load.argtypes = [cp]
    lib.wrfmi_tuning_get.argtypes = [cp, ci, ci, ci] + [ctypes.POINTER(ci)] * 3


def current_shape(lib, kernel, nx, ny, nz):
    r, l, p = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    runtime.check(lib.wrfmi_tuning_get(kernel.encode(), nx, ny, nz, ctypes.byref(r), ctypes.byref(l), ctypes.byref(p)))
    return r.value, l.value, p.value


def measure(dev, kernel, steps):
    """Average ms per launch of `kernel` over `steps` eager steps (after one untimed)."""
    dev.step(1)
    with runtime.KernelTimer([kernel]) as t:
        dev.step(steps)
        n, ms = t.collect()[kernel]
    return ms / n if n else float("inf")


def tune(spec, steps=3, log=print):
    """Returns {kernel: (rows, levels, prefetch)} for the domain of `spec`."""
    hs = grid.build_initial_state(spec)
    dev = runtime.DeviceDomain(hs)
    lib = dev.lib
    _bind(lib)
    nx, ny, nz = spec.nx, spec.ny, spec.nz
    best = {}
    for kernel, extra in KERNELS.items():
        rows0, lev0, pf0 = current_shape(lib, kernel, nx, ny, nz)
        cands = [(r, lev0, pf0) for r in (TALL_ROWS if kernel in TALL else ROWS)]
        if extra == "levels":
            cands += [(rows0, lv, pf0) for lv in LEVELS]
        if extra == "prefetch":
            cands += [(rows0, lev0, pf) for pf in PREFETCH]
        times = {}
        for c in dict.fromkeys(cands):
            runtime.check(lib.wrfmi_tuning_set(kernel.encode(), nx, ny, nz, *c))
            times[c] = measure(dev, kernel, steps)
        if not any(t < float("inf") for t in times.values()):
            runtime.check(lib.wrfmi_tuning_set(kernel.encode(), nx, ny, nz, rows0, lev0, pf0))
            log(f"{kernel}: not launched on this domain, shape kept")
            continue
        c = min(times, key=times.get)
        if extra:  # the parameter besides rows is chosen at the best row count as well
            r = c[0]
            for v in (LEVELS if extra == "levels" else PREFETCH):
                cc = (r, v, c[2]) if extra == "levels" else (r, c[1], v)
                if cc not in times:
                    runtime.check(lib.wrfmi_tuning_set(kernel.encode(), nx, ny, nz, *cc))
                    times[cc] = measure(dev, kernel, steps)
            c = min(times, key=times.get)
        runtime.check(lib.wrfmi_tuning_set(kernel.encode(), nx, ny, nz, *c))
        best[kernel] = c
        log(f"{kernel}: " + ", ".join(f"{k}={v * 1e3:.1f}us" for k, v in sorted(times.items())) + f" -> {c}")
    dev.release()
    return best


def write_table(path, spec, best, header=""):
    """Replace the entries of this domain shape in `path` (other shapes are kept)."""
    keep = []
    if os.path.exists(path):
        for line in open(path):
            f = line.split("#")[0].split()
            if len(f) == 7 and (int(f[1]), int(f[2]), int(f[3])) == (spec.nx, spec.ny, spec.nz):
                continue
            keep.append(line.rstrip("\n"))
    lines = keep or ["# launch shapes measured by wrf_amd/autotune.py: kernel nx ny nz rows levels prefetch"]
    if header:
        lines.append(f"# {header}")
    for k, (r, lv, pf) in best.items():
        lines.append(f"{k} {spec.nx} {spec.ny} {spec.nz} {r} {lv} {pf}")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


PRESETS = {"large_c1": grid.large_c1, "supercell_c0": grid.supercell_c0}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--preset", choices=sorted(PRESETS))
    ap.add_argument("--nx", type=int)
    ap.add_argument("--ny", type=int)
    ap.add_argument("--nz", type=int)
    ap.add_argument("--bc", default="periodic")
    ap.add_argument("--n-passive", type=int, default=6)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--out", default=os.environ.get("WRFMI_TUNING_FILE") or DEFAULT_TABLE)
    a = ap.parse_args(argv)
    if a.preset:
        spec = PRESETS[a.preset]()
    elif a.nx and a.ny and a.nz:
        spec = grid.DomainSpec(nx=a.nx, ny=a.ny, nz=a.nz, bc=a.bc, n_passive=a.n_passive, n_bubbles=2, seed=7)
    else:
        ap.error("give --preset or --nx/--ny/--nz")
    log = lambda m: print(m, flush=True)  # noqa: E731
    log(f"tuning {spec.nx}x{spec.ny}x{spec.nz} ({a.bc if not a.preset else a.preset}), {a.steps} steps per candidate")
    best = tune(spec, steps=a.steps, log=log)
    write_table(a.out, spec, best, header=f"{spec.nx}x{spec.ny}x{spec.nz}: {a.preset or a.bc}, {a.steps} timed steps")
    log(f"wrote {a.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
