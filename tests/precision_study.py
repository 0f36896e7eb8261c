"""Precision study of the GPU step against the FP64 oracle (the -r8 analogue).

    python tests/precision_study.py [--steps 300] [--out FILE]

On the 41x41x40 supercell case (dx = 2 km, dt = 12 s, open boundaries) it advances the
same seeded state with
  * the GPU, FP32 form (precision_mode 0),
  * the GPU, precision form (precision_mode 1: FP64 tridiagonal coefficients and solve,
    compensated column mass-divergence sums),
  * the FP32 oracle (reference-order FP32, libm LOG/**),
  * the FP64 oracle,
and reports max-abs and relative-L2 errors against FP64 for u, v, w, theta, mu, p and the
domain dry-mass drift, at the requested step count (300 steps = 1 h).
Test infrastructure: it imports the oracle, and needs a GPU."""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), HERE]

import numpy as np  # noqa: E402

from parity import interior, rel_l2  # noqa: E402
from pyoracle import Oracle  # noqa: E402
from wrf_amd import grid, runtime  # noqa: E402

FIELDS = (("u_2", "u", "u"), ("v_2", "v", "v"), ("w_2", "w", "w"), ("t_2", "mass", "theta'"),
          ("mu_2", "mass", "mu'"), ("p", "mass", "p'"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    hs0 = grid.build_initial_state(grid.supercell_c0())
    hs1 = grid.build_initial_state(grid.supercell_c0(precision_mode=1))
    d = hs0.dims
    runs = {}
    t0 = time.perf_counter()
    for name, hs in (("gpu_fp32", hs0), ("gpu_precision", hs1)):
        dev = runtime.DeviceDomain(hs)
        done = 0
        while done < args.steps:
            n = min(50, args.steps - done)
            dev.step(n)
            done += n
            print(f"{name}: {done} steps", flush=True)
        runs[name] = dev.download()
    print(f"GPU runs {time.perf_counter() - t0:.1f} s", flush=True)
    for name, prec in (("oracle_fp32", "f32"), ("oracle_fp64", "f64")):
        t0 = time.perf_counter()
        sess = Oracle(prec).session(hs0)
        done = 0
        while done < args.steps:
            n = min(50, args.steps - done)
            sess.step(n)
            done += n
            print(f"{name}: {done} steps ({time.perf_counter() - t0:.0f} s)", flush=True)
        runs[name] = sess.arrays
    ref = runs["oracle_fp64"]
    mub = interior(hs0.a2["mub"], d).astype(np.float64)
    m0 = float(np.sum(interior(hs0.a2["mu_2"], d).astype(np.float64)))
    lines = [f"# Precision study: 41x41x40 supercell, {args.steps} RK3 steps (dt = 12 s), errors against the FP64 oracle",
             "",
             "| run | field | max abs | rel L2 |", "|---|---|---|---|"]
    for run in ("gpu_fp32", "gpu_precision", "oracle_fp32"):
        for key, kind, label in FIELDS:
            a = interior(runs[run][key], d, kind).astype(np.float64)
            b = interior(ref[key], d, kind).astype(np.float64)
            lines.append(f"| {run} | {label} | {np.max(np.abs(a - b)):.3e} | {rel_l2(a, b):.3e} |")
    lines += ["", "Dry-mass drift, sum(mu') change relative to sum(mub) over the run (open boundaries let mass through):", ""]
    for run in ("gpu_fp32", "gpu_precision", "oracle_fp32", "oracle_fp64"):
        m1 = float(np.sum(interior(runs[run]["mu_2"], d).astype(np.float64)))
        lines.append(f"- {run}: {(m1 - m0) / float(np.sum(mub)):.3e}")
    text = "\n".join(lines) + "\n"
    print(text)
    if args.out:
        open(args.out, "w").write(text)


if __name__ == "__main__":
    main()
