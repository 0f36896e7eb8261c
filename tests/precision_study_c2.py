"""Precision study on the 400x400x80 domain (dx = 1 km, dt = 6 s, doubly periodic, qv + 6
passive scalars, seed 1002): the GPU's FP32 form and precision form against the FP64 oracle
(the -r8 analogue) after 600 RK3 steps (1 h).  Test infrastructure: it runs the oracle.

The FP64 oracle takes several seconds per step even on many cores, so the study runs in two
stages, possibly on two machines:

    python tests/precision_study_c2.py reference --steps 600 --ref-out FILE.npz
        FP64 oracle (OpenMP over j-tiles, OMP_NUM_THREADS threads); saves the final u, v,
        w, theta', p' (as float32: their rounding is far below the differences measured)
        and mu' (float64) over the domain, and the domain dry-mass sums.
    python tests/precision_study_c2.py gpu --steps 600 --ref FILE.npz --out REPORT.md
        the GPU FP32 form (precision_mode 0, bit for bit the reference-order FP32
        restatement at every step) and the precision form (precision_mode 1: FP64
        tridiagonal coefficients and solve, compensated column sums) from the same seeded
        state; writes max-abs and relative-L2 errors against the FP64 result for u, v, w,
        theta, mu, p, and the domain dry-mass drift of each run.
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), HERE]

import numpy as np  # noqa: E402

from parity import interior, rel_l2  # noqa: E402
from wrf_amd import grid  # noqa: E402

FIELDS = (("u_2", "u", "u (m/s)"), ("v_2", "v", "v (m/s)"), ("w_2", "w", "w (m/s)"), ("t_2", "mass", "theta' (K)"),
          ("mu_2", "mass", "mu' (Pa)"), ("p", "mass", "p' (Pa)"))


def dry_mass(mu, d) -> float:
    """Domain sum of mu' (Pa): with doubly periodic boundaries the flux-form step conserves it."""
    return float(np.sum(interior(mu, d).astype(np.float64)))


def reference(args):
    from pyoracle import Oracle
    hs = grid.build_initial_state(grid.large_c1())
    d = hs.dims
    m0 = dry_mass(hs.a2["mu_2"], d)
    sess = Oracle("f64").session(hs)
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        n = min(10, args.steps - done)
        sess.step(n)
        done += n
        print(f"fp64 oracle: {done} steps ({time.perf_counter() - t0:.0f} s)", flush=True)
    out = {}
    for key, kind, _ in FIELDS:
        a = interior(sess.arrays[key], d, kind)
        out[key] = a.astype(np.float64 if key == "mu_2" else np.float32)
    out["dry_mass_0"] = np.float64(m0)
    out["dry_mass_1"] = np.float64(dry_mass(sess.arrays["mu_2"], d))
    out["mub_sum"] = np.float64(np.sum(interior(hs.a2["mub"], d).astype(np.float64)))
    out["steps"] = np.int64(args.steps)
    out["seconds"] = np.float64(time.perf_counter() - t0)
    np.savez(args.ref_out, **out)
    print(f"saved {args.ref_out}", flush=True)


def gpu(args):
    from wrf_amd import runtime
    ref = np.load(args.ref)
    if int(ref["steps"]) != args.steps:
        raise SystemExit(f"{args.ref} holds {int(ref['steps'])} steps, not {args.steps}")
    runs = {}
    d = None
    for name, mode in (("gpu_fp32", 0), ("gpu_precision", 1)):
        hs = grid.build_initial_state(grid.large_c1(precision_mode=mode))
        d = hs.dims
        dev = runtime.DeviceDomain(hs)
        t0 = time.perf_counter()
        done = 0
        while done < args.steps:
            n = min(100, args.steps - done)
            dev.step(n)
            done += n
            print(f"{name}: {done} steps ({time.perf_counter() - t0:.1f} s)", flush=True)
        a = dev.to_host([k for k, _, _ in FIELDS])
        runs[name] = {k: interior(a[k], d, kind) for k, kind, _ in FIELDS}
        assert dev.check_finite() == 0
        dev.release()
        del dev
    mub = float(ref["mub_sum"])
    m0 = float(ref["dry_mass_0"])
    lines = [f"# Precision study: 400x400x80, {args.steps} RK3 steps (dt = 6 s, 1 h), errors against the FP64 oracle", "",
             f"Command: `python tests/precision_study_c2.py reference --steps {args.steps} --ref-out REF.npz` "
             f"(FP64 oracle, {float(ref['seconds']):.0f} s), then "
             f"`python tests/precision_study_c2.py gpu --steps {args.steps} --ref REF.npz --out REPORT.md` on the MI355X.",
             "", "gpu_fp32 is precision_mode 0 (FP32 throughout, reference operation order: bit for bit the FP32 "
             "restatement at every step); gpu_precision is precision_mode 1 (FP64 calc_coef_w and tridiagonal "
             "solve, compensated column sums).", "",
             "| run | field | max abs | rel L2 |", "|---|---|---|---|"]
    for run in ("gpu_fp32", "gpu_precision"):
        for key, kind, label in FIELDS:
            a = runs[run][key].astype(np.float64)
            b = ref[key].astype(np.float64)
            lines.append(f"| {run} | {label} | {np.max(np.abs(a - b)):.3e} | {rel_l2(a, b):.3e} |")
    lines += ["", "Domain dry-mass drift, (sum(mu') after - before) / sum(mub) (doubly periodic: conserved up to "
              "rounding):", ""]
    for run in ("gpu_fp32", "gpu_precision"):
        m1 = float(np.sum(runs[run]["mu_2"].astype(np.float64)))
        lines.append(f"- {run}: {(m1 - m0) / mub:.3e}")
    lines.append(f"- oracle_fp64: {(float(ref['dry_mass_1']) - m0) / mub:.3e}")
    text = "\n".join(lines) + "\n"
    print(text, flush=True)
    if args.out:
        open(args.out, "w").write(text)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("stage", choices=("reference", "gpu"))
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--ref-out", default=None)
    ap.add_argument("--ref", default=None)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    if args.stage == "reference":
        if not args.ref_out:
            ap.error("--ref-out is required")
        reference(args)
    else:
        if not args.ref:
            ap.error("--ref is required")
        gpu(args)


if __name__ == "__main__":
    main()
