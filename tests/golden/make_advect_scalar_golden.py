"""Writes tests/golden/advect_scalar_ref.npz: inputs and outputs of the reference's own
advect_scalar (dyn_em/module_advect_em.F:3029-4359, compiled in place by
`make -C oracle ref` into oracle/_ref/libref_advect.so) on small seeded states.

    python tests/golden/make_advect_scalar_golden.py

Cases 0-3: periodic and open lateral boundaries, each with time_step > 0 and < 0, with
field_old = field as rk_scalar_tend and rk_tendency pass them (dyn_em/module_em.F:625,
1331-1332).  Cases 4-5: open boundaries with a distinct field_old (the open-boundary
inflow terms read it, module_advect_em.F:4104-4177).
The tendency starts at zero, as rk_scalar_tend leaves it before calling advect_scalar
(dyn_em/module_em.F:1248)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

from pyoracle import RefAdvect  # noqa: E402
from wrf_amd import grid  # noqa: E402

MSF = ("msfux", "msfuy", "msfvx", "msfvy", "msftx", "msfty")


def case_arrays(bc, ts, seed, same_old):
    spec = grid.DomainSpec(nx=9, ny=8, nz=6, bc=bc, pad_i=1, seed=seed)
    d = grid.make_dims(spec)
    c = grid.make_config(spec)
    c.time_step = ts
    s3, s2, s1 = grid.shape3(d), grid.shape2(d), grid.shape1(d)
    rng = np.random.default_rng(seed)
    u = lambda sh, lo, hi: rng.uniform(lo, hi, sh).astype(np.float32)
    out = dict(field=u(s3, 280, 320), field_old=u(s3, 280, 320), ru=u(s3, -3e4, 3e4), rv=u(s3, -3e4, 3e4),
               rom=u(s3, -50, 50), mut=u(s2, 9e4, 9.5e4), c1=u(s1, 0.5, 1.0), c2=u(s1, 0.0, 1e3))
    if same_old:
        out["field_old"] = out["field"].copy()
    for n in MSF:
        out[n] = u(s2, 0.95, 1.05)
    out["fzm"] = u(s1, 0.4, 0.6)
    out["fzp"] = (1 - out["fzm"]).astype(np.float32)
    out["rdzw"] = u(s1, -30, -10)
    out["rdx"] = np.float32(1 / 2000.)
    out["rdy"] = np.float32(1 / 2500.)
    tend = np.zeros(s3, np.float32)
    RefAdvect().advect_scalar(out["field"], out["field_old"], tend, out["ru"], out["rv"], out["rom"], out["c1"],
                              out["c2"], out["mut"], ts, c, {n: out[n] for n in MSF}, out["fzm"], out["fzp"],
                              float(out["rdx"]), float(out["rdy"]), out["rdzw"], d)
    out["tendency"] = tend
    out["time_step"] = np.int32(ts)
    out["dims"] = np.array([getattr(d, n) for n, _ in d._fields_], np.int32)
    return out


def main():
    arrays = {}
    cases = [("periodic", 12, 21, True), ("periodic", -12, 22, True), ("open", 12, 23, True), ("open", -6, 24, True),
             ("open", 12, 25, False), ("open", -12, 26, False)]
    for n, (bc, ts, seed, same_old) in enumerate(cases):
        for k, v in case_arrays(bc, ts, seed, same_old).items():
            arrays[f"c{n}_{k}"] = v
        arrays[f"c{n}_bc"] = np.array(0 if bc == "periodic" else 1, np.int32)
    path = os.path.join(HERE, "advect_scalar_ref.npz")
    np.savez_compressed(path, **arrays)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
