"""Writes tests/golden/advect_scalar_pd_ref.npz: inputs and outputs of the reference's own
advect_scalar_pd (dyn_em/module_advect_em.F:6069-7885, compiled in place by
`make -C oracle ref` into oracle/_ref/libref_advect.so) on small seeded states.

    python tests/golden/make_advect_scalar_pd_golden.py

Cases 0-3: periodic and open lateral boundaries, each with time_step > 0 and < 0, with a
field_old distinct from field as the last RK substep passes them (dyn_em/module_em.F:1264-
1280; field_old is the time-n value, field the substep's).  The inputs are a non-negative
tracer in blobs with many zero cells, so that the positive-definite limiter (7681-7779)
engages.  The tendency starts at zero (module_em.F:1248); h_tendency and z_tendency are
requested (tenddec), so the fixture holds the split of the tendency as well."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.dirname(HERE)]

from parity import MSF6, pd_advect_inputs  # noqa: E402
from pyoracle import RefAdvect  # noqa: E402

CASES = [("periodic", 12, 31), ("periodic", -12, 32), ("open", 12, 33), ("open", -6, 34)]
SHAPE = (9, 8, 6)


def case_arrays(bc, ts, seed):
    d, c, a = pd_advect_inputs(bc, *SHAPE, seed)
    c.time_step = ts
    s3 = a["field"].shape
    tend, h, z = (np.zeros(s3, np.float32) for _ in range(3))
    RefAdvect().advect_scalar_pd(a["field"], a["field_old"], tend, h, z, a["ru"], a["rv"], a["rom"], a["c1"], a["c2"],
                                 a["mut"], a["mub"], a["mu_old"], ts, c, 1, {m: a[m] for m in MSF6}, a["fzm"], a["fzp"],
                                 float(a["rdx"]), float(a["rdy"]), a["rdzw"], float(a["dt"]), d)
    out = dict(a)
    out.update(tendency=tend, h_tendency=h, z_tendency=z, time_step=np.int32(ts),
               dims=np.array([getattr(d, n) for n, _ in d._fields_], np.int32),
               bc=np.array(0 if bc == "periodic" else 1, np.int32))
    return out


def main():
    arrays = {}
    for n, (bc, ts, seed) in enumerate(CASES):
        for k, v in case_arrays(bc, ts, seed).items():
            arrays[f"c{n}_{k}"] = v
    path = os.path.join(HERE, "advect_scalar_pd_ref.npz")
    np.savez_compressed(path, **arrays)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
