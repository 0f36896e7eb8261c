"""Shared comparison helpers of the parity tests (test infrastructure)."""
from __future__ import annotations

import numpy as np

# The reference evaluates LOG and ** (vspow) in FP32 libm; the GPU evaluates them in
# double and rounds once.  Against the libm oracle, outputs of calc_p_rho_phi may differ by
# a few ulp of the full (base + perturbation) value at isolated points: 1 ulp of LOG moves
# alpha by up to ~1 ulp, which ** (exponent cp/cv = 1.4) and the product with p0 carry into
# p at up to ~8 ulp.  Against the once-rounded oracle the bound is the rare case where a
# double result rounds differently to float on the two sides.
EOS_FIELDS = {"p": "pb", "al": "alb"}
EOS_ULPS = 8


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sqrt(np.sum(b * b))
    num = np.sqrt(np.sum((a - b) ** 2))
    return float(num / den) if den > 0 else float(num)


def interior(a, d, kind="mass", nk=None):
    """Domain part of a (nj, nk, ni) or (nj, ni) array: i = ids..ide(-1), j = jds..jde(-1)."""
    i0, j0 = d.ids - d.ims, d.jds - d.jms
    ni = d.ide - d.ids + (1 if kind == "u" else 0)
    nj = d.jde - d.jds + (1 if kind == "v" else 0)
    if a.ndim == 3:
        nkk = nk if nk is not None else (d.kde if kind == "w" else d.kde - 1)
        return a[j0:j0 + nj, :nkk, i0:i0 + ni]
    return a[j0:j0 + nj, i0:i0 + ni]


def compare_exact(got: dict, exp: dict, names, eos_tol: bool = False, base: dict | None = None, ulps: int = EOS_ULPS):
    """Bitwise comparison; returns a list of failure descriptions."""
    fails = []
    for k in names:
        g, e = got[k], exp[k]
        if g.shape != e.shape:
            fails.append(f"{k}: shape {g.shape} vs {e.shape}")
            continue
        bad = g.view(np.uint8).reshape(g.shape + (-1,)) != e.view(np.uint8).reshape(e.shape + (-1,))
        bad = bad.any(axis=-1)
        # +0 and -0 compare equal in value; keep bitwise strictness except for signed zeros
        bad &= ~((g == 0) & (e == 0))
        if eos_tol and k in EOS_FIELDS and base is not None:
            full = np.abs(e.astype(np.float64) + base[EOS_FIELDS[k]].astype(np.float64))
            ulp = np.spacing(full.astype(np.float32)).astype(np.float64)
            bad &= np.abs(g.astype(np.float64) - e.astype(np.float64)) > ulps * ulp
        n = int(bad.sum())
        if n:
            idx = tuple(int(x[0]) for x in np.nonzero(bad))
            fails.append(f"{k}: {n} values differ, first at [j,k,i]={idx} got {g[idx]!r} expected {e[idx]!r}, "
                         f"max|diff|={float(np.max(np.abs(g.astype(np.float64) - e.astype(np.float64))[bad])):.3e}")
    return fails


def phase_parity(backend, session, phases, nsteps: int = 1, stop_at_first: bool = True, eos_ulps: int = EOS_ULPS):
    """Run the per-routine calls on the backend and on the oracle from identical inputs
    (the oracle's state is uploaded before every call) and compare every array bitwise.
    Returns a list of failure descriptions."""
    names = list(session.arrays)
    fails = []
    for step in range(nsteps):
        for n, (name, args) in enumerate(phases):
            backend.upload(session.arrays)
            backend.phase(name, *args)
            session.phase(name, *args)
            got = backend.download()
            f = compare_exact(got, session.arrays, names, eos_tol=(name == "calc_p_rho_phi"), base=session.arrays,
                              ulps=eos_ulps)
            if f:
                fails.append(f"step {step + 1} call {n} {name}{args}: " + "; ".join(f))
                if stop_at_first:
                    return fails
    return fails
