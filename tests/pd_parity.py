"""Parity of the positive-definite entry points with the reference's argument lists
(wrfmi_advect_scalar_pd, wrfmi_rk_update_scalar_pd) -- test infrastructure, shared by the
CPU-emulated library's checks and the GPU tests.

Arrays are host numpy arrays registered with the library (wrfmi_register_field), so the same
code drives the emulated and the GPU library: the library works on its device mirrors and
wrfmi_sync_to_host brings the results back.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from wrf_amd import grid

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
MSF6 = ("msfux", "msfuy", "msfvx", "msfvy", "msftx", "msfty")
VP = ctypes.c_void_p


class _Reg:
    """Registers host arrays with `lib` for the duration of a with-block."""

    def __init__(self, lib, arrays):
        self.lib, self.arrays = lib, [a for a in arrays if isinstance(a, np.ndarray) and a.ndim > 0]
        for f in (lib.wrfmi_register_field, lib.wrfmi_unregister_field, lib.wrfmi_sync_to_host):
            f.restype = ctypes.c_int
        lib.wrfmi_register_field.argtypes = [VP, ctypes.c_size_t]
        lib.wrfmi_unregister_field.argtypes = [VP]
        lib.wrfmi_sync_to_host.argtypes = [VP]
        lib.wrfmi_last_error.restype = ctypes.c_char_p

    def __enter__(self):
        for a in self.arrays:
            self.chk(self.lib.wrfmi_register_field(a.ctypes.data, a.nbytes))
        return self

    def chk(self, rc, what="call"):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.lib.wrfmi_last_error().decode()}")

    def sync(self, a):
        self.chk(self.lib.wrfmi_sync_to_host(a.ctypes.data), "sync")

    def __exit__(self, *exc):
        for a in self.arrays:
            self.lib.wrfmi_unregister_field(VP(a.ctypes.data))


def _refs(keep, vals, kind):
    out = []
    for v in vals:
        c = ctypes.c_int(int(v)) if kind == "i" else ctypes.c_float(float(v))
        keep.append(c)
        out.append(ctypes.byref(c))
    return out


def _dims_refs(keep, d):
    return _refs(keep, [getattr(d, n) for n, _ in d._fields_], "i")


def advect_scalar_pd_call(lib, a, d, cfg, ts, tend):
    """wrfmi_advect_scalar_pd on the arrays `a` (pd_advect_inputs' names) adding into `tend`;
    returns the tendency after the call."""
    keep = []
    arrays = [a[k] for k in ("field", "field_old", "ru", "rv", "rom", "c1", "c2", "mut", "mub", "mu_old", "fzm", "fzp",
                             "rdzw") + MSF6] + [tend]
    with _Reg(lib, arrays) as R:
        p = lambda x: VP(x.ctypes.data)  # noqa: E731
        fn = lib.wrfmi_advect_scalar_pd
        fn.restype = ctypes.c_int
        ii = lambda v: _refs(keep, [v], "i")[0]  # noqa: E731
        ff = lambda v: _refs(keep, [v], "f")[0]  # noqa: E731
        args = [p(a["field"]), p(a["field_old"]), p(tend), p(tend), p(tend), p(a["ru"]), p(a["rv"]), p(a["rom"]),
                p(a["c1"]), p(a["c2"]), p(a["mut"]), p(a["mub"]), p(a["mu_old"]), ii(ts), ctypes.byref(cfg), ii(0)]
        args += [p(a[m]) for m in MSF6] + [p(a["fzm"]), p(a["fzp"]), ff(a["rdx"]), ff(a["rdy"]), p(a["rdzw"]),
                                           ff(a["dt"])] + _dims_refs(keep, d)
        R.chk(fn(*args), "wrfmi_advect_scalar_pd")
        R.sync(tend)
    return tend


def oracle_advect_scalar_pd(oracle, a, d, cfg, ts, tend):
    """The FP32 oracle's advect_scalar_pd on the same arguments, adding into `tend`."""
    L = oracle.lib
    L.o_advect_scalar_pd.argtypes = ([VP] * 13 + [ctypes.c_int, VP, ctypes.c_int] + [VP] * 8
                                     + [ctypes.c_float, ctypes.c_float, VP, ctypes.c_float, VP])
    L.o_advect_scalar_pd.restype = ctypes.c_int
    p = lambda x: VP(x.ctypes.data)  # noqa: E731
    rc = L.o_advect_scalar_pd(p(a["field"]), p(a["field_old"]), p(tend), None, None, p(a["ru"]), p(a["rv"]),
                              p(a["rom"]), p(a["c1"]), p(a["c2"]), p(a["mut"]), p(a["mub"]), p(a["mu_old"]), ts,
                              ctypes.byref(cfg), 0, *[p(a[m]) for m in MSF6], p(a["fzm"]), p(a["fzp"]),
                              float(a["rdx"]), float(a["rdy"]), p(a["rdzw"]), float(a["dt"]), ctypes.byref(d))
    assert rc == 0
    return tend


def fixture_failures(lib, oracle):
    """wrfmi_advect_scalar_pd against the reference's own advect_scalar_pd (the committed
    fixture tests/golden/advect_scalar_pd_ref.npz), bit for bit from a zero tendency, and
    against the FP32 oracle (pinned to that routine) from a non-zero one, which the routine
    adds to as the reference does (tendency = tendency - ...)."""
    g = np.load(os.path.join(GOLDEN, "advect_scalar_pd_ref.npz"))
    fails = []
    for n in range(4):
        d = grid.Dims(*[int(x) for x in g[f"c{n}_dims"]])
        cfg = grid.make_config(grid.DomainSpec(bc="periodic" if int(g[f"c{n}_bc"]) == 0 else "open"))
        cfg.moist_adv_opt = cfg.scalar_adv_opt = 1
        a = {k[len(f"c{n}_"):]: (np.ascontiguousarray(v) if v.ndim else v[()]) for k, v in g.items()
             if k.startswith(f"c{n}_")}
        exp = a["tendency"]
        got = advect_scalar_pd_call(lib, a, d, cfg, int(a["time_step"]), np.zeros_like(exp))
        if not np.array_equal(got.view(np.int32), exp.view(np.int32)):
            fails.append(f"fixture case {n}: {int(np.count_nonzero(got != exp))} points differ")
        base = np.random.default_rng(n).uniform(-1e-3, 1e-3, exp.shape).astype(np.float32)
        got2 = advect_scalar_pd_call(lib, a, d, cfg, int(a["time_step"]), base.copy())
        exp2 = oracle_advect_scalar_pd(oracle, a, d, cfg, int(a["time_step"]), base.copy())
        if not np.array_equal(got2.view(np.int32), exp2.view(np.int32)):
            fails.append(f"fixture case {n} from a non-zero tendency: {int(np.count_nonzero(got2 != exp2))} points differ")
    return fails


def update_pd_failures(lib, oracle):
    """wrfmi_rk_update_scalar_pd against the oracle's rk_update_scalar_pd (module_em.F:
    1803-1916) on seeded arrays, with distinct old and new masses and two species."""
    spec = grid.DomainSpec(nx=11, ny=9, nz=7, bc="periodic", pad_i=1)
    d, cfg = grid.make_dims(spec), grid.make_config(spec)
    s3, s2, s1 = grid.shape3(d), grid.shape2(d), grid.shape1(d)
    rng = np.random.default_rng(61)
    u = lambda sh, lo, hi: rng.uniform(lo, hi, sh).astype(np.float32)  # noqa: E731
    nsp = 2
    scalar = u((nsp,) + s3, 0.0, 1e-2)
    sc_tend = u((nsp,) + s3, -1e-2, 1e-2)
    c1, c2 = u(s1, 0.6, 1.0), u(s1, 0.0, 5e2)
    mu_old, mu_new, mu_base = u(s2, 1e3, 3e3), u(s2, 1e3, 3e3), u(s2, 8.6e4, 9.0e4)
    dt = np.float32(6.0)
    exp_s, exp_t = scalar.copy(), sc_tend.copy()
    L = oracle.lib
    L.o_rk_update_scalar_pd.argtypes = [VP] * 7 + [ctypes.c_float, VP]
    L.o_rk_update_scalar_pd.restype = None
    p = lambda x: VP(x.ctypes.data)  # noqa: E731
    for m in range(nsp):
        L.o_rk_update_scalar_pd(p(exp_s[m]), p(exp_t[m]), p(c1), p(c2), p(mu_old), p(mu_new), p(mu_base), float(dt),
                                ctypes.byref(d))
    keep = []
    with _Reg(lib, [scalar, sc_tend, c1, c2, mu_old, mu_new, mu_base]) as R:
        fn = lib.wrfmi_rk_update_scalar_pd
        fn.restype = ctypes.c_int
        args = _refs(keep, [1, nsp], "i") + [p(scalar), p(sc_tend), p(c1), p(c2), p(mu_old), p(mu_new), p(mu_base)]
        args += _refs(keep, [3], "i") + _refs(keep, [dt], "f") + _refs(keep, [1], "i") + [ctypes.byref(cfg)]
        R.chk(fn(*(args + _dims_refs(keep, d))), "wrfmi_rk_update_scalar_pd")
        R.sync(scalar)
        R.sync(sc_tend)
    fails = []
    if not np.array_equal(scalar.view(np.int32), exp_s.view(np.int32)):
        fails.append(f"rk_update_scalar_pd: {int(np.count_nonzero(scalar != exp_s))} scalar points differ")
    if not np.array_equal(sc_tend.view(np.int32), exp_t.view(np.int32)):
        fails.append("rk_update_scalar_pd: sc_tend differs (not zeroed where the reference zeroes it)")
    return fails


def pd_state(case, moist_opt: int = 1, scalar_opt: int = 1, seed: int = 5):
    """A whole-step state for the positive-definite path from a DomainSpec or its keyword
    dict: the passive scalars emptied in about half of the cells (so that the limiter
    engages) and seeded physics tendencies in q_tend (which rk_update_scalar_pd takes in on
    the last substep): signed for the moist species, non-negative for the passive scalars,
    whose values the transport then keeps non-negative (a negative physics tendency can
    empty a cell below zero by itself, which no transport scheme prevents)."""
    hs = grid.build_initial_state(case if isinstance(case, grid.DomainSpec) else grid.DomainSpec(**case))
    hs.cfg.moist_adv_opt, hs.cfg.scalar_adv_opt = moist_opt, scalar_opt
    rng = np.random.default_rng(seed)
    per = (bool(hs.cfg.periodic_x), bool(hs.cfg.periodic_y))
    for m in range(hs.n_species):
        if m >= hs.n_moist:
            q = hs.q[m]
            q *= (rng.uniform(0.0, 1.0, q.shape) < 0.5)
            grid._halo_fill_np(q, hs.dims, per, False, False)
            hs.q_old[m][...] = q
        lo = 0.0 if m >= hs.n_moist else -1e-3
        hs.q_tend[m][...] = rng.uniform(lo, 1e-3, hs.q_tend[m].shape).astype(np.float32)
    return hs
