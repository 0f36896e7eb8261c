"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle_f32.so (reference-order FP32 restatement),
liboracle_f32cr.so (the same with the EOS's LOG and ** rounded once from double),
liboracle_f64.so (same source in FP64, the -r8 analogue), and the reference's own
advect_scalar and advect_scalar_pd from oracle/_ref/libref_advect.so when that was built.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from wrf_amd.fields import (FIELDS_1D, FIELDS_2D, FIELDS_3D, MAX_SPECIES, SCALARS_R, State32, State64,
                             make_lbc_struct)
from wrf_amd.grid import lbc_struct

HERE = os.path.dirname(os.path.abspath(__file__))


def _load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
    return ctypes.CDLL(path)


class Oracle:
    def __init__(self, precision: str = "f32", path: str | None = None):
        """`path` loads another build of the same sources (the mutation tests' builds)."""
        self.precision = precision
        libs = {"f32": "liboracle_f32.so", "f32cr": "liboracle_f32cr.so", "f64": "liboracle_f64.so"}
        self.lib = ctypes.CDLL(path) if path else _load(libs[precision])
        single = precision != "f64"
        self.dtype = np.float32 if single else np.float64
        self.ctype = ctypes.c_float if single else ctypes.c_double
        self.State = State32 if single else State64
        self.Lbc = make_lbc_struct(self.ctype, "OLbc32" if single else "OLbc64")
        self.lib.oracle_state_sizeof.restype = ctypes.c_size_t
        if self.lib.oracle_state_sizeof() != ctypes.sizeof(self.State):
            raise RuntimeError("oracle state layout does not match wrf_amd.fields")
        self.lib.oracle_solve_em_steps.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]

    # ------------------------------------------------------------------ state plumbing
    def to_native(self, hs):
        """Copy a wrf_amd.grid.HostState into oracle-precision arrays; returns (arrays, struct)."""
        arrs = {k: np.array(v, dtype=self.dtype, copy=True, order="C") for k, v in hs.arrays().items()}
        st = self.State()
        P = ctypes.POINTER(self.ctype)
        for n in FIELDS_3D + FIELDS_2D + FIELDS_1D:
            setattr(st, n, arrs[n].ctypes.data_as(P))
        for n in SCALARS_R:
            setattr(st, n, hs.scal[n])
        st.a = arrs["a"].ctypes.data
        st.alpha = arrs["alpha"].ctypes.data
        st.gamma = arrs["gamma"].ctypes.data
        st.n_moist, st.n_species = hs.n_moist, hs.n_species
        for m in range(hs.n_species):
            st.q[m] = arrs[f"q{m}"].ctypes.data_as(P)
            st.q_old[m] = arrs[f"q_old{m}"].ctypes.data_as(P)
            st.q_tend[m] = arrs[f"q_tend{m}"].ctypes.data_as(P)
        if hs.lbc is not None:
            # the olbc struct lives as long as the state struct (kept on it)
            st._lbc = lbc_struct(hs.lbc, self.Lbc, lambda n: arrs[n].ctypes.data, self.ctype)
            st.lbc = ctypes.addressof(st._lbc)
        return arrs, st

    def session(self, hs):
        """A persistent oracle state (copy of `hs`) advanced routine by routine."""
        return OracleSession(self, hs)

    def run(self, hs, nsteps: int = 1):
        """Advance a copy of `hs` by `nsteps` RK3 steps; returns the arrays dict."""
        arrs, st = self.to_native(hs)
        rc = self.lib.oracle_solve_em_steps(ctypes.byref(st), ctypes.byref(hs.cfg), ctypes.byref(hs.dims), nsteps)
        if rc != 0:
            raise RuntimeError(f"oracle solve_em step failed with status {rc}")
        return arrs


class OracleSession:
    """Oracle arrays + native struct kept alive between per-routine calls
    (oracle_<name> in solve_step.c, same names as the wrfmi_<name> entry points)."""

    _ARGS = {"rk_step_prep": "i", "rk_tendency": "i", "small_step_prep": "id", "advance_uv": "d",
             "advance_mu_t_w": "dii", "small_step_finish": "iid", "rk_scalar_step": "iid",
             "rk_scalar_pd_prep": "d", "calc_p_rho_phi": "", "phys_bc_dry_2": "i", "set_w_surface": ""}

    def __init__(self, oracle: Oracle, hs):
        self.o = oracle
        self.cfg, self.dims = hs.cfg, hs.dims
        self.arrays, self.st = oracle.to_native(hs)
        for name, kinds in self._ARGS.items():
            fn = getattr(oracle.lib, "oracle_" + name)
            fn.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int if k == "i" else ctypes.c_double for k in kinds]
            fn.restype = ctypes.c_int

    # the library's state-struct form of advance_uv is wrfmi_advance_uv_state (its
    # wrfmi_advance_uv keeps the reference's argument list)
    _ALIAS = {"advance_uv_state": "advance_uv"}

    def phase(self, name: str, *args):
        name = self._ALIAS.get(name, name)
        rc = getattr(self.o.lib, "oracle_" + name)(ctypes.byref(self.st), ctypes.byref(self.cfg),
                                                   ctypes.byref(self.dims), *args)
        if rc != 0:
            raise RuntimeError(f"oracle {name} failed with status {rc}")

    def step(self, nsteps: int = 1):
        rc = self.o.lib.oracle_solve_em_steps(ctypes.byref(self.st), ctypes.byref(self.cfg), ctypes.byref(self.dims),
                                              nsteps)
        if rc != 0:
            raise RuntimeError(f"oracle solve_em step failed with status {rc}")


def _flags(cfg):
    return (ctypes.c_int * 15)(cfg.periodic_x, cfg.periodic_y, cfg.symmetric_xs, cfg.symmetric_xe, cfg.symmetric_ys,
                               cfg.symmetric_ye, cfg.open_xs, cfg.open_xe, cfg.open_ys, cfg.open_ye, cfg.specified,
                               cfg.nested, cfg.polar, cfg.h_sca_adv_order, cfg.v_sca_adv_order)


class RefAdvect:
    """The reference's own advect_scalar and advect_scalar_pd (compiled in place,
    ADVECT_KERNEL mode)."""

    def __init__(self):
        self.lib = _load(os.path.join("_ref", "libref_advect.so"))
        self.fn = self.lib.ref_advect_scalar
        self.fn.restype = None
        self.fn_pd = self.lib.ref_advect_scalar_pd
        self.fn_pd.restype = None

    def advect_scalar_pd(self, field, field_old, tend, h_tend, z_tend, ru, rv, rom, c1, c2, mut, mub, mu_old,
                         time_step, cfg, tenddec, msf, fzm, fzp, rdx, rdy, rdzw, dt, dims):
        d = (ctypes.c_int * 18)(*[getattr(dims, n) for n, _ in dims._fields_])
        f = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        self.fn_pd(f(field), f(field_old), f(tend), f(h_tend), f(z_tend), f(ru), f(rv), f(rom), f(c1), f(c2), f(mut),
                   f(mub), f(mu_old), ctypes.c_int(time_step), _flags(cfg), ctypes.c_int(int(tenddec)),
                   f(msf["msfux"]), f(msf["msfuy"]), f(msf["msfvx"]), f(msf["msfvy"]), f(msf["msftx"]),
                   f(msf["msfty"]), f(fzm), f(fzp), ctypes.c_float(rdx), ctypes.c_float(rdy), f(rdzw),
                   ctypes.c_float(dt), d)

    def advect_scalar(self, field, field_old, tend, ru, rv, rom, c1, c2, mut, time_step, cfg, msf, fzm, fzp,
                      rdx, rdy, rdzw, dims):
        flags = _flags(cfg)
        d = (ctypes.c_int * 18)(*[getattr(dims, n) for n, _ in dims._fields_])
        f = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        self.fn(f(field), f(field_old), f(tend), f(ru), f(rv), f(rom), f(c1), f(c2), f(mut),
                ctypes.c_int(time_step), flags, f(msf["msfux"]), f(msf["msfuy"]), f(msf["msfvx"]), f(msf["msfvy"]),
                f(msf["msftx"]), f(msf["msfty"]), f(fzm), f(fzp), ctypes.c_float(rdx), ctypes.c_float(rdy),
                f(rdzw), d)
