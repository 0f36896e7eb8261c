"""Parity of the entry points that keep the reference's argument lists (include/wrfmi_routines.h)
with the oracle's routines of the same name (test infrastructure).

A backend is the CPU-emulated library (emu_backend.EmuDomain) or the GPU library
(wrf_amd.runtime.DeviceDomain): `lib`, `upload(arrays)`, `download()` and `addr(name)`, the
address to pass for the array `name` (a registered host array or a device array).  Each check
uploads the oracle's arrays, calls one entry point with the grid fields solve_em passes at its
call site, runs the oracle's routine on the same arrays and compares every array bit for bit.
"""
from __future__ import annotations

import ctypes

import numpy as np

from parity import compare_exact

VP = ctypes.c_void_p
DIMS = ("ids", "ide", "jds", "jde", "kds", "kde", "ims", "ime", "jms", "jme", "kms", "kme",
        "its", "ite", "jts", "jte", "kts", "kte")


class Caller:
    """Calls wrfmi_<routine> of `backend.lib` with pointers for every argument."""

    def __init__(self, backend, cfg, dims, scal):
        self.b, self.cfg, self.dims, self.scal = backend, cfg, dims, scal
        self._keep = []

    def a(self, name):  # array argument
        return VP(self.b.addr(name))

    def f(self, value):  # REAL scalar argument, by reference
        v = ctypes.c_float(value)
        self._keep.append(v)
        return ctypes.byref(v)

    def i(self, value):  # INTEGER / LOGICAL scalar argument, by reference
        v = ctypes.c_int(int(value))
        self._keep.append(v)
        return ctypes.byref(v)

    def dims_args(self):
        return [self.i(getattr(self.dims, n)) for n in DIMS]

    def call(self, routine, args):
        fn = getattr(self.b.lib, "wrfmi_" + routine)
        fn.restype = ctypes.c_int
        rc = fn(*(args + self.dims_args()))
        self._keep.clear()
        if rc != 0:
            msg = self.b.lib.wrfmi_last_error
            msg.restype = ctypes.c_char_p
            raise RuntimeError(f"wrfmi_{routine} failed ({rc}): {msg().decode()}")

    # ---- the call sites (field bindings of solve_em.F / module_em.F)
    def advect(self, which, field, tend, flux=("ru", "rv", "ww")):
        a, f, i, s = self.a, self.f, self.i, self.scal
        rdz = "rdn" if which == "w" else "rdnw"
        args = [a(field), a(field), a(tend), a(flux[0]), a(flux[1]), a(flux[2]), a("c1h"), a("c2h"), a("mut"),
                i(self.cfg.time_step), ctypes.byref(self.cfg), a("msfux"), a("msfuy"), a("msfvx"), a("msfvy"),
                a("msftx"), a("msfty"), a("fnm"), a("fnp"), f(s["rdx"]), f(s["rdy"]), a(rdz)]
        self.call("advect_" + which, args)

    def advance_uv(self, dts):
        a, f, i, s, c = self.a, self.f, self.i, self.scal, self.cfg
        args = [a("u_2"), a("ru_tend"), a("v_2"), a("rv_tend"), a("p"), a("pb"), a("ph_2"), a("php"), a("alt"),
                a("al"), a("mu_2"), a("muu"), a("cqu"), a("muv"), a("cqv"), a("mudf"), a("c1h"), a("c2h"), a("c1f"),
                a("c2f"), a("c3h"), a("c4h"), a("c3f"), a("c4f"), a("msfux"), a("msfuy"), a("msfvx"),
                a("msfvx_inv"), a("msfvy"), f(s["rdx"]), f(s["rdy"]), f(dts), f(s["cf1"]), f(s["cf2"]),
                f(s["cf3"]), a("fnm"), a("fnp"), f(c.emdiv), a("rdnw"), ctypes.byref(c), i(1), i(c.non_hydrostatic),
                i(c.top_lid)]
        self.call("advance_uv", args)

    def advance_mu_t(self, dts, iteration):
        a, f, i, s, c = self.a, self.f, self.i, self.scal, self.cfg
        args = [a("ww"), a("ww1"), a("u_2"), a("u_save"), a("v_2"), a("v_save"), a("mu_2"), a("mut"), a("muave"),
                a("muts"), a("muu"), a("muv"), a("mudf"), a("c1h"), a("c2h"), a("c1f"), a("c2f"), a("c3h"), a("c4h"),
                a("c3f"), a("c4f"), a("ru_m"), a("rv_m"), a("ww_m"), a("t_2"), a("t_save"), a("t_2save"),
                a("t_tend"), a("mu_tend"), f(s["rdx"]), f(s["rdy"]), f(dts), f(c.epssm), a("dnw"), a("fnm"),
                a("fnp"), a("rdnw"), a("msfux"), a("msfuy"), a("msfvx"), a("msfvx_inv"), a("msfvy"), a("msftx"),
                a("msfty"), i(iteration), ctypes.byref(c)]
        self.call("advance_mu_t", args)

    def advance_w(self, dts):
        a, f, i, s, c = self.a, self.f, self.i, self.scal, self.cfg
        args = [a("w_2"), a("rw_tend"), a("ww"), a("w_save"), a("u_2"), a("v_2"), a("mu_2"), a("mut"), a("muave"),
                a("muts"), a("c1h"), a("c2h"), a("c1f"), a("c2f"), a("c3h"), a("c4h"), a("c3f"), a("c4f"),
                a("t_2save"), a("t_2"), a("t_save"), a("ph_2"), a("ph_save"), a("phb"), a("ph_tend"), a("ht"),
                a("c2a"), a("cqw"), a("alt"), a("alb"), a("a"), a("alpha"), a("gamma"), f(s["rdx"]), f(s["rdy"]),
                f(dts), f(300.0), f(c.epssm), a("dnw"), a("fnm"), a("fnp"), a("rdnw"), a("rdn"), f(s["cf1"]),
                f(s["cf2"]), f(s["cf3"]), a("msftx"), a("msfty"), ctypes.byref(c), i(c.top_lid)]
        self.call("advance_w", args)

    def sumflux(self, iteration, nsmall):
        a, f, i, c = self.a, self.f, self.i, self.cfg
        args = [a("u_2"), a("v_2"), a("ww"), a("u_save"), a("v_save"), a("ww1"), a("muu"), a("muv"), a("c1h"),
                a("c2h"), a("c1f"), a("c2f"), a("c3h"), a("c4h"), a("c3f"), a("c4f"), a("ru_m"), a("rv_m"),
                a("ww_m"), f(c.epssm), a("msfux"), a("msfuy"), a("msfvx"), a("msfvx_inv"), a("msfvy"), i(iteration),
                i(nsmall)]
        self.call("sumflux", args)

    def calc_p_rho(self, step):
        a, f, i, c = self.a, self.f, self.i, self.cfg
        args = [a("al"), a("p"), a("ph_2"), a("alt"), a("t_2"), a("t_save"), a("c2a"), a("pm1"), a("mu_2"),
                a("muts"), a("c1h"), a("c2h"), a("c1f"), a("c2f"), a("c3h"), a("c4h"), a("c3f"), a("c4f"), a("znu"),
                f(300.0), a("rdnw"), a("dnw"), f(c.smdiv), i(c.non_hydrostatic), i(step)]
        self.call("calc_p_rho", args)

    def calc_coef_w(self, dts):
        a, f, i, c = self.a, self.f, self.i, self.cfg
        args = [a("a"), a("alpha"), a("gamma"), a("mut"), a("c1h"), a("c2h"), a("c1f"), a("c2f"), a("c3h"),
                a("c4h"), a("c3f"), a("c4f"), a("cqw"), a("rdn"), a("rdnw"), a("c2a"), f(dts), f(9.81), f(c.epssm),
                i(c.top_lid)]
        self.call("calc_coef_w", args)

    def rk_update_scalar(self, m, rk_step, dt_rk):
        a, f, i = self.a, self.f, self.i
        args = [i(1), i(1), a(f"q_old{m}"), a(f"q{m}"), a(f"q_tend{m}"), None, None, a("advect_tend"),
                a("advect_tend"), a("advect_tend"), a("msftx"), a("msfty"), a("c1h"), a("c2h"), a("mu_1"), a("mu_2"),
                a("mub"), i(rk_step), f(dt_rk), i(1), ctypes.byref(self.cfg), i(0)]
        self.call("rk_update_scalar", args)


class OracleRoutines:
    """The oracle's routines on an OracleSession's arrays (same bindings as solve_em)."""

    def __init__(self, session):
        self.s = session
        self.lib = session.o.lib
        self.R = ctypes.c_float if session.o.dtype == np.float32 else ctypes.c_double

    def _st(self):
        return ctypes.byref(self.s.st), ctypes.byref(self.s.cfg), ctypes.byref(self.s.dims)

    def advect(self, which, field, tend, flux=("ru", "rv", "ww")):
        A, s, c = self.s.arrays, self.s.st, self.s.cfg
        p = lambda n: A[n].ctypes.data_as(VP)  # noqa: E731
        fn = getattr(self.lib, "o_advect_" + which)
        fn.restype = ctypes.c_int
        rdz = "rdn" if which == "w" else "rdnw"
        rc = fn(p(field), p(field), p(tend), p(flux[0]), p(flux[1]), p(flux[2]), p("c1h"), p("c2h"), p("mut"),
                ctypes.c_int(c.time_step), ctypes.byref(c), p("msfux"), p("msfuy"), p("msfvx"), p("msfvy"),
                p("msftx"), p("msfty"), p("fnm"), p("fnp"), self.R(s.rdx), self.R(s.rdy), p(rdz),
                ctypes.byref(self.s.dims))
        assert rc == 0

    def state_routine(self, name, *args):
        fn = getattr(self.lib, "o_" + name)
        fn.restype = None
        fn(*self._st(), *args)

    def advance_uv(self, dts):
        self.state_routine("advance_uv", self.R(dts))

    def advance_mu_t(self, dts, iteration):
        self.state_routine("advance_mu_t", self.R(dts))

    def advance_w(self, dts):
        self.state_routine("advance_w", self.R(dts))

    def calc_coef_w(self, dts):
        self.state_routine("calc_coef_w", self.R(dts))

    def calc_p_rho(self, step):
        self.state_routine("calc_p_rho", ctypes.c_int(step))

    def sumflux(self, iteration, nsmall):
        fn = self.lib.o_sumflux
        fn.restype = None
        fn(ctypes.byref(self.s.st), ctypes.byref(self.s.dims), ctypes.c_int(iteration), ctypes.c_int(nsmall))

    def rk_update_scalar(self, m, rk_step, dt_rk):
        fn = self.lib.o_rk_update_scalar
        fn.restype = None
        fn(ctypes.byref(self.s.st), ctypes.byref(self.s.dims), ctypes.c_int(m), ctypes.c_int(rk_step), self.R(dt_rk))


def routine_parity(backend, session, hs, rk_step=2):
    """From the oracle session's state at the start of RK substep `rk_step` (after
    rk_step_prep, rk_tendency and small_step_prep), every reference-argument entry point
    against the oracle routine of the same name.  Returns a list of failure descriptions."""
    f32 = np.float32
    dt = f32(hs.scal["dt"])
    nsound = hs.cfg.time_step_sound
    dts = float(dt / f32(nsound)) if rk_step > 1 else float(dt / f32(3.0))
    nsmall = {1: 1, 2: nsound // 2, 3: nsound}[rk_step]
    dt_rk = {1: dt / f32(3.0), 2: f32(0.5) * dt, 3: dt}[rk_step]
    for rk in range(1, rk_step + 1):
        session.phase("rk_step_prep", rk)
        session.phase("rk_tendency", rk)
        if rk < rk_step:
            session.phase("small_step_prep", rk, dts)
    C = Caller(backend, hs.cfg, hs.dims, hs.scal)
    O = OracleRoutines(session)
    names = list(session.arrays)
    rng = np.random.default_rng(5)
    fails = []

    def check(label, fn_b, fn_o, prep=None):
        if prep:
            prep()
        backend.upload(session.arrays)
        fn_b()
        fn_o()
        got = backend.download()
        f = compare_exact(got, session.arrays, names)
        if f:
            fails.append(f"{label}: " + "; ".join(f))

    def noisy_tendency(name):
        def prep():
            a = session.arrays[name]
            a[...] = (rng.standard_normal(a.shape) * 1e-3).astype(a.dtype)
        return prep

    # the advection routines add to whatever the tendency array holds
    for which, field, tend in (("u", "u_2", "ru_tend"), ("v", "v_2", "rv_tend"), ("w", "w_2", "rw_tend"),
                               ("scalar", "t_2", "t_tend")):
        check(f"advect_{which}", lambda: C.advect(which, field, tend), lambda: O.advect(which, field, tend),
              noisy_tendency(tend))
    check("advect_scalar(q0, ru_m)", lambda: C.advect("scalar", "q0", "advect_tend", ("ru", "rv", "ww")),
          lambda: O.advect("scalar", "q0", "advect_tend", ("ru", "rv", "ww")), noisy_tendency("advect_tend"))
    # this substep's small_step_prep, then calc_p_rho (step 0) and calc_coef_w on their own
    session.phase("rk_tendency", rk_step)
    session.phase("small_step_prep", rk_step, dts)
    check("calc_p_rho(0)", lambda: C.calc_p_rho(0), lambda: O.calc_p_rho(0))
    check("calc_coef_w", lambda: C.calc_coef_w(dts), lambda: O.calc_coef_w(dts))
    for it in range(1, nsmall + 1):
        check(f"advance_uv it {it}", lambda: C.advance_uv(dts), lambda: O.advance_uv(dts))
        check(f"advance_mu_t it {it}", lambda: C.advance_mu_t(dts, it), lambda: O.advance_mu_t(dts, it))
        check(f"advance_w it {it}", lambda: C.advance_w(dts), lambda: O.advance_w(dts))
        check(f"sumflux it {it}", lambda: C.sumflux(it, nsmall), lambda: O.sumflux(it, nsmall))
        check(f"calc_p_rho it {it}", lambda: C.calc_p_rho(it), lambda: O.calc_p_rho(it))
        if fails:
            return fails
    session.phase("small_step_finish", rk_step, nsmall, dts)
    for m in range(hs.n_species):
        O.state_routine("rk_scalar_tend", ctypes.c_int(m), ctypes.c_int(rk_step), O.R(dt_rk))  # advect_tend of m
        check(f"rk_update_scalar species {m}", lambda: C.rk_update_scalar(m, rk_step, float(dt_rk)),
              lambda: O.rk_update_scalar(m, rk_step, float(dt_rk)))
    return fails
