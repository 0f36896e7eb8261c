"""Pins of the CPU oracle (oracle/*.c) -- the checks that make it a trustworthy checker.

* advect_scalar and advect_scalar_pd: bit for bit against the reference's own routines,
  compiled in place (oracle/_ref, when the reference tree is mounted) and through the
  committed fixtures tests/golden/advect_scalar_ref.npz and advect_scalar_pd_ref.npz that the
  same build wrote.
* The whole step and the routines the reference cannot be compiled for here (they need
  the Registry-generated module_configure): properties the equations fix --
  a resting hydrostatic atmosphere stays at rest; a horizontally uniform state stays
  uniform; mirror and x/y-transpose symmetry of a centred bubble without wind;
  translation invariance on a periodic domain; conservation of dry-air and species mass
  on a periodic domain; the equation of state reproduces the independently built
  hydrostatic initial state; FP32 and FP64 builds agree to FP32 accuracy.
* The analytic sounding: against the reference's test/em_quarter_ss/input_sounding.
"""
import ctypes
import os

import numpy as np
import pytest

from parity import interior, rel_l2
from wrf_amd import grid

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
MSF = ("msfux", "msfuy", "msfvx", "msfvy", "msftx", "msfty")


def _oracle_advect_scalar(o, g, n, tend):
    d = grid.Dims(*[int(x) for x in g[f"c{n}_dims"]])
    spec = grid.DomainSpec(bc="periodic" if int(g[f"c{n}_bc"]) == 0 else "open")
    c = grid.make_config(spec)
    ts = int(g[f"c{n}_time_step"])
    c.time_step = ts
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    L = o.lib
    L.o_advect_scalar.argtypes = ([ctypes.c_void_p] * 9 + [ctypes.c_int, ctypes.c_void_p] + [ctypes.c_void_p] * 8
                                  + [ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p])
    a = {k: np.ascontiguousarray(g[f"c{n}_{k}"]) for k in
         ("field", "field_old", "ru", "rv", "rom", "c1", "c2", "mut", "fzm", "fzp", "rdzw") + MSF}
    rc = L.o_advect_scalar(P(a["field"]), P(a["field_old"]), P(tend), P(a["ru"]), P(a["rv"]), P(a["rom"]), P(a["c1"]),
                           P(a["c2"]), P(a["mut"]), ts, ctypes.byref(c), *[P(a[m]) for m in MSF], P(a["fzm"]),
                           P(a["fzp"]), float(g[f"c{n}_rdx"]), float(g[f"c{n}_rdy"]), P(a["rdzw"]), ctypes.byref(d))
    assert rc == 0


@pytest.mark.parametrize("n", range(6))
def test_advect_scalar_matches_reference_fixture(oracle32, n):
    g = np.load(os.path.join(GOLDEN, "advect_scalar_ref.npz"))
    tend = np.zeros_like(g[f"c{n}_tendency"])
    _oracle_advect_scalar(oracle32, g, n, tend)
    assert np.array_equal(tend.view(np.int32), g[f"c{n}_tendency"].view(np.int32))
    assert np.count_nonzero(tend) > 100


def test_advect_scalar_matches_reference_build(oracle32):
    """Fresh random inputs through the reference build itself (needs oracle/_ref)."""
    ref_lib = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_advect.so")
    if not os.path.exists(ref_lib):
        pytest.skip("reference advection kernel not built (no reference tree)")
    from pyoracle import RefAdvect
    ref = RefAdvect()
    rng = np.random.default_rng(77)
    for bc, ts in (("periodic", 6), ("open", -12)):
        spec = grid.DomainSpec(nx=15, ny=12, nz=9, bc=bc)
        d, c = grid.make_dims(spec), grid.make_config(spec)
        c.time_step = ts
        s3, s2, s1 = grid.shape3(d), grid.shape2(d), grid.shape1(d)
        u = lambda sh, lo, hi: rng.uniform(lo, hi, sh).astype(np.float32)
        g = {"c0_field": u(s3, 0, 1e-2), "c0_field_old": u(s3, 0, 1e-2), "c0_ru": u(s3, -5e4, 5e4),
             "c0_rv": u(s3, -5e4, 5e4), "c0_rom": u(s3, -80, 80), "c0_mut": u(s2, 8e4, 9e4),
             "c0_c1": u(s1, 0.5, 1), "c0_c2": u(s1, 0, 1e3), "c0_fzm": u(s1, 0.4, 0.6), "c0_rdzw": u(s1, -40, -10),
             "c0_time_step": np.int32(ts), "c0_rdx": np.float32(1e-3), "c0_rdy": np.float32(1e-3),
             "c0_bc": np.int32(0 if bc == "periodic" else 1),
             "c0_dims": np.array([getattr(d, n) for n, _ in d._fields_], np.int32)}
        g["c0_fzp"] = (1 - g["c0_fzm"]).astype(np.float32)
        for m in MSF:
            g["c0_" + m] = u(s2, 0.9, 1.1)
        t_ref = np.zeros(s3, np.float32)
        ref.advect_scalar(g["c0_field"], g["c0_field_old"], t_ref, g["c0_ru"], g["c0_rv"], g["c0_rom"], g["c0_c1"],
                          g["c0_c2"], g["c0_mut"], ts, c, {m: g["c0_" + m] for m in MSF}, g["c0_fzm"], g["c0_fzp"],
                          1e-3, 1e-3, g["c0_rdzw"], d)
        t_or = np.zeros(s3, np.float32)
        _oracle_advect_scalar(oracle32, g, 0, t_or)
        assert np.array_equal(t_ref.view(np.int32), t_or.view(np.int32)), bc


def _oracle_advect_scalar_pd(o, g, n):
    """o_advect_scalar_pd on case n of a fixture dict `g`; returns (tendency, h, z)."""
    d = grid.Dims(*[int(x) for x in g[f"c{n}_dims"]])
    c = grid.make_config(grid.DomainSpec(bc="periodic" if int(g[f"c{n}_bc"]) == 0 else "open"))
    ts = int(g[f"c{n}_time_step"])
    c.time_step = ts
    vp = ctypes.c_void_p
    P = lambda a: a.ctypes.data_as(vp)
    L = o.lib
    L.o_advect_scalar_pd.argtypes = ([vp] * 13 + [ctypes.c_int, vp, ctypes.c_int] + [vp] * 8
                                     + [ctypes.c_float, ctypes.c_float, vp, ctypes.c_float, vp])
    a = {k: np.ascontiguousarray(g[f"c{n}_{k}"]) for k in
         ("field", "field_old", "ru", "rv", "rom", "c1", "c2", "mut", "mub", "mu_old", "fzm", "fzp", "rdzw") + MSF}
    out = [np.zeros_like(a["field"]) for _ in range(3)]
    rc = L.o_advect_scalar_pd(P(a["field"]), P(a["field_old"]), *[P(x) for x in out], P(a["ru"]), P(a["rv"]),
                              P(a["rom"]), P(a["c1"]), P(a["c2"]), P(a["mut"]), P(a["mub"]), P(a["mu_old"]), ts,
                              ctypes.byref(c), 1, *[P(a[m]) for m in MSF], P(a["fzm"]), P(a["fzp"]),
                              float(g[f"c{n}_rdx"]), float(g[f"c{n}_rdy"]), P(a["rdzw"]), float(g[f"c{n}_dt"]),
                              ctypes.byref(d))
    assert rc == 0
    return out


@pytest.mark.parametrize("n", range(4))
def test_advect_scalar_pd_matches_reference_fixture(oracle32, n):
    """The positive-definite form against the reference's own advect_scalar_pd (fixture
    tests/golden/advect_scalar_pd_ref.npz): tendency and its horizontal/vertical split."""
    g = np.load(os.path.join(GOLDEN, "advect_scalar_pd_ref.npz"))
    for name, got in zip(("tendency", "h_tendency", "z_tendency"), _oracle_advect_scalar_pd(oracle32, g, n)):
        assert np.array_equal(got.view(np.int32), g[f"c{n}_{name}"].view(np.int32)), name
        assert np.count_nonzero(got) > 100, name


def test_advect_scalar_pd_matches_reference_build(oracle32):
    """Fresh inputs of other shapes through the reference build itself (needs oracle/_ref)."""
    ref_lib = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_advect.so")
    if not os.path.exists(ref_lib):
        pytest.skip("reference advection kernel not built (no reference tree)")
    from parity import pd_advect_inputs
    from pyoracle import RefAdvect
    ref = RefAdvect()
    for bc, ts, shape, seed in (("periodic", 6, (15, 12, 9), 51), ("open", -12, (13, 14, 7), 52),
                                ("open", 6, (16, 11, 10), 53)):
        d, c, a = pd_advect_inputs(bc, *shape, seed)
        c.time_step = ts
        outs = [np.zeros_like(a["field"]) for _ in range(3)]
        ref.advect_scalar_pd(a["field"], a["field_old"], *outs, a["ru"], a["rv"], a["rom"], a["c1"], a["c2"], a["mut"],
                             a["mub"], a["mu_old"], ts, c, 1, {m: a[m] for m in MSF}, a["fzm"], a["fzp"],
                             float(a["rdx"]), float(a["rdy"]), a["rdzw"], float(a["dt"]), d)
        g = {f"c0_{k}": v for k, v in a.items()}
        g.update(c0_dims=np.array([getattr(d, n) for n, _ in d._fields_], np.int32), c0_time_step=np.int32(ts),
                 c0_bc=np.int32(0 if bc == "periodic" else 1))
        for name, r, o in zip(("tendency", "h", "z"), outs, _oracle_advect_scalar_pd(oracle32, g, 0)):
            assert np.array_equal(r.view(np.int32), o.view(np.int32)), (bc, name)


def test_sounding_matches_reference_input_sounding():
    a = np.loadtxt(os.path.join(GOLDEN, "em_quarter_ss_input_sounding.txt"), skiprows=1)
    z, th = a[:, 0], a[:, 1]
    trop = z <= 12000.0
    # the file holds theta to 4 decimals
    assert np.abs(grid.wk_theta(z[trop]) - th[trop]).max() < 1e-4
    assert abs(float(np.loadtxt(os.path.join(GOLDEN, "em_quarter_ss_input_sounding.txt"), max_rows=1)[1]) - 300.0) < 1e-9


def _run(o, spec, nsteps):
    hs = grid.build_initial_state(spec)
    return hs, o.run(hs, nsteps)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_resting_atmosphere_stays_at_rest(prec):
    from pyoracle import Oracle
    spec = grid.DomainSpec(nx=12, ny=10, nz=20, bc="periodic", bubble_amp=0.0, noise_amp=0.0, wind_scale=0.0, seed=1)
    hs, a = _run(Oracle(prec), spec, 10)
    d = hs.dims
    w = interior(a["w_2"], d, "w")
    u = interior(a["u_2"], d, "u")
    assert np.abs(w).max() < 1e-3
    assert np.abs(u).max() < 1e-3
    assert rel_l2(interior(a["t_2"], d), interior(hs.a3["t_2"], d)) < 1e-6


def test_uniform_state_stays_uniform(oracle32):
    spec = grid.DomainSpec(nx=10, ny=9, nz=14, bc="periodic", bubble_amp=0.0, noise_amp=0.0, seed=2)
    hs, a = _run(oracle32, spec, 3)
    d = hs.dims
    for k, kind in (("u_2", "u"), ("v_2", "v"), ("w_2", "w"), ("t_2", "mass"), ("ph_2", "w"), ("p", "mass")):
        x = interior(a[k], d, kind)
        col = x[:1, :, :1]
        assert np.array_equal(x, np.broadcast_to(col, x.shape)), k
    assert not np.array_equal(a["u_2"], hs.a3["u_2"])


def _sym_state(bc):
    spec = grid.DomainSpec(nx=16, ny=16, nz=12, dx=2000.0, dy=2000.0, bc=bc, wind_scale=0.0, noise_amp=0.0,
                           bubble_amp=3.0, n_bubbles=1, seed=3)
    return spec


@pytest.mark.parametrize("bc", ["periodic", "open"])
def test_mirror_and_transpose_symmetry(oracle32, bc):
    """One step: rounding-level asymmetries later grow fast, because the upwind sign of
    the 5th-order fluxes flips with the sign of near-zero velocities on the mirror plane."""
    hs, a = _run(oracle32, _sym_state(bc), 1)
    d = hs.dims
    t = interior(a["t_2"], d) - interior(hs.a3["t_2"], d)
    w = interior(a["w_2"], d, "w")
    u, v = interior(a["u_2"], d, "u"), interior(a["v_2"], d, "v")
    p = interior(a["p"], d)
    scale_t, scale_w, scale_u = np.abs(t).max(), np.abs(w).max(), np.abs(u).max()
    assert scale_w > 1e-3 and scale_u > 1e-4
    tol = 1e-4
    # mirror in x: mass points i -> nx-1-i; u at face i -> -u at face nx-i
    assert np.abs(t - t[:, :, ::-1]).max() <= tol * scale_t
    assert np.abs(w - w[:, :, ::-1]).max() <= tol * scale_w
    assert np.abs(u + u[:, :, ::-1]).max() <= tol * scale_u
    # mirror in y
    assert np.abs(t - t[::-1]).max() <= tol * scale_t
    assert np.abs(v + v[::-1]).max() <= tol * scale_u
    # x <-> y transpose: u(j, k, i) = v(i, k, j)
    assert np.abs(t - t.transpose(2, 1, 0)).max() <= tol * scale_t
    assert np.abs(u - v.transpose(2, 1, 0)).max() <= tol * scale_u
    assert np.abs(p - p.transpose(2, 1, 0)).max() <= tol * np.abs(p).max()


def test_periodic_translation_invariance(oracle32):
    spec = grid.DomainSpec(nx=14, ny=12, nz=10, bc="periodic", n_passive=1, n_bubbles=2, seed=4)
    hs = grid.build_initial_state(spec)
    d = hs.dims
    di, dj = 5, 3
    sh = hs.copy()
    i0, j0, nx, ny = d.ids - d.ims, d.jds - d.jms, spec.nx, spec.ny

    def roll(x):
        if x.ndim == 1:
            return
        blk = x[j0:j0 + ny, ..., i0:i0 + nx] if x.ndim == 3 else x[j0:j0 + ny, i0:i0 + nx]
        blk[...] = np.roll(np.roll(blk, dj, axis=0), di, axis=-1)

    stag = {"u_1": (1, 0), "u_2": (1, 0), "v_1": (0, 1), "v_2": (0, 1)}
    for name, x in sh.arrays().items():
        if x.ndim > 1:
            roll(x)
            si, sj = stag.get(name, (0, 0))
            grid._halo_fill_np(x, d, True, bool(si), bool(sj))
    a = oracle32.run(hs, 2)
    b = oracle32.run(sh, 2)
    for k, kind in (("u_2", "u"), ("v_2", "v"), ("w_2", "w"), ("t_2", "mass"), ("mu_2", "mass"), ("q1", "mass")):
        xa = interior(a[k], d, kind)
        xb = interior(b[k], d, kind)
        xa = xa[:ny, ..., :nx] if xa.ndim == 3 else xa[:ny, :nx]
        xb = xb[:ny, ..., :nx] if xb.ndim == 3 else xb[:ny, :nx]
        assert np.array_equal(np.roll(np.roll(xa, dj, axis=0), di, axis=-1), xb), k


def test_periodic_mass_conservation(oracle32):
    spec = grid.DomainSpec(nx=24, ny=20, nz=16, dx=1000.0, dy=1000.0, dt=6.0, bc="periodic", n_passive=2,
                           n_bubbles=2, seed=1002)
    hs = grid.build_initial_state(spec)
    d = hs.dims

    def budgets(a):
        mu = interior(a["mu_2"], d).astype(np.float64)
        mub = interior(a["mub"], d).astype(np.float64)
        c1h = a["c1h"][: d.kde - 1].astype(np.float64)
        c2h = a["c2h"][: d.kde - 1].astype(np.float64)
        dnw = a["dnw"][: d.kde - 1].astype(np.float64)
        col = (c1h[None, :, None] * (mu + mub)[:, None, :] + c2h[None, :, None]) * (-dnw)[None, :, None]
        return float(np.sum(mu)), [np.sum(col * interior(a[f"q{m}"], d).astype(np.float64)) for m in range(hs.n_species)]

    m0, q0 = budgets(hs.arrays())
    a = oracle32.run(hs, 5)
    m1, q1 = budgets(a)
    mtot = float(np.sum(interior(hs.a2["mub"], d).astype(np.float64)))
    assert abs(m1 - m0) / mtot < 1e-9
    for x0, x1 in zip(q0, q1):
        assert abs(x1 - x0) / abs(x0) < 1e-6
    assert not np.array_equal(a["q1"], hs.q[1])


def test_eos_reproduces_hydrostatic_initial_state(oracle32):
    """calc_p_rho_phi (C, FP32) on the initial state built in float64 numpy from the
    sounding (an independent evaluation of the same relations) returns its p and al."""
    hs = grid.build_initial_state(grid.supercell_c0())
    # calc_p_rho_phi reads the full column mass of the acoustic steps (muts = mub + mu)
    hs.a2["muts"][...] = hs.a2["mub"] + hs.a2["mu_2"]
    sess = oracle32.session(hs)
    sess.phase("calc_p_rho_phi")
    d = hs.dims
    p0, pb = interior(hs.a3["p"], d).astype(np.float64), interior(hs.a3["pb"], d).astype(np.float64)
    p1 = interior(sess.arrays["p"], d).astype(np.float64)
    ulp = np.spacing(np.abs(p0 + pb).astype(np.float32)).astype(np.float64)
    # FP32 conditioning: Delta(ph+phb) over a layer and log(pfd/pfu) near 1 each lose
    # ~4-5 bits, so tens of ulps are expected; a wrong level, sign or exponent gives 1e4+
    assert np.max(np.abs(p1 - p0) / ulp) < 256
    al0, alb = interior(hs.a3["al"], d).astype(np.float64), interior(hs.a3["alb"], d).astype(np.float64)
    al1 = interior(sess.arrays["al"], d).astype(np.float64)
    ulpa = np.spacing(np.abs(al0 + alb).astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(al1 - al0) / ulpa) < 256


def test_fp32_and_fp64_builds_agree(oracle32, oracle64):
    hs = grid.build_initial_state(grid.supercell_c0())
    a, b = oracle32.run(hs, 1), oracle64.run(hs, 1)
    d = hs.dims
    bounds = dict(u_2=1e-4, v_2=1e-4, t_2=1e-6, mu_2=1e-3, ph_2=1e-4, p=1e-2, al=1e-3, w_2=1e-1, q0=1e-5)
    kinds = dict(u_2="u", v_2="v", w_2="w", ph_2="w")
    for k, bound in bounds.items():
        err = rel_l2(interior(a[k], d, kinds.get(k, "mass")), interior(b[k], d, kinds.get(k, "mass")))
        assert err < bound, (k, err)


def _advect(o, name, f, ru, rv, rom, ts, cfg, d, rdzw, fzm, fzp, msf1, one1, zero1, mut):
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    fn = getattr(o.lib, name)
    fn.argtypes = ([ctypes.c_void_p] * 9 + [ctypes.c_int, ctypes.c_void_p] + [ctypes.c_void_p] * 8
                   + [ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p])
    tend = np.zeros_like(f)
    assert fn(P(f), P(f), P(tend), P(ru), P(rv), P(rom), P(one1), P(zero1), P(mut), ts, ctypes.byref(cfg),
              *[P(msf1)] * 6, P(fzm), P(fzp), 1 / 1500., 1 / 1200., P(rdzw), ctypes.byref(d)) == 0
    return tend


@pytest.mark.parametrize("ts,vel", [(6, (3.0e4, -2.0e4, 35.0)), (-6, (-1.5e4, 2.5e4, -20.0))])
def test_advect_u_v_equal_pinned_advect_scalar_in_uniform_flow(oracle32, ts, vel):
    """With spatially uniform ru, rv, rom and unit map factors, advect_u's face velocities
    0.5*(ru(i)+ru(i-1)) equal advect_scalar's ru(i), and its fluxes at mass points i-1, i
    are advect_scalar's fluxes at faces i, i+1 of the same array: on a periodic domain
    advect_u(f) and advect_v(f) must equal advect_scalar(f) -- which is pinned bit for bit
    to the reference -- point for point.  advect_w's horizontal fluxes see ru, rv
    interpolated to w levels, so for advect_w the same holds to rounding in the interior."""
    spec = grid.DomainSpec(nx=14, ny=12, nz=10, bc="periodic")
    d, cfg = grid.make_dims(spec), grid.make_config(spec)
    cfg.time_step = ts
    s3, s2, s1 = grid.shape3(d), grid.shape2(d), grid.shape1(d)
    rng = np.random.default_rng(9)
    f = rng.uniform(-5, 5, s3).astype(np.float32)
    grid._halo_fill_np(f, d, True, False, False)
    ru, rv, rom = (np.full(s3, v, np.float32) for v in vel)
    msf1 = np.ones(s2, np.float32)
    mut = np.full(s2, 9.0e4, np.float32)
    one1, zero1 = np.ones(s1, np.float32), np.zeros(s1, np.float32)
    fzm = np.full(s1, 0.5, np.float32)
    fzp = np.full(s1, 0.5, np.float32)
    rdzw = rng.uniform(-40, -10, s1).astype(np.float32)
    args = (f, ru, rv, rom, ts, cfg, d, rdzw, fzm, fzp, msf1, one1, zero1, mut)
    ts_ = _advect(oracle32, "o_advect_scalar", *args)
    tu = _advect(oracle32, "o_advect_u", *args)
    tv = _advect(oracle32, "o_advect_v", *args)
    tw = _advect(oracle32, "o_advect_w", *args)
    j0, i0 = d.jds - d.jms, d.ids - d.ims
    sl = (slice(j0, j0 + spec.ny), slice(0, d.kde - 1), slice(i0, i0 + spec.nx))
    assert np.count_nonzero(ts_[sl]) > 0.9 * ts_[sl].size
    assert np.array_equal(tu[sl], ts_[sl])
    assert np.array_equal(tv[sl], ts_[sl])
    # advect_w on w levels 3..kde-2: horizontal part equal to rounding, vertical part differs
    # (staggered levels), so compare with the vertical flux switched off
    rom0 = np.zeros(s3, np.float32)
    args0 = (f, ru, rv, rom0, ts, cfg, d, rdzw, fzm, fzp, msf1, one1, zero1, mut)
    ts0 = _advect(oracle32, "o_advect_scalar", *args0)
    tw0 = _advect(oracle32, "o_advect_w", *args0)
    slw = (sl[0], slice(2, d.kde - 2), sl[2])
    assert np.abs(tw0[slw] - ts0[slw]).max() <= 1e-5 * np.abs(ts0[slw]).max()
    assert np.count_nonzero(tw) > 0


@pytest.mark.parametrize("ts,W", [(6, 35.0), (-6, -20.0), (6, -20.0)])
def test_advect_w_vertical_equals_pinned_advect_scalar(oracle32, ts, W):
    """ru = rv = 0 and uniform rom: advect_w's vertical flux at mass level k (rom averaged
    to the level, w(k-1..k+2)) is advect_scalar's flux at w level k+1 of the same array,
    so the tendencies agree bit for bit on levels 2..kde-3 (the lowest and the top three
    levels use the reference's special boundary stencils)."""
    spec = grid.DomainSpec(nx=14, ny=12, nz=10, bc="periodic")
    d, cfg = grid.make_dims(spec), grid.make_config(spec)
    cfg.time_step = ts
    s3, s2, s1 = grid.shape3(d), grid.shape2(d), grid.shape1(d)
    rng = np.random.default_rng(10)
    f = rng.uniform(-5, 5, s3).astype(np.float32)
    grid._halo_fill_np(f, d, True, False, False)
    z3 = np.zeros(s3, np.float32)
    rom = np.full(s3, W, np.float32)
    args = (f, z3, z3, rom, ts, cfg, d, rng.uniform(-40, -10, s1).astype(np.float32), np.full(s1, 0.5, np.float32),
            np.full(s1, 0.5, np.float32), np.ones(s2, np.float32), np.ones(s1, np.float32), np.zeros(s1, np.float32),
            np.full(s2, 9.0e4, np.float32))
    tsc = _advect(oracle32, "o_advect_scalar", *args)
    tw = _advect(oracle32, "o_advect_w", *args)
    j0, i0 = d.jds - d.jms, d.ids - d.ims
    sl = (slice(j0, j0 + spec.ny), slice(1, d.kde - 3), slice(i0, i0 + spec.nx))
    assert np.count_nonzero(tsc[sl]) > 0.9 * tsc[sl].size
    assert np.array_equal(tw[sl], tsc[sl])


@pytest.mark.parametrize("bc_x,bc_y", [("open", "periodic"), ("periodic", "open")])
def test_set_w_surface_closed_form(oracle32, bc_x, bc_y):
    """set_w_surface with u = U, v = 0 and unit map factors reduces to the centred slope
    w(1) = U * (ht(i+1) - ht(i-1)) / (2 dx), and across an open edge to the one-sided
    U * (ht(ids+1) - ht(ids)) / (2 dx); cf1 + cf2 + cf3 = 1 extrapolates a uniform u
    exactly up to rounding.  The terrain initial state's w (start_em's call with the fill
    w(k) = w(1) znw(k)^2) is the oracle's level-1 value bit for bit."""
    spec = grid.terrain_c4(nx=30, ny=24, nz=10, dx=1000.0, dy=1000.0, bc_x=bc_x, bc_y=bc_y, ridge_a=(3000.0, 8000.0),
                           n_passive=0, u0=12.5)
    hs = grid.build_initial_state(spec)
    d = hs.dims
    sess = oracle32.session(hs)
    sess.arrays["w_2"][...] = 0.0
    sess.phase("set_w_surface")
    w1 = interior(sess.arrays["w_2"], d, "w")[:, 0, :].astype(np.float64)
    ht = hs.a2["ht"].astype(np.float64)
    i0, j0 = d.ids - d.ims, d.jds - d.jms
    nx, ny = spec.nx, spec.ny
    ii = np.arange(nx) + i0
    hp = ht[j0:j0 + ny][:, ii + 1]
    hm = ht[j0:j0 + ny][:, ii - 1]
    if bc_x == "open":
        hm[:, 0] = ht[j0:j0 + ny, i0]
        hp[:, -1] = ht[j0:j0 + ny, i0 + nx - 1]
    exp = spec.u0 * (hp - hm) / (2.0 * spec.dx)
    assert np.max(np.abs(exp)) > 0.5
    assert np.max(np.abs(w1 - exp)) <= 1e-5 * np.max(np.abs(exp)), np.max(np.abs(w1 - exp))
    # the initial state of both time levels carries the same surface w and its fill
    w_init = interior(hs.a3["w_2"], d, "w")
    assert np.array_equal(w_init[:, 0, :], interior(sess.arrays["w_2"], d, "w")[:, 0, :])
    znw = hs.a1["znw"]
    k = 5
    assert np.array_equal(w_init[:, k, :], w_init[:, 0, :] * znw[k] * znw[k])
    assert np.array_equal(hs.a3["w_1"], hs.a3["w_2"])


def test_advance_w_surface_term_closed_form(oracle32):
    """advance_w's lower boundary (module_small_step_em.F:1383-1395) has the same form as
    set_w_surface on the raw i-1, i+1 and j-1, j+1 neighbours (halo values across an edge).
    With the small-step u = U and v = 0 it is U * (ht(i+1) - ht(i-1)) / (2 dx)."""
    spec = grid.terrain_c4(nx=26, ny=20, nz=10, dx=1000.0, dy=1000.0, ridge_a=(3000.0, 8000.0), n_passive=0)
    hs = grid.build_initial_state(spec)
    d = hs.dims
    sess = oracle32.session(hs)
    U = 7.25
    sess.arrays["u_2"][...] = U
    sess.arrays["v_2"][...] = 0.0
    sess.phase("advance_mu_t_w", 0.5, 1, 1)
    w1 = interior(sess.arrays["w_2"], d, "w")[:, 0, :].astype(np.float64)
    ht = hs.a2["ht"].astype(np.float64)
    i0, j0 = d.ids - d.ims, d.jds - d.jms
    ii = np.arange(spec.nx) + i0
    exp = U * (ht[j0:j0 + spec.ny][:, ii + 1] - ht[j0:j0 + spec.ny][:, ii - 1]) / (2.0 * spec.dx)
    assert np.max(np.abs(exp)) > 0.3
    assert np.max(np.abs(w1 - exp)) <= 1e-5 * np.max(np.abs(exp)), np.max(np.abs(w1 - exp))


TILE_CASES = {
    "periodic": dict(nx=24, ny=20, nz=12, bc="periodic", n_passive=2, n_bubbles=2, bubble_amp=2.0, seed=11),
    "open": dict(nx=22, ny=19, nz=12, bc="open", n_bubbles=1, seed=12),
    "terrain_yopen": dict(nx=20, ny=23, nz=12, dx=1000.0, dy=1000.0, dt=3.0, bc_x="periodic", bc_y="open",
                          hybrid_opt=2, sounding="stable", n_ridges=2, ridge_a=(3000.0, 8000.0), n_passive=1,
                          n_bubbles=0, noise_amp=0.0, seed=22),
}


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("case", sorted(TILE_CASES))
def test_oracle_tiling_invariance(prec, case):
    """The oracle calls every compute routine once per j-tile inside an OpenMP loop, as
    solve_em does; the reference's routines write only their tile's points from inputs of
    earlier phases, so the result of two steps is bit for bit that of a single tile for
    any tile count (3 and 7 tiles here, rows of unequal height)."""
    from pyoracle import Oracle
    o = Oracle(prec)
    o.lib.oracle_set_num_tiles.argtypes = [ctypes.c_int]
    hs = grid.build_initial_state(grid.DomainSpec(**TILE_CASES[case]))
    try:
        o.lib.oracle_set_num_tiles(1)
        one = o.run(hs, 2)
        for nt in (3, 7):
            o.lib.oracle_set_num_tiles(nt)
            many = o.run(hs, 2)
            bad = [k for k in one if not np.array_equal(one[k], many[k])]
            assert not bad, f"{nt} tiles: {bad}"
    finally:
        o.lib.oracle_set_num_tiles(0)
