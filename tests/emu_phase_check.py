"""Phase-by-phase comparison of the CPU-emulated kernel sources with the FP32 oracle.
Run by test_cpu_emu.py in a subprocess with libasan preloaded:
    python tests/emu_phase_check.py <case> <launch-order> <nsteps>
Exit status 0 = every array bitwise equal after every call (EOS outputs within 4 ulp)."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(HERE), os.path.join(os.path.dirname(HERE), "oracle"), HERE]

import numpy as np  # noqa: E402
from emu_backend import EmuDomain  # noqa: E402
from parity import phase_parity  # noqa: E402
from pyoracle import Oracle  # noqa: E402
from wrf_amd import grid, runtime  # noqa: E402

CASES = {
    "periodic": dict(nx=14, ny=11, nz=10, bc="periodic", n_passive=1, n_bubbles=2, bubble_amp=2.0, seed=11),
    "open": dict(nx=13, ny=12, nz=9, bc="open", n_bubbles=1, seed=12),
    "periodic_pm1": dict(nx=12, ny=10, nz=9, bc="periodic", n_bubbles=1, seed=13, precision_mode=1),
    # rows of 23 and 22 floats: not a multiple of four, so the one-point kernels run throughout
    "periodic_unaligned": dict(nx=14, ny=11, nz=10, bc="periodic", n_passive=1, n_bubbles=2, bubble_amp=2.0, seed=14, pad_i=1),
    "open_unaligned": dict(nx=13, ny=12, nz=9, bc="open", n_bubbles=1, seed=15, pad_i=1),
    # rows of 32 floats ending 6 points after ide: the four-point windows run past the row
    "periodic_tight": dict(nx=21, ny=9, nz=8, bc="periodic", n_passive=1, n_bubbles=1, seed=17, pad_i=4),
    "open_tight": dict(nx=21, ny=10, nz=8, bc="open", n_bubbles=1, seed=18, pad_i=4),
    # rows of 101 and 256 points (one and exactly one 256-point wave row per grid row)
    "periodic_wide": dict(nx=100, ny=7, nz=6, bc="periodic", n_passive=1, n_bubbles=1, seed=23),
    "open_wide": dict(nx=255, ny=10, nz=5, bc="open", n_bubbles=1, seed=24),
    # an open direction of 9 points: too narrow for a boundary zone on each side, so the
    # tendency and scalar-advection kernels keep their one-point forms there
    "open_narrow": dict(nx=8, ny=12, nz=6, bc="open", n_passive=1, n_bubbles=1, seed=25),
    # terrain on hybrid levels with one periodic and one open direction, both ways round
    "terrain": dict(nx=16, ny=12, nz=10, dx=2000.0, dy=2000.0, dt=6.0, bc_x="open", bc_y="periodic",
                    hybrid_opt=2, sounding="stable", n_ridges=3, ridge_a=(3000.0, 8000.0), n_passive=1,
                    n_bubbles=0, noise_amp=0.0, seed=19),
    "terrain_yopen": dict(nx=13, ny=14, nz=9, dx=2000.0, dy=2000.0, dt=6.0, bc_x="periodic", bc_y="open",
                          hybrid_opt=2, sounding="stable", n_ridges=2, ridge_a=(3000.0, 8000.0), n_passive=2,
                          n_bubbles=0, noise_amp=0.0, seed=20),
    # positive-definite transport on the last substep (moist_adv_opt, scalar_adv_opt), passive
    # scalars half empty and seeded physics tendencies (pd_parity.pd_state)
    "pd_periodic": dict(nx=14, ny=11, nz=10, bc="periodic", n_passive=2, n_bubbles=2, bubble_amp=2.0, seed=26,
                        pd=(1, 1)),
    "pd_open": dict(nx=13, ny=12, nz=9, bc="open", n_passive=1, n_bubbles=1, seed=27, pd=(1, 1)),
    # rows of 32 floats: the last four-point group holds ide-1 and three points past it
    "pd_periodic_tight": dict(nx=21, ny=9, nz=9, bc="periodic", n_passive=1, n_bubbles=1, seed=29, pad_i=4,
                              pd=(1, 1)),
    # forward-step mixing and damping of the first substep: 6th-order diffusion (monotone form,
    # slope tapering), the Rayleigh damper toward the base profiles, advance_w's damp_opt 3
    "mix_periodic": dict(nx=14, ny=11, nz=12, bc="periodic", n_passive=1, n_bubbles=2, bubble_amp=2.0, seed=35,
                         opts=dict(diff_6th_opt=2, diff_6th_slopeopt=1, damp_opt=2, zdamp=6000.0)),
    "mix_open": dict(nx=15, ny=13, nz=10, bc="open", n_passive=1, n_bubbles=1, seed=36,
                     opts=dict(diff_6th_opt=1, damp_opt=3, zdamp=5000.0, dampcoef=0.3)),
    "mix_terrain": dict(nx=16, ny=12, nz=10, dx=2000.0, dy=2000.0, dt=6.0, bc_x="open", bc_y="periodic",
                        hybrid_opt=2, sounding="stable", n_ridges=3, ridge_a=(3000.0, 8000.0), n_passive=1,
                        n_bubbles=0, noise_amp=0.0, seed=37,
                        opts=dict(diff_6th_opt=2, diff_6th_slopeopt=1, diff_6th_thresh=0.02, damp_opt=2,
                                  zdamp=6000.0)),
    # diff_opt 1 with constant eddy coefficients (km_opt 1): second-order horizontal and
    # vertical diffusion of u, v, w, theta (relative to t_init) and the species (QV relative
    # to qv_base) on the first substep, with 6th-order diffusion and damping on top
    "diff_periodic": dict(nx=14, ny=11, nz=12, bc="periodic", n_passive=1, n_bubbles=2, bubble_amp=2.0, seed=38,
                          opts=dict(diff_opt=1, km_opt=1, khdif=300.0, kvdif=50.0)),
    "diff_open": dict(nx=15, ny=13, nz=10, bc="open", n_passive=1, n_bubbles=1, seed=39,
                      opts=dict(diff_opt=1, km_opt=1, khdif=200.0, kvdif=80.0, diff_6th_opt=2, damp_opt=3,
                                zdamp=5000.0)),
    "diff_terrain": dict(nx=16, ny=12, nz=10, dx=2000.0, dy=2000.0, dt=6.0, bc_x="open", bc_y="periodic",
                         hybrid_opt=2, sounding="stable", n_ridges=3, ridge_a=(3000.0, 8000.0), n_passive=1,
                         n_bubbles=0, noise_amp=0.0, seed=40,
                         opts=dict(diff_opt=1, km_opt=1, khdif=500.0, kvdif=30.0, damp_opt=2, zdamp=6000.0)),
    # specified lateral boundaries (relaxation and spec zones, flow-dependent species)
    "spec_all": dict(nx=16, ny=14, nz=10, bc="specified", n_passive=1, n_bubbles=2, bubble_amp=2.0, seed=41, lbc=51),
    "spec_channel": dict(nx=15, ny=13, nz=9, bc_x="periodic", bc_y="specified", n_passive=1, n_bubbles=1, seed=42,
                         lbc=52),
    "spec_mixed": dict(nx=17, ny=12, nz=10, bc="specified", n_passive=2, n_bubbles=1, seed=43, lbc=53, pd=(1, 1),
                       opts=dict(have_bcs_scalar=1, diff_6th_opt=2, damp_opt=3, zdamp=5000.0)),
    "pd_mixed_unaligned": dict(nx=14, ny=11, nz=8, bc_x="open", bc_y="periodic", n_passive=1, n_bubbles=1, seed=28,
                               pad_i=1, pd=(0, 1)),
}


def build_case(case: str):
    """The initial state of CASES[case] (pd_parity.pd_state for the positive-definite cases)."""
    kw = dict(CASES[case])
    pd = kw.pop("pd", None)
    opts = kw.pop("opts", {})
    lbc = kw.pop("lbc", None)
    if pd is None:
        hs = grid.build_initial_state(grid.DomainSpec(**kw))
    else:
        from pd_parity import pd_state
        hs = pd_state(kw, *pd, seed=kw["seed"])
    for k, v in opts.items():
        setattr(hs.cfg, k, v)
    if lbc is not None:
        grid.build_lbc(hs, seed=lbc)
    return hs


def pd_routines_check(order: int):
    """wrfmi_advect_scalar_pd against the reference's fixture and the oracle, and
    wrfmi_rk_update_scalar_pd against the oracle."""
    import ctypes
    from emu_backend import EMU_LIB
    from pd_parity import fixture_failures, update_pd_failures
    lib = ctypes.CDLL(EMU_LIB)
    lib.emu_set_launch_order(order)
    fails = fixture_failures(lib, Oracle("f32")) + update_pd_failures(lib, Oracle("f32"))
    for f in fails:
        print("MISMATCH", f)
    print("OK pd_routines" if not fails else "FAILED pd_routines", "order", order, flush=True)
    sys.exit(1 if fails else 0)


def registry_check():
    """Host-array registry: a second registration of the same address refreshes the
    mirror; an overlapping range is refused; unregistering frees the address."""
    import ctypes

    import numpy as np
    from emu_backend import EMU_LIB
    lib = ctypes.CDLL(EMU_LIB)
    vp = ctypes.c_void_p
    lib.wrfmi_register_field.argtypes = [vp, ctypes.c_size_t]
    lib.wrfmi_unregister_field.argtypes = [vp]
    lib.wrfmi_sync_to_host.argtypes = [vp]
    lib.wrfmi_sync_to_device.argtypes = [vp]
    a = np.arange(100, dtype=np.float32)
    assert lib.wrfmi_register_field(a.ctypes.data, a.nbytes) == 0
    a[:] = 7.0
    assert lib.wrfmi_register_field(a.ctypes.data, a.nbytes) == 0  # refresh from the host
    a[:] = 0.0
    assert lib.wrfmi_sync_to_host(a.ctypes.data) == 0
    assert np.all(a == 7.0)
    assert lib.wrfmi_register_field(a[10:].ctypes.data, 40) == 4  # overlaps: WRFMI_E_POINTER
    assert lib.wrfmi_unregister_field(a.ctypes.data) == 0
    assert lib.wrfmi_unregister_field(a.ctypes.data) == 4
    assert lib.wrfmi_register_field(a[10:].ctypes.data, 40) == 0
    print("OK registry", flush=True)


def ensemble_check(order: int, nsteps: int, pd: bool = False):
    """Batched ensemble (wrfmi_ensemble_*): members of different sizes advanced together
    are bit for bit the members advanced one by one; mismatched members are refused.  With
    `pd`, every member's species take the positive-definite transport."""
    import ctypes

    from wrf_amd.fields import Config, Dims, State32
    specs = [grid.DomainSpec(nx=14, ny=11, nz=10, bc="open", n_bubbles=1, seed=21),
             grid.DomainSpec(nx=9, ny=13, nz=10, bc="open", n_bubbles=1, bubble_amp=1.5, seed=22),
             grid.DomainSpec(nx=12, ny=9, nz=12, bc="open", n_bubbles=1, bubble_amp=3.5, seed=23)]
    if pd:
        from pd_parity import pd_state
        for s in specs:
            s.n_passive = 1
        hss = [pd_state(s, 1, 1, seed=s.seed) for s in specs]
    else:
        hss = [grid.build_initial_state(s) for s in specs]
    singles = []
    for hs in hss:
        e = EmuDomain(hs, order=order)
        e.step(nsteps)
        singles.append(e.download())
        del e
    members = [EmuDomain(hs, order=order) for hs in hss]
    lib = members[0].lib
    vp = ctypes.c_void_p
    lib.wrfmi_ensemble_create.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.POINTER(vp)]
    lib.wrfmi_ensemble_step.argtypes = [vp, vp, ctypes.c_int]
    lib.wrfmi_ensemble_destroy.argtypes = [vp]
    n = len(members)
    states = (State32 * n)(*[m.state for m in members])
    cfgs = (Config * n)(*[m.cfg for m in members])
    dims = (Dims * n)(*[m.dims for m in members])
    ens = vp()
    rc = lib.wrfmi_ensemble_create(n, ctypes.byref(states), ctypes.byref(cfgs), ctypes.byref(dims), ctypes.byref(ens))
    assert rc == 0, lib.wrfmi_last_error().decode()
    for _ in range(nsteps):
        assert lib.wrfmi_ensemble_step(ens, None, 1) == 0, lib.wrfmi_last_error().decode()
    fails = []
    for m, (mem, exp) in enumerate(zip(members, singles)):
        got = mem.download()
        # a and alpha are calc_coef_w's work arrays: the per-routine entry points store them,
        # the whole step forms them inside the column solve without storing them
        bad = [k for k in exp if k not in ("a", "alpha") and not np.array_equal(got[k], exp[k])]
        if bad:
            fails.append(f"member {m} differs from its single-domain step in {bad}")
    assert lib.wrfmi_ensemble_destroy(ens) == 0
    assert lib.wrfmi_ensemble_destroy(ens) == 4  # no longer known: WRFMI_E_POINTER
    # a periodic member among open ones runs other kernels: refused
    hp = grid.build_initial_state(grid.DomainSpec(nx=10, ny=10, nz=10, bc="periodic", n_bubbles=1, seed=24))
    mp = EmuDomain(hp, order=order)
    states2 = (State32 * 2)(members[0].state, mp.state)
    cfgs2 = (Config * 2)(members[0].cfg, mp.cfg)
    dims2 = (Dims * 2)(members[0].dims, mp.dims)
    rc = lib.wrfmi_ensemble_create(2, ctypes.byref(states2), ctypes.byref(cfgs2), ctypes.byref(dims2), ctypes.byref(ens))
    if rc != 1:
        fails.append(f"mismatched members: expected WRFMI_E_BADOPT, got {rc}")
    if pd:
        # the same boundaries and species with the unlimited transport: other kernels, refused
        cfgs3 = (Config * 2)(members[0].cfg, members[1].cfg)
        cfgs3[1].moist_adv_opt = cfgs3[1].scalar_adv_opt = 0
        states3 = (State32 * 2)(members[0].state, members[1].state)
        dims3 = (Dims * 2)(members[0].dims, members[1].dims)
        rc = lib.wrfmi_ensemble_create(2, ctypes.byref(states3), ctypes.byref(cfgs3), ctypes.byref(dims3),
                                       ctypes.byref(ens))
        if rc != 1 or "transport" not in lib.wrfmi_last_error().decode():
            fails.append(f"members with different transport options: expected WRFMI_E_BADOPT, got {rc}")
    for f in fails:
        print("MISMATCH", f)
    print("OK ensemble" if not fails else "FAILED ensemble", "order", order, flush=True)
    sys.exit(1 if fails else 0)


def noalloc_check(order: int):
    """The allocation-free whole step (work arrays in the library's static array) is bit for
    bit the allocating whole step, on two domains in a row, the second smaller."""
    import ctypes
    cases = [dict(nx=22, ny=14, nz=10, bc="periodic", n_passive=2, n_bubbles=2, bubble_amp=2.0, seed=31),
             dict(nx=13, ny=12, nz=8, bc="open", n_bubbles=1, seed=32)]
    fails = []
    for n, kw in enumerate(cases):
        hs = grid.build_initial_state(grid.DomainSpec(**kw))
        ref = EmuDomain(hs, order=order)
        ref.step(2)
        want = ref.download()
        del ref
        e = EmuDomain(hs, order=order)
        fn = e.lib.wrfmi_solve_em_step_noalloc
        fn.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int]
        for _ in range(2):
            e._chk(fn(ctypes.byref(e.state), ctypes.byref(e.cfg), ctypes.byref(e.dims), None, 1))
        got = e.download()
        bad = [k for k in want if not np.array_equal(got[k], want[k])]
        if bad:
            fails.append(f"case {n}: the allocation-free step differs in {bad}")
        del e
    for f in fails:
        print("MISMATCH", f)
    print("OK noalloc" if not fails else "FAILED noalloc", "order", order, flush=True)
    sys.exit(1 if fails else 0)


def routines_check(case: str, order: int):
    """The entry points with the reference's argument lists against the oracle routines of
    the same name, then one RK3 step of the Fortran solve_em-shaped host (which calls them)
    against the oracle's whole step."""
    import ctypes
    from routine_parity import routine_parity
    hs = build_case(case)
    fails = []
    for rk in (1, 2):
        emu = EmuDomain(hs, order=order)
        fails += [f"rk_step {rk}: {f}" for f in routine_parity(emu, Oracle("f32cr").session(hs), hs, rk_step=rk)]
        del emu
    emu = EmuDomain(hs, order=order)
    flib = ctypes.CDLL(os.path.join(HERE, "cpu_emu", "libwrfmi_fortran_emu.so"))
    flib.wrfmi_fortran_solve_em_step.argtypes = [ctypes.c_void_p] * 3
    for _ in range(2):
        rc = flib.wrfmi_fortran_solve_em_step(ctypes.byref(emu.state), ctypes.byref(emu.cfg), ctypes.byref(emu.dims))
        if rc != 0:
            fails.append(f"Fortran host step failed ({rc}): {emu.lib.wrfmi_last_error().decode()}")
            break
    if not fails:
        got = emu.download()
        exp = Oracle("f32cr").run(hs, 2)
        bad = [k for k in exp if not np.array_equal(got[k], exp[k])]
        if bad:
            fails.append(f"Fortran host: two steps differ from the oracle in {bad}")
    for f in fails:
        print("MISMATCH", f)
    print("OK routines" if not fails else "FAILED routines", case, "order", order, flush=True)
    sys.exit(1 if fails else 0)


def coef_width_check():
    """A state whose calc_coef_w arrays hold 4-byte elements is refused in precision mode 1
    (whose column solve writes 8-byte gamma) before any launch, and accepted in mode 0."""
    import ctypes
    hs = grid.build_initial_state(grid.DomainSpec(nx=12, ny=10, nz=8, bc="periodic", n_bubbles=1, seed=41))
    emu = EmuDomain(hs, order=0)
    emu.lib.wrfmi_solve_em_step.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int]
    emu.cfg.precision_mode = 1
    rc = emu.lib.wrfmi_solve_em_step(ctypes.byref(emu.state), ctypes.byref(emu.cfg), ctypes.byref(emu.dims), None, 1)
    msg = emu.lib.wrfmi_last_error().decode()
    ok = rc == 4 and "8-byte" in msg
    emu.cfg.precision_mode = 0
    rc0 = emu.lib.wrfmi_solve_em_step(ctypes.byref(emu.state), ctypes.byref(emu.cfg), ctypes.byref(emu.dims), None, 1)
    ok = ok and rc0 == 0
    print("OK coefwidth" if ok else f"FAILED coefwidth rc={rc} msg={msg!r} rc0={rc0}", flush=True)
    sys.exit(0 if ok else 1)


def totals_check():
    """wrfmi_domain_totals against the same sums in float64 numpy, before and after two
    steps of a doubly periodic domain, over which dry mass and the theta budget are
    conserved (no diabatic heating, flux-form transport)."""
    import ctypes
    from parity import domain_totals_numpy
    hs = grid.build_initial_state(grid.DomainSpec(nx=18, ny=14, nz=10, bc="periodic", n_passive=1, n_bubbles=2,
                                                  bubble_amp=2.0, seed=43))
    emu = EmuDomain(hs, order=0)
    emu.lib.wrfmi_domain_totals.argtypes = [ctypes.c_void_p] * 5
    fails = []
    res = []
    for stage in range(2):
        out = (ctypes.c_double * 6)()
        emu._chk(emu.lib.wrfmi_domain_totals(ctypes.byref(emu.state), ctypes.byref(emu.cfg), ctypes.byref(emu.dims),
                                             None, out))
        got = np.array(out[:])
        exp = domain_totals_numpy(emu.download(), hs)
        tol = np.array([1e-12, 1e-6, 1e-12, 1e-12, 1e-12, 1e-12])
        rel = np.abs(got - exp) / np.abs(exp)
        if np.any(rel > tol):
            fails.append(f"stage {stage}: relative differences {rel} exceed {tol}")
        res.append(got)
        emu.step(2)
    drift_mass = abs(res[1][0] - res[0][0]) / res[0][0]
    drift_theta = abs(res[1][1] - res[0][1]) / res[0][1]
    if drift_mass > 1e-6 or drift_theta > 1e-5:
        fails.append(f"periodic drifts: dry mass {drift_mass:.2e}, theta budget {drift_theta:.2e}")
    for f in fails:
        print("MISMATCH", f)
    print("OK totals" if not fails else "FAILED totals", f"drift mass {drift_mass:.2e} theta {drift_theta:.2e}",
          flush=True)
    sys.exit(1 if fails else 0)


def history_check_emu(case: str, order: int):
    """wrfmi_history_* on the emulated kernels: two records (the initial state and the state
    after one step) read back with scipy's netCDF reader and compared with the arrays."""
    import tempfile
    from history_check import history_failures
    from wrf_amd.runtime import History
    hs = build_case(case)
    emu = EmuDomain(hs, order=order)
    start = "2026-06-01_12:00:00"
    recs = []
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "wrfout_d01")
        with History(emu.lib, emu.state, emu.cfg, emu.dims, path, start) as h:
            for r in range(2):
                date = f"2026-06-01_12:00:{6 * r:02d}"
                h.write(date, 0.1 * r)
                recs.append((date, 0.1 * r, emu.download()))
                if r == 0:
                    emu.step(1)
            assert h.records == 2
        fails = history_failures(path, recs, hs.dims, hs.cfg, hs.scal, hs.n_moist, hs.n_species, start)
    for f in fails:
        print("MISMATCH", f)
    print("OK history" if not fails else "FAILED history", case, flush=True)
    sys.exit(1 if fails else 0)


def shapes_check(case: str):
    """Tile-shape invariance: with every kernel's rows per workgroup at 2 and at 4 (and the
    stencil kernels' at 8), and the levels per thread and prefetch depth at each
    tunable value, one step is bit for bit the oracle's (the launch shapes the autotuner
    chooses among)."""
    import ctypes
    kernels = ["k_step_prep", "k_ww_scan", "k_tend_u", "k_tend_v", "k_tend_w", "k_tend_t", "k_small_prep", "k_p_rho0",
               "k_coef_w", "k_advance_uv", "k_ac_div", "k_ac_mu_ww", "k_ac_theta", "k_ac_wsolve", "k_small_finish",
               "k_adv_scalar", "k_update_scalar", "k_p_rho_phi"]
    default = {k: (1, 1, 0) for k in kernels}
    default.update({"k_tend_t": (4, 1, 0), "k_adv_scalar": (4, 1, 0), "k_advance_uv": (1, 4, 0), "k_ac_div": (1, 4, 0),
                    "k_ac_wsolve": (1, 1, 4)})
    sets = [("rows 2", {k: (2, v[1], v[2]) for k, v in default.items()}),
            ("rows 4", {k: (4, v[1], v[2]) for k, v in default.items()})]
    tall = ("k_tend_u", "k_tend_v", "k_tend_w", "k_tend_t", "k_adv_scalar")
    sets.append(("rows 8", {k: (8, 1, 0) for k in tall}))
    for lv in (2, 8):
        sets.append((f"levels {lv}", {"k_advance_uv": (1, lv, 0), "k_ac_div": (2, lv, 0)}))
    for pf in (2, 8):
        sets.append((f"prefetch {pf}", {"k_ac_wsolve": (2, 1, pf)}))
    hs = grid.build_initial_state(grid.DomainSpec(**CASES[case]))
    exp = Oracle("f32cr").run(hs, 1)
    fails = []
    for label, table in sets:
        emu = EmuDomain(hs, order=1)
        lib = emu.lib
        lib.wrfmi_tuning_set.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 6
        lib.wrfmi_tuning_clear()
        for k, (r, lv, pf) in table.items():
            emu._chk(lib.wrfmi_tuning_set(k.encode(), 0, 0, 0, r, lv, pf))
        emu.step(1)
        got = emu.download()
        lib.wrfmi_tuning_clear()
        bad = [k for k in exp if k not in ("a", "alpha") and not np.array_equal(got[k], exp[k])]
        if bad:
            fails.append(f"{label}: differs from the oracle in {bad}")
        del emu
    # shapes outside the offered set are refused
    import ctypes as C
    lib = C.CDLL(os.path.join(HERE, "cpu_emu", "libwrfmi_emu.so"))
    lib.wrfmi_tuning_set.argtypes = [C.c_char_p] + [C.c_int] * 6
    if lib.wrfmi_tuning_set(b"k_ac_wsolve", 0, 0, 0, 3, 1, 4) != 1 or lib.wrfmi_tuning_set(b"k_tend_u", 0, 0, 0, 1, 2, 0) != 1 \
            or lib.wrfmi_tuning_set(b"k_nonexistent", 0, 0, 0, 1, 1, 0) != 1 \
            or lib.wrfmi_tuning_set(b"k_ac_theta", 0, 0, 0, 8, 1, 0) != 1 or lib.wrfmi_tuning_set(b"k_tend_v", 0, 0, 0, 16, 1, 0) != 1:
        fails.append("a shape outside the offered set was accepted")
    for f in fails:
        print("MISMATCH", f)
    print("OK shapes" if not fails else "FAILED shapes", case, flush=True)
    sys.exit(1 if fails else 0)


def main():
    if sys.argv[1] == "shapes":
        shapes_check(sys.argv[2])
        return
    if sys.argv[1] == "history":
        history_check_emu(sys.argv[2], int(sys.argv[3]))
    if sys.argv[1] == "totals":
        totals_check()
        return
    if sys.argv[1] == "coefwidth":
        coef_width_check()
        return
    if sys.argv[1] == "routines":
        routines_check(sys.argv[2], int(sys.argv[3]))
        return
    if sys.argv[1] == "noalloc":
        noalloc_check(int(sys.argv[2]))
        return
    if sys.argv[1] == "registry":
        registry_check()
        return
    if sys.argv[1] in ("ensemble", "ensemble_pd"):
        ensemble_check(int(sys.argv[2]), int(sys.argv[3]), pd=sys.argv[1] == "ensemble_pd")
        return
    if sys.argv[1] == "pd_routines":
        pd_routines_check(int(sys.argv[2]))
        return
    case, order, nsteps = sys.argv[1], int(sys.argv[2]), int(sys.argv[3])
    hs = build_case(case)
    emu = EmuDomain(hs, order=order)
    phases = runtime.step_phases(hs.cfg, hs.scal, hs.n_species, hs.n_moist)
    # bitwise against the oracle whose EOS rounds LOG and ** once from double, as the
    # kernels do; then the whole step through wrfmi_solve_em_step
    sess = Oracle("f32cr").session(hs)
    fails = phase_parity(emu, sess, phases, nsteps=nsteps, eos_ulps=0)
    if not fails:
        emu2 = EmuDomain(hs, order=order)
        emu2.step(nsteps)
        got = emu2.download()
        exp = Oracle("f32cr").run(hs, nsteps)
        # a and alpha are calc_coef_w's work arrays: the per-routine entry points store them,
        # the whole step forms them inside the column solve without storing them
        bad = [k for k in exp if k not in ("a", "alpha") and not np.array_equal(got[k], exp[k])]
        if bad:
            fails.append(f"whole step differs in {bad}")
    for f in fails:
        print("MISMATCH", f)
    print("OK" if not fails else "FAILED", case, "order", order, flush=True)
    sys.exit(1 if fails else 0)


if __name__ == "__main__":
    main()
