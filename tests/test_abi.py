"""The C-ABI library loads on a host without a GPU, exports every entry point that
include/wrfmi.h declares, agrees with the Python struct mirrors, and rejects options
outside the implemented set before touching the device."""
import ctypes
import os
import re

import pytest

from wrf_amd import fields, grid, runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = set()
    for h in ("wrfmi.h", "wrfmi_routines.h", "wrfmi_history.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(wrfmi_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_entry_points()
    assert "wrfmi_solve_em_step" in names and "wrfmi_advance_mu_t_w" in names
    for routine in ("advect_u", "advect_v", "advect_w", "advect_scalar", "rk_update_scalar", "advance_uv",
                    "advance_mu_t", "advance_w", "sumflux", "calc_p_rho", "calc_coef_w"):
        assert "wrfmi_" + routine in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    lib = runtime.load_library()
    missing = [n for n in declared_entry_points() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_struct_layouts_match():
    lib = runtime.load_library()
    assert lib.wrfmi_state_sizeof() == ctypes.sizeof(fields.State32)
    assert lib.wrfmi_state_field_count() == fields.state_field_count()
    assert lib.wrfmi_version() == b"1.0"
    names = [lib.wrfmi_kernel_name(k).decode() for k in range(lib.wrfmi_kernel_count())]
    assert {"k_advance_uv", "k_ac_wsolve", "k_adv_scalar", "k_halo"} <= set(names)


def test_ensemble_arguments_checked_before_the_device():
    lib = runtime.load_library()
    out = ctypes.c_void_p()
    st, cfg, dims = fields.State32(), fields.Config(), fields.Dims()
    rc = lib.wrfmi_ensemble_create(0, ctypes.byref(st), ctypes.byref(cfg), ctypes.byref(dims), ctypes.byref(out))
    assert rc == runtime.WRFMI_E_BADOPT
    rc = lib.wrfmi_ensemble_create(1, None, ctypes.byref(cfg), ctypes.byref(dims), ctypes.byref(out))
    assert rc == runtime.WRFMI_E_POINTER
    assert lib.wrfmi_ensemble_step(ctypes.c_void_p(1234), None, 0) == runtime.WRFMI_E_POINTER


def test_constant_coefficient_diffusion_only():
    """diff_opt 1 is implemented with km_opt 1 (constant khdif, kvdif); other eddy-coefficient
    schemes are refused before any launch."""
    lib = runtime.load_library()
    spec = grid.DomainSpec(nx=10, ny=10, nz=6)
    cfg, dims = grid.make_config(spec), grid.make_dims(spec)
    cfg.diff_opt, cfg.km_opt = 1, 4
    st = fields.State32()
    rc = lib.wrfmi_solve_em_step(ctypes.byref(st), ctypes.byref(cfg), ctypes.byref(dims), None, 1)
    assert rc == runtime.WRFMI_E_BADOPT
    assert b"km_opt" in lib.wrfmi_last_error()


def _specified(lib):
    spec = grid.DomainSpec(nx=12, ny=10, nz=6, bc="specified")
    cfg, dims = grid.make_config(spec), grid.make_dims(spec)
    return cfg, dims


def test_specified_options_checked():
    """Specified boundaries: spec_zone 1 and relax_zone >= 2 only; the boundary arrays are
    required; graph replay and ensembles refuse them (dtbc advances every step)."""
    lib = runtime.load_library()
    cfg, dims = _specified(lib)
    st = fields.State32()
    cfg.spec_zone = 2
    rc = lib.wrfmi_solve_em_step(ctypes.byref(st), ctypes.byref(cfg), ctypes.byref(dims), None, 1)
    assert rc == runtime.WRFMI_E_BADOPT and b"spec_zone" in lib.wrfmi_last_error()
    cfg.spec_zone = 1
    rc = lib.wrfmi_solve_em_step_graph(ctypes.byref(st), ctypes.byref(cfg), ctypes.byref(dims), None, 1)
    assert rc == runtime.WRFMI_E_BADOPT and b"graph" in lib.wrfmi_last_error()
    h = ctypes.c_void_p()
    rc = lib.wrfmi_ensemble_create(1, ctypes.byref(st), ctypes.byref(cfg), ctypes.byref(dims), ctypes.byref(h))
    assert rc == runtime.WRFMI_E_BADOPT and b"specified" in lib.wrfmi_last_error()


@pytest.mark.parametrize("field,value,msg", [
    ("rk_ord", 2, "rk_ord"),
    ("h_mom_adv_order", 6, "orders"),
    ("specified", 1, "specified"),
    ("nested", 1, "lateral"),
    ("diff_opt", 2, "TKE"),
    ("damp_opt", 1, "damp_opt"),
    ("moist_adv_opt", 2, "advection options"),
    ("scalar_adv_opt", 3, "advection options"),
])
def test_unsupported_options_are_rejected(field, value, msg):
    lib = runtime.load_library()
    spec = grid.DomainSpec(nx=10, ny=10, nz=6)
    cfg = grid.make_config(spec)
    dims = grid.make_dims(spec)
    setattr(cfg, field, value)
    st = fields.State32()
    rc = lib.wrfmi_solve_em_step(ctypes.byref(st), ctypes.byref(cfg), ctypes.byref(dims), None, 1)
    assert rc == runtime.WRFMI_E_BADOPT
    assert msg in lib.wrfmi_last_error().decode()


def test_bad_dims_are_rejected():
    lib = runtime.load_library()
    spec = grid.DomainSpec(nx=10, ny=10, nz=6)
    cfg = grid.make_config(spec)
    dims = grid.make_dims(spec)
    dims.ims = dims.ids - 2  # halo narrower than the 4 points the stencils need
    st = fields.State32()
    rc = lib.wrfmi_solve_em_step(ctypes.byref(st), ctypes.byref(cfg), ctypes.byref(dims), None, 1)
    assert rc == runtime.WRFMI_E_DIMS


def test_step_phase_sequence():
    spec = grid.supercell_c0()
    hs_cfg = grid.make_config(spec)
    ph = runtime.step_phases(hs_cfg, dict(dt=12.0, dx=2000.0, dy=2000.0), 1)
    names = [n for n, _ in ph]
    # 1 + 3 + 6 acoustic substeps for time_step_sound = 6
    assert names.count("advance_mu_t_w") == 10
    assert names.count("rk_scalar_step") == 3
    assert names[-3:] == ["calc_p_rho_phi", "phys_bc_dry_2", "set_w_surface"]
    assert runtime.num_sound_steps(type(hs_cfg)(), 6.0, 1000.0, 1000.0) == 4
    # positive-definite species: rk_update_scalar_pd's group once, before the last substep's species
    hs_cfg.scalar_adv_opt = 1
    ph = runtime.step_phases(hs_cfg, dict(dt=12.0, dx=2000.0, dy=2000.0), 3, n_moist=1)
    names = [n for n, _ in ph]
    assert names.count("rk_scalar_pd_prep") == 1
    i = names.index("rk_scalar_pd_prep")
    assert names[i - 1] == "small_step_finish" and names[i + 1:i + 4] == ["rk_scalar_step"] * 3
    assert [a for n, a in ph if n == "rk_scalar_step"][-1][1] == 3
    hs_cfg.scalar_adv_opt = 0
    assert "rk_scalar_pd_prep" not in [n for n, _ in runtime.step_phases(hs_cfg, dict(dt=12.0, dx=2000.0, dy=2000.0), 3)]


def test_reference_argument_entry_points_check_before_the_device():
    """A null array argument is refused with WRFMI_E_POINTER, and an option outside the
    implemented set with WRFMI_E_BADOPT, before any launch (no GPU needed)."""
    spec = grid.DomainSpec(nx=10, ny=10, nz=6)
    cfg, dims = grid.make_config(spec), grid.make_dims(spec)
    lib = runtime.load_library()
    d = [ctypes.c_int(getattr(dims, n)) for n, _ in dims._fields_]
    dref = [ctypes.byref(x) for x in d]
    one = ctypes.c_float(1.0)
    ts = ctypes.c_int(6)
    args = [None] * 9 + [ctypes.byref(ts), ctypes.byref(cfg)] + [None] * 8 + [ctypes.byref(one)] * 2 + [None]
    assert lib.wrfmi_advect_u(*(args + dref)) == runtime.WRFMI_E_POINTER
    cfg.h_mom_adv_order = 3
    assert lib.wrfmi_advect_u(*(args + dref)) == runtime.WRFMI_E_BADOPT
