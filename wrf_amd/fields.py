"""Field lists of one WRF-ARW dynamics domain and the ctypes mirror of the C-ABI state.

The single source of truth is ``include/wrfmi_fields.h`` (X-macro lists of the
reference's grid%<name> arrays, Registry/Registry.EM_COMMON:158-463); this module
parses it so that the Python host, the C-ABI struct ``wrfmi_state`` and any other
consumer agree on member order.
"""
from __future__ import annotations

import ctypes
import os
import re

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIELDS_HEADER = os.path.join(_ROOT, "include", "wrfmi_fields.h")


def _parse_fields(path: str = FIELDS_HEADER) -> dict:
    text = open(path).read()
    out = {}
    for name in ("WRFMI_FIELDS_3D", "WRFMI_FIELDS_2D", "WRFMI_FIELDS_1D", "WRFMI_SCALARS_R"):
        m = re.search(r"#define\s+%s\(X\)(.*?)(?:\n\s*\n|\n/\*)" % name, text, re.S)
        if not m:
            raise RuntimeError(f"{name} not found in {path}")
        out[name] = re.findall(r"X\((\w+)\)", m.group(1))
    m = re.search(r"#define\s+WRFMI_MAX_SPECIES\s+(\d+)", text)
    out["MAX_SPECIES"] = int(m.group(1))
    return out


_F = _parse_fields()
FIELDS_3D: list = _F["WRFMI_FIELDS_3D"]
FIELDS_2D: list = _F["WRFMI_FIELDS_2D"]
FIELDS_1D: list = _F["WRFMI_FIELDS_1D"]
SCALARS_R: list = _F["WRFMI_SCALARS_R"]
MAX_SPECIES: int = _F["MAX_SPECIES"]
SPECIES_ARRAYS = ("q", "q_old", "q_tend")


class Dims(ctypes.Structure):
    """wrfmi_dims: the reference's ids..kte index sextets (dyn_em/solve_em.F:101-114)."""

    _fields_ = [(n, ctypes.c_int) for n in (
        "ids", "ide", "jds", "jde", "kds", "kde",
        "ims", "ime", "jms", "jme", "kms", "kme",
        "its", "ite", "jts", "jte", "kts", "kte")]


_CFG_INT = ("periodic_x", "periodic_y", "symmetric_xs", "symmetric_xe", "symmetric_ys", "symmetric_ye",
            "open_xs", "open_xe", "open_ys", "open_ye", "specified", "nested", "polar", "rk_ord",
            "h_mom_adv_order", "v_mom_adv_order", "h_sca_adv_order", "v_sca_adv_order",
            "moist_adv_opt", "scalar_adv_opt", "momentum_adv_opt",
            "non_hydrostatic", "top_lid", "phi_adv_z", "hypsometric_opt", "use_theta_m",
            "damp_opt", "w_damping", "diff_opt", "diff_6th_opt", "zadvect_implicit", "pert_coriolis",
            "map_proj", "time_step_sound", "time_step", "diff_6th_slopeopt", "km_opt", "bl_pbl_physics")
_CFG_INT2 = ("spec_bdy_width", "spec_zone", "relax_zone", "have_bcs_moist", "have_bcs_scalar")
_CFG_REAL = ("epssm", "smdiv", "emdiv", "dampcoef", "zdamp", "w_crit_cfl", "diff_6th_factor", "diff_6th_thresh",
             "khdif", "kvdif")


class Config(ctypes.Structure):
    """wrfmi_config: the config_flags fields the in-scope routines read."""

    _fields_ = ([(n, ctypes.c_int) for n in _CFG_INT] + [(n, ctypes.c_float) for n in _CFG_REAL]
                + [(n, ctypes.c_int) for n in _CFG_INT2] + [("precision_mode", ctypes.c_int)])


def make_state_struct(real_ctype, name: str):
    """Build the ctypes mirror of ``wrfmi_state`` (real_ctype=c_float) or of a struct
    with the same member order whose arrays/scalars have another REAL type."""
    P = ctypes.POINTER(real_ctype)
    fields = [(n, P) for n in FIELDS_3D + FIELDS_2D + FIELDS_1D]
    fields += [(n, real_ctype) for n in SCALARS_R]
    fields += [("a", ctypes.c_void_p), ("alpha", ctypes.c_void_p), ("gamma", ctypes.c_void_p)]
    fields += [("n_moist", ctypes.c_int), ("n_species", ctypes.c_int)]
    fields += [(n, P * MAX_SPECIES) for n in SPECIES_ARRAYS]
    fields += [("lbc", ctypes.c_void_p)]
    return type(name, (ctypes.Structure,), {"_fields_": fields})


MAX_BDY_WIDTH = 16
LBC_FIELDS = ("u", "v", "ph", "t", "w", "mu")


def make_lbc_struct(real_ctype, name: str):
    """The ctypes mirror of ``wrfmi_lbc`` (real_ctype=c_float) or of the oracle's olbc."""
    P = ctypes.POINTER(real_ctype)
    fields = [("dtbc", real_ctype), ("fcx", real_ctype * MAX_BDY_WIDTH), ("gcx", real_ctype * MAX_BDY_WIDTH)]
    for n in LBC_FIELDS:
        fields += [(n + "_b", P * 4), (n + "_bt", P * 4)]
    fields += [("q_b", (P * 4) * MAX_SPECIES), ("q_bt", (P * 4) * MAX_SPECIES)]
    return type(name, (ctypes.Structure,), {"_fields_": fields})


State32 = make_state_struct(ctypes.c_float, "State32")
State64 = make_state_struct(ctypes.c_double, "State64")


def state_field_count() -> int:
    return len(FIELDS_3D) + len(FIELDS_2D) + len(FIELDS_1D)


Lbc32 = make_lbc_struct(ctypes.c_float, "Lbc32")

