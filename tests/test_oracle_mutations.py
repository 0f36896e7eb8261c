"""Mutation test of the oracle's pins (CPU only) -- TEST INFRASTRUCTURE.

Each mutant is the FP64 oracle built from its own sources with one plausible mistake planted
in a copy (a flipped off-centring weight, a dropped 1/2, a face average shifted by one point,
interpolation weights swapped, a wrong level index, ...).  The test builds every mutant in a
temporary directory and runs the pins of tests/oracle_pins.py on it: each mutant must fail at
least one pin, which shows that the pins can tell a wrong oracle from the right one.  The
unmutated build must pass them all.  Nothing here touches the product library.
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import oracle_pins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SRCS = ["bc_prep.c", "advect.c", "advect_pd.c", "mixing.c", "lbc.c", "tend.c", "small.c", "scalar_eos.c", "solve_step.c"]
CFLAGS = ["-O1", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-w", "-DWREAL=double",
          "-DWREAL_IS_DOUBLE=1"]

# (name, file, text in the oracle source, replacement, occurrence: 1 = first, 0 = every one)
MUTANTS = [
    ("advance_w_epssm_flipped", "small.c",
     "((R(1.) + epssm) * (RHS(k + 1) - RHS(k)) + (R(1.) - epssm) * (A3(ph, i, k + 1, j) - A3(ph, i, k, j)))",
     "((R(1.) - epssm) * (RHS(k + 1) - RHS(k)) + (R(1.) + epssm) * (A3(ph, i, k + 1, j) - A3(ph, i, k, j)))", 1),
    ("coef_w_cof_1_minus_epssm", "small.c",
     "wreal cof = (R(.5) * dts * K_G * (R(1.) + epssm));", "wreal cof = (R(.5) * dts * K_G * (R(1.) - epssm));", 0),
    ("coef_w_a_level_index", "small.c",
     "A1(rdn, kk) * A1(rdnw, kk - 1) * A3(c2a, i, kk - 1, j)", "A1(rdn, kk) * A1(rdnw, kk) * A3(c2a, i, kk - 1, j)", 1),
    ("advance_w_buoyancy_level", "small.c",
     "- A3(c2a, i, k - 1, j) * A3(alt, i, k - 1, j) * A3(t_2ave, i, k - 1, j))",
     "- A3(c2a, i, k, j) * A3(alt, i, k - 1, j) * A3(t_2ave, i, k - 1, j))", 1),
    ("advance_uv_half_dropped", "small.c",
     "DPXY(k) = (A2(s->msfux, i, j) / A2(s->msfuy, i, j)) * R(.5) * rdx *",
     "DPXY(k) = (A2(s->msfux, i, j) / A2(s->msfuy, i, j)) * rdx *", 1),
    ("calc_p_rho_mu_theta_term_dropped", "small.c",
     "(A3(t_2, i, k, j) - (A1(c1h, k) * A2(mu, i, j)) * A3(t_1, i, k, j))", "(A3(t_2, i, k, j))", 1),
    ("advance_mu_t_dn_for_dnw", "small.c",
     "dmdt = dmdt + A1(s->dnw, k) * DV(k);", "dmdt = dmdt + A1(s->dn, k) * DV(k);", 1),
    ("small_step_finish_w_c1h", "small.c",
     "(A1(c1f, k) * A2(s->muts, i, j) + A1(c2f, k));", "(A1(c1h, k) * A2(s->muts, i, j) + A1(c2f, k));", 1),
    ("small_step_prep_t_mut_for_muts", "small.c",
     "A3(s->t_2, i, k, j) = (A1(c1h, k) * A2(s->muts, i, j) + A1(c2h, k)) * A3(s->t_1, i, k, j) -",
     "A3(s->t_2, i, k, j) = (A1(c1h, k) * A2(s->mut, i, j) + A1(c2h, k)) * A3(s->t_1, i, k, j) -", 1),
    ("sumflux_average_over_n_plus_1", "small.c",
     "A3(ru_m, i, k, j) = A3(ru_m, i, k, j) / ns +", "A3(ru_m, i, k, j) = A3(ru_m, i, k, j) / (ns + R(1.)) +", 1),
    ("advect_u_face_average_shifted", "advect.c",
     "wreal vel = R(0.5) * (A3(ru, i, k, j) + A3(ru, i - 1, k, j));",
     "wreal vel = R(0.5) * (A3(ru, i + 1, k, j) + A3(ru, i, k, j));", 1),
    ("advect_v_face_average_shifted", "advect.c",
     "wreal vel = R(0.5) * (A3(rv, i, k, j) + A3(rv, i, k, j - 1));",
     "wreal vel = R(0.5) * (A3(rv, i, k, j + 1) + A3(rv, i, k, j));", 1),
    ("advect_u_boundary_fzm_fzp_swapped", "advect.c",
     "(A1(fzm, k) * A3(u, i, k, j) + A1(fzp, k) * A3(u, i, k - 1, j))",
     "(A1(fzp, k) * A3(u, i, k, j) + A1(fzm, k) * A3(u, i, k - 1, j))", 1),
    ("advect_w_fzm_fzp_swapped", "advect.c",
     "((k) <= ktf ? A1(fzm, k) * A3(rr, i, k, j) + A1(fzp, k) * A3(rr, i, (k) - 1, j)",
     "((k) <= ktf ? A1(fzp, k) * A3(rr, i, k, j) + A1(fzm, k) * A3(rr, i, (k) - 1, j)", 1),
    ("advect_w_lid_full_cell", "advect.c",
     "A3(tend, i, k, j) + R(2.) * A1(rdzu, k - 1) * (T2(vflux, i, k));",
     "A3(tend, i, k, j) + A1(rdzu, k - 1) * (T2(vflux, i, k));", 1),
    ("rk_update_scalar_old_mu_for_new", "scalar_eos.c",
     "wreal munew = A2(s->mu_2, i, j) + A2(s->mub, i, j);", "wreal munew = A2(s->mu_1, i, j) + A2(s->mub, i, j);", 1),
    ("pd_limiter_scales_inflow_face", "advect_pd.c",
     "if (T3(fqx, i + 1, k, j) > R(0.)) T3(fqx, i + 1, k, j) = scale * T3(fqx, i + 1, k, j);",
     "if (T3(fqx, i + 1, k, j) < R(0.)) T3(fqx, i + 1, k, j) = scale * T3(fqx, i + 1, k, j);", 1),
    ("pd_low_order_base_mass_perturbation_dropped", "advect_pd.c",
     "+ (A1(c1, k) * A2(mu_old, i, j))) * A3(field_old, i, k, j)", "+ R(0.)) * A3(field_old, i, k, j)", 1),
    ("pd_upwind_flux_downwind", "advect_pd.c",
     "* qm1 + R(0.5) * rmax(R(-1.0), (cr - (wreal)fabs((double)cr))) * q0;",
     "* q0 + R(0.5) * rmax(R(-1.0), (cr - (wreal)fabs((double)cr))) * qm1;", 1),
    ("pd_flux5_upwind_term_sign", "advect_pd.c",
     "return pflux6(qm3, qm2, qm1, q0, qp1, qp2) -", "return pflux6(qm3, qm2, qm1, q0, qp1, qp2) +", 1),
    ("pd_prep_tendency_not_zeroed", "advect_pd.c", "A3(sc_tend, i, k, j) = R(0.);", "", 1),
    ("pd_prep_new_mass_for_old", "solve_step.c",
     "s->c1h, s->c2h, s->mu_1, s->mu_1, s->mub, dt_rk, &dt_);", "s->c1h, s->c2h, s->mu_2, s->mu_1, s->mub, dt_rk, &dt_);",
     1),
    ("sixth_order_coefficient_half_dropped", "mixing.c",
     "const wreal coef = diff_6th_factor * R(0.015625) / (R(2.0) * dt);",
     "const wreal coef = diff_6th_factor * R(0.015625) / dt;", 1),
    ("sixth_order_u_face_mass_shifted", "mixing.c", "m0 = MU(i - 1, j);", "m0 = MU(i, j);", 1),
    ("sixth_order_limiter_sign", "mixing.c",
     "if (fx0 * (F(i, k, j) - F(i - 1, k, j)) <= R(0.0)) fx0 = R(0.0);",
     "if (fx0 * (F(i, k, j) - F(i - 1, k, j)) >= R(0.0)) fx0 = R(0.0);", 1),
    ("rayleigh_interpolation_offset", "mixing.c",
     "(z00[k] - A1(z_base, k2)) / (A1(z_base, k2 + 1) - A1(z_base, k2));",
     "(z00[k] - A1(z_base, k2 + 1)) / (A1(z_base, k2 + 1) - A1(z_base, k2));", 1),
    ("rayleigh_full_layer", "mixing.c",
     "wreal dcoef = R(1.0) - rminw(R(1.0), (ztop - z00[k]) / zdamp);", "wreal dcoef = R(1.0);", 1),
    ("hdiff_u_mass_at_own_point", "mixing.c",
     "mkrdxm = (A2(msftx, i - 1, j) / A2(msfty, i - 1, j)) * CM(i - 1, j) * XK(i - 1, k, j) * rdx;",
     "mkrdxm = (A2(msftx, i - 1, j) / A2(msfty, i - 1, j)) * CM(i, j) * XK(i - 1, k, j) * rdx;", 1),
    ("hdiff_mass_k_average_half_dropped", "mixing.c",
     "mkrdxm = (A2(msfux, i, j) / A2(msfuy, i, j)) * R(0.5) * (XK(i, k, j) + XK(i - 1, k, j)) * R(0.5) *\n"
     "                   (CM(i, j) + CM(i - 1, j)) * rdx;",
     "mkrdxm = (A2(msfux, i, j) / A2(msfuy, i, j)) * (XK(i, k, j) + XK(i - 1, k, j)) * R(0.5) *\n"
     "                   (CM(i, j) + CM(i - 1, j)) * rdx;", 1),
    ("vdiff_flux_below_lowest_level_zeroed", "mixing.c", "vflux[kts - 1] = vflux[kts];\n      vflux[ktf] = R(0.);\n      for (int k = kts; k <= ktf; k++)",
     "vflux[kts - 1] = R(0.);\n      vflux[ktf] = R(0.);\n      for (int k = kts; k <= ktf; k++)", 1),
    ("vdiff_uv_base_sign", "mixing.c", "- A1(base, k + 1) + A1(base, k));", "+ A1(base, k + 1) - A1(base, k));", 1),
    ("vdiff_uv_level_density_unaveraged", "mixing.c", "(R(0.5) * (A3(alt, im, k, jm) + A3(alt, i, k, j))) *",
     "(R(0.5) * (A3(alt, i, k, j) + A3(alt, i, k, j))) *", 1),
    ("vdiff_w_top_level_skipped", "mixing.c", "for (int k = kts + 1; k <= ktf; k++)\n        A3(tend, i, k, j) = A3(tend, i, k, j) + A1(rdn, k)",
     "for (int k = kts + 1; k <= ktf - 1; k++)\n        A3(tend, i, k, j) = A3(tend, i, k, j) + A1(rdn, k)", 1),
    ("relax_row_below_taken_from_row_above", "lbc.c", "f3 = FLS(2, i, b_dist, i, k, j - 1),",
     "f3 = FLS(2, i, b_dist + 2, i, k, j - 1),", 1),
    ("relax_smoother_sign", "lbc.c", "+ fcx[(bd)] * (f0) - gcx[(bd)] *", "+ fcx[(bd)] * (f0) + gcx[(bd)] *", 1),
    ("spec_zone_y_sides_take_the_corners", "lbc.c", "b_limit = px_ ? 0 : b_dist, side = 2;", "b_limit = 0, side = 2;", 1),
    ("zero_grad_copies_the_frame_itself", "lbc.c",
     "A3(field, i, k, j) = A3(field, ii, k, side == 2 ? z.jbs + spec_zone : z.jbe - spec_zone);",
     "A3(field, i, k, j) = A3(field, ii, k, side == 2 ? z.jbs + spec_zone - 1 : z.jbe - spec_zone);", 1),
    ("spec_ph_old_mass_sign", "lbc.c", "A2(muts, i, j) - dt * A2(mu_tend, i, j);", "A2(muts, i, j) + dt * A2(mu_tend, i, j);", 1),
    ("flow_dep_inflow_kept", "lbc.c", "side == 2 ? A3(v, i, k, j) < R(0.)", "side == 2 ? A3(v, i, k, j) > R(0.)", 1),
    ("damp_opt3_half_depth", "small.c", "wreal hbot = htop - hdepth;", "wreal hbot = htop - R(0.5) * hdepth;", 1),
    ("pressure_gradient_half_dropped", "tend.c",
     "wreal dpx = (A2(s->msfux, i, j) / A2(s->msfuy, i, j)) * R(.5) * rdx *",
     "wreal dpx = (A2(s->msfux, i, j) / A2(s->msfuy, i, j)) * rdx *", 1),
]


def _build(tmp, name, patch):
    d = os.path.join(tmp, name)
    os.makedirs(os.path.join(d, "oracle"))
    shutil.copytree(os.path.join(ROOT, "include"), os.path.join(d, "include"))
    for f in SRCS + ["oracle.h"]:
        shutil.copy(os.path.join(ORACLE, f), os.path.join(d, "oracle", f))
    if patch is not None:
        fname, old, new, occ = patch
        path = os.path.join(d, "oracle", fname)
        text = open(path).read()
        n = text.count(old)
        if n == 0:
            return None, f"mutation text of {name} not found in oracle/{fname}"
        text = text.replace(old, new) if occ == 0 else text.replace(old, new, 1)
        open(path, "w").write(text)
    out = os.path.join(d, "liboracle_mut.so")
    cmd = ["gcc"] + CFLAGS + ["-shared", "-o", out] + [os.path.join(d, "oracle", f) for f in SRCS] + ["-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return None, r.stderr[-2000:]
    return out, None


def _failing_pins(path):
    from pyoracle import Oracle
    o = Oracle("f64", path=path)
    failed = []
    for pin in oracle_pins.PINS:
        try:
            pin(o)
        except AssertionError:
            failed.append(pin.__name__)
        except (RuntimeError, FloatingPointError, ValueError, ZeroDivisionError):
            failed.append(pin.__name__)
    return failed


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("oracle_mutants"))
    jobs = [("unmutated", None)] + [(m[0], m[1:]) for m in MUTANTS]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        res = list(ex.map(lambda j: _build(tmp, *j), jobs))
    return {name: r for (name, _), r in zip(jobs, res)}


def test_unmutated_build_passes_every_pin(builds):
    path, err = builds["unmutated"]
    assert path, err
    assert _failing_pins(path) == []


@pytest.mark.parametrize("mutant", [m[0] for m in MUTANTS])
def test_every_mutant_fails_a_pin(builds, mutant):
    path, err = builds[mutant]
    assert path, err
    failed = _failing_pins(path)
    assert failed, f"no pin catches the mutant {mutant}"
    print(f"{mutant}: caught by {', '.join(failed)}")
