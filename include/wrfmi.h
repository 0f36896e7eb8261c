/*
 * wrfmi.h -- C-ABI of the MI355X (gfx950) WRF-ARW dynamics step.
 *
 * The boundary replaced is the reference's Fortran model-layer interface that
 * dyn_em/solve_em.F calls once per tile inside its OpenMP tile loops
 * (e.g. advect_u, dyn_em/module_advect_em.F:126-136).  The reference has no
 * plugin or FFI layer; every routine takes its arrays by reference with the
 * memory dims (ims:ime, kms:kme, jms:jme) and the ids..kte index sextets.
 * This header keeps that contract:
 *   - arrays are plain pointers to the first element of the Fortran array
 *     (column-major, i fastest, then k, then j), laid out exactly as the host
 *     allocated them, memory dims included;
 *   - pointers are device pointers, or host pointers previously registered with
 *     wrfmi_register_field (the library then uses the resident device mirror);
 *   - every entry point returns an int status: 0 = ok, otherwise a WRFMI_E_*
 *     code, with a message from wrfmi_last_error().  A Fortran host turns a
 *     non-zero status into wrf_error_fatal, as the reference does for bad
 *     options (dyn_em/module_advect_em.F:1237-1238).
 * All entry points are synchronous with respect to the host unless stated.
 */
#ifndef WRFMI_H
#define WRFMI_H

#include <stddef.h>
#include <stdint.h>
#include "wrfmi_fields.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  WRFMI_OK = 0,
  WRFMI_E_BADOPT = 1,    /* option outside the implemented set (order, BC kind, ...) */
  WRFMI_E_DIMS = 2,      /* index sextets inconsistent with each other or the halo */
  WRFMI_E_DEVICE = 3,    /* HIP runtime error */
  WRFMI_E_POINTER = 4,   /* null, or neither device nor registered host pointer */
  WRFMI_E_NONFINITE = 5  /* non-finite value found by wrfmi_check_finite */
};

/* Index sextets of the reference (dyn_em/solve_em.F:101-114):
 * domain ids..ide (ide = e_we, the u-staggered extent), memory ims..ime,
 * tile its..ite.  The GPU path processes the whole domain in one call, so the
 * tile is normally the patch (its = ids, ite = ide, ...). */
typedef struct wrfmi_dims {
  int ids, ide, jds, jde, kds, kde;
  int ims, ime, jms, jme, kms, kme;
  int its, ite, jts, jte, kts, kte;
} wrfmi_dims;

/* The grid_config_rec_type fields the in-scope routines read
 * (config_flags%<name>; defaults from Registry/Registry.EM_COMMON:2840-2902).
 * LOGICALs are ints (0/1). */
typedef struct wrfmi_config {
  int periodic_x, periodic_y;
  int symmetric_xs, symmetric_xe, symmetric_ys, symmetric_ye;
  int open_xs, open_xe, open_ys, open_ye;
  int specified, nested, polar;
  int rk_ord;
  int h_mom_adv_order, v_mom_adv_order, h_sca_adv_order, v_sca_adv_order;
  int moist_adv_opt, scalar_adv_opt, momentum_adv_opt;
  int non_hydrostatic, top_lid, phi_adv_z, hypsometric_opt, use_theta_m;
  int damp_opt, w_damping, diff_opt, diff_6th_opt, zadvect_implicit, pert_coriolis;
  int map_proj;
  int time_step_sound;   /* 0 = derive from dt and dx as solve_em.F:446-463 */
  int time_step;         /* namelist integer time_step (its sign enters the upwind fluxes) */
  float epssm, smdiv, emdiv, dampcoef, zdamp, w_crit_cfl;
  /* 0: FP32 everywhere, reference operation order.
   * 1: FP32 bulk, FP64 calc_coef_w + advance_w tridiagonal solve,
   *    compensated (two-sum) column mass-divergence integrals. */
  int precision_mode;
} wrfmi_config;

/* One domain's arrays.  Members are named after the reference's grid fields. */
typedef struct wrfmi_state {
#define WRFMI_DECL_PTR(n) float *n;
  WRFMI_FIELDS_3D(WRFMI_DECL_PTR)
  WRFMI_FIELDS_2D(WRFMI_DECL_PTR)
  WRFMI_FIELDS_1D(WRFMI_DECL_PTR)
#undef WRFMI_DECL_PTR
#define WRFMI_DECL_R(n) float n;
  WRFMI_SCALARS_R(WRFMI_DECL_R)
#undef WRFMI_DECL_R
  /* tridiagonal coefficients of calc_coef_w: float arrays when
   * precision_mode == 0, double arrays when precision_mode == 1 */
  void *a, *alpha, *gamma;
  /* advected species: moist species first (q[0] is QV), then passive scalars.
   * q = moist / scalar, q_old = moist_old / scalar_old, q_tend = *_tend */
  int n_moist, n_species;
  float *q[WRFMI_MAX_SPECIES];
  float *q_old[WRFMI_MAX_SPECIES];
  float *q_tend[WRFMI_MAX_SPECIES];
} wrfmi_state;

/* ---- library ----------------------------------------------------------- */
int wrfmi_init(void);
int wrfmi_finalize(void);
const char *wrfmi_last_error(void);
/* wrfmi_version: "<major>.<minor>" of this ABI */
const char *wrfmi_version(void);
/* number of fields in the state struct's X-macro lists, for ABI checks */
int wrfmi_state_field_count(void);
size_t wrfmi_state_sizeof(void);

/* ---- host-array residency (device mirrors keyed by host address) ---------
 * Replaces the per-call data movement a drop-in would otherwise need: a host
 * that owns the Fortran arrays registers each once, syncs at solve start/end
 * and at I/O steps, and passes its own host pointers to every entry point. */
int wrfmi_register_field(void *host_ptr, size_t bytes);
int wrfmi_unregister_field(void *host_ptr);
int wrfmi_sync_to_device(void *host_ptr);
int wrfmi_sync_to_host(void *host_ptr);

/* ---- whole dynamics step ------------------------------------------------
 * One Runge-Kutta (rk_ord = 3) dynamics time step of dyn_em/solve_em.F:578-3459
 * and the end-of-step diagnostics/boundary fills (solve_em.F:4180-4470) for the
 * in-scope options: rk_step_prep, rk_tendency (advect_u/v/w/scalar(theta),
 * rhs_ph, horizontal_pressure_gradient, pg_buoy_w, w_damp, coriolis, curvature),
 * rk_addtend_dry, small_step_prep, calc_p_rho, calc_coef_w, the acoustic loop
 * (advance_uv, advance_mu_t, advance_w, sumflux, calc_p_rho), calc_mu_uv_1,
 * small_step_finish, rk_scalar_tend/rk_update_scalar for every species,
 * calc_p_rho_phi and the set_physical_bc3d/2d halo fills.  The physics
 * tendencies *_tendf, mu_tendf and h_diabatic are read as given (the reference
 * zeroes them at rk_step 1 when physics is off: module_first_rk_step_part1.F:182).
 * `stream` is a hipStream_t (NULL = the default stream); the call returns
 * after enqueueing unless `sync` is non-zero. */
int wrfmi_solve_em_step(wrfmi_state *st, const wrfmi_config *cfg,
                        const wrfmi_dims *d, void *stream, int sync);

/* Same step, replayed from a hipGraph captured on the first call for this
 * (state, config, dims) triple.  Later calls must pass the same pointers. */
int wrfmi_solve_em_step_graph(wrfmi_state *st, const wrfmi_config *cfg,
                              const wrfmi_dims *d, void *stream, int sync);

/* Non-finite scan of the prognostic state (u_2, v_2, w_2, t_2, ph_2, mu_2);
 * returns WRFMI_E_NONFINITE and the count in *n_bad when any is found. */
int wrfmi_check_finite(const wrfmi_state *st, const wrfmi_dims *d, long long *n_bad);

#ifdef __cplusplus
}
#endif
#endif
