/*
 * Oracle: one RK3 dynamics time step in the call order of dyn_em/solve_em.F.
 * TEST INFRASTRUCTURE.
 *   acoustic step count   solve_em.F:446-465
 *   RK substep sizes      solve_em.F:601-615
 *   rk_step_prep + bc     solve_em.F:652-782 (the rho halo fill is omitted: rho is not read)
 *   rk_tendency           solve_em.F:876-911
 *   rk_addtend_dry        solve_em.F:967-980
 *   small_step_prep, calc_p_rho(0), calc_coef_w   solve_em.F:1089-1147
 *   halo fills                                    solve_em.F:1181-1264
 *   acoustic loop          solve_em.F:1266-1719
 *   calc_mu_uv_1, small_step_finish, rk3 fills    solve_em.F:1732-1800
 *   positive-definite scalar pre-update + fills   solve_em.F:1850-1975
 *   scalar advection/update                       solve_em.F:2198-2364, 2777-2950
 *   calc_p_rho_phi                                solve_em.F:2959-2979
 *   rk_phys_bc_dry_2 + species fills (rk < 3)     solve_em.F:3240-3340
 *   end-of-step calc_p_rho_phi + set_phys_bc_dry_2 + species fills  solve_em.F:4180-4540
 *   set_w_surface (surface w from the final u, v and the terrain)   solve_em.F:4734-4750
 * Physics is off: the *_tendf, mu_tendf, h_diabatic and species tendencies are read
 * as given (module_first_rk_step_part1.F:182 zeroes them before the physics fills them).
 */
#include <stdlib.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

/* Tiling as in solve_em: every compute routine is called once per tile inside an OpenMP
 * loop over the tiles (e.g. solve_em.F:876-911 for rk_tendency), with the tile's
 * its..ite / jts..jte in the index sextets.  The tiles split the domain's j range into
 * num_tiles rows of tiles (the 1-D "Y" strategy of frame/module_tiles.F:334-347, with
 * num_tiles = the OpenMP thread count, module_tiles.F:271-278); the last tile ends at
 * jde, the staggered end.  Results do not depend on the tiling (each tile writes only its
 * own points from inputs of earlier phases).  The halo fills stay whole-domain calls. */
static int g_num_tiles = 0;  /* 0: one tile per OpenMP thread */
int oracle_set_num_tiles(int n) { g_num_tiles = n < 0 ? 0 : n; return 0; }
int oracle_num_tiles(const dims_t *d) {
  int n = g_num_tiles;
#ifdef _OPENMP
  if (n == 0) n = omp_get_max_threads();
#else
  if (n == 0) n = 1;
#endif
  const int rows = d->jde - d->jds + 1;
  if (n > rows / 2) n = rows / 2 > 0 ? rows / 2 : 1;
  return n;
}
static dims_t tile_dims(const dims_t *d, int t, int n) {
  dims_t r = *d;
  const int rows = d->jde - d->jds + 1, base = rows / n, extra = rows % n;
  r.jts = d->jds + t * base + (t < extra ? t : extra);
  r.jte = r.jts + base + (t < extra ? 1 : 0) - 1;
  return r;
}
/* run CALL once per tile with `d` bound to the tile's dims; status codes are OR-ed */
#define TILED(CALL)                                                     \
  do {                                                                  \
    const dims_t *dfull_ = d;                                           \
    const int nt_ = oracle_num_tiles(dfull_);                           \
    int rct_ = 0;                                                       \
    _Pragma("omp parallel for schedule(static) reduction(|:rct_)")      \
    for (int t_ = 0; t_ < nt_; t_++) {                                  \
      const dims_t dt_ = tile_dims(dfull_, t_, nt_);                    \
      const dims_t *d = &dt_;                                           \
      rct_ |= (CALL);                                                   \
    }                                                                   \
    if (rct_) return rct_;                                              \
  } while (0)
#define TILEDV(CALL) TILED(((CALL), 0))

int o_num_sound_steps(const cfg_t *c, const ostate *s) {
  if (c->time_step_sound != 0) return c->time_step_sound;
  wreal spacing = s->dx < s->dy ? s->dx : s->dy;
  int n = 2 * ((int)(R(300.) * s->dt / spacing - R(0.01)) + 1);
  return n > 4 ? n : 4;
}

static void bc_species(ostate *s, const cfg_t *c, const dims_t *d) {
  for (int m = 0; m < s->n_species; m++) o_set_physical_bc3d(s->q[m], 'p', c, d);
}

static int pd_species(const ostate *s, const cfg_t *c, int m) {
  return (m < s->n_moist ? c->moist_adv_opt : c->scalar_adv_opt) == 1;
}

/* solve_em.F:1850-1975: before the last RK substep, a species advected in positive-definite
 * form takes its physics tendency into its time-n values (rk_update_scalar_pd with mu_1 as
 * both masses, tendency zeroed), then those values get their halo */
static void pd_prep(ostate *s, const cfg_t *c, const dims_t *d, wreal dt_rk) {
  for (int m = 0; m < s->n_species; m++) {
    if (!pd_species(s, c, m)) continue;
    for (int t_ = 0, nt_ = oracle_num_tiles(d); t_ < nt_; t_++) {
      const dims_t dt_ = tile_dims(d, t_, nt_);
      o_rk_update_scalar_pd(s->q_old[m], s->q_tend[m], s->c1h, s->c2h, s->mu_1, s->mu_1, s->mub, dt_rk, &dt_);
    }
    o_set_physical_bc3d(s->q_old[m], 'p', c, d);
  }
}

/* the spec-zone updates of the acoustic substep (solve_em.F:1343-1365, 1459-1490,
 * 1583-1620) */
static void spec_uv(ostate *s, const cfg_t *c, const dims_t *d, wreal dts_rk) {
  o_spec_bdyupdate(s->u_2, s->ru_tend, dts_rk, 'u', c, c->spec_zone, d);
  o_spec_bdyupdate(s->v_2, s->rv_tend, dts_rk, 'v', c, c->spec_zone, d);
}
static void spec_mu_t(ostate *s, const cfg_t *c, const dims_t *d, wreal dts_rk) {
  o_spec_bdyupdate(s->t_2, s->t_tend, dts_rk, 't', c, c->spec_zone, d);
  o_spec_bdyupdate(s->mu_2, s->mu_tend, dts_rk, 'm', c, c->spec_zone, d);
  o_spec_bdyupdate(s->muts, s->mu_tend, dts_rk, 'm', c, c->spec_zone, d);
}
static void spec_w_ph(ostate *s, const cfg_t *c, const dims_t *d, wreal dts_rk) {
  if (!c->non_hydrostatic) return;
  o_spec_bdyupdate_ph(s->ph_save, s->ph_2, s->ph_tend, s->mu_tend, s->muts, s->c1f, s->c2f, dts_rk, c, c->spec_zone, d);
  if (c->specified) o_zero_grad_bdy(s->w_2, 'w', c, c->spec_zone, d);
  else o_spec_bdyupdate(s->w_2, s->rw_tend, dts_rk, 'h', c, c->spec_zone, d);
}

int o_solve_em_step(ostate *s, const cfg_t *c, const dims_t *d) {
  int rc;
  const int lbc = c->specified || c->nested;
  if (c->rk_ord != 3 || c->polar || (lbc && !s->lbc)) return WRFMI_E_BADOPT;
  /* solve_em.F:377 */
  if (lbc) s->lbc->dtbc = s->lbc->dtbc + s->dt;
  const int rk_order = 3;
  int nsound = o_num_sound_steps(c, s);
  wreal dts = s->dt / (wreal)nsound;
  for (int rk_step = 1; rk_step <= rk_order; rk_step++) {
    wreal dt_rk, dts_rk;
    int nsmall;
    if (rk_step == 1) { dt_rk = s->dt / R(3.); dts_rk = dt_rk; nsmall = 1; }
    else if (rk_step == 2) { dt_rk = R(0.5) * s->dt; dts_rk = dts; nsmall = nsound / 2; }
    else { dt_rk = s->dt; dts_rk = dts; nsmall = nsound; }

    TILEDV(o_rk_step_prep(s, c, d, rk_step));
    o_rk_phys_bc_dry_1(s, c, d);
    o_set_physical_bc3d(s->al, 'p', c, d);
    o_set_physical_bc3d(s->ph_2, 'w', c, d);

    TILED(o_rk_tendency(s, c, d, rk_step));
    /* solve_em.F:937-1008: relaxation at rk_step 1 into the *_save arrays and mu_tend,
     * then the spec zone's tendencies */
    if (lbc && rk_step == 1) o_relax_bdy_dry(s, c, d);
    TILEDV(o_rk_addtend_dry(s, d, rk_step));
    if (lbc) o_spec_bdy_dry(s, c, d);

    TILEDV(o_small_step_prep(s, d, rk_step));
    TILEDV(o_calc_p_rho(s, c, d, 0));
    if (c->non_hydrostatic) TILEDV(o_calc_coef_w(s, c, d, dts_rk));

    o_set_physical_bc3d(s->ru_tend, 'u', c, d);
    o_set_physical_bc3d(s->rv_tend, 'v', c, d);
    o_set_physical_bc3d(s->ph_2, 'w', c, d);
    o_set_physical_bc3d(s->al, 'p', c, d);
    o_set_physical_bc3d(s->p, 'p', c, d);
    o_set_physical_bc3d(s->t_1, 'p', c, d);
    o_set_physical_bc3d(s->t_save, 't', c, d);
    o_set_physical_bc2d(s->mu_1, 't', c, d);
    o_set_physical_bc2d(s->mu_2, 't', c, d);
    o_set_physical_bc2d(s->mudf, 't', c, d);

    for (int it = 1; it <= nsmall; it++) {
      TILEDV(o_advance_uv(s, c, d, dts_rk));
      if (lbc) spec_uv(s, c, d, dts_rk);
      TILEDV(o_advance_mu_t(s, c, d, dts_rk));
      if (lbc) spec_mu_t(s, c, d, dts_rk);
      if (c->non_hydrostatic) TILEDV(o_advance_w(s, c, d, dts_rk));
      TILEDV(o_sumflux(s, d, it, nsmall));
      if (lbc) spec_w_ph(s, c, d, dts_rk);
      TILEDV(o_calc_p_rho(s, c, d, it));
      o_set_physical_bc3d(s->ph_2, 'w', c, d);
      o_set_physical_bc3d(s->al, 'p', c, d);
      o_set_physical_bc3d(s->p, 'p', c, d);
      o_set_physical_bc2d(s->muts, 't', c, d);
      o_set_physical_bc2d(s->mu_2, 't', c, d);
      o_set_physical_bc2d(s->mudf, 't', c, d);
    }

    TILEDV(o_calc_mu_uv_1(s, c, d));
    TILEDV(o_small_step_finish(s, d, nsmall, dts_rk, rk_step, rk_order));
    if (rk_step == rk_order) {
      o_set_physical_bc3d(s->ru_m, 'u', c, d);
      o_set_physical_bc3d(s->rv_m, 'v', c, d);
      o_set_physical_bc3d(s->ww_m, 'w', c, d);
      o_set_physical_bc2d(s->mut, 't', c, d);
      o_set_physical_bc2d(s->muts, 't', c, d);
    }

    if (rk_step == rk_order) pd_prep(s, c, d, dt_rk);
    for (int m = 0; m < s->n_species; m++) {
      TILED(o_rk_scalar_tend(s, c, d, m, rk_step, dt_rk));
      /* solve_em.F:2261-2296, 2808-2845 */
      if (lbc && rk_step == 1 && o_species_has_bcs(s, c, m)) o_relax_spec_bdy_scalar(s, c, d, m);
      TILEDV(o_rk_update_scalar(s, d, m, rk_step, dt_rk));
      /* solve_em.F:2342-2355, 2886-2930 */
      if (c->specified && !o_species_has_bcs(s, c, m)) o_flow_dep_bdy(s->q[m], s->ru_m, s->rv_m, c, c->spec_zone, d);
    }

    TILEDV(o_calc_p_rho_phi(s, c, d));

    if (rk_step < rk_order) {
      o_set_physical_bc3d(s->u_2, 'u', c, d);
      o_set_physical_bc3d(s->v_2, 'v', c, d);
      o_set_physical_bc3d(s->w_2, 'w', c, d);
      o_set_physical_bc3d(s->t_2, 'p', c, d);
      o_set_physical_bc3d(s->ph_2, 'w', c, d);
      o_set_physical_bc2d(s->mu_2, 't', c, d);
      bc_species(s, c, d);
    }
  }
  /* solve_em.F:4180-4207 recomputes calc_p_rho_phi from unchanged inputs (same values) */
  TILEDV(o_calc_p_rho_phi(s, c, d));
  o_set_physical_bc3d(s->u_1, 'u', c, d);
  o_set_physical_bc3d(s->u_2, 'u', c, d);
  o_set_physical_bc3d(s->v_1, 'v', c, d);
  o_set_physical_bc3d(s->v_2, 'v', c, d);
  o_set_physical_bc3d(s->w_1, 'w', c, d);
  o_set_physical_bc3d(s->w_2, 'w', c, d);
  o_set_physical_bc3d(s->t_1, 'p', c, d);
  o_set_physical_bc3d(s->t_2, 'p', c, d);
  o_set_physical_bc3d(s->ph_1, 'w', c, d);
  o_set_physical_bc3d(s->ph_2, 'w', c, d);
  o_set_physical_bc2d(s->mu_1, 't', c, d);
  o_set_physical_bc2d(s->mu_2, 't', c, d);
  bc_species(s, c, d);
  /* solve_em.F:4734-4750: the surface w is reset from the final u, v (interior only) */
  TILEDV(o_set_w_surface(s->w_2, s, c, d));
  return 0;
}

/* ctypes helpers */
size_t oracle_state_sizeof(void) { return sizeof(ostate); }
int oracle_wreal_bytes(void) { return (int)sizeof(wreal); }
int oracle_solve_em_steps(ostate *s, const cfg_t *c, const dims_t *d, int nsteps) {
  for (int n = 0; n < nsteps; n++) {
    int rc = o_solve_em_step(s, c, d);
    if (rc) return rc;
  }
  return 0;
}

/* Per-routine groups, in the order and with the halo fills of o_solve_em_step, for the
 * phase-by-phase parity tests of the GPU entry points of the same names. */
int oracle_rk_step_prep(ostate *s, const cfg_t *c, const dims_t *d, int rk_step) {
  TILEDV(o_rk_step_prep(s, c, d, rk_step));
  o_rk_phys_bc_dry_1(s, c, d);
  o_set_physical_bc3d(s->al, 'p', c, d);
  o_set_physical_bc3d(s->ph_2, 'w', c, d);
  return 0;
}
int oracle_rk_tendency(ostate *s, const cfg_t *c, const dims_t *d, int rk_step) {
  const int lbc = c->specified || c->nested;
  if (lbc && !s->lbc) return WRFMI_E_BADOPT;
  TILED(o_rk_tendency(s, c, d, rk_step));
  if (lbc && rk_step == 1) o_relax_bdy_dry(s, c, d);
  TILEDV(o_rk_addtend_dry(s, d, rk_step));
  if (lbc) o_spec_bdy_dry(s, c, d);
  return 0;
}
int oracle_small_step_prep(ostate *s, const cfg_t *c, const dims_t *d, int rk_step, double dts) {
  TILEDV(o_small_step_prep(s, d, rk_step));
  TILEDV(o_calc_p_rho(s, c, d, 0));
  if (c->non_hydrostatic) TILEDV(o_calc_coef_w(s, c, d, (wreal)dts));
  o_set_physical_bc3d(s->ru_tend, 'u', c, d);
  o_set_physical_bc3d(s->rv_tend, 'v', c, d);
  o_set_physical_bc3d(s->ph_2, 'w', c, d);
  o_set_physical_bc3d(s->al, 'p', c, d);
  o_set_physical_bc3d(s->p, 'p', c, d);
  o_set_physical_bc3d(s->t_1, 'p', c, d);
  o_set_physical_bc3d(s->t_save, 't', c, d);
  o_set_physical_bc2d(s->mu_1, 't', c, d);
  o_set_physical_bc2d(s->mu_2, 't', c, d);
  o_set_physical_bc2d(s->mudf, 't', c, d);
  return 0;
}
int oracle_advance_uv(ostate *s, const cfg_t *c, const dims_t *d, double dts) {
  TILEDV(o_advance_uv(s, c, d, (wreal)dts));
  if (c->specified || c->nested) spec_uv(s, c, d, (wreal)dts);
  return 0;
}
int oracle_advance_mu_t_w(ostate *s, const cfg_t *c, const dims_t *d, double dts, int iteration, int n_small) {
  const int lbc = c->specified || c->nested;
  TILEDV(o_advance_mu_t(s, c, d, (wreal)dts));
  if (lbc) spec_mu_t(s, c, d, (wreal)dts);
  if (c->non_hydrostatic) TILEDV(o_advance_w(s, c, d, (wreal)dts));
  TILEDV(o_sumflux(s, d, iteration, n_small));
  if (lbc) spec_w_ph(s, c, d, (wreal)dts);
  TILEDV(o_calc_p_rho(s, c, d, iteration));
  o_set_physical_bc3d(s->ph_2, 'w', c, d);
  o_set_physical_bc3d(s->al, 'p', c, d);
  o_set_physical_bc3d(s->p, 'p', c, d);
  o_set_physical_bc2d(s->muts, 't', c, d);
  o_set_physical_bc2d(s->mu_2, 't', c, d);
  o_set_physical_bc2d(s->mudf, 't', c, d);
  return 0;
}
int oracle_small_step_finish(ostate *s, const cfg_t *c, const dims_t *d, int rk_step, int n_small, double dts) {
  TILEDV(o_calc_mu_uv_1(s, c, d));
  TILEDV(o_small_step_finish(s, d, n_small, (wreal)dts, rk_step, 3));
  if (rk_step == 3) {
    o_set_physical_bc3d(s->ru_m, 'u', c, d);
    o_set_physical_bc3d(s->rv_m, 'v', c, d);
    o_set_physical_bc3d(s->ww_m, 'w', c, d);
    o_set_physical_bc2d(s->mut, 't', c, d);
    o_set_physical_bc2d(s->muts, 't', c, d);
  }
  return 0;
}
int oracle_rk_scalar_pd_prep(ostate *s, const cfg_t *c, const dims_t *d, double dt_rk) {
  pd_prep(s, c, d, (wreal)dt_rk);
  return 0;
}
int oracle_rk_scalar_step(ostate *s, const cfg_t *c, const dims_t *d, int ispec, int rk_step, double dt_rk) {
  const int lbc = c->specified || c->nested;
  TILED(o_rk_scalar_tend(s, c, d, ispec, rk_step, (wreal)dt_rk));
  if (lbc && rk_step == 1 && o_species_has_bcs(s, c, ispec)) o_relax_spec_bdy_scalar(s, c, d, ispec);
  TILEDV(o_rk_update_scalar(s, d, ispec, rk_step, (wreal)dt_rk));
  if (c->specified && !o_species_has_bcs(s, c, ispec))
    o_flow_dep_bdy(s->q[ispec], s->ru_m, s->rv_m, c, c->spec_zone, d);
  return 0;
}
int oracle_calc_p_rho_phi(ostate *s, const cfg_t *c, const dims_t *d) {
  TILEDV(o_calc_p_rho_phi(s, c, d));
  return 0;
}
int oracle_set_w_surface(ostate *s, const cfg_t *c, const dims_t *d) {
  TILEDV(o_set_w_surface(s->w_2, s, c, d));
  return 0;
}
int oracle_phys_bc_dry_2(ostate *s, const cfg_t *c, const dims_t *d, int kind) {
  if (kind == 0) {
    o_set_physical_bc3d(s->u_2, 'u', c, d);
    o_set_physical_bc3d(s->v_2, 'v', c, d);
    o_set_physical_bc3d(s->w_2, 'w', c, d);
    o_set_physical_bc3d(s->t_2, 'p', c, d);
    o_set_physical_bc3d(s->ph_2, 'w', c, d);
    o_set_physical_bc2d(s->mu_2, 't', c, d);
  } else {
    o_set_physical_bc3d(s->u_1, 'u', c, d);
    o_set_physical_bc3d(s->u_2, 'u', c, d);
    o_set_physical_bc3d(s->v_1, 'v', c, d);
    o_set_physical_bc3d(s->v_2, 'v', c, d);
    o_set_physical_bc3d(s->w_1, 'w', c, d);
    o_set_physical_bc3d(s->w_2, 'w', c, d);
    o_set_physical_bc3d(s->t_1, 'p', c, d);
    o_set_physical_bc3d(s->t_2, 'p', c, d);
    o_set_physical_bc3d(s->ph_1, 'w', c, d);
    o_set_physical_bc3d(s->ph_2, 'w', c, d);
    o_set_physical_bc2d(s->mu_1, 't', c, d);
    o_set_physical_bc2d(s->mu_2, 't', c, d);
  }
  bc_species(s, c, d);
  return 0;
}
