/*
 * wrfmi_routines.h -- entry points of the MI355X dynamics step that keep the reference's
 * argument lists.
 *
 * Each function replaces one Fortran subroutine of the reference and takes exactly its
 * arguments, in its order, all by reference as a Fortran caller passes them (a bind(C)
 * interface without VALUE attributes; wrf_amd/fortran/wrfmi_bindc.f90 declares them):
 *   - arrays: pointers to the first element of the Fortran array (ims:ime, kms:kme,
 *     jms:jme), device pointers or host arrays registered with wrfmi_register_field;
 *   - REAL scalars: const float *; INTEGER scalars: const int *; LOGICAL scalars:
 *     const int * holding 0 or 1 (a Fortran caller passes merge(1, 0, flag));
 *   - config_flags: the wrfmi_config struct (the grid_config_rec_type fields the path reads);
 *   - the index sextets ids..kde, ims..kme, its..kte as 18 separate INTEGERs; the tile is the
 *     whole patch (its = ids, ite = ide, ...), since the GPU processes a domain in one call.
 * Each returns 0 or a WRFMI_E_* status (the reference calls wrf_error_fatal instead) and
 * enqueues its kernels on the stream wrfmi_set_stream selected for the calling thread (the
 * null stream by default), returning once they are enqueued; wrfmi_sync_to_host waits for
 * them.  Arguments the routine reads only under options outside the implemented set (the
 * c3h/c4h/c3f/c4f hybrid coefficients of hydrostatic mode, spec_zone of specified
 * boundaries, ...) are accepted and not read.  The old-time field of the advection routines
 * must be the advected field itself when a boundary is open, as at every solve_em call site.
 */
#ifndef WRFMI_ROUTINES_H
#define WRFMI_ROUTINES_H

#include "wrfmi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define WRFMI_DIMS_REF                                                                               \
  const int *ids, const int *ide, const int *jds, const int *jde, const int *kds, const int *kde,    \
      const int *ims, const int *ime, const int *jms, const int *jme, const int *kms, const int *kme, \
      const int *its, const int *ite, const int *jts, const int *jte, const int *kts, const int *kte

/* The stream the entry points of this header enqueue on, for the calling thread
 * (a hipStream_t; NULL = the null stream).  No counterpart in the reference. */
int wrfmi_set_stream(void *stream);

/* advect_u, dyn_em/module_advect_em.F:126-136 (called from rk_tendency, module_em.F:505):
 * tendency += the flux divergence of u (5th-order horizontal, 3rd-order vertical). */
int wrfmi_advect_u(const float *u, const float *u_old, float *tendency, const float *ru, const float *rv,
                   const float *rom, const float *c1, const float *c2, const float *mut, const int *time_step,
                   const wrfmi_config *config_flags, const float *msfux, const float *msfuy, const float *msfvx,
                   const float *msfvy, const float *msftx, const float *msfty, const float *fzm, const float *fzp,
                   const float *rdx, const float *rdy, const float *rdzw, WRFMI_DIMS_REF);
/* advect_v, module_advect_em.F:1530-1540 (module_em.F:546). */
int wrfmi_advect_v(const float *v, const float *v_old, float *tendency, const float *ru, const float *rv,
                   const float *rom, const float *c1, const float *c2, const float *mut, const int *time_step,
                   const wrfmi_config *config_flags, const float *msfux, const float *msfuy, const float *msfvx,
                   const float *msfvy, const float *msftx, const float *msfty, const float *fzm, const float *fzp,
                   const float *rdx, const float *rdy, const float *rdzw, WRFMI_DIMS_REF);
/* advect_w, module_advect_em.F:4364-4374 (module_em.F:589). */
int wrfmi_advect_w(const float *w, const float *w_old, float *tendency, const float *ru, const float *rv,
                   const float *rom, const float *c1, const float *c2, const float *mut, const int *time_step,
                   const wrfmi_config *config_flags, const float *msfux, const float *msfuy, const float *msfvx,
                   const float *msfvy, const float *msftx, const float *msfty, const float *fzm, const float *fzp,
                   const float *rdx, const float *rdy, const float *rdzu, WRFMI_DIMS_REF);
/* advect_scalar, module_advect_em.F:3029-3039 (module_em.F:625 for theta, 1331 for the
 * species in rk_scalar_tend). */
int wrfmi_advect_scalar(const float *field, const float *field_old, float *tendency, const float *ru,
                        const float *rv, const float *rom, const float *c1, const float *c2, const float *mut,
                        const int *time_step, const wrfmi_config *config_flags, const float *msfux,
                        const float *msfuy, const float *msfvx, const float *msfvy, const float *msftx,
                        const float *msfty, const float *fzm, const float *fzp, const float *rdx, const float *rdy,
                        const float *rdzw, WRFMI_DIMS_REF);
/* rk_update_scalar, dyn_em/module_em.F:1587-1599 (solve_em.F:2309).  scalar_1, scalar_2 and
 * sc_tend hold species scs..sce (4-D arrays (ims:ime, kms:kme, jms:jme, scs:sce)).  The
 * optional advh_t/advz_t may be NULL; tenddec's split diagnostics are refused. */
int wrfmi_rk_update_scalar(const int *scs, const int *sce, float *scalar_1, float *scalar_2, const float *sc_tend,
                           float *advh_t, float *advz_t, const float *advect_tend, const float *h_tendency,
                           const float *z_tendency, const float *msftx, const float *msfty, const float *c1,
                           const float *c2, const float *mu_old, const float *mu_new, const float *mu_base,
                           const int *rk_step, const float *dt, const int *spec_zone,
                           const wrfmi_config *config_flags, const int *tenddec, WRFMI_DIMS_REF);
/* advect_scalar_pd, dyn_em/module_advect_em.F:6069-6083 (module_em.F:1266, rk_scalar_tend on
 * the last RK substep for moist_adv_opt / scalar_adv_opt 1): tendency += the positive-definite
 * flux divergence (5th-order horizontal, 3rd-order vertical, Smolarkiewicz upwind fluxes from
 * field_old and the limiter with mub + mu_old).  tenddec's split diagnostics are refused.
 * The limiter's scales use a work array of the domain known by `tendency`. */
int wrfmi_advect_scalar_pd(const float *field, const float *field_old, float *tendency, float *h_tendency,
                           float *z_tendency, const float *ru, const float *rv, const float *rom, const float *c1,
                           const float *c2, const float *mut, const float *mub, const float *mu_old,
                           const int *time_step, const wrfmi_config *config_flags, const int *tenddec,
                           const float *msfux, const float *msfuy, const float *msfvx, const float *msfvy,
                           const float *msftx, const float *msfty, const float *fzm, const float *fzp,
                           const float *rdx, const float *rdy, const float *rdzw, const float *dt, WRFMI_DIMS_REF);
/* rk_update_scalar_pd, dyn_em/module_em.F:1803-1811 (solve_em.F:1858, 1921 with mu_1 as both
 * mu_old and mu_new): scalar = ((c1 (mu_old + mu_base) + c2) scalar + dt sc_tend) /
 * (c1 (mu_new + mu_base) + c2) over the mass points, then sc_tend = 0; species scs..sce. */
int wrfmi_rk_update_scalar_pd(const int *scs, const int *sce, float *scalar, float *sc_tend, const float *c1,
                              const float *c2, const float *mu_old, const float *mu_new, const float *mu_base,
                              const int *rk_step, const float *dt, const int *spec_zone,
                              const wrfmi_config *config_flags, WRFMI_DIMS_REF);
/* advance_uv, dyn_em/module_small_step_em.F:654-672 (solve_em.F:1279). */
int wrfmi_advance_uv(float *u, const float *ru_tend, float *v, const float *rv_tend, const float *p, const float *pb,
                     const float *ph, const float *php, const float *alt, const float *al, const float *mu,
                     const float *muu, const float *cqu, const float *muv, const float *cqv, const float *mudf,
                     const float *c1h, const float *c2h, const float *c1f, const float *c2f, const float *c3h,
                     const float *c4h, const float *c3f, const float *c4f, const float *msfux, const float *msfuy,
                     const float *msfvx, const float *msfvx_inv, const float *msfvy, const float *rdx,
                     const float *rdy, const float *dts, const float *cf1, const float *cf2, const float *cf3,
                     const float *fnm, const float *fnp, const float *emdiv, const float *rdnw,
                     const wrfmi_config *config_flags, const int *spec_zone, const int *non_hydrostatic,
                     const int *top_lid, WRFMI_DIMS_REF);
/* advance_mu_t, module_small_step_em.F:969-982 (solve_em.F:1392).  uam, vam and wwam are
 * not read (as in the reference). */
int wrfmi_advance_mu_t(float *ww, const float *ww_1, const float *u, const float *u_1, const float *v, const float *v_1,
                       float *mu, const float *mut, float *muave, float *muts, const float *muu, const float *muv,
                       float *mudf, const float *c1h, const float *c2h, const float *c1f, const float *c2f,
                       const float *c3h, const float *c4h, const float *c3f, const float *c4f, const float *uam,
                       const float *vam, const float *wwam, float *t, const float *t_1, float *t_ave, const float *ft,
                       const float *mu_tend, const float *rdx, const float *rdy, const float *dts, const float *epssm,
                       const float *dnw, const float *fnm, const float *fnp, const float *rdnw, const float *msfux,
                       const float *msfuy, const float *msfvx, const float *msfvx_inv, const float *msfvy,
                       const float *msftx, const float *msfty, const int *step, const wrfmi_config *config_flags,
                       WRFMI_DIMS_REF);
/* advance_w, module_small_step_em.F:1178-1192 (solve_em.F:1499).  a, alpha and gamma are
 * REAL arrays as in the reference (precision_mode 0 only); a and alpha are formed in the
 * column solve and not read, gamma is written. */
int wrfmi_advance_w(float *w, const float *rw_tend, const float *ww, const float *w_save, const float *u,
                    const float *v, const float *mu1, const float *mut, const float *muave, const float *muts,
                    const float *c1h, const float *c2h, const float *c1f, const float *c2f, const float *c3h,
                    const float *c4h, const float *c3f, const float *c4f, float *t_2ave, const float *t_2,
                    const float *t_1, float *ph, const float *ph_1, const float *phb, const float *ph_tend,
                    const float *ht, const float *c2a, const float *cqw, const float *alt, const float *alb, void *a,
                    void *alpha, void *gamma, const float *rdx, const float *rdy, const float *dts, const float *t0,
                    const float *epssm, const float *dnw, const float *fnm, const float *fnp, const float *rdnw,
                    const float *rdn, const float *cf1, const float *cf2, const float *cf3, const float *msftx,
                    const float *msfty, const wrfmi_config *config_flags, const int *top_lid, WRFMI_DIMS_REF);
/* sumflux, module_small_step_em.F:1473-1483 (solve_em.F:1566). */
int wrfmi_sumflux(const float *ru, const float *rv, const float *ww, const float *u_lin, const float *v_lin,
                  const float *ww_lin, const float *muu, const float *muv, const float *c1h, const float *c2h,
                  const float *c1f, const float *c2f, const float *c3h, const float *c4h, const float *c3f,
                  const float *c4f, float *ru_m, float *rv_m, float *ww_m, const float *epssm, const float *msfux,
                  const float *msfuy, const float *msfvx, const float *msfvx_inv, const float *msfvy,
                  const int *iteration, const int *number_of_small_timesteps, WRFMI_DIMS_REF);
/* calc_p_rho, module_small_step_em.F:438-447 (solve_em.F:1109 with step 0, 1623). */
int wrfmi_calc_p_rho(float *al, float *p, float *ph, const float *alt, const float *t_2, const float *t_1,
                     const float *c2a, float *pm1, const float *mu, const float *mut, const float *c1h,
                     const float *c2h, const float *c1f, const float *c2f, const float *c3h, const float *c4h,
                     const float *c3f, const float *c4f, const float *znu, const float *t0, const float *rdnw,
                     const float *dnw, const float *smdiv, const int *non_hydrostatic, const int *step,
                     WRFMI_DIMS_REF);
/* calc_coef_w, module_small_step_em.F:570-580 (solve_em.F:1124): REAL a, alpha, gamma. */
int wrfmi_calc_coef_w(void *a, void *alpha, void *gamma, const float *mut, const float *c1h, const float *c2h,
                      const float *c1f, const float *c2f, const float *c3h, const float *c4h, const float *c3f,
                      const float *c4f, const float *cqw, const float *rdn, const float *rdnw, const float *c2a,
                      const float *dts, const float *g, const float *epssm, const int *top_lid, WRFMI_DIMS_REF);

#ifdef __cplusplus
}
#endif
#endif
