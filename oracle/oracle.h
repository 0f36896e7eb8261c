/*
 * CPU oracle for the WRF-ARW dynamics step -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference routines on the path
 * (dyn_em/module_advect_em.F, module_small_step_em.F, module_em.F,
 *  module_big_step_utilities_em.F, share/module_bc.F, dyn_em/solve_em.F),
 * loop by loop and in the reference's floating-point operation order, so
 * that an IEEE FP32 build reproduces the reference's default-REAL arithmetic
 * and a FP64 build (-DWREAL=double) plays the role of the reference's -r8
 * build.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only as the checker; the product path never calls it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fno-fast-math).
 * Every function cites the reference file:line it follows.
 *
 * How each part is pinned (tests/test_oracle.py, tests/oracle_pins.py; the mutation test
 * tests/test_oracle_mutations.py builds this source with 41 plausible mistakes planted and
 * shows that every one fails at least one pin):
 *   o_advect_scalar  bit for bit against the reference's own advect_scalar, compiled in
 *                    place (oracle/_ref) and through tests/golden/advect_scalar_ref.npz:
 *                    periodic and open, both signs of time_step, field_old = field and not.
 *   o_advect_scalar_pd  bit for bit against the reference's own advect_scalar_pd the same way
 *                    (tests/golden/advect_scalar_pd_ref.npz, tendency and its h/z split);
 *                    unlimited it is the pinned o_advect_scalar; the limiter keeps the
 *                    updated mass non-negative and the periodic total conserved.
 *   o_rk_update_scalar_pd + the last-substep flow  a physics tendency is applied once over the
 *                    step; an empty-celled tracer stays non-negative through whole steps.
 *   o_advect_u/v/w   bit for bit equal to o_advect_scalar in uniform flow (same fluxes,
 *                    shifted stencil); a constant field in a non-uniform mass flux on a
 *                    stretched grid gives -C div(mass flux) on every level, advect_w's
 *                    lowest level and lid half-cell included; a field linear in eta is
 *                    carried exactly through advect_u/v's second-order boundary faces.
 *   o_calc_coef_w    eigenvalues of the implicit operator against the continuum vertical
 *                    acoustic modes of an isothermal column; row sums; self-adjointness.
 *   o_advance_w      the off-centred trapezoidal scheme of the w-phi pair with
 *                    calc_coef_w's operator; a hydrostatically balanced perturbation
 *                    exerts no force.
 *   o_calc_p_rho     the first-order change of the full equation of state.
 *   o_advance_uv, horizontal_pressure_gradient  the four-term eta-coordinate pressure
 *                    gradient in closed form (terrain-slope term included), plus the
 *                    external-mode filter.
 *   o_advance_mu_t   mass budget, omega'' from continuity, uniform theta preserved.
 *   o_small_step_prep/finish  the coupling round trip returns time level n.
 *   rk_update_scalar, sumflux  a uniform mixing ratio stays uniform through whole steps.
 *   o_sixth_order_diffusion  the checkerboard damping rate, flux form (periodic sum zero),
 *                    the monotone limiter on a spike, the slope tapering factor.
 *   o_rk_rayleigh_damp  the relaxation rate sin^2 profile toward linear base profiles.
 *   advance_w damp_opt 3  the damped w over 1 + dts dampcoef sin^2(...) in the layer.
 *   o_horizontal_diffusion(_3dmp)  the discrete Laplacian's eigenvalue on cosines (v's
 *                    y part without the column mass, as the reference has it), flux form,
 *                    the base subtracted.
 *   o_vertical_diffusion*  2 a mu K for a z^2 profile on interior levels, zero gradient
 *                    below the lowest level, no flux through the top, bases subtracted.
 *   o_relax_bdytend  boundary arrays equal to the field give no tendency; a deviation at
 *                    one relaxation point relaxes at fcx + 4 gcx and feeds gcx to its four
 *                    neighbours (the 5-point smoother), on every side.
 *   o_spec_bdytend/_bdyupdate, o_zero_grad_bdy, o_flow_dep_bdy  exactly the frame of width
 *                    spec_zone changes, by dt times its side's tendency (corners split along
 *                    the diagonal); zero gradient from the first inner row; outflow kept, inflow
 *                    zeroed.
 *   o_spec_bdyupdate_ph  the coupled tendency decoupled by (c1 mu + c2); the coupled value
 *                    kept under a mass change.
 *   specified whole step  zero boundary tendencies hold the spec zone at time level n.
 *   whole step       resting atmosphere stays at rest (flat and over terrain); uniform
 *                    state stays uniform; mirror and x/y-transpose symmetry of a centred
 *                    bubble (periodic and open); periodic translation invariance (bitwise);
 *                    dry-air and species mass conservation; the EOS reproduces the
 *                    independently built hydrostatic state; FP32 and FP64 builds agree.
 * Parity unpinned (no reference build, fixture or closed form covers them, and the
 * benchmark configurations do not exercise them): the w_damping = 1 branch of w_damp,
 * the open- and specified-boundary loop bounds of the diffusion routines (their closed
 * forms run on periodic domains), the relaxation and spec zones of nested boundaries
 * (refused by the GPU library), top_lid = 1, phi_adv_z = 2, hypsometric_opt = 2's
 * log-pressure form of calc_p_rho_phi beyond the initial state it reproduces,
 * use_theta_m = 0, Coriolis with f or e non-zero, curvature with map factors other than
 * 1.  The terrain terms (the surface-w term of advance_w and set_w_surface, with their
 * cf1..cf3 extrapolations) are pinned by their closed form in uniform flow,
 * U * (ht(i+1) - ht(i-1)) / (2 dx) (tests/test_oracle.py).
 */
#ifndef WRF_ORACLE_H
#define WRF_ORACLE_H

#include <stddef.h>
#include <math.h>
#include "../include/wrfmi.h"

#ifndef WREAL
#define WREAL float
#endif
typedef WREAL wreal;
#define R(x) ((wreal)(x))

typedef wrfmi_dims dims_t;
typedef wrfmi_config cfg_t;

/* wrfmi_lbc with WREAL arrays and weights */
typedef struct olbc {
  wreal dtbc;
  wreal fcx[WRFMI_MAX_BDY_WIDTH], gcx[WRFMI_MAX_BDY_WIDTH];
#define O_LBC(n) wreal *n##_b[4]; wreal *n##_bt[4];
  WRFMI_LBC_FIELDS(O_LBC)
#undef O_LBC
  wreal *q_b[WRFMI_MAX_SPECIES][4];
  wreal *q_bt[WRFMI_MAX_SPECIES][4];
} olbc;

/* state with WREAL arrays; member order identical to wrfmi_state */
typedef struct ostate {
#define O_PTR(n) wreal *n;
  WRFMI_FIELDS_3D(O_PTR)
  WRFMI_FIELDS_2D(O_PTR)
  WRFMI_FIELDS_1D(O_PTR)
#undef O_PTR
#define O_R(n) wreal n;
  WRFMI_SCALARS_R(O_R)
#undef O_R
  void *a, *alpha, *gamma;
  int n_moist, n_species;
  wreal *q[WRFMI_MAX_SPECIES];
  wreal *q_old[WRFMI_MAX_SPECIES];
  wreal *q_tend[WRFMI_MAX_SPECIES];
  olbc *lbc;
} ostate;

#define NI_ (d->ime - d->ims + 1)
#define NK_ (d->kme - d->kms + 1)
#define A3(a, i, k, j) ((a)[(((size_t)((j) - d->jms)) * NK_ + (size_t)((k) - d->kms)) * NI_ + (size_t)((i) - d->ims)])
#define A2(a, i, j) ((a)[((size_t)((j) - d->jms)) * NI_ + (size_t)((i) - d->ims)])
#define A1(a, k) ((a)[(k) - d->kms])

#define IMIN(a, b) ((a) < (b) ? (a) : (b))
#define IMAX(a, b) ((a) > (b) ? (a) : (b))

/* share/module_model_constants.F:17-47 (REAL parameters, evaluated in WREAL) */
#define K_G R(9.81)
#define K_RD R(287.)
#define K_CP (R(7.) * K_RD / R(2.))
#define K_RV R(461.6)
#define K_CV (K_CP - K_RD)
#define K_CVPM (-K_CV / K_CP)
#define K_P1000 R(100000.)
#define K_T0 R(300.)
#define K_P0 K_P1000
#define K_CPOVCV (K_CP / (K_CP - K_RD))
#define K_RVOVRD (K_RV / K_RD)
#define K_RERADIUS (R(1.) / R(6370.0e03))
#define K_CB R(25.)
#define K_W_ALPHA R(0.3)
#define K_W_BETA R(1.0)

/* Fortran SIGN(1.,x): magnitude 1 with the sign bit of x */
static inline wreal sgn1(wreal x) { return (wreal)copysign(1.0, (double)x); }

/* ---- share/module_bc.F ---- */
void o_set_physical_bc3d(wreal *dat, char var, const cfg_t *c, const dims_t *d);
void o_set_physical_bc2d(wreal *dat, char var, const cfg_t *c, const dims_t *d);

/* ---- rk_step_prep and callees ---- */
void o_rk_step_prep(ostate *s, const cfg_t *c, const dims_t *d, int rk_step);
void o_rk_phys_bc_dry_1(ostate *s, const cfg_t *c, const dims_t *d);
void o_set_w_surface(wreal *w, const ostate *s, const cfg_t *c, const dims_t *d);

/* ---- advection (order 5 horizontal / 3 vertical, plus 2..6 rejected) ---- */
int o_advect_u(const wreal *u, const wreal *u_old, wreal *tend, const wreal *ru, const wreal *rv,
               const wreal *rom, const wreal *c1, const wreal *c2, const wreal *mut, int time_step,
               const cfg_t *c, const wreal *msfux, const wreal *msfuy, const wreal *msfvx,
               const wreal *msfvy, const wreal *msftx, const wreal *msfty, const wreal *fzm,
               const wreal *fzp, wreal rdx, wreal rdy, const wreal *rdzw, const dims_t *d);
int o_advect_v(const wreal *v, const wreal *v_old, wreal *tend, const wreal *ru, const wreal *rv,
               const wreal *rom, const wreal *c1, const wreal *c2, const wreal *mut, int time_step,
               const cfg_t *c, const wreal *msfux, const wreal *msfuy, const wreal *msfvx,
               const wreal *msfvy, const wreal *msftx, const wreal *msfty, const wreal *fzm,
               const wreal *fzp, wreal rdx, wreal rdy, const wreal *rdzw, const dims_t *d);
int o_advect_w(const wreal *w, const wreal *w_old, wreal *tend, const wreal *ru, const wreal *rv,
               const wreal *rom, const wreal *c1, const wreal *c2, const wreal *mut, int time_step,
               const cfg_t *c, const wreal *msfux, const wreal *msfuy, const wreal *msfvx,
               const wreal *msfvy, const wreal *msftx, const wreal *msfty, const wreal *fzm,
               const wreal *fzp, wreal rdx, wreal rdy, const wreal *rdzu, const dims_t *d);
int o_advect_scalar(const wreal *field, const wreal *field_old, wreal *tend, const wreal *ru,
                    const wreal *rv, const wreal *rom, const wreal *c1, const wreal *c2,
                    const wreal *mut, int time_step, const cfg_t *c, const wreal *msfux,
                    const wreal *msfuy, const wreal *msfvx, const wreal *msfvy, const wreal *msftx,
                    const wreal *msfty, const wreal *fzm, const wreal *fzp, wreal rdx, wreal rdy,
                    const wreal *rdzw, const dims_t *d);

/* positive-definite form (order 5 horizontal / 3 vertical); h_tend and z_tend are written
 * when tenddec is non-zero */
int o_advect_scalar_pd(const wreal *field, const wreal *field_old, wreal *tend, wreal *h_tend, wreal *z_tend,
                       const wreal *ru, const wreal *rv, const wreal *rom, const wreal *c1, const wreal *c2,
                       const wreal *mut, const wreal *mub, const wreal *mu_old, int time_step, const cfg_t *c,
                       int tenddec, const wreal *msfux, const wreal *msfuy, const wreal *msfvx, const wreal *msfvy,
                       const wreal *msftx, const wreal *msfty, const wreal *fzm, const wreal *fzp, wreal rdx, wreal rdy,
                       const wreal *rdzw, wreal dt, const dims_t *d);
void o_rk_update_scalar_pd(wreal *scalar, wreal *sc_tend, const wreal *c1, const wreal *c2, const wreal *mu_old,
                           const wreal *mu_new, const wreal *mu_base, wreal dt, const dims_t *d);

/* ---- forward-step mixing and damping of the first RK substep (mixing.c) ---- */
#define WRFMI_ORACLE_MAXK 1024
void o_sixth_order_diffusion(char name, const wreal *field, wreal *tend, const wreal *mut, wreal dt, const cfg_t *c,
                             const wreal *c1, const wreal *c2, int diff_6th_opt, wreal diff_6th_factor,
                             const wreal *phb, const wreal *ph, wreal rdx, wreal rdy, const wreal *msftx,
                             const wreal *msfty, const wreal *msfux, const wreal *msfuy, const wreal *msfvx,
                             const wreal *msfvy, const dims_t *d);
void o_rk_rayleigh_damp(ostate *s, const cfg_t *c, const dims_t *d);
void o_horizontal_diffusion(char name, const wreal *field, wreal *tend, const wreal *mut, const wreal *c1,
                            const wreal *c2, const cfg_t *c, const wreal *msfux, const wreal *msfuy,
                            const wreal *msfvx, const wreal *msfvx_inv, const wreal *msfvy, const wreal *msftx,
                            const wreal *msfty, wreal khdif, const wreal *xkmhd, wreal rdx, wreal rdy,
                            const dims_t *d);
void o_horizontal_diffusion_3dmp(char name, const wreal *field, wreal *tend, const wreal *mut, const wreal *c1,
                                 const wreal *c2, const cfg_t *c, const wreal *base_3d, const wreal *msfux,
                                 const wreal *msfuy, const wreal *msfvx, const wreal *msfvx_inv, const wreal *msfvy,
                                 const wreal *msftx, const wreal *msfty, wreal khdif, const wreal *xkmhd, wreal rdx,
                                 wreal rdy, const dims_t *d);
void o_vertical_diffusion(char name, const wreal *field, wreal *tend, const cfg_t *c, const wreal *c1,
                          const wreal *c2, const wreal *alt, const wreal *mut, const wreal *rdn, const wreal *rdnw,
                          wreal kvdif, const dims_t *d);
void o_vertical_diffusion_mp(const wreal *field, wreal *tend, const cfg_t *c, const wreal *base, const wreal *c1,
                             const wreal *c2, const wreal *alt, const wreal *mut, const wreal *rdn, const wreal *rdnw,
                             wreal kvdif, const dims_t *d);
void o_vertical_diffusion_3dmp(const wreal *field, wreal *tend, const cfg_t *c, const wreal *base_3d,
                               const wreal *c1, const wreal *c2, const wreal *alt, const wreal *mut, const wreal *rdn,
                               const wreal *rdnw, wreal kvdif, const dims_t *d);
void o_vertical_diffusion_u(const wreal *field, wreal *tend, const cfg_t *c, const wreal *u_base, const wreal *c1h,
                            const wreal *c2h, const wreal *alt, const wreal *muu, const wreal *rdn, const wreal *rdnw,
                            wreal kvdif, const dims_t *d);
void o_vertical_diffusion_v(const wreal *field, wreal *tend, const cfg_t *c, const wreal *v_base, const wreal *c1h,
                            const wreal *c2h, const wreal *alt, const wreal *muv, const wreal *rdn, const wreal *rdnw,
                            wreal kvdif, const dims_t *d);

/* ---- specified / nested lateral boundaries (lbc.c) ---- */
void o_relax_bdytend(const wreal *field, wreal *tend, wreal *const *b, wreal *const *bt, char var, const cfg_t *c,
                     int spec_zone, int relax_zone, wreal dtbc, const wreal *fcx, const wreal *gcx, const dims_t *d);
void o_spec_bdytend(wreal *tend, wreal *const *bt, char var, const cfg_t *c, int spec_zone, const dims_t *d);
void o_spec_bdyupdate(wreal *field, const wreal *tend, wreal dt, char var, const cfg_t *c, int spec_zone,
                      const dims_t *d);
void o_spec_bdyupdate_ph(const wreal *ph_save, wreal *field, const wreal *tend, const wreal *mu_tend,
                         const wreal *muts, const wreal *c1, const wreal *c2, wreal dt, const cfg_t *c, int spec_zone,
                         const dims_t *d);
void o_zero_grad_bdy(wreal *field, char var, const cfg_t *c, int spec_zone, const dims_t *d);
void o_flow_dep_bdy(wreal *field, const wreal *u, const wreal *v, const cfg_t *c, int spec_zone, const dims_t *d);
void o_relax_bdy_dry(ostate *s, const cfg_t *c, const dims_t *d);
void o_spec_bdy_dry(ostate *s, const cfg_t *c, const dims_t *d);
void o_relax_spec_bdy_scalar(ostate *s, const cfg_t *c, const dims_t *d, int m);
int o_species_has_bcs(const ostate *s, const cfg_t *c, int m);

/* ---- rk_tendency and rk_addtend_dry ---- */
int o_rk_tendency(ostate *s, const cfg_t *c, const dims_t *d, int rk_step);
void o_rk_addtend_dry(ostate *s, const dims_t *d, int rk_step);

/* ---- small step ---- */
void o_small_step_prep(ostate *s, const dims_t *d, int rk_step);
void o_calc_p_rho(ostate *s, const cfg_t *c, const dims_t *d, int step);
void o_calc_coef_w(ostate *s, const cfg_t *c, const dims_t *d, wreal dts);
void o_advance_uv(ostate *s, const cfg_t *c, const dims_t *d, wreal dts);
void o_advance_mu_t(ostate *s, const cfg_t *c, const dims_t *d, wreal dts);
void o_advance_w(ostate *s, const cfg_t *c, const dims_t *d, wreal dts);
void o_sumflux(ostate *s, const dims_t *d, int iteration, int number_of_small_timesteps);
void o_calc_mu_uv_1(ostate *s, const cfg_t *c, const dims_t *d);
void o_small_step_finish(ostate *s, const dims_t *d, int nsmall, wreal dts, int rk_step, int rk_order);

/* ---- scalars and EOS ---- */
int o_rk_scalar_tend(ostate *s, const cfg_t *c, const dims_t *d, int ispec, int rk_step, wreal dt_rk);
void o_rk_update_scalar(ostate *s, const dims_t *d, int ispec, int rk_step, wreal dt_rk);
void o_calc_p_rho_phi(ostate *s, const cfg_t *c, const dims_t *d);

/* ---- one RK3 step (dyn_em/solve_em.F) ---- */
int o_solve_em_step(ostate *s, const cfg_t *c, const dims_t *d);

#endif
