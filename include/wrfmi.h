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
  WRFMI_E_NONFINITE = 5, /* non-finite value found by wrfmi_check_finite */
  WRFMI_E_IO = 6         /* a history file could not be created or written */
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
  int diff_6th_slopeopt; /* 6th-order diffusion tapered on steep terrain (1) or not (0) */
  int km_opt;            /* eddy coefficients of diff_opt 1: 1 = constant khdif, kvdif */
  int bl_pbl_physics;    /* 0: diff_opt 1 also diffuses vertically (no boundary-layer scheme) */
  float epssm, smdiv, emdiv, dampcoef, zdamp, w_crit_cfl;
  float diff_6th_factor, diff_6th_thresh;
  float khdif, kvdif;    /* horizontal and vertical eddy viscosity of km_opt 1 (m2/s) */
  /* specified / nested lateral boundaries: the boundary zone's width and its split into
   * the specified rows and the relaxation rows (namelist spec_bdy_width, spec_zone,
   * relax_zone); have_bcs_moist / have_bcs_scalar: the moist species other than QV / the
   * passive scalars take boundary arrays too (otherwise flow_dep_bdy) */
  int spec_bdy_width, spec_zone, relax_zone, have_bcs_moist, have_bcs_scalar;
  /* 0: FP32 everywhere, reference operation order.
   * 1: FP32 bulk, FP64 calc_coef_w + advance_w tridiagonal solve,
   *    compensated (two-sum) column mass-divergence integrals. */
  int precision_mode;
} wrfmi_config;

/* Lateral boundary arrays of a domain with specified (or nested) boundaries: the
 * Registry's <field>_b[xs|xe|ys|ye] values and <field>_bt[..] tendencies
 * (Registry/Registry.EM_COMMON), side index 0..3 = xs, xe, ys, ye.  x sides are Fortran
 * arrays (jms:jme, kds:kde, spec_bdy_width), y sides (ims:ime, kds:kde, spec_bdy_width); mu
 * has one level.  The values are the coupled ones solve_em relaxes toward (u, v: the
 * mu-coupled ru, rv; ph, t: times c1 mu + c2; mu itself).  dtbc is grid%dtbc, the time since
 * the boundary values' valid time: the whole step adds dt to it first, as solve_em.F:377
 * does; fcx, gcx are lbc_fcx_gcx's relaxation weights (module_bc_em.F:1297-1341).
 * Implemented: specified boundaries with spec_zone 1 and at most 256 levels; nested
 * boundaries, and specified ones under graph capture or in ensembles, return
 * WRFMI_E_BADOPT. */
#define WRFMI_MAX_BDY_WIDTH 16
#define WRFMI_LBC_FIELDS(X) X(u) X(v) X(ph) X(t) X(w) X(mu)
typedef struct wrfmi_lbc {
  float dtbc;
  float fcx[WRFMI_MAX_BDY_WIDTH], gcx[WRFMI_MAX_BDY_WIDTH];
#define WRFMI_LBC_DECL(n) float *n##_b[4]; float *n##_bt[4];
  WRFMI_LBC_FIELDS(WRFMI_LBC_DECL)
#undef WRFMI_LBC_DECL
  float *q_b[WRFMI_MAX_SPECIES][4];
  float *q_bt[WRFMI_MAX_SPECIES][4];
} wrfmi_lbc;

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
  /* lateral boundary arrays (host memory struct; its arrays device or registered host
   * memory), read when cfg->specified or cfg->nested; NULL otherwise */
  wrfmi_lbc *lbc;
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
 * and at I/O steps, and passes its own host pointers to every entry point.
 * Registering an address again refreshes its mirror from the host (or replaces
 * it when the size changed); a range overlapping another registered array is
 * refused.  Unregister an array before freeing it. */
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
 * small_step_finish, rk_scalar_tend/rk_update_scalar for every species (with
 * rk_update_scalar_pd and advect_scalar_pd on the last substep for the species whose
 * moist_adv_opt / scalar_adv_opt is 1),
 * calc_p_rho_phi, the set_physical_bc3d/2d halo fills and set_w_surface
 * (solve_em.F:4740).  The physics tendencies *_tendf, mu_tendf and h_diabatic are
 * read as given (the reference zeroes them at rk_step 1 when physics is off:
 * module_first_rk_step_part1.F:182).  calc_coef_w's a and alpha are formed inside
 * the column solve and not stored by this entry point (the per-routine
 * wrfmi_small_step_prep stores them); gamma is stored as the reference does.
 * `stream` is a hipStream_t (NULL = the default stream); the call returns
 * after enqueueing unless `sync` is non-zero. */
int wrfmi_solve_em_step(wrfmi_state *st, const wrfmi_config *cfg,
                        const wrfmi_dims *d, void *stream, int sync);

/* Same step, replayed from a hipGraph captured on the first call for this
 * (state, config, dims) triple, compared through the device pointers the state
 * translates to (a registered array whose mirror was replaced gets a new graph). */
int wrfmi_solve_em_step_graph(wrfmi_state *st, const wrfmi_config *cfg,
                              const wrfmi_dims *d, void *stream, int sync);

/* Same step (bit for bit), for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Its work arrays live in a static device array of the
 * library, 48 MiB on every GPU it runs on, held until the process ends; the call makes no
 * hipMalloc, hipFree or other allocating call.  Limit: memory dims of at most 4,194,304
 * (2^22) elements per field (WRFMI_E_DIMS beyond).  The species are advected one by one.
 * Kernel timing mode is refused.  All calls on one GPU share the array: calls that
 * overlap in time on one GPU (different streams or threads) race. */
int wrfmi_solve_em_step_noalloc(wrfmi_state *st, const wrfmi_config *cfg,
                                const wrfmi_dims *d, void *stream, int sync);

/* Work arrays.  The library keeps a few grid-sized work arrays per domain (a domain
 * is known by the device address of its u_2 array and its field size), allocated on
 * the domain's first call: steps of different domains may run concurrently on
 * different streams.  wrfmi_domain_release frees a domain's work arrays and cached
 * graphs (waiting for the device first); a later call allocates them again.  No
 * counterpart in the reference, whose work arrays are solve_em's stack and i1 locals
 * (Registry/Registry.EM_COMMON:186-348). */
int wrfmi_domain_release(const wrfmi_state *st);

/* ---- batched ensembles ---------------------------------------------------
 * No counterpart in the reference, which advances one domain per solve_em call
 * (ensembles run as separate executions).  An ensemble is n_members independent
 * domains, of any sizes, each advanced by exactly the step of
 * wrfmi_solve_em_step (bit for bit), with one launch per kernel of the step
 * covering the tiles of every member.  states/cfgs/dims are arrays of
 * n_members structs, copied at creation: later steps use those pointers.
 * Members must share the boundary type, precision_mode, n_species and n_moist with
 * their transport options (moist_adv_opt, scalar_adv_opt), and the number of acoustic
 * steps (WRFMI_E_BADOPT otherwise).  Each member gets its
 * own work arrays.  _step enqueues the step eagerly; _step_graph replays a
 * hipGraph captured on its first call. */
typedef struct wrfmi_ensemble wrfmi_ensemble;
int wrfmi_ensemble_create(int n_members, const wrfmi_state *states,
                          const wrfmi_config *cfgs, const wrfmi_dims *dims,
                          wrfmi_ensemble **out);
int wrfmi_ensemble_step(wrfmi_ensemble *e, void *stream, int sync);
int wrfmi_ensemble_step_graph(wrfmi_ensemble *e, void *stream, int sync);
int wrfmi_ensemble_destroy(wrfmi_ensemble *e);

/* ---- per-routine entry points -------------------------------------------
 * The calls solve_em makes inside one RK substep, each with the halo fills the
 * reference makes right after it.  Each enqueues on `stream` and returns; the
 * whole step above is exactly this sequence.  rk_step is 1..3; dts is the
 * acoustic step of the substep (dt/3 at rk_step 1, dt/time_step_sound after). */

/* rk_step_prep (dyn_em/module_em.F:37-186, called at solve_em.F:657),
 * rk_phys_bc_dry_1 (module_bc_em.F:960-1041, called at solve_em.F:726) and
 * the al, ph_2 fills of solve_em.F:760-782. */
int wrfmi_rk_step_prep(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d,
                       int rk_step, void *stream);
/* rk_tendency (module_em.F:190-955; advect_u/v/w, advect_scalar(theta), rhs_ph,
 * horizontal_pressure_gradient, pg_buoy_w, w_damp, coriolis, curvature) and
 * rk_addtend_dry (module_em.F:959-1092), called at solve_em.F:881 and 967. */
int wrfmi_rk_tendency(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d,
                      int rk_step, void *stream);
/* small_step_prep (module_small_step_em.F:16-290), calc_p_rho step 0
 * (438-568), calc_coef_w (570-652), called at solve_em.F:1089-1124, and the
 * fills of solve_em.F:1181-1264. */
int wrfmi_small_step_prep(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d,
                          int rk_step, float dts, void *stream);
/* advance_uv (module_small_step_em.F:654-967), called at solve_em.F:1279; the entry point
 * with advance_uv's own argument list is wrfmi_advance_uv (wrfmi_routines.h). */
int wrfmi_advance_uv_state(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d,
                           float dts, void *stream);
/* One column pass: advance_mu_t (969-1175), advance_w (1178-1469), sumflux
 * (1473-1633), calc_p_rho (438-568) for acoustic substep `iteration` of
 * `n_small` (called at solve_em.F:1392, 1499, 1566, 1623), and the fills
 * that follow them (solve_em.F:1640-1719). */
int wrfmi_advance_mu_t_w(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d,
                         float dts, int iteration, int n_small, void *stream);
/* calc_mu_uv_1 (module_big_step_utilities_em.F:184-321), small_step_finish
 * (module_small_step_em.F:295-434), and at rk_step 3 the ru_m, rv_m, ww_m,
 * mut, muts fills (called at solve_em.F:1732-1800). */
int wrfmi_small_step_finish(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d,
                            int rk_step, int n_small, float dts, void *stream);
/* rk_scalar_tend (module_em.F:1096-1442, advect_scalar with ru_m, rv_m, ww_m, or at
 * rk_step 3 advect_scalar_pd as for wrfmi_rk_scalar_tend below) followed by
 * rk_update_scalar (1587-1799) for species `ispec` (called at solve_em.F:2215-2309 and
 * 2781-2856). */
int wrfmi_rk_scalar_step(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d,
                         int ispec, int rk_step, float dt_rk, void *stream);
/* rk_scalar_tend alone (module_em.F:1096-1442: advect_tend = advect_scalar of species
 * `ispec` with ru_m, rv_m, ww_m, from zero), called at solve_em.F:2215 and 2781; a host
 * then calls rk_update_scalar with its own argument list (wrfmi_rk_update_scalar).  At
 * rk_step 3 a species whose option (moist_adv_opt for the n_moist moist species,
 * scalar_adv_opt for the others) is 1 takes advect_scalar_pd (module_advect_em.F:
 * 6069-7885) with muts, mub, mu_1 and dt_rk instead (module_em.F:1264-1280). */
int wrfmi_rk_scalar_tend(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d,
                         int ispec, int rk_step, float dt_rk, void *stream);
/* Before the last substep's species transport (solve_em.F:1850-1975): for every species in
 * positive-definite form, rk_update_scalar_pd (module_em.F:1803-1916) takes q_tend into
 * q_old with mu_1 + mub as both masses and zeroes q_tend, then q_old gets its halo.
 * dt_rk is the last substep's length (dt). */
int wrfmi_rk_scalar_pd_prep(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d,
                            float dt_rk, void *stream);
/* calc_p_rho_phi (module_big_step_utilities_em.F:953-1223), called at solve_em.F:2965. */
int wrfmi_calc_p_rho_phi(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d,
                         void *stream);
/* rk_phys_bc_dry_2 (module_bc_em.F:1045-1104, called at solve_em.F:3245) plus
 * the species fills that follow it (kind 0), or set_phys_bc_dry_2 and the
 * species fills at the end of the step (solve_em.F:4461-4540, kind 1). */
int wrfmi_phys_bc_dry_2(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d,
                        int kind, void *stream);
/* set_w_surface with fill_w_flag false (module_bc_em.F:1196-1295), called at
 * solve_em.F:4740 after the RK loop: w_2 at level 1 from u_2, v_2 and ht, interior only. */
int wrfmi_set_w_surface(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d,
                        void *stream);
/* set_physical_bc3d / set_physical_bc2d (share/module_bc.F:651-1113, 202-647)
 * on one field; var is the reference's staggering letter 'u','v','w','t','p'. */
int wrfmi_set_physical_bc3d(float *field, char var, const wrfmi_config *cfg,
                            const wrfmi_dims *d, void *stream);
int wrfmi_set_physical_bc2d(float *field, char var, const wrfmi_config *cfg,
                            const wrfmi_dims *d, void *stream);

/* ---- launch shapes (no counterpart in the reference, whose tile decomposition is
 * frame/module_tiles.F:200-460) ----------------------------------------------------
 * Rows per workgroup of the four-point and column kernels (1, 2 or 4; 8 too for the 5th-order
 * stencil kernels k_tend_u/v/w/t and k_adv_scalar), levels per thread of the
 * four-point advance_uv and column divergence (k_advance_uv, k_ac_div: 2, 4 or 8), and the
 * column solve's prefetch depth (k_ac_wsolve: 2, 4 or 8), per kernel (the names of
 * wrfmi_kernel_name) and domain shape in mass points (nx, ny, nz; 0 matches any).  Every
 * shape gives the same results bit for bit; wrf_amd/autotune.py measures the candidates and
 * writes the table.  wrfmi_init loads $WRFMI_TUNING_FILE, else wrfmi_tuning.txt beside the
 * library if present (lines "kernel nx ny nz rows levels prefetch").  A later entry takes
 * precedence; ensembles use the default shapes.  A step already captured in a graph keeps
 * the shapes it was captured with. */
int wrfmi_tuning_set(const char *kernel, int nx, int ny, int nz, int rows, int levels, int prefetch);
int wrfmi_tuning_clear(void);
int wrfmi_tuning_load(const char *path);
int wrfmi_tuning_get(const char *kernel, int nx, int ny, int nz, int *rows, int *levels, int *prefetch);

/* ---- kernel timing (diagnostic; no counterpart in the reference) ----------
 * While enabled, every launch of a selected kernel by the eager entry points
 * above is bracketed by a HIP event pair on its own stream; the graph entry
 * point is refused.  wrfmi_timing_enable(mask) clears the accumulators and
 * selects kernel n when bit n of mask is set (0 switches timing off);
 * wrfmi_timing_collect() waits for the recorded events and adds their elapsed
 * times; wrfmi_timing_get() returns launches and total milliseconds of kernel
 * `id` (0 .. wrfmi_kernel_count()-1, named by wrfmi_kernel_name). */
int wrfmi_kernel_count(void);
const char *wrfmi_kernel_name(int id);
int wrfmi_timing_enable(unsigned kernel_mask);
int wrfmi_timing_collect(void);
int wrfmi_timing_get(int id, long long *launches, double *total_ms);

/* Non-finite scan of the prognostic state (u_2, v_2, w_2, t_2, ph_2, mu_2);
 * returns WRFMI_E_NONFINITE and the count in *n_bad when any is found.  Waits for all
 * work queued on the device before it scans. */
int wrfmi_check_finite(const wrfmi_state *st, const wrfmi_dims *d, long long *n_bad);

/* Domain totals in FP64 (diagnostic; no counterpart routine in the reference):
 * out[0] dry-air mass [kg], out[1] theta budget sum(dm * theta) [kg K], out[2] kinetic,
 * out[3] internal, out[4] potential and out[5] total energy [J] of the mass columns
 * i = ids..ide-1, j = jds..jde-1, with layer dry mass dm = (c1h*mut + c2h)*(-dnw)/g times the
 * cell area dx*dy/(msftx*msfty).  FP64 accumulation throughout, except the theta budget's
 * column sums, which are FP32 with compensated (two-sum) accumulation; reductions run in a
 * fixed order (reproducible).  Enqueued on `stream`, then waits for it. */
int wrfmi_domain_totals(const wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d,
                        void *stream, double *out);

#ifdef __cplusplus
}
#endif
#endif
