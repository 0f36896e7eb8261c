// Per-routine entry points that keep the reference's argument lists (declared in
// include/wrfmi_routines.h): advect_u/v/w/scalar, advect_scalar_pd, rk_update_scalar,
// rk_update_scalar_pd, advance_uv, advance_mu_t,
// advance_w, sumflux, calc_p_rho and calc_coef_w, each with the arguments, in the order, of
// the Fortran subroutine it replaces, all by reference (as a Fortran caller passes them).
//
// Each binds its array arguments to the state slots the kernels read (the slots named after
// the grid fields solve_em passes at its call sites, e.g. advect_u's rom -> ww, fzm -> fnm,
// rdzw -> rdnw), leaves every other slot null, and launches the same kernels as the whole
// step in the routine's own form (KArgs::routine, or k_ac_wsolve without calc_p_rho), so the
// result is the reference routine's alone.  Scalar arguments that are namelist values
// (time_step, emdiv, epssm, smdiv, top_lid, non_hydrostatic) override the corresponding
// config field for the call.  The work goes on the stream wrfmi_set_stream chose for the
// calling thread and the call returns once it is enqueued.
#include <hip/hip_runtime.h>
#include <string.h>
#include "host_internal.h"
#include "wrfmi_routines.h"

using namespace wrfmi;
using namespace wrfmi::host;

namespace {

const dim3 kPtBlk(64, 4, 1);     // one-point kernels: 64 x 4 (i, j) columns
const dim3 kColBlk1(64, 1, 1);   // one column per lane, one row of 64 columns per workgroup

dim3 col_grid(const Dm &d, int i0, int j0, dim3 blk) {
  const int ni = d.ide - i0 + 1, nj = d.jde - j0 + 1;
  return dim3((ni + blk.x - 1) / blk.x, (nj + blk.y - 1) / blk.y, 1);
}
dim3 pt3(const Dm &d, int nk) {
  const dim3 g = col_grid(d, d.ids, d.jds, kPtBlk);
  return dim3(g.x, g.y, nk > 0 ? nk : 1);
}
dim3 columns(const Dm &d) { return col_grid(d, col_ia(d, false), d.jds, kColBlk1); }

// One call: the config copy, the dims, and the state slots bound so far.
struct Call {
  wrfmi_state s;
  wrfmi_config c;
  wrfmi_dims w;
  Dm d;
  int rc = 0;
  std::unique_lock<std::mutex> lk;
  Call(const wrfmi_config *cfg, const int *const dims[18]) : lk(lock()) {
    memset(&s, 0, sizeof s);
    memset(&c, 0, sizeof c);
    memset(&w, 0, sizeof w);
    for (int n = 0; n < 18; n++)
      if (!dims[n]) { rc = fail(WRFMI_E_POINTER, "null index argument"); return; }
    if (!cfg) { rc = fail(WRFMI_E_POINTER, "null config_flags"); return; }
    c = *cfg;
    int *p = &w.ids;
    for (int n = 0; n < 18; n++) p[n] = *dims[n];
    if ((rc = check_cfg(&c))) return;
    // the boundary-zone updates between these routines (spec_bdyupdate, relax_bdy_dry, ...)
    // run inside the state-struct entry points
    if (c.specified || c.nested) {
      rc = fail(WRFMI_E_BADOPT, "specified boundaries: use the state-struct entry points (wrfmi_solve_em_step, "
                "wrfmi_advance_uv_state, ...), which carry the boundary arrays");
      return;
    }
    rc = make_dm(&w, &d);
  }
  // bind array argument `name` to a state slot (device address after registry translation)
  template <typename T>
  void bind(T *&slot, const void *p, const char *name) {
    if (rc) return;
    if (!p) { rc = fail(WRFMI_E_POINTER, "argument %s is null", name); return; }
    slot = (T *)translate(const_cast<void *>(p));
    // a kernel given host memory would fault the GPU: check every array before any launch
    rc = device_array((const void *)slot, name);
  }
  void scalar(float &slot, const float *p, const char *name) {
    if (rc) return;
    if (!p) { rc = fail(WRFMI_E_POINTER, "argument %s is null", name); return; }
    slot = *p;
  }
  int iscalar(const int *p, const char *name) {
    if (rc) return 0;
    if (!p) { rc = fail(WRFMI_E_POINTER, "argument %s is null", name); return 0; }
    return *p;
  }
  // a physical constant the kernels hold at compile time
  void constant(const float *p, float want, const char *name) {
    if (rc) return;
    if (!p) { rc = fail(WRFMI_E_POINTER, "argument %s is null", name); return; }
    if (*p != want) rc = fail(WRFMI_E_BADOPT, "%s = %g: the kernels are built with %g", name, (double)*p, (double)want);
  }
  // the kernels take the old-time field of an advection routine to be the field itself, as
  // every solve_em call site passes it (module_em.F:505,546,589,625,1331); it is read only
  // by the radiation terms of open boundaries
  void same_old(const void *f, const void *f_old, const char *name) {
    if (rc || f == f_old || (c.periodic_x && c.periodic_y)) return;
    rc = fail(WRFMI_E_BADOPT, "%s must be the advected field itself on a domain with open boundaries", name);
  }
  KArgs args(int routine) const {
    KArgs a;
    memset(&a, 0, sizeof a);
    a.s = s;
    a.c = c;
    a.d = d;
    a.routine = routine;
    return a;
  }
  // the work arrays of the domain known by the device array `key` (by default its u_2 or w_2)
  void scratch(KArgs &a, const void *key = nullptr) {
    if (rc) return;
    float *dv = nullptr, *rhs = nullptr, *lo = nullptr;
    if (!key) key = s.u_2 ? (const void *)s.u_2 : (const void *)s.w_2;
    if ((rc = domain_scratch(key, d, &dv, &rhs, &lo))) return;
    a.scr_dv = FPtr<float>(dv);
    a.scr_rhs = FPtr<float>(rhs);
    a.scr_lo = FPtr<float>(lo);
  }
};

// the options of a call whose reference routine takes no config_flags: the index checks and
// FP32 (precision_mode 0)
wrfmi_config plain_config() {
  wrfmi_config c;
  memset(&c, 0, sizeof c);
  c.rk_ord = 3;
  c.h_mom_adv_order = c.h_sca_adv_order = 5;
  c.v_mom_adv_order = c.v_sca_adv_order = 3;
  c.momentum_adv_opt = 1;
  c.periodic_x = c.periodic_y = 1;
  c.non_hydrostatic = 1;
  c.hypsometric_opt = 2;
  return c;
}

#define DIMS_PTRS ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte

int finish(Call &C) {
  if (C.rc) return C.rc;
  C.lk.unlock();
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(WRFMI_E_DEVICE, "kernel launch: %s", hipGetErrorString(e));
  return 0;
}

// advect_u / advect_v / advect_w / advect_scalar: the same argument list
enum Adv { ADV_U, ADV_V, ADV_W, ADV_SCALAR };
int advect(Adv which, const float *f, const float *f_old, float *tendency, const float *ru, const float *rv,
           const float *rom, const float *c1, const float *c2, const float *mut, const int *time_step,
           const wrfmi_config *config_flags, const float *msfux, const float *msfuy, const float *msfvx,
           const float *msfvy, const float *msftx, const float *msfty, const float *fzm, const float *fzp,
           const float *rdx, const float *rdy, const float *rdz, const int *const dims[18]) {
  Call C(config_flags, dims);
  wrfmi_state &s = C.s;
  const char *fname = which == ADV_U ? "u_old" : which == ADV_V ? "v_old" : which == ADV_W ? "w_old" : "field_old";
  C.same_old(f, f_old, fname);
  switch (which) {
    case ADV_U: C.bind(s.u_2, f, "u"); C.bind(s.ru_tend, tendency, "tendency"); break;
    case ADV_V: C.bind(s.v_2, f, "v"); C.bind(s.rv_tend, tendency, "tendency"); break;
    case ADV_W: C.bind(s.w_2, f, "w"); C.bind(s.rw_tend, tendency, "tendency"); break;
    default: C.bind(s.q[0], f, "field"); C.bind(s.advect_tend, tendency, "tendency"); break;
  }
  if (which == ADV_SCALAR) {
    // advect_scalar's kernel reads its face velocities from the time-averaged flux slots,
    // as rk_scalar_tend binds them (module_em.F:1331)
    C.bind(s.ru_m, ru, "ru"); C.bind(s.rv_m, rv, "rv"); C.bind(s.ww_m, rom, "rom");
    s.n_species = 1;
  } else {
    C.bind(s.ru, ru, "ru"); C.bind(s.rv, rv, "rv"); C.bind(s.ww, rom, "rom");
  }
  C.bind(s.c1h, c1, "c1"); C.bind(s.c2h, c2, "c2"); C.bind(s.mut, mut, "mut");
  C.bind(s.msfux, msfux, "msfux"); C.bind(s.msfuy, msfuy, "msfuy"); C.bind(s.msfvx, msfvx, "msfvx");
  C.bind(s.msfvy, msfvy, "msfvy"); C.bind(s.msftx, msftx, "msftx"); C.bind(s.msfty, msfty, "msfty");
  C.bind(s.fnm, fzm, "fzm"); C.bind(s.fnp, fzp, "fzp");
  if (which == ADV_W) C.bind(s.rdn, rdz, "rdzu");
  else C.bind(s.rdnw, rdz, "rdzw");
  C.scalar(s.rdx, rdx, "rdx");
  C.scalar(s.rdy, rdy, "rdy");
  C.c.time_step = C.iscalar(time_step, "time_step");
  if (C.rc) return C.rc;
  KArgs a = C.args(R_ADVECT);
  const Dm &d = C.d;
  const bool per = C.c.periodic_x && C.c.periodic_y;
  const hipStream_t st = routine_stream();
  const int nk = d.kde - d.kds + 1;
  if (which == ADV_SCALAR) {
    const dim3 g((d.ide - d.ids + 1 + kAdvTileI - 1) / kAdvTileI, (d.jde - d.jds + 1 + kAdvRows - 1) / kAdvRows,
                 (nk + kAdvK - 1) / kAdvK);
    launch_tiles(per ? k_adv_scalar<KArgs, true, false> : k_adv_scalar<KArgs, false, false>, g, dim3(64, kAdvRows, 1), a, st);
  } else {
    const dim3 g = pt3(d, (nk + kTendK - 1) / kTendK);
    void (*k)(KArgs) = which == ADV_U ? (per ? k_tend_u<KArgs, true, false> : k_tend_u<KArgs, false, false>)
                       : which == ADV_V ? (per ? k_tend_v<KArgs, true, false> : k_tend_v<KArgs, false, false>)
                                        : (per ? k_tend_w<KArgs, true, false> : k_tend_w<KArgs, false, false>);
    launch_tiles(k, g, kPtBlk, a, st);
  }
  return finish(C);
}

}  // namespace

extern "C" int wrfmi_advect_u(const float *u, const float *u_old, float *tendency, const float *ru, const float *rv,
                              const float *rom, const float *c1, const float *c2, const float *mut, const int *time_step,
                              const wrfmi_config *config_flags, const float *msfux, const float *msfuy,
                              const float *msfvx, const float *msfvy, const float *msftx, const float *msfty,
                              const float *fzm, const float *fzp, const float *rdx, const float *rdy, const float *rdzw,
                              WRFMI_DIMS_REF) {
  const int *const dims[18] = {DIMS_PTRS};
  return advect(ADV_U, u, u_old, tendency, ru, rv, rom, c1, c2, mut, time_step, config_flags, msfux, msfuy, msfvx, msfvy,
                msftx, msfty, fzm, fzp, rdx, rdy, rdzw, dims);
}
extern "C" int wrfmi_advect_v(const float *v, const float *v_old, float *tendency, const float *ru, const float *rv,
                              const float *rom, const float *c1, const float *c2, const float *mut, const int *time_step,
                              const wrfmi_config *config_flags, const float *msfux, const float *msfuy,
                              const float *msfvx, const float *msfvy, const float *msftx, const float *msfty,
                              const float *fzm, const float *fzp, const float *rdx, const float *rdy, const float *rdzw,
                              WRFMI_DIMS_REF) {
  const int *const dims[18] = {DIMS_PTRS};
  return advect(ADV_V, v, v_old, tendency, ru, rv, rom, c1, c2, mut, time_step, config_flags, msfux, msfuy, msfvx, msfvy,
                msftx, msfty, fzm, fzp, rdx, rdy, rdzw, dims);
}
extern "C" int wrfmi_advect_w(const float *w, const float *w_old, float *tendency, const float *ru, const float *rv,
                              const float *rom, const float *c1, const float *c2, const float *mut, const int *time_step,
                              const wrfmi_config *config_flags, const float *msfux, const float *msfuy,
                              const float *msfvx, const float *msfvy, const float *msftx, const float *msfty,
                              const float *fzm, const float *fzp, const float *rdx, const float *rdy, const float *rdzu,
                              WRFMI_DIMS_REF) {
  const int *const dims[18] = {DIMS_PTRS};
  return advect(ADV_W, w, w_old, tendency, ru, rv, rom, c1, c2, mut, time_step, config_flags, msfux, msfuy, msfvx, msfvy,
                msftx, msfty, fzm, fzp, rdx, rdy, rdzu, dims);
}
extern "C" int wrfmi_advect_scalar(const float *field, const float *field_old, float *tendency, const float *ru,
                                   const float *rv, const float *rom, const float *c1, const float *c2, const float *mut,
                                   const int *time_step, const wrfmi_config *config_flags, const float *msfux,
                                   const float *msfuy, const float *msfvx, const float *msfvy, const float *msftx,
                                   const float *msfty, const float *fzm, const float *fzp, const float *rdx,
                                   const float *rdy, const float *rdzw, WRFMI_DIMS_REF) {
  const int *const dims[18] = {DIMS_PTRS};
  return advect(ADV_SCALAR, field, field_old, tendency, ru, rv, rom, c1, c2, mut, time_step, config_flags, msfux, msfuy,
                msfvx, msfvy, msftx, msfty, fzm, fzp, rdx, rdy, rdzw, dims);
}

extern "C" int wrfmi_rk_update_scalar(const int *scs, const int *sce, float *scalar_1, float *scalar_2,
                                      const float *sc_tend, float *advh_t, float *advz_t, const float *advect_tend,
                                      const float *h_tendency, const float *z_tendency, const float *msftx,
                                      const float *msfty, const float *c1, const float *c2, const float *mu_old,
                                      const float *mu_new, const float *mu_base, const int *rk_step, const float *dt,
                                      const int *spec_zone, const wrfmi_config *config_flags, const int *tenddec,
                                      WRFMI_DIMS_REF) {
  const int *const dims[18] = {DIMS_PTRS};
  Call C(config_flags, dims);
  wrfmi_state &s = C.s;
  (void)advh_t; (void)advz_t; (void)h_tendency; (void)z_tendency; (void)msftx;
  const int s0 = C.iscalar(scs, "scs"), s1 = C.iscalar(sce, "sce"), rk = C.iscalar(rk_step, "rk_step");
  C.iscalar(spec_zone, "spec_zone");  // its loop bounds apply to specified boundaries only
  if (!C.rc && C.iscalar(tenddec, "tenddec") && advh_t && advz_t && rk == C.c.rk_ord)
    C.rc = fail(WRFMI_E_BADOPT, "tenddec: the split horizontal/vertical tendency diagnostics are not implemented");
  if (!C.rc && (s1 < s0 || s1 - s0 >= WRFMI_MAX_SPECIES))
    C.rc = fail(WRFMI_E_BADOPT, "scs..sce = %d..%d: between 1 and %d species", s0, s1, WRFMI_MAX_SPECIES);
  if (!C.rc && (rk < 1 || rk > 3)) C.rc = fail(WRFMI_E_BADOPT, "rk_step=%d outside 1..3", rk);
  C.bind(s.advect_tend, advect_tend, "advect_tend");
  C.bind(s.msfty, msfty, "msfty");
  C.bind(s.c1h, c1, "c1"); C.bind(s.c2h, c2, "c2");
  C.bind(s.mu_1, mu_old, "mu_old"); C.bind(s.mu_2, mu_new, "mu_new"); C.bind(s.mub, mu_base, "mu_base");
  float dt_rk = 0.f;
  C.scalar(dt_rk, dt, "dt");
  if (C.rc) return C.rc;
  // scalar_1(ims:ime, kms:kme, jms:jme, scs:sce): species im at (im - scs) fields
  const size_t field = (size_t)C.d.sj * (size_t)(C.d.jme - C.d.jms + 1);
  const int n = s1 - s0 + 1;
  for (int m = 0; m < n; m++) {
    C.bind(s.q_old[m], scalar_1 ? scalar_1 + m * field : nullptr, "scalar_1");
    C.bind(s.q[m], scalar_2 ? scalar_2 + m * field : nullptr, "scalar_2");
    C.bind(s.q_tend[m], sc_tend ? sc_tend + m * field : nullptr, "sc_tend");
  }
  s.n_species = n;
  if (C.rc) return C.rc;
  KArgs a = C.args(R_FUSED);
  a.rk_step = rk;
  a.dt_rk = dt_rk;
  const hipStream_t st = routine_stream();
  for (int m = 0; m < n; m++) {
    a.species = m;
    launch_tiles(k_update_scalar<KArgs>, pt3(C.d, C.d.kde - C.d.kds), kPtBlk, a, st);
  }
  return finish(C);
}

extern "C" int wrfmi_rk_update_scalar_pd(const int *scs, const int *sce, float *scalar, float *sc_tend, const float *c1,
                                         const float *c2, const float *mu_old, const float *mu_new, const float *mu_base,
                                         const int *rk_step, const float *dt, const int *spec_zone,
                                         const wrfmi_config *config_flags, WRFMI_DIMS_REF) {
  const int *const dims[18] = {DIMS_PTRS};
  Call C(config_flags, dims);
  wrfmi_state &s = C.s;
  const int s0 = C.iscalar(scs, "scs"), s1 = C.iscalar(sce, "sce");
  C.iscalar(rk_step, "rk_step");      // not read by the routine
  C.iscalar(spec_zone, "spec_zone");  // its loop bounds apply to specified boundaries only
  if (!C.rc && (s1 < s0 || s1 - s0 >= WRFMI_MAX_SPECIES))
    C.rc = fail(WRFMI_E_BADOPT, "scs..sce = %d..%d: between 1 and %d species", s0, s1, WRFMI_MAX_SPECIES);
  C.bind(s.c1h, c1, "c1"); C.bind(s.c2h, c2, "c2");
  C.bind(s.mu_1, mu_old, "mu_old"); C.bind(s.mu_2, mu_new, "mu_new"); C.bind(s.mub, mu_base, "mu_base");
  float dt_rk = 0.f;
  C.scalar(dt_rk, dt, "dt");
  if (C.rc) return C.rc;
  // scalar(ims:ime, kms:kme, jms:jme, scs:sce): species im at (im - scs) fields
  const size_t field = (size_t)C.d.sj * (size_t)(C.d.jme - C.d.jms + 1);
  const int n = s1 - s0 + 1;
  for (int m = 0; m < n; m++) {
    C.bind(s.q_old[m], scalar ? scalar + m * field : nullptr, "scalar");
    C.bind(s.q_tend[m], sc_tend ? sc_tend + m * field : nullptr, "sc_tend");
  }
  s.n_species = n;
  if (C.rc) return C.rc;
  KArgs a = C.args(R_FUSED);
  a.dt_rk = dt_rk;
  const hipStream_t st = routine_stream();
  for (int m = 0; m < n; m++) {
    a.species = m;
    launch_tiles(k_update_scalar_pd<KArgs>, pt3(C.d, C.d.kde - C.d.kds), kPtBlk, a, st);
  }
  return finish(C);
}

extern "C" int wrfmi_advect_scalar_pd(const float *field, const float *field_old, float *tendency, float *h_tendency,
                                      float *z_tendency, const float *ru, const float *rv, const float *rom,
                                      const float *c1, const float *c2, const float *mut, const float *mub,
                                      const float *mu_old, const int *time_step, const wrfmi_config *config_flags,
                                      const int *tenddec, const float *msfux, const float *msfuy, const float *msfvx,
                                      const float *msfvy, const float *msftx, const float *msfty, const float *fzm,
                                      const float *fzp, const float *rdx, const float *rdy, const float *rdzw,
                                      const float *dt, WRFMI_DIMS_REF) {
  const int *const dims[18] = {DIMS_PTRS};
  Call C(config_flags, dims);
  wrfmi_state &s = C.s;
  (void)h_tendency; (void)z_tendency; (void)msfux; (void)msfuy; (void)msfvx; (void)msfvy;
  if (!C.rc && C.iscalar(tenddec, "tenddec"))
    C.rc = fail(WRFMI_E_BADOPT, "tenddec: the split horizontal/vertical tendency diagnostics are not implemented");
  C.bind(s.q[0], field, "field"); C.bind(s.q_old[0], field_old, "field_old");
  C.bind(s.advect_tend, tendency, "tendency");
  C.bind(s.ru_m, ru, "ru"); C.bind(s.rv_m, rv, "rv"); C.bind(s.ww_m, rom, "rom");
  C.bind(s.c1h, c1, "c1"); C.bind(s.c2h, c2, "c2");
  C.bind(s.muts, mut, "mut"); C.bind(s.mub, mub, "mub"); C.bind(s.mu_1, mu_old, "mu_old");
  C.bind(s.msftx, msftx, "msftx"); C.bind(s.msfty, msfty, "msfty");
  C.bind(s.fnm, fzm, "fzm"); C.bind(s.fnp, fzp, "fzp"); C.bind(s.rdnw, rdzw, "rdzw");
  C.scalar(s.rdx, rdx, "rdx");
  C.scalar(s.rdy, rdy, "rdy");
  float dt_step = 0.f;
  C.scalar(dt_step, dt, "dt");
  C.c.time_step = C.iscalar(time_step, "time_step");
  s.n_species = 1;
  if (C.rc) return C.rc;
  KArgs a = C.args(R_ADVECT);
  a.dt_rk = dt_step;
  // the limiter's scales go to the dvdxi work array of the domain known by its tendency array
  // (advect_tend at every call site)
  C.scratch(a, s.advect_tend);
  if (C.rc) return C.rc;
  const Dm &d = C.d;
  const hipStream_t st = routine_stream();
  const dim3 gs = col_grid(d, d.ids - 1, d.jds - 1, kPtBlk);
  launch_tiles(k_pd_scale<KArgs>, dim3(gs.x, gs.y, (d.kde - d.kds + kPdK - 1) / kPdK), kPtBlk, a, st);
  launch_tiles(k_pd_tend<KArgs>, pt3(d, (d.kde - d.kds + 1 + kPdK - 1) / kPdK), kPtBlk, a, st);
  return finish(C);
}

extern "C" int wrfmi_advance_uv(float *u, const float *ru_tend, float *v, const float *rv_tend, const float *p,
                                const float *pb, const float *ph, const float *php, const float *alt, const float *al,
                                const float *mu, const float *muu, const float *cqu, const float *muv, const float *cqv,
                                const float *mudf, const float *c1h, const float *c2h, const float *c1f,
                                const float *c2f, const float *c3h, const float *c4h, const float *c3f, const float *c4f,
                                const float *msfux, const float *msfuy, const float *msfvx, const float *msfvx_inv,
                                const float *msfvy, const float *rdx, const float *rdy, const float *dts,
                                const float *cf1, const float *cf2, const float *cf3, const float *fnm, const float *fnp,
                                const float *emdiv, const float *rdnw, const wrfmi_config *config_flags,
                                const int *spec_zone, const int *non_hydrostatic, const int *top_lid, WRFMI_DIMS_REF) {
  const int *const dims[18] = {DIMS_PTRS};
  Call C(config_flags, dims);
  wrfmi_state &s = C.s;
  (void)c1f; (void)c2f; (void)c3h; (void)c4h; (void)c3f; (void)c4f;
  C.bind(s.u_2, u, "u"); C.bind(s.ru_tend, ru_tend, "ru_tend"); C.bind(s.v_2, v, "v"); C.bind(s.rv_tend, rv_tend, "rv_tend");
  C.bind(s.p, p, "p"); C.bind(s.pb, pb, "pb"); C.bind(s.ph_2, ph, "ph"); C.bind(s.php, php, "php");
  C.bind(s.alt, alt, "alt"); C.bind(s.al, al, "al"); C.bind(s.mu_2, mu, "mu"); C.bind(s.muu, muu, "muu");
  C.bind(s.cqu, cqu, "cqu"); C.bind(s.muv, muv, "muv"); C.bind(s.cqv, cqv, "cqv"); C.bind(s.mudf, mudf, "mudf");
  C.bind(s.c1h, c1h, "c1h"); C.bind(s.c2h, c2h, "c2h");
  C.bind(s.msfux, msfux, "msfux"); C.bind(s.msfuy, msfuy, "msfuy"); C.bind(s.msfvx, msfvx, "msfvx");
  C.bind(s.msfvx_inv, msfvx_inv, "msfvx_inv"); C.bind(s.msfvy, msfvy, "msfvy");
  C.bind(s.fnm, fnm, "fnm"); C.bind(s.fnp, fnp, "fnp"); C.bind(s.rdnw, rdnw, "rdnw");
  C.scalar(s.rdx, rdx, "rdx"); C.scalar(s.rdy, rdy, "rdy");
  C.scalar(s.cf1, cf1, "cf1"); C.scalar(s.cf2, cf2, "cf2"); C.scalar(s.cf3, cf3, "cf3");
  C.scalar(C.c.emdiv, emdiv, "emdiv");
  float ts = 0.f;
  C.scalar(ts, dts, "dts");
  C.iscalar(spec_zone, "spec_zone");  // its loop bounds apply to specified/nested boundaries only
  C.c.non_hydrostatic = C.iscalar(non_hydrostatic, "non_hydrostatic");
  C.c.top_lid = C.iscalar(top_lid, "top_lid");
  if (!C.rc && !C.c.non_hydrostatic) C.rc = fail(WRFMI_E_BADOPT, "hydrostatic mode is not implemented");
  if (C.rc) return C.rc;
  KArgs a = C.args(R_FUSED);
  a.dts = ts;
  const bool per = C.c.periodic_x && C.c.periodic_y;
  const int nkb = (C.d.kde - C.d.kds + kAdvUvK - 1) / kAdvUvK;
  launch_tiles(per ? k_advance_uv<KArgs, true> : k_advance_uv<KArgs, false>, pt3(C.d, nkb), kPtBlk, a, routine_stream());
  return finish(C);
}

extern "C" int wrfmi_advance_mu_t(float *ww, const float *ww_1, const float *u, const float *u_1, const float *v,
                                  const float *v_1, float *mu, const float *mut, float *muave, float *muts,
                                  const float *muu, const float *muv, float *mudf, const float *c1h, const float *c2h,
                                  const float *c1f, const float *c2f, const float *c3h, const float *c4h,
                                  const float *c3f, const float *c4f, const float *uam, const float *vam,
                                  const float *wwam, float *t, const float *t_1, float *t_ave, const float *ft,
                                  const float *mu_tend, const float *rdx, const float *rdy, const float *dts,
                                  const float *epssm, const float *dnw, const float *fnm, const float *fnp,
                                  const float *rdnw, const float *msfux, const float *msfuy, const float *msfvx,
                                  const float *msfvx_inv, const float *msfvy, const float *msftx, const float *msfty,
                                  const int *step, const wrfmi_config *config_flags, WRFMI_DIMS_REF) {
  const int *const dims[18] = {DIMS_PTRS};
  Call C(config_flags, dims);
  wrfmi_state &s = C.s;
  (void)c1f; (void)c2f; (void)c3h; (void)c4h; (void)c3f; (void)c4f; (void)uam; (void)vam; (void)wwam;
  (void)msfux; (void)msfvx; (void)msfvy;
  C.bind(s.ww, ww, "ww"); C.bind(s.ww1, ww_1, "ww_1"); C.bind(s.u_2, u, "u"); C.bind(s.u_save, u_1, "u_1");
  C.bind(s.v_2, v, "v"); C.bind(s.v_save, v_1, "v_1"); C.bind(s.mu_2, mu, "mu"); C.bind(s.mut, mut, "mut");
  C.bind(s.muave, muave, "muave"); C.bind(s.muts, muts, "muts"); C.bind(s.muu, muu, "muu"); C.bind(s.muv, muv, "muv");
  C.bind(s.mudf, mudf, "mudf"); C.bind(s.c1h, c1h, "c1h"); C.bind(s.c2h, c2h, "c2h");
  C.bind(s.t_2, t, "t"); C.bind(s.t_save, t_1, "t_1"); C.bind(s.t_2save, t_ave, "t_ave"); C.bind(s.t_tend, ft, "ft");
  C.bind(s.mu_tend, mu_tend, "mu_tend"); C.bind(s.dnw, dnw, "dnw"); C.bind(s.fnm, fnm, "fnm"); C.bind(s.fnp, fnp, "fnp");
  C.bind(s.rdnw, rdnw, "rdnw"); C.bind(s.msfuy, msfuy, "msfuy"); C.bind(s.msfvx_inv, msfvx_inv, "msfvx_inv");
  C.bind(s.msftx, msftx, "msftx"); C.bind(s.msfty, msfty, "msfty");
  C.scalar(s.rdx, rdx, "rdx"); C.scalar(s.rdy, rdy, "rdy");
  C.scalar(C.c.epssm, epssm, "epssm");
  float ts = 0.f;
  C.scalar(ts, dts, "dts");
  C.iscalar(step, "step");
  if (C.rc) return C.rc;
  KArgs a = C.args(R_MU_T);
  a.dts = ts;
  C.scratch(a);
  if (C.rc) return C.rc;
  const Dm &d = C.d;
  const hipStream_t st = routine_stream();
  const int nk = d.kde - d.kds + 1;
  launch_tiles(k_ac_div<KArgs>, pt3(d, nk), kPtBlk, a, st);
  launch_tiles(C.c.precision_mode == 1 ? k_ac_mu_ww<KArgs, true> : k_ac_mu_ww<KArgs, false>, columns(d), kColBlk1, a, st);
  launch_tiles(k_ac_theta<KArgs>, pt3(d, nk), kPtBlk, a, st);
  return finish(C);
}

extern "C" int wrfmi_advance_w(float *w, const float *rw_tend, const float *ww, const float *w_save, const float *u,
                               const float *v, const float *mu1, const float *mut, const float *muave,
                               const float *muts, const float *c1h, const float *c2h, const float *c1f,
                               const float *c2f, const float *c3h, const float *c4h, const float *c3f, const float *c4f,
                               float *t_2ave, const float *t_2, const float *t_1, float *ph, const float *ph_1,
                               const float *phb, const float *ph_tend, const float *ht, const float *c2a,
                               const float *cqw, const float *alt, const float *alb, void *a, void *alpha, void *gamma,
                               const float *rdx, const float *rdy, const float *dts, const float *t0,
                               const float *epssm, const float *dnw, const float *fnm, const float *fnp,
                               const float *rdnw, const float *rdn, const float *cf1, const float *cf2,
                               const float *cf3, const float *msftx, const float *msfty,
                               const wrfmi_config *config_flags, const int *top_lid, WRFMI_DIMS_REF) {
  const int *const dims[18] = {DIMS_PTRS};
  Call C(config_flags, dims);
  wrfmi_state &s = C.s;
  (void)c3h; (void)c4h; (void)c3f; (void)c4f; (void)alb; (void)dnw; (void)a; (void)alpha;
  C.bind(s.w_2, w, "w"); C.bind(s.rw_tend, rw_tend, "rw_tend"); C.bind(s.ww, ww, "ww");
  C.bind(s.w_save, w_save, "w_save");  // read by the damp_opt = 3 branch
  C.bind(s.u_2, u, "u"); C.bind(s.v_2, v, "v"); C.bind(s.mu_2, mu1, "mu1"); C.bind(s.mut, mut, "mut");
  C.bind(s.muave, muave, "muave"); C.bind(s.muts, muts, "muts");
  C.bind(s.c1h, c1h, "c1h"); C.bind(s.c2h, c2h, "c2h"); C.bind(s.c1f, c1f, "c1f"); C.bind(s.c2f, c2f, "c2f");
  C.bind(s.t_2save, t_2ave, "t_2ave"); C.bind(s.t_2, t_2, "t_2"); C.bind(s.t_save, t_1, "t_1");
  C.bind(s.ph_2, ph, "ph"); C.bind(s.ph_save, ph_1, "ph_1"); C.bind(s.phb, phb, "phb"); C.bind(s.ph_tend, ph_tend, "ph_tend");
  C.bind(s.ht, ht, "ht"); C.bind(s.c2a, c2a, "c2a"); C.bind(s.cqw, cqw, "cqw"); C.bind(s.alt, alt, "alt");
  C.bind(s.gamma, gamma, "gamma");
  C.bind(s.fnm, fnm, "fnm"); C.bind(s.fnp, fnp, "fnp"); C.bind(s.rdnw, rdnw, "rdnw"); C.bind(s.rdn, rdn, "rdn");
  C.bind(s.msftx, msftx, "msftx"); C.bind(s.msfty, msfty, "msfty");
  C.scalar(s.rdx, rdx, "rdx"); C.scalar(s.rdy, rdy, "rdy");
  C.scalar(s.cf1, cf1, "cf1"); C.scalar(s.cf2, cf2, "cf2"); C.scalar(s.cf3, cf3, "cf3");
  C.scalar(C.c.epssm, epssm, "epssm");
  C.constant(t0, 300.f, "t0");
  float ts = 0.f;
  C.scalar(ts, dts, "dts");
  C.c.top_lid = C.iscalar(top_lid, "top_lid");
  // a, alpha and gamma are REAL arrays here, as in the reference; the precision form's
  // FP64 coefficients need arrays of doubles (the state-struct entry points)
  if (!C.rc && C.c.precision_mode != 0)
    C.rc = fail(WRFMI_E_BADOPT, "advance_w with REAL a, alpha, gamma: precision_mode must be 0");
  if (C.rc) return C.rc;
  KArgs k = C.args(R_W_TAVE);
  k.dts = ts;
  C.scratch(k);
  if (C.rc) return C.rc;
  const Dm &d = C.d;
  const hipStream_t st = routine_stream();
  // t_2ave first (advance_w 1312-1320), then the column solve without calc_p_rho
  launch_tiles(k_ac_theta<KArgs>, pt3(d, d.kde - d.kds + 1), kPtBlk, k, st);
  k.routine = R_FUSED;
  launch_tiles(k_ac_wsolve<KArgs, float, false>, columns(d), kColBlk1, k, st);
  return finish(C);
}

extern "C" int wrfmi_sumflux(const float *ru, const float *rv, const float *ww, const float *u_lin, const float *v_lin,
                             const float *ww_lin, const float *muu, const float *muv, const float *c1h,
                             const float *c2h, const float *c1f, const float *c2f, const float *c3h, const float *c4h,
                             const float *c3f, const float *c4f, float *ru_m, float *rv_m, float *ww_m,
                             const float *epssm, const float *msfux, const float *msfuy, const float *msfvx,
                             const float *msfvx_inv, const float *msfvy, const int *iteration,
                             const int *number_of_small_timesteps, WRFMI_DIMS_REF) {
  const int *const dims[18] = {DIMS_PTRS};
  const wrfmi_config cfg = plain_config();  // sumflux takes no config_flags
  Call C(&cfg, dims);
  wrfmi_state &s = C.s;
  (void)c1f; (void)c2f; (void)c3h; (void)c4h; (void)c3f; (void)c4f; (void)epssm; (void)msfux; (void)msfvx; (void)msfvy;
  C.bind(s.u_2, ru, "ru"); C.bind(s.v_2, rv, "rv"); C.bind(s.ww, ww, "ww");
  C.bind(s.u_save, u_lin, "u_lin"); C.bind(s.v_save, v_lin, "v_lin"); C.bind(s.ww1, ww_lin, "ww_lin");
  C.bind(s.muu, muu, "muu"); C.bind(s.muv, muv, "muv"); C.bind(s.c1h, c1h, "c1h"); C.bind(s.c2h, c2h, "c2h");
  C.bind(s.ru_m, ru_m, "ru_m"); C.bind(s.rv_m, rv_m, "rv_m"); C.bind(s.ww_m, ww_m, "ww_m");
  C.bind(s.msfuy, msfuy, "msfuy"); C.bind(s.msfvx_inv, msfvx_inv, "msfvx_inv");
  const int it = C.iscalar(iteration, "iteration"), ns = C.iscalar(number_of_small_timesteps, "number_of_small_timesteps");
  if (!C.rc && (it < 1 || it > ns)) C.rc = fail(WRFMI_E_BADOPT, "iteration %d outside 1..%d", it, ns);
  if (C.rc) return C.rc;
  KArgs a = C.args(R_SUMFLUX);
  a.iteration = it;
  a.nsmall = ns;
  const hipStream_t st = routine_stream();
  const int nk = C.d.kde - C.d.kds + 1;
  launch_tiles(k_ac_div<KArgs>, pt3(C.d, nk), kPtBlk, a, st);
  launch_tiles(k_ac_theta<KArgs>, pt3(C.d, nk), kPtBlk, a, st);
  return finish(C);
}

extern "C" int wrfmi_calc_p_rho(float *al, float *p, float *ph, const float *alt, const float *t_2, const float *t_1,
                                const float *c2a, float *pm1, const float *mu, const float *mut, const float *c1h,
                                const float *c2h, const float *c1f, const float *c2f, const float *c3h,
                                const float *c4h, const float *c3f, const float *c4f, const float *znu, const float *t0,
                                const float *rdnw, const float *dnw, const float *smdiv, const int *non_hydrostatic,
                                const int *step, WRFMI_DIMS_REF) {
  const int *const dims[18] = {DIMS_PTRS};
  const wrfmi_config cfg = plain_config();  // calc_p_rho takes no config_flags
  Call C(&cfg, dims);
  wrfmi_state &s = C.s;
  (void)c1f; (void)c2f; (void)c3h; (void)c4h; (void)c3f; (void)c4f; (void)znu; (void)dnw;
  C.bind(s.al, al, "al"); C.bind(s.p, p, "p"); C.bind(s.ph_2, ph, "ph"); C.bind(s.alt, alt, "alt");
  C.bind(s.t_2, t_2, "t_2"); C.bind(s.t_save, t_1, "t_1"); C.bind(s.c2a, c2a, "c2a"); C.bind(s.pm1, pm1, "pm1");
  C.bind(s.mu_2, mu, "mu"); C.bind(s.muts, mut, "mut"); C.bind(s.c1h, c1h, "c1h"); C.bind(s.c2h, c2h, "c2h");
  C.bind(s.rdnw, rdnw, "rdnw");
  C.constant(t0, 300.f, "t0");
  C.scalar(C.c.smdiv, smdiv, "smdiv");
  if (!C.rc && !C.iscalar(non_hydrostatic, "non_hydrostatic")) C.rc = fail(WRFMI_E_BADOPT, "hydrostatic mode is not implemented");
  const int stp = C.iscalar(step, "step");
  if (C.rc) return C.rc;
  KArgs a = C.args(R_FUSED);
  a.iteration = stp;
  launch_tiles(k_calc_p_rho<KArgs>, pt3(C.d, C.d.kde - C.d.kds), kPtBlk, a, routine_stream());
  return finish(C);
}

extern "C" int wrfmi_calc_coef_w(void *a, void *alpha, void *gamma, const float *mut, const float *c1h, const float *c2h,
                                 const float *c1f, const float *c2f, const float *c3h, const float *c4h,
                                 const float *c3f, const float *c4f, const float *cqw, const float *rdn,
                                 const float *rdnw, const float *c2a, const float *dts, const float *g,
                                 const float *epssm, const int *top_lid, WRFMI_DIMS_REF) {
  const int *const dims[18] = {DIMS_PTRS};
  const wrfmi_config cfg = plain_config();  // calc_coef_w takes no config_flags: REAL (FP32) coefficients
  Call C(&cfg, dims);
  wrfmi_state &s = C.s;
  (void)c3h; (void)c4h; (void)c3f; (void)c4f;
  C.bind(s.a, a, "a"); C.bind(s.alpha, alpha, "alpha"); C.bind(s.gamma, gamma, "gamma");
  C.bind(s.mut, mut, "mut"); C.bind(s.c1h, c1h, "c1h"); C.bind(s.c2h, c2h, "c2h"); C.bind(s.c1f, c1f, "c1f");
  C.bind(s.c2f, c2f, "c2f"); C.bind(s.cqw, cqw, "cqw"); C.bind(s.rdn, rdn, "rdn"); C.bind(s.rdnw, rdnw, "rdnw");
  C.bind(s.c2a, c2a, "c2a");
  C.constant(g, 9.81f, "g");
  C.scalar(C.c.epssm, epssm, "epssm");
  C.c.top_lid = C.iscalar(top_lid, "top_lid");
  float ts = 0.f;
  C.scalar(ts, dts, "dts");
  if (C.rc) return C.rc;
  KArgs k = C.args(R_COEF_W);
  k.dts = ts;
  launch_tiles(k_coef_w<KArgs, float, true>, columns(C.d), kColBlk1, k, routine_stream());
  return finish(C);
}
