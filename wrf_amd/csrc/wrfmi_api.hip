// C-ABI of the MI355X WRF-ARW dynamics step (declarations in include/wrfmi.h).
// Host driver: the call order of dyn_em/solve_em.F:578-4540 for the implemented options,
// expressed as a sequence of kernel launches on one HIP stream (optionally replayed from a
// captured hipGraph), with halo fills at the points where the reference calls
// set_physical_bc3d/2d.
#include <hip/hip_runtime.h>
#include <dlfcn.h>
#include <stdarg.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <initializer_list>
#include <iterator>
#include <map>
#include <mutex>
#include <string>
#include <vector>
#include "host_internal.h"


using namespace wrfmi;

static thread_local std::string g_err;
static std::mutex g_mu;

static int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
#define HIPCHK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) return fail(WRFMI_E_DEVICE, "%s: %s", #x, hipGetErrorString(e_));   \
  } while (0)

// ------------------------------------------------------------------ registry of host arrays
struct Mirror { void *dev; size_t bytes; };
// Before a mirror is freed: drop the cached graphs and work arrays of domains that use it,
// and mark the ensembles that use it stale (g_mu held, device idle).
static void drop_device_range(uintptr_t lo, uintptr_t hi);
static std::map<uintptr_t, Mirror> g_reg;  // host base address -> mirror

static void *translate(void *p) {
  if (!p) return p;
  uintptr_t a = (uintptr_t)p;
  auto it = g_reg.upper_bound(a);
  if (it == g_reg.begin()) return p;
  --it;
  if (a >= it->first && a < it->first + it->second.bytes) return (char *)it->second.dev + (a - it->first);
  return p;
}

extern "C" int wrfmi_register_field(void *host_ptr, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!host_ptr || !bytes) return fail(WRFMI_E_POINTER, "register: null pointer or zero size");
  const uintptr_t a = (uintptr_t)host_ptr;
  auto it = g_reg.find(a);
  if (it != g_reg.end()) {
    // registering the same array again refreshes its mirror from the host
    if (it->second.bytes == bytes) {
      HIPCHK(hipMemcpy(it->second.dev, host_ptr, bytes, hipMemcpyHostToDevice));
      return 0;
    }
    HIPCHK(hipDeviceSynchronize());
    drop_device_range((uintptr_t)it->second.dev, (uintptr_t)it->second.dev + it->second.bytes);
    HIPCHK(hipFree(it->second.dev));
    g_reg.erase(it);
  }
  auto nx = g_reg.lower_bound(a);
  if (nx != g_reg.end() && nx->first < a + bytes)
    return fail(WRFMI_E_POINTER, "register: range overlaps another registered array (unregister it first)");
  if (nx != g_reg.begin()) {
    auto pv = std::prev(nx);
    if (pv->first + pv->second.bytes > a)
      return fail(WRFMI_E_POINTER, "register: range overlaps another registered array (unregister it first)");
  }
  void *dev = nullptr;
  HIPCHK(hipMalloc(&dev, bytes));
  HIPCHK(hipMemcpy(dev, host_ptr, bytes, hipMemcpyHostToDevice));
  g_reg[(uintptr_t)host_ptr] = Mirror{dev, bytes};
  return 0;
}
extern "C" int wrfmi_unregister_field(void *host_ptr) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_reg.find((uintptr_t)host_ptr);
  if (it == g_reg.end()) return fail(WRFMI_E_POINTER, "unregister: pointer was not registered");
  HIPCHK(hipDeviceSynchronize());
  drop_device_range((uintptr_t)it->second.dev, (uintptr_t)it->second.dev + it->second.bytes);
  HIPCHK(hipFree(it->second.dev));
  g_reg.erase(it);
  return 0;
}
extern "C" int wrfmi_sync_to_device(void *host_ptr) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_reg.find((uintptr_t)host_ptr);
  if (it == g_reg.end()) return fail(WRFMI_E_POINTER, "sync_to_device: pointer was not registered");
  HIPCHK(hipMemcpy(it->second.dev, host_ptr, it->second.bytes, hipMemcpyHostToDevice));
  return 0;
}
extern "C" int wrfmi_sync_to_host(void *host_ptr) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_reg.find((uintptr_t)host_ptr);
  if (it == g_reg.end()) return fail(WRFMI_E_POINTER, "sync_to_host: pointer was not registered");
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(host_ptr, it->second.dev, it->second.bytes, hipMemcpyDeviceToHost));
  return 0;
}

// ------------------------------------------------------------------ work arrays
// Library-owned 3-D work arrays (one grid field each) of one domain.  A domain is known by
// the device address of its u_2 array and its field size, so two domains never share work
// arrays, and steps of different domains may run concurrently on different streams.  They
// are allocated on the domain's first call, before any launch is enqueued or captured, and
// kept until wrfmi_domain_release or wrfmi_finalize.
struct Scratch { float *dv, *rhs, *lo; };
typedef std::pair<uintptr_t, size_t> DomainKey;  // (device address of u_2, elements per field)
static std::map<DomainKey, Scratch> g_scratch;

static int get_scratch(DomainKey key, Scratch *out) {
  auto it = g_scratch.find(key);
  if (it == g_scratch.end()) {
    Scratch sc{nullptr, nullptr, nullptr};
    const size_t bytes = key.second * sizeof(float);
    if (hipMalloc(&sc.dv, bytes) != hipSuccess || hipMalloc(&sc.rhs, bytes) != hipSuccess ||
        hipMalloc(&sc.lo, bytes) != hipSuccess) {
      (void)hipGetLastError();
      if (sc.dv) (void)hipFree(sc.dv);
      if (sc.rhs) (void)hipFree(sc.rhs);
      return fail(WRFMI_E_DEVICE, "cannot allocate %zu bytes of work arrays", 3 * bytes);
    }
    it = g_scratch.emplace(key, sc).first;
  }
  *out = it->second;
  return 0;
}
// Advection tendencies of species 0 .. n-2 for the all-species scalar kernels: n-1 grid
// fields per (domain, species count), allocated on first use.
static std::map<std::pair<DomainKey, int>, float *> g_qt;
static int get_qt(DomainKey key, int nspecies, float **out) {
  const auto qk = std::make_pair(key, nspecies);
  auto it = g_qt.find(qk);
  if (it == g_qt.end()) {
    float *p = nullptr;
    const size_t bytes = key.second * sizeof(float) * (size_t)(nspecies - 1);
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(WRFMI_E_DEVICE, "cannot allocate %zu bytes of species work arrays", bytes);
    }
    it = g_qt.emplace(qk, p).first;
  }
  *out = it->second;
  return 0;
}
// free the work arrays of domains whose u_2 lies in [lo, hi)
static void free_scratch_in(uintptr_t lo, uintptr_t hi) {
  for (auto it = g_scratch.begin(); it != g_scratch.end();) {
    if (it->first.first >= lo && it->first.first < hi) {
      (void)hipFree(it->second.dv);
      (void)hipFree(it->second.rhs);
      (void)hipFree(it->second.lo);
      it = g_scratch.erase(it);
    } else {
      ++it;
    }
  }
  for (auto it = g_qt.begin(); it != g_qt.end();) {
    if (it->first.first.first >= lo && it->first.first.first < hi) {
      (void)hipFree(it->second);
      it = g_qt.erase(it);
    } else {
      ++it;
    }
  }
}
static void free_scratch() { free_scratch_in(0, ~(uintptr_t)0); }

// ------------------------------------------------------------------ library
static hipStream_t g_stream = nullptr;
static unsigned long long *g_counter = nullptr;
static double *g_diag = nullptr;  // wrfmi_domain_totals' partial sums (g_mu)
static size_t g_diag_bytes = 0;
static int noalloc_scratch(float **out);

static int tune_load(const char *path);
static bool g_tune_loaded = false;
extern "C" int wrfmi_init(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_tune_loaded) {
    // the autotuner's table: $WRFMI_TUNING_FILE, else wrfmi_tuning.txt beside this library
    g_tune_loaded = true;
    const char *env = getenv("WRFMI_TUNING_FILE");
    std::string path;
    if (env && *env) {
      path = env;
    } else {
      Dl_info info;
      if (dladdr((void *)&wrfmi_init, &info) && info.dli_fname) {
        path = info.dli_fname;
        const size_t slash = path.rfind('/');
        path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/wrfmi_tuning.txt";
        FILE *f = fopen(path.c_str(), "r");
        if (f) fclose(f);
        else path.clear();
      }
    }
    if (!path.empty())
      if (int rc = tune_load(path.c_str())) return rc;
  }
  if (!g_stream) HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  if (!g_counter) HIPCHK(hipMalloc(&g_counter, sizeof(unsigned long long)));
  float *p = nullptr;
  return noalloc_scratch(&p);  // the allocation-free step's work array on this device
}
struct GraphEntry;
static std::vector<GraphEntry *> g_graphs;
static void free_graphs();
static void free_ensembles();
extern "C" int wrfmi_finalize(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  free_graphs();
  free_ensembles();
  free_scratch();
  for (auto &kv : g_reg) (void)hipFree(kv.second.dev);
  g_reg.clear();
  if (g_counter) { (void)hipFree(g_counter); g_counter = nullptr; }
  if (g_diag) { (void)hipFree(g_diag); g_diag = nullptr; g_diag_bytes = 0; }
  if (g_stream) { (void)hipStreamDestroy(g_stream); g_stream = nullptr; }
  return 0;
}
extern "C" const char *wrfmi_last_error(void) { return g_err.c_str(); }
extern "C" const char *wrfmi_version(void) { return "1.0"; }
extern "C" int wrfmi_state_field_count(void) {
  int n = 0;
#define CNT(x) n++;
  WRFMI_FIELDS_3D(CNT) WRFMI_FIELDS_2D(CNT) WRFMI_FIELDS_1D(CNT)
#undef CNT
  return n;
}
extern "C" size_t wrfmi_state_sizeof(void) { return sizeof(wrfmi_state); }

// ------------------------------------------------------------------ validation
static int check_cfg(const wrfmi_config *c) {
  if (c->rk_ord != 3) return fail(WRFMI_E_BADOPT, "rk_ord=%d: only the third-order Runge-Kutta step is implemented", c->rk_ord);
  if (c->h_mom_adv_order != 5 || c->v_mom_adv_order != 3 || c->h_sca_adv_order != 5 || c->v_sca_adv_order != 3)
    return fail(WRFMI_E_BADOPT, "advection orders h/v = %d/%d (momentum), %d/%d (scalar): only 5/3 is implemented",
                c->h_mom_adv_order, c->v_mom_adv_order, c->h_sca_adv_order, c->v_sca_adv_order);
  if (c->nested || c->polar || c->symmetric_xs || c->symmetric_xe || c->symmetric_ys || c->symmetric_ye)
    return fail(WRFMI_E_BADOPT, "lateral boundaries: periodic, open and specified are implemented (not nested/polar/symmetric)");
  if (c->specified) {
    if (c->periodic_y || c->open_xs || c->open_xe || c->open_ys || c->open_ye)
      return fail(WRFMI_E_BADOPT, "specified boundaries hold all four sides (or the y sides with a periodic x)");
    if (c->spec_zone != 1 || c->relax_zone < 2 || c->spec_bdy_width < c->spec_zone + c->relax_zone ||
        c->spec_bdy_width > WRFMI_MAX_BDY_WIDTH)
      return fail(WRFMI_E_BADOPT, "specified boundaries: spec_zone 1, relax_zone >= 2 and spec_bdy_width >= spec_zone + "
                  "relax_zone (at most %d) are implemented", WRFMI_MAX_BDY_WIDTH);
  } else {
    if (!c->periodic_x && !(c->open_xs && c->open_xe)) return fail(WRFMI_E_BADOPT, "x boundaries must be periodic or open on both sides");
    if (!c->periodic_y && !(c->open_ys && c->open_ye)) return fail(WRFMI_E_BADOPT, "y boundaries must be periodic or open on both sides");
  }
  if (c->moist_adv_opt < 0 || c->moist_adv_opt > 1 || c->scalar_adv_opt < 0 || c->scalar_adv_opt > 1 ||
      c->momentum_adv_opt != 1)
    return fail(WRFMI_E_BADOPT, "advection options: the original fluxes and positive-definite scalar transport "
                "(moist/scalar_adv_opt 0 or 1, momentum_adv_opt 1) are implemented, not monotonic or WENO");
  if (!c->non_hydrostatic) return fail(WRFMI_E_BADOPT, "hydrostatic mode is not implemented");
  if (c->diff_opt < 0 || c->diff_opt > 1 || (c->diff_opt == 1 && c->km_opt != 1) || c->diff_6th_opt < 0 ||
      c->diff_6th_opt > 2 || c->damp_opt == 1 || c->damp_opt < 0 || c->damp_opt > 3 || c->zadvect_implicit ||
      c->pert_coriolis || c->map_proj == 6)
    return fail(WRFMI_E_BADOPT, "diff_opt 2 (TKE diffusion), km_opt other than 1, damp_opt 1, IEVA, perturbation "
                "Coriolis and lat-lon grids are outside the implemented set (diff_opt 0 and 1 with km_opt 1, "
                "diff_6th_opt 0-2 and damp_opt 0, 2, 3 are implemented)");
  if (c->hypsometric_opt != 1 && c->hypsometric_opt != 2) return fail(WRFMI_E_BADOPT, "hypsometric_opt must be 1 or 2");
  if (c->precision_mode != 0 && c->precision_mode != 1) return fail(WRFMI_E_BADOPT, "precision_mode must be 0 or 1");
  return 0;
}

static int make_dm(const wrfmi_dims *w, Dm *d) {
  if (w->its != w->ids || w->ite != w->ide || w->jts != w->jds || w->jte != w->jde || w->kts != w->kds || w->kte != w->kde)
    return fail(WRFMI_E_DIMS, "the GPU path processes the whole patch: tile must equal the domain");
  if (w->ids - w->ims < 4 || w->jds - w->jms < 4 || w->ime - w->ide < 4 || w->jme - w->jde < 4)
    return fail(WRFMI_E_DIMS, "memory dims need a halo of at least 4 (bdyzone) around the domain");
  // the kernels take level 1 as the surface (k_ac_wsolve's surface w from u, v at levels
  // 1..3, calc_coef_w's bottom row), as the reference's vertical loops do
  if (w->kds != 1 || w->kms != 1) return fail(WRFMI_E_DIMS, "kds and kms must be 1 (level 1 is the surface)");
  if (w->kme < w->kde) return fail(WRFMI_E_DIMS, "memory k range must reach kde");
  if (w->kde - w->kds < 4 || w->ide - w->ids < 7 || w->jde - w->jds < 7) return fail(WRFMI_E_DIMS, "domain too small");
  d->ids = w->ids; d->ide = w->ide; d->jds = w->jds; d->jde = w->jde; d->kds = w->kds; d->kde = w->kde;
  d->ims = w->ims; d->ime = w->ime; d->jms = w->jms; d->jme = w->jme; d->kms = w->kms; d->kme = w->kme;
  long long ni = w->ime - w->ims + 1, nk = w->kme - w->kms + 1, nj = w->jme - w->jms + 1;
  // kernels address elements by 32-bit byte offsets, twice the element offset for the FP64
  // tridiagonal coefficients: fields stay below 2^29 elements
  if (ni * nk * nj >= (1LL << 29)) return fail(WRFMI_E_DIMS, "domain of 2^29 or more elements per field");
  d->ni = (int)ni; d->nk = (int)nk; d->sj = (int)(ni * nk);
  return 0;
}

static int translate_state(const wrfmi_state *in, wrfmi_state *out) {
  *out = *in;
#define TR(n) out->n = (float *)translate((void *)in->n); if (!out->n) return fail(WRFMI_E_POINTER, "state field %s is null", #n);
  WRFMI_FIELDS_3D(TR) WRFMI_FIELDS_2D(TR) WRFMI_FIELDS_1D(TR)
#undef TR
  out->a = translate(in->a); out->alpha = translate(in->alpha); out->gamma = translate(in->gamma);
  if (!out->a || !out->alpha || !out->gamma) return fail(WRFMI_E_POINTER, "tridiagonal coefficient arrays are null");
  if (in->n_moist < 0 || in->n_species < in->n_moist || in->n_species > WRFMI_MAX_SPECIES)
    return fail(WRFMI_E_BADOPT, "n_moist=%d n_species=%d out of range", in->n_moist, in->n_species);
  for (int m = 0; m < in->n_species; m++) {
    out->q[m] = (float *)translate(in->q[m]);
    out->q_old[m] = (float *)translate(in->q_old[m]);
    out->q_tend[m] = (float *)translate(in->q_tend[m]);
    if (!out->q[m] || !out->q_old[m] || !out->q_tend[m]) return fail(WRFMI_E_POINTER, "species %d arrays are null", m);
  }
  // every pointer must now be device-accessible
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, out->u_2) != hipSuccess || at.type != hipMemoryTypeDevice) {
    (void)hipGetLastError();
    return fail(WRFMI_E_POINTER, "u_2 is neither a device pointer nor a registered host array");
  }
  return 0;
}

// ------------------------------------------------------------------ launch helpers
static inline dim3 col_grid(const Dm &d, int i0, int j0, dim3 blk) {
  int ni = d.ide - i0 + 1, nj = d.jde - j0 + 1;
  return dim3((ni + blk.x - 1) / blk.x, (nj + blk.y - 1) / blk.y, 1);
}
static const dim3 kColBlk(64, 4, 1);
// workgroup of the four-points-per-lane kernels: one 256-point row per wave, WRFMI_VEC_ROWS
// rows.  With one row, the tiles of neighbouring rows follow each other more closely in the
// XCD-contiguous tile order, and a launch's waves spread evenly over the CUs: at 400x400x80
// the step takes 17.45 ms with 1 row, 17.46 with 2 and 17.67 with 4 (8 and 16 were slower
// still).
#ifndef WRFMI_VEC_ROWS
#define WRFMI_VEC_ROWS 1
#endif
// k_adv_scalar_all_x4 and k_tend_t_x4 keep four rows: their 5th-order y stencils read the
// j-3..j+3 rows of several fields, which the four waves of a workgroup share in L2 (with
// one row: 1.41 and 1.11 ms per step at 400x400x80, with four: 1.33 and 1.08)
#ifndef WRFMI_ADVALL_ROWS
#define WRFMI_ADVALL_ROWS 4
#endif
#ifndef WRFMI_TENDT_ROWS
#define WRFMI_TENDT_ROWS 4
#endif
static const dim3 kVecBlk(64, WRFMI_VEC_ROWS, 1);
// workgroup of the one-column-per-lane kernels: 64 columns by WRFMI_COL_ROWS rows (see g_col)
#ifndef WRFMI_COL_ROWS
#define WRFMI_COL_ROWS 1
#endif

// ------------------------------------------------------------------ kernel timing (diagnostic)
enum KernelId {
  KID_STEP_PREP, KID_TEND_U, KID_TEND_V, KID_TEND_W, KID_TEND_T, KID_SMALL_PREP, KID_ADVANCE_UV,
  KID_AC_DIV, KID_SMALL_FINISH, KID_ADV_SCALAR, KID_UPDATE_SCALAR, KID_P_RHO_PHI, KID_HALO, KID_P_RHO0,
  KID_COEF_W, KID_WW_SCAN, KID_AC_MU_WW, KID_AC_THETA, KID_AC_WSOLVE, KID_W_SURFACE, KID_PD_SCALE, KID_PD_TEND,
  KID_UPDATE_SCALAR_PD, KID_DIFF6, KID_RAYLEIGH, KID_DIFF2, KID_LBC, KID_COUNT
};
static const char *kKernelNames[KID_COUNT] = {
  "k_step_prep", "k_tend_u", "k_tend_v", "k_tend_w", "k_tend_t", "k_small_prep", "k_advance_uv",
  "k_ac_div", "k_small_finish", "k_adv_scalar", "k_update_scalar", "k_p_rho_phi", "k_halo", "k_p_rho0",
  "k_coef_w", "k_ww_scan", "k_ac_mu_ww", "k_ac_theta", "k_ac_wsolve", "k_set_w_surface", "k_pd_scale", "k_pd_tend",
  "k_update_scalar_pd", "k_diff6", "k_rayleigh", "k_diff2", "k_lbc"};
struct EvPair { int id; hipEvent_t a, b; };
static unsigned g_timing = 0;  // bit n set: time launches of kernel n
static std::vector<EvPair> g_ev_pending;
static std::vector<hipEvent_t> g_ev_free;
static long long g_t_count[KID_COUNT];
static double g_t_ms[KID_COUNT];

static hipEvent_t ev_get() {
  if (!g_ev_free.empty()) {
    hipEvent_t e = g_ev_free.back();
    g_ev_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// The two compiled forms of a state kernel: one domain, and an ensemble launch.
struct KRef {
  void (*one)(KArgs);
  void (*batch)(BatchArgs);
};
#define KR(K) KRef{K<KArgs>, K<BatchArgs>}
#define KRP(K, P) KRef{K<KArgs, P>, K<BatchArgs, P>}
#define KRZ(K) KRef{K<KArgs, false, true>, K<BatchArgs, false, true>}
#define KRC(K, P, Q) KRef{K<KArgs, P, Q>, K<BatchArgs, P, Q>}
#define KRWS(CT, PF) KRef{k_ac_wsolve<KArgs, CT, true, PF>, k_ac_wsolve<BatchArgs, CT, true, PF>}

// One launch of an ensemble member, as recorded for a batched launch (wrfmi_ensemble_*).
struct RecOp {
  int id;
  bool halo;
  void *kern;      // the BatchArgs / BatchHalo form
  dim3 blk;
  unsigned wgs;    // workgroups of this member
  size_t arg;      // index into RecMember::k or ::h
};
struct RecMember {
  std::vector<RecOp> ops;
  std::vector<KArgs> k;
  std::vector<HaloArgs> h;
};

// Launch shapes: rows per workgroup of the four-point and column kernels, levels per thread
// of the four-point advance_uv and column divergence, and the column solve's prefetch depth.
// The autotuner (wrf_amd/autotune.py) picks them per kernel and domain shape; every shape
// computes the same values in the same order (only the assignment of points to threads and
// the distance loads run ahead change), so the results do not depend on them.
struct Shape {
  int rows, levels, prefetch;
};
static Shape default_shape(int kid);

struct Launcher {
  hipStream_t st;
  KArgs a;
  HaloArgs h;
  Shape sh[KID_COUNT];
  int err = 0;
  RecMember *rec = nullptr;  // set: record the launches instead of enqueueing them
  bool vec = false;          // the layout allows the four-points-per-lane kernels
  bool scal_all = false;     // species advected and updated by the all-species kernels
  bool zone_ok = false;      // each open direction is wide enough for a boundary zone on each side
  bool whole_step = false;   // enqueued by the whole-step driver (work arrays stay internal)
  bool lbc_on = false;       // specified boundaries: lbc holds the boundary arrays' device addresses
  wrfmi_lbc lbc;
  // launch with an optional event pair around it (kernel timing mode)
  template <typename K, typename... A>
  void launch(int id, K kern, dim3 g, dim3 b, A... args) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (g_timing & (1u << id)) {
      e0 = ev_get();
      e1 = ev_get();
      if (e0 && e1) (void)hipEventRecord(e0, st);
    }
    hipLaunchKernelGGL(kern, g, b, 0, st, args...);
    if (e0 && e1) {
      (void)hipEventRecord(e1, st);
      g_ev_pending.push_back(EvPair{id, e0, e1});
    }
  }
  // a state kernel on a tnx x tny x tnz tile grid, launched 1-D for the XCD-aware order
  void launchK(int id, KRef kern, dim3 g, dim3 b, KArgs args) {
    args.tnx = (int)g.x;
    args.tny = (int)g.y;
    args.tnz = (int)g.z;
    const unsigned n = g.x * g.y * g.z;
    if (rec) {
      rec->ops.push_back(RecOp{id, false, (void *)kern.batch, b, n, rec->k.size()});
      rec->k.push_back(args);
      return;
    }
    launch(id, kern.one, dim3(n, 1, 1), b, args);
  }
  void halo_begin() { h.n = 0; }
  void halo_add(float *p, char kind, int is3d) {
    int k = kind == 'u' ? 1 : kind == 'v' ? 2 : kind == 'w' ? 3 : 0;
    h.f[h.n++] = HaloField{p, k, is3d};
    if (h.n == 16) halo_flush();
  }
  void halo_flush() {
    if (h.n == 0) return;
    const Dm &d = h.d;
    const long ring = (long)(d.jds - d.jms + d.jme - d.jde + 1) * d.ni +
                      (long)(d.jde - d.jds) * (d.ids - d.ims + d.ime - d.ide + 1);
    const long nthreads = ring * ((d.kde + kHaloK - 1) / kHaloK) * h.n;
    const unsigned wgs = (unsigned)((nthreads + 255) / 256);
    if (rec) {
      rec->ops.push_back(RecOp{KID_HALO, true, (void *)k_halo<BatchHalo>, dim3(256, 1, 1), wgs, rec->h.size()});
      rec->h.push_back(h);
    } else {
      launch(KID_HALO, k_halo<HaloArgs>, dim3(wgs, 1, 1), dim3(256, 1, 1), h);
    }
    h.n = 0;
  }
};

// ------------------------------------------------------------------ the launch-shape table
static Shape default_shape(int kid) {
  switch (kid) {
    case KID_TEND_T: return Shape{WRFMI_TENDT_ROWS, 1, 0};
    case KID_ADV_SCALAR: return Shape{WRFMI_ADVALL_ROWS, 1, 0};
    case KID_ADVANCE_UV: return Shape{WRFMI_VEC_ROWS, kAdvUvK, 0};
    case KID_AC_DIV: return Shape{WRFMI_VEC_ROWS, kDivK, 0};
    case KID_AC_WSOLVE: return Shape{WRFMI_COL_ROWS, 1, 4};
    case KID_WW_SCAN:
    case KID_AC_MU_WW:
    case KID_COEF_W: return Shape{WRFMI_COL_ROWS, 1, 0};
    default: return Shape{WRFMI_VEC_ROWS, 1, 0};
  }
}
// which of a kernel's shape parameters are tunable, and the values each may take
static bool tunable_rows(int kid) {
  return kid != KID_HALO && kid != KID_W_SURFACE && kid != KID_PD_SCALE && kid != KID_PD_TEND && kid != KID_UPDATE_SCALAR_PD &&
         kid != KID_DIFF6 && kid != KID_RAYLEIGH && kid != KID_DIFF2 && kid != KID_LBC;
}
static bool tunable_levels(int kid) { return kid == KID_ADVANCE_UV || kid == KID_AC_DIV; }
// the four-point stencil kernels built for workgroups of up to kTallBlk threads
static bool tall_rows(int kid) {
  return kid == KID_TEND_U || kid == KID_TEND_V || kid == KID_TEND_W || kid == KID_TEND_T || kid == KID_ADV_SCALAR;
}
static bool tunable_prefetch(int kid) { return kid == KID_AC_WSOLVE; }
struct TuneEntry {
  int kid, nx, ny, nz;  // domain shape in mass points; 0 matches any
  Shape s;
};
static std::vector<TuneEntry> g_tune;  // (g_mu) a later entry takes precedence
static Shape shape_for(int kid, const Dm &d) {
  Shape s = default_shape(kid);
  const int nx = d.ide - d.ids, ny = d.jde - d.jds, nz = d.kde - d.kds;
  for (const TuneEntry &e : g_tune)
    if (e.kid == kid && (!e.nx || e.nx == nx) && (!e.ny || e.ny == ny) && (!e.nz || e.nz == nz)) s = e.s;
  return s;
}
static int kernel_id(const char *name) {
  for (int n = 0; n < KID_COUNT; n++)
    if (name && !strcmp(name, kKernelNames[n])) return n;
  return -1;
}
static int tune_add(const char *kernel, int nx, int ny, int nz, int rows, int levels, int prefetch) {
  const int kid = kernel_id(kernel);
  if (kid < 0) return fail(WRFMI_E_BADOPT, "no kernel named '%s'", kernel ? kernel : "(null)");
  if (!tunable_rows(kid)) return fail(WRFMI_E_BADOPT, "%s has no tunable launch shape", kernel);
  if (tall_rows(kid) ? (rows != 1 && rows != 2 && rows != 4 && rows != 8) : (rows != 1 && rows != 2 && rows != 4))
    return fail(WRFMI_E_BADOPT, "%s: rows per workgroup %d (1, 2, 4%s)", kernel, rows, tall_rows(kid) ? " or 8" : "");
  const Shape def = default_shape(kid);
  if (tunable_levels(kid) ? (levels != 2 && levels != 4 && levels != 8) : levels != def.levels)
    return fail(WRFMI_E_BADOPT, "%s: levels per thread %d not available", kernel, levels);
  if (tunable_prefetch(kid) ? (prefetch != 2 && prefetch != 4 && prefetch != 8) : prefetch != def.prefetch)
    return fail(WRFMI_E_BADOPT, "%s: prefetch depth %d not available", kernel, prefetch);
  if (nx < 0 || ny < 0 || nz < 0) return fail(WRFMI_E_BADOPT, "negative domain shape");
  g_tune.push_back(TuneEntry{kid, nx, ny, nz, Shape{rows, levels, prefetch}});
  return 0;
}
// Lines "kernel nx ny nz rows levels prefetch" ('#' starts a comment).  The whole file is
// checked before any entry is kept.
static int tune_load(const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) return fail(WRFMI_E_BADOPT, "cannot open launch-shape table %s", path);
  std::vector<TuneEntry> saved = g_tune;
  char line[256];
  int lineno = 0, rc = 0;
  while (!rc && fgets(line, sizeof line, f)) {
    lineno++;
    char *hash = strchr(line, '#');
    if (hash) *hash = 0;
    char name[64];
    int nx, ny, nz, rows, levels, prefetch;
    const int n = sscanf(line, "%63s %d %d %d %d %d %d", name, &nx, &ny, &nz, &rows, &levels, &prefetch);
    if (n <= 0) continue;
    if (n != 7) {
      rc = fail(WRFMI_E_BADOPT, "%s:%d: expected 'kernel nx ny nz rows levels prefetch'", path, lineno);
      break;
    }
    if ((rc = tune_add(name, nx, ny, nz, rows, levels, prefetch))) {
      std::string msg = g_err;
      rc = fail(rc, "%s:%d: %s", path, lineno, msg.c_str());
    }
  }
  fclose(f);
  if (rc) g_tune = saved;
  return rc;
}

// ------------------------------------------------------------------ phases of one RK substep
// Each enqueues the kernels of one group of solve_em calls, then the halo fills the
// reference makes after that group.
static int num_sound_steps(const wrfmi_config &c, const wrfmi_state &s) {
  // solve_em.F:446-463
  if (c.time_step_sound != 0) return c.time_step_sound;
  float spacing = s.dx < s.dy ? s.dx : s.dy;
  int n = 2 * ((int)(300.f * s.dt / spacing - 0.01f) + 1);
  return n < 4 ? 4 : n;
}
static inline dim3 g_prep(const Dm &d) { return col_grid(d, d.ids - 1, d.jds - 1, kColBlk); }
// workgroup of the one-column-per-lane kernels: 64 columns by WRFMI_COL_ROWS rows.  One
// row per workgroup spreads the few waves of a column launch (2800 at 400x400) evenly over
// the CUs; at 400x400x80 the w solve takes 2.97 ms per step with 1 row, 3.12 with 2 and
// 3.05 with 4.
static const dim3 kColumnBlk(64, WRFMI_COL_ROWS, 1);
static inline dim3 g_col(const Launcher &L) { return col_grid(L.a.d, col_ia(L.a.d, L.rec != nullptr), L.a.d.jds, kColumnBlk); }
// the tuned forms: workgroup and grid of kernel `kid` with its launch shape
static inline dim3 cblk(const Launcher &L, int kid) { return dim3(64, (unsigned)L.sh[kid].rows, 1); }
static inline dim3 gcol(const Launcher &L, int kid) {
  return col_grid(L.a.d, col_ia(L.a.d, L.rec != nullptr), L.a.d.jds, cblk(L, kid));
}
// (i, j) tiles of kColBlk over i = ids..ide, j = jds..jde, and nk levels in z
static inline dim3 g_pt3(const Dm &d, int nk) {
  const dim3 g = col_grid(d, d.ids, d.jds, kColBlk);
  return dim3(g.x, g.y, nk);
}

// (i, j) tiles of four-point kernels covering the points from i = ibase and the rows from
// jds - jlow through jde, nk levels: each workgroup takes 256 points of `rows` rows
static inline dim3 g_vec(const KArgs &a, int ibase, int jlow, int rows, int nk) {
  const Dm &d = a.d;
  const int w = kVecTileI;
  return dim3((d.ide - ibase + w) / w, (d.jde - d.jds + 1 + jlow + rows - 1) / rows, nk);
}
static inline dim3 g_vec3(const KArgs &a, int nk) {
  return g_vec(a, a.d.ims + 4 * ((a.d.ids - a.d.ims) >> 2), 0, WRFMI_VEC_ROWS, nk);
}
static inline dim3 vblk(const Launcher &L, int kid) { return dim3(64, (unsigned)L.sh[kid].rows, 1); }
static inline dim3 gvec(const Launcher &L, int kid, int nk) { return g_vec(L.a, vec_ia0(L.a.d), 0, L.sh[kid].rows, nk); }

// whether species m takes boundary arrays (solve_em.F:2261-2262, 2808-2810)
static bool species_bcs(const wrfmi_config &c, const wrfmi_state &s, int m) {
  if (c.nested) return true;
  if (m == 0 && s.n_moist > 0) return true;
  return m < s.n_moist ? c.have_bcs_moist : c.have_bcs_scalar;
}
static void ph_step_prep(Launcher &L) {
  const KArgs &a = L.a;
  const wrfmi_state &s = a.s;
  const dim3 gp = g_prep(a.d);
  if (L.vec) {
    // one extra group of four below ids (mut at ids-1), rows from jds-1
    const int ia0 = a.d.ims + 4 * ((a.d.ids - a.d.ims) >> 2) - 4;
    L.launchK(KID_STEP_PREP, KR(k_step_prep_x4), g_vec(a, ia0, 1, L.sh[KID_STEP_PREP].rows, a.d.kde - a.d.kds + 1),
              vblk(L, KID_STEP_PREP), a);
  } else {
    L.launchK(KID_STEP_PREP, KR(k_step_prep), dim3(gp.x, gp.y, a.d.kde - a.d.kds + 1), kColBlk, a);
  }
  if (a.c.precision_mode == 1) L.launchK(KID_WW_SCAN, KRP(k_ww_scan, true), gcol(L, KID_WW_SCAN), cblk(L, KID_WW_SCAN), a);
  else L.launchK(KID_WW_SCAN, KRP(k_ww_scan, false), gcol(L, KID_WW_SCAN), cblk(L, KID_WW_SCAN), a);
  L.halo_begin();
  L.halo_add(s.ru, 'u', 1); L.halo_add(s.rv, 'v', 1); L.halo_add(s.rw, 'w', 1); L.halo_add(s.ww, 'w', 1);
  L.halo_add(s.php, 'w', 1); L.halo_add(s.alt, 't', 1); L.halo_add(s.p, 'p', 1);
  L.halo_add(s.muu, 'u', 0); L.halo_add(s.muv, 'v', 0); L.halo_add(s.mut, 't', 0);
  L.halo_add(s.al, 'p', 1); L.halo_add(s.ph_2, 'w', 1);
  L.halo_flush();
}
// the one-point zone form's grid: 256 zone points per workgroup, nkz level groups
static inline dim3 g_zone(const KArgs &a, int nkz) {
  const unsigned n = zone_count(a.d, a.c.periodic_x, a.c.periodic_y);
  return dim3((n + 255) / 256, 1, nkz);
}
// rk_tendency's forward-step terms at rk_step 1 (module_em.F:797-950), into the *_tendf
// arrays before rk_addtend_dry (fused into k_tend_*) reads them: diff_opt 1's second-order
// diffusion of u, v, w, theta, then 6th-order diffusion, then the Rayleigh damper
static void ph_mixing(Launcher &L) {
  const KArgs &a = L.a;
  const Dm &d = a.d;
  if (a.c.diff_opt == 1) {
    const dim3 g = g_pt3(d, d.kde - d.kds);
    L.launchK(KID_DIFF2, KRP(k_diff2, 0), g, kColBlk, a);
    L.launchK(KID_DIFF2, KRP(k_diff2, 1), g, kColBlk, a);
    L.launchK(KID_DIFF2, KRP(k_diff2, 2), g, kColBlk, a);
    L.launchK(KID_DIFF2, KRP(k_diff2, 3), g, kColBlk, a);
  }
  if (a.c.diff_6th_opt != 0) {
    const dim3 g = g_pt3(d, d.kde - d.kds);
    L.launchK(KID_DIFF6, KRP(k_diff6, 0), g, kColBlk, a);
    L.launchK(KID_DIFF6, KRP(k_diff6, 1), g, kColBlk, a);
    if (a.c.non_hydrostatic) L.launchK(KID_DIFF6, KRP(k_diff6, 2), g, kColBlk, a);
    L.launchK(KID_DIFF6, KRP(k_diff6, 3), g, kColBlk, a);
  }
  if (a.c.damp_opt == 2) L.launchK(KID_RAYLEIGH, KR(k_rayleigh), g_pt3(d, 1), kColBlk, a);
}
// rk_scalar_tend's diff_opt 1 and 6th-order diffusion of species a.species into its
// tendency at rk_step 1 (module_em.F:1380-1436)
static void ph_scalar_mixing(Launcher &L) {
  if (L.a.rk_step == 1 && L.a.c.diff_opt == 1)
    L.launchK(KID_DIFF2, KRP(k_diff2, 4), g_pt3(L.a.d, L.a.d.kde - L.a.d.kds), kColBlk, L.a);
  if (L.a.rk_step == 1 && L.a.c.diff_6th_opt != 0)
    L.launchK(KID_DIFF6, KRP(k_diff6, 4), g_pt3(L.a.d, L.a.d.kde - L.a.d.kds), kColBlk, L.a);
}
// k_lbc over the columns of the domain (see k_lbc.hip for the modes)
static LbcJob lbc_job(int var, float *dst, const float *src = nullptr, const float *tend = nullptr,
                      float *const *b = nullptr, float *const *bt = nullptr, int massw = 0) {
  LbcJob J;
  memset(&J, 0, sizeof J);
  for (int n = 0; n < 4; n++) {
    J.b[n] = b ? b[n] : nullptr;
    J.bt[n] = bt ? bt[n] : nullptr;
  }
  J.src = src;
  J.tend = tend;
  J.dst = dst;
  J.var = var;
  J.massw = massw;
  return J;
}
// one boundary launch over up to kLbcJobs fields: the whole extent (whole), else the frame
// of `width` rows; LBC_PH_PRHO is k_lbc_ph, one workgroup per spec-zone column
static void lbc_launch(Launcher &L, int mode, float dt, int whole, int width, std::initializer_list<LbcJob> jobs,
                       int zero_grad_w = 0) {
  LbcArgs A;
  memset(&A, 0, sizeof A);
  A.a = L.a;
  for (const LbcJob &J : jobs) A.job[A.njobs++] = J;
  memcpy(A.fcx, L.lbc.fcx, sizeof A.fcx);
  memcpy(A.gcx, L.lbc.gcx, sizeof A.gcx);
  A.dtbc = L.lbc.dtbc;
  A.dt = dt;
  A.mode = mode;
  A.whole = whole;
  A.width = width;
  A.zero_grad_w = zero_grad_w;
  const Dm &d = L.a.d;
  // the frame of the widest staggering bounds every field's frame
  const LbcRing r = lbc_ring(d.ids, d.ide, d.jds, d.jde, L.a.c.periodic_x, width);
  if (mode == LBC_PH_PRHO) {
    L.launch(KID_LBC, k_lbc_ph, dim3((unsigned)r.count, 1, 1), dim3(128, 1, 1), A);
    return;
  }
  const int nk = d.kde - d.kds + 1;
  if (whole)
    L.launch(KID_LBC, k_lbc, dim3((d.ide - d.ids + 64) / 64, (d.jde - d.jds + 4) / 4, (mode == LBC_ZERO ? 1 : nk) * A.njobs),
             dim3(64, 4, 1), A);
  else
    L.launch(KID_LBC, k_lbc, dim3((unsigned)((r.count + 255) / 256), nk, A.njobs), dim3(256, 1, 1), A);
}
// relax_bdy_dry at rk_step 1 (solve_em.F:937-962), into the *_save arrays that k_tend_*
// add to the forward tendencies (rk_addtend_dry) and, for mu, into mu_save (k_tend_t adds it
// to mu_tend): the save arrays are zeroed over the whole extent, as rk_tendency's zero_tend
// leaves them, then the relaxation rows are written
static void lbc_relax_dry(Launcher &L) {
  const wrfmi_state &s = L.a.s;
  wrfmi_lbc &B = L.lbc;
  const bool nest = L.a.c.nested;
  lbc_launch(L, LBC_ZERO, 0.f, 1, 0,
             {lbc_job('u', s.u_save), lbc_job('v', s.v_save), lbc_job('h', s.ph_save), lbc_job('t', s.t_save),
              lbc_job('m', s.mu_save), lbc_job('h', s.w_save)});
  lbc_launch(L, LBC_RELAX, 0.f, 0, L.a.c.relax_zone,
             {lbc_job('u', s.u_save, s.ru, nullptr, B.u_b, B.u_bt), lbc_job('v', s.v_save, s.rv, nullptr, B.v_b, B.v_bt),
              lbc_job('h', s.ph_save, s.ph_2, nullptr, B.ph_b, B.ph_bt, 1),
              lbc_job('t', s.t_save, s.t_2, nullptr, B.t_b, B.t_bt, 1), lbc_job('m', s.mu_save, s.mu_2, nullptr, B.mu_b, B.mu_bt)});
  if (nest) lbc_launch(L, LBC_RELAX, 0.f, 0, L.a.c.relax_zone, {lbc_job('h', s.w_save, s.w_2, nullptr, B.w_b, B.w_bt, 1)});
}
// spec_bdy_dry after rk_addtend_dry (solve_em.F:982-1008)
static void lbc_spec_dry(Launcher &L) {
  const wrfmi_state &s = L.a.s;
  wrfmi_lbc &B = L.lbc;
  const int sz = L.a.c.spec_zone;
  lbc_launch(L, LBC_SPEC_TEND, 0.f, 0, sz,
             {lbc_job('u', s.ru_tend, nullptr, nullptr, nullptr, B.u_bt), lbc_job('v', s.rv_tend, nullptr, nullptr, nullptr, B.v_bt),
              lbc_job('h', s.ph_tend, nullptr, nullptr, nullptr, B.ph_bt), lbc_job('t', s.t_tend, nullptr, nullptr, nullptr, B.t_bt),
              lbc_job('m', s.mu_tend, nullptr, nullptr, nullptr, B.mu_bt)});
  if (L.a.c.nested) lbc_launch(L, LBC_SPEC_TEND, 0.f, 0, sz, {lbc_job('h', s.rw_tend, nullptr, nullptr, nullptr, B.w_bt)});
}
static void ph_tendency(Launcher &L) {
  if (L.a.rk_step == 1) ph_mixing(L);
  if (L.lbc_on && L.a.rk_step == 1) lbc_relax_dry(L);
  const KArgs &a = L.a;
  const int nk = a.d.kde - a.d.kds + 1, nkt = (nk + kTendK - 1) / kTendK;
  const bool per = a.c.periodic_x && a.c.periodic_y;
  const dim3 gt = g_pt3(a.d, nkt);
  if (L.vec && (per || L.zone_ok)) {
    // four-point kernels; on a domain with an open direction they leave the boundary zone
    // to the one-point kernels' zone form
    L.launchK(KID_TEND_U, KR(k_tend_u_x4), gvec(L, KID_TEND_U, nk), vblk(L, KID_TEND_U), a);
    L.launchK(KID_TEND_V, KR(k_tend_v_x4), gvec(L, KID_TEND_V, nk), vblk(L, KID_TEND_V), a);
    L.launchK(KID_TEND_W, KR(k_tend_w_x4), gvec(L, KID_TEND_W, nk), vblk(L, KID_TEND_W), a);
    L.launchK(KID_TEND_T, KR(k_tend_t_x4), gvec(L, KID_TEND_T, nk), vblk(L, KID_TEND_T), a);
    if (!per) {
      const dim3 gz = g_zone(a, nk);  // one level per thread
      L.launchK(KID_TEND_U, KRZ(k_tend_u), gz, kColBlk, a);
      L.launchK(KID_TEND_V, KRZ(k_tend_v), gz, kColBlk, a);
      L.launchK(KID_TEND_W, KRZ(k_tend_w), gz, kColBlk, a);
      L.launchK(KID_TEND_T, KRZ(k_tend_t), gz, kColBlk, a);
    }
    if (L.lbc_on) lbc_spec_dry(L);
    return;
  }
  L.launchK(KID_TEND_U, per ? KRP(k_tend_u, true) : KRP(k_tend_u, false), gt, kColBlk, a);
  L.launchK(KID_TEND_V, per ? KRP(k_tend_v, true) : KRP(k_tend_v, false), gt, kColBlk, a);
  L.launchK(KID_TEND_W, per ? KRP(k_tend_w, true) : KRP(k_tend_w, false), gt, kColBlk, a);
  L.launchK(KID_TEND_T, per ? KRP(k_tend_t, true) : KRP(k_tend_t, false), gt, kColBlk, a);
  if (L.lbc_on) lbc_spec_dry(L);
}
static void ph_small_prep(Launcher &L) {
  const KArgs &a = L.a;
  const wrfmi_state &s = a.s;
  const int nk = a.d.kde - a.d.kds + 1;
  if (L.vec) {
    L.launchK(KID_SMALL_PREP, KR(k_small_prep_x4), gvec(L, KID_SMALL_PREP, nk), vblk(L, KID_SMALL_PREP), a);
    L.launchK(KID_P_RHO0, KR(k_p_rho0_x4), gvec(L, KID_P_RHO0, nk - 1), vblk(L, KID_P_RHO0), a);
  } else {
    L.launchK(KID_SMALL_PREP, KR(k_small_prep), g_pt3(a.d, nk), kColBlk, a);
    L.launchK(KID_P_RHO0, KR(k_p_rho0), g_pt3(a.d, nk - 1), kColBlk, a);
  }
  // the per-routine entry point stores calc_coef_w's a, alpha, gamma as the reference does;
  // the whole step leaves them to k_ac_wsolve, which forms them on its way up each column
  const bool p64 = a.c.precision_mode == 1;
  if (L.whole_step)
    L.launchK(KID_COEF_W, p64 ? KRC(k_coef_w, double, false) : KRC(k_coef_w, float, false), gcol(L, KID_COEF_W),
              cblk(L, KID_COEF_W), a);
  else
    L.launchK(KID_COEF_W, p64 ? KRC(k_coef_w, double, true) : KRC(k_coef_w, float, true), gcol(L, KID_COEF_W),
              cblk(L, KID_COEF_W), a);
  L.halo_begin();
  L.halo_add(s.ru_tend, 'u', 1); L.halo_add(s.rv_tend, 'v', 1); L.halo_add(s.ph_2, 'w', 1);
  L.halo_add(s.al, 'p', 1); L.halo_add(s.p, 'p', 1); L.halo_add(s.t_1, 'p', 1); L.halo_add(s.t_save, 't', 1);
  L.halo_add(s.mu_1, 't', 0); L.halo_add(s.mu_2, 't', 0); L.halo_add(s.mudf, 't', 0);
  L.halo_flush();
}
static void ph_advance_uv(Launcher &L) {
  const bool per = L.a.c.periodic_x && L.a.c.periodic_y;
  const int nkb = (L.a.d.kde - L.a.d.kds + kAdvUvK - 1) / kAdvUvK;
  if (L.vec) {
    const int ks = L.sh[KID_ADVANCE_UV].levels;
    const int nkv = (L.a.d.kde - L.a.d.kds + ks - 1) / ks;
    const KRef k = ks == 2 ? (per ? KRC(k_advance_uv_x4, true, 2) : KRC(k_advance_uv_x4, false, 2))
                   : ks == 8 ? (per ? KRC(k_advance_uv_x4, true, 8) : KRC(k_advance_uv_x4, false, 8))
                             : (per ? KRC(k_advance_uv_x4, true, 4) : KRC(k_advance_uv_x4, false, 4));
    L.launchK(KID_ADVANCE_UV, k, gvec(L, KID_ADVANCE_UV, nkv), vblk(L, KID_ADVANCE_UV), L.a);
  } else
    L.launchK(KID_ADVANCE_UV, per ? KRP(k_advance_uv, true) : KRP(k_advance_uv, false), g_pt3(L.a.d, nkb), kColBlk, L.a);
  if (L.lbc_on) {  // spec_bdyupdate of u and v (solve_em.F:1343-1365)
    const wrfmi_state &s = L.a.s;
    lbc_launch(L, LBC_UPDATE, L.a.dts, 0, L.a.c.spec_zone,
               {lbc_job('u', s.u_2, nullptr, s.ru_tend), lbc_job('v', s.v_2, nullptr, s.rv_tend)});
  }
}
static void ph_advance_mu_t_w(Launcher &L) {
  const KArgs &a = L.a;
  const wrfmi_state &s = a.s;
  const bool p64 = a.c.precision_mode == 1;
  const int nk = a.d.kde - a.d.kds + 1;
  if (L.vec) {
    const int ks = L.sh[KID_AC_DIV].levels;
    const KRef k = ks == 2 ? KRP(k_ac_div_x4, 2) : ks == 8 ? KRP(k_ac_div_x4, 8) : KRP(k_ac_div_x4, 4);
    L.launchK(KID_AC_DIV, k, gvec(L, KID_AC_DIV, (nk + ks - 1) / ks), vblk(L, KID_AC_DIV), a);
  }
  else L.launchK(KID_AC_DIV, KR(k_ac_div), g_pt3(a.d, nk), kColBlk, a);
  if (p64) L.launchK(KID_AC_MU_WW, KRP(k_ac_mu_ww, true), gcol(L, KID_AC_MU_WW), cblk(L, KID_AC_MU_WW), a);
  else L.launchK(KID_AC_MU_WW, KRP(k_ac_mu_ww, false), gcol(L, KID_AC_MU_WW), cblk(L, KID_AC_MU_WW), a);
  if (L.vec) L.launchK(KID_AC_THETA, KR(k_ac_theta_x4), gvec(L, KID_AC_THETA, nk), vblk(L, KID_AC_THETA), a);
  else L.launchK(KID_AC_THETA, KR(k_ac_theta), g_pt3(a.d, nk), kColBlk, a);
  if (L.lbc_on) {  // spec_bdyupdate of theta, mu and muts (solve_em.F:1459-1490)
    lbc_launch(L, LBC_UPDATE, a.dts, 0, a.c.spec_zone,
               {lbc_job('t', s.t_2, nullptr, s.t_tend), lbc_job('m', s.mu_2, nullptr, s.mu_tend),
                lbc_job('m', s.muts, nullptr, s.mu_tend)});
  }
  {
    const int pf = L.sh[KID_AC_WSOLVE].prefetch;
    KRef k;
    if (p64) k = pf == 2 ? KRWS(double, 2) : pf == 8 ? KRWS(double, 8) : KRWS(double, 4);
    else k = pf == 2 ? KRWS(float, 2) : pf == 8 ? KRWS(float, 8) : KRWS(float, 4);
    L.launchK(KID_AC_WSOLVE, k, gcol(L, KID_AC_WSOLVE), cblk(L, KID_AC_WSOLVE), a);
  }
  if (L.lbc_on && a.c.non_hydrostatic) {
    // spec_bdyupdate_ph, zero_grad_bdy of w (solve_em.F:1583-1620), and calc_p_rho on the
    // spec-zone columns k_ac_wsolve left out
    lbc_launch(L, LBC_PH_PRHO, a.dts, 0, a.c.spec_zone, {lbc_job('t', s.ph_2)}, a.c.specified);
    if (!a.c.specified) lbc_launch(L, LBC_UPDATE, a.dts, 0, a.c.spec_zone, {lbc_job('h', s.w_2, nullptr, s.rw_tend)});
  }
  L.halo_begin();
  L.halo_add(s.ph_2, 'w', 1); L.halo_add(s.al, 'p', 1); L.halo_add(s.p, 'p', 1);
  L.halo_add(s.muts, 't', 0); L.halo_add(s.mu_2, 't', 0); L.halo_add(s.mudf, 't', 0);
  L.halo_flush();
}
static void ph_small_finish(Launcher &L) {
  const KArgs &a = L.a;
  const wrfmi_state &s = a.s;
  if (L.vec)
    L.launchK(KID_SMALL_FINISH, KR(k_small_finish_x4), gvec(L, KID_SMALL_FINISH, a.d.kde - a.d.kds + 1),
              vblk(L, KID_SMALL_FINISH), a);
  else L.launchK(KID_SMALL_FINISH, KR(k_small_finish), g_pt3(a.d, a.d.kde - a.d.kds + 1), kColBlk, a);
  if (a.rk_step == 3) {
    L.halo_begin();
    L.halo_add(s.ru_m, 'u', 1); L.halo_add(s.rv_m, 'v', 1); L.halo_add(s.ww_m, 'w', 1);
    L.halo_add(s.mut, 't', 0); L.halo_add(s.muts, 't', 0);
    L.halo_flush();
  }
}
// species m's transport option: moist species follow moist_adv_opt, the others
// scalar_adv_opt; 1 is the positive-definite form on the last substep
static bool pd_species(const KArgs &a, int m) {
  return (m < a.s.n_moist ? a.c.moist_adv_opt : a.c.scalar_adv_opt) == 1;
}
static bool any_pd(const KArgs &a) {
  for (int m = 0; m < a.s.n_species; m++)
    if (pd_species(a, m)) return true;
  return false;
}
// solve_em.F:1850-1975: before the last substep's transport, rk_update_scalar_pd takes each
// positive-definite species' physics tendency into its time-n values (mu_1 as both masses,
// dt_rk, tendency zeroed), then those values get their halo.  a.dt_rk is the substep's dt.
static void ph_scalar_pd_prep(Launcher &L) {
  const wrfmi_state &s = L.a.s;
  const Dm &d = L.a.d;
  KArgs a = L.a;
  a.s.mu_2 = a.s.mu_1;  // solve_em passes mu_1 as both masses
  for (int m = 0; m < s.n_species; m++) {
    if (!pd_species(a, m)) continue;
    a.species = m;
    L.launchK(KID_UPDATE_SCALAR_PD, KR(k_update_scalar_pd), g_pt3(d, d.kde - d.kds), kColBlk, a);
  }
  L.halo_begin();
  for (int m = 0; m < s.n_species; m++)
    if (pd_species(L.a, m)) L.halo_add(s.q_old[m], 'p', 1);
  L.halo_flush();
}
// advect_scalar_pd of species L.a.species into advect_tend (module_em.F:1264-1280): the
// limiter's scales into the dvdxi work array, then the limited divergence.  A doubly
// periodic aligned domain takes the four-point forms, which read the halo cells' scales
// from the interior cells they copy (the phases before this one have filled every halo).
static void ph_scalar_tend_pd(Launcher &L) {
  const Dm &d = L.a.d;
  if (L.vec && L.a.c.periodic_x && L.a.c.periodic_y) {
    const int nks = (d.kde - d.kds + kPdVK - 1) / kPdVK, nkt = (d.kde - d.kds + 1 + kPdVK - 1) / kPdVK;
    L.launchK(KID_PD_SCALE, KR(k_pd_scale_x4), g_vec(L.a, vec_ia0(d), -1, WRFMI_VEC_ROWS, nks), kVecBlk, L.a);
    L.launchK(KID_PD_TEND, KR(k_pd_tend_x4), g_vec(L.a, vec_ia0(d), 0, WRFMI_VEC_ROWS, nkt), kVecBlk, L.a);
    return;
  }
  const int nkz = (d.kde - d.kds + kPdK - 1) / kPdK;
  const dim3 gs = col_grid(d, d.ids - 1, d.jds - 1, kColBlk);
  L.launchK(KID_PD_SCALE, KR(k_pd_scale), dim3(gs.x, gs.y, nkz), kColBlk, L.a);
  const int nkt = (d.kde - d.kds + 1 + kPdK - 1) / kPdK;
  L.launchK(KID_PD_TEND, KR(k_pd_tend), g_pt3(d, nkt), kColBlk, L.a);
}
// rk_scalar_tend's advect_scalar of species L.a.species into advect_tend (the
// positive-definite form on the last substep for a species that asks for it)
static void ph_scalar_tend(Launcher &L) {
  ph_scalar_mixing(L);
  if (L.a.rk_step == 3 && pd_species(L.a, L.a.species)) {
    ph_scalar_tend_pd(L);
    return;
  }
  const bool per = L.a.c.periodic_x && L.a.c.periodic_y;
  const Dm &d = L.a.d;
  const dim3 ga((d.ide - d.ids + 1 + kAdvTileI - 1) / kAdvTileI, (d.jde - d.jds + 1 + kAdvRows - 1) / kAdvRows,
                (d.kde - d.kds + 1 + kAdvK - 1) / kAdvK);
  if (L.vec && (per || L.zone_ok)) {
    const int ia0 = d.ims + 4 * ((d.ids - d.ims) >> 2);
    const dim3 gv = g_vec(L.a, ia0, 0, L.sh[KID_ADV_SCALAR].rows, (d.kde - d.kds + kAdvVK) / kAdvVK);
    L.launchK(KID_ADV_SCALAR, KR(k_adv_scalar_x4), gv, vblk(L, KID_ADV_SCALAR), L.a);
    if (!per) L.launchK(KID_ADV_SCALAR, KRZ(k_adv_scalar), g_zone(L.a, d.kde - d.kds + 1), dim3(64, kAdvRows, 1), L.a);
  } else {
    L.launchK(KID_ADV_SCALAR, per ? KRP(k_adv_scalar, true) : KRP(k_adv_scalar, false), ga, dim3(64, kAdvRows, 1), L.a);
  }
}
static void ph_scalar(Launcher &L) {
  ph_scalar_tend(L);
  const int m = L.a.species;
  const wrfmi_state &s = L.a.s;
  const bool bcs = L.lbc_on && species_bcs(L.a.c, s, m);
  if (bcs && L.a.rk_step == 1) {  // relax_bdy_scalar, spec_bdy_scalar (solve_em.F:2261-2296)
    lbc_launch(L, LBC_RELAX, 0.f, 0, L.a.c.relax_zone, {lbc_job('q', s.q_tend[m], s.q[m], nullptr, L.lbc.q_b[m], L.lbc.q_bt[m], 1)});
    lbc_launch(L, LBC_SPEC_TEND, 0.f, 0, L.a.c.spec_zone, {lbc_job('q', s.q_tend[m], nullptr, nullptr, nullptr, L.lbc.q_bt[m])});
  }
  if (L.vec)
    L.launchK(KID_UPDATE_SCALAR, KR(k_update_scalar_x4), gvec(L, KID_UPDATE_SCALAR, L.a.d.kde - L.a.d.kds),
              vblk(L, KID_UPDATE_SCALAR), L.a);
  else L.launchK(KID_UPDATE_SCALAR, KR(k_update_scalar), g_pt3(L.a.d, L.a.d.kde - L.a.d.kds), kColBlk, L.a);
  // flow_dep_bdy of a species without boundary arrays (solve_em.F:2342-2355, 2886-2930)
  if (L.a.c.specified && !bcs) lbc_launch(L, LBC_FLOW_DEP, 0.f, 0, L.a.c.spec_zone, {lbc_job('q', s.q[m])});
}
// every species' rk_scalar_tend and rk_update_scalar (species are independent: the order
// over species does not change any value)
static void ph_scalar_all(Launcher &L) {
  const Dm &d = L.a.d;
  L.launchK(KID_ADV_SCALAR, KR(k_adv_scalar_all_x4), gvec(L, KID_ADV_SCALAR, d.kde - d.kds + 1), vblk(L, KID_ADV_SCALAR), L.a);
  for (int m = 0; m < L.a.s.n_species; m++) {
    L.a.species = m;
    ph_scalar_mixing(L);
  }
  L.launchK(KID_UPDATE_SCALAR, KR(k_update_scalar_all_x4), gvec(L, KID_UPDATE_SCALAR, d.kde - d.kds),
            vblk(L, KID_UPDATE_SCALAR), L.a);
}
static void ph_p_rho_phi(Launcher &L) {
  if (L.vec)
    L.launchK(KID_P_RHO_PHI, KR(k_p_rho_phi_x4), gvec(L, KID_P_RHO_PHI, L.a.d.kde - L.a.d.kds), vblk(L, KID_P_RHO_PHI), L.a);
  else L.launchK(KID_P_RHO_PHI, KR(k_p_rho_phi), g_pt3(L.a.d, L.a.d.kde - L.a.d.kds), kColBlk, L.a);
}
static void ph_bc_dry_2(Launcher &L, int kind) {
  const wrfmi_state &s = L.a.s;
  L.halo_begin();
  if (kind == 0) {
    L.halo_add(s.u_2, 'u', 1); L.halo_add(s.v_2, 'v', 1); L.halo_add(s.w_2, 'w', 1);
    L.halo_add(s.t_2, 'p', 1); L.halo_add(s.ph_2, 'w', 1); L.halo_add(s.mu_2, 't', 0);
  } else {
    L.halo_add(s.u_1, 'u', 1); L.halo_add(s.u_2, 'u', 1); L.halo_add(s.v_1, 'v', 1); L.halo_add(s.v_2, 'v', 1);
    L.halo_add(s.w_1, 'w', 1); L.halo_add(s.w_2, 'w', 1); L.halo_add(s.t_1, 'p', 1); L.halo_add(s.t_2, 'p', 1);
    L.halo_add(s.ph_1, 'w', 1); L.halo_add(s.ph_2, 'w', 1); L.halo_add(s.mu_1, 't', 0); L.halo_add(s.mu_2, 't', 0);
  }
  for (int m = 0; m < s.n_species; m++) L.halo_add(s.q[m], 'p', 1);
  L.halo_flush();
}

// set_w_surface after the RK loop (solve_em.F:4734-4750): interior columns only, no fill
static void ph_w_surface(Launcher &L) {
  L.launchK(KID_W_SURFACE, KR(k_set_w_surface), g_pt3(L.a.d, 1), kColBlk, L.a);
}

// RK substep sizes of solve_em.F:601-615
static void set_substep(KArgs &a, int rk, int nsound) {
  const float dt = a.s.dt;
  a.rk_step = rk;
  a.dt = dt;
  if (rk == 1) { a.dt_rk = dt / 3.f; a.dts = a.dt_rk; a.nsmall = 1; }
  else if (rk == 2) { a.dt_rk = 0.5f * dt; a.dts = dt / (float)nsound; a.nsmall = nsound / 2; }
  else { a.dt_rk = dt; a.dts = dt / (float)nsound; a.nsmall = nsound; }
}

static void enqueue_step(Launcher &L) {
  KArgs &a = L.a;
  L.whole_step = true;
  const int nsound = num_sound_steps(a.c, a.s);
  for (int rk = 1; rk <= 3; rk++) {
    set_substep(a, rk, nsound);
    ph_step_prep(L);
    ph_tendency(L);
    ph_small_prep(L);
    for (int it = 1; it <= a.nsmall; it++) {
      a.iteration = it;
      ph_advance_uv(L);
      ph_advance_mu_t_w(L);
    }
    ph_small_finish(L);
    const bool pd = rk == 3 && any_pd(a);
    if (pd) ph_scalar_pd_prep(L);
    if (L.scal_all && !pd) {
      ph_scalar_all(L);
    } else {
      for (int m = 0; m < a.s.n_species; m++) {
        a.species = m;
        ph_scalar(L);
      }
    }
    ph_p_rho_phi(L);
    if (rk < 3) ph_bc_dry_2(L, 0);
  }
  // solve_em.F:4195 recomputes calc_p_rho_phi from inputs that have not changed since the
  // last substep's call (same values), then set_phys_bc_dry_2 and the species fills follow,
  // and the surface w is reset from the final u, v (solve_em.F:4740).
  ph_bc_dry_2(L, 1);
  ph_w_surface(L);
}

// The four-points-per-lane kernels read 16-byte groups starting at an i whose offset in the
// row is a multiple of four: rows must be a multiple of four floats long and every array
// 16-byte aligned.  Their windows reach up to 7 points past ide, which may run past the end
// of the row into the next one (inside the array: no kernel reads the last row j = jme); such
// values only feed points that are not stored.
static bool vec_layout(const wrfmi_state &s, const Dm &d, const KArgs &a) {
  if (d.ni % 4 != 0) return false;
  bool ok = true;
  auto al = [&](const void *p) { ok = ok && ((uintptr_t)p & 15) == 0; };
#define AL(n) al(s.n);
  WRFMI_FIELDS_3D(AL) WRFMI_FIELDS_2D(AL)
#undef AL
  for (int m = 0; m < s.n_species; m++) {
    al(s.q[m]);
    al(s.q_old[m]);
    al(s.q_tend[m]);
  }
  al(a.scr_dv.p);
  al(a.scr_rhs.p);
  al(a.scr_lo.p);
  return ok;
}

static inline size_t field_elems(const Dm &d) { return (size_t)d.sj * (size_t)(d.jme - d.jms + 1); }

// calc_coef_w's a, alpha and gamma hold 4-byte elements in precision mode 0 and 8-byte
// elements in mode 1 (the column solve writes gamma at that width): each must lie in a device
// allocation that reaches the end of the field at that width, or the kernels would write past
// it.  hipMemGetAddressRange gives the allocation that holds the pointer.
static int check_coef_arrays(const wrfmi_state &s, const wrfmi_config &c, const Dm &d) {
  const size_t need = field_elems(d) * (c.precision_mode == 1 ? 8 : 4);
  const void *arrs[3] = {s.a, s.alpha, s.gamma};
  const char *names[3] = {"a", "alpha", "gamma"};
  for (int n = 0; n < 3; n++) {
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)arrs[n]) != hipSuccess) {
      (void)hipGetLastError();
      return fail(WRFMI_E_POINTER, "%s is not in a device allocation", names[n]);
    }
    if ((uintptr_t)arrs[n] + need > (uintptr_t)base + size)
      return fail(WRFMI_E_POINTER, "%s: its allocation ends %zu bytes after it; precision_mode %d needs %zu (%s elements)",
                  names[n], (size_t)((uintptr_t)base + size - (uintptr_t)arrs[n]), c.precision_mode, need,
                  c.precision_mode == 1 ? "8-byte" : "4-byte");
  }
  return 0;
}

// Validate and translate the arguments into L (g_mu held).  The work arrays are this
// domain's own (get_scratch) unless own_scratch is given: ensemble members each have theirs,
// and the allocation-free entry point passes the static array.  With own_scratch and no
// own_qt, the species are advected one by one (no species work arrays).
// the state's boundary arrays at their device addresses
static int translate_lbc(const wrfmi_state *st, const wrfmi_config &c, wrfmi_lbc *out) {
  if (!st->lbc) return fail(WRFMI_E_POINTER, "specified boundaries: the state's lbc (boundary arrays) is null");
  *out = *st->lbc;
  auto tr = [&](float *&p, const char *name, int side) -> int {
    p = p ? (float *)translate((void *)p) : nullptr;
    if (!p) return fail(WRFMI_E_POINTER, "specified boundaries: %s[%d] is null or not device memory", name, side);
    return 0;
  };
  int rc;
  for (int sd = 0; sd < 4; sd++) {
#define TRL(n) if ((rc = tr(out->n##_b[sd], #n "_b", sd)) || (rc = tr(out->n##_bt[sd], #n "_bt", sd))) return rc;
    TRL(u) TRL(v) TRL(ph) TRL(t) TRL(mu)
    if (c.nested) { TRL(w) }
#undef TRL
    for (int m = 0; m < st->n_species; m++)
      if (species_bcs(c, *st, m))
        if ((rc = tr(out->q_b[m][sd], "q_b", sd)) || (rc = tr(out->q_bt[m][sd], "q_bt", sd))) return rc;
  }
  return 0;
}

static int prepare(const wrfmi_state *st_in, const wrfmi_config *cfg, const wrfmi_dims *dims, Launcher &L, wrfmi_state &s,
                   const Scratch *own_scratch = nullptr, float *own_qt = nullptr, bool default_shapes = false) {
  int rc;
  if ((rc = check_cfg(cfg))) return rc;
  Dm d;
  if ((rc = make_dm(dims, &d))) return rc;
  if ((rc = translate_state(st_in, &s))) return rc;
  memset(&L.a, 0, sizeof(L.a));
  L.a.s = s;
  L.a.c = *cfg;
  L.a.d = d;
  L.a.dt = s.dt;  // grid%dt, as rk_tendency receives it (the whole step sets it per substep too)
  if ((rc = check_coef_arrays(s, *cfg, d))) return rc;
  const DomainKey key((uintptr_t)s.u_2, field_elems(d));
  Scratch sc;
  if (own_scratch) sc = *own_scratch;
  else if ((rc = get_scratch(key, &sc))) return rc;
  L.a.scr_dv = FPtr<float>(sc.dv);
  L.a.scr_rhs = FPtr<float>(sc.rhs);
  L.a.scr_lo = FPtr<float>(sc.lo);
  L.h.d = d;
  L.h.periodic_x = cfg->periodic_x; L.h.periodic_y = cfg->periodic_y;
  // set_physical_bc3d/2d fill a specified side as an open one (share/module_bc.F:217-242)
  const int spx = cfg->specified && !cfg->periodic_x, spy = cfg->specified;
  L.h.open_xs = cfg->open_xs || spx; L.h.open_xe = cfg->open_xe || spx;
  L.h.open_ys = cfg->open_ys || spy; L.h.open_ye = cfg->open_ye || spy;
  L.lbc_on = cfg->specified || cfg->nested;
  if (L.lbc_on && d.kde > kLbcMaxK)
    return fail(WRFMI_E_DIMS, "specified boundaries: at most %d levels (kde %d)", kLbcMaxK, d.kde);
  if (L.lbc_on && (rc = translate_lbc(st_in, *cfg, &L.lbc))) return rc;
  // ensemble members are recorded with the default shapes (one launch sequence for all)
  for (int k = 0; k < KID_COUNT; k++) L.sh[k] = default_shapes ? default_shape(k) : shape_for(k, d);
  L.vec = vec_layout(s, d, L.a);
  // the all-species scalar kernels (doubly periodic, aligned layout, more than one species)
  L.scal_all = L.vec && cfg->periodic_x && cfg->periodic_y && s.n_species > 1 && (own_qt || !own_scratch);
  L.zone_ok = (cfg->periodic_x || d.ide - d.ids + 1 >= 2 * kZone + 2) && (cfg->periodic_y || d.jde - d.jds + 1 >= 2 * kZone + 2);
  // specified boundaries: the four-point kernels carry the spec-zone bounds; the boundary
  // zone's one-point kernels take the relaxation rows (added to the tendencies), so those
  // must lie inside it.  With a periodic x, rhs_ph still narrows its x range near the x
  // edges (module_big_step_utilities_em.F:1937-1938, 2023-2048 test `specified` alone), which only the
  // one-point k_tend_t has: the tendencies then run one-point over the whole domain.
  if (L.lbc_on) {
    L.scal_all = false;
    if (cfg->relax_zone > kZone || cfg->periodic_x) L.zone_ok = false;
  }
  if (L.scal_all) {
    float *qt = own_qt;
    if (!qt && (rc = get_qt(key, s.n_species, &qt))) return rc;
    L.a.scr_qt = qt;
    L.a.qt_stride = field_elems(d);
  }
  return 0;
}

// solve_em.F:377: the step starts by advancing grid%dtbc by dt (the host's boundary struct)
static void lbc_begin_step(Launcher &L, wrfmi_state *st) {
  if (!L.lbc_on) return;
  st->lbc->dtbc = st->lbc->dtbc + st->dt;
  L.lbc.dtbc = st->lbc->dtbc;
}

extern "C" int wrfmi_solve_em_step(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *dims, void *stream, int sync) {
  if (!st || !cfg || !dims) return fail(WRFMI_E_POINTER, "null argument");
  Launcher L;
  wrfmi_state s;
  int rc;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = prepare(st, cfg, dims, L, s))) return rc;
    lbc_begin_step(L, st);
  }
  L.st = (hipStream_t)stream;
  enqueue_step(L);
  HIPCHK(hipGetLastError());
  if (sync) HIPCHK(hipStreamSynchronize(L.st));
  return 0;
}

// ------------------------------------------------------------------ allocation-free step
// The whole step with its three work arrays (dvdxi, rhs, and the FP64 low part of the
// precision form) in `scratch`, a static device array of three fields of at most
// kNoAllocElems elements each (48 MiB on every GPU the library runs on, held until the
// process ends), so that no call allocates.  In host code `scratch` names a host-side
// stand-in: its device address comes from hipGetSymbolAddress, once per device (at
// wrfmi_init, or at the entry point's first call on a device), and only that address is
// passed to kernels.  Nothing in it needs to start at zero: every kernel writes its part of
// a work array before it or a later kernel reads that part.  The species are advected one
// by one, so no species work arrays are needed.  Calls on one GPU share the array, so calls
// that overlap in time on one GPU (two streams, two threads) race.
constexpr size_t kNoAllocElems = (size_t)1 << 22;
static __device__ __align__(256) float scratch[3 * kNoAllocElems];
constexpr int kMaxDevices = 64;
static float *g_scratch_dev[kMaxDevices];  // device address of `scratch` per device (g_mu)

// device address of `scratch` on the current device (g_mu held)
static int noalloc_scratch(float **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(WRFMI_E_DEVICE, "device %d beyond the %d the library supports", dev, kMaxDevices);
  if (!g_scratch_dev[dev]) {
    float *p = nullptr;
    HIPCHK(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(WRFMI_E_DEVICE, "no device address for the static work array");
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

extern "C" int wrfmi_solve_em_step_noalloc(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *dims, void *stream,
                                           int sync) {
  if (!st || !cfg || !dims) return fail(WRFMI_E_POINTER, "null argument");
  Launcher L;
  wrfmi_state s;
  int rc;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_timing) return fail(WRFMI_E_BADOPT, "kernel timing mode creates events: not available without allocation");
    Dm d;
    if ((rc = make_dm(dims, &d))) return rc;
    if (field_elems(d) > kNoAllocElems)
      return fail(WRFMI_E_DIMS, "memory dims hold %zu elements per field; the allocation-free step takes at most %zu",
                  field_elems(d), kNoAllocElems);
    float *p = nullptr;
    if ((rc = noalloc_scratch(&p))) return rc;
    const Scratch sc{p, p + kNoAllocElems, p + 2 * kNoAllocElems};
    if ((rc = prepare(st, cfg, dims, L, s, &sc, nullptr))) return rc;
    lbc_begin_step(L, st);
  }
  L.st = (hipStream_t)stream;
  enqueue_step(L);
  HIPCHK(hipGetLastError());
  if (sync) HIPCHK(hipStreamSynchronize(L.st));
  return 0;
}

// ------------------------------------------------------------------ hipGraph replay
struct GraphEntry {
  wrfmi_state s;
  wrfmi_config c;
  wrfmi_dims d;
  hipGraph_t g;
  hipGraphExec_t x;
};
static void free_graphs() {
  for (GraphEntry *e : g_graphs) {
    (void)hipGraphExecDestroy(e->x);
    (void)hipGraphDestroy(e->g);
    delete e;
  }
  g_graphs.clear();
}

extern "C" int wrfmi_solve_em_step_graph(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *dims, void *stream, int sync) {
  if (!st || !cfg || !dims) return fail(WRFMI_E_POINTER, "null argument");
  if (cfg->specified || cfg->nested)
    return fail(WRFMI_E_BADOPT, "specified boundaries: dtbc changes every step, so the step is not replayed as a graph "
                "(use wrfmi_solve_em_step)");
  if (g_timing) return fail(WRFMI_E_BADOPT, "kernel timing mode records events on eager launches only");
  hipStream_t hs = (hipStream_t)stream;
  GraphEntry *hit = nullptr;
  {
    // graphs are cached under the translated (device) pointers: a registered host array
    // whose mirror was replaced no longer matches the graph recorded on the old mirror
    std::lock_guard<std::mutex> lk(g_mu);
    wrfmi_state tr;
    int rc;
    if ((rc = translate_state(st, &tr))) return rc;
    for (GraphEntry *e : g_graphs)
      if (!memcmp(&e->s, &tr, sizeof(tr)) && !memcmp(&e->c, cfg, sizeof(*cfg)) && !memcmp(&e->d, dims, sizeof(*dims))) hit = e;
  }
  if (!hit) {
    Launcher L;
    wrfmi_state s;
    int rc;
    {
      std::lock_guard<std::mutex> lk(g_mu);
      if ((rc = prepare(st, cfg, dims, L, s))) return rc;
      if (!g_stream) HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    }
    hipStream_t cs = g_stream;
    L.st = cs;
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    enqueue_step(L);
    HIPCHK(hipStreamEndCapture(cs, &g));
    hipGraphExec_t x;
    HIPCHK(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
    hit = new GraphEntry{s, *cfg, *dims, g, x};
    std::lock_guard<std::mutex> lk(g_mu);
    g_graphs.push_back(hit);
  }
  HIPCHK(hipGraphLaunch(hit->x, hs));
  if (sync) HIPCHK(hipStreamSynchronize(hs));
  return 0;
}

// ------------------------------------------------------------------ batched ensembles
// An ensemble advances many independent domains (members, of any sizes) with one launch per
// kernel of the step: the launches of every member are recorded once, at creation, and each
// group of corresponding launches becomes one launch over the concatenated tiles of all
// members.  The per-member arguments, the workgroup offsets and the workgroup -> member table
// live in one device allocation that the kernels read (Batch<T> in dev_common.h).
struct EnsOp {
  int id;
  bool halo;
  void *kern;
  dim3 blk;
  unsigned total;
  BatchArgs ka;
  BatchHalo ha;
};
struct wrfmi_ensemble {
  int n = 0;
  std::vector<EnsOp> ops;
  std::vector<float *> work;  // per-member work arrays
  void *table = nullptr;      // device: member arguments, offsets, tables
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  std::vector<wrfmi_state> members;  // translated (device) pointers of every member
  bool stale = false;         // a registered array a member uses was re-registered or unregistered
};
static void ens_free(wrfmi_ensemble *e) {
  if (e->x) (void)hipGraphExecDestroy(e->x);
  if (e->g) (void)hipGraphDestroy(e->g);
  for (float *p : e->work) (void)hipFree(p);
  if (e->table) (void)hipFree(e->table);
  delete e;
}
static std::vector<wrfmi_ensemble *> g_ensembles;
static void free_ensembles() {
  for (wrfmi_ensemble *e : g_ensembles) ens_free(e);
  g_ensembles.clear();
}

extern "C" int wrfmi_ensemble_create(int n_members, const wrfmi_state *states, const wrfmi_config *cfgs,
                                     const wrfmi_dims *dims, wrfmi_ensemble **out) {
  if (!states || !cfgs || !dims || !out) return fail(WRFMI_E_POINTER, "null argument");
  if (n_members < 1) return fail(WRFMI_E_BADOPT, "an ensemble needs at least one member");
  for (int m = 0; m < n_members; m++)
    if (cfgs[m].specified || cfgs[m].nested)
      return fail(WRFMI_E_BADOPT, "member %d: specified boundaries are stepped one domain at a time (wrfmi_solve_em_step)", m);
  *out = nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  wrfmi_ensemble *e = new wrfmi_ensemble;
  e->n = n_members;
  std::vector<RecMember> rec(n_members);
  std::vector<Launcher> Ls(n_members);
  int rc = 0;
  for (int m = 0; m < n_members && !rc; m++) {
    Dm d;
    if ((rc = make_dm(&dims[m], &d))) break;
    const size_t bytes = (size_t)d.sj * (size_t)(d.jme - d.jms + 1) * sizeof(float);
    Scratch sc{nullptr, nullptr, nullptr};
    for (float **p : {&sc.dv, &sc.rhs, &sc.lo}) {
      if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        rc = fail(WRFMI_E_DEVICE, "member %d: cannot allocate %zu bytes of work arrays", m, bytes);
        break;
      }
      e->work.push_back(*p);
    }
    if (rc) break;
    float *qt = nullptr;  // species work arrays of a doubly periodic member with several species
    if (cfgs[m].periodic_x && cfgs[m].periodic_y && states[m].n_species > 1) {
      const size_t qb = bytes * (size_t)(states[m].n_species - 1);
      if (hipMalloc(&qt, qb) != hipSuccess) {
        (void)hipGetLastError();
        rc = fail(WRFMI_E_DEVICE, "member %d: cannot allocate %zu bytes of species work arrays", m, qb);
        break;
      }
      e->work.push_back(qt);
    }
    wrfmi_state s;
    if ((rc = prepare(&states[m], &cfgs[m], &dims[m], Ls[m], s, &sc, qt, true))) {
      std::string msg = g_err;
      rc = fail(rc, "member %d: %s", m, msg.c_str());
      break;
    }
    e->members.push_back(s);
  }
  // one kernel form for all members: the four-points-per-lane kernels only if every
  // member's layout allows them
  bool all_vec = true;
  for (int m = 0; m < n_members && !rc; m++) all_vec = all_vec && Ls[m].vec;
  bool all_zone = true;
  for (int m = 0; m < n_members && !rc; m++) all_zone = all_zone && Ls[m].zone_ok;
  for (int m = 0; m < n_members && !rc; m++) {
    Ls[m].vec = all_vec;
    Ls[m].scal_all = Ls[m].scal_all && all_vec;
    Ls[m].zone_ok = all_zone;
    Ls[m].rec = &rec[m];
    enqueue_step(Ls[m]);
  }
  // every member must have produced the same launch sequence (same kernels and blocks):
  // members share the boundary type, precision mode, species count, species transport
  // options and acoustic steps
  for (int m = 1; m < n_members && !rc; m++) {
    bool same = rec[m].ops.size() == rec[0].ops.size();
    for (size_t o = 0; same && o < rec[0].ops.size(); o++)
      same = rec[m].ops[o].kern == rec[0].ops[o].kern && rec[m].ops[o].blk.x == rec[0].ops[o].blk.x &&
             rec[m].ops[o].blk.y == rec[0].ops[o].blk.y && rec[m].ops[o].halo == rec[0].ops[o].halo &&
             (!rec[m].ops[o].halo || rec[m].h[rec[m].ops[o].arg].n == rec[0].h[rec[0].ops[o].arg].n);
    if (!same)
      rc = fail(WRFMI_E_BADOPT, "member %d: members must share the boundary type, precision mode, number of species, "
                "species transport options and number of acoustic steps", m);
  }
  if (rc) {
    ens_free(e);
    return rc;
  }
  // host image of the device table
  std::vector<char> img;
  auto put = [&](const void *p, size_t n) {
    size_t at = (img.size() + 255) & ~(size_t)255;
    img.resize(at + n);
    memcpy(img.data() + at, p, n);
    return at;
  };
  std::map<std::vector<unsigned>, std::pair<size_t, size_t>> shapes;  // member wgs -> (off, lut)
  struct Pending { size_t args, off, lut; };
  std::vector<Pending> pend;
  const size_t nops = rec[0].ops.size();
  for (size_t o = 0; o < nops; o++) {
    const RecOp &op0 = rec[0].ops[o];
    std::vector<unsigned> w(n_members);
    for (int m = 0; m < n_members; m++) w[m] = rec[m].ops[o].wgs;
    auto it = shapes.find(w);
    if (it == shapes.end()) {
      std::vector<unsigned> off(n_members + 1, 0), lut;
      for (int m = 0; m < n_members; m++) {
        off[m + 1] = off[m] + w[m];
        lut.insert(lut.end(), w[m], (unsigned)m);
      }
      const size_t oo = put(off.data(), off.size() * sizeof(unsigned));
      const size_t lo = put(lut.data(), lut.size() * sizeof(unsigned));
      it = shapes.emplace(w, std::make_pair(oo, lo)).first;
    }
    size_t ao;
    if (op0.halo) {
      std::vector<HaloArgs> hv(n_members);
      for (int m = 0; m < n_members; m++) hv[m] = rec[m].h[rec[m].ops[o].arg];
      ao = put(hv.data(), hv.size() * sizeof(HaloArgs));
    } else {
      std::vector<KArgs> kv(n_members);
      for (int m = 0; m < n_members; m++) kv[m] = rec[m].k[rec[m].ops[o].arg];
      ao = put(kv.data(), kv.size() * sizeof(KArgs));
    }
    unsigned total = 0;
    for (unsigned v : w) total += v;
    EnsOp eo{op0.id, op0.halo, op0.kern, op0.blk, total, {}, {}};
    e->ops.push_back(eo);
    pend.push_back(Pending{ao, it->second.first, it->second.second});
  }
  if (hipMalloc(&e->table, img.size()) != hipSuccess || hipMemcpy(e->table, img.data(), img.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    ens_free(e);
    return fail(WRFMI_E_DEVICE, "cannot set up the ensemble table (%zu bytes)", img.size());
  }
  char *base = (char *)e->table;
  for (size_t o = 0; o < nops; o++) {
    EnsOp &eo = e->ops[o];
    auto off = (const WRFMI_CONST_AS unsigned *)(base + pend[o].off);
    auto lut = (const WRFMI_CONST_AS unsigned *)(base + pend[o].lut);
    if (eo.halo) eo.ha = BatchHalo{(const WRFMI_CONST_AS HaloArgs *)(base + pend[o].args), off, lut, n_members, eo.total};
    else eo.ka = BatchArgs{(const WRFMI_CONST_AS KArgs *)(base + pend[o].args), off, lut, n_members, eo.total};
  }
  g_ensembles.push_back(e);
  *out = e;
  return 0;
}

static void ens_enqueue(wrfmi_ensemble *e, hipStream_t st) {
  Launcher L;
  L.st = st;
  for (const EnsOp &o : e->ops) {
    if (o.halo) L.launch(o.id, (void (*)(BatchHalo))o.kern, dim3(o.total, 1, 1), o.blk, o.ha);
    else L.launch(o.id, (void (*)(BatchArgs))o.kern, dim3(o.total, 1, 1), o.blk, o.ka);
  }
}
static bool ens_known(wrfmi_ensemble *e) {
  for (wrfmi_ensemble *k : g_ensembles)
    if (k == e) return true;
  return false;
}

extern "C" int wrfmi_ensemble_step(wrfmi_ensemble *e, void *stream, int sync) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!e || !ens_known(e)) return fail(WRFMI_E_POINTER, "not an ensemble of this library");
    if (e->stale) return fail(WRFMI_E_POINTER, "a registered array of this ensemble was replaced or unregistered");
  }
  ens_enqueue(e, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  if (sync) HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}

extern "C" int wrfmi_ensemble_step_graph(wrfmi_ensemble *e, void *stream, int sync) {
  if (g_timing) return fail(WRFMI_E_BADOPT, "kernel timing mode records events on eager launches only");
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!e || !ens_known(e)) return fail(WRFMI_E_POINTER, "not an ensemble of this library");
    if (e->stale) return fail(WRFMI_E_POINTER, "a registered array of this ensemble was replaced or unregistered");
    if (!e->x) {
      if (!g_stream) HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
      HIPCHK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
      ens_enqueue(e, g_stream);
      HIPCHK(hipStreamEndCapture(g_stream, &e->g));
      HIPCHK(hipGraphInstantiate(&e->x, e->g, nullptr, nullptr, 0));
    }
  }
  HIPCHK(hipGraphLaunch(e->x, (hipStream_t)stream));
  if (sync) HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}

extern "C" int wrfmi_ensemble_destroy(wrfmi_ensemble *e) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (size_t n = 0; n < g_ensembles.size(); n++)
    if (g_ensembles[n] == e) {
      HIPCHK(hipDeviceSynchronize());
      ens_free(e);
      g_ensembles.erase(g_ensembles.begin() + n);
      return 0;
    }
  return fail(WRFMI_E_POINTER, "not an ensemble of this library");
}

static bool state_touches(const wrfmi_state &s, uintptr_t lo, uintptr_t hi) {
  auto in = [&](const void *p) { return (uintptr_t)p >= lo && (uintptr_t)p < hi; };
  bool t = in(s.a) || in(s.alpha) || in(s.gamma);
#define TOUCH(n) t = t || in(s.n);
  WRFMI_FIELDS_3D(TOUCH) WRFMI_FIELDS_2D(TOUCH) WRFMI_FIELDS_1D(TOUCH)
#undef TOUCH
  for (int m = 0; m < s.n_species && m < WRFMI_MAX_SPECIES; m++) t = t || in(s.q[m]) || in(s.q_old[m]) || in(s.q_tend[m]);
  return t;
}
static void drop_device_range(uintptr_t lo, uintptr_t hi) {
  for (size_t n = 0; n < g_graphs.size();) {
    GraphEntry *e = g_graphs[n];
    if (state_touches(e->s, lo, hi)) {
      (void)hipGraphExecDestroy(e->x);
      (void)hipGraphDestroy(e->g);
      delete e;
      g_graphs.erase(g_graphs.begin() + n);
    } else {
      n++;
    }
  }
  free_scratch_in(lo, hi);
  for (wrfmi_ensemble *e : g_ensembles)
    for (const wrfmi_state &m : e->members) e->stale = e->stale || state_touches(m, lo, hi);
}

// Release what the library holds for one domain: its work arrays and its cached graphs.
extern "C" int wrfmi_domain_release(const wrfmi_state *st) {
  if (!st) return fail(WRFMI_E_POINTER, "null argument");
  std::lock_guard<std::mutex> lk(g_mu);
  const uintptr_t u2 = (uintptr_t)translate((void *)st->u_2);
  HIPCHK(hipDeviceSynchronize());
  for (size_t n = 0; n < g_graphs.size();) {
    GraphEntry *e = g_graphs[n];
    if ((uintptr_t)e->s.u_2 == u2) {
      (void)hipGraphExecDestroy(e->x);
      (void)hipGraphDestroy(e->g);
      delete e;
      g_graphs.erase(g_graphs.begin() + n);
    } else {
      n++;
    }
  }
  free_scratch_in(u2, u2 + 1);
  return 0;
}

// ------------------------------------------------------------------ per-routine entry points
// Common wrapper: validate, translate pointers, set the substep fields, enqueue one phase.
template <typename F>
static int run_phase(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *dims, void *stream, F &&body) {
  if (!st || !cfg || !dims) return fail(WRFMI_E_POINTER, "null argument");
  Launcher L;
  wrfmi_state s;
  int rc;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = prepare(st, cfg, dims, L, s))) return rc;
  }
  L.st = (hipStream_t)stream;
  if ((rc = body(L))) return rc;
  HIPCHK(hipGetLastError());
  return 0;
}
static int check_rk(int rk_step) {
  return (rk_step < 1 || rk_step > 3) ? fail(WRFMI_E_BADOPT, "rk_step=%d outside 1..3", rk_step) : 0;
}

extern "C" int wrfmi_rk_step_prep(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d, int rk_step, void *stream) {
  return run_phase(st, cfg, d, stream, [&](Launcher &L) {
    if (int rc = check_rk(rk_step)) return rc;
    L.a.rk_step = rk_step;
    ph_step_prep(L);
    return 0;
  });
}
extern "C" int wrfmi_rk_tendency(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d, int rk_step, void *stream) {
  return run_phase(st, cfg, d, stream, [&](Launcher &L) {
    if (int rc = check_rk(rk_step)) return rc;
    L.a.rk_step = rk_step;
    ph_tendency(L);
    return 0;
  });
}
extern "C" int wrfmi_small_step_prep(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d, int rk_step, float dts,
                                     void *stream) {
  return run_phase(st, cfg, d, stream, [&](Launcher &L) {
    if (int rc = check_rk(rk_step)) return rc;
    L.a.rk_step = rk_step;
    L.a.dts = dts;
    ph_small_prep(L);
    return 0;
  });
}
extern "C" int wrfmi_advance_uv_state(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d, float dts, void *stream) {
  return run_phase(st, cfg, d, stream, [&](Launcher &L) {
    L.a.dts = dts;
    ph_advance_uv(L);
    return 0;
  });
}
extern "C" int wrfmi_advance_mu_t_w(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d, float dts, int iteration,
                                    int n_small, void *stream) {
  return run_phase(st, cfg, d, stream, [&](Launcher &L) {
    if (iteration < 1 || iteration > n_small) return fail(WRFMI_E_BADOPT, "iteration %d outside 1..%d", iteration, n_small);
    L.a.dts = dts;
    L.a.iteration = iteration;
    L.a.nsmall = n_small;
    ph_advance_mu_t_w(L);
    return 0;
  });
}
extern "C" int wrfmi_small_step_finish(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d, int rk_step, int n_small,
                                       float dts, void *stream) {
  return run_phase(st, cfg, d, stream, [&](Launcher &L) {
    if (int rc = check_rk(rk_step)) return rc;
    L.a.rk_step = rk_step;
    L.a.nsmall = n_small;
    L.a.dts = dts;
    ph_small_finish(L);
    return 0;
  });
}
extern "C" int wrfmi_rk_scalar_step(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d, int ispec, int rk_step,
                                    float dt_rk, void *stream) {
  return run_phase(st, cfg, d, stream, [&](Launcher &L) {
    if (int rc = check_rk(rk_step)) return rc;
    if (ispec < 0 || ispec >= L.a.s.n_species) return fail(WRFMI_E_BADOPT, "species %d outside 0..%d", ispec, L.a.s.n_species - 1);
    L.a.rk_step = rk_step;
    L.a.species = ispec;
    L.a.dt_rk = dt_rk;
    ph_scalar(L);
    return 0;
  });
}
extern "C" int wrfmi_rk_scalar_tend(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d, int ispec, int rk_step,
                                    float dt_rk, void *stream) {
  return run_phase(st, cfg, d, stream, [&](Launcher &L) {
    if (int rc = check_rk(rk_step)) return rc;
    if (ispec < 0 || ispec >= L.a.s.n_species) return fail(WRFMI_E_BADOPT, "species %d outside 0..%d", ispec, L.a.s.n_species - 1);
    L.a.rk_step = rk_step;
    L.a.species = ispec;
    L.a.dt_rk = dt_rk;
    ph_scalar_tend(L);
    return 0;
  });
}
extern "C" int wrfmi_rk_scalar_pd_prep(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d, float dt_rk,
                                       void *stream) {
  return run_phase(st, cfg, d, stream, [&](Launcher &L) {
    L.a.dt_rk = dt_rk;
    ph_scalar_pd_prep(L);
    return 0;
  });
}
extern "C" int wrfmi_calc_p_rho_phi(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d, void *stream) {
  return run_phase(st, cfg, d, stream, [&](Launcher &L) {
    ph_p_rho_phi(L);
    return 0;
  });
}
extern "C" int wrfmi_phys_bc_dry_2(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d, int kind, void *stream) {
  return run_phase(st, cfg, d, stream, [&](Launcher &L) {
    if (kind != 0 && kind != 1) return fail(WRFMI_E_BADOPT, "kind must be 0 or 1");
    ph_bc_dry_2(L, kind);
    return 0;
  });
}

extern "C" int wrfmi_set_w_surface(wrfmi_state *st, const wrfmi_config *cfg, const wrfmi_dims *d, void *stream) {
  return run_phase(st, cfg, d, stream, [&](Launcher &L) {
    ph_w_surface(L);
    return 0;
  });
}

static int bc_single(float *field, char var, const wrfmi_config *cfg, const wrfmi_dims *dims, void *stream, int is3d) {
  if (!field || !cfg || !dims) return fail(WRFMI_E_POINTER, "null argument");
  if (var != 'u' && var != 'v' && var != 'w' && var != 't' && var != 'p')
    return fail(WRFMI_E_BADOPT, "staggering letter '%c' is not one of u v w t p", var);
  int rc;
  Launcher L;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = check_cfg(cfg))) return rc;
    if ((rc = make_dm(dims, &L.a.d))) return rc;
    field = (float *)translate(field);
  }
  L.h.d = L.a.d;
  L.h.periodic_x = cfg->periodic_x; L.h.periodic_y = cfg->periodic_y;
  L.h.open_xs = cfg->open_xs; L.h.open_xe = cfg->open_xe; L.h.open_ys = cfg->open_ys; L.h.open_ye = cfg->open_ye;
  L.st = (hipStream_t)stream;
  L.halo_begin();
  L.halo_add(field, var, is3d);
  L.halo_flush();
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int wrfmi_set_physical_bc3d(float *field, char var, const wrfmi_config *cfg, const wrfmi_dims *d, void *stream) {
  return bc_single(field, var, cfg, d, stream, 1);
}
extern "C" int wrfmi_set_physical_bc2d(float *field, char var, const wrfmi_config *cfg, const wrfmi_dims *d, void *stream) {
  return bc_single(field, var, cfg, d, stream, 0);
}

// ------------------------------------------------------------------ diagnostics
namespace wrfmi {
__global__ void k_count_nonfinite(const float *a, Dm d, int kend, int is3d, unsigned long long *cnt) {
  const int i = d.ids + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)blockIdx.y;
  if (i > d.ide - 1 || j > d.jde - 1) return;
  unsigned long long n = 0;
  if (is3d) {
    for (int k = d.kds; k <= kend; k++) {
      float v = a[X3i(d, i, k, j)];
      if (!isfinite(v)) n++;
    }
  } else if (!isfinite(a[X2i(d, i, j)])) {
    n = 1;
  }
  if (n) atomicAdd(cnt, n);
}
}  // namespace wrfmi

extern "C" int wrfmi_check_finite(const wrfmi_state *st_in, const wrfmi_dims *dims, long long *n_bad) {
  if (!st_in || !dims || !n_bad) return fail(WRFMI_E_POINTER, "null argument");
  int rc;
  Dm d;
  wrfmi_state s;
  // the scan waits for all work queued on the device (steps may sit on non-blocking
  // streams), and holds the library lock so that concurrent scans do not share the counter
  std::lock_guard<std::mutex> lk(g_mu);
  if ((rc = make_dm(dims, &d))) return rc;
  if ((rc = translate_state(st_in, &s))) return rc;
  if (!g_counter) HIPCHK(hipMalloc(&g_counter, sizeof(unsigned long long)));
  HIPCHK(hipDeviceSynchronize());
  unsigned long long zero = 0, cnt = 0;
  HIPCHK(hipMemcpy(g_counter, &zero, sizeof zero, hipMemcpyHostToDevice));
  dim3 blk(64, 1, 1), grd((d.ide - d.ids + 63) / 64, d.jde - d.jds, 1);
  const float *f3[] = {s.u_2, s.v_2, s.w_2, s.t_2, s.ph_2};
  const int ke[] = {d.kde - 1, d.kde - 1, d.kde, d.kde - 1, d.kde};
  for (int n = 0; n < 5; n++) hipLaunchKernelGGL(k_count_nonfinite, grd, blk, 0, 0, f3[n], d, ke[n], 1, g_counter);
  hipLaunchKernelGGL(k_count_nonfinite, grd, blk, 0, 0, (const float *)s.mu_2, d, 1, 0, g_counter);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(&cnt, g_counter, sizeof cnt, hipMemcpyDeviceToHost));
  *n_bad = (long long)cnt;
  return cnt ? fail(WRFMI_E_NONFINITE, "%llu non-finite values in the prognostic state", cnt) : 0;
}

// ------------------------------------------------------------------ domain totals (FP64)

extern "C" int wrfmi_domain_totals(const wrfmi_state *st_in, const wrfmi_config *cfg, const wrfmi_dims *dims,
                                   void *stream, double *out) {
  if (!st_in || !cfg || !dims || !out) return fail(WRFMI_E_POINTER, "null argument");
  std::lock_guard<std::mutex> lk(g_mu);
  int rc;
  Dm d;
  wrfmi_state s;
  if ((rc = make_dm(dims, &d))) return rc;
  if ((rc = translate_state(st_in, &s))) return rc;
  const int ncol = (d.ide - d.ids) * (d.jde - d.jds);
  const int nblocks = (ncol + 255) / 256;
  const size_t need = (size_t)(nblocks + 1) * kNTotals * sizeof(double);
  if (g_diag_bytes < need) {
    if (g_diag) HIPCHK(hipFree(g_diag));
    g_diag = nullptr;
    g_diag_bytes = 0;
    HIPCHK(hipMalloc(&g_diag, need));
    g_diag_bytes = need;
  }
  KArgs a;
  memset(&a, 0, sizeof a);
  a.s = s;
  a.c = *cfg;
  a.d = d;
  const hipStream_t hs = (hipStream_t)stream;
  hipLaunchKernelGGL(k_domain_totals, dim3(nblocks, 1, 1), dim3(256, 1, 1), 0, hs, a, g_diag + kNTotals);
  hipLaunchKernelGGL(k_domain_totals_final, dim3(1, 1, 1), dim3(256, 1, 1), 0, hs, (const double *)(g_diag + kNTotals),
                     nblocks, g_diag);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, g_diag, kNTotals * sizeof(double), hipMemcpyDeviceToHost, hs));
  HIPCHK(hipStreamSynchronize(hs));
  return 0;
}

// ------------------------------------------------------------------ launch-shape table entry points
extern "C" int wrfmi_tuning_set(const char *kernel, int nx, int ny, int nz, int rows, int levels, int prefetch) {
  std::lock_guard<std::mutex> lk(g_mu);
  return tune_add(kernel, nx, ny, nz, rows, levels, prefetch);
}
extern "C" int wrfmi_tuning_clear(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_tune.clear();
  return 0;
}
extern "C" int wrfmi_tuning_load(const char *path) {
  if (!path) return fail(WRFMI_E_POINTER, "null path");
  std::lock_guard<std::mutex> lk(g_mu);
  return tune_load(path);
}
extern "C" int wrfmi_tuning_get(const char *kernel, int nx, int ny, int nz, int *rows, int *levels, int *prefetch) {
  if (!rows || !levels || !prefetch) return fail(WRFMI_E_POINTER, "null output");
  std::lock_guard<std::mutex> lk(g_mu);
  const int kid = kernel_id(kernel);
  if (kid < 0) return fail(WRFMI_E_BADOPT, "no kernel named '%s'", kernel ? kernel : "(null)");
  Dm d;
  memset(&d, 0, sizeof d);
  d.ids = 1; d.ide = nx + 1; d.jds = 1; d.jde = ny + 1; d.kds = 1; d.kde = nz + 1;
  const Shape s = shape_for(kid, d);
  *rows = s.rows;
  *levels = s.levels;
  *prefetch = s.prefetch;
  return 0;
}

// ------------------------------------------------------------------ kernel timing entry points
extern "C" int wrfmi_kernel_count(void) { return KID_COUNT; }
extern "C" const char *wrfmi_kernel_name(int id) { return (id >= 0 && id < KID_COUNT) ? kKernelNames[id] : ""; }
extern "C" int wrfmi_timing_enable(unsigned kernel_mask) {
  std::lock_guard<std::mutex> lk(g_mu);
  HIPCHK(hipDeviceSynchronize());
  for (EvPair &e : g_ev_pending) { g_ev_free.push_back(e.a); g_ev_free.push_back(e.b); }
  g_ev_pending.clear();
  for (int n = 0; n < KID_COUNT; n++) { g_t_count[n] = 0; g_t_ms[n] = 0.0; }
  g_timing = kernel_mask;
  return 0;
}
extern "C" int wrfmi_timing_collect(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (EvPair &e : g_ev_pending) {
    HIPCHK(hipEventSynchronize(e.b));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
    g_t_count[e.id]++;
    g_t_ms[e.id] += ms;
    g_ev_free.push_back(e.a);
    g_ev_free.push_back(e.b);
  }
  g_ev_pending.clear();
  return 0;
}
extern "C" int wrfmi_timing_get(int id, long long *launches, double *total_ms) {
  if (id < 0 || id >= KID_COUNT || !launches || !total_ms) return fail(WRFMI_E_BADOPT, "bad kernel id or null output");
  std::lock_guard<std::mutex> lk(g_mu);
  *launches = g_t_count[id];
  *total_ms = g_t_ms[id];
  return 0;
}

// ------------------------------------------------------------------ services of host_internal.h
// for the entry points with the reference's argument lists (wrfmi_routines.hip)
static thread_local hipStream_t t_routine_stream = nullptr;
extern "C" int wrfmi_set_stream(void *stream) {
  t_routine_stream = (hipStream_t)stream;
  return 0;
}
namespace wrfmi {
namespace host {
int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
std::mutex &lock() { return g_mu; }
void *translate(void *p) { return ::translate(p); }
int check_cfg(const wrfmi_config *c) { return ::check_cfg(c); }
int make_dm(const wrfmi_dims *w, Dm *d) { return ::make_dm(w, d); }
int domain_scratch(const void *key, const Dm &d, float **dv, float **rhs, float **lo) {
  Scratch sc;
  if (int rc = get_scratch(DomainKey((uintptr_t)key, field_elems(d)), &sc)) return rc;
  *dv = sc.dv;
  *rhs = sc.rhs;
  *lo = sc.lo;
  return 0;
}
int device_array(const void *p, const char *name) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess || at.type != hipMemoryTypeDevice) {
    (void)hipGetLastError();
    return wrfmi::host::fail(WRFMI_E_POINTER, "argument %s is neither a device array nor a registered host array", name);
  }
  return 0;
}
hipStream_t routine_stream() { return t_routine_stream; }
void launch_tiles(void (*kern)(KArgs), dim3 g, dim3 b, KArgs a, hipStream_t st) {
  a.tnx = (int)g.x;
  a.tny = (int)g.y;
  a.tnz = (int)g.z;
  hipLaunchKernelGGL(kern, dim3(g.x * g.y * g.z, 1, 1), b, 0, st, a);
}
}  // namespace host
}  // namespace wrfmi
