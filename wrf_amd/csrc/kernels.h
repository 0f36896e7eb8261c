// Kernel entry points of the dynamics step, shared by the kernel sources and the host driver.
#pragma once
#include "dev_common.h"

namespace wrfmi {
constexpr int kAdvUvK = 4;  // levels per thread in k_advance_uv (the four-point form: a launch shape)
constexpr int kAdvK = 8;     // levels per block in k_adv_scalar
constexpr int kTendK = 8;    // levels per thread in k_tend_u/v/w/t
constexpr int kAdvTileI = 63;  // i points per wave in k_adv_scalar (lane 63 supplies a face)
constexpr int kAdvRows = 4;    // j rows (waves) per block in k_adv_scalar
constexpr int kVecW = 4;       // i points per lane in the vector kernels (16-byte accesses)
constexpr int kVecTileI = 64 * kVecW;  // i points per wave in the vector kernels
constexpr int kAdvVK = 4;      // levels per thread in k_adv_scalar_x4
constexpr int kHaloK = 8;      // levels per thread in k_halo
constexpr int kDivK = 4;       // default levels per thread in k_ac_div_x4 (a launch shape)
constexpr int kPdK = 8;        // levels per thread in k_pd_scale and k_pd_tend
constexpr int kPdVK = 4;       // levels per thread in k_pd_scale_x4 and k_pd_tend_x4
// The four-point 5th-order stencil kernels (k_tend_u/v/w/t_x4, k_adv_scalar_x4,
// k_adv_scalar_all_x4) take workgroups of up to 8 rows of 256 points (a launch shape): a
// taller tile re-reads fewer of the j+-3 halo rows its stencils share with the tiles above
// and below.  (At 1024 threads the compiler would cap them at 128 VGPRs and spill.)
constexpr int kTallBlk = 512;
// The reference routine a launch computes on its own, for the per-routine entry points
// with the reference's argument lists (KArgs::routine); R_FUSED is the whole step's form,
// with each kernel's fused parts.
enum Routine : int {
  R_FUSED = 0,
  R_ADVECT = 1,   // k_tend_u/v/w, k_adv_scalar: advect_u/v/w/scalar only, added to the tendency array's values
  R_MU_T = 2,     // k_ac_div, k_ac_theta: advance_mu_t only (no sumflux; t_ave = theta before the update)
  R_W_TAVE = 3,   // k_ac_theta: advance_w's time-averaged theta only (from t_2 and advance_mu_t's t_ave)
  R_SUMFLUX = 4,  // k_ac_div, k_ac_theta: sumflux only
  R_COEF_W = 5,   // k_coef_w: calc_coef_w only (no small_step_prep 2-D updates)
};
// Every state kernel is a template on its argument type: KArgs for one domain, BatchArgs
// for an ensemble launch over the tiles of all members (k_halo: HaloArgs / BatchHalo).
// ZONE = true is the one-point kernel's form over the boundary zone of a domain with an
// open direction (zone_point), launched after the four-point form.
template <typename A> __global__ void k_step_prep(A);
template <typename A> __global__ void k_step_prep_x4(A);
template <typename A, bool COMP> __global__ void k_ww_scan(A);
template <typename A, bool PER, bool ZONE = false> __global__ void k_tend_u(A);
template <typename A> __global__ void k_tend_u_x4(A);
template <typename A, bool PER, bool ZONE = false> __global__ void k_tend_v(A);
template <typename A> __global__ void k_tend_v_x4(A);
template <typename A, bool PER, bool ZONE = false> __global__ void k_tend_w(A);
template <typename A> __global__ void k_tend_w_x4(A);
template <typename A, bool PER, bool ZONE = false> __global__ void k_tend_t(A);
template <typename A> __global__ void k_tend_t_x4(A);
template <typename A> __global__ void k_small_prep(A);
template <typename A> __global__ void k_p_rho0(A);
template <typename A, typename CT, bool COEF> __global__ void k_coef_w(A);
template <typename A, bool PER> __global__ void k_advance_uv(A);
template <typename A, bool PER, int KS> __global__ void k_advance_uv_x4(A);
template <typename A> __global__ void k_ac_div(A);
template <typename A, bool COMP> __global__ void k_ac_mu_ww(A);
template <typename A> __global__ void k_ac_theta(A);
template <typename A, int KS> __global__ void k_ac_div_x4(A);
template <typename A> __global__ void k_small_prep_x4(A);
template <typename A> __global__ void k_p_rho0_x4(A);
template <typename A> __global__ void k_ac_theta_x4(A);
template <typename A, typename CT, bool PRHO = true, int PF = 4> __global__ void k_ac_wsolve(A);
template <typename A> __global__ void k_calc_p_rho(A);
template <typename A> __global__ void k_small_finish(A);
template <typename A> __global__ void k_small_finish_x4(A);
template <typename A, bool PER, bool ZONE = false> __global__ void k_adv_scalar(A);
template <typename A> __global__ void k_adv_scalar_x4(A);
template <typename A> __global__ void k_update_scalar(A);
template <typename A> __global__ void k_update_scalar_x4(A);
template <typename A> __global__ void k_adv_scalar_all_x4(A);
template <typename A> __global__ void k_update_scalar_all_x4(A);
template <typename A> __global__ void k_pd_scale(A);
template <typename A> __global__ void k_pd_tend(A);
template <typename A> __global__ void k_pd_scale_x4(A);
template <typename A> __global__ void k_pd_tend_x4(A);
template <typename A> __global__ void k_update_scalar_pd(A);
template <typename A, int K> __global__ void k_diff6(A);
template <typename A> __global__ void k_rayleigh(A);

// the outermost mass row of a specified (or nested) domain, which advance_mu_t and
// advance_w leave to the boundary updates (module_small_step_em.F:1054-1063, 1273-1282)
DEVI bool spec_frame(const wrfmi_config &c, const Dm &d, int i, int j) {
  if (!(c.specified || c.nested)) return false;
  return (!c.periodic_x && (i == d.ids || i == d.ide - 1)) || j == d.jds || j == d.jde - 1;
}

// specified lateral boundaries (k_lbc.hip)
enum LbcMode : int { LBC_RELAX, LBC_SPEC_TEND, LBC_UPDATE, LBC_PH_PRHO, LBC_ZERO_GRAD, LBC_FLOW_DEP, LBC_ZERO };
// one field of a boundary launch
struct LbcJob {
  const float *b[4], *bt[4];  // the field's boundary values and tendencies (xs, xe, ys, ye)
  const float *src;   // LBC_RELAX: the relaxed field
  const float *tend;  // LBC_UPDATE: the tendency
  float *dst;         // the array written (LBC_ZERO_GRAD, LBC_FLOW_DEP: also read)
  int var;            // staggering: 'u', 'v', 'h' (w levels), 'w', 'm' (2-D), other mass levels
  int massw;          // LBC_RELAX: weight the field by c1 mut + c2 (c1f, c2f for 'h')
};
constexpr int kLbcJobs = 6;     // fields per launch
constexpr int kLbcMaxK = 256;   // levels of the k_lbc_ph column (LDS)
struct LbcArgs {
  KArgs a;
  LbcJob job[kLbcJobs];
  float fcx[WRFMI_MAX_BDY_WIDTH], gcx[WRFMI_MAX_BDY_WIDTH];
  float dtbc, dt;
  int mode;           // LbcMode
  int njobs;
  int whole;          // LBC_ZERO: the whole extent (ids..ide, jds..jde, every level)
  int width;          // rows from the edges the launch covers (spec_zone, or relax_zone)
  int zero_grad_w;    // k_lbc_ph: also zero_grad_bdy of w
};
static_assert(sizeof(LbcArgs) <= 4096, "kernel arguments");
// the frame of `width` rows along the edges of the extent [ibs, ibe] x [jbs, jbe] (no
// x columns when periodic in x), enumerated y rows first: rb rows at jbs, rt at jbe, then
// the nmid rows between with cl columns at ibs and cr at ibe
struct LbcRing {
  int ni, rb, rt, cl, cr;
  long count;
};
HDI LbcRing lbc_ring(int ibs, int ibe, int jbs, int jbe, bool px, int width) {
  LbcRing r;
  const int nj = jbe - jbs + 1;
  r.ni = ibe - ibs + 1;
  r.rb = width < nj ? width : nj;
  r.rt = width < nj - r.rb ? width : nj - r.rb;
  r.cl = px ? 0 : (width < r.ni ? width : r.ni);
  r.cr = px ? 0 : (width < r.ni - r.cl ? width : r.ni - r.cl);
  r.count = (long)(r.rb + r.rt) * r.ni + (long)(nj - r.rb - r.rt) * (r.cl + r.cr);
  return r;
}
__global__ void k_lbc(LbcArgs);
__global__ void k_lbc_ph(LbcArgs);
template <typename A, int K> __global__ void k_diff2(A);
template <typename A> __global__ void k_p_rho_phi(A);
template <typename A> __global__ void k_p_rho_phi_x4(A);
template <typename A> __global__ void k_halo(A);
template <typename A> __global__ void k_set_w_surface(A);
__global__ void k_count_nonfinite(const float *a, Dm d, int kend, int is3d, unsigned long long *cnt);
// domain totals (k_diag.hip): dry mass, theta budget, kinetic, internal, potential, total
constexpr int kNTotals = 6;
__global__ void k_domain_totals(KArgs a, double *part);
__global__ void k_domain_totals_final(const double *part, int nblocks, double *out);
}  // namespace wrfmi

// explicit instantiations, in the source file that defines the kernel
#define WRFMI_INST(K) \
  template __global__ void K<KArgs>(KArgs); \
  template __global__ void K<BatchArgs>(BatchArgs);
#define WRFMI_INST_P(K, P) \
  template __global__ void K<KArgs, P>(KArgs); \
  template __global__ void K<BatchArgs, P>(BatchArgs);
#define WRFMI_INST_P2(K, P, Q) \
  template __global__ void K<KArgs, P, Q>(KArgs); \
  template __global__ void K<BatchArgs, P, Q>(BatchArgs);
