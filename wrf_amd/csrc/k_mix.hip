// Forward-step mixing and damping of the first RK substep, added into the *_tendf arrays
// before rk_addtend_dry reads them (dyn_em/module_em.F:797-950, 1380-1436):
//   k_diff6   : sixth_order_diffusion (module_big_step_utilities_em.F:6220-6636), one point
//               per thread, for u, v, w, theta (rk_tendency's call with the full dt) and a
//               species (rk_scalar_tend's call into its tendency, with muts and dt_rk)
//   k_diff2   : diff_opt 1 with km_opt 1, horizontal_diffusion and the vertical_diffusion
//               routines (module_big_step_utilities_em.F:2715-3581), one point per thread,
//               for u, v, w, theta and a species
//   k_rayleigh: rk_rayleigh_damp (module_big_step_utilities_em.F:5836-6110), one column per
//               thread: u, v, theta relaxed toward the base profiles interpolated to the
//               heights of the levels in the damping layer, w toward zero
// Every expression is in the reference's operation order.  SIN is evaluated in double and
// rounded once (as LOG and ** of the equation of state are).
#include "kernels.h"

namespace wrfmi {

DEVI float rmax_f(float a, float b) { return a > b ? a : b; }
DEVI float rmin_f(float a, float b) { return a < b ? a : b; }
DEVI float sin_r(float x) { return (float)sin((double)x); }
// 2*ASIN(1.) of a REAL: twice pi/2 rounded to FP32
constexpr float kPiR = 2.0f * 1.57079637f;

// K: 0 = u, 1 = v, 2 = w, 3 = theta (mass point, rk_tendency), 4 = species a.species
// (mass point, rk_scalar_tend)
template <typename ARGS, int K>
__global__ void __launch_bounds__(256) k_diff6(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  const wrfmi_config &c = a.c;
  const int ids = d.ids, ide = d.ide, jds = d.jds, jde = d.jde, ktf = d.kde - 1;
  // loop bounds of the reference for the whole-patch tile (periodic or open boundaries)
  int i0 = ids, i1 = ide - 1, j0 = jds, j1 = jde - 1, k0 = d.kds;
  const bool spec = c.specified || c.nested;
  if (K == 0) {
    i1 = ide;
    if (c.open_xs) i0 = ids + 3;
    if (c.open_ys) j0 = jds + 3;
    if (c.open_xe) i1 = ide - 3;
    if (c.open_ye) j1 = jde - 4;
    if (spec) { i0 = ids + 3; i1 = ide - 3; j0 = jds + 3; j1 = jde - 4; }
  } else if (K == 1) {
    j1 = jde;
    if (c.open_xs) i0 = ids + 3;
    if (c.open_ys) j0 = jds + 3;
    if (c.open_xe) i1 = ide - 4;
    if (c.open_ye) j1 = jde - 3;
    if (c.open_xs || spec) { i0 = ids + 3; i1 = ide - 4; j0 = jds + 3; j1 = jde - 3; }  // 6380-6385
  } else {
    if (c.open_xs) i0 = ids + 3;
    if (c.open_ys) j0 = jds + 3;
    if (c.open_xe) i1 = ide - 4;
    if (c.open_ye) j1 = jde - 4;
    if (spec) { i0 = ids + 3; i1 = ide - 4; j0 = jds + 3; j1 = jde - 4; }
    if (K == 2) k0 = d.kds + 1;
  }
  if (i < i0 || i > i1 || j < j0 || j > j1 || k < k0 || k > ktf) return;
  const FPtr<float> f = K == 0 ? s.u_2 : K == 1 ? s.v_2 : K == 2 ? s.w_2 : K == 3 ? s.t_2 : s.q[a.species];
  const FPtr<float> tend = K == 0 ? s.ru_tendf : K == 1 ? s.rv_tendf : K == 2 ? s.rw_tendf : K == 3 ? s.t_tendf
                                                                                                      : s.q_tend[a.species];
  const FPtr<float> mut = K == 4 ? s.muts : s.mut;
  const float dt = K == 4 ? a.dt_rk : a.dt;
  const float c1 = K == 2 ? s.c1f[X1(d, k)] : s.c1h[X1(d, k)], c2 = K == 2 ? s.c2f[X1(d, k)] : s.c2h[X1(d, k)];
  const float coef = c.diff_6th_factor * 0.015625f / (2.0f * dt);
  const int opt = c.diff_6th_opt;
  const Off o = X3(d, i, k, j);
  const int sj = d.sj;
  auto F = [&](int di, int dj) { return f[o + di + dj * sj]; };
  auto MU = [&](int di, int dj) { return c1 * mut[X2(d, i + di, j + dj)] + c2; };
  auto phb = [&](int di, int dj) { return s.phb[o + di + dj * sj]; };
  auto M2 = [&](const FPtr<float> &m, int di, int dj) { return m[X2(d, i + di, j + dj)]; };
  // x
  float fx0 = (10.0f * (F(0, 0) - F(-1, 0)) - 5.0f * (F(1, 0) - F(-2, 0)) + (F(2, 0) - F(-3, 0)));
  float fx1 = (10.0f * (F(1, 0) - F(0, 0)) - 5.0f * (F(2, 0) - F(-1, 0)) + (F(3, 0) - F(-2, 0)));
  if (opt == 2) {
    if (fx0 * (F(0, 0) - F(-1, 0)) <= 0.0f) fx0 = 0.0f;
    if (fx1 * (F(1, 0) - F(0, 0)) <= 0.0f) fx1 = 0.0f;
  }
  float sd0 = 1.f, sd1 = 1.f;
  if (c.diff_6th_slopeopt >= 1) {
    const float dx = 1.f / s.rdx;
    const float dzthresh = c.diff_6th_thresh * 9.81f * dx;
    float dz0, dz1;
    if (K == 0) {
      dz0 = rmax_f(__builtin_fabsf(phb(0, 0) - phb(-1, 0)) * M2(s.msfux, 0, 0), __builtin_fabsf(phb(-1, 0) - phb(-2, 0)) * M2(s.msfux, -1, 0));
      dz1 = rmax_f(__builtin_fabsf(phb(1, 0) - phb(0, 0)) * M2(s.msfux, 1, 0), __builtin_fabsf(phb(0, 0) - phb(-1, 0)) * M2(s.msfux, 0, 0));
    } else if (K == 1) {
      dz0 = rmax_f(__builtin_fabsf(phb(0, 0) - phb(-1, 0)) * M2(s.msfux, 0, 0), __builtin_fabsf(phb(0, -1) - phb(-1, -1)) * M2(s.msfux, 0, -1));
      dz1 = rmax_f(__builtin_fabsf(phb(1, 0) - phb(0, 0)) * M2(s.msfux, 1, 0), __builtin_fabsf(phb(1, -1) - phb(0, -1)) * M2(s.msfux, 1, -1));
    } else {
      dz0 = __builtin_fabsf(phb(0, 0) - phb(-1, 0)) * M2(s.msfux, 0, 0);
      dz1 = __builtin_fabsf(phb(1, 0) - phb(0, 0)) * M2(s.msfux, 1, 0);
    }
    sd0 = rmax_f(1.f - (dz0 / dzthresh), 0.f);
    sd1 = rmax_f(1.f - (dz1 / dzthresh), 0.f);
  }
  float m0, m1, tx;
  if (K == 0) {
    m0 = MU(-1, 0);
    m1 = MU(0, 0);
    tx = coef * M2(s.msfux, 0, 0) * ((sd1 * m1 * fx1) - (sd0 * m0 * fx0));
  } else if (K == 1) {
    m0 = 0.25f * (MU(-1, -1) + MU(0, -1) + MU(-1, 0) + MU(0, 0));
    m1 = 0.25f * (MU(0, -1) + MU(1, -1) + MU(0, 0) + MU(1, 0));
    tx = coef * M2(s.msfvx, 0, 0) * ((sd1 * m1 * fx1) - (sd0 * m0 * fx0));
  } else {
    m0 = 0.5f * (MU(-1, 0) + MU(0, 0));
    m1 = 0.5f * (MU(0, 0) + MU(1, 0));
    tx = coef * M2(s.msftx, 0, 0) * ((sd1 * m1 * fx1) - (sd0 * m0 * fx0));
  }
  // y
  float fy0 = (10.0f * (F(0, 0) - F(0, -1)) - 5.0f * (F(0, 1) - F(0, -2)) + (F(0, 2) - F(0, -3)));
  float fy1 = (10.0f * (F(0, 1) - F(0, 0)) - 5.0f * (F(0, 2) - F(0, -1)) + (F(0, 3) - F(0, -2)));
  if (opt == 2) {
    if (fy0 * (F(0, 0) - F(0, -1)) <= 0.0f) fy0 = 0.0f;
    if (fy1 * (F(0, 1) - F(0, 0)) <= 0.0f) fy1 = 0.0f;
  }
  sd0 = 1.f;
  sd1 = 1.f;
  if (c.diff_6th_slopeopt >= 1) {
    const float dy = 1.f / s.rdy;
    const float dzthresh = c.diff_6th_thresh * 9.81f * dy;
    float dz0, dz1;
    if (K == 0) {
      dz0 = rmax_f(__builtin_fabsf(phb(0, 0) - phb(0, -1)) * M2(s.msfvy, 0, 0), __builtin_fabsf(phb(-1, 0) - phb(-1, -1)) * M2(s.msfvy, -1, 0));
      dz1 = rmax_f(__builtin_fabsf(phb(0, 1) - phb(0, 0)) * M2(s.msfvy, 0, 1), __builtin_fabsf(phb(-1, 1) - phb(-1, 0)) * M2(s.msfvy, -1, 1));
    } else if (K == 1) {
      dz0 = rmax_f(__builtin_fabsf(phb(0, 0) - phb(0, -1)) * M2(s.msfvy, 0, 0), __builtin_fabsf(phb(0, -1) - phb(0, -2)) * M2(s.msfvy, 0, -1));
      dz1 = rmax_f(__builtin_fabsf(phb(0, 1) - phb(0, 0)) * M2(s.msfvy, 0, 1), __builtin_fabsf(phb(0, 0) - phb(0, -1)) * M2(s.msfvy, 0, 0));
    } else {
      dz0 = __builtin_fabsf(phb(0, 0) - phb(0, -1)) * M2(s.msfvy, 0, 0);
      dz1 = __builtin_fabsf(phb(0, 1) - phb(0, 0)) * M2(s.msfvy, 0, 1);
    }
    sd0 = rmax_f(1.f - (dz0 / dzthresh), 0.f);
    sd1 = rmax_f(1.f - (dz1 / dzthresh), 0.f);
  }
  float ty;
  if (K == 0) {
    m0 = 0.25f * (MU(-1, -1) + MU(0, -1) + MU(-1, 0) + MU(0, 0));
    m1 = 0.25f * (MU(-1, 0) + MU(0, 0) + MU(-1, 1) + MU(0, 1));
    ty = coef * M2(s.msfuy, 0, 0) * ((sd1 * m1 * fy1) - (sd0 * m0 * fy0));
  } else if (K == 1) {
    m0 = MU(0, -1);
    m1 = MU(0, 0);
    ty = coef * M2(s.msfvy, 0, 0) * ((sd1 * m1 * fy1) - (sd0 * m0 * fy0));
  } else {
    m0 = 0.5f * (MU(0, -1) + MU(0, 0));
    m1 = 0.5f * (MU(0, 0) + MU(0, 1));
    ty = coef * M2(s.msfty, 0, 0) * ((sd1 * m1 * fy1) - (sd0 * m0 * fy0));
  }
  tend[o] = tend[o] + tx + ty;
}

// K: 0 = u, 1 = v, 2 = w, 3 = theta (rk_tendency, module_em.F:799-877), 4 = species
// a.species (rk_scalar_tend, module_em.F:1380-1415).  diff_opt 1 with km_opt 1: the
// horizontal diffusion (horizontal_diffusion, and horizontal_diffusion_3dmp of theta -
// t_init) with the constant eddy viscosity of isotropic_km, then, with no boundary-layer
// scheme, the vertical diffusion (vertical_diffusion_u/_v, vertical_diffusion 'w' and 'm',
// _3dmp of theta - t_init, _mp of QV - qv_base), both added to the point's tendency in the
// reference's order.  The coefficient averages are formed from the constant as the
// reference forms them from its (constant) coefficient arrays.
template <typename ARGS, int K>
__global__ void __launch_bounds__(256) k_diff2(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  const wrfmi_config &c = a.c;
  const int ids = d.ids, ide = d.ide, jds = d.jds, jde = d.jde, kts = d.kds, ktf = d.kde - 1;
  const bool spec = c.specified || c.nested;
  // horizontal_diffusion's bounds (module_big_step_utilities_em.F:2772-2914)
  int hi0 = ids, hi1 = ide - 1, hj0 = jds, hj1 = jde - 1, hk0 = kts;
  if (K == 0) {
    hi1 = ide;
    if (c.open_xs || spec) hi0 = ids + 1;
    if (c.open_xe || spec) hi1 = ide - 1;
    if (c.open_ys || spec) hj0 = jds + 1;
    if (c.open_ye || spec) hj1 = jde - 2;
    if (c.periodic_x) { hi0 = ids; hi1 = ide; }
  } else if (K == 1) {
    hj1 = jde;
    if (c.open_xs || spec) hi0 = ids + 1;
    if (c.open_xe || spec) hi1 = ide - 2;
    if (c.open_ys || spec) hj0 = jds + 1;
    if (c.open_ye || spec) hj1 = jde - 1;
    if (c.periodic_x) { hi0 = ids; hi1 = ide - 1; }
    if (c.polar) { hj0 = jds + 1; hj1 = jde - 1; }
  } else {
    if (c.open_xs || spec) hi0 = ids + 1;
    if (c.open_xe || spec) hi1 = ide - 2;
    if (c.open_ys || spec) hj0 = jds + 1;
    if (c.open_ye || spec) hj1 = jde - 2;
    if (c.periodic_x) { hi0 = ids; hi1 = ide - 1; }
    if (K == 2) hk0 = kts + 1;
  }
  // the vertical routines' bounds (3064-3581)
  int vi0 = ids, vi1 = ide - 1, vj0 = jds, vj1 = jde - 1, vk0 = kts, vk1 = ktf;
  if (K == 0) {
    vi1 = ide;
    if (c.open_xs || spec) vi0 = ids + 1;
    if (c.open_xe || spec) vi1 = ide - 1;
    if (c.periodic_x) { vi0 = ids; vi1 = ide; }
    vk1 = ktf - 1;
  } else if (K == 1) {
    if (c.open_ys || spec) vj0 = jds + 1;
    if (c.open_ye || spec) vj1 = jde - 1;
    vk1 = ktf - 1;
  } else if (K == 2) {
    vk0 = kts + 1;
  }
  const bool hor = i >= hi0 && i <= hi1 && j >= hj0 && j <= hj1 && k >= hk0 && k <= ktf;
  const bool ver = c.bl_pbl_physics == 0 && (K != 2 || c.non_hydrostatic) && i >= vi0 && i <= vi1 && j >= vj0 &&
                   j <= vj1 && k >= vk0 && k <= vk1;
  if (!hor && !ver) return;
  const FPtr<float> f = K == 0 ? s.u_2 : K == 1 ? s.v_2 : K == 2 ? s.w_2 : K == 3 ? s.t_2 : s.q[a.species];
  FPtr<float> tend = K == 0 ? s.ru_tendf : K == 1 ? s.rv_tendf : K == 2 ? s.rw_tendf : K == 3 ? s.t_tendf
                                                                                            : s.q_tend[a.species];
  const FPtr<float> mut = K == 4 ? s.muts : s.mut;
  const float c1 = K == 2 ? s.c1f[X1(d, k)] : s.c1h[X1(d, k)], c2 = K == 2 ? s.c2f[X1(d, k)] : s.c2h[X1(d, k)];
  // prandtl = 1./3.0 (share/module_model_constants.F:86); xkmh = khdif for u, v, w and
  // xkhh = khdif/prandtl for theta and the species (isotropic_km)
  const float prandtl = 1.0f / 3.0f;
  const float xk = K <= 2 ? c.khdif : c.khdif / prandtl;
  const float kv = K <= 2 ? c.kvdif : K == 3 ? 3.0f * c.kvdif : c.kvdif / prandtl;
  const Off o = X3(d, i, k, j);
  const int sj = d.sj, ni = d.ni;
  const bool qv = K == 4 && a.species == 0 && s.n_moist > 0;
  float t = tend[o];
  if (hor) {
    auto F = [&](int di, int dj) { return f[o + di + dj * sj]; };
    auto CM = [&](int di, int dj) { return c1 * mut[X2(d, i + di, j + dj)] + c2; };
    auto M2 = [&](const FPtr<float> &m, int di, int dj) { return m[X2(d, i + di, j + dj)]; };
    const float rdx = s.rdx, rdy = s.rdy;
    float mkrdxm, mkrdxp, mrdx, mkrdym, mkrdyp, mrdy;
    if (K == 0) {
      mkrdxm = (M2(s.msftx, -1, 0) / M2(s.msfty, -1, 0)) * CM(-1, 0) * xk * rdx;
      mkrdxp = (M2(s.msftx, 0, 0) / M2(s.msfty, 0, 0)) * CM(0, 0) * xk * rdx;
      mrdx = M2(s.msfux, 0, 0) * M2(s.msfuy, 0, 0) * rdx;
      mkrdym = ((M2(s.msfuy, 0, 0) + M2(s.msfuy, 0, -1)) / (M2(s.msfux, 0, 0) + M2(s.msfux, 0, -1))) * 0.25f *
               (CM(0, 0) + CM(0, -1) + CM(-1, -1) + CM(-1, 0)) * 0.25f * (xk + xk + xk + xk) * rdy;
      mkrdyp = ((M2(s.msfuy, 0, 0) + M2(s.msfuy, 0, 1)) / (M2(s.msfux, 0, 0) + M2(s.msfux, 0, 1))) * 0.25f *
               (CM(0, 0) + CM(0, 1) + CM(-1, 1) + CM(-1, 0)) * 0.25f * (xk + xk + xk + xk) * rdy;
      mrdy = M2(s.msfux, 0, 0) * M2(s.msfuy, 0, 0) * rdy;
    } else if (K == 1) {
      mkrdxm = ((M2(s.msfvx, 0, 0) + M2(s.msfvx, -1, 0)) / (M2(s.msfvy, 0, 0) + M2(s.msfvy, -1, 0))) * 0.25f *
               (CM(0, 0) + CM(0, -1) + CM(-1, -1) + CM(-1, 0)) * 0.25f * (xk + xk + xk + xk) * rdx;
      mkrdxp = ((M2(s.msfvx, 0, 0) + M2(s.msfvx, 1, 0)) / (M2(s.msfvy, 0, 0) + M2(s.msfvy, 1, 0))) * 0.25f *
               (CM(0, 0) + CM(0, -1) + CM(1, -1) + CM(1, 0)) * 0.25f * (xk + xk + xk + xk) * rdx;
      mrdx = M2(s.msfvx, 0, 0) * M2(s.msfvy, 0, 0) * rdx;
      // the reference's y part of v carries no column mass (2822-2823)
      mkrdym = (M2(s.msfty, 0, -1) / M2(s.msftx, 0, -1)) * xk * rdy;
      mkrdyp = (M2(s.msfty, 0, 0) / M2(s.msftx, 0, 0)) * xk * rdy;
      mrdy = M2(s.msfvx, 0, 0) * M2(s.msfvy, 0, 0) * rdy;
    } else if (K == 2) {
      mkrdxm = (M2(s.msfux, 0, 0) / M2(s.msfuy, 0, 0)) * 0.25f * (CM(0, 0) + CM(-1, 0) + CM(0, 0) + CM(-1, 0)) *
               0.25f * (xk + xk + xk + xk) * rdx;
      mkrdxp = (M2(s.msfux, 1, 0) / M2(s.msfuy, 1, 0)) * 0.25f * (CM(1, 0) + CM(0, 0) + CM(1, 0) + CM(0, 0)) * 0.25f *
               (xk + xk + xk + xk) * rdx;
      mrdx = M2(s.msftx, 0, 0) * M2(s.msfty, 0, 0) * rdx;
      mkrdym = (M2(s.msfvy, 0, 0) * M2(s.msfvx_inv, 0, 0)) * 0.25f * (CM(0, 0) + CM(0, -1) + CM(0, 0) + CM(0, -1)) *
               0.25f * (xk + xk + xk + xk) * rdy;
      mkrdyp = (M2(s.msfvy, 0, 1) * M2(s.msfvx_inv, 0, 1)) * 0.25f * (CM(0, 1) + CM(0, 0) + CM(0, 1) + CM(0, 0)) *
               0.25f * (xk + xk + xk + xk) * rdy;
      mrdy = M2(s.msftx, 0, 0) * M2(s.msfty, 0, 0) * rdy;
    } else {
      mkrdxm = (M2(s.msfux, 0, 0) / M2(s.msfuy, 0, 0)) * 0.5f * (xk + xk) * 0.5f * (CM(0, 0) + CM(-1, 0)) * rdx;
      mkrdxp = (M2(s.msfux, 1, 0) / M2(s.msfuy, 1, 0)) * 0.5f * (xk + xk) * 0.5f * (CM(1, 0) + CM(0, 0)) * rdx;
      mrdx = M2(s.msftx, 0, 0) * M2(s.msfty, 0, 0) * rdx;
      mkrdym = (M2(s.msfvy, 0, 0) * M2(s.msfvx_inv, 0, 0)) * 0.5f * (xk + xk) * 0.5f * (CM(0, 0) + CM(0, -1)) * rdy;
      mkrdyp = (M2(s.msfvy, 0, 1) * M2(s.msfvx_inv, 0, 1)) * 0.5f * (xk + xk) * 0.5f * (CM(0, 1) + CM(0, 0)) * rdy;
      mrdy = M2(s.msftx, 0, 0) * M2(s.msfty, 0, 0) * rdy;
    }
    if (K == 3) {
      // field - t_init (horizontal_diffusion_3dmp)
      auto B = [&](int di, int dj) { return s.t_init[o + di + dj * sj]; };
      t = t + (mrdx * (mkrdxp * (F(1, 0) - F(0, 0) - B(1, 0) + B(0, 0)) - mkrdxm * (F(0, 0) - F(-1, 0) - B(0, 0) + B(-1, 0))) +
               mrdy * (mkrdyp * (F(0, 1) - F(0, 0) - B(0, 1) + B(0, 0)) - mkrdym * (F(0, 0) - F(0, -1) - B(0, 0) + B(0, -1))));
    } else {
      t = t + (mrdx * (mkrdxp * (F(1, 0) - F(0, 0)) - mkrdxm * (F(0, 0) - F(-1, 0))) +
               mrdy * (mkrdyp * (F(0, 1) - F(0, 0)) - mkrdym * (F(0, 0) - F(0, -1))));
    }
  }
  if (ver) {
    const float g = K_G;
    auto AL = [&](int dk, int di, int dj) { return s.alt[o + dk * ni + di + dj * sj]; };
    if (K == 0 || K == 1) {
      const int di = K == 0 ? -1 : 0, dj = K == 1 ? -1 : 0;
      const CPtr<float> &base = K == 0 ? s.u_base : s.v_base;
      // vflux(kk) between levels kk and kk+1 (kk = kts-1 takes vflux(kts))
      auto vflux = [&](int kk) {
        if (kk < kts) kk = kts;
        const int dk = kk - k;
        return kv * s.rdn[X1(d, kk + 1)] / (0.25f * (AL(dk, 0, 0) + AL(dk, di, dj) + AL(dk + 1, 0, 0) + AL(dk + 1, di, dj))) *
               (f[o + (dk + 1) * ni] - f[o + dk * ni] - base[X1(d, kk + 1)] + base[X1(d, kk)]);
      };
      const float mu = K == 0 ? s.muu[X2(d, i, j)] : s.muv[X2(d, i, j)];
      t = t + g * g * s.rdnw[X1(d, k)] / (s.c1h[X1(d, k)] * mu + s.c2h[X1(d, k)]) / (0.5f * (AL(0, di, dj) + AL(0, 0, 0))) *
                  (vflux(k) - vflux(k - 1));
    } else if (K == 2) {
      auto vflux = [&](int kk) {
        if (kk >= ktf) return 0.0f;
        const int dk = kk - k;
        return (kv / AL(dk, 0, 0)) * s.rdnw[X1(d, kk)] * (f[o + (dk + 1) * ni] - f[o + dk * ni]);
      };
      t = t + s.rdn[X1(d, k)] * g * g / (c1 * mut[X2(d, i, j)] + c2) / (0.5f * (AL(0, 0, 0) + AL(-1, 0, 0))) *
                  (vflux(k) - vflux(k - 1));
    } else {
      // 'm', _mp (QV - qv_base) and _3dmp (theta - t_init); zero gradient below the lowest
      // level, no flux through the top
      auto vflux = [&](int kk) {
        if (kk < kts) kk = kts;
        if (kk >= ktf) return 0.0f;
        const int dk = kk - k;
        float df = f[o + (dk + 1) * ni] - f[o + dk * ni];
        if (qv) df = df - s.qv_base[X1(d, kk + 1)] + s.qv_base[X1(d, kk)];
        if (K == 3) df = df - s.t_init[o + (dk + 1) * ni] + s.t_init[o + dk * ni];
        return kv * s.rdn[X1(d, kk + 1)] / (0.5f * (AL(dk, 0, 0) + AL(dk + 1, 0, 0))) * df;
      };
      t = t + g * g / (c1 * mut[X2(d, i, j)] + c2) / AL(0, 0, 0) * s.rdnw[X1(d, k)] * (vflux(k) - vflux(k - 1));
    }
  }
  tend[o] = t;
}

// the damping of one field in the layer of one column: levels ktf downward while the
// level's height z00(k) is within zdamp of the top (the contiguous layer the reference's
// DO WHILE finds), the reference value interpolated from the base profile `base`
template <typename Z>
DEVI void damp_layer(const KArgs &a, FPtr<float> tnd, const FPtr<float> &fld, const CPtr<float> &base, float mu,
                     float ztop, Off ob, Z &&z00) {
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int ktf = d.kde - 1;
  const float zdamp = a.c.zdamp, dampcoef = a.c.dampcoef;
  for (int k = ktf; k >= d.kds; k--) {
    const float z = z00(k);
    if (!(z >= (ztop - zdamp))) break;
    int k2 = ktf;
    while (s.z_base[X1(d, k2)] > z) k2 = k2 - 1;
    float ref;
    if (k2 + 1 > ktf)
      ref = base[X1(d, k2)] + (base[X1(d, k2)] - base[X1(d, k2 - 1)]) * (z - s.z_base[X1(d, k2)]) /
                                  (s.z_base[X1(d, k2)] - s.z_base[X1(d, k2 - 1)]);
    else
      ref = base[X1(d, k2)] + (base[X1(d, k2 + 1)] - base[X1(d, k2)]) * (z - s.z_base[X1(d, k2)]) /
                                  (s.z_base[X1(d, k2 + 1)] - s.z_base[X1(d, k2)]);
    float dcoef = 1.0f - rmin_f(1.0f, (ztop - z) / zdamp);
    dcoef = sin_r(0.5f * kPiR * dcoef);
    dcoef = dcoef * dcoef;
    const Off o = ob + k * d.ni;
    tnd[o] = tnd[o] - (s.c1h[X1(d, k)] * mu + s.c2h[X1(d, k)]) * (dcoef * dampcoef) * (fld[o] - ref);
  }
}

// one column (i, j) of rk_rayleigh_damp: u at (i, j) for j <= jde-1, v for i <= ide-1, w
// and theta at mass columns
template <typename ARGS>
__global__ void __launch_bounds__(256) k_rayleigh(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  if (i > d.ide || j > d.jde) return;
  const int kde = d.kde, ni = d.ni, sj = d.sj;
  const float g = K_G, zdamp = a.c.zdamp, dampcoef = a.c.dampcoef;
  const FPtr<float> phb = s.phb, ph = s.ph_2;
  const Off ob = X3(d, i, 0, j);
  const Off ot = ob + kde * ni;
  if (j <= d.jde - 1) {  // u
    const float ztop = 0.5f * (phb[ot] + phb[ot - 1] + ph[ot] + ph[ot - 1]) / g;
    damp_layer(a, s.ru_tendf, s.u_2, s.u_base, s.muu[X2(d, i, j)], ztop, ob, [&](int k) {
      const Off o = ob + k * ni;
      return 0.25f * (phb[o] + phb[o + ni] + phb[o - 1] + phb[o - 1 + ni] + ph[o] + ph[o + ni] + ph[o - 1] + ph[o - 1 + ni]) / g;
    });
  }
  if (i <= d.ide - 1) {  // v
    const float ztop = 0.5f * (phb[ot] + phb[ot - sj] + ph[ot] + ph[ot - sj]) / g;
    damp_layer(a, s.rv_tendf, s.v_2, s.v_base, s.muv[X2(d, i, j)], ztop, ob, [&](int k) {
      const Off o = ob + k * ni;
      return 0.25f * (phb[o] + phb[o + ni] + phb[o - sj] + phb[o - sj + ni] + ph[o] + ph[o + ni] + ph[o - sj] + ph[o - sj + ni]) / g;
    });
  }
  if (i <= d.ide - 1 && j <= d.jde - 1) {
    const float ztop = (phb[ot] + ph[ot]) / g;
    const float mut = s.mut[X2(d, i, j)];
    // w: every level within zdamp of the top
    for (int k = d.kds; k <= kde; k++) {
      const Off o = ob + k * ni;
      const float z = (phb[o] + ph[o]) / g;
      if (z >= (ztop - zdamp)) {
        float dcoef = 1.0f - rmin_f(1.0f, (ztop - z) / zdamp);
        dcoef = sin_r(0.5f * kPiR * dcoef);
        dcoef = dcoef * dcoef;
        s.rw_tendf[o] = s.rw_tendf[o] - (s.c1f[X1(d, k)] * mut + s.c2f[X1(d, k)]) * (dcoef * dampcoef) * s.w_2[o];
      }
    }
    // theta
    damp_layer(a, s.t_tendf, s.t_2, s.t_base, mut, ztop, ob, [&](int k) {
      const Off o = ob + k * ni;
      return 0.5f * (phb[o] + phb[o + ni] + ph[o] + ph[o + ni]) / g;
    });
  }
}

WRFMI_INST_P(k_diff6, 0)
WRFMI_INST_P(k_diff6, 1)
WRFMI_INST_P(k_diff6, 2)
WRFMI_INST_P(k_diff6, 3)
WRFMI_INST_P(k_diff6, 4)
WRFMI_INST(k_rayleigh)
WRFMI_INST_P(k_diff2, 0)
WRFMI_INST_P(k_diff2, 1)
WRFMI_INST_P(k_diff2, 2)
WRFMI_INST_P(k_diff2, 3)
WRFMI_INST_P(k_diff2, 4)

}  // namespace wrfmi
