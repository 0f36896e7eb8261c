// Forward-step mixing and damping of the first RK substep, added into the *_tendf arrays
// before rk_addtend_dry reads them (dyn_em/module_em.F:800-950, 1421-1436):
//   k_diff6   : sixth_order_diffusion (module_big_step_utilities_em.F:6220-6636), one point
//               per thread, for u, v, w, theta (rk_tendency's call with the full dt) and a
//               species (rk_scalar_tend's call into its tendency, with muts and dt_rk)
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
  if (K == 0) {
    i1 = ide;
    if (c.open_xs) i0 = ids + 3;
    if (c.open_ys) j0 = jds + 3;
    if (c.open_xe) i1 = ide - 3;
    if (c.open_ye) j1 = jde - 4;
  } else if (K == 1) {
    j1 = jde;
    if (c.open_xs) i0 = ids + 3;
    if (c.open_ys) j0 = jds + 3;
    if (c.open_xe) i1 = ide - 4;
    if (c.open_ye) j1 = jde - 3;
    if (c.open_xs) { i0 = ids + 3; i1 = ide - 4; j0 = jds + 3; j1 = jde - 3; }  // 6380-6385
  } else {
    if (c.open_xs) i0 = ids + 3;
    if (c.open_ys) j0 = jds + 3;
    if (c.open_xe) i1 = ide - 4;
    if (c.open_ye) j1 = jde - 4;
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

}  // namespace wrfmi
