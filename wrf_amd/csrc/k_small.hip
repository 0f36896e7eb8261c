// The acoustic (small) step on the GPU, dyn_em/module_small_step_em.F.
//   k_small_prep   : small_step_prep (16-290), one thread per (i,k,j) point
//   k_p_rho0       : calc_p_rho at step 0 (438-568), one thread per point
//   k_coef_w       : small_step_prep's 2-D updates + calc_coef_w (570-652), one thread per column
//   k_advance_uv   : advance_uv (654-967), one thread per (i,k,j) point
//   k_ac_*         : advance_mu_t (969-1175) + advance_w (1178-1469) + sumflux (1473-1633)
//                    + calc_p_rho (438-568), in six kernels (see below)
//   k_small_finish : calc_mu_uv_1 (module_big_step_utilities_em.F:184-321) + small_step_finish (295-434)
// Arguments are bound as at the solve_em.F call sites (1089-1757): the coupled
// perturbations live in u_2..ph_2, the time-t values in *_save, ww1 is the large-step
// omega, t_2save holds t_2ave.  CT is the type of the tridiagonal coefficients
// a, alpha, gamma (float: reference order; double: FP64 coefficient build and solve).
#include "kernels.h"

namespace wrfmi {

// small_step_prep, point part: one thread per (i, k, j).  The 2-D fields it updates
// (muts, muus, muvs, mu_1, mu_2, mu_save, mudf) are written afterwards by k_coef_w, so that
// every thread here reads their values from before the update.
struct MuS {  // small_step_prep's column masses of one (i, j)
  float muts, muus, muvs, mu_new;
};
DEVI MuS small_mu(const KArgs &a, int i, int j) {
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const Off o2 = X2(d, i, j);
  MuS m;
  if (a.rk_step == 1) {
    m.muts = s.mub[o2] + s.mu_2[o2];
    m.muus = s.muu[o2];
    m.muvs = s.muv[o2];
    m.mu_new = 0.f;
  } else {
    m.muts = s.mub[o2] + s.mu_1[o2];
    m.muus = 0.5f * (((s.mub[o2] + s.mu_1[o2]) + s.mub[X2(d, i - 1, j)]) + s.mu_1[X2(d, i - 1, j)]);
    m.muvs = 0.5f * (((s.mub[o2] + s.mu_1[o2]) + s.mub[X2(d, i, j - 1)]) + s.mu_1[X2(d, i, j - 1)]);
    m.mu_new = s.mu_1[o2] - s.mu_2[o2];
  }
  return m;
}

template <typename ARGS>
__global__ void __launch_bounds__(256) k_small_prep(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  const int ide = d.ide, jde = d.jde, kde = d.kde;
  if (i > ide || j > jde || k > kde) return;
  const bool mi = i <= ide - 1, mj = j <= jde - 1, mass = mi && mj;
  const int k_end = kde - 1;
  const CPtr<float> c1h = s.c1h, c2h = s.c2h, c1f = s.c1f, c2f = s.c2f;
  const Off o2 = X2(d, i, j), o = X3(d, i, k, j);
  const bool rk1 = a.rk_step == 1;
  const MuS m = small_mu(a, i, j);
  if (mj && k <= k_end) {
    const float u2 = s.u_2[o];
    float u1;
    if (rk1) { u1 = u2; s.u_1[o] = u1; } else { u1 = s.u_1[o]; }
    s.u_save[o] = u2;
    s.u_2[o] = ((c1h[X1(d, k)] * m.muus + c2h[X1(d, k)]) * u1 - (c1h[X1(d, k)] * s.muu[o2] + c2h[X1(d, k)]) * u2) / s.msfuy[o2];
  }
  if (mi && k <= k_end) {
    const float v2 = s.v_2[o];
    float v1;
    if (rk1) { v1 = v2; s.v_1[o] = v1; } else { v1 = s.v_1[o]; }
    s.v_save[o] = v2;
    s.v_2[o] = ((c1h[X1(d, k)] * m.muvs + c2h[X1(d, k)]) * v1 - (c1h[X1(d, k)] * s.muv[o2] + c2h[X1(d, k)]) * v2) * s.msfvx_inv[o2];
  }
  if (!mass) return;
  const float mut = s.mut[o2], msfty = s.msfty[o2];
  if (k <= k_end) {
    s.c2a[o] = K_CPOVCV * (s.pb[o] + s.p[o]) / s.alt[o];
    const float t2 = s.t_2[o];
    float t1;
    if (rk1) { t1 = t2; s.t_1[o] = t1; } else { t1 = s.t_1[o]; }
    s.t_save[o] = t2;
    s.t_2[o] = (c1h[X1(d, k)] * m.muts + c2h[X1(d, k)]) * t1 - (c1h[X1(d, k)] * mut + c2h[X1(d, k)]) * t2;
  }
  {
    const float w2 = s.w_2[o], ph2 = s.ph_2[o];
    float w1, ph1;
    if (rk1) { w1 = w2; ph1 = ph2; s.w_1[o] = w1; s.ph_1[o] = ph1; } else { w1 = s.w_1[o]; ph1 = s.ph_1[o]; }
    s.w_save[o] = w2;
    s.w_2[o] = ((c1f[X1(d, k)] * m.muts + c2f[X1(d, k)]) * w1 - (c1f[X1(d, k)] * mut + c2f[X1(d, k)]) * w2) / msfty;
    s.ph_save[o] = ph2;
    s.ph_2[o] = ph1 - ph2;
    s.ww1[o] = s.ww[o];
  }
}

// calc_p_rho at step 0 (mu = the new mu_2, mut = muts, t_1 = t_save): one thread per point
template <typename ARGS>
__global__ void __launch_bounds__(256) k_p_rho0(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  if (i > d.ide - 1 || j > d.jde - 1 || k > d.kde - 1) return;
  if (!a.c.non_hydrostatic) return;
  const Off o = X3(d, i, k, j);
  const MuS m = small_mu(a, i, j);
  const float mu = m.mu_new, muts = m.muts;
  const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
  const float alt = s.alt[o];
  const float al = -1.f / (c1 * muts + c2) * (alt * (c1 * mu) + s.rdnw[X1(d, k)] * (s.ph_2[o + d.ni] - s.ph_2[o]));
  s.al[o] = al;
  const float t1 = s.t_save[o];
  const float pp = s.c2a[o] * (alt * (s.t_2[o] - (c1 * mu) * t1) / ((c1 * muts + c2) * (K_T0 + t1)) - al);
  s.p[o] = pp;
  s.pm1[o] = pp;
}

// calc_p_rho alone (module_small_step_em.F:438-568, non-hydrostatic), for the per-routine
// entry point with the reference's argument list: mu and mut are the arrays bound to
// mu_2 and muts, t_1 to t_save, step = the iteration field (0 at small_step_prep).  The
// whole step computes the same values in k_p_rho0 (step 0) and k_ac_wsolve (steps >= 1).
template <typename ARGS>
__global__ void __launch_bounds__(256) k_calc_p_rho(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  if (i > d.ide - 1 || j > d.jde - 1 || k > d.kde - 1) return;
  const Off o = X3(d, i, k, j), o2 = X2(d, i, j);
  const float mu = s.mu_2[o2], mut = s.muts[o2];
  const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
  const float alt = s.alt[o];
  const float al = -1.f / (c1 * mut + c2) * (alt * (c1 * mu) + s.rdnw[X1(d, k)] * (s.ph_2[o + d.ni] - s.ph_2[o]));
  s.al[o] = al;
  const float t1 = s.t_save[o];
  float pp = s.c2a[o] * (alt * (s.t_2[o] - (c1 * mu) * t1) / ((c1 * mut + c2) * (K_T0 + t1)) - al);
  if (a.iteration == 0) {
    s.p[o] = pp;
    s.pm1[o] = pp;
  } else {
    const float ptmp = pp;
    pp = pp + a.c.smdiv * (pp - s.pm1[o]);
    s.p[o] = pp;
    s.pm1[o] = ptmp;
  }
}

// small_step_prep's 2-D updates and, with COEF, calc_coef_w: one thread per (i, j) column.
// The whole-step driver launches it without COEF: k_ac_wsolve forms the same coefficients
// on its way up each column from values it loads anyway.
template <typename ARGS, typename CT, bool COEF>
__global__ void __launch_bounds__(256) k_coef_w(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = col_ia(d, !std::is_same<ARGS, KArgs>::value) + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int ide = d.ide, jde = d.jde, kde = d.kde;
  if (i < d.ids || i > ide || j > jde) return;
  const bool mi = i <= ide - 1, mj = j <= jde - 1, mass = mi && mj;
  const CPtr<float> c1h = s.c1h, c2h = s.c2h, c1f = s.c1f, c2f = s.c2f;
  const Off o2 = X2(d, i, j);
  // calc_coef_w alone (a per-routine entry point): no small_step_prep updates
  if (a.routine != R_COEF_W) {
    const MuS m = small_mu(a, i, j);
    if (mass) {
      if (a.rk_step == 1) {
        s.mu_1[o2] = s.mu_2[o2];
        s.mudf[o2] = 0.f;
      }
      s.muts[o2] = m.muts;
      s.mu_save[o2] = s.mu_2[o2];
      s.mu_2[o2] = m.mu_new;
    }
    if (mj) s.muus[o2] = m.muus;
    if (mi) s.muvs[o2] = m.muvs;
  }
  if (!COEF || !mass || !a.c.non_hydrostatic) return;
  const float mut = s.mut[o2];
  const FPtr<CT> A((CT *)s.a), AL((CT *)s.alpha), GA((CT *)s.gamma);
  const CT lid = a.c.top_lid ? CT(0) : CT(1);
  const CT epssm = (CT)a.c.epssm;
  CT cof = (CT(.5) * (CT)a.dts * (CT)K_G * (CT(1) + epssm));
  cof = cof * cof;
  const CT MUT = (CT)mut;
  auto cH = [&](int k) -> CT { return (CT)c1h[X1(d, k)] * MUT + (CT)c2h[X1(d, k)]; };
  auto cF = [&](int k) -> CT { return (CT)c1f[X1(d, k)] * MUT + (CT)c2f[X1(d, k)]; };
  auto rdnw = [&](int k) -> CT { return (CT)s.rdnw[X1(d, k)]; };
  auto rdn = [&](int k) -> CT { return (CT)s.rdn[X1(d, k)]; };
  auto C2A = [&](int k) -> CT { return (CT)s.c2a[X3(d, i, k, j)]; };
  auto CQW = [&](int k) -> CT { return (CT)s.cqw[X3(d, i, k, j)]; };
  {
    const int k = kde - 1;
    A[X3(d, i, 2, j)] = CT(0);
    A[X3(d, i, kde, j)] = -CT(2) * cof * (rdnw(kde - 1) * rdnw(kde - 1)) * C2A(kde - 1) * lid / (cH(k) * cF(k));
    GA[X3(d, i, 1, j)] = CT(0);
  }
  CT gam = CT(0);
  for (int k = 2; k <= kde - 1; k++) {
    const CT ak = k >= 3 ? -CQW(k) * cof * rdn(k) * rdnw(k - 1) * C2A(k - 1) / (cH(k - 1) * cF(k - 1)) : CT(0);
    if (k >= 3) A[X3(d, i, k, j)] = ak;
    const CT b = CT(1) + CQW(k) * cof * rdn(k) *
                             (rdnw(k) * C2A(k) / (cH(k) * cF(k)) + rdnw(k - 1) * C2A(k - 1) / (cH(k - 1) * cF(k)));
    const CT c = -CQW(k) * cof * rdn(k) * rdnw(k) * C2A(k) / (cH(k) * cF(k + 1));
    const CT alpha = CT(1) / (b - ak * gam);
    AL[X3(d, i, k, j)] = alpha;
    gam = c * alpha;
    GA[X3(d, i, k, j)] = gam;
  }
  {
    const int k = kde;
    const CT b = CT(1) + CT(2) * cof * (rdnw(kde - 1) * rdnw(kde - 1)) * C2A(kde - 1) / (cH(k - 1) * cF(k));
    const CT c = CT(0);
    const CT alpha = CT(1) / (b - A[X3(d, i, kde, j)] * gam);
    AL[X3(d, i, kde, j)] = alpha;
    GA[X3(d, i, kde, j)] = c * alpha;
  }
}

// advance_uv, one thread per (i, j) column segment of kAdvUvK levels: the column constants
// (map-factor ratios, column masses, the mudf term) are formed once, and the vertically
// interpolated pressure dpn at the lower w level is carried up from the level below.
template <typename ARGS, bool PER>
__global__ void __launch_bounds__(256) k_advance_uv(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k0 = d.kds + (int)T.bz * kAdvUvK;
  const int ids = d.ids, ide = d.ide, jds = d.jds, jde = d.jde, kde = d.kde;
  const int its = ids, ite = ide, jts = jds, jte = jde;
  if (i > ide || j > jde || k0 > kde - 1) return;
  const int k1 = imin(k0 + kAdvUvK - 1, kde - 1);
  const wrfmi_config &c = a.c;
  const Geo g = geo<PER>(c);
  int i_start = its, i_end = imin(ite, ide - 1), j_start = jts, j_end = imin(jte, jde - 1), i_endu = ite, j_endv = jte;
  if (c.specified || c.nested) {  // module_small_step_em.F:734-747: the spec zone is spec_bdyupdate's
    const int sz = c.spec_zone;
    i_start = imax(its, ids + sz);
    i_end = imin(ite, ide - sz - 1);
    j_start = imax(jts, jds + sz);
    j_end = imin(jte, jde - sz - 1);
    i_endu = imin(ite, ide - sz);
    j_endv = imin(jte, jde - sz);
    if (c.periodic_x) {
      i_start = its;
      i_end = imin(ite, ide - 1);
      i_endu = ite;
    }
  }
  int i_start_up = i_start, i_end_up = i_endu, j_start_vp = j_start, j_end_vp = j_endv;
  if ((g.open_xs || g.sym_xs) && its == ids) i_start_up++;
  if ((g.open_xe || g.sym_xe) && ite == ide) i_end_up--;
  if ((g.open_ys || g.sym_ys || g.polar) && jts == jds) j_start_vp++;
  if ((g.open_ye || g.sym_ye || g.polar) && jte == jde) j_end_vp--;
  int i_start_u_tend = i_start, i_end_u_tend = i_endu, j_start_v_tend = j_start, j_end_v_tend = j_endv;
  if (g.sym_xs && its == ids) i_start_u_tend++;
  if (g.sym_xe && ite == ide) i_end_u_tend--;
  if (g.sym_ys && jts == jds) j_start_v_tend++;
  if (g.sym_ye && jte == jde) j_end_v_tend--;
  const float rdx = s.rdx, rdy = s.rdy, dx = 1.f / rdx, dy = 1.f / rdy, dts = a.dts;
  const float cf1 = s.cf1, cf2 = s.cf2, cf3 = s.cf3, emdiv = c.emdiv;
  const FPtr<float> p = s.p, pb = s.pb, ph = s.ph_2, php = s.php, alt = s.alt, al = s.al, mu = s.mu_2, mudf = s.mudf;
  const bool nh = c.non_hydrostatic;
  const Off o2 = X2(d, i, j);
  // Every lane of the (i <= ide, j <= jde) patch evaluates both updates with loads that
  // stay inside the memory patch, and only the stores are masked: loads under per-lane
  // conditions would make the compiler wait for all loads in flight at each join.
  const bool u_on = j >= j_start && j <= j_end;
  const bool v_on = i >= i_start && i <= i_end && j >= j_start_v_tend && j <= j_end_v_tend;
  const bool utend_on = i >= i_start_u_tend && i <= i_end_u_tend, upg_on = i >= i_start_up && i <= i_end_up;
  const bool vpg_on = j >= j_start_vp && j <= j_end_vp;
  // column constants (map-factor ratios, column masses, the mudf term)
  const float mudf_xu = -emdiv * dx * (mudf[o2] - mudf[X2(d, i - 1, j)]) / s.msfuy[o2];
  const float mu_m = s.msfux[o2] / s.msfuy[o2], muu = s.muu[o2], mu0u = mu[X2(d, i - 1, j)], mu1 = mu[o2];
  const float mudf_yv = -emdiv * dy * (mudf[o2] - mudf[X2(d, i, j - 1)]) * s.msfvx_inv[o2];
  const float mv_m = s.msfvy[o2] / s.msfvx[o2], muv = s.muv[o2], mu0v = mu[X2(d, i, j - 1)];
  // dpn at w level kk averaged to the u (sh = 1) or v (sh = sj) point, general form
  auto dpn_gen = [&](int kk, int sh) -> float {
    const Off q = X3(d, i, kk, j);
    return .5f * (s.fnm[X1(d, kk)] * (p[q] + p[q - sh]) + s.fnp[X1(d, kk)] * (p[q - d.ni] + p[q - d.ni - sh]));
  };
  // kk = 1: extrapolation with cf1..cf3; kk = kde: lid (zero without top_lid)
  auto dpn_bot = [&](int sh) -> float {
    const Off q = X3(d, i, 1, j);
    return .5f * (cf1 * (p[q] + p[q - sh]) + cf2 * (p[q + d.ni] + p[q + d.ni - sh]) + cf3 * (p[q + 2 * d.ni] + p[q + 2 * d.ni - sh]));
  };
  auto dpn_top = [&](int sh) -> float {
    const Off q = X3(d, i, kde - 1, j);
    return .5f * (cf1 * (p[q - sh] + p[q]) + cf2 * (p[q - d.ni - sh] + p[q - d.ni]) + cf3 * (p[q - 2 * d.ni - sh] + p[q - 2 * d.ni]));
  };
  float dpnu_lo, dpnv_lo, dpnu_top = 0.f, dpnv_top = 0.f;
  if (k0 == 1) {
    dpnu_lo = dpn_bot(1);
    dpnv_lo = dpn_bot(d.sj);
  } else {
    dpnu_lo = dpn_gen(k0, 1);
    dpnv_lo = dpn_gen(k0, d.sj);
  }
  if (c.top_lid && k1 == kde - 1) {
    dpnu_top = dpn_top(1);
    dpnv_top = dpn_top(d.sj);
  }
#pragma unroll
  for (int kk = 0; kk < kAdvUvK; kk++) {
    const int k = imin(k0 + kk, k1);  // levels past k1 repeat k1 and are not stored
    const Off o = X3(d, i, k, j);
    const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)], rdnw = s.rdnw[X1(d, k)];
    const int kp = imin(k + 1, kde);
    const bool top = k + 1 == kde;
    // u
    {
      float u = s.u_2[o];
      const float ut = s.ru_tend[o];
      u = utend_on ? u + dts * ut : u;
      float dpxy = mu_m * .5f * rdx * (c1 * muu + c2) *
                   (((ph[o + d.ni] - ph[X3(d, i - 1, k + 1, j)]) + (ph[o] - ph[o - 1])) +
                    (alt[o] + alt[o - 1]) * (p[o] - p[o - 1]) +
                    (al[o] + al[o - 1]) * (pb[o] - pb[o - 1]));
      const float g_hi = dpn_gen(kp, 1);
      const float dpn_hi = top ? dpnu_top : g_hi;
      const float dpxn = dpxy + mu_m * rdx * (php[o] - php[o - 1]) * (rdnw * (dpn_hi - dpnu_lo) - .5f * ((c1 * mu0u) + (c1 * mu1)));
      dpxy = nh ? dpxn : dpxy;
      dpnu_lo = dpn_hi;
      const float upg = u - dts * s.cqu[o] * dpxy + (c1 * mudf_xu);
      u = upg_on ? upg : u;
      if (u_on && k0 + kk <= k1) s.u_2[o] = u;
    }
    // v
    {
      float v = s.v_2[o];
      v = v + dts * s.rv_tend[o];
      float dpxy = mv_m * .5f * rdy * (c1 * muv + c2) *
                   (((ph[o + d.ni] - ph[X3(d, i, k + 1, j - 1)]) + (ph[o] - ph[o - d.sj])) +
                    (alt[o] + alt[o - d.sj]) * (p[o] - p[o - d.sj]) +
                    (al[o] + al[o - d.sj]) * (pb[o] - pb[o - d.sj]));
      const float g_hi = dpn_gen(kp, d.sj);
      const float dpn_hi = top ? dpnv_top : g_hi;
      const float dpyn = dpxy + mv_m * rdy * (php[o] - php[o - d.sj]) * (rdnw * (dpn_hi - dpnv_lo) - .5f * ((c1 * mu0v) + (c1 * mu1)));
      dpxy = nh ? dpyn : dpxy;
      dpnv_lo = dpn_hi;
      const float vpg = v - dts * s.cqv[o] * dpxy + (c1 * mudf_yv);
      v = vpg_on ? vpg : v;
      if (v_on && k0 + kk <= k1) s.v_2[o] = v;
    }
  }
}

// advance_uv with four i points per lane (16-byte loads): the same expressions as
// k_advance_uv element by element.  Values at i-1 come from the group of four below (o-4),
// values at j-1 from the row below; p and ph at level k+1 are carried to the next level.
template <typename ARGS, bool PER, int KS>
__global__ void __launch_bounds__(256) k_advance_uv_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k0 = d.kds + (int)T.bz * KS;
  const int ids = d.ids, ide = d.ide, jds = d.jds, jde = d.jde, kde = d.kde;
  const int its = ids, ite = ide, jts = jds, jte = jde;
  if (i0 > ide || j > jde || k0 > kde - 1) return;
  const int k1 = imin(k0 + KS - 1, kde - 1);
  const wrfmi_config &c = a.c;
  const Geo g = geo<PER>(c);
  int i_start = its, i_end = imin(ite, ide - 1), j_start = jts, j_end = imin(jte, jde - 1), i_endu = ite, j_endv = jte;
  if (c.specified || c.nested) {  // module_small_step_em.F:734-747: the spec zone is spec_bdyupdate's
    const int sz = c.spec_zone;
    j_start = imax(jts, jds + sz);
    j_end = imin(jte, jde - sz - 1);
    j_endv = imin(jte, jde - sz);
    if (!c.periodic_x) {
      i_start = imax(its, ids + sz);
      i_end = imin(ite, ide - sz - 1);
      i_endu = imin(ite, ide - sz);
    }
  }
  int i_start_up = i_start, i_end_up = i_endu, j_start_vp = j_start, j_end_vp = j_endv;
  if ((g.open_xs || g.sym_xs) && its == ids) i_start_up++;
  if ((g.open_xe || g.sym_xe) && ite == ide) i_end_up--;
  if ((g.open_ys || g.sym_ys || g.polar) && jts == jds) j_start_vp++;
  if ((g.open_ye || g.sym_ye || g.polar) && jte == jde) j_end_vp--;
  int i_start_u_tend = i_start, i_end_u_tend = i_endu, j_start_v_tend = j_start, j_end_v_tend = j_endv;
  if (g.sym_xs && its == ids) i_start_u_tend++;
  if (g.sym_xe && ite == ide) i_end_u_tend--;
  if (g.sym_ys && jts == jds) j_start_v_tend++;
  if (g.sym_ye && jte == jde) j_end_v_tend--;
  const float rdx = s.rdx, rdy = s.rdy, dx = 1.f / rdx, dy = 1.f / rdy, dts = a.dts;
  const float cf1 = s.cf1, cf2 = s.cf2, cf3 = s.cf3, emdiv = c.emdiv;
  const FPtr<float> p = s.p, pb = s.pb, ph = s.ph_2, php = s.php, alt = s.alt, al = s.al, mu = s.mu_2, mudf = s.mudf;
  const bool nh = c.non_hydrostatic;
  const int ni = d.ni, sj = d.sj;
  const Off o2 = X2(d, i0, j), ob = X3(d, i0, 0, j);  // level k at ob + k*ni
  const bool u_on = j >= j_start && j <= j_end;
  const bool v_on = j >= j_start_v_tend && j <= j_end_v_tend;
  const bool vpg_on = j >= j_start_vp && j <= j_end_vp;
  // column constants
  const F4 mdf = mudf.ld<F4>(o2), mdfl = mudf.ld<F4>(o2 - 4), mdfs = mudf.ld<F4>(o2 - ni);
  const F4 msfuy = s.msfuy.ld<F4>(o2), msfux = s.msfux.ld<F4>(o2), muu = s.muu.ld<F4>(o2);
  const F4 muc = mu.ld<F4>(o2), mul = mu.ld<F4>(o2 - 4), mus = mu.ld<F4>(o2 - ni);
  const F4 mvi = s.msfvx_inv.ld<F4>(o2), msfvy = s.msfvy.ld<F4>(o2), msfvx = s.msfvx.ld<F4>(o2), muv = s.muv.ld<F4>(o2);
  float mudf_xu[4], mu_m[4], mu0u[4], mudf_yv[4], mv_m[4];
#pragma unroll
  for (int n = 0; n < 4; n++) {
    mudf_xu[n] = -emdiv * dx * (mdf.v[n] - xm1(mdfl, mdf, n)) / msfuy.v[n];
    mu_m[n] = msfux.v[n] / msfuy.v[n];
    mu0u[n] = xm1(mul, muc, n);
    mudf_yv[n] = -emdiv * dy * (mdf.v[n] - mdfs.v[n]) * mvi.v[n];
    mv_m[n] = msfvy.v[n] / msfvx.v[n];
  }
  // dpn at the u points (from the groups c and l) and at the v points (c and s)
  F4 dpnu_lo, dpnv_lo, dpnu_top, dpnv_top;
  if (k0 == 1) {
    const Off q = ob + ni;
    const F4 c1 = p.ld<F4>(q), c2 = p.ld<F4>(q + ni), c3 = p.ld<F4>(q + 2 * ni);
    const F4 l1 = p.ld<F4>(q - 4), l2 = p.ld<F4>(q + ni - 4), l3 = p.ld<F4>(q + 2 * ni - 4);
    const F4 s1 = p.ld<F4>(q - sj), s2 = p.ld<F4>(q + ni - sj), s3 = p.ld<F4>(q + 2 * ni - sj);
#pragma unroll
    for (int n = 0; n < 4; n++) {
      dpnu_lo.v[n] = .5f * (cf1 * (c1.v[n] + xm1(l1, c1, n)) + cf2 * (c2.v[n] + xm1(l2, c2, n)) + cf3 * (c3.v[n] + xm1(l3, c3, n)));
      dpnv_lo.v[n] = .5f * (cf1 * (c1.v[n] + s1.v[n]) + cf2 * (c2.v[n] + s2.v[n]) + cf3 * (c3.v[n] + s3.v[n]));
    }
  } else {
    const Off q = ob + k0 * ni;
    const F4 c0 = p.ld<F4>(q), cm = p.ld<F4>(q - ni), l0 = p.ld<F4>(q - 4), lm = p.ld<F4>(q - ni - 4);
    const F4 s0 = p.ld<F4>(q - sj), sm = p.ld<F4>(q - ni - sj);
    const float fm = s.fnm[X1(d, k0)], fp = s.fnp[X1(d, k0)];
#pragma unroll
    for (int n = 0; n < 4; n++) {
      dpnu_lo.v[n] = .5f * (fm * (c0.v[n] + xm1(l0, c0, n)) + fp * (cm.v[n] + xm1(lm, cm, n)));
      dpnv_lo.v[n] = .5f * (fm * (c0.v[n] + s0.v[n]) + fp * (cm.v[n] + sm.v[n]));
    }
  }
#pragma unroll
  for (int n = 0; n < 4; n++) dpnu_top.v[n] = dpnv_top.v[n] = 0.f;
  if (c.top_lid && k1 == kde - 1) {
    const Off q = ob + (kde - 1) * ni;
    const F4 c1 = p.ld<F4>(q), c2 = p.ld<F4>(q - ni), c3 = p.ld<F4>(q - 2 * ni);
    const F4 l1 = p.ld<F4>(q - 4), l2 = p.ld<F4>(q - ni - 4), l3 = p.ld<F4>(q - 2 * ni - 4);
    const F4 s1 = p.ld<F4>(q - sj), s2 = p.ld<F4>(q - ni - sj), s3 = p.ld<F4>(q - 2 * ni - sj);
#pragma unroll
    for (int n = 0; n < 4; n++) {
      dpnu_top.v[n] = .5f * (cf1 * (xm1(l1, c1, n) + c1.v[n]) + cf2 * (xm1(l2, c2, n) + c2.v[n]) + cf3 * (xm1(l3, c3, n) + c3.v[n]));
      dpnv_top.v[n] = .5f * (cf1 * (s1.v[n] + c1.v[n]) + cf2 * (s2.v[n] + c2.v[n]) + cf3 * (s3.v[n] + c3.v[n]));
    }
  }
  auto lv = [&](int k) { return ob + imin(k, kde) * ni; };
  // p and ph of the level itself (centre, i-1 and j-1 groups), carried from the level below
  F4 pc = p.ld<F4>(lv(k0)), hc = ph.ld<F4>(lv(k0)), pl = p.ld<F4>(lv(k0) - 4), hl = ph.ld<F4>(lv(k0) - 4);
  F4 ps = p.ld<F4>(lv(k0) - sj), hs = ph.ld<F4>(lv(k0) - sj);
#pragma unroll 1
  for (int k = k0; k <= k1; k++) {
    const Off o = lv(k), op = lv(k + 1);
    const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)], rdnw = s.rdnw[X1(d, k)];
    const int kp = imin(k + 1, kde);
    const float fm = s.fnm[X1(d, kp)], fp = s.fnp[X1(d, kp)];
    const bool top = k + 1 == kde;
    const F4 pcp = p.ld<F4>(op), hcp = ph.ld<F4>(op);
    const F4 altc = alt.ld<F4>(o), alc = al.ld<F4>(o), pbc = pb.ld<F4>(o), phpc = php.ld<F4>(o);
    const F4 plp = p.ld<F4>(op - 4), hlp = ph.ld<F4>(op - 4), psp = p.ld<F4>(op - sj), hsp = ph.ld<F4>(op - sj);
    // u: neighbours at i-1
    {
      const F4 altl = alt.ld<F4>(o - 4), all = al.ld<F4>(o - 4), pbl = pb.ld<F4>(o - 4), phpl = php.ld<F4>(o - 4);
      const F4 u0 = s.u_2.ld<F4>(o), ut = s.ru_tend.ld<F4>(o), cqu = s.cqu.ld<F4>(o);
      F4 un;
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const int i = i0 + n;
        const bool utend_on = i >= i_start_u_tend && i <= i_end_u_tend, upg_on = i >= i_start_up && i <= i_end_up;
        float u = u0.v[n];
        u = utend_on ? u + dts * ut.v[n] : u;
        float dpxy = mu_m[n] * .5f * rdx * (c1 * muu.v[n] + c2) *
                     (((hcp.v[n] - xm1(hlp, hcp, n)) + (hc.v[n] - xm1(hl, hc, n))) +
                      (altc.v[n] + xm1(altl, altc, n)) * (pc.v[n] - xm1(pl, pc, n)) +
                      (alc.v[n] + xm1(all, alc, n)) * (pbc.v[n] - xm1(pbl, pbc, n)));
        const float g_hi = .5f * (fm * (pcp.v[n] + xm1(plp, pcp, n)) + fp * (pc.v[n] + xm1(pl, pc, n)));
        const float dpn_hi = top ? dpnu_top.v[n] : g_hi;
        const float dpxn = dpxy + mu_m[n] * rdx * (phpc.v[n] - xm1(phpl, phpc, n)) *
                                      (rdnw * (dpn_hi - dpnu_lo.v[n]) - .5f * ((c1 * mu0u[n]) + (c1 * muc.v[n])));
        dpxy = nh ? dpxn : dpxy;
        dpnu_lo.v[n] = dpn_hi;
        const float upg = u - dts * cqu.v[n] * dpxy + (c1 * mudf_xu[n]);
        un.v[n] = upg_on ? upg : u;
      }
      if (u_on) st4m(s.u_2, o, un, i0, ids, ide);
    }
    // v: neighbours at j-1
    {
      const F4 alts = alt.ld<F4>(o - sj), als = al.ld<F4>(o - sj), pbs = pb.ld<F4>(o - sj), phps = php.ld<F4>(o - sj);
      const F4 v0 = s.v_2.ld<F4>(o), vt = s.rv_tend.ld<F4>(o), cqv = s.cqv.ld<F4>(o);
      F4 vn;
#pragma unroll
      for (int n = 0; n < 4; n++) {
        float v = v0.v[n];
        v = v + dts * vt.v[n];
        float dpxy = mv_m[n] * .5f * rdy * (c1 * muv.v[n] + c2) *
                     (((hcp.v[n] - hsp.v[n]) + (hc.v[n] - hs.v[n])) +
                      (altc.v[n] + alts.v[n]) * (pc.v[n] - ps.v[n]) +
                      (alc.v[n] + als.v[n]) * (pbc.v[n] - pbs.v[n]));
        const float g_hi = .5f * (fm * (pcp.v[n] + psp.v[n]) + fp * (pc.v[n] + ps.v[n]));
        const float dpn_hi = top ? dpnv_top.v[n] : g_hi;
        const float dpyn = dpxy + mv_m[n] * rdy * (phpc.v[n] - phps.v[n]) *
                                      (rdnw * (dpn_hi - dpnv_lo.v[n]) - .5f * ((c1 * mus.v[n]) + (c1 * muc.v[n])));
        dpxy = nh ? dpyn : dpxy;
        dpnv_lo.v[n] = dpn_hi;
        const float vpg = v - dts * cqv.v[n] * dpxy + (c1 * mudf_yv[n]);
        vn.v[n] = vpg_on ? vpg : v;
      }
      if (v_on) st4m(s.v_2, o, vn, i0, i_start, i_end);
    }
    pc = pcp; hc = hcp; pl = plp; hl = hlp; ps = psp; hs = hsp;
  }
}

// ---------------------------------------------------------------------------------------
// The column part of one acoustic substep: advance_mu_t + advance_w + sumflux + calc_p_rho,
// split at the two column dependencies (the mass-divergence integral before the omega scan,
// and the tridiagonal forward/backward sweeps) into point kernels and two column kernels.
// Per substep, in order:
//   k_ac_div    (point)  dvdxi -> scr_dv; sumflux of ru_m, rv_m
//   k_ac_mu_ww  (column) dmdt, mu_2/mudf/muts/muave, omega scan -> ww
//   k_ac_theta  (point)  theta update, t_2ave -> t_2, t_2save; sumflux of ww_m
//   k_ac_wsolve (column) surface w, advance_w's rhs, explicit w predictor (levels 2..kde), calc_coef_w's
//                        coefficients, tridiagonal forward elimination and back
//                        substitution, ph, calc_p_rho -> w, ph, al, p, pm1 (gamma)
// Each point kernel only reads what earlier kernels wrote, so a thread never depends on
// another thread of its own launch.

// sumflux's running sums at (i, k, j) (module_small_step_em.F:1473-1633) for the x and y
// mass fluxes; ww_m at mass points is summed in k_ac_theta, once omega is final
DEVI void sumflux_uv(const KArgs &a, const DState &s, int i, int j, int k, Off o, Off o2) {
  const Dm &d = a.d;
  const bool mi = i <= d.ide - 1, mj = j <= d.jde - 1, kin = k <= d.kde - 1;
  const int it = a.iteration, ns = a.nsmall;
  const float nsf = (float)ns;
  float rum = it == 1 ? 0.f : s.ru_m[o];
  float rvm = it == 1 ? 0.f : s.rv_m[o];
  if (mj && kin) rum = rum + s.u_2[o];
  if (mi && kin) rvm = rvm + s.v_2[o];
  if (it == ns) {
    if (mj && kin) rum = rum / nsf + (s.c1h[X1(d, k)] * s.muu[o2] + s.c2h[X1(d, k)]) * s.u_save[o] / s.msfuy[o2];
    if (mi && kin) rvm = rvm / nsf + (s.c1h[X1(d, k)] * s.muv[o2] + s.c2h[X1(d, k)]) * s.v_save[o] * s.msfvx_inv[o2];
  }
  s.ru_m[o] = rum;
  s.rv_m[o] = rvm;
  if (!(mi && mj)) s.ww_m[o] = it == 1 ? 0.f : s.ww_m[o];
}

// dvdxi of advance_mu_t (module_small_step_em.F:1092-1099) at mass points, and the x/y parts
// of sumflux; one thread per (i, k, j) with i <= ide, j <= jde, k <= kde
template <typename ARGS>
__global__ void __launch_bounds__(256) k_ac_div(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  if (i > d.ide || j > d.jde || k > d.kde) return;
  const Off o2 = X2(d, i, j), o = X3(d, i, k, j);
  // per-routine entry points: advance_mu_t without sumflux, or sumflux alone
  if (a.routine != R_MU_T) sumflux_uv(a, s, i, j, k, o, o2);
  if (a.routine == R_SUMFLUX) return;
  if (i > d.ide - 1 || j > d.jde - 1 || k > d.kde - 1) return;
  const FPtr<float> u = s.u_2, u_1 = s.u_save, v = s.v_2, v_1 = s.v_save;
  const float msftx = s.msftx[o2], msfty = s.msfty[o2], rdx = s.rdx, rdy = s.rdy;
  const float muu0 = s.muu[o2], muu1 = s.muu[X2(d, i + 1, j)], muv0 = s.muv[o2], muv1 = s.muv[X2(d, i, j + 1)];
  const float msfuy0 = s.msfuy[o2], msfuy1 = s.msfuy[X2(d, i + 1, j)];
  const float mvi0 = s.msfvx_inv[o2], mvi1 = s.msfvx_inv[X2(d, i, j + 1)];
  const float cc = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
  a.scr_dv[o] = msftx * msfty *
                (rdy * ((v[o + d.sj] + (cc * muv1 + c2) * v_1[o + d.sj] * mvi1) - (v[o] + (cc * muv0 + c2) * v_1[o] * mvi0)) +
                 rdx * ((u[o + 1] + (cc * muu1 + c2) * u_1[o + 1] / msfuy1) - (u[o] + (cc * muu0 + c2) * u_1[o] / msfuy0)));
}

// advance_mu_t's column integral, mass update and omega scan (module_small_step_em.F:1100-1121)
template <typename ARGS, bool COMP>
__global__ void __launch_bounds__(256) k_ac_mu_ww(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = col_ia(d, !std::is_same<ARGS, KArgs>::value) + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  if (i < d.ids || i > d.ide - 1 || j > d.jde - 1) return;
  if (spec_frame(a.c, d, i, j)) return;
  const int k_end = d.kde - 1;
  const Off o2 = X2(d, i, j);
  const Off ob = X3(d, i, 0, j);  // level k of the column is at ob + k*ni
  const FPtr<float> dvp = a.scr_dv, ww1 = s.ww1;
  const CPtr<float> dnw = s.dnw, c1h = s.c1h;
  FPtr<float> ww = s.ww;
  const int ni = d.ni, x1 = -d.kms;
  const float dts = a.dts, epssm = a.c.epssm, msfty = s.msfty[o2];
  float dmdt = 0.f, comp = 0.f;
  column_march_up<1, float>(1, k_end, [&](int, int k) { return dvp[ob + k * ni]; }, [&](int k, const float *v) {
    const float x = dnw[k + x1] * v[0];
    if (COMP) two_sum_acc(dmdt, comp, x);
    else dmdt = dmdt + x;
  });
  if (COMP) dmdt = dmdt + comp;
  const float mu_tend = s.mu_tend[o2];
  float mu = s.mu_2[o2];
  float muave = mu;
  mu = mu + dts * (dmdt + mu_tend);
  s.mu_2[o2] = mu;
  s.mudf[o2] = (dmdt + mu_tend);
  s.muts[o2] = s.mut[o2] + mu;
  muave = .5f * ((1.f + epssm) * mu + (1.f - epssm) * muave);
  s.muave[o2] = muave;
  // scan: level kk reads dvdxi(kk-1); fields: 0 = dvdxi(kk-1), 1 = ww1(kk)
  float prev = ww[ob + ni];
  ww[ob + ni] = prev - ww1[ob + ni];
  column_march_up<2, float>(
      2, k_end, [&](int f, int kk) { return f == 0 ? dvp[ob + (kk - 1) * ni] : ww1[ob + kk * ni]; },
      [&](int kk, const float *v) {
        const int k = kk - 1;
        const float cur = prev - dnw[kk - 1 + x1] * ((c1h[k + x1] * dmdt) + v[0] + (c1h[k + x1] * mu_tend)) / msfty;
        prev = cur;
        ww[ob + kk * ni] = cur - v[1];
      });
}

// theta update of advance_mu_t (1125-1174), t_2ave of advance_w (1312-1320) and the ww_m
// part of sumflux; in the per-routine forms (KArgs::routine) one of them
template <typename ARGS>
__global__ void __launch_bounds__(256) k_ac_theta(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  const int kde = d.kde, k_end = kde - 1;
  if (i > d.ide - 1 || j > d.jde - 1 || k > kde) return;
  const wrfmi_config &c = a.c;
  const Off o2 = X2(d, i, j), o = X3(d, i, k, j);
  const float dts = a.dts, epssm = c.epssm;
  const float rdx = s.rdx, rdy = s.rdy;
  const FPtr<float> ww = s.ww, t_1 = s.t_save;
  const int routine = a.routine;
  if (routine == R_FUSED || routine == R_SUMFLUX) {  // sumflux of omega (module_small_step_em.F:1473-1633) at this mass point
    const int it = a.iteration, ns = a.nsmall;
    float wwm = it == 1 ? 0.f : s.ww_m[o];
    wwm = wwm + ww[o];
    if (it == ns) wwm = wwm / (float)ns + s.ww1[o];
    s.ww_m[o] = wwm;
  }
  if (routine == R_W_TAVE && k <= k_end && !spec_frame(c, d, i, j)) {
    // advance_w alone: t_2ave from the updated theta and advance_mu_t's t_ave (1312-1320),
    // the expressions of the fused form below
    const float t = s.t_2[o], tv = s.t_2save[o];
    const float muave = s.muave[o2], muts = s.muts[o2];
    float v1 = .5f * ((1.f + epssm) * t + (1.f - epssm) * tv);
    v1 = (v1 + (s.c1h[X1(d, k)] * muave) * K_T0) / ((s.c1h[X1(d, k)] * muts + s.c2h[X1(d, k)]) * (K_T0 + t_1[o]));
    s.t_2save[o] = v1;
  }
  if ((routine == R_FUSED || routine == R_MU_T) && k <= k_end && !spec_frame(c, d, i, j)) {
    const float msftx = s.msftx[o2], msfty = s.msfty[o2];
    const FPtr<float> u = s.u_2, v = s.v_2;
    const float tv = s.t_2[o];
    float t = tv + msfty * dts * s.t_tend[o];
    const float wdtn_hi = (k + 1 <= k_end) ? ww[o + d.ni] * (s.fnm[X1(d, k + 1)] * t_1[o + d.ni] + s.fnp[X1(d, k + 1)] * t_1[o]) : 0.f;
    const float wdtn_lo = (k >= 2) ? ww[o] * (s.fnm[X1(d, k)] * t_1[o] + s.fnp[X1(d, k)] * t_1[o - d.ni]) : 0.f;
    t = t - dts * msfty *
                (msftx * (.5f * rdy * (v[o + d.sj] * (t_1[o + d.sj] + t_1[o]) - v[o] * (t_1[o] + t_1[o - d.sj])) +
                          .5f * rdx * (u[o + 1] * (t_1[o + 1] + t_1[o]) - u[o] * (t_1[o] + t_1[o - 1]))) +
                 s.rdnw[X1(d, k)] * (wdtn_hi - wdtn_lo));
    s.t_2[o] = t;
    if (c.non_hydrostatic && routine == R_FUSED) {
      // t_2ave (advance_w 1312-1320): the old theta is t_2save's value from advance_mu_t
      const float muave = s.muave[o2], muts = s.muts[o2];
      float v1 = .5f * ((1.f + epssm) * t + (1.f - epssm) * tv);
      v1 = (v1 + (s.c1h[X1(d, k)] * muave) * K_T0) / ((s.c1h[X1(d, k)] * muts + s.c2h[X1(d, k)]) * (K_T0 + t_1[o]));
      s.t_2save[o] = v1;
    } else {
      s.t_2save[o] = tv;
    }
  }
}


// calc_mu_uv_1 + small_step_finish with four i points per lane (16-byte accesses; element by
// element the expressions of k_small_finish)
template <typename ARGS>
__global__ void __launch_bounds__(256) k_small_finish_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  const int ids = d.ids, ide = d.ide, jds = d.jds, jde = d.jde, kde = d.kde;
  if (i0 > ide || j > jde || k > kde) return;
  const bool mj = j <= jde - 1, kin = k <= kde - 1;
  const int kc = imin(k, kde - 1);
  const float c1 = s.c1h[X1(d, kc)], c2 = s.c2h[X1(d, kc)], c1f = s.c1f[X1(d, k)], c2f = s.c2f[X1(d, k)];
  const Off o2 = X2(d, i0, j), o = X3(d, i0, k, j);
  const bool top = k == d.kds;
  const F4 mts = s.muts.ld<F4>(o2), mtsl = s.muts.ld<F4>(o2 - 4), mtss = s.muts.ld<F4>(o2 - d.ni);
  const bool px = a.c.periodic_x, py = a.c.periodic_y;
  F4 muus, muvs;
#pragma unroll
  for (int n = 0; n < 4; n++) {
    const int i = i0 + n;
    const float c = mts.v[n], l = xm1(mtsl, mts, n);
    if (i > ids && i < ide) muus.v[n] = 0.5f * (c + l);
    else if (i == ids) muus.v[n] = 0.5f * (c + (px ? l : c));
    else muus.v[n] = 0.5f * (l + (px ? c : l));
    const float sv = mtss.v[n];
    if (j > jds && j < jde) muvs.v[n] = 0.5f * (c + sv);
    else if (j == jds) muvs.v[n] = 0.5f * (c + (py ? sv : c));
    else muvs.v[n] = 0.5f * (sv + (py ? c : sv));
  }
  if (top) {
    if (mj) st4m(s.muus, o2, muus, i0, ids, ide);
    st4m(s.muvs, o2, muvs, i0, ids, ide - 1);
  }
  if (kin) {
    const F4 muv = s.muv.ld<F4>(o2), msfvx = s.msfvx.ld<F4>(o2), v2 = s.v_2.ld<F4>(o), vs = s.v_save.ld<F4>(o);
    F4 vn;
#pragma unroll
    for (int n = 0; n < 4; n++) vn.v[n] = (msfvx.v[n] * v2.v[n] + vs.v[n] * (c1 * muv.v[n] + c2)) / (c1 * muvs.v[n] + c2);
    st4m(s.v_2, o, vn, i0, ids, ide - 1);
  }
  if (mj && kin) {
    const F4 muu = s.muu.ld<F4>(o2), msfuy = s.msfuy.ld<F4>(o2), u2 = s.u_2.ld<F4>(o), us = s.u_save.ld<F4>(o);
    F4 un;
#pragma unroll
    for (int n = 0; n < 4; n++) un.v[n] = (msfuy.v[n] * u2.v[n] + us.v[n] * (c1 * muu.v[n] + c2)) / (c1 * muus.v[n] + c2);
    st4m(s.u_2, o, un, i0, ids, ide);
  }
  if (!mj || i0 > ide - 1) return;
  const F4 mut = s.mut.ld<F4>(o2), msfty = s.msfty.ld<F4>(o2);
  {
    const F4 w2 = s.w_2.ld<F4>(o), ws = s.w_save.ld<F4>(o), ph2 = s.ph_2.ld<F4>(o), phs = s.ph_save.ld<F4>(o);
    const F4 ww = s.ww.ld<F4>(o), ww1 = s.ww1.ld<F4>(o);
    F4 wn, phn, wwn;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      wn.v[n] = (msfty.v[n] * w2.v[n] + ws.v[n] * (c1f * mut.v[n] + c2f)) / (c1f * mts.v[n] + c2f);
      phn.v[n] = ph2.v[n] + phs.v[n];
      wwn.v[n] = ww.v[n] + ww1.v[n];
    }
    st4m(s.w_2, o, wn, i0, ids, ide - 1);
    st4m(s.ph_2, o, phn, i0, ids, ide - 1);
    st4m(s.ww, o, wwn, i0, ids, ide - 1);
  }
  if (kin) {
    const bool last = a.rk_step >= a.c.rk_ord;
    // h_diabatic enters at the last RK substep only; before that the load re-reads t_save
    const F4 t2 = s.t_2.ld<F4>(o), tsv = s.t_save.ld<F4>(o), hd = (last ? s.h_diabatic : s.t_save).ld<F4>(o);
    F4 tn;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const float cm = c1 * mut.v[n] + c2;
      const float nl = (t2.v[n] + tsv.v[n] * cm) / (c1 * mts.v[n] + c2);
      const float lt = (t2.v[n] - a.dts * (float)a.nsmall * cm * hd.v[n] + tsv.v[n] * cm) / (c1 * mts.v[n] + c2);
      tn.v[n] = last ? lt : nl;
    }
    st4m(s.t_2, o, tn, i0, ids, ide - 1);
  }
  if (top) {
    const F4 mu2 = s.mu_2.ld<F4>(o2), mus = s.mu_save.ld<F4>(o2);
    F4 mn;
#pragma unroll
    for (int n = 0; n < 4; n++) mn.v[n] = mu2.v[n] + mus.v[n];
    st4m(s.mu_2, o2, mn, i0, ids, ide - 1);
  }
}


// ---------------------------------------------------------------------------------------
// Four-points-per-lane forms of small_step_prep and calc_p_rho step 0 (16-byte accesses;
// element by element the expressions of k_small_prep and k_p_rho0).
struct MuS4 {
  F4 muts, muus, muvs, mu_new;
};
DEVI MuS4 small_mu4(const KArgs &a, int i0, int j) {
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const Off o2 = X2(d, i0, j);
  const F4 mub = s.mub.ld<F4>(o2), mu1 = s.mu_1.ld<F4>(o2), mu2 = s.mu_2.ld<F4>(o2);
  const F4 mubl = s.mub.ld<F4>(o2 - 4), mu1l = s.mu_1.ld<F4>(o2 - 4), mubs = s.mub.ld<F4>(o2 - d.ni), mu1s = s.mu_1.ld<F4>(o2 - d.ni);
  const F4 muu = s.muu.ld<F4>(o2), muv = s.muv.ld<F4>(o2);
  const bool rk1 = a.rk_step == 1;
  MuS4 m;
#pragma unroll
  for (int n = 0; n < 4; n++) {
    m.muts.v[n] = rk1 ? mub.v[n] + mu2.v[n] : mub.v[n] + mu1.v[n];
    m.muus.v[n] = rk1 ? muu.v[n] : 0.5f * (((mub.v[n] + mu1.v[n]) + xm1(mubl, mub, n)) + xm1(mu1l, mu1, n));
    m.muvs.v[n] = rk1 ? muv.v[n] : 0.5f * (((mub.v[n] + mu1.v[n]) + mubs.v[n]) + mu1s.v[n]);
    m.mu_new.v[n] = rk1 ? 0.f : mu1.v[n] - mu2.v[n];
  }
  return m;
}

template <typename ARGS>
__global__ void __launch_bounds__(256) k_small_prep_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  const int ids = d.ids, ide = d.ide, jde = d.jde, kde = d.kde, k_end = kde - 1;
  if (i0 > ide || j > jde || k > kde) return;
  const bool mj = j <= jde - 1, kin = k <= k_end;
  const int kc = imin(k, k_end);
  const float c1 = s.c1h[X1(d, kc)], c2 = s.c2h[X1(d, kc)], c1f = s.c1f[X1(d, k)], c2f = s.c2f[X1(d, k)];
  const Off o2 = X2(d, i0, j), o = X3(d, i0, k, j);
  const bool rk1 = a.rk_step == 1;
  const MuS4 m = small_mu4(a, i0, j);
  // at rk_step 1 the *_1 time level is the *_2 one (then saved): its loads read *_2 again,
  // which the cache serves, instead of a second array from HBM
  if (mj && kin) {
    const F4 u2 = s.u_2.ld<F4>(o), u1o = (rk1 ? s.u_2 : s.u_1).ld<F4>(o), muu = s.muu.ld<F4>(o2), msfuy = s.msfuy.ld<F4>(o2);
    F4 un;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const float u1 = rk1 ? u2.v[n] : u1o.v[n];
      un.v[n] = ((c1 * m.muus.v[n] + c2) * u1 - (c1 * muu.v[n] + c2) * u2.v[n]) / msfuy.v[n];
    }
    if (rk1) st4m(s.u_1, o, u2, i0, ids, ide);
    st4m(s.u_save, o, u2, i0, ids, ide);
    st4m(s.u_2, o, un, i0, ids, ide);
  }
  if (kin) {
    const F4 v2 = s.v_2.ld<F4>(o), v1o = (rk1 ? s.v_2 : s.v_1).ld<F4>(o), muv = s.muv.ld<F4>(o2), mvi = s.msfvx_inv.ld<F4>(o2);
    F4 vn;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const float v1 = rk1 ? v2.v[n] : v1o.v[n];
      vn.v[n] = ((c1 * m.muvs.v[n] + c2) * v1 - (c1 * muv.v[n] + c2) * v2.v[n]) * mvi.v[n];
    }
    if (rk1) st4m(s.v_1, o, v2, i0, ids, ide - 1);
    st4m(s.v_save, o, v2, i0, ids, ide - 1);
    st4m(s.v_2, o, vn, i0, ids, ide - 1);
  }
  if (!mj || i0 > ide - 1) return;
  const F4 mut = s.mut.ld<F4>(o2), msfty = s.msfty.ld<F4>(o2);
  if (kin) {
    const F4 pb = s.pb.ld<F4>(o), pp = s.p.ld<F4>(o), alt = s.alt.ld<F4>(o);
    const F4 t2 = s.t_2.ld<F4>(o), t1o = (rk1 ? s.t_2 : s.t_1).ld<F4>(o);
    F4 c2a, tn;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      c2a.v[n] = K_CPOVCV * (pb.v[n] + pp.v[n]) / alt.v[n];
      const float t1 = rk1 ? t2.v[n] : t1o.v[n];
      tn.v[n] = (c1 * m.muts.v[n] + c2) * t1 - (c1 * mut.v[n] + c2) * t2.v[n];
    }
    st4m(s.c2a, o, c2a, i0, ids, ide - 1);
    if (rk1) st4m(s.t_1, o, t2, i0, ids, ide - 1);
    st4m(s.t_save, o, t2, i0, ids, ide - 1);
    st4m(s.t_2, o, tn, i0, ids, ide - 1);
  }
  {
    const F4 w2 = s.w_2.ld<F4>(o), ph2 = s.ph_2.ld<F4>(o), w1o = (rk1 ? s.w_2 : s.w_1).ld<F4>(o), ph1o = (rk1 ? s.ph_2 : s.ph_1).ld<F4>(o), ww = s.ww.ld<F4>(o);
    F4 wn, phn;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const float w1 = rk1 ? w2.v[n] : w1o.v[n];
      const float ph1 = rk1 ? ph2.v[n] : ph1o.v[n];
      wn.v[n] = ((c1f * m.muts.v[n] + c2f) * w1 - (c1f * mut.v[n] + c2f) * w2.v[n]) / msfty.v[n];
      phn.v[n] = ph1 - ph2.v[n];
    }
    if (rk1) {
      st4m(s.w_1, o, w2, i0, ids, ide - 1);
      st4m(s.ph_1, o, ph2, i0, ids, ide - 1);
    }
    st4m(s.w_save, o, w2, i0, ids, ide - 1);
    st4m(s.w_2, o, wn, i0, ids, ide - 1);
    st4m(s.ph_save, o, ph2, i0, ids, ide - 1);
    st4m(s.ph_2, o, phn, i0, ids, ide - 1);
    st4m(s.ww1, o, ww, i0, ids, ide - 1);
  }
}


template <typename ARGS>
__global__ void __launch_bounds__(256) k_p_rho0_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  if (i0 > d.ide - 1 || j > d.jde - 1 || k > d.kde - 1) return;
  if (!a.c.non_hydrostatic) return;
  const Off o = X3(d, i0, k, j);
  const MuS4 m = small_mu4(a, i0, j);
  const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)], rdnw = s.rdnw[X1(d, k)];
  const F4 alt = s.alt.ld<F4>(o), php = s.ph_2.ld<F4>(o + d.ni), ph = s.ph_2.ld<F4>(o), t1 = s.t_save.ld<F4>(o);
  const F4 t2 = s.t_2.ld<F4>(o), c2a = s.c2a.ld<F4>(o);
  F4 aln, pn;
#pragma unroll
  for (int n = 0; n < 4; n++) {
    const float mu = m.mu_new.v[n], muts = m.muts.v[n];
    const float al = -1.f / (c1 * muts + c2) * (alt.v[n] * (c1 * mu) + rdnw * (php.v[n] - ph.v[n]));
    aln.v[n] = al;
    pn.v[n] = c2a.v[n] * (alt.v[n] * (t2.v[n] - (c1 * mu) * t1.v[n]) / ((c1 * muts + c2) * (K_T0 + t1.v[n])) - al);
  }
  st4m(s.al, o, aln, i0, d.ids, d.ide - 1);
  st4m(s.p, o, pn, i0, d.ids, d.ide - 1);
  st4m(s.pm1, o, pn, i0, d.ids, d.ide - 1);
}

// ---------------------------------------------------------------------------------------
// Four-points-per-lane forms of k_ac_div and k_ac_theta (16-byte loads and
// stores; used when the host finds the layout aligned).  Element by element they evaluate
// the same expressions in the same order as the one-point kernels above; i neighbours come
// from the groups of four on either side.

template <typename ARGS, int KS>
__global__ void __launch_bounds__(256) k_ac_div_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k0 = d.kds + (int)T.bz * KS;
  const int ids = d.ids, ide = d.ide, jde = d.jde, kde = d.kde;
  if (i0 > ide || j > jde || k0 > kde) return;
  const int k1 = imin(k0 + KS - 1, kde);
  const Off o2 = X2(d, i0, j), ob = X3(d, i0, 0, j);  // level k at ob + k*ni
  const int ni = d.ni;
  const bool mj = j <= jde - 1;
  const int it = a.iteration, ns = a.nsmall;
  const float nsf = (float)ns;
  const float rdx = s.rdx, rdy = s.rdy;
  // the column's 2-D fields, loaded once for its KS levels
  const F4 muu = s.muu.ld<F4>(o2), msfuy = s.msfuy.ld<F4>(o2), muv = s.muv.ld<F4>(o2), mvi = s.msfvx_inv.ld<F4>(o2);
  const F4 muur = s.muu.ld<F4>(o2 + 4), msfuyr = s.msfuy.ld<F4>(o2 + 4);
  const F4 muvn = s.muv.ld<F4>(o2 + ni), mvin = s.msfvx_inv.ld<F4>(o2 + ni);
  const F4 msftx = s.msftx.ld<F4>(o2), msfty = s.msfty.ld<F4>(o2);
#pragma unroll 1
  for (int k = k0; k <= k1; k++) {
    const Off o = ob + k * ni;
    const bool kin = k <= kde - 1;
    const float c1 = s.c1h[X1(d, imin(k, kde - 1))], c2 = s.c2h[X1(d, imin(k, kde - 1))];
    const F4 u = s.u_2.ld<F4>(o), v = s.v_2.ld<F4>(o), us = s.u_save.ld<F4>(o), vs = s.v_save.ld<F4>(o);
    const F4 rum0 = s.ru_m.ld<F4>(o), rvm0 = s.rv_m.ld<F4>(o);
    // sumflux of ru_m, rv_m (and ww_m outside the mass points), as sumflux_uv
    F4 rum, rvm;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const bool mi = i0 + n <= ide - 1;
      float ru_ = it == 1 ? 0.f : rum0.v[n];
      float rv_ = it == 1 ? 0.f : rvm0.v[n];
      if (mj && kin) ru_ = ru_ + u.v[n];
      if (mi && kin) rv_ = rv_ + v.v[n];
      if (it == ns) {
        if (mj && kin) ru_ = ru_ / nsf + (c1 * muu.v[n] + c2) * us.v[n] / msfuy.v[n];
        if (mi && kin) rv_ = rv_ / nsf + (c1 * muv.v[n] + c2) * vs.v[n] * mvi.v[n];
      }
      rum.v[n] = ru_;
      rvm.v[n] = rv_;
    }
    st4m(s.ru_m, o, rum, i0, ids, ide);
    st4m(s.rv_m, o, rvm, i0, ids, ide);
    if (it == 1) {
#pragma unroll
      for (int n = 0; n < 4; n++)
        if (i0 + n >= ids && i0 + n <= ide && !(i0 + n <= ide - 1 && mj)) s.ww_m[o + n] = 0.f;
    }
    if (!mj || !kin) continue;
    // dvdxi at the mass points
    const F4 ur = s.u_2.ld<F4>(o + 4), usr = s.u_save.ld<F4>(o + 4);
    const F4 vn = s.v_2.ld<F4>(o + d.sj), vsn = s.v_save.ld<F4>(o + d.sj);
    F4 dv;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const float u1 = xp1(u, ur, n), us1 = xp1(us, usr, n), muu1 = xp1(muu, muur, n), msfuy1 = xp1(msfuy, msfuyr, n);
      dv.v[n] = msftx.v[n] * msfty.v[n] *
                (rdy * ((vn.v[n] + (c1 * muvn.v[n] + c2) * vsn.v[n] * mvin.v[n]) - (v.v[n] + (c1 * muv.v[n] + c2) * vs.v[n] * mvi.v[n])) +
                 rdx * ((u1 + (c1 * muu1 + c2) * us1 / msfuy1) - (u.v[n] + (c1 * muu.v[n] + c2) * us.v[n] / msfuy.v[n])));
    }
    st4m(a.scr_dv, o, dv, i0, ids, ide - 1);
  }
}


template <typename ARGS>
__global__ void __launch_bounds__(256) k_ac_theta_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  const int ids = d.ids, ide = d.ide, jde = d.jde, kde = d.kde, k_end = kde - 1;
  if (i0 > ide - 1 || j > jde - 1 || k > kde) return;
  const wrfmi_config &c = a.c;
  const Off o2 = X2(d, i0, j), o = X3(d, i0, k, j);
  const int ni = d.ni, sj = d.sj;
  const int kp = imin(k + 1, kde), km = imax(k - 1, d.kds);  // clamped neighbour levels
  const Off op = o + (kp - k) * ni, om = o - (k - km) * ni;
  const float dts = a.dts, epssm = c.epssm, rdx = s.rdx, rdy = s.rdy;
  const F4 msftx = s.msftx.ld<F4>(o2), msfty = s.msfty.ld<F4>(o2);
  const F4 ww = s.ww.ld<F4>(o);
  {  // sumflux of omega at the mass points
    const int it = a.iteration, ns = a.nsmall;
    const F4 wwm0 = s.ww_m.ld<F4>(o), ww1 = s.ww1.ld<F4>(o);
    F4 wwm;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      float x = it == 1 ? 0.f : wwm0.v[n];
      x = x + ww.v[n];
      if (it == ns) x = x / (float)ns + ww1.v[n];
      wwm.v[n] = x;
    }
    st4m(s.ww_m, o, wwm, i0, ids, ide - 1);
  }
  const F4 t1 = s.t_save.ld<F4>(o);
  if (k <= k_end) {
    const F4 tv = s.t_2.ld<F4>(o), tt = s.t_tend.ld<F4>(o);
    const F4 wwp = s.ww.ld<F4>(op), t1p = s.t_save.ld<F4>(op), t1m = s.t_save.ld<F4>(om);
    const F4 t1l = s.t_save.ld<F4>(o - 4), t1r = s.t_save.ld<F4>(o + 4);
    const F4 t1n = s.t_save.ld<F4>(o + sj), t1s = s.t_save.ld<F4>(o - sj);
    const F4 u = s.u_2.ld<F4>(o), ur = s.u_2.ld<F4>(o + 4), v = s.v_2.ld<F4>(o), vn = s.v_2.ld<F4>(o + sj);
    const float fnm_p = s.fnm[X1(d, kp)], fnp_p = s.fnp[X1(d, kp)], fnm_0 = s.fnm[X1(d, k)], fnp_0 = s.fnp[X1(d, k)];
    const float rdnw = s.rdnw[X1(d, k)], c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
    F4 tn, t2s;
    const F4 muave = s.muave.ld<F4>(o2), muts = s.muts.ld<F4>(o2);
#pragma unroll
    for (int n = 0; n < 4; n++) {
      float t = tv.v[n] + msfty.v[n] * dts * tt.v[n];
      const float wdtn_hi = (k + 1 <= k_end) ? wwp.v[n] * (fnm_p * t1p.v[n] + fnp_p * t1.v[n]) : 0.f;
      const float wdtn_lo = (k >= 2) ? ww.v[n] * (fnm_0 * t1.v[n] + fnp_0 * t1m.v[n]) : 0.f;
      t = t - dts * msfty.v[n] *
                  (msftx.v[n] * (.5f * rdy * (vn.v[n] * (t1n.v[n] + t1.v[n]) - v.v[n] * (t1.v[n] + t1s.v[n])) +
                                 .5f * rdx * (xp1(u, ur, n) * (xp1(t1, t1r, n) + t1.v[n]) - u.v[n] * (t1.v[n] + xm1(t1l, t1, n)))) +
                   rdnw * (wdtn_hi - wdtn_lo));
      tn.v[n] = t;
      float v1 = .5f * ((1.f + epssm) * t + (1.f - epssm) * tv.v[n]);
      v1 = (v1 + (c1 * muave.v[n]) * K_T0) / ((c1 * muts.v[n] + c2) * (K_T0 + t1.v[n]));
      t2s.v[n] = c.non_hydrostatic ? v1 : tv.v[n];
    }
    // a specified domain's outermost rows and columns are spec_bdyupdate's
    const bool spec = c.specified || c.nested, sx = spec && !c.periodic_x;
    if (!(spec && (j == d.jds || j == jde - 1))) {
      st4m(s.t_2, o, tn, i0, sx ? ids + 1 : ids, sx ? ide - 2 : ide - 1);
      st4m(s.t_2save, o, t2s, i0, sx ? ids + 1 : ids, sx ? ide - 2 : ide - 1);
    }
  }
}


// Per column: the explicit w predictor (advance_w 1402-1431), the tridiagonal forward
// elimination and back substitution (1433-1443), ph from rhs and the final w (1460-1465)
// and calc_p_rho for step >= 1
// (438-568).  CT = double keeps the forward value in FP64, its low-order part parked in scr_lo.
// levels of loads in flight ahead of each march of k_ac_wsolve (2, 4, 6 and 8 time within
// 2 % of each other at 400x400x80: the kernel is bound by the bytes it moves)
template <typename ARGS, typename CT, bool PRHO, int PF>
__global__ void __launch_bounds__(256) k_ac_wsolve(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = col_ia(d, !std::is_same<ARGS, KArgs>::value) + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int kde = d.kde, k_end = kde - 1;
  if (i < d.ids || i > d.ide - 1 || j > d.jde - 1) return;
  const wrfmi_config &c = a.c;
  if (!c.non_hydrostatic) return;
  // a specified domain's outermost columns: spec_bdyupdate_ph and calc_p_rho in k_lbc
  if (spec_frame(c, d, i, j)) return;
  FPtr<float> w = s.w_2, ph = s.ph_2;
  FPtr<float> lo = a.scr_lo;
  const FPtr<float> rhs = a.scr_rhs;
  FPtr<CT> GA((CT *)s.gamma);
  const Off ob = X3(d, i, 0, j), o2 = X2(d, i, j);
  const int ni = d.ni;
  // advance_w's lower boundary (1383-1395): w(1) from u, v extrapolated to the surface by
  // cf1..cf3 and the terrain slopes on both sides of the point
  float wsfc;
  {
    const FPtr<float> u = s.u_2, v = s.v_2, ht = s.ht;
    const float cf1 = s.cf1, cf2 = s.cf2, cf3 = s.cf3, msftx = s.msftx[o2], msfty_s = s.msfty[o2];
    wsfc = msfty_s * .5f * s.rdy *
               ((ht[X2(d, i, j + 1)] - ht[o2]) * (cf1 * v[X3(d, i, 1, j + 1)] + cf2 * v[X3(d, i, 2, j + 1)] + cf3 * v[X3(d, i, 3, j + 1)]) +
                (ht[o2] - ht[X2(d, i, j - 1)]) * (cf1 * v[X3(d, i, 1, j)] + cf2 * v[X3(d, i, 2, j)] + cf3 * v[X3(d, i, 3, j)])) +
           msftx * .5f * s.rdx *
               ((ht[X2(d, i + 1, j)] - ht[o2]) * (cf1 * u[X3(d, i + 1, 1, j)] + cf2 * u[X3(d, i + 1, 2, j)] + cf3 * u[X3(d, i + 1, 3, j)]) +
                (ht[o2] - ht[X2(d, i - 1, j)]) * (cf1 * u[X3(d, i, 1, j)] + cf2 * u[X3(d, i, 2, j)] + cf3 * u[X3(d, i, 3, j)]));
    w[ob + ni] = wsfc;
  }
  CT wfe = (CT)wsfc;
  // Upward: advance_w's right-hand side rhs(k+1) (1324-1371: phi advection by omega, the
  // g*w term, rhs = ph + msfty*rhs/(c1f*mut+c2f); stored for the way down), the explicit
  // w predictor (1402-1431) at level k, calc_coef_w's coefficients a(k), alpha(k),
  // gamma(k) (570-652, the same expressions in the same order as k_coef_w; gamma stored
  // for the way down), then the forward elimination.  Fields at level k: 0 = t_2ave,
  // 1 = c2a, 2 = alt, 3 = rw_tend, 4 = cqw; at k+1 (clamped to kde): 5 = w, 6 = ph,
  // 7 = ph_tend; at k+2 (clamped): 8 = ph_save, 9 = phb, 10 = omega.  w at k, rhs and ph at
  // k and k-1, t_2ave, c2a and alt at k-1 and ph_save, phb and omega at k and k+1 are
  // carried from the levels below.
  const float dts = a.dts, epssm = c.epssm;
  const FPtr<float> ph_c = s.ph_2, t2ave = s.t_2save, c2a = s.c2a, alt_f = s.alt;
  const FPtr<float> ph1f = s.ph_save, phbf = s.phb, wwf = s.ww;
  const float msfty = s.msfty[o2], mut = s.mut[o2], muave = s.muave[o2];
  const CPtr<float> c1h = s.c1h, c2h = s.c2h, c1f = s.c1f, c2f = s.c2f, rdnw = s.rdnw, rdn = s.rdn;
  const CPtr<float> fnm = s.fnm, fnp = s.fnp;
  const float msft_inv = 1.f / msfty;
  const bool z2 = c.phi_adv_z == 2;
  // rhs(kk) for kk = 2..kde from ph_save, phb, omega at kk-1, kk, min(kk+1, kde)
  auto rhs_at = [&](int kk, float w2, float pht, float ph2, float ph1m, float ph1, float ph1p, float pbm, float pb0,
                    float pbp, float wwm, float ww, float wwp) -> float {
    const float rdnw_k = rdnw[X1(d, imin(kk, k_end))], rdnw_m = rdnw[X1(d, kk - 1)];
    const float dph_hi = ph1p - ph1 + pbp - pb0;
    const float dph_lo = ph1 - ph1m + pb0 - pbm;
    const float wd_hi = z2 ? rdnw_k * dph_hi : .5f * (wwp + ww) * rdnw_k * dph_hi;
    const float wd_lo = z2 ? rdnw_m * dph_lo : .5f * (ww + wwm) * rdnw_m * dph_lo;
    float r = dts * (pht + .5f * K_G * (1.f - epssm) * w2);
    if (kk <= k_end) {
      if (z2) r = r - dts * ww * (fnm[X1(d, kk)] * wd_hi + fnp[X1(d, kk)] * wd_lo);
      else r = r - dts * (fnm[X1(d, kk)] * wd_hi + fnp[X1(d, kk)] * wd_lo);
    }
    r = ph2 + msfty * r / (c1f[X1(d, kk)] * mut + c2f[X1(d, kk)]);
    if (c.top_lid && kk == k_end + 1) r = 0.f;
    return r;
  };
  const Off o1 = ob + ni, o2l = ob + 2 * ni, o3 = ob + imin(3, kde) * ni;
  float w_k = w[o2l];  // w at the march's level (before the predictor)
  float ph1_k = ph1f[o2l], ph1_k1 = ph1f[o3], pb_k = phbf[o2l], pb_k1 = phbf[o3], ww_k = wwf[o2l], ww_k1 = wwf[o3];
  float rhs_m1 = 0.f;  // rhs(1) = 0
  float rhs_0 = rhs_at(2, w_k, s.ph_tend[o2l], ph_c[o2l], ph1f[o1], ph1_k, ph1_k1, phbf[o1], pb_k, pb_k1, wwf[o1], ww_k, ww_k1);
  rhs[o2l] = rhs_0;
  float phl_m1 = ph_c[o1], phl_0 = ph_c[o2l];
  float t2a_lo = t2ave[o1], c2a_lo = c2a[o1], alt_lo = alt_f[o1];
  const CT lid = c.top_lid ? CT(0) : CT(1);
  const CT epc = (CT)c.epssm;
  CT cof = (CT(.5) * (CT)a.dts * (CT)K_G * (CT(1) + epc));
  cof = cof * cof;
  const CT MUT = (CT)mut;
  auto cH = [&](int k) -> CT { return (CT)c1h[X1(d, k)] * MUT + (CT)c2h[X1(d, k)]; };
  auto cF = [&](int k) -> CT { return (CT)c1f[X1(d, k)] * MUT + (CT)c2f[X1(d, k)]; };
  auto rdnwC = [&](int k) -> CT { return (CT)rdnw[X1(d, k)]; };
  auto rdnC = [&](int k) -> CT { return (CT)rdn[X1(d, k)]; };
  CT gam = CT(0);
  column_march_up<11, float, PF>(
      2, k_end + 1,
      [&](int f, int k) -> float {
        const Off o = ob + k * ni, op = ob + imin(k + 1, kde) * ni, opp = ob + imin(k + 2, kde) * ni;
        switch (f) {
          case 0: return t2ave[o];
          case 1: return c2a[o];
          case 2: return alt_f[o];
          case 3: return s.rw_tend[o];
          case 4: return s.cqw[o];
          case 5: return w[op];
          case 6: return ph_c[op];
          case 7: return s.ph_tend[op];
          case 8: return ph1f[opp];
          case 9: return phbf[opp];
          default: return wwf[opp];
        }
      },
      [&](int k, const float *v) {
        const Off o = ob + k * ni;
        const float w0 = w_k, t2a_hi = v[0], c2a0 = v[1], alt0 = v[2], rwt = v[3], cqw = v[4];
        const float w_p1 = v[5], phl_p1 = v[6];
        float rhs_p1 = 0.f;
        if (k + 1 <= kde) {
          rhs_p1 = rhs_at(k + 1, w_p1, v[7], phl_p1, ph1_k, ph1_k1, v[8], pb_k, pb_k1, v[9], ww_k, ww_k1, v[10]);
          rhs[o + ni] = rhs_p1;
        }
        float wv;
        CT ak, alpha;
        if (k <= k_end) {
          wv = w0 + dts * rwt +
               msft_inv * cqw *
                   (+.5f * dts * K_G * rdn[X1(d, k)] *
                    (c2a0 * rdnw[X1(d, k)] / (c1h[X1(d, k)] * mut + c2h[X1(d, k)]) *
                         ((1.f + epssm) * (rhs_p1 - rhs_0) + (1.f - epssm) * (phl_p1 - phl_0)) -
                     c2a_lo * rdnw[X1(d, k - 1)] / (c1h[X1(d, k - 1)] * mut + c2h[X1(d, k - 1)]) *
                         ((1.f + epssm) * (rhs_0 - rhs_m1) + (1.f - epssm) * (phl_0 - phl_m1)))) +
               dts * K_G * msft_inv *
                   (rdn[X1(d, k)] * (c2a0 * alt0 * t2a_hi - c2a_lo * alt_lo * t2a_lo) - (c1f[X1(d, k)] * muave));
          const CT CQW = (CT)cqw, C2A = (CT)c2a0, C2Al = (CT)c2a_lo;
          ak = k >= 3 ? -CQW * cof * rdnC(k) * rdnwC(k - 1) * C2Al / (cH(k - 1) * cF(k - 1)) : CT(0);
          const CT b = CT(1) + CQW * cof * rdnC(k) *
                                   (rdnwC(k) * C2A / (cH(k) * cF(k)) + rdnwC(k - 1) * C2Al / (cH(k - 1) * cF(k)));
          const CT cc = -CQW * cof * rdnC(k) * rdnwC(k) * C2A / (cH(k) * cF(k + 1));
          alpha = CT(1) / (b - ak * gam);
          gam = cc * alpha;
          GA[o] = gam;
        } else {
          wv = w0 + dts * rwt +
               msft_inv * (-.5f * dts * K_G / (c1h[X1(d, k - 1)] * mut + c2h[X1(d, k - 1)]) * (rdnw[X1(d, k - 1)] * rdnw[X1(d, k - 1)]) * 2.f *
                               c2a_lo * ((1.f + epssm) * (rhs_0 - rhs_m1) + (1.f - epssm) * (phl_0 - phl_m1)) -
                           dts * K_G * (2.f * rdnw[X1(d, k - 1)] * c2a_lo * alt_lo * t2a_lo + (c1f[X1(d, k)] * muave)));
          if (c.top_lid) wv = 0.f;
          const CT C2Al = (CT)c2a_lo;
          ak = -CT(2) * cof * (rdnwC(kde - 1) * rdnwC(kde - 1)) * C2Al * lid / (cH(kde - 1) * cF(kde - 1));
          const CT b = CT(1) + CT(2) * cof * (rdnwC(kde - 1) * rdnwC(kde - 1)) * C2Al / (cH(kde - 1) * cF(kde));
          alpha = CT(1) / (b - ak * gam);
        }
        rhs_m1 = rhs_0;
        rhs_0 = rhs_p1;
        phl_m1 = phl_0;
        phl_0 = phl_p1;
        t2a_lo = t2a_hi;
        c2a_lo = c2a0;
        alt_lo = alt0;
        w_k = w_p1;
        ph1_k = ph1_k1;
        ph1_k1 = v[8];
        pb_k = pb_k1;
        pb_k1 = v[9];
        ww_k = ww_k1;
        ww_k1 = v[10];
        if (sizeof(CT) == sizeof(float)) {
          const float wf = (wv - (float)ak * (float)wfe) * (float)alpha;
          wfe = (CT)wf;
          w[o] = wf;
        } else {
          wfe = ((CT)wv - ak * wfe) * alpha;
          w[o] = (float)wfe;
          lo[o] = (float)((double)wfe - (double)(float)wfe);
        }
      });
  // downward: back substitution, ph, calc_p_rho.  Fields: 0 = forward value of w (plus its
  // low part in FP64), 1 = gamma, 2 = rhs, 3 = t_2, 4 = t_save, 5 = alt, 6 = c2a, 7 = pm1.
  // Without PRHO (advance_w alone, a per-routine entry point) calc_p_rho is left out and
  // fields 3 to 7 are not loaded.
  const float muts = s.muts[o2], mu = s.mu_2[o2];
  const float smdiv = c.smdiv;
  // damp_opt = 3 (1445-1458): the back-substituted w relaxed toward w_save in the layer
  // within zdamp of the model top (heights from ph_save + phb), after the whole sweep
  const bool damp3 = c.damp_opt == 3;
  const float dampmag = dts * c.dampcoef, hdepth = c.zdamp;
  const float htop = damp3 ? (ph1f[ob + kde * ni] + phbf[ob + kde * ni]) / K_G : 0.f;
  CT wup = wfe;
  float ph_up = 0.f;  // final ph one level above
  column_march_down<8, CT, PF>(
      1, k_end + 1,
      [&](int f, int k) -> CT {
        const Off o = ob + k * ni;
        switch (f) {
          case 0:
            if (sizeof(CT) == sizeof(float) || k > k_end || k < 2) return (CT)w[o];
            return (CT)w[o] + (CT)lo[o];
          case 1: return GA[o];
          case 2: return (CT)rhs[o];
          case 3: return PRHO ? (CT)s.t_2[o] : CT(0);
          case 4: return PRHO ? (CT)s.t_save[o] : CT(0);
          case 5: return PRHO ? (CT)s.alt[o] : CT(0);
          case 6: return PRHO ? (CT)s.c2a[o] : CT(0);
          default: return PRHO ? (CT)s.pm1[o] : CT(0);
        }
      },
      [&](int k, const CT *v) {
        const Off o = ob + k * ni;
        float phk;
        if (k >= 2) {
          float wk;
          if (k == k_end + 1) {
            wk = (float)wup;  // top: the forward value
          } else if (sizeof(CT) == sizeof(float)) {
            wk = (float)v[0] - (float)v[1] * (float)wup;
            wup = (CT)wk;
            w[o] = wk;
          } else {
            const CT wn = v[0] - v[1] * wup;
            wup = wn;
            wk = (float)wn;
            w[o] = wk;
          }
          if (damp3) {
            const float hk = (ph1f[o] + phbf[o]) / K_G;
            const float hbot = htop - hdepth;
            float dampwt = 0.f;
            if (hk >= hbot) {
              const float sn = (float)sin((double)(.5f * (4.f * 0.785398185f) * (hk - hbot) / hdepth));
              dampwt = dampmag * sn * sn;
            }
            wk = (wk - dampwt * (c1f[X1(d, k)] * mut + c2f[X1(d, k)]) * s.w_save[o]) / (1.f + dampwt);
            w[o] = wk;
          }
          phk = (float)v[2] + msfty * .5f * dts * K_G * (1.f + epssm) * wk / (c1f[X1(d, k)] * muts + c2f[X1(d, k)]);
          ph[o] = phk;
        } else {
          phk = ph[o];
        }
        if (PRHO && k <= k_end) {
          const float cc = c1h[X1(d, k)], c2 = c2h[X1(d, k)];
          const float alt = (float)v[5];
          const float al = -1.f / (cc * muts + c2) * (alt * (cc * mu) + rdnw[X1(d, k)] * (ph_up - phk));
          s.al[o] = al;
          const float t1 = (float)v[4];
          float pp = (float)v[6] * (alt * ((float)v[3] - (cc * mu) * t1) / ((cc * muts + c2) * (K_T0 + t1)) - al);
          const float ptmp = pp;
          pp = pp + smdiv * (pp - (float)v[7]);
          s.p[o] = pp;
          s.pm1[o] = ptmp;
        }
        ph_up = phk;
      });
}

// One thread per (i, k, j); the thread of level kds also writes the 2-D fields.
template <typename ARGS>
__global__ void __launch_bounds__(256) k_small_finish(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  const int ide = d.ide, jde = d.jde, kde = d.kde;
  if (i > ide || j > jde || k > kde) return;
  const bool mi = i <= ide - 1, mj = j <= jde - 1, mass = mi && mj;
  const CPtr<float> c1h = s.c1h, c2h = s.c2h, c1f = s.c1f, c2f = s.c2f;
  const FPtr<float> muts = s.muts;
  const Off o2 = X2(d, i, j);
  const Off o = X3(d, i, k, j);
  const bool top = k == d.kds;
  // calc_mu_uv_1(muts -> muus, muvs)
  float muus = 0.f, muvs = 0.f;
  if (mj) {
    if (i > d.ids && i < ide) muus = 0.5f * (muts[o2] + muts[X2(d, i - 1, j)]);
    else if (i == d.ids) muus = 0.5f * (muts[o2] + muts[X2(d, a.c.periodic_x ? i - 1 : i, j)]);
    else muus = 0.5f * (muts[X2(d, i - 1, j)] + muts[X2(d, a.c.periodic_x ? i : i - 1, j)]);
    if (top) s.muus[o2] = muus;
  }
  if (mi) {
    if (j > d.jds && j < jde) muvs = 0.5f * (muts[o2] + muts[X2(d, i, j - 1)]);
    else if (j == d.jds) muvs = 0.5f * (muts[o2] + muts[X2(d, i, a.c.periodic_y ? j - 1 : j)]);
    else muvs = 0.5f * (muts[X2(d, i, j - 1)] + muts[X2(d, i, a.c.periodic_y ? j : j - 1)]);
    if (top) s.muvs[o2] = muvs;
  }
  // small_step_finish
  if (mi && k <= kde - 1) {
    const float muv = s.muv[o2], msfvx = s.msfvx[o2];
    s.v_2[o] = (msfvx * s.v_2[o] + s.v_save[o] * (c1h[X1(d, k)] * muv + c2h[X1(d, k)])) / (c1h[X1(d, k)] * muvs + c2h[X1(d, k)]);
  }
  if (mj && k <= kde - 1) {
    const float muu = s.muu[o2], msfuy = s.msfuy[o2];
    s.u_2[o] = (msfuy * s.u_2[o] + s.u_save[o] * (c1h[X1(d, k)] * muu + c2h[X1(d, k)])) / (c1h[X1(d, k)] * muus + c2h[X1(d, k)]);
  }
  if (!mass) return;
  const float mut = s.mut[o2], mts = muts[o2], msfty = s.msfty[o2];
  s.w_2[o] = (msfty * s.w_2[o] + s.w_save[o] * (c1f[X1(d, k)] * mut + c2f[X1(d, k)])) / (c1f[X1(d, k)] * mts + c2f[X1(d, k)]);
  s.ph_2[o] = s.ph_2[o] + s.ph_save[o];
  s.ww[o] = s.ww[o] + s.ww1[o];
  if (k <= kde - 1) {
    const bool last = a.rk_step >= a.c.rk_ord;
    const float cm = c1h[X1(d, k)] * mut + c2h[X1(d, k)];
    if (!last)
      s.t_2[o] = (s.t_2[o] + s.t_save[o] * cm) / (c1h[X1(d, k)] * mts + c2h[X1(d, k)]);
    else
      s.t_2[o] = (s.t_2[o] - a.dts * (float)a.nsmall * cm * s.h_diabatic[o] + s.t_save[o] * cm) / (c1h[X1(d, k)] * mts + c2h[X1(d, k)]);
  }
  if (top) s.mu_2[o2] = s.mu_2[o2] + s.mu_save[o2];
}

WRFMI_INST_P(k_advance_uv, false)
WRFMI_INST_P(k_advance_uv, true)
// four-point advance_uv and the column divergence: 2, 4 or 8 levels per thread (a launch
// shape the autotuner chooses); the column solve: prefetch depth 2, 4 or 8
#define WRFMI_INST_UV(P, KS) \
  template __global__ void k_advance_uv_x4<KArgs, P, KS>(KArgs); \
  template __global__ void k_advance_uv_x4<BatchArgs, P, KS>(BatchArgs);
WRFMI_INST_UV(false, 2)
WRFMI_INST_UV(false, 4)
WRFMI_INST_UV(false, 8)
WRFMI_INST_UV(true, 2)
WRFMI_INST_UV(true, 4)
WRFMI_INST_UV(true, 8)
WRFMI_INST_P2(k_coef_w, float, true)
WRFMI_INST_P2(k_coef_w, double, true)
WRFMI_INST_P2(k_coef_w, float, false)
WRFMI_INST_P2(k_coef_w, double, false)
WRFMI_INST_P(k_ac_mu_ww, false)
WRFMI_INST_P(k_ac_mu_ww, true)
#define WRFMI_INST_WS(CT, PRHO, PF) \
  template __global__ void k_ac_wsolve<KArgs, CT, PRHO, PF>(KArgs); \
  template __global__ void k_ac_wsolve<BatchArgs, CT, PRHO, PF>(BatchArgs);
WRFMI_INST_WS(float, true, 2)
WRFMI_INST_WS(float, true, 4)
WRFMI_INST_WS(float, true, 8)
WRFMI_INST_WS(double, true, 2)
WRFMI_INST_WS(double, true, 4)
WRFMI_INST_WS(double, true, 8)
WRFMI_INST_WS(float, false, 4)
WRFMI_INST(k_calc_p_rho)

WRFMI_INST(k_small_prep)
WRFMI_INST(k_p_rho0)
WRFMI_INST(k_ac_div)
WRFMI_INST(k_ac_theta)
WRFMI_INST_P(k_ac_div_x4, 2)
WRFMI_INST_P(k_ac_div_x4, 4)
WRFMI_INST_P(k_ac_div_x4, 8)
WRFMI_INST(k_small_prep_x4)
WRFMI_INST(k_p_rho0_x4)
WRFMI_INST(k_ac_theta_x4)
WRFMI_INST(k_small_finish)
WRFMI_INST(k_small_finish_x4)

}  // namespace wrfmi
