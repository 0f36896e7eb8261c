// rk_step_prep on the GPU.
//   k_step_prep : every point-wise part, one thread per (i,k,j) point
//   k_ww_scan   : calc_ww_cp's column sum and upward scan, one thread per (i,j) column
//   calculate_full   dyn_em/module_big_step_utilities_em.F:3587-3636
//   calc_mu_uv       dyn_em/module_big_step_utilities_em.F:26-180
//   couple_momentum  dyn_em/module_big_step_utilities_em.F:329-399
//   calc_ww_cp       dyn_em/module_big_step_utilities_em.F:640-782 (column scan)
//   calc_cq          dyn_em/module_big_step_utilities_em.F:787-906
//   calc_alt/calc_php dyn_em/module_big_step_utilities_em.F:910-949, 1227-1266
#include "kernels.h"

namespace wrfmi {

// calc_mu_uv's muu, muv at (i, j) (whole-patch tile: its == ids, ite == ide)
DEVI float prep_muu(const KArgs &a, int i, int j) {
  const Dm &d = a.d;
  const FPtr<float> mu = dstate(a).mu_2, mub = dstate(a).mub;
  if (i > d.ids && i < d.ide) return 0.5f * (((mu[X2(d, i, j)] + mu[X2(d, i - 1, j)]) + mub[X2(d, i, j)]) + mub[X2(d, i - 1, j)]);
  if (i == d.ids) {
    const int im = a.c.periodic_x ? i - 1 : i;
    return 0.5f * (((mu[X2(d, i, j)] + mu[X2(d, im, j)]) + mub[X2(d, i, j)]) + mub[X2(d, im, j)]);
  }
  const int im = a.c.periodic_x ? i : i - 1;  // i == ide
  return 0.5f * (((mu[X2(d, i - 1, j)] + mu[X2(d, im, j)]) + mub[X2(d, i - 1, j)]) + mub[X2(d, im, j)]);
}
DEVI float prep_muv(const KArgs &a, int i, int j) {
  const Dm &d = a.d;
  const FPtr<float> mu = dstate(a).mu_2, mub = dstate(a).mub;
  if (j > d.jds && j < d.jde) return 0.5f * (((mu[X2(d, i, j)] + mu[X2(d, i, j - 1)]) + mub[X2(d, i, j)]) + mub[X2(d, i, j - 1)]);
  if (j == d.jds) {
    const int jm = a.c.periodic_y ? j - 1 : j;
    return 0.5f * (((mu[X2(d, i, j)] + mu[X2(d, i, jm)]) + mub[X2(d, i, j)]) + mub[X2(d, i, jm)]);
  }
  const int jm = a.c.periodic_y ? j : j - 1;
  return 0.5f * (((mu[X2(d, i, j - 1)] + mu[X2(d, i, jm)]) + mub[X2(d, i, j - 1)]) + mub[X2(d, i, jm)]);
}

// Point part: one thread per (i, k, j) with i = ids-1..ide, j = jds-1..jde (mut is also
// needed one point into the halo, as calculate_full computes it there); the thread of
// level 1 writes the 2-D fields.
template <typename ARGS>
__global__ void __launch_bounds__(256) k_step_prep(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids - 1 + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds - 1 + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  const int ide = d.ide, jde = d.jde, kde = d.kde;
  if (i > ide || j > jde || k > kde) return;
  const bool mi = i <= ide - 1, mj = j <= jde - 1;
  const bool top = k == d.kds;
  float mut = 0.f;
  if (mi && mj) {
    mut = s.mub[X2(d, i, j)] + s.mu_2[X2(d, i, j)];
    if (top) s.mut[X2(d, i, j)] = mut;
  }
  if (i < d.ids || j < d.jds) return;
  const int ktf = kde - 1;
  const Off o = X3(d, i, k, j);
  const int nm = s.n_moist;
  if (mj) {
    const float muu = prep_muu(a, i, j);
    if (top) s.muu[X2(d, i, j)] = muu;
    if (k <= ktf) {
      s.ru[o] = s.u_2[o] * (s.c1h[X1(d, k)] * muu + s.c2h[X1(d, k)]) / s.msfuy[X2(d, i, j)];
      float cq;
      if (nm >= 1) {
        float qtot = 0.f;
        for (int m = 0; m < nm; m++) qtot = qtot + s.q[m][o] + s.q[m][o - 1];
        cq = 1.f / (1.f + 0.5f * qtot);
      } else {
        cq = 1.f;
      }
      s.cqu[o] = cq;
    }
  }
  if (mi) {
    const float muv = prep_muv(a, i, j);
    if (top) s.muv[X2(d, i, j)] = muv;
    if (k <= ktf) {
      s.rv[o] = s.v_2[o] * (s.c1h[X1(d, k)] * muv + s.c2h[X1(d, k)]) * s.msfvx_inv[X2(d, i, j)];
      float cq;
      if (nm >= 1) {
        float qtot = 0.f;
        for (int m = 0; m < nm; m++) qtot = qtot + s.q[m][o] + s.q[m][o - d.sj];
        cq = 1.f / (1.f + 0.5f * qtot);
      } else {
        cq = 1.f;
      }
      s.cqv[o] = cq;
    }
  }
  if (mi && mj) {
    if (k >= 2 && k <= ktf) {
      float cw;
      if (nm >= 1) {
        float qtot = 0.f;
        for (int m = 0; m < nm; m++) qtot = qtot + s.q[m][o] + s.q[m][o - d.ni];
        cw = 0.5f * qtot;
      } else {
        cw = 0.f;
      }
      s.cqw[o] = cw;
    }
    if (k <= ktf) {
      s.alt[o] = s.al[o] + s.alb[o];
      s.php[o] = 0.5f * (((s.phb[o] + s.phb[o + d.ni]) + s.ph_2[o]) + s.ph_2[o + d.ni]);
    }
    s.rw[o] = s.w_2[o] * (s.c1f[X1(d, k)] * mut + s.c2f[X1(d, k)]) / s.msfty[X2(d, i, j)];
  }
}

// k_step_prep with four i points per lane (16-byte accesses; element by element the
// expressions of k_step_prep).  The lanes start one group of four below the first group
// holding ids, so that mut is also formed at i = ids-1.
template <typename ARGS>
__global__ void __launch_bounds__(256) k_step_prep_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) - 4 + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds - 1 + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  const int ids = d.ids, ide = d.ide, jds = d.jds, jde = d.jde, kde = d.kde;
  if (i0 > ide || j > jde || k > kde) return;
  const bool mj = j <= jde - 1;
  const bool top = k == d.kds;
  const Off o2 = X2(d, i0, j), o = X3(d, i0, k, j);
  const F4 mub = s.mub.ld<F4>(o2), mu = s.mu_2.ld<F4>(o2);
  F4 mut;
#pragma unroll
  for (int n = 0; n < 4; n++) mut.v[n] = mub.v[n] + mu.v[n];
  if (top && mj) st4m(s.mut, o2, mut, i0, ids - 1, ide - 1);
  if (j < jds || i0 + 3 < ids) return;
  const int ktf = kde - 1;
  const int nm = s.n_moist;
  const int kc = imin(k, ktf);
  const float c1 = s.c1h[X1(d, kc)], c2 = s.c2h[X1(d, kc)], c1f = s.c1f[X1(d, k)], c2f = s.c2f[X1(d, k)];
  const bool px = a.c.periodic_x, py = a.c.periodic_y;
  if (mj) {
    // prep_muu at the u points
    const F4 mul = s.mu_2.ld<F4>(o2 - 4), mubl = s.mub.ld<F4>(o2 - 4);
    F4 muu;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const int i = i0 + n;
      const float m0 = mu.v[n], ml = xm1(mul, mu, n), b0 = mub.v[n], bl = xm1(mubl, mub, n);
      if (i > ids && i < ide) muu.v[n] = 0.5f * (((m0 + ml) + b0) + bl);
      else if (i == ids) muu.v[n] = 0.5f * (((m0 + (px ? ml : m0)) + b0) + (px ? bl : b0));
      else muu.v[n] = 0.5f * (((ml + (px ? m0 : ml)) + bl) + (px ? b0 : bl));
    }
    if (top) st4m(s.muu, o2, muu, i0, ids, ide);
    if (k <= ktf) {
      const F4 u = s.u_2.ld<F4>(o), msfuy = s.msfuy.ld<F4>(o2);
      F4 qtot, ru;
#pragma unroll
      for (int n = 0; n < 4; n++) qtot.v[n] = 0.f;
      for (int m = 0; m < nm; m++) {
        const F4 qc = s.q[m].ld<F4>(o), ql = s.q[m].ld<F4>(o - 4);
#pragma unroll
        for (int n = 0; n < 4; n++) qtot.v[n] = qtot.v[n] + qc.v[n] + xm1(ql, qc, n);
      }
      F4 cq;
#pragma unroll
      for (int n = 0; n < 4; n++) {
        ru.v[n] = u.v[n] * (c1 * muu.v[n] + c2) / msfuy.v[n];
        cq.v[n] = nm >= 1 ? 1.f / (1.f + 0.5f * qtot.v[n]) : 1.f;
      }
      st4m(s.ru, o, ru, i0, ids, ide);
      st4m(s.cqu, o, cq, i0, ids, ide);
    }
  }
  {
    // prep_muv at the v points (i <= ide-1)
    const F4 mus = s.mu_2.ld<F4>(o2 - d.ni), mubs = s.mub.ld<F4>(o2 - d.ni);
    F4 muv;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const float m0 = mu.v[n], ms = mus.v[n], b0 = mub.v[n], bs = mubs.v[n];
      if (j > jds && j < jde) muv.v[n] = 0.5f * (((m0 + ms) + b0) + bs);
      else if (j == jds) muv.v[n] = 0.5f * (((m0 + (py ? ms : m0)) + b0) + (py ? bs : b0));
      else muv.v[n] = 0.5f * (((ms + (py ? m0 : ms)) + bs) + (py ? b0 : bs));
    }
    if (top) st4m(s.muv, o2, muv, i0, ids, ide - 1);
    if (k <= ktf) {
      const F4 v = s.v_2.ld<F4>(o), mvi = s.msfvx_inv.ld<F4>(o2);
      F4 qtot, rv;
#pragma unroll
      for (int n = 0; n < 4; n++) qtot.v[n] = 0.f;
      for (int m = 0; m < nm; m++) {
        const F4 qc = s.q[m].ld<F4>(o), qs = s.q[m].ld<F4>(o - d.sj);
#pragma unroll
        for (int n = 0; n < 4; n++) qtot.v[n] = qtot.v[n] + qc.v[n] + qs.v[n];
      }
      F4 cq;
#pragma unroll
      for (int n = 0; n < 4; n++) {
        rv.v[n] = v.v[n] * (c1 * muv.v[n] + c2) * mvi.v[n];
        cq.v[n] = nm >= 1 ? 1.f / (1.f + 0.5f * qtot.v[n]) : 1.f;
      }
      st4m(s.rv, o, rv, i0, ids, ide - 1);
      st4m(s.cqv, o, cq, i0, ids, ide - 1);
    }
  }
  if (!mj) return;
  if (k >= 2 && k <= ktf) {
    F4 qtot, cw;
#pragma unroll
    for (int n = 0; n < 4; n++) qtot.v[n] = 0.f;
    for (int m = 0; m < nm; m++) {
      const F4 qc = s.q[m].ld<F4>(o), qb = s.q[m].ld<F4>(o - d.ni);
#pragma unroll
      for (int n = 0; n < 4; n++) qtot.v[n] = qtot.v[n] + qc.v[n] + qb.v[n];
    }
#pragma unroll
    for (int n = 0; n < 4; n++) cw.v[n] = nm >= 1 ? 0.5f * qtot.v[n] : 0.f;
    st4m(s.cqw, o, cw, i0, ids, ide - 1);
  }
  if (k <= ktf) {
    const F4 al = s.al.ld<F4>(o), alb = s.alb.ld<F4>(o), phb = s.phb.ld<F4>(o), phbp = s.phb.ld<F4>(o + d.ni);
    const F4 ph = s.ph_2.ld<F4>(o), php = s.ph_2.ld<F4>(o + d.ni);
    F4 alt, phpn;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      alt.v[n] = al.v[n] + alb.v[n];
      phpn.v[n] = 0.5f * (((phb.v[n] + phbp.v[n]) + ph.v[n]) + php.v[n]);
    }
    st4m(s.alt, o, alt, i0, ids, ide - 1);
    st4m(s.php, o, phpn, i0, ids, ide - 1);
  }
  {
    const F4 w = s.w_2.ld<F4>(o), msfty = s.msfty.ld<F4>(o2);
    F4 rw;
#pragma unroll
    for (int n = 0; n < 4; n++) rw.v[n] = w.v[n] * (c1f * mut.v[n] + c2f) / msfty.v[n];
    st4m(s.rw, o, rw, i0, ids, ide - 1);
  }
}

// calc_ww_cp (module_big_step_utilities_em.F:640-782): column divergence of the coupled
// horizontal momentum and the upward scan for omega; one thread per (i, j) column.
template <typename ARGS, bool COMP>
__global__ void __launch_bounds__(256) k_ww_scan(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = col_ia(d, !std::is_same<ARGS, KArgs>::value) + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int ide = d.ide, jde = d.jde, kde = d.kde;
  if (i < d.ids || i > ide || j > jde) return;
  const bool mi = i <= ide - 1, mj = j <= jde - 1;
  const FPtr<float> mu = s.mu_2, mub = s.mub;
  const CPtr<float> c1h = s.c1h, c2h = s.c2h;
  const int ktf = kde - 1;
  // ww(1) = ww(kte) = 0 on i = its..ite
  if (mj) {
    s.ww[X3(d, i, 1, j)] = 0.f;
    s.ww[X3(d, i, kde, j)] = 0.f;
  }
  if (!(mi && mj)) return;
  // full column mass at the u points i, i+1 and v points j, j+1 (calc_ww_cp's own muu/muv)
  const float lmuu0 = 0.5f * (((mu[X2(d, i, j)] + mub[X2(d, i, j)]) + mu[X2(d, i - 1, j)]) + mub[X2(d, i - 1, j)]);
  const float lmuu1 = 0.5f * (((mu[X2(d, i + 1, j)] + mub[X2(d, i + 1, j)]) + mu[X2(d, i, j)]) + mub[X2(d, i, j)]);
  const float lmuv0 = 0.5f * (((mu[X2(d, i, j)] + mub[X2(d, i, j)]) + mu[X2(d, i, j - 1)]) + mub[X2(d, i, j - 1)]);
  const float lmuv1 = 0.5f * (((mu[X2(d, i, j + 1)] + mub[X2(d, i, j + 1)]) + mu[X2(d, i, j)]) + mub[X2(d, i, j)]);
  const float msfuy0 = s.msfuy[X2(d, i, j)], msfuy1 = s.msfuy[X2(d, i + 1, j)];
  const float mvi0 = s.msfvx_inv[X2(d, i, j)], mvi1 = s.msfvx_inv[X2(d, i, j + 1)];
  const float msftx = s.msftx[X2(d, i, j)];
  const float rdx = s.rdx, rdy = s.rdy;
  float dmdt = 0.f, comp = 0.f;
  // pass 1: divv(k) and its column sum; divv(k) is parked in ww(k+1) until the scan reads it
  for (int k = 1; k <= ktf; k++) {
    const Off o = X3(d, i, k, j);
    const float c1 = c1h[X1(d, k)], c2 = c2h[X1(d, k)];
    const float tx = (c1 * lmuu1 + c2) * s.u_2[o + 1] / msfuy1 - (c1 * lmuu0 + c2) * s.u_2[o] / msfuy0;
    const float ty = (c1 * lmuv1 + c2) * s.v_2[o + d.sj] * mvi1 - (c1 * lmuv0 + c2) * s.v_2[o] * mvi0;
    const float divv = msftx * s.dnw[X1(d, k)] * (rdx * tx + rdy * ty);
    if (COMP) two_sum_acc(dmdt, comp, divv);
    else dmdt = dmdt + divv;
    if (k <= ktf - 1) s.ww[o + d.ni] = divv;
  }
  if (COMP) dmdt = dmdt + comp;
  float prev = 0.f;
  for (int k = 2; k <= ktf; k++) {
    const Off o = X3(d, i, k, j);
    const float dv = s.ww[o];
    const float cur = prev - s.dnw[X1(d, k - 1)] * c1h[X1(d, k - 1)] * dmdt - dv;
    s.ww[o] = cur;
    prev = cur;
  }
}

// set_w_surface with fill_w_flag false (dyn_em/module_bc_em.F:1196-1295): the kinematic
// lower boundary condition w(1) = msfty v.dz/dy + msftx u.dz/dx, with u and v extrapolated
// to the surface by cf1..cf3; across a non-periodic edge the neighbour is the point itself.
// One thread per interior mass column.
template <typename ARGS>
__global__ void __launch_bounds__(256) k_set_w_surface(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  if (i > d.ide - 1 || j > d.jde - 1) return;
  const bool px = a.c.periodic_x, py = a.c.periodic_y;
  const int jm1 = imax(j - 1, py ? d.jds - 1 : d.jds), jp1 = imin(j + 1, py ? d.jde : d.jde - 1);
  const int im1 = imax(i - 1, px ? d.ids - 1 : d.ids), ip1 = imin(i + 1, px ? d.ide : d.ide - 1);
  const FPtr<float> u = s.u_2, v = s.v_2, ht = s.ht;
  const float cf1 = s.cf1, cf2 = s.cf2, cf3 = s.cf3;
  const Off o2 = X2(d, i, j);
  const float h0 = ht[o2];
  const float vs_n = cf1 * v[X3(d, i, 1, j + 1)] + cf2 * v[X3(d, i, 2, j + 1)] + cf3 * v[X3(d, i, 3, j + 1)];
  const float vs_s = cf1 * v[X3(d, i, 1, j)] + cf2 * v[X3(d, i, 2, j)] + cf3 * v[X3(d, i, 3, j)];
  const float us_e = cf1 * u[X3(d, i + 1, 1, j)] + cf2 * u[X3(d, i + 1, 2, j)] + cf3 * u[X3(d, i + 1, 3, j)];
  const float us_w = cf1 * u[X3(d, i, 1, j)] + cf2 * u[X3(d, i, 2, j)] + cf3 * u[X3(d, i, 3, j)];
  s.w_2[X3(d, i, 1, j)] = s.msfty[o2] * .5f * s.rdy * ((ht[X2(d, i, jp1)] - h0) * vs_n + (h0 - ht[X2(d, i, jm1)]) * vs_s) +
                          s.msftx[o2] * .5f * s.rdx * ((ht[X2(d, ip1, j)] - h0) * us_e + (h0 - ht[X2(d, im1, j)]) * us_w);
}

WRFMI_INST(k_set_w_surface)
WRFMI_INST_P(k_ww_scan, false)
WRFMI_INST_P(k_ww_scan, true)

WRFMI_INST(k_step_prep)
WRFMI_INST(k_step_prep_x4)

}  // namespace wrfmi
