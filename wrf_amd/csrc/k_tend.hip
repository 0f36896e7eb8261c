// rk_tendency + rk_addtend_dry on the GPU (dyn_em/module_em.F:190-1092), physics off.
// Four column kernels (one thread per (i,j), marching in k), one per tendency family:
//   k_tend_u : ru_tend = advect_u + horizontal_pressure_gradient(x) + coriolis + curvature + addtend
//   k_tend_v : rv_tend = advect_v + horizontal_pressure_gradient(y) + coriolis + curvature + addtend
//   k_tend_w : rw_tend = advect_w + pg_buoy_w (+ cqw := 1/(1+cqw)) + w_damp + coriolis + curvature + addtend
//   k_tend_t : t_tend = advect_scalar(t) + addtend;  ph_tend = rhs_ph + addtend;
//              plus zero_tend of u/v/w/ph/t_save, mu_tend, mu_save
// Every output point accumulates its contributions in the reference's order
// (zero_tend, then each routine of rk_tendency in call order, then rk_addtend_dry),
// so the result does not depend on which thread computes which face flux.
// Routines and line ranges: advect_u module_advect_em.F:126-1526, advect_v 1530-3024,
// advect_scalar 3029-4359, advect_w 4364-6064; rhs_ph module_big_step_utilities_em.F:1365-2178,
// horizontal_pressure_gradient 2183-2415, pg_buoy_w 2419-2499, w_damp 2503-2711,
// coriolis 3640-3850, curvature 4175-4469.
#include "kernels.h"

namespace wrfmi {

// ------------------------------------------------------------------------------ u
template <typename ARGS, bool PER, bool ZONE>
__global__ void __launch_bounds__(256) k_tend_u(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  int i, j;
  if (!point_ij<ZONE>(a, T, i, j)) return;
  constexpr int KS = ZONE ? 1 : kTendK;  // levels per thread (the zone form: one)
  const int k0 = d.kds + (int)T.bz * KS;
  const int ids = d.ids, ide = d.ide, jds = d.jds, jde = d.jde, kde = d.kde;
  const int its = ids, ite = ide, jts = jds, jte = jde, kts = d.kds, kte = kde;
  if (i > ide || j > jde || k0 > kte) return;
  const int k1 = imin(k0 + KS - 1, kte);
  const Geo g = geo<PER>(a.c);
  const float ts = a.c.time_step >= 0 ? 1.f : -1.f;
  const int ktf = kte - 1;
  const FPtr<float> u = s.u_2, ru = s.ru, rv = s.rv, rom = s.ww, rw = s.rw;
  const float rdx = s.rdx, rdy = s.rdy;
  const bool dxs = !(g.per_x || g.sym_xs || its > ids + 3), dxe = !(g.per_x || g.sym_xe || ite < ide - 2);
  const bool dys = !(g.per_y || g.sym_ys || jts > jds + 3), dye = !(g.per_y || g.sym_ye || jte < jde - 4);
  // y flux: columns iy0..iy1, tendency rows jy0..jy1, full-order faces jf0..jf1
  int iy0 = its, iy1 = ite;
  if (g.open_xs || g.spec) iy0 = imax(ids + 1, its);
  if (g.open_xe || g.spec) iy1 = imin(ide - 1, ite);
  if (g.per_x) { iy0 = its; iy1 = ite; }
  int jy0 = jts, jy1 = imin(jte, jde - 1), jf0 = jy0, jf1 = jy1 + 1;
  if (dys) { jy0 = imax(jts, jds + 1); jf0 = jds + 3; }
  if (dye) { jy1 = imin(jte, jde - 2); jf1 = jde - 3; }
  // x flux
  int ix0 = its, ix1 = ite, if0 = ix0, if1 = ix1 + 1;
  if (dxs) { ix0 = imax(ids + 1, its); if0 = ids + 3; }
  if (dxe) { ix1 = imin(ide - 1, ite); if1 = ide - 2; }
  // vertical
  int iv0 = its, iv1 = ite;
  if (g.open_ys || g.spec) iv0 = imax(ids + 1, its);
  if (g.open_ye || g.spec) iv1 = imin(ide - 1, ite);
  if (g.per_x) { iv0 = its; iv1 = ite; }
  // pressure gradient
  int ih0 = its, ih1 = ite;
  if ((g.open_xs || g.spec || g.nested) && its == ids) ih0 = its + 1;
  if ((g.open_xe || g.spec || g.nested) && ite == ide) ih1 = ite - 1;
  if (g.per_x) { ih0 = its; ih1 = ite; }
  // coriolis
  int ic0 = its, ic1 = ite;
  if (g.open_xs || g.spec || g.nested) ic0 = imax(ids + 1, its);
  if (g.open_xe || g.spec || g.nested) ic1 = imin(ide - 1, ite);
  if (g.per_x) { ic0 = its; ic1 = ite; }
  // curvature: vxgm stencil ranges and its copy rules
  const bool cpx_s = (g.open_xs || (g.spec && !g.per_x) || g.nested) && its == ids;
  const bool cpx_e = (g.open_xe || (g.spec && !g.per_x) || g.nested) && ite == ide;
  const bool cpy_s = (g.open_ys || g.spec || g.nested || g.polar) && jts == jds;
  const bool cpy_e = (g.open_ye || g.spec || g.nested || g.polar) && jte == jde;
  int ik0 = its, ik1 = ite;
  if (g.open_xs || g.spec || g.nested) ik0 = imax(ids + 1, its);
  if (g.open_xe || g.spec || g.nested) ik1 = imin(ide - 1, ite);
  if (g.per_x) { ik0 = its; ik1 = ite; }
  const FPtr<float> msfvx = s.msfvx, msfuy = s.msfuy, vv = s.v_2;
  auto vx_raw = [&](int ii, int k, int jj) -> float {
    return 0.5f * (u[X3(d, ii, k, jj)] + u[X3(d, ii + 1, k, jj)]) * (msfvx[X2(d, ii, jj + 1)] - msfvx[X2(d, ii, jj)]) * rdy -
           0.5f * (vv[X3(d, ii, k, jj)] + vv[X3(d, ii, k, jj + 1)]) * (msfuy[X2(d, ii + 1, jj)] - msfuy[X2(d, ii, jj)]) * rdx;
  };
  auto vx = [&](int ii, int k, int jj) -> float {
    if (cpy_s && jj == jts - 1) jj = jts;
    else if (cpy_e && jj == jte) jj = jte - 1;
    if (jj >= jts && jj <= jte - 1) {
      if (cpx_s && ii == its - 1) ii = its;
      else if (cpx_e && ii == ite) ii = ite - 1;
    }
    return vx_raw(ii, k, jj);
  };
  auto fqy = [&](int k, int jj) -> float {
    if (jj >= jf0 && jj <= jf1) {
      const float vel = 0.5f * (rv[X3(d, i, k, jj)] + rv[X3(d, i - 1, k, jj)]);
      return vel * flux5(u[X3(d, i, k, jj - 3)], u[X3(d, i, k, jj - 2)], u[X3(d, i, k, jj - 1)], u[X3(d, i, k, jj)],
                         u[X3(d, i, k, jj + 1)], u[X3(d, i, k, jj + 2)], vel, ts);
    } else if (jj == jds + 1 || jj == jde - 1) {
      return 0.25f * (rv[X3(d, i, k, jj)] + rv[X3(d, i - 1, k, jj)]) * (u[X3(d, i, k, jj)] + u[X3(d, i, k, jj - 1)]);
    } else {
      const float vel = 0.5f * (rv[X3(d, i, k, jj)] + rv[X3(d, i - 1, k, jj)]);
      return vel * flux3(u[X3(d, i, k, jj - 2)], u[X3(d, i, k, jj - 1)], u[X3(d, i, k, jj)], u[X3(d, i, k, jj + 1)], vel, ts);
    }
  };
  auto fqx = [&](int k, int ii) -> float {
    if (ii >= if0 && ii <= if1) {
      const float vel = 0.5f * (ru[X3(d, ii, k, j)] + ru[X3(d, ii - 1, k, j)]);
      return vel * flux5(u[X3(d, ii - 3, k, j)], u[X3(d, ii - 2, k, j)], u[X3(d, ii - 1, k, j)], u[X3(d, ii, k, j)],
                         u[X3(d, ii + 1, k, j)], u[X3(d, ii + 2, k, j)], vel, ts);
    } else if ((dxs && ix0 == ids + 1 && ii == ids + 1)) {
      float ub = u[X3(d, ii - 1, k, j)];
      if (g.spec && u[X3(d, ii, k, j)] < 0.f) ub = u[X3(d, ii, k, j)];
      return 0.25f * (ru[X3(d, ii, k, j)] + ru[X3(d, ii - 1, k, j)]) * (u[X3(d, ii, k, j)] + ub);
    } else if (dxe && ix1 == ide - 1 && ii == ide) {
      float ub = u[X3(d, ii, k, j)];
      if (g.spec && u[X3(d, ii - 1, k, j)] > 0.f) ub = u[X3(d, ii - 1, k, j)];
      return 0.25f * (ru[X3(d, ii, k, j)] + ru[X3(d, ii - 1, k, j)]) * (u[X3(d, ii - 1, k, j)] + ub);
    } else {
      const float vel = 0.5f * (ru[X3(d, ii, k, j)] + ru[X3(d, ii - 1, k, j)]);
      return vel * flux3(u[X3(d, ii - 2, k, j)], u[X3(d, ii - 1, k, j)], u[X3(d, ii, k, j)], u[X3(d, ii + 1, k, j)], vel, ts);
    }
  };
  auto vflux = [&](int k) -> float {
    if (k == kts || k == kte) return 0.f;
    if (k == kts + 1 || k == ktf)
      return 0.5f * (rom[X3(d, i, k, j)] + rom[X3(d, i - 1, k, j)]) *
             (s.fnm[X1(d, k)] * u[X3(d, i, k, j)] + s.fnp[X1(d, k)] * u[X3(d, i, k - 1, j)]);
    const float vel = 0.5f * (rom[X3(d, i - 1, k, j)] + rom[X3(d, i, k, j)]);
    return vel * flux3(u[X3(d, i, k - 2, j)], u[X3(d, i, k - 1, j)], u[X3(d, i, k, j)], u[X3(d, i, k + 1, j)], -vel, ts);
  };
  const bool row_ok = j <= imin(jte, jde - 1);
  const FPtr<float> p = s.p, ph = s.ph_2, al = s.al, alt = s.alt, pb = s.pb, php = s.php, mu = s.mu_2;
  const float cf1 = s.cf1, cf2 = s.cf2, cf3 = s.cf3, cfn = s.cfn, cfn1 = s.cfn1;
  const bool nh = a.c.non_hydrostatic;
  auto dpn = [&](int k) -> float {  // vertically interpolated p at w level k, averaged to the u point
    if (k == 1)
      return .5f * (cf1 * (p[X3(d, i - 1, 1, j)] + p[X3(d, i, 1, j)]) + cf2 * (p[X3(d, i - 1, 2, j)] + p[X3(d, i, 2, j)]) +
                    cf3 * (p[X3(d, i - 1, 3, j)] + p[X3(d, i, 3, j)]));
    if (k == kde) {
      if (!a.c.top_lid) return 0.f;
      return .5f * (cfn * (p[X3(d, i - 1, kde - 1, j)] + p[X3(d, i, kde - 1, j)]) + cfn1 * (p[X3(d, i - 1, kde - 2, j)] + p[X3(d, i, kde - 2, j)]));
    }
    return .5f * (s.fnm[X1(d, k)] * (p[X3(d, i - 1, k, j)] + p[X3(d, i, k, j)]) + s.fnp[X1(d, k)] * (p[X3(d, i - 1, k - 1, j)] + p[X3(d, i, k - 1, j)]));
  };
  // the vertical flux and dpn at the lower w level are carried up the column segment
  float vf_lo = (row_ok && k0 <= ktf && i >= iv0 && i <= iv1) ? vflux(k0) : 0.f;
  // advect_u alone (a per-routine entry point): added to the tendency array's values
  const bool adv_only = a.routine == R_ADVECT;
  float dpn_lo = (row_ok && k0 <= ktf && nh && !adv_only && i >= ih0 && i <= ih1) ? dpn(k0) : 0.f;
  for (int k = k0; k <= k1; k++) {
    const Off o = X3(d, i, k, j);
    float t = adv_only ? s.ru_tend[o] : 0.f;
    float vf_hi = 0.f, dpn_hi = 0.f;
    if (row_ok && k <= ktf) {
      // advect_u: y flux divergence
      if (i >= iy0 && i <= iy1 && j >= jy0 && j <= jy1) {
        const float mrdy = s.msfux[X2(d, i, j)] * rdy;
        t = t - mrdy * (fqy(k, j + 1) - fqy(k, j));
      }
      // x flux divergence
      if (i >= ix0 && i <= ix1) {
        const float mrdx = s.msfux[X2(d, i, j)] * rdx;
        t = t - mrdx * (fqx(k, i + 1) - fqx(k, i));
      }
      const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
      // radiative lateral boundary condition in x
      if (g.open_xs && i == its) {
        float ub = ru[o] - K_CB * (c1 * s.mut[X2(d, its, j)] + c2);
        if (ub > 0.f) ub = 0.f;
        t = t - rdx * ub * (u[X3(d, its + 1, k, j)] - u[o]);
      }
      if (g.open_xe && i == ite) {
        float ub = ru[o] + K_CB * (c1 * s.mut[X2(d, ite - 1, j)] + c2);
        if (ub < 0.f) ub = 0.f;
        t = t - rdx * ub * (u[o] - u[X3(d, ite - 1, k, j)]);
      }
      {
        int io0 = its, io1 = imin(ite, ide), imn = ids, imx = ide - 1;
        if (g.open_xs) io0 = imax(ids + 1, its);
        if (g.open_xe) io1 = imin(ite, ide - 1);
        const int ip = imin(imx, i), im = imax(imn, i - 1);
        if (g.open_ys && j == jts && i >= io0 && i <= io1) {
          const float mrdy = s.msfux[X2(d, i, jts)] * rdy;
          const float vw = 0.5f * (rv[X3(d, ip, k, jts)] + rv[X3(d, im, k, jts)]);
          const float vb = vw < 0.f ? vw : 0.f;
          const float dvm = rv[X3(d, ip, k, jts + 1)] - rv[X3(d, ip, k, jts)];
          const float dvp = rv[X3(d, im, k, jts + 1)] - rv[X3(d, im, k, jts)];
          t = t - mrdy * (vb * (u[X3(d, i, k, jts + 1)] - u[X3(d, i, k, jts)]) + 0.5f * u[X3(d, i, k, jts)] * (dvm + dvp));
        }
        if (g.open_ye && j == jte - 1 && i >= io0 && i <= io1) {
          const float mrdy = s.msfux[X2(d, i, jte - 1)] * rdy;
          const float vw = 0.5f * (rv[X3(d, ip, k, jte)] + rv[X3(d, im, k, jte)]);
          const float vb = vw > 0.f ? vw : 0.f;
          const float dvm = rv[X3(d, ip, k, jte)] - rv[X3(d, ip, k, jte - 1)];
          const float dvp = rv[X3(d, im, k, jte)] - rv[X3(d, im, k, jte - 1)];
          t = t - mrdy * (vb * (u[X3(d, i, k, jte - 1)] - u[X3(d, i, k, jte - 2)]) + 0.5f * u[X3(d, i, k, jte - 1)] * (dvm + dvp));
        }
      }
      // vertical flux divergence (order 3)
      if (i >= iv0 && i <= iv1) {
        vf_hi = vflux(k + 1);
        t = t - s.rdnw[X1(d, k)] * (vf_hi - vf_lo);
      }
      // horizontal_pressure_gradient (x)
      if (!adv_only && i >= ih0 && i <= ih1) {
        const float m = s.msfux[X2(d, i, j)] / s.msfuy[X2(d, i, j)];
        float dpx = m * .5f * rdx * (c1 * s.muu[X2(d, i, j)] + c2) *
                    ((ph[X3(d, i, k + 1, j)] - ph[X3(d, i - 1, k + 1, j)] + ph[o] - ph[X3(d, i - 1, k, j)]) +
                     (alt[o] + alt[X3(d, i - 1, k, j)]) * (p[o] - p[X3(d, i - 1, k, j)]) +
                     (al[o] + al[X3(d, i - 1, k, j)]) * (pb[o] - pb[X3(d, i - 1, k, j)]));
        if (nh) {
          dpn_hi = dpn(k + 1);
          dpx = dpx + m * rdx * (php[o] - php[X3(d, i - 1, k, j)]) *
                          (s.rdnw[X1(d, k)] * (dpn_hi - dpn_lo) - .5f * ((c1 * mu[X2(d, i - 1, j)]) + (c1 * mu[X2(d, i, j)])));
          t = t - s.cqu[o] * dpx;
        } else {
          t = t - dpx;
        }
      }
      // coriolis
      if (!adv_only && i >= ic0 && i <= ic1) {
        t = t + (s.msfux[X2(d, i, j)] / s.msfuy[X2(d, i, j)]) * 0.5f * (s.f[X2(d, i, j)] + s.f[X2(d, i - 1, j)]) * 0.25f *
                    (rv[X3(d, i - 1, k, j + 1)] + rv[X3(d, i, k, j + 1)] + rv[X3(d, i - 1, k, j)] + rv[X3(d, i, k, j)]) -
            0.5f * (s.e[X2(d, i, j)] + s.e[X2(d, i - 1, j)]) * 0.5f * (s.cosa[X2(d, i, j)] + s.cosa[X2(d, i - 1, j)]) * 0.25f *
                (rw[X3(d, i - 1, k + 1, j)] + rw[X3(d, i - 1, k, j)] + rw[X3(d, i, k + 1, j)] + rw[X3(d, i, k, j)]);
      }
      // curvature
      if (!adv_only && i >= ik0 && i <= ik1) {
        t = t + 0.5f * (vx(i, k, j) + vx(i - 1, k, j)) * 0.25f *
                    (rv[X3(d, i - 1, k, j + 1)] + rv[X3(d, i, k, j + 1)] + rv[X3(d, i - 1, k, j)] + rv[X3(d, i, k, j)]) -
            u[o] * K_RERADIUS * 0.25f * (rw[X3(d, i - 1, k + 1, j)] + rw[X3(d, i - 1, k, j)] + rw[X3(d, i, k + 1, j)] + rw[X3(d, i, k, j)]);
      }
    }
    // rk_addtend_dry
    if (!adv_only && row_ok && k <= kte - 1) {
      const float msf = s.msfuy[X2(d, i, j)];
      float tf = s.ru_tendf[o];
      if (a.rk_step == 1) {
        // u_save: zeroed by rk_tendency (module_em.F:395), or relax_bdy_dry's relaxation
        tf = tf + ((a.c.specified || a.c.nested) ? s.u_save[o] : 0.f) * msf;
        s.ru_tendf[o] = tf;
      }
      t = t + tf / msf;
    }
    s.ru_tend[o] = t;
    vf_lo = vf_hi;
    dpn_lo = dpn_hi;
  }
}

// k_tend_u for the doubly periodic case with four i points per lane (16-byte loads), one
// level per thread.  Element by element the same terms in the same order as k_tend_u with
// PER = true (5th-order fluxes on every face, no lateral-boundary terms); the vertical
// flux and dpn at both w levels of the level are formed directly.
template <typename ARGS>
__global__ void __launch_bounds__(kTallBlk) k_tend_u_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  if (!zrow(a, j)) return;  // rows of an open y edge: the zone form
  const int k = d.kds + (int)T.bz;
  const int ids = d.ids, ide = d.ide, jde = d.jde, kde = d.kde, kts = d.kds, kte = kde, ktf = kte - 1;
  if (i0 > ide || j > jde || k > kte) return;
  const float ts = a.c.time_step >= 0 ? 1.f : -1.f;
  const float rdx = s.rdx, rdy = s.rdy;
  const int ni = d.ni, sj = d.sj;
  const FPtr<float> u = s.u_2, ru = s.ru, rv = s.rv, rom = s.ww, rw = s.rw;
  const Off o2 = X2(d, i0, j), ob = X3(d, i0, 0, j);  // level kk at ob + kk*ni
  auto lv = [&](int kk) { return ob + imax(kts, imin(kk, kte)) * ni; };
  const Off o = lv(k);
  const bool row_ok = j <= jde - 1;
  const bool rk1 = a.rk_step == 1;
  F4 t;
#pragma unroll
  for (int n = 0; n < 4; n++) t.v[n] = 0.f;
  if (row_ok && k <= ktf) {
    const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)], rdnw = s.rdnw[X1(d, k)];
    const F4 msfux = s.msfux.ld<F4>(o2), msfuy = s.msfuy.ld<F4>(o2), msfuyl = s.msfuy.ld<F4>(o2 - 4), msfuyr = s.msfuy.ld<F4>(o2 + 4);
    // advect_u: y faces j and j+1 at the u points
    {
      const F4 ym3 = u.ld<F4>(o - 3 * sj), ym2 = u.ld<F4>(o - 2 * sj), ym1 = u.ld<F4>(o - sj), y0 = u.ld<F4>(o);
      const F4 yp1 = u.ld<F4>(o + sj), yp2 = u.ld<F4>(o + 2 * sj), yp3 = u.ld<F4>(o + 3 * sj);
      const F4 rv0 = rv.ld<F4>(o), rv0l = rv.ld<F4>(o - 4), rv1 = rv.ld<F4>(o + sj), rv1l = rv.ld<F4>(o + sj - 4);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float vel0 = 0.5f * (rv0.v[n] + xm1(rv0l, rv0, n));
        const float vel1 = 0.5f * (rv1.v[n] + xm1(rv1l, rv1, n));
        const float f0 = vel0 * flux5(ym3.v[n], ym2.v[n], ym1.v[n], y0.v[n], yp1.v[n], yp2.v[n], vel0, ts);
        const float f1 = vel1 * flux5(ym2.v[n], ym1.v[n], y0.v[n], yp1.v[n], yp2.v[n], yp3.v[n], vel1, ts);
        const float mrdy = msfux.v[n] * rdy;
        t.v[n] = t.v[n] - mrdy * (f1 - f0);
      }
    }
    // x faces i0 .. i0+4
    {
      const F4 wl = u.ld<F4>(o - 4), wc = u.ld<F4>(o), wr = u.ld<F4>(o + 4);
      const F4 rl = ru.ld<F4>(o - 4), rc = ru.ld<F4>(o), rr = ru.ld<F4>(o + 4);
      float W[12], R[6];
#pragma unroll
      for (int n = 0; n < 4; n++) {
        W[n] = wl.v[n];
        W[n + 4] = wc.v[n];
        W[n + 8] = wr.v[n];
        R[n + 1] = rc.v[n];
      }
      R[0] = rl.v[3];
      R[5] = rr.v[0];
      float fx[5];
#pragma unroll
      for (int n = 0; n < 5; n++) {
        const float vel = 0.5f * (R[n + 1] + R[n]);
        fx[n] = vel * flux5(W[n + 1], W[n + 2], W[n + 3], W[n + 4], W[n + 5], W[n + 6], vel, ts);
      }
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float mrdx = msfux.v[n] * rdx;
        t.v[n] = t.v[n] - mrdx * (fx[n + 1] - fx[n]);
      }
    }
    // vertical flux divergence (order 3) from the fluxes at w levels k and k+1
    {
      const F4 qm2 = u.ld<F4>(lv(k - 2)), qm1 = u.ld<F4>(lv(k - 1)), q0 = u.ld<F4>(o), qp1 = u.ld<F4>(lv(k + 1)),
               qp2 = u.ld<F4>(lv(k + 2));
      const F4 w0 = rom.ld<F4>(o), w0l = rom.ld<F4>(o - 4), w1 = rom.ld<F4>(lv(k + 1)), w1l = rom.ld<F4>(lv(k + 1) - 4);
      auto vflux = [&](int kk, float wc_, float wl_, float qm2_, float qm1_, float q0_, float qp1_) {
        const int kc = imin(kk, kte);
        const float lo = 0.5f * (wc_ + wl_) * (s.fnm[X1(d, kc)] * q0_ + s.fnp[X1(d, kc)] * qm1_);
        const float vel = 0.5f * (wl_ + wc_);
        const float hi = vel * flux3(qm2_, qm1_, q0_, qp1_, -vel, ts);
        return (kk == kts || kk == kte) ? 0.f : ((kk == kts + 1 || kk == ktf) ? lo : hi);
      };
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float vf_lo = vflux(k, w0.v[n], xm1(w0l, w0, n), qm2.v[n], qm1.v[n], q0.v[n], qp1.v[n]);
        const float vf_hi = vflux(k + 1, w1.v[n], xm1(w1l, w1, n), qm1.v[n], q0.v[n], qp1.v[n], qp2.v[n]);
        t.v[n] = t.v[n] - rdnw * (vf_hi - vf_lo);
      }
    }
    // horizontal_pressure_gradient (x)
    {
      const FPtr<float> p = s.p, ph = s.ph_2, al = s.al, alt = s.alt, pb = s.pb, php = s.php;
      const F4 muu = s.muu.ld<F4>(o2), mu = s.mu_2.ld<F4>(o2), mul = s.mu_2.ld<F4>(o2 - 4);
      const Off op = lv(k + 1), om = lv(k - 1);
      const F4 php1 = ph.ld<F4>(op), php1l = ph.ld<F4>(op - 4), ph0 = ph.ld<F4>(o), ph0l = ph.ld<F4>(o - 4);
      const F4 altc = alt.ld<F4>(o), altl = alt.ld<F4>(o - 4), pc = p.ld<F4>(o), pl = p.ld<F4>(o - 4);
      const F4 alc = al.ld<F4>(o), all = al.ld<F4>(o - 4), pbc = pb.ld<F4>(o), pbl = pb.ld<F4>(o - 4);
      const F4 phpc = php.ld<F4>(o), phpl = php.ld<F4>(o - 4), cqu = s.cqu.ld<F4>(o);
      const F4 pp = p.ld<F4>(op), ppl = p.ld<F4>(op - 4), pm = p.ld<F4>(om), pml = p.ld<F4>(om - 4);
      const bool nh = a.c.non_hydrostatic;
      const float fm0 = s.fnm[X1(d, k)], fp0 = s.fnp[X1(d, k)], fm1 = s.fnm[X1(d, imin(k + 1, kte))], fp1 = s.fnp[X1(d, imin(k + 1, kte))];
      // dpn at w level k (extrapolated at k = 1) and k+1 (the lid at kde)
      F4 dlo, dhi;
#pragma unroll
      for (int n = 0; n < 4; n++) {
        dlo.v[n] = .5f * (fm0 * (xm1(pl, pc, n) + pc.v[n]) + fp0 * (xm1(pml, pm, n) + pm.v[n]));
        dhi.v[n] = .5f * (fm1 * (xm1(ppl, pp, n) + pp.v[n]) + fp1 * (xm1(pl, pc, n) + pc.v[n]));
      }
      if (k == 1) {
        const float cf1 = s.cf1, cf2 = s.cf2, cf3 = s.cf3;
        const Off q = ob + ni;
        const F4 c1_ = p.ld<F4>(q), c2_ = p.ld<F4>(q + ni), c3_ = p.ld<F4>(q + 2 * ni);
        const F4 l1 = p.ld<F4>(q - 4), l2 = p.ld<F4>(q + ni - 4), l3 = p.ld<F4>(q + 2 * ni - 4);
#pragma unroll
        for (int n = 0; n < 4; n++)
          dlo.v[n] = .5f * (cf1 * (xm1(l1, c1_, n) + c1_.v[n]) + cf2 * (xm1(l2, c2_, n) + c2_.v[n]) + cf3 * (xm1(l3, c3_, n) + c3_.v[n]));
      }
      if (k + 1 == kde) {
        F4 top;
#pragma unroll
        for (int n = 0; n < 4; n++) top.v[n] = 0.f;
        if (a.c.top_lid) {
          const float cfn = s.cfn, cfn1 = s.cfn1;
          const Off q = ob + (kde - 1) * ni;
          const F4 c1_ = p.ld<F4>(q), c2_ = p.ld<F4>(q - ni), l1 = p.ld<F4>(q - 4), l2 = p.ld<F4>(q - ni - 4);
#pragma unroll
          for (int n = 0; n < 4; n++) top.v[n] = .5f * (cfn * (xm1(l1, c1_, n) + c1_.v[n]) + cfn1 * (xm1(l2, c2_, n) + c2_.v[n]));
        }
        dhi = top;
      }
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float m = msfux.v[n] / msfuy.v[n];
        float dpx = m * .5f * rdx * (c1 * muu.v[n] + c2) *
                    ((php1.v[n] - xm1(php1l, php1, n) + ph0.v[n] - xm1(ph0l, ph0, n)) +
                     (altc.v[n] + xm1(altl, altc, n)) * (pc.v[n] - xm1(pl, pc, n)) +
                     (alc.v[n] + xm1(all, alc, n)) * (pbc.v[n] - xm1(pbl, pbc, n)));
        if (nh) {
          dpx = dpx + m * rdx * (phpc.v[n] - xm1(phpl, phpc, n)) *
                          (rdnw * (dhi.v[n] - dlo.v[n]) - .5f * ((c1 * xm1(mul, mu, n)) + (c1 * mu.v[n])));
          t.v[n] = t.v[n] - cqu.v[n] * dpx;
        } else {
          t.v[n] = t.v[n] - dpx;
        }
      }
    }
    // coriolis and curvature
    {
      const F4 f = s.f.ld<F4>(o2), fl = s.f.ld<F4>(o2 - 4), e = s.e.ld<F4>(o2), el = s.e.ld<F4>(o2 - 4);
      const F4 ca = s.cosa.ld<F4>(o2), cal = s.cosa.ld<F4>(o2 - 4);
      const F4 rv0 = rv.ld<F4>(o), rv0l = rv.ld<F4>(o - 4), rv1 = rv.ld<F4>(o + sj), rv1l = rv.ld<F4>(o + sj - 4);
      const F4 rw0 = rw.ld<F4>(o), rw0l = rw.ld<F4>(o - 4), rw1 = rw.ld<F4>(lv(k + 1)), rw1l = rw.ld<F4>(lv(k + 1) - 4);
      const F4 wl = u.ld<F4>(o - 4), wc = u.ld<F4>(o), wr = u.ld<F4>(o + 4);
      const F4 v0 = s.v_2.ld<F4>(o), v0l = s.v_2.ld<F4>(o - 4), v1 = s.v_2.ld<F4>(o + sj), v1l = s.v_2.ld<F4>(o + sj - 4);
      const F4 mx0 = s.msfvx.ld<F4>(o2), mx0l = s.msfvx.ld<F4>(o2 - 4), mx1 = s.msfvx.ld<F4>(o2 + ni), mx1l = s.msfvx.ld<F4>(o2 + ni - 4);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float rvs = xm1(rv1l, rv1, n) + rv1.v[n] + xm1(rv0l, rv0, n) + rv0.v[n];
        const float rws = xm1(rw1l, rw1, n) + xm1(rw0l, rw0, n) + rw1.v[n] + rw0.v[n];
        t.v[n] = t.v[n] + (msfux.v[n] / msfuy.v[n]) * 0.5f * (f.v[n] + xm1(fl, f, n)) * 0.25f * rvs -
                 0.5f * (e.v[n] + xm1(el, e, n)) * 0.5f * (ca.v[n] + xm1(cal, ca, n)) * 0.25f * rws;
        // vx at i and i-1
        const float ui = wc.v[n], uim = xm1(wl, wc, n), uip = xp1(wc, wr, n);
        const float vx_i = 0.5f * (ui + uip) * (mx1.v[n] - mx0.v[n]) * rdy -
                           0.5f * (v0.v[n] + v1.v[n]) * (xp1(msfuy, msfuyr, n) - msfuy.v[n]) * rdx;
        const float vx_im = 0.5f * (uim + ui) * (xm1(mx1l, mx1, n) - xm1(mx0l, mx0, n)) * rdy -
                            0.5f * (xm1(v0l, v0, n) + xm1(v1l, v1, n)) * (msfuy.v[n] - xm1(msfuyl, msfuy, n)) * rdx;
        t.v[n] = t.v[n] + 0.5f * (vx_i + vx_im) * 0.25f * rvs - ui * K_RERADIUS * 0.25f * rws;
      }
    }
  }
  // rk_addtend_dry
  if (row_ok && k <= kte - 1) {
    const F4 msf = s.msfuy.ld<F4>(o2);
    F4 tf = s.ru_tendf.ld<F4>(o);
#pragma unroll
    for (int n = 0; n < 4; n++) {
      if (rk1) tf.v[n] = tf.v[n] + 0.f * msf.v[n];  // u_save was zeroed by rk_tendency (module_em.F:395)
      t.v[n] = t.v[n] + tf.v[n] / msf.v[n];
    }
    if (rk1) st4m(s.ru_tendf, o, tf, i0, zlo(a, ids), zhi(a, ide));
  }
  st4m(s.ru_tend, o, t, i0, zlo(a, ids), zhi(a, ide));
}

// ------------------------------------------------------------------------------ v
template <typename ARGS, bool PER, bool ZONE>
__global__ void __launch_bounds__(256) k_tend_v(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  int i, j;
  if (!point_ij<ZONE>(a, T, i, j)) return;
  constexpr int KS = ZONE ? 1 : kTendK;  // levels per thread (the zone form: one)
  const int k0 = d.kds + (int)T.bz * KS;
  const int ids = d.ids, ide = d.ide, jds = d.jds, jde = d.jde, kde = d.kde;
  const int its = ids, ite = ide, jts = jds, jte = jde, kts = d.kds, kte = kde;
  if (i > ide || j > jde || k0 > kte) return;
  const int k1 = imin(k0 + KS - 1, kte);
  const Geo g = geo<PER>(a.c);
  const float ts = a.c.time_step >= 0 ? 1.f : -1.f;
  const int ktf = kte - 1;
  const FPtr<float> v = s.v_2, ru = s.ru, rv = s.rv, rom = s.ww, rw = s.rw;
  const float rdx = s.rdx, rdy = s.rdy;
  const bool dxs = !(g.per_x || g.sym_xs || its > ids + 3), dxe = !(g.per_x || g.sym_xe || ite < ide - 3);
  const bool dys = !(g.per_y || g.sym_ys || jts > jds + 3), dye = !(g.per_y || g.sym_ye || jte < jde - 3);
  const bool col_ok = i <= imin(ite, ide - 1);
  // y flux
  int jy0 = jts, jy1 = jte, jf0 = jts, jf1 = jte + 1;
  if (dys) { jy0 = imax(jts, jds + 1); jf0 = jds + 3; }
  if (dye) { jy1 = imin(jte, jde - 1); jf1 = jde - 2; }
  // x flux
  int jx0 = jts, jx1 = jte;
  if (g.open_ys || g.spec) jx0 = imax(jds + 1, jts);
  if (g.open_ye || g.spec) jx1 = imin(jde - 1, jte);
  int ix0 = its, ix1 = imin(ite, ide - 1), if0 = ix0, if1 = ix1 + 1;
  if (dxs) { ix0 = imax(ids + 1, its); if0 = imin(ix0 + 2, ids + 3); }
  if (dxe) { ix1 = imin(ide - 2, ite); if1 = ide - 3; }
  // vertical
  int jv0 = jts, jv1 = jte;
  if (g.open_ys || g.spec) jv0 = imax(jds + 1, jts);
  if (g.open_ye || g.spec) jv1 = imin(jde - 1, jte);
  // pressure gradient
  int jh0 = jts, jh1 = jte;
  if ((g.open_ys || g.spec || g.nested || g.polar) && jts == jds) jh0 = jts + 1;
  if ((g.open_ye || g.spec || g.nested || g.polar) && jte == jde) jh1 = jte - 1;
  // coriolis / curvature rows
  int jc0 = jts, jc1 = jte;
  if (g.open_ys || g.spec || g.nested || g.polar) jc0 = imax(jds + 1, jts);
  if (g.open_ye || g.spec || g.nested || g.polar) jc1 = imin(jde - 1, jte);
  // curvature vxgm
  const bool cpx_s = (g.open_xs || (g.spec && !g.per_x) || g.nested) && its == ids;
  const bool cpx_e = (g.open_xe || (g.spec && !g.per_x) || g.nested) && ite == ide;
  const bool cpy_s = (g.open_ys || g.spec || g.nested || g.polar) && jts == jds;
  const bool cpy_e = (g.open_ye || g.spec || g.nested || g.polar) && jte == jde;
  const FPtr<float> msfvx = s.msfvx, msfuy = s.msfuy, uu = s.u_2;
  auto vx_raw = [&](int ii, int k, int jj) -> float {
    return 0.5f * (uu[X3(d, ii, k, jj)] + uu[X3(d, ii + 1, k, jj)]) * (msfvx[X2(d, ii, jj + 1)] - msfvx[X2(d, ii, jj)]) * rdy -
           0.5f * (v[X3(d, ii, k, jj)] + v[X3(d, ii, k, jj + 1)]) * (msfuy[X2(d, ii + 1, jj)] - msfuy[X2(d, ii, jj)]) * rdx;
  };
  auto vx = [&](int ii, int k, int jj) -> float {
    if (cpy_s && jj == jts - 1) jj = jts;
    else if (cpy_e && jj == jte) jj = jte - 1;
    if (jj >= jts && jj <= jte - 1) {
      if (cpx_s && ii == its - 1) ii = its;
      else if (cpx_e && ii == ite) ii = ite - 1;
    }
    return vx_raw(ii, k, jj);
  };
  auto fqy = [&](int k, int jj) -> float {
    if (jj >= jf0 && jj <= jf1) {
      const float vel = 0.5f * (rv[X3(d, i, k, jj)] + rv[X3(d, i, k, jj - 1)]);
      return vel * flux5(v[X3(d, i, k, jj - 3)], v[X3(d, i, k, jj - 2)], v[X3(d, i, k, jj - 1)], v[X3(d, i, k, jj)],
                         v[X3(d, i, k, jj + 1)], v[X3(d, i, k, jj + 2)], vel, ts);
    } else if (jj == jds + 1) {
      float vb = v[X3(d, i, k, jj - 1)];
      if (g.spec && v[X3(d, i, k, jj)] < 0.f) vb = v[X3(d, i, k, jj)];
      return 0.25f * (rv[X3(d, i, k, jj)] + rv[X3(d, i, k, jj - 1)]) * (v[X3(d, i, k, jj)] + vb);
    } else if (jj == jde) {
      float vb = v[X3(d, i, k, jj)];
      if (g.spec && v[X3(d, i, k, jj - 1)] > 0.f) vb = v[X3(d, i, k, jj - 1)];
      return 0.25f * (rv[X3(d, i, k, jj)] + rv[X3(d, i, k, jj - 1)]) * (vb + v[X3(d, i, k, jj - 1)]);
    } else {
      const float vel = 0.5f * (rv[X3(d, i, k, jj)] + rv[X3(d, i, k, jj - 1)]);
      return vel * flux3(v[X3(d, i, k, jj - 2)], v[X3(d, i, k, jj - 1)], v[X3(d, i, k, jj)], v[X3(d, i, k, jj + 1)], vel, ts);
    }
  };
  auto fqx = [&](int k, int ii) -> float {
    if (ii >= if0 && ii <= if1) {
      const float vel = 0.5f * (ru[X3(d, ii, k, j)] + ru[X3(d, ii, k, j - 1)]);
      return vel * flux5(v[X3(d, ii - 3, k, j)], v[X3(d, ii - 2, k, j)], v[X3(d, ii - 1, k, j)], v[X3(d, ii, k, j)],
                         v[X3(d, ii + 1, k, j)], v[X3(d, ii + 2, k, j)], vel, ts);
    } else if (ii == ids + 1 || ii == ide - 1) {
      return 0.25f * (ru[X3(d, ii, k, j)] + ru[X3(d, ii, k, j - 1)]) * (v[X3(d, ii, k, j)] + v[X3(d, ii - 1, k, j)]);
    } else {
      const float vel = 0.5f * (ru[X3(d, ii, k, j)] + ru[X3(d, ii, k, j - 1)]);
      return vel * flux3(v[X3(d, ii - 2, k, j)], v[X3(d, ii - 1, k, j)], v[X3(d, ii, k, j)], v[X3(d, ii + 1, k, j)], vel, ts);
    }
  };
  auto vflux = [&](int k) -> float {
    if (k == kts || k == kte) return 0.f;
    if (k == kts + 1 || k == ktf)
      return 0.5f * (rom[X3(d, i, k, j)] + rom[X3(d, i, k, j - 1)]) *
             (s.fnm[X1(d, k)] * v[X3(d, i, k, j)] + s.fnp[X1(d, k)] * v[X3(d, i, k - 1, j)]);
    const float vel = 0.5f * (rom[X3(d, i, k, j)] + rom[X3(d, i, k, j - 1)]);
    return vel * flux3(v[X3(d, i, k - 2, j)], v[X3(d, i, k - 1, j)], v[X3(d, i, k, j)], v[X3(d, i, k + 1, j)], -vel, ts);
  };
  const FPtr<float> p = s.p, ph = s.ph_2, al = s.al, alt = s.alt, pb = s.pb, php = s.php, mu = s.mu_2;
  const float cf1 = s.cf1, cf2 = s.cf2, cf3 = s.cf3, cfn = s.cfn, cfn1 = s.cfn1;
  const bool nh = a.c.non_hydrostatic;
  auto dpn = [&](int k) -> float {
    if (k == 1)
      return .5f * (cf1 * (p[X3(d, i, 1, j - 1)] + p[X3(d, i, 1, j)]) + cf2 * (p[X3(d, i, 2, j - 1)] + p[X3(d, i, 2, j)]) +
                    cf3 * (p[X3(d, i, 3, j - 1)] + p[X3(d, i, 3, j)]));
    if (k == kde) {
      if (!a.c.top_lid) return 0.f;
      return .5f * (cfn * (p[X3(d, i, kde - 1, j - 1)] + p[X3(d, i, kde - 1, j)]) + cfn1 * (p[X3(d, i, kde - 2, j - 1)] + p[X3(d, i, kde - 2, j)]));
    }
    return .5f * (s.fnm[X1(d, k)] * (p[X3(d, i, k, j - 1)] + p[X3(d, i, k, j)]) + s.fnp[X1(d, k)] * (p[X3(d, i, k - 1, j - 1)] + p[X3(d, i, k - 1, j)]));
  };
  float vf_lo = (col_ok && k0 <= ktf && j >= jv0 && j <= jv1) ? vflux(k0) : 0.f;
  // advect_v alone (a per-routine entry point): added to the tendency array's values
  const bool adv_only = a.routine == R_ADVECT;
  float dpn_lo = (col_ok && k0 <= ktf && nh && !adv_only && j >= jh0 && j <= jh1) ? dpn(k0) : 0.f;
  for (int k = k0; k <= k1; k++) {
    const Off o = X3(d, i, k, j);
    float t = adv_only ? s.rv_tend[o] : 0.f, vf_hi = 0.f, dpn_hi = 0.f;
    if (col_ok && k <= ktf) {
      if (j >= jy0 && j <= jy1) {
        const float mrdy = s.msfvy[X2(d, i, j)] * rdy;
        t = t - mrdy * (fqy(k, j + 1) - fqy(k, j));
      }
      if (j >= jx0 && j <= jx1 && i >= ix0 && i <= ix1) {
        const float mrdx = s.msfvy[X2(d, i, j)] * rdx;
        t = t - mrdx * (fqx(k, i + 1) - fqx(k, i));
      }
      const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
      if (g.open_ys && j == jts) {
        float vb = rv[o] - K_CB * (c1 * s.mut[X2(d, i, jts)] + c2);
        if (vb > 0.f) vb = 0.f;
        t = t - rdy * vb * (v[X3(d, i, k, jts + 1)] - v[o]);
      }
      if (g.open_ye && j == jte) {
        float vb = rv[o] + K_CB * (c1 * s.mut[X2(d, i, jte - 1)] + c2);
        if (vb < 0.f) vb = 0.f;
        t = t - rdy * vb * (v[o] - v[X3(d, i, k, jte - 1)]);
      }
      {
        int jo0 = jts, jo1 = imin(jte, jde), jmn = jds, jmx = jde - 1;
        if (g.open_ys) jo0 = imax(jds + 1, jts);
        if (g.open_ye) jo1 = imin(jte, jde - 1);
        const int jp = imin(jmx, j), jm = imax(jmn, j - 1);
        if (g.open_xs && i == its && j >= jo0 && j <= jo1) {
          const float mrdx = s.msfvy[X2(d, its, j)] * rdx;
          const float uw = 0.5f * (ru[X3(d, its, k, jp)] + ru[X3(d, its, k, jm)]);
          const float ub = uw < 0.f ? uw : 0.f;
          const float dup = ru[X3(d, its + 1, k, jp)] - ru[X3(d, its, k, jp)];
          const float dum = ru[X3(d, its + 1, k, jm)] - ru[X3(d, its, k, jm)];
          t = t - mrdx * (ub * (v[X3(d, its + 1, k, j)] - v[X3(d, its, k, j)]) + 0.5f * v[X3(d, its, k, j)] * (dup + dum));
        }
        if (g.open_xe && i == ite - 1 && j >= jo0 && j <= jo1) {
          const float mrdx = s.msfvy[X2(d, ite - 1, j)] * rdx;
          const float uw = 0.5f * (ru[X3(d, ite, k, jp)] + ru[X3(d, ite, k, jm)]);
          const float ub = uw > 0.f ? uw : 0.f;
          const float dup = ru[X3(d, ite, k, jp)] - ru[X3(d, ite - 1, k, jp)];
          const float dum = ru[X3(d, ite, k, jm)] - ru[X3(d, ite - 1, k, jm)];
          t = t - mrdx * (ub * (v[X3(d, ite - 1, k, j)] - v[X3(d, ite - 2, k, j)]) + 0.5f * v[X3(d, ite - 1, k, j)] * (dup + dum));
        }
      }
      if (j >= jv0 && j <= jv1) {
        vf_hi = vflux(k + 1);
        t = t - (s.msfvy[X2(d, i, j)] / s.msfvx[X2(d, i, j)]) * s.rdnw[X1(d, k)] * (vf_hi - vf_lo);
      }
      if (!adv_only && j >= jh0 && j <= jh1) {
        const float m = s.msfvy[X2(d, i, j)] / s.msfvx[X2(d, i, j)];
        float dpy = m * .5f * rdy * (c1 * s.muv[X2(d, i, j)] + c2) *
                    ((ph[X3(d, i, k + 1, j)] - ph[X3(d, i, k + 1, j - 1)] + ph[o] - ph[X3(d, i, k, j - 1)]) +
                     (alt[o] + alt[X3(d, i, k, j - 1)]) * (p[o] - p[X3(d, i, k, j - 1)]) +
                     (al[o] + al[X3(d, i, k, j - 1)]) * (pb[o] - pb[X3(d, i, k, j - 1)]));
        if (nh) {
          dpn_hi = dpn(k + 1);
          dpy = dpy + m * rdy * (php[o] - php[X3(d, i, k, j - 1)]) *
                          (s.rdnw[X1(d, k)] * (dpn_hi - dpn_lo) - .5f * ((c1 * mu[X2(d, i, j - 1)]) + (c1 * mu[X2(d, i, j)])));
          t = t - s.cqv[o] * dpy;
        } else {
          t = t - dpy;
        }
      }
      if (!adv_only && j >= jc0 && j <= jc1) {
        const float m = s.msfvy[X2(d, i, j)] / s.msfvx[X2(d, i, j)];
        t = t - m * 0.5f * (s.f[X2(d, i, j)] + s.f[X2(d, i, j - 1)]) * 0.25f *
                    (ru[X3(d, i, k, j)] + ru[X3(d, i + 1, k, j)] + ru[X3(d, i, k, j - 1)] + ru[X3(d, i + 1, k, j - 1)]) +
            m * 0.5f * (s.e[X2(d, i, j)] + s.e[X2(d, i, j - 1)]) * 0.5f * (s.sina[X2(d, i, j)] + s.sina[X2(d, i, j - 1)]) * 0.25f *
                (rw[X3(d, i, k + 1, j - 1)] + rw[X3(d, i, k, j - 1)] + rw[X3(d, i, k + 1, j)] + rw[X3(d, i, k, j)]);
        t = t - 0.5f * (vx(i, k, j) + vx(i, k, j - 1)) * 0.25f *
                    (ru[X3(d, i, k, j)] + ru[X3(d, i + 1, k, j)] + ru[X3(d, i, k, j - 1)] + ru[X3(d, i + 1, k, j - 1)]) -
            m * v[o] * K_RERADIUS * 0.25f * (rw[X3(d, i, k + 1, j - 1)] + rw[X3(d, i, k, j - 1)] + rw[X3(d, i, k + 1, j)] + rw[X3(d, i, k, j)]);
      }
    }
    if (!adv_only && col_ok && k <= kte - 1) {
      float tf = s.rv_tendf[o];
      if (a.rk_step == 1) {
        tf = tf + ((a.c.specified || a.c.nested) ? s.v_save[o] : 0.f) * s.msfvx[X2(d, i, j)];  // v_save: zero or relaxation
        s.rv_tendf[o] = tf;
      }
      t = t + tf * s.msfvx_inv[X2(d, i, j)];
    }
    s.rv_tend[o] = t;
    vf_lo = vf_hi;
    dpn_lo = dpn_hi;
  }
}

// k_tend_v for the doubly periodic case with four i points per lane, one level per thread;
// element by element the terms of k_tend_v with PER = true, in its order.
template <typename ARGS>
__global__ void __launch_bounds__(kTallBlk) k_tend_v_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  if (!zrow(a, j)) return;  // rows of an open y edge: the zone form
  const int k = d.kds + (int)T.bz;
  const int ids = d.ids, ide = d.ide, jde = d.jde, kde = d.kde, kts = d.kds, kte = kde, ktf = kte - 1;
  if (i0 > ide || j > jde || k > kte) return;
  const float ts = a.c.time_step >= 0 ? 1.f : -1.f;
  const float rdx = s.rdx, rdy = s.rdy;
  const int ni = d.ni, sj = d.sj;
  const FPtr<float> v = s.v_2, ru = s.ru, rv = s.rv, rom = s.ww, rw = s.rw;
  const Off o2 = X2(d, i0, j), ob = X3(d, i0, 0, j);
  auto lv = [&](int kk) { return ob + imax(kts, imin(kk, kte)) * ni; };
  const Off o = lv(k);
  const bool rk1 = a.rk_step == 1;
  F4 t;
#pragma unroll
  for (int n = 0; n < 4; n++) t.v[n] = 0.f;
  if (k <= ktf) {
    const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)], rdnw = s.rdnw[X1(d, k)];
    const F4 msfvy = s.msfvy.ld<F4>(o2), msfvx = s.msfvx.ld<F4>(o2);
    // advect_v: y faces j and j+1
    {
      const F4 ym3 = v.ld<F4>(o - 3 * sj), ym2 = v.ld<F4>(o - 2 * sj), ym1 = v.ld<F4>(o - sj), y0 = v.ld<F4>(o);
      const F4 yp1 = v.ld<F4>(o + sj), yp2 = v.ld<F4>(o + 2 * sj), yp3 = v.ld<F4>(o + 3 * sj);
      const F4 rvm = rv.ld<F4>(o - sj), rv0 = rv.ld<F4>(o), rv1 = rv.ld<F4>(o + sj);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float vel0 = 0.5f * (rv0.v[n] + rvm.v[n]);
        const float vel1 = 0.5f * (rv1.v[n] + rv0.v[n]);
        const float f0 = vel0 * flux5(ym3.v[n], ym2.v[n], ym1.v[n], y0.v[n], yp1.v[n], yp2.v[n], vel0, ts);
        const float f1 = vel1 * flux5(ym2.v[n], ym1.v[n], y0.v[n], yp1.v[n], yp2.v[n], yp3.v[n], vel1, ts);
        const float mrdy = msfvy.v[n] * rdy;
        t.v[n] = t.v[n] - mrdy * (f1 - f0);
      }
    }
    // x faces i0 .. i0+4 (face ii between mass points ii-1 and ii)
    {
      const F4 wl = v.ld<F4>(o - 4), wc = v.ld<F4>(o), wr = v.ld<F4>(o + 4);
      const F4 r0 = ru.ld<F4>(o), r0r = ru.ld<F4>(o + 4), rm = ru.ld<F4>(o - sj), rmr = ru.ld<F4>(o - sj + 4);
      float W[12];
#pragma unroll
      for (int n = 0; n < 4; n++) {
        W[n] = wl.v[n];
        W[n + 4] = wc.v[n];
        W[n + 8] = wr.v[n];
      }
      float fx[5];
#pragma unroll
      for (int n = 0; n < 5; n++) {
        const float rj = n < 4 ? r0.v[n] : r0r.v[0], rjm = n < 4 ? rm.v[n] : rmr.v[0];
        const float vel = 0.5f * (rj + rjm);
        fx[n] = vel * flux5(W[n + 1], W[n + 2], W[n + 3], W[n + 4], W[n + 5], W[n + 6], vel, ts);
      }
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float mrdx = msfvy.v[n] * rdx;
        t.v[n] = t.v[n] - mrdx * (fx[n + 1] - fx[n]);
      }
    }
    // vertical flux divergence
    {
      const F4 qm2 = v.ld<F4>(lv(k - 2)), qm1 = v.ld<F4>(lv(k - 1)), q0 = v.ld<F4>(o), qp1 = v.ld<F4>(lv(k + 1)),
               qp2 = v.ld<F4>(lv(k + 2));
      const F4 w0 = rom.ld<F4>(o), w0s = rom.ld<F4>(o - sj), w1 = rom.ld<F4>(lv(k + 1)), w1s = rom.ld<F4>(lv(k + 1) - sj);
      auto vflux = [&](int kk, float wc_, float ws_, float qm2_, float qm1_, float q0_, float qp1_) {
        const int kc = imin(kk, kte);
        const float vel = 0.5f * (wc_ + ws_);
        const float lo = vel * (s.fnm[X1(d, kc)] * q0_ + s.fnp[X1(d, kc)] * qm1_);
        const float hi = vel * flux3(qm2_, qm1_, q0_, qp1_, -vel, ts);
        return (kk == kts || kk == kte) ? 0.f : ((kk == kts + 1 || kk == ktf) ? lo : hi);
      };
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float vf_lo = vflux(k, w0.v[n], w0s.v[n], qm2.v[n], qm1.v[n], q0.v[n], qp1.v[n]);
        const float vf_hi = vflux(k + 1, w1.v[n], w1s.v[n], qm1.v[n], q0.v[n], qp1.v[n], qp2.v[n]);
        t.v[n] = t.v[n] - (msfvy.v[n] / msfvx.v[n]) * rdnw * (vf_hi - vf_lo);
      }
    }
    // horizontal_pressure_gradient (y)
    {
      const FPtr<float> p = s.p, ph = s.ph_2, al = s.al, alt = s.alt, pb = s.pb, php = s.php;
      const F4 muv = s.muv.ld<F4>(o2), mu = s.mu_2.ld<F4>(o2), mus = s.mu_2.ld<F4>(o2 - ni);
      const Off op = lv(k + 1), om = lv(k - 1);
      const F4 h1 = ph.ld<F4>(op), h1s = ph.ld<F4>(op - sj), h0 = ph.ld<F4>(o), h0s = ph.ld<F4>(o - sj);
      const F4 altc = alt.ld<F4>(o), alts = alt.ld<F4>(o - sj), pc = p.ld<F4>(o), ps = p.ld<F4>(o - sj);
      const F4 alc = al.ld<F4>(o), als = al.ld<F4>(o - sj), pbc = pb.ld<F4>(o), pbs = pb.ld<F4>(o - sj);
      const F4 phpc = php.ld<F4>(o), phps = php.ld<F4>(o - sj), cqv = s.cqv.ld<F4>(o);
      const F4 pp = p.ld<F4>(op), pps = p.ld<F4>(op - sj), pm = p.ld<F4>(om), pms = p.ld<F4>(om - sj);
      const bool nh = a.c.non_hydrostatic;
      const float fm0 = s.fnm[X1(d, k)], fp0 = s.fnp[X1(d, k)], fm1 = s.fnm[X1(d, imin(k + 1, kte))], fp1 = s.fnp[X1(d, imin(k + 1, kte))];
      F4 dlo, dhi;
#pragma unroll
      for (int n = 0; n < 4; n++) {
        dlo.v[n] = .5f * (fm0 * (ps.v[n] + pc.v[n]) + fp0 * (pms.v[n] + pm.v[n]));
        dhi.v[n] = .5f * (fm1 * (pps.v[n] + pp.v[n]) + fp1 * (ps.v[n] + pc.v[n]));
      }
      if (k == 1) {
        const float cf1 = s.cf1, cf2 = s.cf2, cf3 = s.cf3;
        const Off q = ob + ni;
        const F4 c1_ = p.ld<F4>(q), c2_ = p.ld<F4>(q + ni), c3_ = p.ld<F4>(q + 2 * ni);
        const F4 s1 = p.ld<F4>(q - sj), s2 = p.ld<F4>(q + ni - sj), s3 = p.ld<F4>(q + 2 * ni - sj);
#pragma unroll
        for (int n = 0; n < 4; n++)
          dlo.v[n] = .5f * (cf1 * (s1.v[n] + c1_.v[n]) + cf2 * (s2.v[n] + c2_.v[n]) + cf3 * (s3.v[n] + c3_.v[n]));
      }
      if (k + 1 == kde) {
        F4 top;
#pragma unroll
        for (int n = 0; n < 4; n++) top.v[n] = 0.f;
        if (a.c.top_lid) {
          const float cfn = s.cfn, cfn1 = s.cfn1;
          const Off q = ob + (kde - 1) * ni;
          const F4 c1_ = p.ld<F4>(q), c2_ = p.ld<F4>(q - ni), s1 = p.ld<F4>(q - sj), s2 = p.ld<F4>(q - ni - sj);
#pragma unroll
          for (int n = 0; n < 4; n++) top.v[n] = .5f * (cfn * (s1.v[n] + c1_.v[n]) + cfn1 * (s2.v[n] + c2_.v[n]));
        }
        dhi = top;
      }
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float m = msfvy.v[n] / msfvx.v[n];
        float dpy = m * .5f * rdy * (c1 * muv.v[n] + c2) *
                    ((h1.v[n] - h1s.v[n] + h0.v[n] - h0s.v[n]) +
                     (altc.v[n] + alts.v[n]) * (pc.v[n] - ps.v[n]) +
                     (alc.v[n] + als.v[n]) * (pbc.v[n] - pbs.v[n]));
        if (nh) {
          dpy = dpy + m * rdy * (phpc.v[n] - phps.v[n]) * (rdnw * (dhi.v[n] - dlo.v[n]) - .5f * ((c1 * mus.v[n]) + (c1 * mu.v[n])));
          t.v[n] = t.v[n] - cqv.v[n] * dpy;
        } else {
          t.v[n] = t.v[n] - dpy;
        }
      }
    }
    // coriolis and curvature
    {
      const F4 f = s.f.ld<F4>(o2), fs = s.f.ld<F4>(o2 - ni), e = s.e.ld<F4>(o2), es = s.e.ld<F4>(o2 - ni);
      const F4 sa = s.sina.ld<F4>(o2), sas = s.sina.ld<F4>(o2 - ni);
      const F4 r0 = ru.ld<F4>(o), r0r = ru.ld<F4>(o + 4), rm = ru.ld<F4>(o - sj), rmr = ru.ld<F4>(o - sj + 4);
      const F4 rw0 = rw.ld<F4>(o), rw0s = rw.ld<F4>(o - sj), rw1 = rw.ld<F4>(lv(k + 1)), rw1s = rw.ld<F4>(lv(k + 1) - sj);
      const F4 u0 = s.u_2.ld<F4>(o), u0r = s.u_2.ld<F4>(o + 4), um = s.u_2.ld<F4>(o - sj), umr = s.u_2.ld<F4>(o - sj + 4);
      const F4 vm = v.ld<F4>(o - sj), v0 = v.ld<F4>(o), vp = v.ld<F4>(o + sj);
      const F4 mxm = s.msfvx.ld<F4>(o2 - ni), mx0 = msfvx, mxp = s.msfvx.ld<F4>(o2 + ni);
      const F4 my0 = s.msfuy.ld<F4>(o2), my0r = s.msfuy.ld<F4>(o2 + 4), mym = s.msfuy.ld<F4>(o2 - ni), mymr = s.msfuy.ld<F4>(o2 - ni + 4);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float m = msfvy.v[n] / msfvx.v[n];
        const float rus = r0.v[n] + xp1(r0, r0r, n) + rm.v[n] + xp1(rm, rmr, n);
        const float rws = rw1s.v[n] + rw0s.v[n] + rw1.v[n] + rw0.v[n];
        t.v[n] = t.v[n] - m * 0.5f * (f.v[n] + fs.v[n]) * 0.25f * rus +
                 m * 0.5f * (e.v[n] + es.v[n]) * 0.5f * (sa.v[n] + sas.v[n]) * 0.25f * rws;
        const float vx_j = 0.5f * (u0.v[n] + xp1(u0, u0r, n)) * (mxp.v[n] - mx0.v[n]) * rdy -
                           0.5f * (v0.v[n] + vp.v[n]) * (xp1(my0, my0r, n) - my0.v[n]) * rdx;
        const float vx_jm = 0.5f * (um.v[n] + xp1(um, umr, n)) * (mx0.v[n] - mxm.v[n]) * rdy -
                            0.5f * (vm.v[n] + v0.v[n]) * (xp1(mym, mymr, n) - mym.v[n]) * rdx;
        t.v[n] = t.v[n] - 0.5f * (vx_j + vx_jm) * 0.25f * rus - m * v0.v[n] * K_RERADIUS * 0.25f * rws;
      }
    }
  }
  // rk_addtend_dry
  if (k <= kte - 1) {
    const F4 mvx = s.msfvx.ld<F4>(o2), mvi = s.msfvx_inv.ld<F4>(o2);
    F4 tf = s.rv_tendf.ld<F4>(o);
#pragma unroll
    for (int n = 0; n < 4; n++) {
      if (rk1) tf.v[n] = tf.v[n] + 0.f * mvx.v[n];  // v_save was zeroed by rk_tendency
      t.v[n] = t.v[n] + tf.v[n] * mvi.v[n];
    }
    if (rk1) st4m(s.rv_tendf, o, tf, i0, zlo(a, ids), zhi(a, ide - 1));
  }
  // col_ok: only i <= ide-1 carry the tendency; i = ide is stored as zero
#pragma unroll
  for (int n = 0; n < 4; n++)
    if (i0 + n > ide - 1) t.v[n] = 0.f;
  st4m(s.rv_tend, o, t, i0, zlo(a, ids), zhi(a, ide));
}

// ------------------------------------------------------------------------------ w
template <typename ARGS, bool PER, bool ZONE>
__global__ void __launch_bounds__(256) k_tend_w(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  int i, j;
  if (!point_ij<ZONE>(a, T, i, j)) return;
  constexpr int KS = ZONE ? 1 : kTendK;  // levels per thread (the zone form: one)
  const int k0 = d.kds + (int)T.bz * KS;
  const int ids = d.ids, ide = d.ide, jds = d.jds, jde = d.jde, kde = d.kde;
  const int its = ids, ite = ide, jts = jds, jte = jde, kts = d.kds, kte = kde;
  if (i > ide || j > jde || k0 > kte) return;
  const int k1 = imin(k0 + KS - 1, kte);
  const Geo g = geo<PER>(a.c);
  const float ts = a.c.time_step >= 0 ? 1.f : -1.f;
  const int ktf = kte - 1;
  const FPtr<float> w = s.w_2, ru = s.ru, rv = s.rv, rom = s.ww;
  const CPtr<float> fzm = s.fnm, fzp = s.fnp;
  const float rdx = s.rdx, rdy = s.rdy;
  const bool mass = i <= imin(ite, ide - 1) && j <= imin(jte, jde - 1);
  const bool dxs = !(g.per_x || g.sym_xs || its > ids + 3), dxe = !(g.per_x || g.sym_xe || ite < ide - 3);
  const bool dys = !(g.per_y || g.sym_ys || jts > jds + 3), dye = !(g.per_y || g.sym_ye || jte < jde - 4);
  int jy0 = jts, jy1 = imin(jte, jde - 1), jf0 = jy0, jf1 = jy1 + 1;
  if (dys) { jy0 = imax(jts, jds + 1); jf0 = jds + 3; }
  if (dye) { jy1 = imin(jte, jde - 2); jf1 = jde - 3; }
  int ix0 = its, ix1 = imin(ite, ide - 1), if0 = ix0, if1 = ix1 + 1;
  if (dxs) { ix0 = imax(ids + 1, its); if0 = imin(ix0 + 2, ids + 3); }
  if (dxe) { ix1 = imin(ide - 2, ite); if1 = ide - 3; }
  // face velocity at w level k (k <= ktf) or at the lid (k = ktf+1)
  auto wvel = [&](FPtr<float> rr, int ii, int k, int jj) -> float {
    if (k <= ktf) return fzm[X1(d, k)] * rr[X3(d, ii, k, jj)] + fzp[X1(d, k)] * rr[X3(d, ii, k - 1, jj)];
    return (2.f - fzm[X1(d, k - 1)]) * rr[X3(d, ii, k - 1, jj)] - fzp[X1(d, k - 1)] * rr[X3(d, ii, k - 2, jj)];
  };
  auto fqy = [&](int k, int jj) -> float {
    if (jj >= jf0 && jj <= jf1) {
      const float vel = wvel(rv, i, k, jj);
      return vel * flux5(w[X3(d, i, k, jj - 3)], w[X3(d, i, k, jj - 2)], w[X3(d, i, k, jj - 1)], w[X3(d, i, k, jj)],
                         w[X3(d, i, k, jj + 1)], w[X3(d, i, k, jj + 2)], vel, ts);
    } else if (jj == jds + 1 || jj == jde - 1) {
      return 0.5f * (wvel(rv, i, k, jj)) * (w[X3(d, i, k, jj)] + w[X3(d, i, k, jj - 1)]);
    } else {
      const float vel = wvel(rv, i, k, jj);
      return vel * flux3(w[X3(d, i, k, jj - 2)], w[X3(d, i, k, jj - 1)], w[X3(d, i, k, jj)], w[X3(d, i, k, jj + 1)], vel, ts);
    }
  };
  auto fqx = [&](int k, int ii) -> float {
    if (ii >= if0 && ii <= if1) {
      const float vel = wvel(ru, ii, k, j);
      return vel * flux5(w[X3(d, ii - 3, k, j)], w[X3(d, ii - 2, k, j)], w[X3(d, ii - 1, k, j)], w[X3(d, ii, k, j)],
                         w[X3(d, ii + 1, k, j)], w[X3(d, ii + 2, k, j)], vel, ts);
    } else if (ii == ids + 1 || ii == ide - 1) {
      return 0.5f * (wvel(ru, ii, k, j)) * (w[X3(d, ii, k, j)] + w[X3(d, ii - 1, k, j)]);
    } else {
      const float vel = wvel(ru, ii, k, j);
      return vel * flux3(w[X3(d, ii - 2, k, j)], w[X3(d, ii - 1, k, j)], w[X3(d, ii, k, j)], w[X3(d, ii + 1, k, j)], vel, ts);
    }
  };
  auto vflux = [&](int k) -> float {  // k in kts+1 .. ktf+1
    if (k == kts + 1 || k == ktf + 1)
      return 0.25f * (rom[X3(d, i, k, j)] + rom[X3(d, i, k - 1, j)]) * (w[X3(d, i, k, j)] + w[X3(d, i, k - 1, j)]);
    const float vel = 0.5f * (rom[X3(d, i, k, j)] + rom[X3(d, i, k - 1, j)]);
    return vel * flux3(w[X3(d, i, k - 2, j)], w[X3(d, i, k - 1, j)], w[X3(d, i, k, j)], w[X3(d, i, k + 1, j)], -vel, ts);
  };
  // advect_w alone (a per-routine entry point): added to the tendency array's values
  const bool adv_only = a.routine == R_ADVECT;
  const float msfty = mass ? s.msfty[X2(d, i, j)] : 1.f;
  const float mu = (mass && !adv_only) ? s.mu_2[X2(d, i, j)] : 0.f, mub = (mass && !adv_only) ? s.mub[X2(d, i, j)] : 0.f;
  const float mut = mass ? s.mut[X2(d, i, j)] : 0.f;
  float vf_c = (mass && imax(k0, kts + 1) <= kte) ? vflux(imax(k0, kts + 1)) : 0.f;  // vflux at level k
  for (int k = k0; k <= k1; k++) {
    const Off o = X3(d, i, k, j);
    float t = adv_only ? s.rw_tend[o] : 0.f;
    if (mass) {
      if (k >= kts + 1) {
        if (j >= jy0 && j <= jy1) {
          const float mrdy = s.msftx[X2(d, i, j)] * rdy;
          t = t - mrdy * (fqy(k, j + 1) - fqy(k, j));
        }
        if (i >= ix0 && i <= ix1) {
          const float mrdx = s.msftx[X2(d, i, j)] * rdx;
          t = t - mrdx * (fqx(k, i + 1) - fqx(k, i));
        }
        // open boundaries
        if (g.open_xs && i == its) {
          float uw, dru;
          if (k <= ktf) {
            uw = 0.5f * (fzm[X1(d, k)] * (ru[X3(d, its, k, j)] + ru[X3(d, its + 1, k, j)]) + fzp[X1(d, k)] * (ru[X3(d, its, k - 1, j)] + ru[X3(d, its + 1, k - 1, j)]));
            dru = fzm[X1(d, k)] * (ru[X3(d, its + 1, k, j)] - ru[X3(d, its, k, j)]) + fzp[X1(d, k)] * (ru[X3(d, its + 1, k - 1, j)] - ru[X3(d, its, k - 1, j)]);
          } else {
            uw = 0.5f * ((2.f - fzm[X1(d, k - 1)]) * (ru[X3(d, its, k - 1, j)] + ru[X3(d, its + 1, k - 1, j)]) - fzp[X1(d, k - 1)] * (ru[X3(d, its, k - 2, j)] + ru[X3(d, its + 1, k - 2, j)]));
            dru = (2.f - fzm[X1(d, k - 1)]) * (ru[X3(d, its + 1, k - 1, j)] - ru[X3(d, its, k - 1, j)]) - fzp[X1(d, k - 1)] * (ru[X3(d, its + 1, k - 2, j)] - ru[X3(d, its, k - 2, j)]);
          }
          const float ub = uw < 0.f ? uw : 0.f;
          t = t - rdx * (ub * (w[X3(d, its + 1, k, j)] - w[X3(d, its, k, j)]) + w[X3(d, its, k, j)] * dru);
        }
        if (g.open_xe && i == imin(ite, ide - 1)) {
          const int ie = i;
          float uw, dru;
          if (k <= ktf) {
            uw = 0.5f * (fzm[X1(d, k)] * (ru[X3(d, ite - 1, k, j)] + ru[X3(d, ite, k, j)]) + fzp[X1(d, k)] * (ru[X3(d, ite - 1, k - 1, j)] + ru[X3(d, ite, k - 1, j)]));
            dru = fzm[X1(d, k)] * (ru[X3(d, ite, k, j)] - ru[X3(d, ite - 1, k, j)]) + fzp[X1(d, k)] * (ru[X3(d, ite, k - 1, j)] - ru[X3(d, ite - 1, k - 1, j)]);
          } else {
            uw = 0.5f * ((2.f - fzm[X1(d, k - 1)]) * (ru[X3(d, ite - 1, k - 1, j)] + ru[X3(d, ite, k - 1, j)]) - fzp[X1(d, k - 1)] * (ru[X3(d, ite - 1, k - 2, j)] + ru[X3(d, ite, k - 2, j)]));
            dru = (2.f - fzm[X1(d, k - 1)]) * (ru[X3(d, ite, k - 1, j)] - ru[X3(d, ite - 1, k - 1, j)]) - fzp[X1(d, k - 1)] * (ru[X3(d, ite, k - 2, j)] - ru[X3(d, ite - 1, k - 2, j)]);
          }
          const float ub = uw > 0.f ? uw : 0.f;
          t = t - rdx * (ub * (w[X3(d, ie, k, j)] - w[X3(d, ie - 1, k, j)]) + w[X3(d, ie, k, j)] * dru);
        }
        if (g.open_ys && j == jts) {
          float vw, drv;
          if (k <= ktf) {
            vw = 0.5f * (fzm[X1(d, k)] * (rv[X3(d, i, k, jts)] + rv[X3(d, i, k, jts + 1)]) + fzp[X1(d, k)] * (rv[X3(d, i, k - 1, jts)] + rv[X3(d, i, k - 1, jts + 1)]));
            drv = fzm[X1(d, k)] * (rv[X3(d, i, k, jts + 1)] - rv[X3(d, i, k, jts)]) + fzp[X1(d, k)] * (rv[X3(d, i, k - 1, jts + 1)] - rv[X3(d, i, k - 1, jts)]);
          } else {
            vw = 0.5f * ((2.f - fzm[X1(d, k - 1)]) * (rv[X3(d, i, k - 1, jts)] + rv[X3(d, i, k - 1, jts + 1)]) - fzp[X1(d, k - 1)] * (rv[X3(d, i, k - 2, jts)] + rv[X3(d, i, k - 2, jts + 1)]));
            drv = (2.f - fzm[X1(d, k - 1)]) * (rv[X3(d, i, k - 1, jts + 1)] - rv[X3(d, i, k - 1, jts)]) - fzp[X1(d, k - 1)] * (rv[X3(d, i, k - 2, jts + 1)] - rv[X3(d, i, k - 2, jts)]);
          }
          const float vb = vw < 0.f ? vw : 0.f;
          t = t - rdy * (vb * (w[X3(d, i, k, jts + 1)] - w[X3(d, i, k, jts)]) + w[X3(d, i, k, jts)] * drv);
        }
        if (g.open_ye && j == imin(jte, jde - 1)) {
          const int je = j;
          float vw, drv;
          if (k <= ktf) {
            vw = 0.5f * (fzm[X1(d, k)] * (rv[X3(d, i, k, jte - 1)] + rv[X3(d, i, k, jte)]) + fzp[X1(d, k)] * (rv[X3(d, i, k - 1, jte - 1)] + rv[X3(d, i, k - 1, jte)]));
            drv = fzm[X1(d, k)] * (rv[X3(d, i, k, jte)] - rv[X3(d, i, k, jte - 1)]) + fzp[X1(d, k)] * (rv[X3(d, i, k - 1, jte)] - rv[X3(d, i, k - 1, jte - 1)]);
          } else {
            vw = 0.5f * ((2.f - fzm[X1(d, k - 1)]) * (rv[X3(d, i, k - 1, jte - 1)] + rv[X3(d, i, k - 1, jte)]) - fzp[X1(d, k - 1)] * (rv[X3(d, i, k - 2, jte - 1)] + rv[X3(d, i, k - 2, jte)]));
            drv = (2.f - fzm[X1(d, k - 1)]) * (rv[X3(d, i, k - 1, jte)] - rv[X3(d, i, k - 1, jte - 1)]) - fzp[X1(d, k - 1)] * (rv[X3(d, i, k - 2, jte)] - rv[X3(d, i, k - 2, jte - 1)]);
          }
          const float vb = vw > 0.f ? vw : 0.f;
          t = t - rdy * (vb * (w[X3(d, i, k, je)] - w[X3(d, i, k, je - 1)]) + w[X3(d, i, k, je)] * drv);
        }
        // vertical advection of w (order 3) with the lid flux term
        if (k <= ktf) {
          const float vf_hi = vflux(k + 1);
          t = t - s.rdn[X1(d, k)] * (vf_hi - vf_c);
          vf_c = vf_hi;
        } else {
          t = t + 2.f * s.rdn[X1(d, k - 1)] * (vf_c);
        }
        // pg_buoy_w
        if (adv_only) {
        } else if (k == kde) {
          // The top row reads the moisture coefficient cqw(kde-1) of calc_cq before the
          // thread of level kde-1 replaces it by 1/(1+cqw): recompute it as calc_cq does
          // (module_big_step_utilities_em.F:787-906) from the moist species.
          float cqw_top = 0.f;
          if (s.n_moist >= 1) {
            const Off ot = X3(d, i, kde - 1, j);
            float qtot = 0.f;
            for (int m = 0; m < s.n_moist; m++) qtot = qtot + s.q[m][ot] + s.q[m][ot - d.ni];
            cqw_top = 0.5f * qtot;
          }
          const float cq1 = 1.f / (1.f + cqw_top);
          const float cq2 = cqw_top * cq1;
          t = t + (1.f / msfty) * K_G *
                      (cq1 * 2.f * s.rdnw[X1(d, k - 1)] * (-s.p[X3(d, i, k - 1, j)]) - (s.c1f[X1(d, k)] * mu) -
                       cq2 * (s.c1f[X1(d, k)] * mub + s.c2f[X1(d, k)]));
        } else {
          const float cqw = s.cqw[o];
          const float cq1 = 1.f / (1.f + cqw);
          const float cq2 = cqw * cq1;
          s.cqw[o] = cq1;
          t = t + (1.f / msfty) * K_G *
                      (cq1 * s.rdn[X1(d, k)] * (s.p[o] - s.p[X3(d, i, k - 1, j)]) - (s.c1f[X1(d, k)] * mu) -
                       cq2 * (s.c1f[X1(d, k)] * mub + s.c2f[X1(d, k)]));
        }
        // w_damp
        if (!adv_only && a.c.w_damping == 1 && k <= kde - 1) {
          const float w_damp_on = a.c.zadvect_implicit > 0 ? a.c.w_crit_cfl : K_W_BETA;
          const float cf = s.c1f[X1(d, k)] * mut + s.c2f[X1(d, k)];
          const float vert_cfl = fabsf(s.ww[o] / cf * s.rdnw[X1(d, k)] * a.dt);
          if (vert_cfl > w_damp_on) t = t - sgn1(w[o]) * K_W_ALPHA * (vert_cfl - a.c.w_crit_cfl) * cf;
        }
        // coriolis and curvature
        if (!adv_only && k <= ktf) {
          const float fm = fzm[X1(d, k)], fp = fzp[X1(d, k)];
          t = t + s.e[X2(d, i, j)] *
                      (s.cosa[X2(d, i, j)] * 0.5f * (fm * (ru[o] + ru[o + 1]) + fp * (ru[o - d.ni] + ru[o - d.ni + 1])) -
                       (s.msftx[X2(d, i, j)] / msfty) * s.sina[X2(d, i, j)] * 0.5f * (fm * (rv[o] + rv[o + d.sj]) + fp * (rv[o - d.ni] + rv[o - d.ni + d.sj])));
          const FPtr<float> uu = s.u_2, vv = s.v_2;
          t = t + K_RERADIUS * (0.5f * (fm * (ru[o] + ru[o + 1]) + fp * (ru[o - d.ni] + ru[o - d.ni + 1])) * 0.5f *
                                    (fm * (uu[o] + uu[o + 1]) + fp * (uu[o - d.ni] + uu[o - d.ni + 1])) +
                                (s.msftx[X2(d, i, j)] / msfty) * 0.5f * (fm * (rv[o] + rv[o + d.sj]) + fp * (rv[o - d.ni] + rv[o - d.ni + d.sj])) *
                                    0.5f * (fm * (vv[o] + vv[o + d.sj]) + fp * (vv[o - d.ni] + vv[o - d.ni + d.sj])));
        }
      }
      // rk_addtend_dry (rw and ph parts share this loop in the reference; ph is in k_tend_t)
      if (!adv_only) {
        float tf = s.rw_tendf[o];
        if (a.rk_step == 1) {
          tf = tf + ((a.c.specified || a.c.nested) ? s.w_save[o] : 0.f) * msfty;  // w_save: zero or relaxation
          s.rw_tendf[o] = tf;
        }
        t = t + tf / msfty;
      }
    }
    s.rw_tend[o] = t;
  }
}

// k_tend_w for the doubly periodic case with four i points per lane, one level per thread;
// element by element the terms of k_tend_w with PER = true, in its order.
template <typename ARGS>
__global__ void __launch_bounds__(kTallBlk) k_tend_w_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  if (!zrow(a, j)) return;  // rows of an open y edge: the zone form
  const int k = d.kds + (int)T.bz;
  const int ids = d.ids, ide = d.ide, jde = d.jde, kde = d.kde, kts = d.kds, kte = kde, ktf = kte - 1;
  if (i0 > ide || j > jde || k > kte) return;
  const float ts = a.c.time_step >= 0 ? 1.f : -1.f;
  const float rdx = s.rdx, rdy = s.rdy;
  const int ni = d.ni, sj = d.sj;
  const FPtr<float> w = s.w_2, ru = s.ru, rv = s.rv, rom = s.ww;
  const CPtr<float> fzm = s.fnm, fzp = s.fnp;
  const Off o2 = X2(d, i0, j), ob = X3(d, i0, 0, j);
  auto lv = [&](int kk) { return ob + imax(kts, imin(kk, kte)) * ni; };
  const Off o = lv(k);
  const bool rk1 = a.rk_step == 1;
  const bool jmass = j <= jde - 1;
  F4 t;
#pragma unroll
  for (int n = 0; n < 4; n++) t.v[n] = 0.f;
  if (!jmass || i0 > ide - 1) {
    st4m(s.rw_tend, o, t, i0, zlo(a, ids), zhi(a, ide));
    return;
  }
  const F4 msfty = s.msfty.ld<F4>(o2), msftx = s.msftx.ld<F4>(o2);
  if (k >= kts + 1) {
    // face velocities at w level k: fzm(k) r(k) + fzp(k) r(k-1) below the lid, the
    // extrapolation (2 - fzm(k-1)) r(k-1) - fzp(k-1) r(k-2) at k = ktf+1
    const bool lid = k > ktf;
    const float ca = lid ? 2.f - fzm[X1(d, k - 1)] : fzm[X1(d, k)];
    const float cb = lid ? -fzp[X1(d, k - 1)] : fzp[X1(d, k)];
    const Off oa = lid ? lv(k - 1) : o, obb = lid ? lv(k - 2) : lv(k - 1);
    // y faces j and j+1
    {
      const F4 ym3 = w.ld<F4>(o - 3 * sj), ym2 = w.ld<F4>(o - 2 * sj), ym1 = w.ld<F4>(o - sj), y0 = w.ld<F4>(o);
      const F4 yp1 = w.ld<F4>(o + sj), yp2 = w.ld<F4>(o + 2 * sj), yp3 = w.ld<F4>(o + 3 * sj);
      const F4 r0a = rv.ld<F4>(oa), r0b = rv.ld<F4>(obb), r1a = rv.ld<F4>(oa + sj), r1b = rv.ld<F4>(obb + sj);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float vel0 = ca * r0a.v[n] + cb * r0b.v[n];
        const float vel1 = ca * r1a.v[n] + cb * r1b.v[n];
        const float f0 = vel0 * flux5(ym3.v[n], ym2.v[n], ym1.v[n], y0.v[n], yp1.v[n], yp2.v[n], vel0, ts);
        const float f1 = vel1 * flux5(ym2.v[n], ym1.v[n], y0.v[n], yp1.v[n], yp2.v[n], yp3.v[n], vel1, ts);
        const float mrdy = msftx.v[n] * rdy;
        t.v[n] = t.v[n] - mrdy * (f1 - f0);
      }
    }
    // x faces i0 .. i0+4
    {
      const F4 wl = w.ld<F4>(o - 4), wc = w.ld<F4>(o), wr = w.ld<F4>(o + 4);
      const F4 rac = ru.ld<F4>(oa), rar = ru.ld<F4>(oa + 4), rbc = ru.ld<F4>(obb), rbr = ru.ld<F4>(obb + 4);
      float W[12];
#pragma unroll
      for (int n = 0; n < 4; n++) {
        W[n] = wl.v[n];
        W[n + 4] = wc.v[n];
        W[n + 8] = wr.v[n];
      }
      float fx[5];
#pragma unroll
      for (int n = 0; n < 5; n++) {
        const float ra = n < 4 ? rac.v[n] : rar.v[0], rb = n < 4 ? rbc.v[n] : rbr.v[0];
        const float vel = ca * ra + cb * rb;
        fx[n] = vel * flux5(W[n + 1], W[n + 2], W[n + 3], W[n + 4], W[n + 5], W[n + 6], vel, ts);
      }
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float mrdx = msftx.v[n] * rdx;
        t.v[n] = t.v[n] - mrdx * (fx[n + 1] - fx[n]);
      }
    }
    // vertical advection of w (order 3) with the lid flux term
    {
      const F4 qm2 = w.ld<F4>(lv(k - 2)), qm1 = w.ld<F4>(lv(k - 1)), q0 = w.ld<F4>(o), qp1 = w.ld<F4>(lv(k + 1)),
               qp2 = w.ld<F4>(lv(k + 2));
      const F4 wm = rom.ld<F4>(lv(k - 1)), w0 = rom.ld<F4>(o), wp = rom.ld<F4>(lv(k + 1));
      // vflux at w level kk in kts+1 .. ktf+1 from rom and w at kk and kk-1
      auto vflux = [&](int kk, float rk, float rkm, float qm2_, float qm1_, float q0_, float qp1_) {
        const float lo = 0.25f * (rk + rkm) * (q0_ + qm1_);
        const float vel = 0.5f * (rk + rkm);
        const float hi = vel * flux3(qm2_, qm1_, q0_, qp1_, -vel, ts);
        return (kk == kts + 1 || kk == ktf + 1) ? lo : hi;
      };
      const float rdnk = s.rdn[X1(d, imin(k, ktf))], rdnm = s.rdn[X1(d, k - 1)];
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float vf_c = vflux(k, w0.v[n], wm.v[n], qm2.v[n], qm1.v[n], q0.v[n], qp1.v[n]);
        if (k <= ktf) {
          const float vf_hi = vflux(k + 1, wp.v[n], w0.v[n], qm1.v[n], q0.v[n], qp1.v[n], qp2.v[n]);
          t.v[n] = t.v[n] - rdnk * (vf_hi - vf_c);
        } else {
          t.v[n] = t.v[n] + 2.f * rdnm * (vf_c);
        }
      }
    }
    // pg_buoy_w
    {
      const F4 mu = s.mu_2.ld<F4>(o2), mub = s.mub.ld<F4>(o2);
      const float c1f = s.c1f[X1(d, k)], c2f = s.c2f[X1(d, k)];
      const F4 pm = s.p.ld<F4>(lv(k - 1));
      if (k == kde) {
        // cqw(kde-1) recomputed from the moist species (k_tend_w reads it before the level
        // kde-1 thread replaces it by 1/(1+cqw))
        F4 qtot;
#pragma unroll
        for (int n = 0; n < 4; n++) qtot.v[n] = 0.f;
        const Off ot = lv(kde - 1);
        for (int m = 0; m < s.n_moist; m++) {
          const F4 qa = s.q[m].ld<F4>(ot), qb = s.q[m].ld<F4>(ot - ni);
#pragma unroll
          for (int n = 0; n < 4; n++) qtot.v[n] = qtot.v[n] + qa.v[n] + qb.v[n];
        }
        const float rdnw = s.rdnw[X1(d, k - 1)];
#pragma unroll
        for (int n = 0; n < 4; n++) {
          const float cqw_top = s.n_moist >= 1 ? 0.5f * qtot.v[n] : 0.f;
          const float cq1 = 1.f / (1.f + cqw_top);
          const float cq2 = cqw_top * cq1;
          t.v[n] = t.v[n] + (1.f / msfty.v[n]) * K_G *
                                (cq1 * 2.f * rdnw * (-pm.v[n]) - (c1f * mu.v[n]) - cq2 * (c1f * mub.v[n] + c2f));
        }
      } else {
        const F4 cqw = s.cqw.ld<F4>(o), pc = s.p.ld<F4>(o);
        const float rdn = s.rdn[X1(d, k)];
        F4 cq1v;
#pragma unroll
        for (int n = 0; n < 4; n++) {
          const float cq1 = 1.f / (1.f + cqw.v[n]);
          const float cq2 = cqw.v[n] * cq1;
          cq1v.v[n] = cq1;
          t.v[n] = t.v[n] + (1.f / msfty.v[n]) * K_G *
                                (cq1 * rdn * (pc.v[n] - pm.v[n]) - (c1f * mu.v[n]) - cq2 * (c1f * mub.v[n] + c2f));
        }
        st4m(s.cqw, o, cq1v, i0, zlo(a, ids), zhi(a, ide - 1));
      }
    }
    // w_damp
    if (a.c.w_damping == 1 && k <= kde - 1) {
      const float w_damp_on = a.c.zadvect_implicit > 0 ? a.c.w_crit_cfl : K_W_BETA;
      const F4 mut = s.mut.ld<F4>(o2), wwc = s.ww.ld<F4>(o), wc = w.ld<F4>(o);
      const float c1f = s.c1f[X1(d, k)], c2f = s.c2f[X1(d, k)], rdnw = s.rdnw[X1(d, k)];
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float cf = c1f * mut.v[n] + c2f;
        const float vert_cfl = fabsf(wwc.v[n] / cf * rdnw * a.dt);
        if (vert_cfl > w_damp_on) t.v[n] = t.v[n] - sgn1(wc.v[n]) * K_W_ALPHA * (vert_cfl - a.c.w_crit_cfl) * cf;
      }
    }
    // coriolis and curvature
    if (k <= ktf) {
      const float fm = fzm[X1(d, k)], fp = fzp[X1(d, k)];
      const Off om = lv(k - 1);
      const F4 e = s.e.ld<F4>(o2), cosa = s.cosa.ld<F4>(o2), sina = s.sina.ld<F4>(o2);
      const F4 r0 = ru.ld<F4>(o), r0r = ru.ld<F4>(o + 4), rm = ru.ld<F4>(om), rmr = ru.ld<F4>(om + 4);
      const F4 v0 = rv.ld<F4>(o), v0n = rv.ld<F4>(o + sj), vm = rv.ld<F4>(om), vmn = rv.ld<F4>(om + sj);
      const F4 u0 = s.u_2.ld<F4>(o), u0r = s.u_2.ld<F4>(o + 4), um = s.u_2.ld<F4>(om), umr = s.u_2.ld<F4>(om + 4);
      const F4 q0 = s.v_2.ld<F4>(o), q0n = s.v_2.ld<F4>(o + sj), qm = s.v_2.ld<F4>(om), qmn = s.v_2.ld<F4>(om + sj);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const float rus = fm * (r0.v[n] + xp1(r0, r0r, n)) + fp * (rm.v[n] + xp1(rm, rmr, n));
        const float rvs = fm * (v0.v[n] + v0n.v[n]) + fp * (vm.v[n] + vmn.v[n]);
        const float mr = msftx.v[n] / msfty.v[n];
        t.v[n] = t.v[n] + e.v[n] * (cosa.v[n] * 0.5f * rus - mr * sina.v[n] * 0.5f * rvs);
        const float uus = fm * (u0.v[n] + xp1(u0, u0r, n)) + fp * (um.v[n] + xp1(um, umr, n));
        const float vvs = fm * (q0.v[n] + q0n.v[n]) + fp * (qm.v[n] + qmn.v[n]);
        t.v[n] = t.v[n] + K_RERADIUS * (0.5f * rus * 0.5f * uus + mr * 0.5f * rvs * 0.5f * vvs);
      }
    }
  }
  // rk_addtend_dry (rw part)
  {
    F4 tf = s.rw_tendf.ld<F4>(o);
#pragma unroll
    for (int n = 0; n < 4; n++) {
      if (rk1) tf.v[n] = tf.v[n] + 0.f * msfty.v[n];  // w_save was zeroed by rk_tendency
      t.v[n] = t.v[n] + tf.v[n] / msfty.v[n];
    }
    if (rk1) st4m(s.rw_tendf, o, tf, i0, zlo(a, ids), zhi(a, ide - 1));
  }
#pragma unroll
  for (int n = 0; n < 4; n++)
    if (i0 + n > ide - 1) t.v[n] = 0.f;
  st4m(s.rw_tend, o, t, i0, zlo(a, ids), zhi(a, ide));
}

// ------------------------------------------------------------------------------ t, ph, zeroing
template <typename ARGS, bool PER, bool ZONE>
__global__ void __launch_bounds__(256) k_tend_t(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  int i, j;
  if (!point_ij<ZONE>(a, T, i, j)) return;
  constexpr int KS = ZONE ? 1 : kTendK;  // levels per thread (the zone form: one)
  const int k0 = d.kds + (int)T.bz * KS;
  const int ids = d.ids, ide = d.ide, jds = d.jds, jde = d.jde, kde = d.kde;
  const int its = ids, ite = ide, jts = jds, jte = jde, kts = d.kds, kte = kde;
  if (i > ide || j > jde || k0 > kte) return;
  const int k1 = imin(k0 + KS - 1, kte);
  const Geo g = geo<PER>(a.c);
  const float ts = a.c.time_step >= 0 ? 1.f : -1.f;
  const int ktf = kte - 1;
  const FPtr<float> f = s.t_2, ru = s.ru, rv = s.rv, rom = s.ww;
  const float rdx = s.rdx, rdy = s.rdy;
  const bool mass = i <= imin(ite, ide - 1) && j <= imin(jte, jde - 1);
  const bool dxs = !(g.per_x || g.sym_xs || its > ids + 3), dxe = !(g.per_x || g.sym_xe || ite < ide - 3);
  const bool dys = !(g.per_y || g.sym_ys || jts > jds + 3), dye = !(g.per_y || g.sym_ye || jte < jde - 4);
  int jy0 = jts, jy1 = imin(jte, jde - 1), jf0 = jy0, jf1 = jy1 + 1;
  if (dys) { jy0 = imax(jts, jds + 1); jf0 = jds + 3; }
  if (dye) { jy1 = imin(jte, jde - 2); jf1 = jde - 3; }
  int ix0 = its, ix1 = imin(ite, ide - 1), if0 = ix0, if1 = ix1 + 1;
  if (dxs) { ix0 = imax(ids + 1, its); if0 = imin(ix0 + 2, ids + 3); }
  if (dxe) { ix1 = imin(ide - 2, ite); if1 = ide - 3; }
  auto fqy = [&](int k, int jj) -> float {
    if (jj >= jf0 && jj <= jf1) {
      const float vel = rv[X3(d, i, k, jj)];
      return vel * flux5(f[X3(d, i, k, jj - 3)], f[X3(d, i, k, jj - 2)], f[X3(d, i, k, jj - 1)], f[X3(d, i, k, jj)],
                         f[X3(d, i, k, jj + 1)], f[X3(d, i, k, jj + 2)], vel, ts);
    } else if (jj == jds + 1 || jj == jde - 1) {
      return 0.5f * rv[X3(d, i, k, jj)] * (f[X3(d, i, k, jj)] + f[X3(d, i, k, jj - 1)]);
    } else {
      const float vel = rv[X3(d, i, k, jj)];
      return vel * flux3(f[X3(d, i, k, jj - 2)], f[X3(d, i, k, jj - 1)], f[X3(d, i, k, jj)], f[X3(d, i, k, jj + 1)], vel, ts);
    }
  };
  auto fqx = [&](int k, int ii) -> float {
    if (ii >= if0 && ii <= if1) {
      const float vel = ru[X3(d, ii, k, j)];
      return vel * flux5(f[X3(d, ii - 3, k, j)], f[X3(d, ii - 2, k, j)], f[X3(d, ii - 1, k, j)], f[X3(d, ii, k, j)],
                         f[X3(d, ii + 1, k, j)], f[X3(d, ii + 2, k, j)], vel, ts);
    } else if (ii == ids + 1 || ii == ide - 1) {
      return 0.5f * (ru[X3(d, ii, k, j)]) * (f[X3(d, ii, k, j)] + f[X3(d, ii - 1, k, j)]);
    } else {
      const float vel = ru[X3(d, ii, k, j)];
      return vel * flux3(f[X3(d, ii - 2, k, j)], f[X3(d, ii - 1, k, j)], f[X3(d, ii, k, j)], f[X3(d, ii + 1, k, j)], vel, ts);
    }
  };
  auto vflux = [&](int k) -> float {
    if (k == kts || k == kte) return 0.f;
    if (k == kts + 1 || k == ktf)
      return rom[X3(d, i, k, j)] * (s.fnm[X1(d, k)] * f[X3(d, i, k, j)] + s.fnp[X1(d, k)] * f[X3(d, i, k - 1, j)]);
    const float vel = rom[X3(d, i, k, j)];
    return vel * flux3(f[X3(d, i, k - 2, j)], f[X3(d, i, k - 1, j)], f[X3(d, i, k, j)], f[X3(d, i, k + 1, j)], -vel, ts);
  };
  // rhs_ph pieces
  const FPtr<float> u = s.u_2, v = s.v_2, ph = s.ph_2, phb = s.phb, w = s.w_2;
  const float cfn = s.cfn, cfn1 = s.cfn1;
  const bool spec = g.spec;
  const bool ord6 = a.c.h_sca_adv_order <= 6 && a.c.h_sca_adv_order > 4;
  int jp0 = jts, jp1 = imin(jte, jde - 1);
  if (g.open_ys || spec) jp0 = imax(jts, jds + 3);
  if (g.open_ye || spec) jp1 = imin(jp1, jde - 4);
  int ip0 = its, ip1 = imin(ite, ide - 1);
  if (g.open_xs || spec) ip0 = imax(its, ids + 3);
  if (g.open_xe || spec) ip1 = imin(ip1, ide - 4);
  auto vface = [&](int k, int jj) -> float {
    const float cc = s.c1f[X1(d, k)] * s.muv[X2(d, i, jj)] + s.c2f[X1(d, k)];
    if (k < kte) return cc * (v[X3(d, i, k, jj)] + v[X3(d, i, k - 1, jj)]) * s.msfvy[X2(d, i, jj)];
    return cc * (cfn * v[X3(d, i, k - 1, jj)] + cfn1 * v[X3(d, i, k - 2, jj)]) * s.msfvy[X2(d, i, jj)];
  };
  auto uface = [&](int k, int ii) -> float {
    const float cc = s.c1f[X1(d, k)] * s.muu[X2(d, ii, j)] + s.c2f[X1(d, k)];
    if (k < kte) return cc * (u[X3(d, ii, k, j)] + u[X3(d, ii, k - 1, j)]) * s.msfux[X2(d, ii, j)];
    return cc * (cfn * u[X3(d, ii, k - 1, j)] + cfn1 * u[X3(d, ii, k - 2, j)]) * s.msfux[X2(d, ii, j)];
  };
  auto P = [&](FPtr<float> arr, int ii, int k, int jj) -> float { return arr[X3(d, ii, k, jj)]; };
  const float msfty = mass ? s.msfty[X2(d, i, j)] : 1.f;
  const float mut = mass ? s.mut[X2(d, i, j)] : 0.f;
  // rhs_ph's wdwn at w level kk (kk = 2..kde-1)
  auto wdwn = [&](int kk) -> float {
    if (a.c.phi_adv_z == 2)
      return s.rdnw[X1(d, kk - 1)] * (P(ph, i, kk, j) - P(ph, i, kk - 1, j) + P(phb, i, kk, j) - P(phb, i, kk - 1, j));
    return .5f * (P(rom, i, kk, j) + P(rom, i, kk - 1, j)) * s.rdnw[X1(d, kk - 1)] *
           (P(ph, i, kk, j) - P(ph, i, kk - 1, j) + P(phb, i, kk, j) - P(phb, i, kk - 1, j));
  };
  float vf_c = mass ? vflux(k0) : 0.f;                                  // vflux at level k
  float wd_c = (mass && k0 >= 2 && k0 <= kte - 1) ? wdwn(k0) : 0.f;     // rhs_ph's wdwn at level k
  for (int k = k0; k <= k1; k++) {
    const Off o = X3(d, i, k, j);
    // relax_bdy_dry's relaxation of theta and phi (specified boundaries, rk_step 1), which
    // rk_addtend_dry adds to the forward tendencies; zero otherwise
    const bool relax = (a.c.specified || a.c.nested) && a.rk_step == 1;
    const float t_sv = relax ? s.t_save[o] : 0.f, ph_sv = relax ? s.ph_save[o] : 0.f;
    // zero_tend of the *_save arrays over the whole tile (with the relaxation they hold its
    // values, zero outside the relaxation rows, as they are after rk_addtend_dry)
    if (!relax) {
      s.u_save[o] = 0.f;
      s.v_save[o] = 0.f;
      s.w_save[o] = 0.f;
      s.ph_save[o] = 0.f;
      s.t_save[o] = 0.f;
    }
    float t = 0.f, pt = 0.f, vf_hi = 0.f;
    if (mass) {
      if (k <= ktf) {
        // advect_scalar(t)
        if (j >= jy0 && j <= jy1) {
          const float mrdy = s.msftx[X2(d, i, j)] * rdy;
          t = t - mrdy * (fqy(k, j + 1) - fqy(k, j));
        }
        if (i >= ix0 && i <= ix1) {
          const float mrdx = s.msftx[X2(d, i, j)] * rdx;
          t = t - mrdx * (fqx(k, i + 1) - fqx(k, i));
        }
        if (g.open_xs && i == its) {
          float ub = 0.5f * (P(ru, its, k, j) + P(ru, its + 1, k, j));
          if (ub > 0.f) ub = 0.f;
          t = t - rdx * (ub * (P(f, its + 1, k, j) - P(f, its, k, j)) + P(f, its, k, j) * (P(ru, its + 1, k, j) - P(ru, its, k, j)));
        }
        if (g.open_xe && i == imin(ite, ide - 1)) {
          float ub = 0.5f * (P(ru, ite - 1, k, j) + P(ru, ite, k, j));
          if (ub < 0.f) ub = 0.f;
          t = t - rdx * (ub * (P(f, i, k, j) - P(f, i - 1, k, j)) + P(f, i, k, j) * (P(ru, ite, k, j) - P(ru, ite - 1, k, j)));
        }
        if (g.open_ys && j == jts) {
          float vb = 0.5f * (P(rv, i, k, jts) + P(rv, i, k, jts + 1));
          if (vb > 0.f) vb = 0.f;
          t = t - rdy * (vb * (P(f, i, k, jts + 1) - P(f, i, k, jts)) + P(f, i, k, jts) * (P(rv, i, k, jts + 1) - P(rv, i, k, jts)));
        }
        if (g.open_ye && j == imin(jte, jde - 1)) {
          float vb = 0.5f * (P(rv, i, k, jte - 1) + P(rv, i, k, jte));
          if (vb < 0.f) vb = 0.f;
          t = t - rdy * (vb * (P(f, i, k, j) - P(f, i, k, j - 1)) + P(f, i, k, j) * (P(rv, i, k, jte) - P(rv, i, k, jte - 1)));
        }
        vf_hi = vflux(k + 1);
        t = t - s.rdnw[X1(d, k)] * (vf_hi - vf_c);
        vf_c = vf_hi;
        // rk_addtend_dry (theta)
        float tf = s.t_tendf[o];
        if (a.rk_step == 1) {
          tf = tf + t_sv;  // t_save: zeroed by rk_tendency, or the relaxation
          s.t_tendf[o] = tf;
        }
        t = t + tf / msfty + (s.c1h[X1(d, k)] * mut + s.c2h[X1(d, k)]) * s.h_diabatic[o] / msfty;
      }
      // rhs_ph
      if (k >= 2 && k <= kte - 1) {
        const float wd_hi = wdwn(k + 1), wd_lo = k == 2 || k == k0 ? wdwn(k) : wd_c;
        wd_c = wd_hi;
        if (a.c.phi_adv_z == 2)
          pt = pt - P(rom, i, k, j) * (s.fnm[X1(d, k)] * wd_hi + s.fnp[X1(d, k)] * wd_lo);
        else
          pt = pt - (s.fnm[X1(d, k)] * wd_hi + s.fnp[X1(d, k)] * wd_lo);
      }
      if (a.c.non_hydrostatic) {
        if (k == kde) pt = 0.f;
        if (k >= 2) pt = pt + (s.c1f[X1(d, k)] * mut + s.c2f[X1(d, k)]) * K_G * w[o] / msfty;
      }
      if (ord6 && k >= 2) {
        const float cy = k < kte ? (0.25f * rdy / msfty) : (0.5f * rdy / msfty);
        const float cx = k < kte ? (0.25f * rdx / msfty) : (0.5f * rdx / msfty);
        if (j >= jp0 && j <= jp1) {
          const float S = 45.f * (P(ph, i, k, j + 1) - P(ph, i, k, j - 1)) - 9.f * (P(ph, i, k, j + 2) - P(ph, i, k, j - 2)) +
                          (P(ph, i, k, j + 3) - P(ph, i, k, j - 3)) + 45.f * (P(phb, i, k, j + 1) - P(phb, i, k, j - 1)) -
                          9.f * (P(phb, i, k, j + 2) - P(phb, i, k, j - 2)) + (P(phb, i, k, j + 3) - P(phb, i, k, j - 3));
          pt = pt - cy * ((vface(k, j + 1) + vface(k, j)) * (1.f / 60.f) * S);
        }
        if (((g.open_ys || spec) && j == jds + 2) || ((g.open_ye || spec) && j == jde - 3)) {
          const float S = 8.f * (P(ph, i, k, j + 1) - P(ph, i, k, j - 1)) - (P(ph, i, k, j + 2) - P(ph, i, k, j - 2)) +
                          8.f * (P(phb, i, k, j + 1) - P(phb, i, k, j - 1)) - (P(phb, i, k, j + 2) - P(phb, i, k, j - 2));
          pt = pt - cy * ((vface(k, j + 1) + vface(k, j)) * (1.f / 12.f) * S);
        }
        if (((g.open_ys || spec) && j == jds + 1) || ((g.open_ye || spec) && j == jde - 2)) {
          const float dA = P(phb, i, k, j + 1) - P(phb, i, k, j) + P(ph, i, k, j + 1) - P(ph, i, k, j);
          const float dB = P(phb, i, k, j) - P(phb, i, k, j - 1) + P(ph, i, k, j) - P(ph, i, k, j - 1);
          pt = pt - cy * (vface(k, j + 1) * dA + vface(k, j) * dB);
        }
        if (i >= ip0 && i <= ip1) {
          const float S = 45.f * (P(ph, i + 1, k, j) - P(ph, i - 1, k, j)) - 9.f * (P(ph, i + 2, k, j) - P(ph, i - 2, k, j)) +
                          (P(ph, i + 3, k, j) - P(ph, i - 3, k, j)) + 45.f * (P(phb, i + 1, k, j) - P(phb, i - 1, k, j)) -
                          9.f * (P(phb, i + 2, k, j) - P(phb, i - 2, k, j)) + (P(phb, i + 3, k, j) - P(phb, i - 3, k, j));
          pt = pt - cx * ((uface(k, i + 1) + uface(k, i)) * (1.f / 60.f) * S);
        }
        if ((g.open_xs && i == ids + 2) || (g.open_xe && i == ide - 3)) {
          const float S = 8.f * (P(ph, i + 1, k, j) - P(ph, i - 1, k, j)) - (P(ph, i + 2, k, j) - P(ph, i - 2, k, j)) +
                          8.f * (P(phb, i + 1, k, j) - P(phb, i - 1, k, j)) - (P(phb, i + 2, k, j) - P(phb, i - 2, k, j));
          pt = pt - cx * ((uface(k, i + 1) + uface(k, i)) * (1.f / 12.f) * S);
        }
        if (((g.open_xs || spec) && i == ids + 1) || ((g.open_xe || spec) && i == ide - 2)) {
          const float dA = P(phb, i + 1, k, j) - P(phb, i, k, j) + P(ph, i + 1, k, j) - P(ph, i, k, j);
          const float dB = P(phb, i, k, j) - P(phb, i - 1, k, j) + P(ph, i, k, j) - P(ph, i - 1, k, j);
          pt = pt - cx * (uface(k, i + 1) * dA + uface(k, i) * dB);
        }
      }
      if (k >= 2) {
        const float cc = s.c1f[X1(d, k)] * mut + s.c2f[X1(d, k)];
        if (g.open_ys && j == jts) {
          const int kz = imin(k, kde - 1);
          const float vb = .5f * (s.fnm[X1(d, kz)] * (P(v, i, kz, j + 1) + P(v, i, kz, j)) + s.fnp[X1(d, kz)] * (P(v, i, kz - 1, j + 1) + P(v, i, kz - 1, j)));
          const float vl = vb < 0.f ? vb : 0.f;
          pt = pt - rdy * cc * (vl * (P(ph, i, k, j + 1) - P(ph, i, k, j)));
        }
        if (g.open_ye && j == jte - 1) {
          const int kz = imin(k, kde - 1);
          const float vb = .5f * (s.fnm[X1(d, kz)] * (P(v, i, kz, j + 1) + P(v, i, kz, j)) + s.fnp[X1(d, kz)] * (P(v, i, kz - 1, j + 1) + P(v, i, kz - 1, j)));
          const float vr = vb > 0.f ? vb : 0.f;
          pt = pt - rdy * cc * (vr * (P(ph, i, k, j) - P(ph, i, k, j - 1)));
        }
        if (g.open_xs && i == its) {
          const int kz = k;
          const float ub = .5f * (s.fnm[X1(d, kz)] * (P(u, i + 1, kz, j) + P(u, i, kz, j)) + s.fnp[X1(d, kz)] * (P(u, i + 1, kz - 1, j) + P(u, i, kz - 1, j)));
          const float ul = ub < 0.f ? ub : 0.f;
          pt = pt - (s.msftx[X2(d, i, j)] / msfty) * rdx * cc * (ul * (P(ph, i + 1, k, j) - P(ph, i, k, j)));
        }
        if (g.open_xe && i == ite - 1) {
          const int kz = k < kde ? k : k - 1;
          const float ub = .5f * (s.fnm[X1(d, kz)] * (P(u, i + 1, kz, j) + P(u, i, kz, j)) + s.fnp[X1(d, kz)] * (P(u, i + 1, kz - 1, j) + P(u, i, kz - 1, j)));
          const float ur = ub > 0.f ? ub : 0.f;
          pt = pt - (s.msftx[X2(d, i, j)] / msfty) * rdx * cc * (ur * (P(ph, i, k, j) - P(ph, i - 1, k, j)));
        }
      }
      float ptf = s.ph_tendf[o];
      if (a.rk_step == 1) {
        ptf = ptf + ph_sv;  // ph_save: zeroed by rk_tendency, or the relaxation
        s.ph_tendf[o] = ptf;
      }
      pt = pt + ptf / msfty;
    }
    s.t_tend[o] = t;
    s.ph_tend[o] = pt;
  }
  if (k0 == kts) {
    // mu_tend: zeroed by rk_tendency, plus relax_bdy_dry's relaxation of mu at rk_step 1
    // (k_lbc leaves it in mu_save), then rk_addtend_dry's mu_tendf
    const float mu_rx = ((a.c.specified || a.c.nested) && a.rk_step == 1) ? s.mu_save[X2(d, i, j)] : 0.f;
    s.mu_tend[X2(d, i, j)] = 0.f;
    s.mu_save[X2(d, i, j)] = 0.f;
    if (mass) s.mu_tend[X2(d, i, j)] = mu_rx + s.mu_tendf[X2(d, i, j)];
  }
}

// k_tend_t for the doubly periodic case with four i points per lane, one level per thread:
// zero_tend of the *_save arrays, advect_scalar(theta) + rk_addtend_dry into t_tend, rhs_ph
// (order-6 horizontal form) + rk_addtend_dry into ph_tend, and mu_tend/mu_save at level
// kts.  Element by element the terms of k_tend_t with PER = true, in its order.
template <typename ARGS>
__global__ void __launch_bounds__(kTallBlk) k_tend_t_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  if (!zrow(a, j)) return;  // rows of an open y edge: the zone form
  const int k = d.kds + (int)T.bz;
  const int ids = d.ids, ide = d.ide, jde = d.jde, kde = d.kde, kts = d.kds, kte = kde, ktf = kte - 1;
  if (i0 > ide || j > jde || k > kte) return;
  const float ts = a.c.time_step >= 0 ? 1.f : -1.f;
  const float rdx = s.rdx, rdy = s.rdy;
  const int ni = d.ni, sj = d.sj;
  const FPtr<float> f = s.t_2, ru = s.ru, rv = s.rv, rom = s.ww, ph = s.ph_2, phb = s.phb;
  const Off o2 = X2(d, i0, j), ob = X3(d, i0, 0, j);
  auto lv = [&](int kk) { return ob + imax(kts, imin(kk, kte)) * ni; };
  const Off o = lv(k);
  const bool rk1 = a.rk_step == 1;
  F4 z;
#pragma unroll
  for (int n = 0; n < 4; n++) z.v[n] = 0.f;
  // zero_tend of the *_save arrays over the whole tile
  st4m(s.u_save, o, z, i0, zlo(a, ids), zhi(a, ide));
  st4m(s.v_save, o, z, i0, zlo(a, ids), zhi(a, ide));
  st4m(s.w_save, o, z, i0, zlo(a, ids), zhi(a, ide));
  st4m(s.ph_save, o, z, i0, zlo(a, ids), zhi(a, ide));
  st4m(s.t_save, o, z, i0, zlo(a, ids), zhi(a, ide));
  if (k == kts) {
    st4m(s.mu_tend, o2, z, i0, zlo(a, ids), zhi(a, ide));
    st4m(s.mu_save, o2, z, i0, zlo(a, ids), zhi(a, ide));
  }
  F4 t = z, pt = z;
  const bool jmass = j <= jde - 1;
  if (jmass && i0 <= ide - 1) {
    const F4 msftx = s.msftx.ld<F4>(o2), msfty = s.msfty.ld<F4>(o2), mut = s.mut.ld<F4>(o2);
    if (k <= ktf) {
      // advect_scalar(t): y faces j and j+1
      {
        const F4 ym3 = f.ld<F4>(o - 3 * sj), ym2 = f.ld<F4>(o - 2 * sj), ym1 = f.ld<F4>(o - sj), y0 = f.ld<F4>(o);
        const F4 yp1 = f.ld<F4>(o + sj), yp2 = f.ld<F4>(o + 2 * sj), yp3 = f.ld<F4>(o + 3 * sj);
        const F4 rv0 = rv.ld<F4>(o), rv1 = rv.ld<F4>(o + sj);
#pragma unroll
        for (int n = 0; n < 4; n++) {
          const float f0 = rv0.v[n] * flux5(ym3.v[n], ym2.v[n], ym1.v[n], y0.v[n], yp1.v[n], yp2.v[n], rv0.v[n], ts);
          const float f1 = rv1.v[n] * flux5(ym2.v[n], ym1.v[n], y0.v[n], yp1.v[n], yp2.v[n], yp3.v[n], rv1.v[n], ts);
          const float mrdy = msftx.v[n] * rdy;
          t.v[n] = t.v[n] - mrdy * (f1 - f0);
        }
      }
      // x faces i0 .. i0+4
      {
        const F4 wl = f.ld<F4>(o - 4), wc = f.ld<F4>(o), wr = f.ld<F4>(o + 4);
        const F4 rc = ru.ld<F4>(o), rr = ru.ld<F4>(o + 4);
        float W[12];
#pragma unroll
        for (int n = 0; n < 4; n++) {
          W[n] = wl.v[n];
          W[n + 4] = wc.v[n];
          W[n + 8] = wr.v[n];
        }
        float fx[5];
#pragma unroll
        for (int n = 0; n < 5; n++) {
          const float vel = n < 4 ? rc.v[n] : rr.v[0];
          fx[n] = vel * flux5(W[n + 1], W[n + 2], W[n + 3], W[n + 4], W[n + 5], W[n + 6], vel, ts);
        }
#pragma unroll
        for (int n = 0; n < 4; n++) {
          const float mrdx = msftx.v[n] * rdx;
          t.v[n] = t.v[n] - mrdx * (fx[n + 1] - fx[n]);
        }
      }
      // vertical
      {
        const F4 qm2 = f.ld<F4>(lv(k - 2)), qm1 = f.ld<F4>(lv(k - 1)), q0 = f.ld<F4>(o), qp1 = f.ld<F4>(lv(k + 1)),
                 qp2 = f.ld<F4>(lv(k + 2));
        const F4 w0 = rom.ld<F4>(o), w1 = rom.ld<F4>(lv(k + 1));
        auto vflux = [&](int kk, float w_, float qm2_, float qm1_, float q0_, float qp1_) {
          const int kc = imin(kk, kte);
          const float lo = w_ * (s.fnm[X1(d, kc)] * q0_ + s.fnp[X1(d, kc)] * qm1_);
          const float hi = w_ * flux3(qm2_, qm1_, q0_, qp1_, -w_, ts);
          return (kk == kts || kk == kte) ? 0.f : ((kk == kts + 1 || kk == ktf) ? lo : hi);
        };
        const float rdnw = s.rdnw[X1(d, k)];
#pragma unroll
        for (int n = 0; n < 4; n++) {
          const float vf_c = vflux(k, w0.v[n], qm2.v[n], qm1.v[n], q0.v[n], qp1.v[n]);
          const float vf_hi = vflux(k + 1, w1.v[n], qm1.v[n], q0.v[n], qp1.v[n], qp2.v[n]);
          t.v[n] = t.v[n] - rdnw * (vf_hi - vf_c);
        }
      }
      // rk_addtend_dry (theta)
      {
        F4 tf = s.t_tendf.ld<F4>(o);
        const F4 hd = s.h_diabatic.ld<F4>(o);
        const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
#pragma unroll
        for (int n = 0; n < 4; n++) {
          if (rk1) tf.v[n] = tf.v[n] + 0.f;  // t_save was zeroed by rk_tendency
          t.v[n] = t.v[n] + tf.v[n] / msfty.v[n] + (c1 * mut.v[n] + c2) * hd.v[n] / msfty.v[n];
        }
        if (rk1) st4m(s.t_tendf, o, tf, i0, zlo(a, ids), zhi(a, ide - 1));
      }
    }
    // rhs_ph
    const bool z2 = a.c.phi_adv_z == 2;
    if (k >= 2 && k <= kte - 1) {
      const F4 hm = ph.ld<F4>(lv(k - 1)), h0 = ph.ld<F4>(o), hp = ph.ld<F4>(lv(k + 1));
      const F4 bm = phb.ld<F4>(lv(k - 1)), b0 = phb.ld<F4>(o), bp = phb.ld<F4>(lv(k + 1));
      const F4 rm = rom.ld<F4>(lv(k - 1)), r0 = rom.ld<F4>(o), rp = rom.ld<F4>(lv(k + 1));
      const float rdnw_k = s.rdnw[X1(d, k)], rdnw_m = s.rdnw[X1(d, k - 1)];
      const float fm = s.fnm[X1(d, k)], fp = s.fnp[X1(d, k)];
#pragma unroll
      for (int n = 0; n < 4; n++) {
        // wdwn at w levels k+1 and k
        const float dhi = hp.v[n] - h0.v[n] + bp.v[n] - b0.v[n], dlo = h0.v[n] - hm.v[n] + b0.v[n] - bm.v[n];
        const float wd_hi = z2 ? rdnw_k * dhi : .5f * (rp.v[n] + r0.v[n]) * rdnw_k * dhi;
        const float wd_lo = z2 ? rdnw_m * dlo : .5f * (r0.v[n] + rm.v[n]) * rdnw_m * dlo;
        if (z2) pt.v[n] = pt.v[n] - r0.v[n] * (fm * wd_hi + fp * wd_lo);
        else pt.v[n] = pt.v[n] - (fm * wd_hi + fp * wd_lo);
      }
    }
    if (a.c.non_hydrostatic) {
      const F4 w = s.w_2.ld<F4>(o);
      const float c1f = s.c1f[X1(d, k)], c2f = s.c2f[X1(d, k)];
#pragma unroll
      for (int n = 0; n < 4; n++) {
        if (k == kde) pt.v[n] = 0.f;
        if (k >= 2) pt.v[n] = pt.v[n] + (c1f * mut.v[n] + c2f) * K_G * w.v[n] / msfty.v[n];
      }
    }
    const bool ord6 = a.c.h_sca_adv_order <= 6 && a.c.h_sca_adv_order > 4;
    if (ord6 && k >= 2) {
      const float c1f = s.c1f[X1(d, k)], c2f = s.c2f[X1(d, k)], cfn = s.cfn, cfn1 = s.cfn1;
      const bool below = k < kte;
      // v and u at the faces: levels k and k-1, or k-1 and k-2 at the lid
      const Off fa = below ? o : lv(k - 1), fb = below ? lv(k - 1) : lv(k - 2);
      const float ca = below ? 1.f : cfn, cb = below ? 1.f : cfn1;
      // y: rows j-3 .. j+3 of ph and phb
      {
        const F4 p1 = ph.ld<F4>(o + sj), m1 = ph.ld<F4>(o - sj), p2 = ph.ld<F4>(o + 2 * sj), m2 = ph.ld<F4>(o - 2 * sj);
        const F4 p3 = ph.ld<F4>(o + 3 * sj), m3 = ph.ld<F4>(o - 3 * sj);
        const F4 bp1 = phb.ld<F4>(o + sj), bm1 = phb.ld<F4>(o - sj), bp2 = phb.ld<F4>(o + 2 * sj), bm2 = phb.ld<F4>(o - 2 * sj);
        const F4 bp3 = phb.ld<F4>(o + 3 * sj), bm3 = phb.ld<F4>(o - 3 * sj);
        const F4 va0 = s.v_2.ld<F4>(fa), vb0 = s.v_2.ld<F4>(fb), va1 = s.v_2.ld<F4>(fa + sj), vb1 = s.v_2.ld<F4>(fb + sj);
        const F4 muv0 = s.muv.ld<F4>(o2), muv1 = s.muv.ld<F4>(o2 + ni), my0 = s.msfvy.ld<F4>(o2), my1 = s.msfvy.ld<F4>(o2 + ni);
#pragma unroll
        for (int n = 0; n < 4; n++) {
          const float cy = below ? (0.25f * rdy / msfty.v[n]) : (0.5f * rdy / msfty.v[n]);
          const float S = 45.f * (p1.v[n] - m1.v[n]) - 9.f * (p2.v[n] - m2.v[n]) + (p3.v[n] - m3.v[n]) +
                          45.f * (bp1.v[n] - bm1.v[n]) - 9.f * (bp2.v[n] - bm2.v[n]) + (bp3.v[n] - bm3.v[n]);
          const float cc1 = c1f * muv1.v[n] + c2f, cc0 = c1f * muv0.v[n] + c2f;
          const float vf1 = below ? cc1 * (va1.v[n] + vb1.v[n]) * my1.v[n] : cc1 * (ca * va1.v[n] + cb * vb1.v[n]) * my1.v[n];
          const float vf0 = below ? cc0 * (va0.v[n] + vb0.v[n]) * my0.v[n] : cc0 * (ca * va0.v[n] + cb * vb0.v[n]) * my0.v[n];
          pt.v[n] = pt.v[n] - cy * ((vf1 + vf0) * (1.f / 60.f) * S);
        }
      }
      // x: points i-3 .. i+3 from the groups l, c, r
      {
        const F4 hl = ph.ld<F4>(o - 4), hc = ph.ld<F4>(o), hr = ph.ld<F4>(o + 4);
        const F4 bl = phb.ld<F4>(o - 4), bc = phb.ld<F4>(o), br = phb.ld<F4>(o + 4);
        const F4 ua = s.u_2.ld<F4>(fa), uar = s.u_2.ld<F4>(fa + 4), ub = s.u_2.ld<F4>(fb), ubr = s.u_2.ld<F4>(fb + 4);
        const F4 muu = s.muu.ld<F4>(o2), muur = s.muu.ld<F4>(o2 + 4), mx = s.msfux.ld<F4>(o2), mxr = s.msfux.ld<F4>(o2 + 4);
        float H[12], B[12];
#pragma unroll
        for (int n = 0; n < 4; n++) {
          H[n] = hl.v[n]; H[n + 4] = hc.v[n]; H[n + 8] = hr.v[n];
          B[n] = bl.v[n]; B[n + 4] = bc.v[n]; B[n + 8] = br.v[n];
        }
#pragma unroll
        for (int n = 0; n < 4; n++) {
          const int c = n + 4;  // index of point i in H/B
          const float cx = below ? (0.25f * rdx / msfty.v[n]) : (0.5f * rdx / msfty.v[n]);
          const float S = 45.f * (H[c + 1] - H[c - 1]) - 9.f * (H[c + 2] - H[c - 2]) + (H[c + 3] - H[c - 3]) +
                          45.f * (B[c + 1] - B[c - 1]) - 9.f * (B[c + 2] - B[c - 2]) + (B[c + 3] - B[c - 3]);
          const float cc1 = c1f * xp1(muu, muur, n) + c2f, cc0 = c1f * muu.v[n] + c2f;
          const float uf1 = below ? cc1 * (xp1(ua, uar, n) + xp1(ub, ubr, n)) * xp1(mx, mxr, n)
                                  : cc1 * (ca * xp1(ua, uar, n) + cb * xp1(ub, ubr, n)) * xp1(mx, mxr, n);
          const float uf0 = below ? cc0 * (ua.v[n] + ub.v[n]) * mx.v[n] : cc0 * (ca * ua.v[n] + cb * ub.v[n]) * mx.v[n];
          pt.v[n] = pt.v[n] - cx * ((uf1 + uf0) * (1.f / 60.f) * S);
        }
      }
    }
    // rk_addtend_dry (ph)
    {
      F4 ptf = s.ph_tendf.ld<F4>(o);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        if (rk1) ptf.v[n] = ptf.v[n] + 0.f;  // ph_save was zeroed by rk_tendency
        pt.v[n] = pt.v[n] + ptf.v[n] / msfty.v[n];
      }
      if (rk1) st4m(s.ph_tendf, o, ptf, i0, zlo(a, ids), zhi(a, ide - 1));
    }
    if (k == kts) {
      const F4 mtf = s.mu_tendf.ld<F4>(o2);
      F4 mt;
#pragma unroll
      for (int n = 0; n < 4; n++) mt.v[n] = 0.f + mtf.v[n];
      st4m(s.mu_tend, o2, mt, i0, zlo(a, ids), zhi(a, ide - 1));
    }
  }
#pragma unroll
  for (int n = 0; n < 4; n++)
    if (i0 + n > ide - 1 || !jmass) {
      t.v[n] = 0.f;
      pt.v[n] = 0.f;
    }
  st4m(s.t_tend, o, t, i0, zlo(a, ids), zhi(a, ide));
  st4m(s.ph_tend, o, pt, i0, zlo(a, ids), zhi(a, ide));
}

WRFMI_INST_P2(k_tend_u, false, false)
WRFMI_INST_P2(k_tend_u, true, false)
WRFMI_INST_P2(k_tend_u, false, true)
WRFMI_INST(k_tend_u_x4)
WRFMI_INST_P2(k_tend_v, false, false)
WRFMI_INST_P2(k_tend_v, true, false)
WRFMI_INST_P2(k_tend_v, false, true)
WRFMI_INST(k_tend_v_x4)
WRFMI_INST_P2(k_tend_w, false, false)
WRFMI_INST_P2(k_tend_w, true, false)
WRFMI_INST_P2(k_tend_w, false, true)
WRFMI_INST(k_tend_w_x4)
WRFMI_INST_P2(k_tend_t, false, false)
WRFMI_INST_P2(k_tend_t, true, false)
WRFMI_INST_P2(k_tend_t, false, true)
WRFMI_INST(k_tend_t_x4)

}  // namespace wrfmi
