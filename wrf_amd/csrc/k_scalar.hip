// Scalar transport, the equation of state and the lateral halo fills on the GPU.
//   k_adv_scalar    : rk_scalar_tend's advect_scalar (dyn_em/module_em.F:1096-1442,
//                     module_advect_em.F:3029-4359) with the time-averaged ru_m/rv_m/ww_m
//   k_update_scalar : rk_update_scalar (dyn_em/module_em.F:1587-1799)
//   k_p_rho_phi     : calc_p_rho_phi, non-hydrostatic (module_big_step_utilities_em.F:953-1223)
//   k_halo          : set_physical_bc3d / set_physical_bc2d (share/module_bc.F:202-1113)
//                     for up to 16 fields per launch
#include "kernels.h"

namespace wrfmi {

// advect_scalar with each face flux computed once.  A block is a tile of 63 x 4 columns
// (wave lane = i, wave = j row) marching up kAdvK levels: lane l computes the x flux at its
// own face i and takes the face i+1 flux from lane l+1 (lane 63 only computes); each row
// computes the y flux at its face j and reads face j+1 from the row above through LDS (the
// top row also computes face j+4); the vertical flux is carried up the column.  Every flux
// is the same expression, in the same order, as in advect_scalar.
template <typename ARGS, bool PER, bool ZONE>
__global__ void __launch_bounds__(256) k_adv_scalar(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int lane = (int)threadIdx.x, row = (int)threadIdx.y;
  int i = d.ids + T.bx * kAdvTileI + lane;
  int j = d.jds + T.by * (int)blockDim.y + row;
  // the zone form: one point per thread, which forms both faces of its divergence itself
  if (ZONE && !point_ij<true>(a, T, i, j)) return;
  constexpr int KS = ZONE ? 1 : kAdvK;  // levels per thread (the zone form: one)
  const int k0 = d.kds + T.bz * KS;
  const int ids = d.ids, ide = d.ide, jds = d.jds, jde = d.jde, kde = d.kde;
  const int its = ids, ite = ide, jts = jds, jte = jde, kts = d.kds, kte = kde;
  const wrfmi_config &c = a.c;
  const Geo g = geo<PER>(c);
  const float ts = c.time_step >= 0 ? 1.f : -1.f;
  const int ktf = kte - 1;
  const FPtr<float> f = s.q[a.species], ru = s.ru_m, rv = s.rv_m, rom = s.ww_m;
  FPtr<float> tend = s.advect_tend;
  const float rdx = s.rdx, rdy = s.rdy;
  const bool inside = i <= ide && j <= jde;          // a point of the u/v-staggered patch
  const bool writer = inside && (ZONE || lane < kAdvTileI);  // lane 63 of each row only computes a face
  const bool mass = i <= ide - 1 && j <= jde - 1;
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
  auto P = [&](FPtr<float> arr, int ii, int k, int jj) -> float { return arr[X3(d, ii, k, jj)]; };
  // which faces this thread supplies: x face i (rows of mass points), y face j (columns of
  // mass points), and for the top row of the block y face j+1 as well
  const bool fx_on = i <= ide && j <= jde - 1 && i >= ix0 && i <= ix1 + 1;
  const bool fy_on = i <= ide - 1 && j <= jde && j >= jy0 && j <= jy1 + 1;
  const bool fy2_on = row == (int)blockDim.y - 1 && i <= ide - 1 && j + 1 <= jde && j + 1 >= jy0 && j + 1 <= jy1 + 1;
  const bool xdiv = mass && i >= ix0 && i <= ix1, ydiv = mass && j >= jy0 && j <= jy1;
  const float msftx = mass ? s.msftx[X2(d, i, j)] : 0.f;
  __shared__ float fy_s[2][kAdvRows + 1][64];
  float vf_lo = (mass && k0 <= ktf) ? vflux(k0) : 0.f;
  for (int kk = 0; kk < KS; kk++) {
    const int k = k0 + kk;
    if (k > kte) break;  // uniform over the block
    float fx, fx_up, fy, fy_up;
    if (ZONE) {
      fx = (xdiv && k <= ktf) ? fqx(k, i) : 0.f;
      fx_up = (xdiv && k <= ktf) ? fqx(k, i + 1) : 0.f;
      fy = (ydiv && k <= ktf) ? fqy(k, j) : 0.f;
      fy_up = (ydiv && k <= ktf) ? fqy(k, j + 1) : 0.f;
    } else {
      fx = (fx_on && k <= ktf) ? fqx(k, i) : 0.f;
      fx_up = __shfl_down(fx, 1, 64);
      fy = (fy_on && k <= ktf) ? fqy(k, j) : 0.f;
      float(*buf)[64] = fy_s[kk & 1];
      buf[row][lane] = fy;
      if (row == (int)blockDim.y - 1) buf[row + 1][lane] = (fy2_on && k <= ktf) ? fqy(k, j + 1) : 0.f;
      __syncthreads();
      fy_up = buf[row + 1][lane];
    }
    if (writer) {
      // advect_scalar alone (a per-routine entry point) adds to the tendency array's values;
      // rk_scalar_tend's call starts from zero
      float t = a.routine == R_ADVECT ? tend[X3(d, i, k, j)] : 0.f;
      if (mass && k <= ktf) {
        if (ydiv) t = t - msftx * rdy * (fy_up - fy);
        if (xdiv) t = t - msftx * rdx * (fx_up - fx);
        if (g.open_xs && i == its) {
          float ub = 0.5f * (P(ru, its, k, j) + P(ru, its + 1, k, j));
          if (ub > 0.f) ub = 0.f;
          t = t - rdx * (ub * (P(f, its + 1, k, j) - P(f, its, k, j)) + P(f, its, k, j) * (P(ru, its + 1, k, j) - P(ru, its, k, j)));
        }
        if (g.open_xe && i == ide - 1) {
          float ub = 0.5f * (P(ru, ite - 1, k, j) + P(ru, ite, k, j));
          if (ub < 0.f) ub = 0.f;
          t = t - rdx * (ub * (P(f, i, k, j) - P(f, i - 1, k, j)) + P(f, i, k, j) * (P(ru, ite, k, j) - P(ru, ite - 1, k, j)));
        }
        if (g.open_ys && j == jts) {
          float vb = 0.5f * (P(rv, i, k, jts) + P(rv, i, k, jts + 1));
          if (vb > 0.f) vb = 0.f;
          t = t - rdy * (vb * (P(f, i, k, jts + 1) - P(f, i, k, jts)) + P(f, i, k, jts) * (P(rv, i, k, jts + 1) - P(rv, i, k, jts)));
        }
        if (g.open_ye && j == jde - 1) {
          float vb = 0.5f * (P(rv, i, k, jte - 1) + P(rv, i, k, jte));
          if (vb < 0.f) vb = 0.f;
          t = t - rdy * (vb * (P(f, i, k, j) - P(f, i, k, j - 1)) + P(f, i, k, j) * (P(rv, i, k, jte) - P(rv, i, k, jte - 1)));
        }
        const float vf_hi = vflux(k + 1);
        t = t - s.rdnw[X1(d, k)] * (vf_hi - vf_lo);
        vf_lo = vf_hi;
      }
      tend[X3(d, i, k, j)] = t;
    }
  }
}

// advect_scalar for the doubly periodic case with four consecutive i points per lane
// (16-byte loads).  Same fluxes and the same operation order as k_adv_scalar, element by
// element: x faces from a 12-point window (the lane's points and four on each side), y
// faces j and j+1 from the rows j-3..j+3, the vertical flux carried up the column with the
// values of q at k-1..k+2 held in registers.  The lane's points start at an i whose memory
// offset is a multiple of four floats; the host checks the alignment and the padding the
// windows need.
template <typename ARGS>
__global__ void __launch_bounds__(kTallBlk) k_adv_scalar_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int ia0 = d.ims + 4 * ((d.ids - d.ims) >> 2);
  const int i0 = ia0 + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k0 = d.kds + T.bz * kAdvVK;
  const int ids = d.ids, ide = d.ide, jde = d.jde, kts = d.kds, kte = d.kde, ktf = kte - 1;
  if (i0 > ide || j > jde || !zrow(a, j)) return;  // rows of an open y edge: the zone form
  const int ilo = zlo(a, ids), ihi = zhi(a, ide);
  const int k1 = imin(k0 + kAdvVK - 1, kte);
  const float ts = a.c.time_step >= 0 ? 1.f : -1.f;
  const FPtr<float> f = s.q[a.species], ru = s.ru_m, rv = s.rv_m, rom = s.ww_m;
  const FPtr<float> tend = s.advect_tend;
  const float rdx = s.rdx, rdy = s.rdy;
  const int ni = d.ni, sj = d.sj;
  const Off ob = X3(d, i0, 0, j);  // level k at ob + k*ni
  auto at = [&](int k) { return ob + k * ni; };
  const F4 msftx = s.msftx.ld<F4>(X2(d, i0, j));
  const bool jmass = j <= jde - 1;
  // vertical flux at w level kp from q at kp-2..kp+1 (vflux of k_adv_scalar)
  auto vflux = [&](int kp, const F4 &qm2, const F4 &qm1, const F4 &q0, const F4 &qp1, const F4 &w) {
    const float fm = s.fnm[X1(d, imin(kp, kte))], fp = s.fnp[X1(d, imin(kp, kte))];
    const bool zero = kp == kts || kp == kte, low = kp == kts + 1 || kp == ktf;
    F4 r;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const float lo = w.v[n] * (fm * q0.v[n] + fp * qm1.v[n]);
      const float hi = w.v[n] * flux3(qm2.v[n], qm1.v[n], q0.v[n], qp1.v[n], -w.v[n], ts);
      r.v[n] = zero ? 0.f : (low ? lo : hi);
    }
    return r;
  };
  auto ldq = [&](int k) { return f.ld<F4>(at(imax(kts, imin(k, ktf)))); };
  F4 qm1 = ldq(k0 - 1), q0 = ldq(k0), qp1 = ldq(k0 + 1);
  F4 vf_lo = vflux(k0, ldq(k0 - 2), qm1, q0, qp1, rom.ld<F4>(at(imin(k0, kte))));
#pragma unroll
  for (int kk = 0; kk < kAdvVK; kk++) {
    const int k = k0 + kk;
    const Off o = at(imin(k, ktf));
    // x window: points i0-4 .. i0+7
    const F4 wl = f.ld<F4>(o - 4), wr = f.ld<F4>(o + 4);
    const F4 ru0 = ru.ld<F4>(o), ru1 = ru.ld<F4>(o + 4);
    // y rows j-3 .. j+3 (row j is q0)
    const F4 ym3 = f.ld<F4>(o - 3 * sj), ym2 = f.ld<F4>(o - 2 * sj), ym1 = f.ld<F4>(o - sj);
    const F4 yp1 = f.ld<F4>(o + sj), yp2 = f.ld<F4>(o + 2 * sj), yp3 = f.ld<F4>(o + 3 * sj);
    const F4 rv0 = rv.ld<F4>(o), rv1 = rv.ld<F4>(o + sj);
    const F4 qp2 = ldq(k + 2);
    const F4 wp = rom.ld<F4>(at(imin(k + 1, kte)));
    float W[12];
#pragma unroll
    for (int n = 0; n < 4; n++) {
      W[n] = wl.v[n];
      W[n + 4] = q0.v[n];
      W[n + 8] = wr.v[n];
    }
    float fx[5];
#pragma unroll
    for (int n = 0; n < 5; n++) {
      const float vel = n < 4 ? ru0.v[n] : ru1.v[0];
      fx[n] = vel * flux5(W[n + 1], W[n + 2], W[n + 3], W[n + 4], W[n + 5], W[n + 6], vel, ts);
    }
    const F4 vf_hi = vflux(k + 1, qm1, q0, qp1, qp2, wp);
    const float rdnw = s.rdnw[X1(d, imin(k, ktf))];
    F4 t;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const float fy0 = rv0.v[n] * flux5(ym3.v[n], ym2.v[n], ym1.v[n], q0.v[n], yp1.v[n], yp2.v[n], rv0.v[n], ts);
      const float fy1 = rv1.v[n] * flux5(ym2.v[n], ym1.v[n], q0.v[n], yp1.v[n], yp2.v[n], yp3.v[n], rv1.v[n], ts);
      float tt = 0.f;
      tt = tt - msftx.v[n] * rdy * (fy1 - fy0);
      tt = tt - msftx.v[n] * rdx * (fx[n + 1] - fx[n]);
      tt = tt - rdnw * (vf_hi.v[n] - vf_lo.v[n]);
      const bool mass = jmass && i0 + n <= ide - 1 && k <= ktf;
      t.v[n] = mass ? tt : 0.f;
    }
    qm1 = q0;
    q0 = qp1;
    qp1 = qp2;
    vf_lo = vf_hi;
    if (k <= k1) {
      const Off os = at(k);
      if (i0 >= ilo && i0 + 3 <= ihi) {
        tend.st<F4>(os, t);
      } else {
#pragma unroll
        for (int n = 0; n < 4; n++)
          if (i0 + n >= ilo && i0 + n <= ihi) tend[os + n] = t.v[n];
      }
    }
  }
}

template <typename ARGS>
__global__ void __launch_bounds__(256) k_update_scalar(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  if (i > d.ide - 1 || j > d.jde - 1 || k > d.kde - 1) return;
  const Off o = X3(d, i, k, j), o2 = X2(d, i, j);
  FPtr<float> sc1 = s.q_old[a.species], sc2 = s.q[a.species];
  const float muold = s.mu_1[o2] + s.mub[o2];
  const float munew = s.mu_2[o2] + s.mub[o2];
  float tendency = s.advect_tend[o] * s.msfty[o2];
  tendency = tendency + s.q_tend[a.species][o];
  float q1;
  if (a.rk_step == 1) {
    q1 = sc2[o];
    sc1[o] = q1;
  } else {
    q1 = sc1[o];
  }
  const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
  sc2[o] = ((c1 * muold + c2) * q1 + a.dt_rk * tendency) / (c1 * munew + c2);
}

// advect_scalar for every species in one pass (doubly periodic, four i points per lane, one
// level per thread): ru_m, rv_m and ww_m are read once for all species.  Species m's
// tendency goes to the work array m (the last species to advect_tend, as after the
// reference's species loop); element by element the fluxes of k_adv_scalar_x4.
template <typename ARGS>
__global__ void __launch_bounds__(kTallBlk) k_adv_scalar_all_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  const int ids = d.ids, ide = d.ide, jde = d.jde, kts = d.kds, kte = d.kde, ktf = kte - 1;
  if (i0 > ide || j > jde || k > kte) return;
  const float ts = a.c.time_step >= 0 ? 1.f : -1.f;
  const float rdx = s.rdx, rdy = s.rdy;
  const int ni = d.ni, sj = d.sj;
  const Off ob = X3(d, i0, 0, j);
  auto lq = [&](int kk) { return ob + imax(kts, imin(kk, ktf)) * ni; };  // q levels kts..ktf
  auto lw = [&](int kk) { return ob + imax(kts, imin(kk, kte)) * ni; };  // w levels kts..kte
  const Off o = lq(k), os = lw(k);
  const F4 msftx = s.msftx.ld<F4>(X2(d, i0, j));
  const F4 ru0 = s.ru_m.ld<F4>(o), ru1 = s.ru_m.ld<F4>(o + 4), rv0 = s.rv_m.ld<F4>(o), rv1 = s.rv_m.ld<F4>(o + sj);
  const F4 w0 = s.ww_m.ld<F4>(lw(k)), w1 = s.ww_m.ld<F4>(lw(k + 1));
  const float rdnw = s.rdnw[X1(d, imin(k, ktf))];
  const float fm0 = s.fnm[X1(d, k)], fp0 = s.fnp[X1(d, k)], fm1 = s.fnm[X1(d, imin(k + 1, kte))], fp1 = s.fnp[X1(d, imin(k + 1, kte))];
  auto vflux = [&](int kp, float fm, float fp, float w, float qm2, float qm1, float q0, float qp1) {
    const bool zero = kp == kts || kp == kte, low = kp == kts + 1 || kp == ktf;
    const float lo = w * (fm * q0 + fp * qm1);
    const float hi = w * flux3(qm2, qm1, q0, qp1, -w, ts);
    return zero ? 0.f : (low ? lo : hi);
  };
  const bool jmass = j <= jde - 1;
  const int ns = s.n_species;
  for (int m = 0; m < ns; m++) {
    const FPtr<float> f = s.q[m];
    const F4 q0 = f.ld<F4>(o), wl = f.ld<F4>(o - 4), wr = f.ld<F4>(o + 4);
    const F4 ym3 = f.ld<F4>(o - 3 * sj), ym2 = f.ld<F4>(o - 2 * sj), ym1 = f.ld<F4>(o - sj);
    const F4 yp1 = f.ld<F4>(o + sj), yp2 = f.ld<F4>(o + 2 * sj), yp3 = f.ld<F4>(o + 3 * sj);
    const F4 qm2 = f.ld<F4>(lq(k - 2)), qm1 = f.ld<F4>(lq(k - 1)), qp1 = f.ld<F4>(lq(k + 1)), qp2 = f.ld<F4>(lq(k + 2));
    float W[12];
#pragma unroll
    for (int n = 0; n < 4; n++) {
      W[n] = wl.v[n];
      W[n + 4] = q0.v[n];
      W[n + 8] = wr.v[n];
    }
    float fx[5];
#pragma unroll
    for (int n = 0; n < 5; n++) {
      const float vel = n < 4 ? ru0.v[n] : ru1.v[0];
      fx[n] = vel * flux5(W[n + 1], W[n + 2], W[n + 3], W[n + 4], W[n + 5], W[n + 6], vel, ts);
    }
    F4 t;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const float vf_lo = vflux(k, fm0, fp0, w0.v[n], qm2.v[n], qm1.v[n], q0.v[n], qp1.v[n]);
      const float vf_hi = vflux(k + 1, fm1, fp1, w1.v[n], qm1.v[n], q0.v[n], qp1.v[n], qp2.v[n]);
      const float fy0 = rv0.v[n] * flux5(ym3.v[n], ym2.v[n], ym1.v[n], q0.v[n], yp1.v[n], yp2.v[n], rv0.v[n], ts);
      const float fy1 = rv1.v[n] * flux5(ym2.v[n], ym1.v[n], q0.v[n], yp1.v[n], yp2.v[n], yp3.v[n], rv1.v[n], ts);
      float tt = 0.f;
      tt = tt - msftx.v[n] * rdy * (fy1 - fy0);
      tt = tt - msftx.v[n] * rdx * (fx[n + 1] - fx[n]);
      tt = tt - rdnw * (vf_hi - vf_lo);
      const bool mass = jmass && i0 + n <= ide - 1 && k <= ktf;
      t.v[n] = mass ? tt : 0.f;
    }
    const FPtr<float> dst = m == ns - 1 ? s.advect_tend : FPtr<float>(a.scr_qt + (size_t)m * a.qt_stride);
    st4m(dst, os, t, i0, ids, ide);
  }
}

// rk_update_scalar for every species in one pass (four i points per lane), reading each
// species' advection tendency where k_adv_scalar_all_x4 left it
template <typename ARGS>
__global__ void __launch_bounds__(256) k_update_scalar_all_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  if (i0 > d.ide - 1 || j > d.jde - 1 || k > d.kde - 1) return;
  const Off o = X3(d, i0, k, j), o2 = X2(d, i0, j);
  const F4 mu1 = s.mu_1.ld<F4>(o2), mu2 = s.mu_2.ld<F4>(o2), mub = s.mub.ld<F4>(o2), msfty = s.msfty.ld<F4>(o2);
  const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
  const bool rk1 = a.rk_step == 1;
  const int ns = s.n_species;
  for (int m = 0; m < ns; m++) {
    const FPtr<float> sc1 = s.q_old[m], sc2 = s.q[m];
    const FPtr<float> src = m == ns - 1 ? s.advect_tend : FPtr<float>(a.scr_qt + (size_t)m * a.qt_stride);
    const F4 adv = src.ld<F4>(o), qt = s.q_tend[m].ld<F4>(o);
    // the old time level: the current values at rk_step 1 (then saved), the saved ones after
    const F4 q1 = (rk1 ? sc2 : sc1).ld<F4>(o);
    F4 qn;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const float muold = mu1.v[n] + mub.v[n];
      const float munew = mu2.v[n] + mub.v[n];
      float tendency = adv.v[n] * msfty.v[n];
      tendency = tendency + qt.v[n];
      qn.v[n] = ((c1 * muold + c2) * q1.v[n] + a.dt_rk * tendency) / (c1 * munew + c2);
    }
    if (rk1) st4m(sc1, o, q1, i0, d.ids, d.ide - 1);
    st4m(sc2, o, qn, i0, d.ids, d.ide - 1);
  }
}

// rk_update_scalar with four i points per lane (16-byte loads and stores)
template <typename ARGS>
__global__ void __launch_bounds__(256) k_update_scalar_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  if (i0 > d.ide - 1 || j > d.jde - 1 || k > d.kde - 1) return;
  const Off o = X3(d, i0, k, j), o2 = X2(d, i0, j);
  const FPtr<float> sc1 = s.q_old[a.species], sc2 = s.q[a.species];
  const F4 mu1 = s.mu_1.ld<F4>(o2), mu2 = s.mu_2.ld<F4>(o2), mub = s.mub.ld<F4>(o2), msfty = s.msfty.ld<F4>(o2);
  const F4 adv = s.advect_tend.ld<F4>(o), qt = s.q_tend[a.species].ld<F4>(o);
  const bool rk1 = a.rk_step == 1;
  // the old time level: the current values at rk_step 1 (then saved), the saved ones after
  const F4 q1 = (rk1 ? sc2 : sc1).ld<F4>(o);
  const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
  F4 qn;
#pragma unroll
  for (int n = 0; n < 4; n++) {
    const float muold = mu1.v[n] + mub.v[n];
    const float munew = mu2.v[n] + mub.v[n];
    float tendency = adv.v[n] * msfty.v[n];
    tendency = tendency + qt.v[n];
    qn.v[n] = ((c1 * muold + c2) * q1.v[n] + a.dt_rk * tendency) / (c1 * munew + c2);
  }
  if (rk1) st4m(sc1, o, q1, i0, d.ids, d.ide - 1);
  st4m(sc2, o, qn, i0, d.ids, d.ide - 1);
}

// calc_p_rho_phi with four i points per lane (element by element as k_p_rho_phi)
template <typename ARGS>
__global__ void __launch_bounds__(256) k_p_rho_phi_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  if (i0 > d.ide - 1 || j > d.jde - 1 || k > d.kde - 1) return;
  const Off o = X3(d, i0, k, j), o2 = X2(d, i0, j);
  const F4 muts = s.muts.ld<F4>(o2), mu = s.mu_2.ld<F4>(o2);
  const F4 php = s.ph_2.ld<F4>(o + d.ni), ph = s.ph_2.ld<F4>(o), alb = s.alb.ld<F4>(o);
  const F4 pbp = s.phb.ld<F4>(o + d.ni), pb0 = s.phb.ld<F4>(o), t = s.t_2.ld<F4>(o), pbs = s.pb.ld<F4>(o);
  const bool hyp1 = a.c.hypsometric_opt == 1;
  const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)], rdnw = s.rdnw[X1(d, k)];
  const float c3fp = s.c3f[X1(d, k + 1)], c4fp = s.c4f[X1(d, k + 1)], c3f = s.c3f[X1(d, k)], c4f = s.c4f[X1(d, k)];
  const float c3h = s.c3h[X1(d, k)], c4h = s.c4h[X1(d, k)], ptop = s.p_top;
  const bool moist = s.n_moist >= 1, thm = a.c.use_theta_m == 1;
  F4 qv;
  if (moist && !thm) qv = s.q[0].ld<F4>(o);
  else
#pragma unroll
    for (int n = 0; n < 4; n++) qv.v[n] = 0.f;
  F4 aln, pn;
#pragma unroll
  for (int n = 0; n < 4; n++) {
    float al;
    if (hyp1) {
      al = -1.f / (c1 * muts.v[n] + c2) * (alb.v[n] * (c1 * mu.v[n]) + rdnw * (php.v[n] - ph.v[n]));
    } else {
      const float pfu = c3fp * muts.v[n] + c4fp + ptop;
      const float pfd = c3f * muts.v[n] + c4f + ptop;
      const float phm = c3h * muts.v[n] + c4h + ptop;
      const float lg = (float)log((double)(pfd / pfu));
      al = (php.v[n] - ph.v[n] + pbp.v[n] - pb0.v[n]) / phm / lg - alb.v[n];
    }
    aln.v[n] = al;
    if (moist) {
      float temp;
      if (thm) {
        temp = (K_RD * (K_T0 + t.v[n])) / (K_P0 * (al + alb.v[n]));
      } else {
        const float qvf = 1.f + K_RVOVRD * qv.v[n];
        temp = (K_RD * (K_T0 + t.v[n]) * qvf) / (K_P0 * (al + alb.v[n]));
      }
      const float pw = (float)pow((double)temp, (double)K_CPOVCV);
      pn.v[n] = pw * K_P0 - pbs.v[n];
    } else {
      const float pw = (float)pow((double)((K_RD * (K_T0 + t.v[n])) / (K_P0 * (al + alb.v[n]))), (double)K_CPOVCV);
      pn.v[n] = K_P0 * pw - pbs.v[n];
    }
  }
  st4m(s.al, o, aln, i0, d.ids, d.ide - 1);
  st4m(s.p, o, pn, i0, d.ids, d.ide - 1);
}

// calc_p_rho_phi, non-hydrostatic; logarithm and power in FP64 then rounded once to FP32
template <typename ARGS>
__global__ void __launch_bounds__(256) k_p_rho_phi(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  if (i > d.ide - 1 || j > d.jde - 1 || k > d.kde - 1) return;
  const Off o = X3(d, i, k, j), o2 = X2(d, i, j);
  const float muts = s.muts[o2], mu = s.mu_2[o2];
  float al;
  if (a.c.hypsometric_opt == 1) {
    const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
    al = -1.f / (c1 * muts + c2) * (s.alb[o] * (c1 * mu) + s.rdnw[X1(d, k)] * (s.ph_2[o + d.ni] - s.ph_2[o]));
  } else {
    const float ptop = s.p_top;
    const float pfu = s.c3f[X1(d, k + 1)] * muts + s.c4f[X1(d, k + 1)] + ptop;
    const float pfd = s.c3f[X1(d, k)] * muts + s.c4f[X1(d, k)] + ptop;
    const float phm = s.c3h[X1(d, k)] * muts + s.c4h[X1(d, k)] + ptop;
    const float lg = (float)log((double)(pfd / pfu));
    al = (s.ph_2[o + d.ni] - s.ph_2[o] + s.phb[o + d.ni] - s.phb[o]) / phm / lg - s.alb[o];
  }
  s.al[o] = al;
  const float t = s.t_2[o];
  float temp;
  if (s.n_moist >= 1) {
    if (a.c.use_theta_m == 1) {
      temp = (K_RD * (K_T0 + t)) / (K_P0 * (al + s.alb[o]));
    } else {
      const float qvf = 1.f + K_RVOVRD * s.q[0][o];
      temp = (K_RD * (K_T0 + t) * qvf) / (K_P0 * (al + s.alb[o]));
    }
    const float pw = (float)pow((double)temp, (double)K_CPOVCV);
    s.p[o] = pw * K_P0 - s.pb[o];
  } else {
    const float pw = (float)pow((double)((K_RD * (K_T0 + t)) / (K_P0 * (al + s.alb[o]))), (double)K_CPOVCV);
    s.p[o] = K_P0 * pw - s.pb[o];
  }
}

// ----------------------------------------------------------------------------- halo fills
// Value a filled halo cell ends up with after set_physical_bc's x pass followed by its
// y pass (periodic wrap, or the open-boundary zero-gradient copy), computed directly
// from the source cell so that all cells of all fields fill in one launch.
// One thread per (halo cell, group of kHaloK levels, field).  The halo ring of the memory patch is
// enumerated as: the rows below jds and from jde up (every i), then the columns left of ids
// and from ide on for the rows jds..jde-1.
template <typename ARGS>
__global__ void __launch_bounds__(256) k_halo(ARGS args) {
  unsigned blk;
  const HaloArgs &h = hargs(args, blk);
  const Dm &d = h.d;
  const unsigned W = (unsigned)d.ni, B = (unsigned)(d.jds - d.jms), Tn = (unsigned)(d.jme - d.jde + 1);
  const unsigned Lw = (unsigned)(d.ids - d.ims), LR = Lw + (unsigned)(d.ime - d.ide + 1);
  const unsigned nring = (B + Tn) * W + (unsigned)(d.jde - d.jds) * LR, nkc = (unsigned)(d.kde + kHaloK - 1) / kHaloK;
  const unsigned t = blk * blockDim.x + threadIdx.x;
  if (t >= nring * nkc * (unsigned)h.n) return;
  const unsigned fidx = t / (nring * nkc), rem = t % (nring * nkc);
  const int k0 = d.kds + (int)(rem / nring) * kHaloK;
  unsigned c = rem % nring;
  int i, j;
  if (c < B * W) {
    j = d.jms + (int)(c / W); i = d.ims + (int)(c % W);
  } else if (c < (B + Tn) * W) {
    c -= B * W;
    j = d.jde + (int)(c / W); i = d.ims + (int)(c % W);
  } else {
    c -= (B + Tn) * W;
    j = d.jds + (int)(c / LR);
    const unsigned cc = c % LR;
    i = cc < Lw ? d.ims + (int)cc : d.ide + (int)(cc - Lw);
  }
  const HaloField F = h.f[fidx];
  const int ids = d.ids, ide = d.ide, jds = d.jds, jde = d.jde;
  const int istag = F.kind == 1 ? 0 : -1, jstag = F.kind == 2 ? 0 : -1;
  const int kend = F.is3d ? (F.kind == 3 ? d.kde : d.kde - 1) : 1;
  const int nx = ide - ids, ny = jde - jds;
  // x pass: which i are written, over which j, and from where
  int xsrc = i;
  bool xfill = false;
  int xj0 = 0, xj1 = -1;
  if (h.periodic_x) {
    if (i <= ids - 1 && i >= ids - 4) { xsrc = i + nx; xfill = true; }
    else if (i >= ide && i <= ide + 4 + istag) { xsrc = i - nx; xfill = true; }
    xj0 = jds; xj1 = jde + jstag;
  } else {
    if (h.open_xs && i <= ids - 1 && i >= ids - 3) { xsrc = ids; xfill = true; }
    if (h.open_xe) {
      if (istag == -1 && i >= ide && i <= ide + 2) { xsrc = ide - 1; xfill = true; }
      if (istag == 0 && i >= ide + 1 && i <= ide + 3) { xsrc = ide; xfill = true; }
    }
    if (F.is3d) {
      xj0 = jds - 4;
      xj1 = jde + jstag + 4;
    } else {
      xj0 = jds; xj1 = jde + jstag;
    }
  }
  // y pass over every i of the memory row
  int ysrc = j;
  bool yfill = false;
  if (h.periodic_y) {
    if (j <= jds - 1 && j >= jds - 4) { ysrc = j + ny; yfill = true; }
    else if (j >= jde && j <= jde + 4 + jstag) { ysrc = j - ny; yfill = true; }
  } else {
    if (h.open_ys && j <= jds - 1 && j >= jds - 3) { ysrc = jds; yfill = true; }
    if (h.open_ye) {
      if (jstag == -1 && j >= jde && j <= jde + 2) { ysrc = jde - 1; yfill = true; }
      if (jstag == 0 && j >= jde + 1 && j <= jde + 3) { ysrc = jde; yfill = true; }
    }
  }
  int si = i, sj = j;
  if (yfill) {
    sj = ysrc;
    if (xfill && sj >= xj0 && sj <= xj1) si = xsrc;
  } else if (xfill && j >= xj0 && j <= xj1) {
    si = xsrc;
  } else {
    return;
  }
  if (!F.is3d) {
    if (k0 == d.kds) F.p[X2i(d, i, j)] = F.p[X2i(d, si, sj)];
    return;
  }
  // kHaloK levels per thread: all loads first (levels past kend repeat kend), then the stores
  float v[kHaloK];
#pragma unroll
  for (int kk = 0; kk < kHaloK; kk++) v[kk] = F.p[X3i(d, si, imin(k0 + kk, kend), sj)];
#pragma unroll
  for (int kk = 0; kk < kHaloK; kk++)
    if (k0 + kk <= kend) F.p[X3i(d, i, k0 + kk, j)] = v[kk];
}

WRFMI_INST_P2(k_adv_scalar, false, false)
WRFMI_INST_P2(k_adv_scalar, true, false)
WRFMI_INST_P2(k_adv_scalar, false, true)

WRFMI_INST(k_adv_scalar_x4)
WRFMI_INST(k_update_scalar)
WRFMI_INST(k_update_scalar_x4)
WRFMI_INST(k_adv_scalar_all_x4)
WRFMI_INST(k_update_scalar_all_x4)
WRFMI_INST(k_p_rho_phi)
WRFMI_INST(k_p_rho_phi_x4)
template __global__ void k_halo<HaloArgs>(HaloArgs);
template __global__ void k_halo<BatchHalo>(BatchHalo);

}  // namespace wrfmi
