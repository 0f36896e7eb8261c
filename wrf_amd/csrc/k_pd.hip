// Positive-definite scalar transport on the last RK substep.
//   k_update_scalar_pd : rk_update_scalar_pd (dyn_em/module_em.F:1803-1916), the species'
//                        physics tendency taken into its time-n values before the substep
//                        (solve_em.F:1850-1975 calls it with mu_1 as both masses)
//   k_pd_scale         : advect_scalar_pd's limiter (dyn_em/module_advect_em.F:7681-7779):
//                        per cell, the scale of its outgoing high-order corrections
//   k_pd_tend          : advect_scalar_pd's limited flux divergence (7781-7883) with the
//                        radiation terms of open boundaries (7255-7326)
// The reference keeps six flux arrays (high-order minus upwind, and upwind, in x, y, z) and
// scales the faces cell by cell.  Each face is scaled by at most one cell -- the one it
// leaves: a positive x flux at face i by cell i-1, a negative one by cell i (z: a positive
// flux at w level k by cell k, a negative one by cell k-1) -- and the scale only shrinks
// it, so the order of the cells does not matter.  Here k_pd_scale stores every cell's
// scale (1 for a cell the limiter leaves alone or does not visit; a face times 1 is itself)
// in a work array, and k_pd_tend recomputes the faces of its cell, scales each by the
// scale of the cell it leaves and forms the divergence: no flux arrays, two passes over the
// inputs.  Every flux is the same expression, in the same order, as in advect_scalar_pd,
// including its own product form of the flux operators (module_advect_em.F:6166-6186).
#include "kernels.h"

namespace wrfmi {

// flux operators of advect_scalar_pd (module_advect_em.F:6166-6186); ts = SIGN(1,time_step)
DEVI float pd_flux4(float qm2, float qm1, float q0, float qp1) {
  return (7.f / 12.f) * (q0 + qm1) - (1.f / 12.f) * (qp1 + qm2);
}
DEVI float pd_flux3(float qm2, float qm1, float q0, float qp1, float ua, float ts) {
  return pd_flux4(qm2, qm1, q0, qp1) + ts * sgn1(ua) * (1.f / 12.f) * ((qp1 - qm2) - 3.f * (q0 - qm1));
}
DEVI float pd_flux6(float qm3, float qm2, float qm1, float q0, float qp1, float qp2) {
  return (37.f / 60.f) * (q0 + qm1) - (2.f / 15.f) * (qp1 + qm2) + (1.f / 60.f) * (qp2 + qm3);
}
DEVI float pd_flux5(float qm3, float qm2, float qm1, float q0, float qp1, float qp2, float ua, float ts) {
  return pd_flux6(qm3, qm2, qm1, q0, qp1, qp2) -
         ts * sgn1(ua) * (1.f / 60.f) * ((qp2 - qm3) - 5.f * (qp1 - qm2) + 10.f * (q0 - qm1));
}
// Fortran MIN/MAX of two REALs as the reference build evaluates them (the first argument
// when the comparison fails)
DEVI float fmin_r(float a, float b) { return a < b ? a : b; }
DEVI float fmax_r(float a, float b) { return a > b ? a : b; }
// first-order upwind flux of Smolarkiewicz (1989) for Courant number cr
DEVI float pd_upwind(float qm1, float q0, float cr) {
  return 0.5f * fmin_r(1.0f, (cr + __builtin_fabsf(cr))) * qm1 + 0.5f * fmax_r(-1.0f, (cr - __builtin_fabsf(cr))) * q0;
}

// The faces of advect_scalar_pd for one domain.  fq is the high-order flux minus the upwind
// flux (the reference's fqx/fqy/fqz before the limiter), fl the upwind flux (fqxl/...).
struct PdFaces {
  const DState &s;
  const Dm &d;
  FPtr<float> f, fo, ru, rv, rom, mut;
  float ts, dt, rdx, rdy;
  bool per_x, per_y;

  // x face ii of row j (between mass points ii-1 and ii): 5th order, or next to an open
  // boundary 2nd order at ids+1 / ide-1 and 3rd at ids+2 / ide-2 (6658-6781)
  DEVI void fx(int ii, int k, int j, float &fq, float &fl) const {
    const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
    const float dx = 2.f / (s.msfty[X2(d, ii, j)] + s.msfty[X2(d, ii - 1, j)]) / rdx;
    const float mu = 0.5f * ((c1 * mut[X2(d, ii, j)] + c2) + (c1 * mut[X2(d, ii - 1, j)] + c2));
    const float r = ru[X3(d, ii, k, j)];
    const float om1 = fo[X3(d, ii - 1, k, j)], o0 = fo[X3(d, ii, k, j)];
    if (!per_x && ii == d.ids + 1) {
      // 6733-6736: here the velocity is formed first and cr has no 1/mu
      const float vel = r / mu;
      const float cr = vel * dt / dx;
      fl = mu * (dx / dt) * pd_upwind(om1, o0, cr);
      fq = 0.5f * (r) * (f[X3(d, ii, k, j)] + f[X3(d, ii - 1, k, j)]);
    } else {
      const float vel = r;
      const float cr = vel * dt / dx / mu;
      fl = mu * (dx / dt) * pd_upwind(om1, o0, cr);
      if (per_x || (ii >= d.ids + 3 && ii <= d.ide - 3))
        fq = vel * pd_flux5(f[X3(d, ii - 3, k, j)], f[X3(d, ii - 2, k, j)], f[X3(d, ii - 1, k, j)], f[X3(d, ii, k, j)],
                            f[X3(d, ii + 1, k, j)], f[X3(d, ii + 2, k, j)], vel, ts);
      else if (ii == d.ide - 1)
        fq = 0.5f * (r) * (f[X3(d, ii, k, j)] + f[X3(d, ii - 1, k, j)]);
      else
        fq = vel * pd_flux3(f[X3(d, ii - 2, k, j)], f[X3(d, ii - 1, k, j)], f[X3(d, ii, k, j)], f[X3(d, ii + 1, k, j)],
                            vel, ts);
    }
    fq = fq - fl;
  }
  // y face jj of column i (6523-6585)
  DEVI void fy(int i, int k, int jj, float &fq, float &fl) const {
    const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
    const float dy = 2.f / (s.msftx[X2(d, i, jj)] + s.msftx[X2(d, i, jj - 1)]) / rdy;
    const float mu = 0.5f * ((c1 * mut[X2(d, i, jj)] + c2) + (c1 * mut[X2(d, i, jj - 1)] + c2));
    const float vel = rv[X3(d, i, k, jj)];
    const float cr = vel * dt / dy / mu;
    fl = mu * (dy / dt) * pd_upwind(fo[X3(d, i, k, jj - 1)], fo[X3(d, i, k, jj)], cr);
    if (per_y || (jj >= d.jds + 3 && jj <= d.jde - 3))
      fq = vel * pd_flux5(f[X3(d, i, k, jj - 3)], f[X3(d, i, k, jj - 2)], f[X3(d, i, k, jj - 1)], f[X3(d, i, k, jj)],
                          f[X3(d, i, k, jj + 1)], f[X3(d, i, k, jj + 2)], vel, ts);
    else if (jj == d.jds + 1 || jj == d.jde - 1)
      fq = 0.5f * rv[X3(d, i, k, jj)] * (f[X3(d, i, k, jj)] + f[X3(d, i, k, jj - 1)]);
    else
      fq = vel * pd_flux3(f[X3(d, i, k, jj - 2)], f[X3(d, i, k, jj - 1)], f[X3(d, i, k, jj)], f[X3(d, i, k, jj + 1)],
                          vel, ts);
    fq = fq - fl;
  }
  // z face at w level kk of column (i, j), order 3 with 2nd-order interpolation next to the
  // ground and the top and no flux through them (7595-7645)
  DEVI void fz(int i, int kk, int j, float &fq, float &fl) const {
    const int kts = d.kds, ktf = d.kde - 1;
    if (kk <= kts || kk >= d.kde) {
      fq = 0.f;
      fl = 0.f;
      return;
    }
    const float c1 = s.c1h[X1(d, kk)], c2 = s.c2h[X1(d, kk)];
    const float dz = 2.f / (s.rdnw[X1(d, kk)] + s.rdnw[X1(d, kk - 1)]);
    const float m = mut[X2(d, i, j)];
    const float mu = 0.5f * ((c1 * m + c2) + (c1 * m + c2));
    const float vel = rom[X3(d, i, kk, j)];
    const float cr = vel * dt / dz / mu;
    fl = mu * (dz / dt) * pd_upwind(fo[X3(d, i, kk - 1, j)], fo[X3(d, i, kk, j)], cr);
    if (kk == kts + 1 || kk == ktf)
      fq = rom[X3(d, i, kk, j)] * (s.fnm[X1(d, kk)] * f[X3(d, i, kk, j)] + s.fnp[X1(d, kk)] * f[X3(d, i, kk - 1, j)]);
    else
      fq = vel * pd_flux3(f[X3(d, i, kk - 2, j)], f[X3(d, i, kk - 1, j)], f[X3(d, i, kk, j)], f[X3(d, i, kk + 1, j)],
                          -vel, ts);
    fq = fq - fl;
  }
};

DEVI PdFaces pd_faces(const KArgs &a) {
  const DState &s = dstate(a);
  const int sp = a.species;
  return PdFaces{s, a.d, s.q[sp], s.q_old[sp], s.ru_m, s.rv_m, s.ww_m, s.muts, a.c.time_step >= 0 ? 1.f : -1.f,
                 a.dt_rk, s.rdx, s.rdy, (bool)a.c.periodic_x, (bool)a.c.periodic_y};
}
// the limiter's cells (7681-7716): all with periodic directions, one point in from an
// open boundary otherwise
DEVI bool pd_limited_cell(const KArgs &a, int i, int j) {
  const Dm &d = a.d;
  const int i0 = a.c.periodic_x ? d.ids - 1 : d.ids + 1, i1 = a.c.periodic_x ? d.ide : d.ide - 2;
  const int j0 = a.c.periodic_y ? d.jds - 1 : d.jds + 1, j1 = a.c.periodic_y ? d.jde : d.jde - 2;
  return i >= i0 && i <= i1 && j >= j0 && j <= j1;
}

// Scales of the cells i = ids-1..ide, j = jds-1..jde (the cells whose scales k_pd_tend
// reads), kPdK levels per thread; cells the limiter does not visit get 1.
template <typename ARGS>
__global__ void __launch_bounds__(256) k_pd_scale(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids - 1 + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds - 1 + (int)(T.by * blockDim.y + threadIdx.y);
  const int k0 = d.kds + T.bz * kPdK, ktf = d.kde - 1, k1 = imin(k0 + kPdK - 1, ktf);
  if (i > d.ide || j > d.jde || k0 > ktf) return;
  const FPtr<float> S = a.scr_dv;
  if (!pd_limited_cell(a, i, j)) {
    for (int k = k0; k <= k1; k++) S[X3(d, i, k, j)] = 1.f;
    return;
  }
  const PdFaces F = pd_faces(a);
  const Off o2 = X2(d, i, j);
  const float msftx = s.msftx[o2], msfty = s.msfty[o2], mub = s.mub[o2], mu_old = s.mu_1[o2];
  const float rdx = s.rdx, rdy = s.rdy, dt = F.dt;
  const float eps = 1.e-20f;
  float zq_lo, zl_lo;
  F.fz(i, k0, j, zq_lo, zl_lo);
  for (int k = k0; k <= k1; k++) {
    float xq0, xl0, xq1, xl1, yq0, yl0, yq1, yl1, zq1, zl1;
    F.fx(i, k, j, xq0, xl0);
    F.fx(i + 1, k, j, xq1, xl1);
    F.fy(i, k, j, yq0, yl0);
    F.fy(i, k, j + 1, yq1, yl1);
    F.fz(i, k + 1, j, zq1, zl1);
    const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)], rdzw = s.rdnw[X1(d, k)];
    const Off o = X3(d, i, k, j);
    const float ph_low = ((c1 * mub + c2) + (c1 * mu_old)) * F.fo[o] -
                         dt * (msftx * msfty * (rdx * (xl1 - xl0) + rdy * (yl1 - yl0)) + msfty * rdzw * (zl1 - zl_lo));
    const float flux_out = dt * ((msftx * msfty) * (rdx * (fmax_r(0.f, xq1) - fmin_r(0.f, xq0)) +
                                                    rdy * (fmax_r(0.f, yq1) - fmin_r(0.f, yq0))) +
                                 msfty * rdzw * (fmin_r(0.f, zq1) - fmax_r(0.f, zq_lo)));
    S[o] = flux_out > ph_low ? fmax_r(0.f, ph_low / (flux_out + eps)) : 1.f;
    zq_lo = zq1;
    zl_lo = zl1;
  }
}

// advect_scalar_pd's tendency of species a.species at the points i = ids..ide,
// j = jds..jde, k = kds..kde (zero where the reference's rk_scalar_tend leaves zero), kPdK
// levels per thread.  The scale array holds k_pd_scale's result.
template <typename ARGS>
__global__ void __launch_bounds__(256) k_pd_tend(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int kts = d.kds, kte = d.kde, ktf = kte - 1;
  const int k0 = kts + T.bz * kPdK, k1 = imin(k0 + kPdK - 1, kte);
  if (i > d.ide || j > d.jde || k0 > kte) return;
  const FPtr<float> S = a.scr_dv, tend = s.advect_tend;
  const bool mass = i <= d.ide - 1 && j <= d.jde - 1;
  const bool add = a.routine == R_ADVECT;  // the routine form adds to the tendency's values
  if (!mass) {
    if (!add)
      for (int k = k0; k <= k1; k++) tend[X3(d, i, k, j)] = 0.f;
    return;
  }
  const PdFaces F = pd_faces(a);
  const bool per_x = F.per_x, per_y = F.per_y;
  const bool xdiv = per_x || (i >= d.ids + 1 && i <= d.ide - 2), ydiv = per_y || (j >= d.jds + 1 && j <= d.jde - 2);
  const float msftx = s.msftx[X2(d, i, j)], rdx = s.rdx, rdy = s.rdy;
  const FPtr<float> f = F.f, fo = F.fo, ru = F.ru, rv = F.rv;
  // a face flux scaled by the cell it leaves: `up` the cell below the face, `dn` above it
  auto lim = [&](float fq, Off up, Off dn) { return fq > 0.f ? S[up] * fq : fq < 0.f ? S[dn] * fq : fq; };
  float zq_lo = 0.f, zl_lo = 0.f;
  if (k0 <= ktf) {
    F.fz(i, k0, j, zq_lo, zl_lo);
    // w level k0 lies between the cells k0-1 and k0; a positive flux leaves cell k0
    if (k0 > kts) zq_lo = zq_lo > 0.f ? S[X3(d, i, k0, j)] * zq_lo : zq_lo < 0.f ? S[X3(d, i, k0 - 1, j)] * zq_lo : zq_lo;
  }
  for (int k = k0; k <= k1; k++) {
    const Off o = X3(d, i, k, j);
    float t = add ? tend[o] : 0.f;
    if (k <= ktf) {
      // radiation terms of open boundaries (7255-7326; not of specified ones), before the
      // flux divergence
      if (a.c.open_xs && i == d.ids) {
        const float ub = fmin_r(0.5f * (ru[o] + ru[o + 1]), 0.f);
        t = t - rdx * (ub * (fo[o + 1] - fo[o]) + f[o] * (ru[o + 1] - ru[o]));
      }
      if (a.c.open_xe && i == d.ide - 1) {
        const Off oe = X3(d, d.ide, k, j);
        const float ub = fmax_r(0.5f * (ru[oe - 1] + ru[oe]), 0.f);
        t = t - rdx * (ub * (fo[o] - fo[o - 1]) + f[o] * (ru[oe] - ru[oe - 1]));
      }
      if (a.c.open_ys && j == d.jds) {
        const float vb = fmin_r(0.5f * (rv[o] + rv[o + d.sj]), 0.f);
        t = t - rdy * (vb * (fo[o + d.sj] - fo[o]) + f[o] * (rv[o + d.sj] - rv[o]));
      }
      if (a.c.open_ye && j == d.jde - 1) {
        const Off on = X3(d, i, k, d.jde);
        const float vb = fmax_r(0.5f * (rv[on - d.sj] + rv[on]), 0.f);
        t = t - rdy * (vb * (fo[o] - fo[o - d.sj]) + f[o] * (rv[on] - rv[on - d.sj]));
      }
      // z: w level k+1 lies between the cells k and k+1; a positive flux leaves cell k+1
      float zq1, zl1;
      F.fz(i, k + 1, j, zq1, zl1);
      if (k + 1 <= ktf) zq1 = lim(zq1, o + d.ni, o);
      const float rdzw = s.rdnw[X1(d, k)];
      t = t - rdzw * (zq1 - zq_lo + zl1 - zl_lo);
      zq_lo = zq1;
      zl_lo = zl1;
      if (xdiv) {
        float xq0, xl0, xq1, xl1;
        F.fx(i, k, j, xq0, xl0);
        F.fx(i + 1, k, j, xq1, xl1);
        xq0 = lim(xq0, o - 1, o);
        xq1 = lim(xq1, o, o + 1);
        t = t - msftx * (rdx * (xq1 - xq0 + xl1 - xl0));
      }
      if (ydiv) {
        float yq0, yl0, yq1, yl1;
        F.fy(i, k, j, yq0, yl0);
        F.fy(i, k, j + 1, yq1, yl1);
        yq0 = lim(yq0, o - d.sj, o);
        yq1 = lim(yq1, o, o + d.sj);
        t = t - msftx * (rdy * (yq1 - yq0 + yl1 - yl0));
      }
    }
    if (k <= ktf || !add) tend[o] = t;
  }
}

// rk_update_scalar_pd of species a.species over the mass points: the physics tendency
// q_tend, coupled, into q_old with mu_1 + mub as the old and mu_2 + mub as the new mass,
// then q_tend = 0 (the whole step launches it with mu_2 bound to mu_1, as solve_em calls it)
template <typename ARGS>
__global__ void __launch_bounds__(256) k_update_scalar_pd(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i = d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int k = d.kds + (int)T.bz;
  if (i > d.ide - 1 || j > d.jde - 1 || k > d.kde - 1) return;
  const Off o = X3(d, i, k, j), o2 = X2(d, i, j);
  const FPtr<float> sc = s.q_old[a.species], sct = s.q_tend[a.species];
  const float muold = s.mu_1[o2] + s.mub[o2];
  const float munew = s.mu_2[o2] + s.mub[o2];
  float tendency = 0.f;
  tendency = tendency + sct[o];
  sct[o] = 0.f;
  const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
  sc[o] = ((c1 * muold + c2) * sc[o] + a.dt_rk * tendency) / (c1 * munew + c2);
}

// ---------------------------------------------------------------------------------------
// Four points per lane, doubly periodic domains of the whole step.  A lane holds the cells
// i0..i0+3 (16-byte groups) of one row and marches kPdVK levels.  The five x faces of its
// cells come from 12-point windows of q and 4-point groups of q_old, ru_m, mut and msfty;
// the y faces from the rows j-3..j+3; the z face is carried up the column.  Every face is
// the same expression, in the same order, as in PdFaces.  The limiter's scales are needed
// on the interior cells only: the halo cells ids-1, ide (and rows jds-1, jde) of a periodic
// domain hold exact copies of the interior inputs (the whole step fills every halo before
// this transport), so the reference's scales there equal those of the cells they copy, and
// k_pd_tend_x4 reads those.

struct PdLane {
  const DState &s;
  const Dm &d;
  FPtr<float> f, fo, ru, rv, rom, mut;
  float ts, dt, rdx, rdy;
  int i0, j;
  Off o2;         // (i0, j) in a 2-D field
  float mutw[6];  // mut at i0-1 .. i0+4 of row j
  float msfw[6];  // msfty at i0-1 .. i0+4 of row j
  F4 mut_s, mut_n, mxw_s, mxw_c, mxw_n;  // mut rows j-1, j+1; msftx rows j-1, j, j+1

  DEVI void load2d() {
    const F4 ml = mut.ld<F4>(o2 - 4), mc = mut.ld<F4>(o2), mr = mut.ld<F4>(o2 + 4);
    const F4 sl = s.msfty.ld<F4>(o2 - 4), sc = s.msfty.ld<F4>(o2), sr = s.msfty.ld<F4>(o2 + 4);
    mutw[0] = ml.v[3];
    msfw[0] = sl.v[3];
#pragma unroll
    for (int n = 0; n < 4; n++) {
      mutw[n + 1] = mc.v[n];
      msfw[n + 1] = sc.v[n];
    }
    mutw[5] = mr.v[0];
    msfw[5] = sr.v[0];
    const int ni = d.ni;
    mut_s = mut.ld<F4>(o2 - ni);
    mut_n = mut.ld<F4>(o2 + ni);
    mxw_s = s.msftx.ld<F4>(o2 - ni);
    mxw_c = s.msftx.ld<F4>(o2);
    mxw_n = s.msftx.ld<F4>(o2 + ni);
  }
  // x faces i0..i0+4 of level k (o = (i0, k, j)): fq = high - upwind, fl = upwind
  DEVI void fx(Off o, float c1, float c2, float fq[5], float fl[5]) const {
    const F4 wl = f.ld<F4>(o - 4), wc = f.ld<F4>(o), wr = f.ld<F4>(o + 4);
    const F4 ol = fo.ld<F4>(o - 4), oc = fo.ld<F4>(o), orr = fo.ld<F4>(o + 4);
    const F4 rc = ru.ld<F4>(o), rr = ru.ld<F4>(o + 4);
    float W[12], FO[6];
#pragma unroll
    for (int n = 0; n < 4; n++) {
      W[n] = wl.v[n];
      W[n + 4] = wc.v[n];
      W[n + 8] = wr.v[n];
      FO[n + 1] = oc.v[n];
    }
    FO[0] = ol.v[3];
    FO[5] = orr.v[0];
#pragma unroll
    for (int n = 0; n < 5; n++) {
      const float dx = 2.f / (msfw[n + 1] + msfw[n]) / rdx;
      const float mu = 0.5f * ((c1 * mutw[n + 1] + c2) + (c1 * mutw[n] + c2));
      const float vel = n < 4 ? rc.v[n] : rr.v[0];
      const float cr = vel * dt / dx / mu;
      fl[n] = mu * (dx / dt) * pd_upwind(FO[n], FO[n + 1], cr);
      fq[n] = vel * pd_flux5(W[n + 1], W[n + 2], W[n + 3], W[n + 4], W[n + 5], W[n + 6], vel, ts);
      fq[n] = fq[n] - fl[n];
    }
  }
  // y faces j (lo) and j+1 (hi) of level k for the lane's four cells
  DEVI void fy(Off o, float c1, float c2, const F4 &q0, F4 &fq0, F4 &fl0, F4 &fq1, F4 &fl1) const {
    const int sj = d.sj;
    const F4 ym3 = f.ld<F4>(o - 3 * sj), ym2 = f.ld<F4>(o - 2 * sj), ym1 = f.ld<F4>(o - sj);
    const F4 yp1 = f.ld<F4>(o + sj), yp2 = f.ld<F4>(o + 2 * sj), yp3 = f.ld<F4>(o + 3 * sj);
    const F4 os = fo.ld<F4>(o - sj), oc = fo.ld<F4>(o), on = fo.ld<F4>(o + sj);
    const F4 v0 = rv.ld<F4>(o), v1 = rv.ld<F4>(o + sj);
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const float mc = mutw[n + 1];
      {
        const float dy = 2.f / (mxw_c.v[n] + mxw_s.v[n]) / rdy;
        const float mu = 0.5f * ((c1 * mc + c2) + (c1 * mut_s.v[n] + c2));
        const float vel = v0.v[n];
        const float cr = vel * dt / dy / mu;
        fl0.v[n] = mu * (dy / dt) * pd_upwind(os.v[n], oc.v[n], cr);
        fq0.v[n] = vel * pd_flux5(ym3.v[n], ym2.v[n], ym1.v[n], q0.v[n], yp1.v[n], yp2.v[n], vel, ts) - fl0.v[n];
      }
      {
        const float dy = 2.f / (mxw_n.v[n] + mxw_c.v[n]) / rdy;
        const float mu = 0.5f * ((c1 * mut_n.v[n] + c2) + (c1 * mc + c2));
        const float vel = v1.v[n];
        const float cr = vel * dt / dy / mu;
        fl1.v[n] = mu * (dy / dt) * pd_upwind(oc.v[n], on.v[n], cr);
        fq1.v[n] = vel * pd_flux5(ym2.v[n], ym1.v[n], q0.v[n], yp1.v[n], yp2.v[n], yp3.v[n], vel, ts) - fl1.v[n];
      }
    }
  }
  // z face at w level kk of the lane's cells, from q at kk-2..kk+1, q_old at kk-1, kk
  DEVI void fz(int kk, const F4 &qm2, const F4 &qm1, const F4 &q0, const F4 &qp1, const F4 &om1, const F4 &o0,
               const F4 &w, F4 &fq, F4 &fl) const {
    const int kts = d.kds, ktf = d.kde - 1;
    const bool zero = kk <= kts || kk >= d.kde, low = kk == kts + 1 || kk == ktf;
    const int kc = imax(kts + 1, imin(kk, d.kde - 1));
    const float c1 = s.c1h[X1(d, kc)], c2 = s.c2h[X1(d, kc)];
    const float dz = 2.f / (s.rdnw[X1(d, kc)] + s.rdnw[X1(d, kc - 1)]);
    const float fm = s.fnm[X1(d, kc)], fp = s.fnp[X1(d, kc)];
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const float m = mutw[n + 1];
      const float mu = 0.5f * ((c1 * m + c2) + (c1 * m + c2));
      const float vel = w.v[n];
      const float cr = vel * dt / dz / mu;
      const float l = mu * (dz / dt) * pd_upwind(om1.v[n], o0.v[n], cr);
      const float h = low ? w.v[n] * (fm * q0.v[n] + fp * qm1.v[n])
                          : vel * pd_flux3(qm2.v[n], qm1.v[n], q0.v[n], qp1.v[n], -vel, ts);
      fl.v[n] = zero ? 0.f : l;
      fq.v[n] = zero ? 0.f : h - l;
    }
  }
};

DEVI PdLane pd_lane(const KArgs &a, int i0, int j) {
  const DState &s = dstate(a);
  const int sp = a.species;
  PdLane L{s, a.d, s.q[sp], s.q_old[sp], s.ru_m, s.rv_m, s.ww_m, s.muts, a.c.time_step >= 0 ? 1.f : -1.f,
           a.dt_rk, s.rdx, s.rdy, i0, j, X2(a.d, i0, j)};
  L.load2d();
  return L;
}

// The interior cells' scales (ids..ide-1, jds..jde-1), kPdVK levels per thread.
template <typename ARGS>
__global__ void __launch_bounds__(256) k_pd_scale_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int kts = d.kds, ktf = d.kde - 1;
  const int k0 = kts + T.bz * kPdVK, k1 = imin(k0 + kPdVK - 1, ktf);
  if (i0 > d.ide - 1 || j > d.jde - 1 || k0 > ktf) return;
  const PdLane P = pd_lane(a, i0, j);
  const FPtr<float> S = a.scr_dv;
  const F4 mub = s.mub.ld<F4>(P.o2), mu_old = s.mu_1.ld<F4>(P.o2), msftx = P.mxw_c;
  const float rdx = s.rdx, rdy = s.rdy, dt = P.dt, eps = 1.e-20f;
  const int ni = d.ni;
  const Off ob = X3(d, i0, 0, j);
  auto lq = [&](const FPtr<float> &arr, int kk) { return arr.ld<F4>(ob + imax(kts, imin(kk, ktf)) * ni); };
  auto lw = [&](int kk) { return P.rom.ld<F4>(ob + imax(kts, imin(kk, d.kde)) * ni); };
  F4 qm2 = lq(P.f, k0 - 2), qm1 = lq(P.f, k0 - 1), q0 = lq(P.f, k0), qp1 = lq(P.f, k0 + 1);
  F4 om1 = lq(P.fo, k0 - 1), o0 = lq(P.fo, k0);
  F4 zq_lo, zl_lo;
  P.fz(k0, qm2, qm1, q0, qp1, om1, o0, lw(k0), zq_lo, zl_lo);
  for (int k = k0; k <= k1; k++) {
    const Off o = ob + k * ni;
    const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)], rdzw = s.rdnw[X1(d, k)];
    float xq[5], xl[5];
    P.fx(o, c1, c2, xq, xl);
    F4 yq0, yl0, yq1, yl1;
    P.fy(o, c1, c2, q0, yq0, yl0, yq1, yl1);
    const F4 qp2 = lq(P.f, k + 2), op1 = lq(P.fo, k + 1);
    F4 zq1, zl1;
    P.fz(k + 1, qm1, q0, qp1, qp2, o0, op1, lw(k + 1), zq1, zl1);
    F4 sc;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const float msfty = P.msfw[n + 1];
      const float ph_low = ((c1 * mub.v[n] + c2) + (c1 * mu_old.v[n])) * o0.v[n] -
                           dt * (msftx.v[n] * msfty * (rdx * (xl[n + 1] - xl[n]) + rdy * (yl1.v[n] - yl0.v[n])) +
                                 msfty * rdzw * (zl1.v[n] - zl_lo.v[n]));
      const float flux_out = dt * ((msftx.v[n] * msfty) * (rdx * (fmax_r(0.f, xq[n + 1]) - fmin_r(0.f, xq[n])) +
                                                           rdy * (fmax_r(0.f, yq1.v[n]) - fmin_r(0.f, yq0.v[n]))) +
                                   msfty * rdzw * (fmin_r(0.f, zq1.v[n]) - fmax_r(0.f, zq_lo.v[n])));
      sc.v[n] = flux_out > ph_low ? fmax_r(0.f, ph_low / (flux_out + eps)) : 1.f;
    }
    st4m(S, o, sc, i0, d.ids, d.ide - 1);
    qm2 = qm1;
    qm1 = q0;
    q0 = qp1;
    qp1 = qp2;
    om1 = o0;
    o0 = op1;
    zq_lo = zq1;
    zl_lo = zl1;
  }
}

// advect_scalar_pd's tendency on a doubly periodic domain of the whole step (rk_scalar_tend's
// form: from zero, zero at the points i = ide, j = jde, k = kde), kPdVK levels per thread.
template <typename ARGS>
__global__ void __launch_bounds__(256) k_pd_tend_x4(ARGS args) {
  Tile3 T;
  const KArgs &a = kargs(args, T);
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int i0 = vec_ia0(d) + kVecW * (T.bx * 64 + (int)threadIdx.x);
  const int j = d.jds + (int)(T.by * blockDim.y + threadIdx.y);
  const int kts = d.kds, kte = d.kde, ktf = kte - 1;
  const int k0 = kts + T.bz * kPdVK, k1 = imin(k0 + kPdVK - 1, kte);
  if (i0 > d.ide || j > d.jde || k0 > kte) return;
  const FPtr<float> S = a.scr_dv, tend = s.advect_tend;
  const int ni = d.ni, nx = d.ide - d.ids, ny = d.jde - d.jds;
  const Off ob = X3(d, i0, 0, j);
  const F4 zero4{{0.f, 0.f, 0.f, 0.f}};
  if (j == d.jde || k0 > ktf) {  // row jde, or the level kde alone: zeros only
    for (int k = k0; k <= k1; k++) st4m(tend, ob + k * ni, zero4, i0, d.ids, d.ide);
    return;
  }
  const PdLane P = pd_lane(a, i0, j);
  const F4 msftx = P.mxw_c;
  const float rdx = s.rdx, rdy = s.rdy;
  // the scale rows j-1 and j+1, wrapped into the interior rows
  const int js = j - 1 < d.jds ? j - 1 + ny : j - 1, jn = j + 1 > d.jde - 1 ? j + 1 - ny : j + 1;
  const Off obs = X3(d, i0, 0, js), obn = X3(d, i0, 0, jn);
  // the scales of the cells i0-1 .. i0+4, each wrapped into the interior columns (a lane
  // whose cells all lie inside reads its own group and the two neighbours as they are)
  const bool inner = i0 - 1 >= d.ids && i0 + 4 <= d.ide - 1;
  int ix[6];
#pragma unroll
  for (int m = 0; m < 6; m++) {
    const int i = i0 - 1 + m;
    ix[m] = i < d.ids ? i + nx : i > d.ide - 1 ? i - nx : i;
  }
  // cells i0..i0+3 past ide-1 store a zero; their own scales are not used
  auto lq = [&](const FPtr<float> &arr, int kk) { return arr.ld<F4>(ob + imax(kts, imin(kk, ktf)) * ni); };
  auto lw = [&](int kk) { return P.rom.ld<F4>(ob + imax(kts, imin(kk, kte)) * ni); };
  auto lim = [](float fq, float s_up, float s_dn) { return fq > 0.f ? s_up * fq : fq < 0.f ? s_dn * fq : fq; };
  const int kz = imin(k0, ktf);
  F4 qm2 = lq(P.f, kz - 2), qm1 = lq(P.f, kz - 1), q0 = lq(P.f, kz), qp1 = lq(P.f, kz + 1);
  F4 om1 = lq(P.fo, kz - 1), o0 = lq(P.fo, kz);
  F4 s0 = S.ld<F4>(ob + kz * ni);
  F4 zq_lo, zl_lo;
  P.fz(kz, qm2, qm1, q0, qp1, om1, o0, lw(kz), zq_lo, zl_lo);
  if (kz > kts) {
    const F4 sm = S.ld<F4>(ob + (kz - 1) * ni);
#pragma unroll
    for (int n = 0; n < 4; n++) zq_lo.v[n] = lim(zq_lo.v[n], s0.v[n], sm.v[n]);
  }
  for (int k = kz; k <= imin(k1, ktf); k++) {
    const Off o = ob + k * ni;
    const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)], rdzw = s.rdnw[X1(d, k)];
    const F4 qp2 = lq(P.f, k + 2), op1 = lq(P.fo, k + 1);
    F4 zq1, zl1;
    P.fz(k + 1, qm1, q0, qp1, qp2, o0, op1, lw(k + 1), zq1, zl1);
    const F4 sp = S.ld<F4>(ob + imin(k + 1, ktf) * ni);
    float xq[5], xl[5];
    P.fx(o, c1, c2, xq, xl);
    F4 yq0, yl0, yq1, yl1;
    P.fy(o, c1, c2, q0, yq0, yl0, yq1, yl1);
    const F4 ss = S.ld<F4>(obs + k * ni), sn = S.ld<F4>(obn + k * ni);
    float SX[6];
    if (inner) {
      SX[0] = S[o - 1];
      SX[5] = S[o + 4];
#pragma unroll
      for (int n = 0; n < 4; n++) SX[n + 1] = s0.v[n];
    } else {
#pragma unroll
      for (int m = 0; m < 6; m++) SX[m] = S[X3(d, ix[m], k, j)];
    }
    F4 t;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      float tt = 0.f;
      const float zhi = k + 1 <= ktf ? lim(zq1.v[n], sp.v[n], s0.v[n]) : zq1.v[n];
      tt = tt - rdzw * (zhi - zq_lo.v[n] + zl1.v[n] - zl_lo.v[n]);
      zq1.v[n] = zhi;
      const float x0 = lim(xq[n], SX[n], SX[n + 1]), x1 = lim(xq[n + 1], SX[n + 1], SX[n + 2]);
      tt = tt - msftx.v[n] * (rdx * (x1 - x0 + xl[n + 1] - xl[n]));
      const float y0 = lim(yq0.v[n], ss.v[n], s0.v[n]), y1 = lim(yq1.v[n], s0.v[n], sn.v[n]);
      tt = tt - msftx.v[n] * (rdy * (y1 - y0 + yl1.v[n] - yl0.v[n]));
      t.v[n] = i0 + n <= d.ide - 1 ? tt : 0.f;
    }
    st4m(tend, o, t, i0, d.ids, d.ide);
    qm2 = qm1;
    qm1 = q0;
    q0 = qp1;
    qp1 = qp2;
    om1 = o0;
    o0 = op1;
    zq_lo = zq1;
    zl_lo = zl1;
    s0 = sp;
  }
  if (k1 == kte) st4m(tend, ob + kte * ni, zero4, i0, d.ids, d.ide);
}

WRFMI_INST(k_pd_scale)
WRFMI_INST(k_pd_tend)
WRFMI_INST(k_pd_scale_x4)
WRFMI_INST(k_pd_tend_x4)
WRFMI_INST(k_update_scalar_pd)

}  // namespace wrfmi
