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
      // radiation terms of open boundaries (7255-7326), before the flux divergence
      if (!per_x && i == d.ids) {
        const float ub = fmin_r(0.5f * (ru[o] + ru[o + 1]), 0.f);
        t = t - rdx * (ub * (fo[o + 1] - fo[o]) + f[o] * (ru[o + 1] - ru[o]));
      }
      if (!per_x && i == d.ide - 1) {
        const Off oe = X3(d, d.ide, k, j);
        const float ub = fmax_r(0.5f * (ru[oe - 1] + ru[oe]), 0.f);
        t = t - rdx * (ub * (fo[o] - fo[o - 1]) + f[o] * (ru[oe] - ru[oe - 1]));
      }
      if (!per_y && j == d.jds) {
        const float vb = fmin_r(0.5f * (rv[o] + rv[o + d.sj]), 0.f);
        t = t - rdy * (vb * (fo[o + d.sj] - fo[o]) + f[o] * (rv[o + d.sj] - rv[o]));
      }
      if (!per_y && j == d.jde - 1) {
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

WRFMI_INST(k_pd_scale)
WRFMI_INST(k_pd_tend)
WRFMI_INST(k_update_scalar_pd)

}  // namespace wrfmi
