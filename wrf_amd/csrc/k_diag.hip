// Domain-total diagnostics of one domain's state, in FP64 (no counterpart routine in the
// reference, whose budgets are computed off line from history output).
//
// Per mass column (i, j), with layer dry-air mass per unit area
//   dm_k = (c1h(k) * mut + c2h(k)) * (-dnw(k)) / g,   mut = mub + mu_2
// (module_big_step_utilities_em.F's column mass; dnw < 0 upward) and cell area
// dx * dy / (msftx * msfty):
//   dry mass       sum dm_k                                                  [kg]
//   theta budget   sum dm_k * (t_2 + t0)                                     [kg K]
//   kinetic        sum dm_k * (ubar^2 + vbar^2 + wbar^2) / 2                 [J]
//   internal       sum dm_k * cv * T,  T = theta_d * ((p + pb) / p0)^(rd/cp) [J]
//   potential      sum dm_k * (phi_k + phi_k+1) / 2,  phi = ph_2 + phb       [J]
//   total          kinetic + internal + potential
// with ubar, vbar, wbar the two-point averages to the mass point and theta_d the dry
// potential temperature (theta / (1 + rv/rd qv) when the state carries moist theta).
// Everything is accumulated in FP64 except the column sums of the theta budget, which are
// FP32 with compensated (two-sum) accumulation, as the model's column budgets are; each
// column's total is then added in FP64.  The block and domain reductions run in a fixed
// order, so the results do not depend on scheduling.
#include "kernels.h"

namespace wrfmi {

__global__ void __launch_bounds__(256) k_domain_totals(KArgs a, double *part) {
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const int nx = d.ide - d.ids, ny = d.jde - d.jds;
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  double v[kNTotals] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (t < nx * ny) {
    const int i = d.ids + t % nx, j = d.jds + t / nx;
    const Off o2 = X2(d, i, j);
    // the model constants as the -r8 build holds them (share/module_model_constants.F:17-47)
    const double g = 9.81, rd = 287.0, cp = 7.0 * 287.0 / 2.0, cv = cp - rd, t0 = 300.0, p0 = 1.0e5;
    const double area = (double)s.dx * (double)s.dy / ((double)s.msftx[o2] * (double)s.msfty[o2]);
    const double mut = (double)s.mub[o2] + (double)s.mu_2[o2];
    const float mutf = s.mub[o2] + s.mu_2[o2];
    const bool moist = a.c.use_theta_m && s.n_moist >= 1;
    float th_s = 0.f, th_c = 0.f;  // compensated FP32 column sum of the theta budget
    double mass = 0.0, ke = 0.0, ie = 0.0, pe = 0.0;
    for (int k = d.kds; k <= d.kde - 1; k++) {
      const Off o = X3(d, i, k, j);
      const double dm = ((double)s.c1h[X1(d, k)] * mut + (double)s.c2h[X1(d, k)]) * (-(double)s.dnw[X1(d, k)]) / g;
      const float dmf = (s.c1h[X1(d, k)] * mutf + s.c2h[X1(d, k)]) * (-s.dnw[X1(d, k)]) / K_G;
      two_sum_acc(th_s, th_c, dmf * (s.t_2[o] + K_T0));
      mass += dm;
      const double ub = 0.5 * ((double)s.u_2[o] + (double)s.u_2[o + 1]);
      const double vb = 0.5 * ((double)s.v_2[o] + (double)s.v_2[o + d.sj]);
      const double wb = 0.5 * ((double)s.w_2[o] + (double)s.w_2[o + d.ni]);
      ke += dm * 0.5 * (ub * ub + vb * vb + wb * wb);
      double theta = (double)s.t_2[o] + t0;
      if (moist) theta = theta / (1.0 + (461.6 / 287.0) * (double)s.q[0][o]);
      const double temp = theta * pow(((double)s.p[o] + (double)s.pb[o]) / p0, rd / cp);
      ie += dm * cv * temp;
      const double phi_lo = (double)s.ph_2[o] + (double)s.phb[o];
      const double phi_hi = (double)s.ph_2[o + d.ni] + (double)s.phb[o + d.ni];
      pe += dm * 0.5 * (phi_lo + phi_hi);
    }
    v[0] = mass * area;
    v[1] = ((double)th_s + (double)th_c) * area;
    v[2] = ke * area;
    v[3] = ie * area;
    v[4] = pe * area;
    v[5] = v[2] + v[3] + v[4];
  }
  // block sum in a fixed tree order
  __shared__ double red[kNTotals][256];
  const int l = (int)threadIdx.x;
  for (int q = 0; q < kNTotals; q++) red[q][l] = v[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (l < w)
      for (int q = 0; q < kNTotals; q++) red[q][l] += red[q][l + w];
    __syncthreads();
  }
  if (l < kNTotals) part[(size_t)blockIdx.x * kNTotals + l] = red[l][0];
}

// the partial sums of nblocks blocks, summed in a fixed order by one workgroup
__global__ void __launch_bounds__(256) k_domain_totals_final(const double *part, int nblocks, double *out) {
  __shared__ double red[kNTotals][256];
  const int l = (int)threadIdx.x;
  for (int q = 0; q < kNTotals; q++) {
    double acc = 0.0;
    for (int b = l; b < nblocks; b += 256) acc += part[(size_t)b * kNTotals + q];
    red[q][l] = acc;
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (l < w)
      for (int q = 0; q < kNTotals; q++) red[q][l] += red[q][l + w];
    __syncthreads();
  }
  if (l < kNTotals) out[l] = red[l][0];
}

}  // namespace wrfmi
