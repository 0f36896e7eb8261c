// Specified lateral boundaries (the boundary-zone routines solve_em calls when
// config_flags%specified).  k_lbc: one thread per point (i, j, k) of the frame of rows
// along the edges that the mode touches (of the whole extent for the relaxation into the
// save arrays), up to six fields per launch (blockIdx.z); k_lbc_ph: one workgroup per
// spec-zone column.  Modes:
//   LBC_RELAX      relax_bdytend (share/module_bc.F:1221-1427) of ru, rv, ph or t (mass
//                  weighted as relax_bdy_dry's mass_weight does, dyn_em/module_bc_em.F:
//                  1716-1745), mu or a species added to the target: the *_save arrays
//                  (after LBC_ZERO), mu_save for mu (which k_tend_t adds to mu_tend), q_tend;
//                  with `whole`, written over the whole extent, zero outside the rows
//   LBC_ZERO       zero_tend of the save arrays over the whole extent, every level
//   LBC_SPEC_TEND  spec_bdytend (1430-1549): the spec zone's tendency is the boundary one
//   LBC_UPDATE     spec_bdyupdate (1955-2064): field += dt * tendency in the spec zone
//   LBC_PH_PRHO    spec_bdyupdate_ph (dyn_em/module_bc_em.F:17-157) on the spec-zone
//                  columns, then calc_p_rho there (k_ac_wsolve leaves these columns alone)
//                  (k_lbc_ph, which also takes zero_grad_bdy of w)
//   LBC_ZERO_GRAD  zero_grad_bdy of w (share/module_bc.F:2219-2332)
//   LBC_FLOW_DEP   flow_dep_bdy of a species (2335-2456)
// Zone membership follows the reference's loops with the tile equal to the domain: side
// order ys, ye, xs, xe (the x sides only when not periodic in x); a point of a y side row at
// distance b lies b or more points from the x ends, a point of an x side column at distance
// b lies b+1 or more rows from the y ends, so each point has at most one side unless the
// domain is narrower than two zones, where the sides act in that order.
#include "kernels.h"

namespace wrfmi {

// the variable's zone geometry (ibs, ibe, jbs, jbe, last level)
struct LbcGeo {
  int ibs, ibe, jbs, jbe, ktf;
};
DEVI LbcGeo lbc_geo(const Dm &d, int var) {
  LbcGeo z;
  z.ibs = d.ids; z.ibe = d.ide - 1; z.jbs = d.jds; z.jbe = d.jde - 1; z.ktf = d.kde - 1;
  if (var == 'u') z.ibe = d.ide;
  if (var == 'v') z.jbe = d.jde;
  if (var == 'm') z.ktf = d.kds;
  if (var == 'h' || var == 'w') z.ktf = d.kde;
  return z;
}

// side (0..3 = xs, xe, ys, ye) whose rows at distances [lo, hi] hold (i, j), after side
// `after` (-1 for the first), with the distance in *bd; -1 if none
DEVI int lbc_side(const LbcGeo &z, bool px, int i, int j, int lo, int hi, int after, int *bd) {
  const int order[4] = {2, 3, 0, 1};
  int start = 0;
  if (after >= 0)
    for (int n = 0; n < 4; n++)
      if (order[n] == after) start = n + 1;
  for (int n = start; n < 4; n++) {
    const int side = order[n];
    if (side == 2 || side == 3) {
      const int b = side == 2 ? j - z.jbs : z.jbe - j;
      const int bl = px ? 0 : b;
      if (b >= lo && b <= hi && i >= z.ibs + bl && i <= z.ibe - bl) { *bd = b; return side; }
    } else if (!px) {
      const int b = side == 0 ? i - z.ibs : z.ibe - i;
      if (b >= lo && b <= hi && j >= z.jbs + b + 1 && j <= z.jbe - b - 1) { *bd = b; return side; }
    }
  }
  return -1;
}

// boundary array element at (ij, k, b) of side `side` (b 1-based)
DEVI float lbc_at(const Dm &d, const float *arr, int side, int ij, int k, int b, bool twod) {
  const int nkb = twod ? 1 : d.kde - d.kds + 1;
  const int kk = twod ? 0 : k - d.kds;
  const long n = side < 2 ? (long)(d.jme - d.jms + 1) : (long)d.ni;
  const int o = side < 2 ? ij - d.jms : ij - d.ims;
  return arr[((long)(b - 1) * nkb + kk) * n + o];
}

__global__ void __launch_bounds__(256) k_lbc(LbcArgs L) {
  const KArgs &a = L.a;
  const Dm &d = a.d;
  const DState &s = dstate(a);
  // whole: blockIdx.z = job * nk + level, (64, 4) threads over (i, j); frame: blockIdx.y =
  // level, blockIdx.z = job, threads over the frame's points
  const int nk = d.kde - d.kds + 1;
  const int jb = L.whole && L.mode != LBC_ZERO ? (int)blockIdx.z / nk : (int)blockIdx.z;
  const LbcJob &J = L.job[jb];
  const LbcGeo z = lbc_geo(d, J.var);
  const bool px = a.c.periodic_x, twod = J.var == 'm';
  const int kb = d.kds + (L.mode == LBC_ZERO ? 0 : L.whole ? (int)blockIdx.z - jb * nk : (int)blockIdx.y);
  int i, j;
  if (L.whole) {
    // a whole-domain target is written over zero_tend's extent (ids..ide, jds..jde, every
    // level); everything else stays inside the variable's staggering
    i = d.ids + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    j = d.jds + (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (i > d.ide || j > d.jde) return;
  } else {
    const LbcRing r = lbc_ring(z.ibs, z.ibe, z.jbs, z.jbe, px, L.width);
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= (int)r.count) return;
    const int ny = (r.rb + r.rt) * r.ni;
    if (t < ny) {
      const int row = t / r.ni;
      i = z.ibs + (t - row * r.ni);
      j = row < r.rb ? z.jbs + row : z.jbe - (row - r.rb);
    } else {
      const int nc = r.cl + r.cr, row = (t - ny) / nc, c = (t - ny) - row * nc;
      j = z.jbs + r.rb + row;
      i = c < r.cl ? z.ibs + c : z.ibe - (c - r.cl);
    }
  }
  if (L.mode == LBC_ZERO) {  // zero_tend of a save array over the whole extent, a column per thread
    FPtr<float> D(J.dst);
    if (twod) D[X2(d, i, j)] = 0.f;
    else
      for (int k = d.kds; k <= d.kde; k++) D[X3(d, i, k, j)] = 0.f;
    return;
  }
  const bool inside = i <= z.ibe && j <= z.jbe;
  const int sz = a.c.spec_zone, rz = a.c.relax_zone;
  const Off o2 = X2(d, i, j);
  // element offsets are byte offsets on a base pointer (FPtr)
  const FPtr<float> SRC((float *)J.src), TEND((float *)J.tend);
  FPtr<float> DST(J.dst);
  auto at = [&](const FPtr<float> &p, int ii, int kk, int jj) -> float {
    return twod ? p[X2(d, ii, jj)] : p[X3(d, ii, kk, jj)];
  };
  if (L.mode == LBC_RELAX) {
    const int kmax = L.whole && !twod ? d.kde : z.ktf;
    for (int k = kb; k <= imin(kb, kmax); k++) {  // this thread's level
      // the relaxed field at a point: ru, rv, mu, or the mass-weighted ph, t, species
      auto fld = [&](int ii, int kk, int jj) -> float {
        const float f = at(SRC, ii, kk, jj);
        if (!J.massw) return f;
        const float c1 = J.var == 'h' ? s.c1f[X1(d, kk)] : s.c1h[X1(d, kk)];
        const float c2 = J.var == 'h' ? s.c2f[X1(d, kk)] : s.c2h[X1(d, kk)];
        return f * (c1 * s.mut[X2(d, ii, jj)] + c2);
      };
      float tv = L.whole ? 0.f : at(DST, i, k, j);
      bool any = false;
      int b, side = -1;
      // only points within relax_zone of an edge can lie in a relaxation row
      const bool near = inside && (j - z.jbs < rz || z.jbe - j < rz || (!px && (i - z.ibs < rz || z.ibe - i < rz)));
      while (near && k <= z.ktf && J.b[0] && (side = lbc_side(z, px, i, j, sz, rz - 1, side, &b)) >= 0) {
        auto fls = [&](int ij, int bb, int fi, int fj) -> float {
          return lbc_at(d, J.b[side], side, ij, k, bb, twod) + L.dtbc * lbc_at(d, J.bt[side], side, ij, k, bb, twod) -
                 fld(fi, k, fj);
        };
        float f0, f1, f2, f3, f4;
        if (side >= 2) {
          const int im1 = imax(i - 1, z.ibs), ip1 = imin(i + 1, z.ibe), dj = side == 2 ? 1 : -1;
          f0 = fls(i, b + 1, i, j);
          f1 = fls(im1, b + 1, im1, j);
          f2 = fls(ip1, b + 1, ip1, j);
          f3 = fls(i, b, i, j - dj);
          f4 = fls(i, b + 2, i, j + dj);
        } else {
          const int di = side == 0 ? 1 : -1;
          f0 = fls(j, b + 1, i, j);
          f1 = fls(j - 1, b + 1, i, j - 1);
          f2 = fls(j + 1, b + 1, i, j + 1);
          f3 = fls(j, b, i - di, j);
          f4 = fls(j, b + 2, i + di, j);
        }
        tv = tv + L.fcx[b] * f0 - L.gcx[b] * (f1 + f2 + f3 + f4 - 4.f * f0);
        any = true;
      }
      if (any || L.whole) {
        if (twod) DST[o2] = tv;
        else DST[X3(d, i, k, j)] = tv;
      }
    }
    return;
  }
  const int k = kb;
  if (!inside || k > z.ktf) return;
  if (L.mode == LBC_FLOW_DEP || L.mode == LBC_ZERO_GRAD) {
    int b, side = -1, last = -1;
    while ((side = lbc_side(z, px, i, j, 0, sz - 1, side, &b)) >= 0) last = side;
    if (last < 0) return;
    int si, sj;
    if (last >= 2) {
      si = imin(imax(i, z.ibs + sz), z.ibe - sz);
      if (px) si = i;
      sj = last == 2 ? z.jbs + sz : z.jbe - sz;
    } else {
      sj = imin(imax(j, z.jbs + sz), z.jbe - sz);
      si = last == 0 ? z.ibs + sz : z.ibe - sz;
    }
    const float inner = DST[X3(d, si, k, sj)];
    float v = inner;
    if (L.mode == LBC_FLOW_DEP) {
      const FPtr<float> u = s.ru_m, vv = s.rv_m;
      const bool out = last == 2 ? vv[X3(d, i, k, j)] < 0.f : last == 3 ? vv[X3(d, i, k, j + 1)] > 0.f
                     : last == 0 ? u[X3(d, i, k, j)] < 0.f : u[X3(d, i + 1, k, j)] > 0.f;
      v = out ? inner : 0.f;
    }
    DST[X3(d, i, k, j)] = v;
    return;
  }
  // LBC_SPEC_TEND, LBC_UPDATE: the spec zone (later sides overwrite earlier)
  int b = 0, side = -1, last = -1, lb = 0;
  while ((side = lbc_side(z, px, i, j, 0, sz - 1, side, &b)) >= 0) {
    last = side;
    lb = b;
    if (L.mode == LBC_UPDATE) {  // each side's loop adds dt * tendency once
      if (twod) DST[o2] = DST[o2] + L.dt * TEND[o2];
      else DST[X3(d, i, k, j)] = DST[X3(d, i, k, j)] + L.dt * TEND[X3(d, i, k, j)];
    }
  }
  if (last < 0 || L.mode == LBC_UPDATE) return;
  const float v = lbc_at(d, J.bt[last], last, last < 2 ? j : i, k, lb + 1, twod);
  if (twod) DST[o2] = v;
  else DST[X3(d, i, k, j)] = v;
}

// calc_coef_w's a, alpha and gamma (module_small_step_em.F:570-652) of column (i, j) at the
// width CT, in k_coef_w's expressions, over the workgroup's lanes: every level's
// coefficients at once, the forward recurrence on lane 0
template <typename CT>
DEVI void coef_w_column(const KArgs &a, const DState &s, int i, int j, CT *sa, CT *sb, CT *sc) {
  const Dm &d = a.d;
  const int kde = d.kde;
  const CPtr<float> c1h = s.c1h, c2h = s.c2h, c1f = s.c1f, c2f = s.c2f;
  const FPtr<CT> A((CT *)s.a), AL((CT *)s.alpha), GA((CT *)s.gamma);
  const CT lid = a.c.top_lid ? CT(0) : CT(1);
  const CT epssm = (CT)a.c.epssm;
  CT cof = (CT(.5) * (CT)a.dts * (CT)K_G * (CT(1) + epssm));
  cof = cof * cof;
  const CT MUT = (CT)s.mut[X2(d, i, j)];
  auto cH = [&](int k) -> CT { return (CT)c1h[X1(d, k)] * MUT + (CT)c2h[X1(d, k)]; };
  auto cF = [&](int k) -> CT { return (CT)c1f[X1(d, k)] * MUT + (CT)c2f[X1(d, k)]; };
  auto rdnw = [&](int k) -> CT { return (CT)s.rdnw[X1(d, k)]; };
  auto rdn = [&](int k) -> CT { return (CT)s.rdn[X1(d, k)]; };
  auto C2A = [&](int k) -> CT { return (CT)s.c2a[X3(d, i, k, j)]; };
  auto CQW = [&](int k) -> CT { return (CT)s.cqw[X3(d, i, k, j)]; };
  for (int k = 2 + (int)threadIdx.x; k <= kde; k += blockDim.x) {
    if (k == kde) {
      const int km = kde - 1;
      const CT top = -CT(2) * cof * (rdnw(kde - 1) * rdnw(kde - 1)) * C2A(kde - 1) * lid / (cH(km) * cF(km));
      A[X3(d, i, kde, j)] = top;
      sa[k] = top;
      sb[k] = CT(1) + CT(2) * cof * (rdnw(kde - 1) * rdnw(kde - 1)) * C2A(kde - 1) / (cH(k - 1) * cF(k));
    } else {
      const CT ak = k >= 3 ? -CQW(k) * cof * rdn(k) * rdnw(k - 1) * C2A(k - 1) / (cH(k - 1) * cF(k - 1)) : CT(0);
      A[X3(d, i, k, j)] = ak;  // A(2) = 0
      sa[k] = ak;
      sb[k] = CT(1) + CQW(k) * cof * rdn(k) *
                          (rdnw(k) * C2A(k) / (cH(k) * cF(k)) + rdnw(k - 1) * C2A(k - 1) / (cH(k - 1) * cF(k)));
      sc[k] = -CQW(k) * cof * rdn(k) * rdnw(k) * C2A(k) / (cH(k) * cF(k + 1));
    }
  }
  if (threadIdx.x == 0) GA[X3(d, i, 1, j)] = CT(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    CT gam = CT(0);
    for (int k = 2; k <= kde - 1; k++) {
      const CT alpha = CT(1) / (sb[k] - sa[k] * gam);
      sb[k] = alpha;
      gam = sc[k] * alpha;
      sc[k] = gam;
    }
    const CT alpha = CT(1) / (sb[kde] - sa[kde] * gam);
    sb[kde] = alpha;
    sc[kde] = CT(0) * alpha;
  }
  __syncthreads();
  for (int k = 2 + (int)threadIdx.x; k <= kde; k += blockDim.x) {
    AL[X3(d, i, k, j)] = sb[k];
    GA[X3(d, i, k, j)] = sc[k];
  }
}

// LBC_PH_PRHO (mass staggering, spec_zone 1): the spec-zone columns advance_w leaves out --
// calc_coef_w, spec_bdyupdate_ph, calc_p_rho and (zero_grad_w) zero_grad_bdy of w
__global__ void __launch_bounds__(128) k_lbc_ph(LbcArgs L) {
  const KArgs &a = L.a;
  const Dm &d = a.d;
  const DState &s = dstate(a);
  const LbcGeo z = lbc_geo(d, L.job[0].var);
  const bool px = a.c.periodic_x;
  const int sz = a.c.spec_zone;
  const LbcRing r = lbc_ring(z.ibs, z.ibe, z.jbs, z.jbe, px, L.width);
  const long t = blockIdx.x;
  if (t >= r.count) return;
  int i, j;
  const long ny = (long)(r.rb + r.rt) * r.ni;
  if (t < ny) {
    const int row = (int)(t / r.ni);
    i = z.ibs + (int)(t % r.ni);
    j = row < r.rb ? z.jbs + row : z.jbe - (row - r.rb);
  } else {
    const int nc = r.cl + r.cr, row = (int)((t - ny) / nc), c = (int)((t - ny) % nc);
    j = z.jbs + r.rb + row;
    i = c < r.cl ? z.ibs + c : z.ibe - (c - r.cl);
  }
  int b, side = -1, last = -1;
  while ((side = lbc_side(z, px, i, j, 0, sz - 1, side, &b)) >= 0) last = side;
  if (last < 0) return;  // uniform over the workgroup
  __shared__ double sa[kLbcMaxK + 1], sb[kLbcMaxK + 1], sc[kLbcMaxK + 1];
  __shared__ float sph[kLbcMaxK + 1];
  if (a.c.precision_mode == 1) coef_w_column<double>(a, s, i, j, sa, sb, sc);
  else coef_w_column<float>(a, s, i, j, (float *)sa, (float *)sb, (float *)sc);
  const Off o2 = X2(d, i, j);
  const float dt = L.dt, muts = s.muts[o2], mu_tend = s.mu_tend[o2];
  for (int k = d.kds + (int)threadIdx.x; k <= d.kde; k += blockDim.x) {
    const Off o = X3(d, i, k, j);
    const float c1 = s.c1f[X1(d, k)], c2 = s.c2f[X1(d, k)];
    const float mu_old = muts - dt * mu_tend;
    const float ph = s.ph_2[o] * (c1 * mu_old + c2) / (c1 * muts + c2) + dt * s.ph_tend[o] / (c1 * muts + c2) +
                     s.ph_save[o] * ((c1 * mu_old + c2) / (c1 * muts + c2) - 1.f);
    s.ph_2[o] = ph;
    sph[k - d.kds] = ph;
  }
  __syncthreads();
  // calc_p_rho (module_small_step_em.F:438-568) for the column, as k_calc_p_rho forms it
  const float mu = s.mu_2[o2];
  for (int k = d.kds + (int)threadIdx.x; k <= d.kde - 1; k += blockDim.x) {
    const Off o = X3(d, i, k, j);
    const float c1 = s.c1h[X1(d, k)], c2 = s.c2h[X1(d, k)];
    const float alt = s.alt[o];
    const float al = -1.f / (c1 * muts + c2) * (alt * (c1 * mu) + s.rdnw[X1(d, k)] * (sph[k + 1 - d.kds] - sph[k - d.kds]));
    s.al[o] = al;
    const float t1 = s.t_save[o];
    float pp = s.c2a[o] * (alt * (s.t_2[o] - (c1 * mu) * t1) / ((c1 * muts + c2) * (K_T0 + t1)) - al);
    const float ptmp = pp;
    pp = pp + a.c.smdiv * (pp - s.pm1[o]);
    s.p[o] = pp;
    s.pm1[o] = ptmp;
  }
  if (!L.zero_grad_w) return;
  // zero_grad_bdy of w (share/module_bc.F:2219-2332): the frame takes the first inner row
  int si, sj;
  if (last >= 2) {
    si = px ? i : imin(imax(i, z.ibs + sz), z.ibe - sz);
    sj = last == 2 ? z.jbs + sz : z.jbe - sz;
  } else {
    sj = imin(imax(j, z.jbs + sz), z.jbe - sz);
    si = last == 0 ? z.ibs + sz : z.ibe - sz;
  }
  FPtr<float> W(s.w_2);
  for (int k = d.kds + (int)threadIdx.x; k <= d.kde; k += blockDim.x) W[X3(d, i, k, j)] = W[X3(d, si, k, sj)];
}

}  // namespace wrfmi
