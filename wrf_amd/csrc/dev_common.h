// Device-side helpers shared by the gfx950 kernels of the WRF-ARW dynamics step.
//
// Layout: every 3-D field is the reference's Fortran array (ims:ime, kms:kme, jms:jme),
// i fastest.  Offsets are 32-bit byte offsets (the host rejects fields of 2^29 elements or more).
// All kernels are compiled with -ffp-contract=off and written in the reference's
// floating-point operation order, so that FP32 results match a reference-order
// IEEE FP32 evaluation bit for bit (transcendentals aside).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "wrfmi.h"

namespace wrfmi {

struct Dm {
  int ids, ide, jds, jde, kds, kde;
  int ims, ime, jms, jme, kms, kme;
  int ni, nk, sj;  // memory extents in i and k, stride of j (= ni*nk)
};

// Element offsets are carried as 32-bit byte offsets (Off) and added to the uniform base
// pointer of each array (FPtr), so that every access is one SGPR-base + VGPR-offset load
// with no 64-bit address arithmetic.  The host keeps fields below 2^29 elements.
struct Off {
  unsigned b;  // byte offset of a 4-byte element
};
#define HDI __host__ __device__ __forceinline__
HDI Off operator+(Off o, int n) { return Off{o.b + 4u * (unsigned)n}; }
HDI Off operator-(Off o, int n) { return Off{o.b - 4u * (unsigned)n}; }
// Every array lives in device global memory (address space 1); saying so lets the compiler
// use global (not flat) loads and stores with an SGPR base.
#ifndef WRFMI_GLOBAL_AS
#define WRFMI_GLOBAL_AS __attribute__((address_space(1)))
#endif
template <typename T>
struct FPtr {
  T *p;
  HDI FPtr() : p(nullptr) {}
  HDI FPtr(T *q) : p(q) {}
  // element at Off o; 8-byte elements (the FP64 tridiagonal coefficients) sit at twice the
  // byte offset of the 4-byte element with the same index
  HDI T &operator[](Off o) const {
    WRFMI_GLOBAL_AS char *g = (WRFMI_GLOBAL_AS char *)(char *)p;
    return *(T *)(WRFMI_GLOBAL_AS T *)(g + (sizeof(T) == 8 ? 2u * o.b : o.b));
  }
  HDI T &operator[](int i) const { return *(T *)((WRFMI_GLOBAL_AS T *)p + i); }
  // 16-byte access of four consecutive i points at Off o (o.b must be a multiple of 16)
  template <typename V>
  HDI V ld(Off o) const {
    return *(const V *)(const WRFMI_GLOBAL_AS V *)((WRFMI_GLOBAL_AS char *)(char *)p + o.b);
  }
  template <typename V>
  HDI void st(Off o, const V &v) const {
    *(V *)(WRFMI_GLOBAL_AS V *)((WRFMI_GLOBAL_AS char *)(char *)p + o.b) = v;
  }
};
// Four consecutive i points of one lane ("vector" kernels: 16-byte loads and stores)
struct alignas(16) F4 {
  float v[4];
};

// Arrays that no kernel writes (the 1-D level coefficients).  Loads through the constant
// address space become scalar loads when the index is uniform across the wave; those
// complete on the scalar counter, so waiting for one does not also wait for the vector
// loads in flight (vector loads retire in order: a wait for the newest waits for all).
#ifndef WRFMI_CONST_AS
#define WRFMI_CONST_AS __attribute__((address_space(4)))
#endif
template <typename T>
struct CPtr {
  const T *p;
  HDI CPtr() : p(nullptr) {}
  HDI CPtr(const T *q) : p(q) {}
  HDI T operator[](Off o) const {
    const WRFMI_CONST_AS char *g = (const WRFMI_CONST_AS char *)(const char *)p;
    return *(const WRFMI_CONST_AS T *)(g + o.b);
  }
  HDI T operator[](int i) const { return ((const WRFMI_CONST_AS T *)p)[i]; }
};

// Device view of wrfmi_state: same layout, arrays as FPtr (1-D level arrays as CPtr)
struct DState {
#define WRFMI_DPTR(n) FPtr<float> n;
#define WRFMI_CPTR(n) CPtr<float> n;
  WRFMI_FIELDS_3D(WRFMI_DPTR)
  WRFMI_FIELDS_2D(WRFMI_DPTR)
  WRFMI_FIELDS_1D(WRFMI_CPTR)
#undef WRFMI_DPTR
#undef WRFMI_CPTR
#define WRFMI_DR(n) float n;
  WRFMI_SCALARS_R(WRFMI_DR)
#undef WRFMI_DR
  void *a, *alpha, *gamma;
  int n_moist, n_species;
  FPtr<float> q[WRFMI_MAX_SPECIES];
  FPtr<float> q_old[WRFMI_MAX_SPECIES];
  FPtr<float> q_tend[WRFMI_MAX_SPECIES];
  const void *lbc;  // host memory: the boundary kernels take their arrays by argument
};
static_assert(sizeof(DState) == sizeof(wrfmi_state), "DState must mirror wrfmi_state");

struct KArgs {
  wrfmi_state s;   // device pointers
  wrfmi_config c;
  Dm d;
  float dts;       // small (acoustic) time step of this RK substep
  float dt_rk;     // RK substep length
  float dt;        // full time step
  int rk_step;
  int iteration;   // acoustic substep number (1-based)
  int nsmall;      // acoustic substeps in this RK substep
  int species;     // species index for the scalar kernels
  // library-owned 3-D work arrays, same layout as a grid field
  FPtr<float> scr_dv;   // advance_mu_t's column-mass divergence dvdxi
  FPtr<float> scr_rhs;  // advance_w's right-hand side rhs
  FPtr<float> scr_lo;   // low-order part of the FP64 forward-elimination value (precision mode 1)
  float *scr_qt;        // advect_tend of species 0 .. n_species-2 (all-species scalar kernels),
  unsigned long long qt_stride;  // one grid field (elements) apart
  int tnx, tny, tnz;  // tiles of the current launch in i, j and k (see tile3)
  int routine;        // Routine: R_FUSED in the whole step and the grouped entry points
};

// One field of a batched halo fill (set_physical_bc3d/2d on up to 16 fields per launch)
struct HaloField {
  float *p;
  int kind;  // 0 mass ('p','t'), 1 u-staggered ('u'), 2 v-staggered ('v'), 3 w levels ('w')
  int is3d;
};
struct HaloArgs {
  HaloField f[16];
  int n;
  Dm d;
  int periodic_x, periodic_y, open_xs, open_xe, open_ys, open_ye;
};

// Ensemble launches: one launch covers the workgroups of every member.  m[] holds each
// member's arguments, off[] the prefix sums of the members' workgroup counts (n + 1
// entries) and lut[] the member of every logical workgroup (total entries).
// The tables are read through the constant address space: the kernels store to the state
// arrays, and a table in ordinary global memory would have to be read again after every
// store (it might alias them); constant memory cannot, so its values stay in registers.
template <typename T>
struct Batch {
  const WRFMI_CONST_AS T *m;
  const WRFMI_CONST_AS unsigned *off;
  const WRFMI_CONST_AS unsigned *lut;
  int n;
  unsigned total;
};

#define DEVI __device__ __forceinline__

// XCD-aware tile order.  Kernels are launched on a 1-D grid of tnx*tny*tnz workgroups and
// the hardware deals workgroup b to XCD b % 8.  Each XCD is given a contiguous run of the
// logical tile sequence, ordered k fastest, then i, then j, so that the tiles which read
// each other's halo rows and levels run on one XCD close together in time and share its L2.
struct Tile3 {
  int bx, by, bz;
};
constexpr unsigned kXCD = 8;
// logical index of workgroup b among n, contiguous per XCD
DEVI unsigned xcd_logical(unsigned b, unsigned n) {
  const unsigned x = b % kXCD, r = b / kXCD;
  return x * (n / kXCD) + (x < n % kXCD ? x : n % kXCD) + r;
}
DEVI Tile3 tile_of_logical(unsigned L, int tnx, int tnz) {
  Tile3 t;
  t.bz = (int)(L % (unsigned)tnz);
  const unsigned q = L / (unsigned)tnz;
  t.bx = (int)(q % (unsigned)tnx);
  t.by = (int)(q / (unsigned)tnx);
  return t;
}
DEVI Tile3 tile3_of(unsigned b, int tnx, int tny, int tnz) {
  return tile_of_logical(xcd_logical(b, (unsigned)tnx * (unsigned)tny * (unsigned)tnz), tnx, tnz);
}

// First i of the one-column-per-lane kernels' tiles (64 columns wide): ids rounded down to a
// 32-point (128-byte) boundary of the row, so that a wave's 64 points of one level start on
// a cache line (rows are padded to a multiple of 32 points) and lanes below ids return at
// once.  A launch over one domain keeps ids instead when the rounding takes one more tile to
// cover the mass columns ids..ide-1 (widths that are multiples of 64): a tile row holding a
// few columns only, marching the whole column, was measured to double the launch time of
// the w solve on small domains.  A batched launch over many domains always rounds (measured
// 3 % faster on the 64-member ensemble, where such tiles overlap other members' work).
HDI int col_ia(const Dm &d, bool batched) {
  const int ia = d.ims + ((d.ids - d.ims) & ~31), n = d.ide - d.ids;
  return batched || ((n + d.ids - ia + 63) >> 6) == ((n + 63) >> 6) ? ia : d.ids;
}
DEVI int X3i(const Dm &d, int i, int k, int j) { return (i - d.ims) + d.ni * (k - d.kms) + d.sj * (j - d.jms); }
DEVI int X2i(const Dm &d, int i, int j) { return (i - d.ims) + d.ni * (j - d.jms); }
DEVI Off X3(const Dm &d, int i, int k, int j) { return Off{4u * (unsigned)X3i(d, i, k, j)}; }
DEVI Off X2(const Dm &d, int i, int j) { return Off{4u * (unsigned)X2i(d, i, j)}; }
DEVI Off X1(const Dm &d, int k) { return Off{4u * (unsigned)(k - d.kms)}; }
DEVI const DState &dstate(const KArgs &a) { return *reinterpret_cast<const DState *>(&a.s); }

DEVI Tile3 tile3(const KArgs &a) { return tile3_of(blockIdx.x, a.tnx, a.tny, a.tnz); }

typedef Batch<KArgs> BatchArgs;
typedef Batch<HaloArgs> BatchHalo;

DEVI const KArgs &kargs(const KArgs &a, Tile3 &T) {
  T = tile3(a);
  return a;
}
// Halo launches: the arguments and the workgroup index within the member
DEVI const HaloArgs &hargs(const HaloArgs &h, unsigned &blk) {
  blk = blockIdx.x;
  return h;
}
DEVI const HaloArgs &hargs(const BatchHalo &b, unsigned &blk) {
  const unsigned m = b.lut[blockIdx.x];
  blk = blockIdx.x - b.off[m];
  return *(const HaloArgs *)(b.m + m);
}
// A member of an ensemble launch: the XCD-aware order runs over the concatenated tiles of
// all members, so each XCD takes a contiguous run of members.
DEVI const KArgs &kargs(const BatchArgs &b, Tile3 &T) {
  const unsigned L = xcd_logical(blockIdx.x, b.total);
  const unsigned m = b.lut[L];
  const KArgs &a = *(const KArgs *)(b.m + m);
  T = tile_of_logical(L - b.off[m], a.tnx, a.tnz);
  return a;
}

// Lateral-boundary flags of the whole-patch tile (its = ids, ite = ide, ...).  The host
// accepts periodic, open and specified boundaries, so symmetric and polar are compile-time
// false here; PER = true is the doubly periodic specialisation, in which every
// open/specified-boundary and degraded-order branch is compiled out.
struct Geo {
  bool per_x, per_y, sym_xs, sym_xe, sym_ys, sym_ye, open_xs, open_xe, open_ys, open_ye, spec, nested, polar;
};
template <bool PER>
DEVI Geo geo(const wrfmi_config &c) {
  Geo g;
  g.per_x = PER ? true : (bool)c.periodic_x;
  g.per_y = PER ? true : (bool)c.periodic_y;
  g.sym_xs = g.sym_xe = g.sym_ys = g.sym_ye = false;
  g.open_xs = PER ? false : (bool)c.open_xs;
  g.open_xe = PER ? false : (bool)c.open_xe;
  g.open_ys = PER ? false : (bool)c.open_ys;
  g.open_ye = PER ? false : (bool)c.open_ye;
  g.spec = PER ? false : (bool)(c.specified || c.nested);
  g.nested = PER ? false : (bool)c.nested;
  g.polar = false;
  return g;
}

template <typename T> DEVI T imin(T a, T b) { return a < b ? a : b; }
template <typename T> DEVI T imax(T a, T b) { return a > b ? a : b; }

// Vector kernels: a lane holds the four points i0..i0+3, where i0 - ims is a multiple of four
// (so that the lane's memory is a 16-byte group); vec_ia0 is the first such i at or below ids.
HDI int vec_ia0(const Dm &d) { return d.ims + 4 * ((d.ids - d.ims) >> 2); }
// store of the lane's points that lie in [lo, hi]
DEVI void st4m(const FPtr<float> &p, Off o, const F4 &v, int i0, int lo, int hi) {
  if (i0 >= lo && i0 + 3 <= hi) {
    p.st<F4>(o, v);
  } else {
#pragma unroll
    for (int n = 0; n < 4; n++)
      if (i0 + n >= lo && i0 + n <= hi) p[o + n] = v.v[n];
  }
}
// Boundary zone of a domain with an open direction.  Within kZone points of an open edge
// the reference switches to lower-order fluxes and adds radiation terms; a four-point
// kernel stores only the points outside the zone (zlo/zhi in i, zrow in j), and the
// one-point kernel's zone form computes the zone itself: columns ids..ids+kZone-1 and
// ide-kZone..ide (every row) if x is open, then rows jds..jds+kZone-1 and jde-kZone..jde
// (the remaining columns) if y is open.  Point n of the zone is zone_point(n).
constexpr int kZone = 4;
DEVI int zlo(const KArgs &a, int lo) { return a.c.periodic_x ? lo : a.d.ids + kZone; }
DEVI int zhi(const KArgs &a, int hi) { return a.c.periodic_x ? hi : a.d.ide - 1 - kZone; }
DEVI bool zrow(const KArgs &a, int j) { return a.c.periodic_y || (j >= a.d.jds + kZone && j <= a.d.jde - 1 - kZone); }
HDI unsigned zone_count(const Dm &d, bool px, bool py) {
  const int zx = px ? 0 : kZone, zy = py ? 0 : kZone;
  const int cx = zx ? 2 * zx + 1 : 0, ry = zy ? 2 * zy + 1 : 0;
  const int ncol = d.ide - d.ids + 1, nrow = d.jde - d.jds + 1;
  return (unsigned)(cx * nrow + ry * (ncol - cx));
}
DEVI bool zone_point(const Dm &d, bool px, bool py, unsigned n, int &i, int &j) {
  const int zx = px ? 0 : kZone, zy = py ? 0 : kZone;
  const int cx = zx ? 2 * zx + 1 : 0, ry = zy ? 2 * zy + 1 : 0;
  const int ncol = d.ide - d.ids + 1, nrow = d.jde - d.jds + 1;
  const unsigned na = (unsigned)(cx * nrow);
  if (n < na) {
    const int c = (int)(n % (unsigned)cx), r = (int)(n / (unsigned)cx);
    i = c < zx ? d.ids + c : d.ide - zx + (c - zx);
    j = d.jds + r;
    return true;
  }
  n -= na;
  const int ic = ncol - cx;
  if (ry == 0 || ic <= 0 || n >= (unsigned)(ry * ic)) return false;
  const int c = (int)(n % (unsigned)ic), r = (int)(n / (unsigned)ic);
  i = d.ids + zx + c;
  j = r < zy ? d.jds + r : d.jde - zy + (r - zy);
  return true;
}
// (i, j) of a one-point kernel's thread: its tile position, or point n of the zone
template <bool ZONE>
DEVI bool point_ij(const KArgs &a, const Tile3 &T, int &i, int &j) {
  if (!ZONE) {
    i = a.d.ids + (int)(T.bx * blockDim.x + threadIdx.x);
    j = a.d.jds + (int)(T.by * blockDim.y + threadIdx.y);
    return true;
  }
  const unsigned n = ((unsigned)T.bx * blockDim.y + threadIdx.y) * blockDim.x + threadIdx.x;
  return zone_point(a.d, a.c.periodic_x, a.c.periodic_y, n, i, j);
}
// neighbours in i of point n from the lane's group c and the groups l (i0-4..) and r (i0+4..)
DEVI float xm1(const F4 &l, const F4 &c, int n) { return n ? c.v[n - 1] : l.v[3]; }
DEVI float xp1(const F4 &c, const F4 &r, int n) { return n < 3 ? c.v[n + 1] : r.v[0]; }

// share/module_model_constants.F:17-47, evaluated in FP32 as the reference's REAL parameters
constexpr float K_G = 9.81f;
constexpr float K_RD = 287.f;
constexpr float K_CP = 7.f * 287.f / 2.f;
constexpr float K_RV = 461.6f;
constexpr float K_CPOVCV = K_CP / (K_CP - K_RD);
constexpr float K_RVOVRD = K_RV / K_RD;
constexpr float K_T0 = 300.f;
constexpr float K_P0 = 100000.f;
constexpr float K_RERADIUS = 1.f / 6370.0e03f;
constexpr float K_CB = 25.f;
constexpr float K_W_ALPHA = 0.3f;
constexpr float K_W_BETA = 1.0f;

// Fortran SIGN(1.,x) for REAL x: 1 with the sign bit of x
DEVI float sgn1(float x) { return __builtin_copysignf(1.f, x); }

// x/12 and x/60 correctly rounded, as a product in double rounded once to float.
// For float x the exact quotient x/c (c = 12 = 3*4, c = 60 = 15*4) lies at least
// 2^-28 (relative) away from every midpoint between adjacent normal floats: x - c*m is a
// non-zero multiple of a power of two no smaller than 2^(E-22) for a midpoint m of binade
// E.  The double product x*(1/c) is within 2^-52 (relative) of x/c, so it rounds to the
// same float as the IEEE division (results in the subnormal range, |x/c| < 2^-126, excepted;
// zero is exact).  Three instructions instead of the IEEE division sequence.
DEVI float div12(float x) { return (float)((double)x * (1.0 / 12.0)); }
DEVI float div60(float x) { return (float)((double)x * (1.0 / 60.0)); }

// Flux operators of module_advect_em.F:199-213.  ts = SIGN(1,time_step) as a REAL.
DEVI float flux4(float qm2, float qm1, float q0, float qp1) {
  return div12(7.f * (q0 + qm1) - (qp1 + qm2));
}
DEVI float flux3(float qm2, float qm1, float q0, float qp1, float ua, float ts) {
  return flux4(qm2, qm1, q0, qp1) + div12(ts * sgn1(ua) * ((qp1 - qm2) - 3.f * (q0 - qm1)));
}
DEVI float flux6(float qm3, float qm2, float qm1, float q0, float qp1, float qp2) {
  return div60(37.f * (q0 + qm1) - 8.f * (qp1 + qm2) + (qp2 + qm3));
}
DEVI float flux5(float qm3, float qm2, float qm1, float q0, float qp1, float qp2, float ua, float ts) {
  return flux6(qm3, qm2, qm1, q0, qp1, qp2) - div60(ts * sgn1(ua) * ((qp2 - qm3) - 5.f * (qp1 - qm2) + 10.f * (q0 - qm1)));
}

// Error-free transformation used for the compensated column sums (precision mode 1)
DEVI void two_sum_acc(float &s, float &comp, float x) {
  float t = s + x;
  float bp = t - s;
  float e = (s - (t - bp)) + (x - bp);
  s = t;
  comp = comp + e;
}

// Software-pipelined column march: visits k = k0..k1 in order and calls body(k, vals),
// where vals[f] is field f's value at level k, loaded kPF levels ahead of its use.  The
// loads of one level do not depend on the body's results, so up to kPF levels of loads are
// in flight while the body works through the column (a thread per column otherwise waits
// a full memory latency per level).  load(f, k) must only read data the body never writes.
// The loads are unconditional (levels past k1 are clamped to k1): a load under a branch
// makes the compiler copy its destination at the join, and such a copy waits for every
// load in flight.  Full blocks of kPF levels run without any branch; the last partial
// block only calls the body.
constexpr int kPFdef = 8;
template <int NF, typename T, int kPF = kPFdef, typename Load, typename Body>
DEVI void column_march_up(int k0, int k1, Load &&load, Body &&body) {
  T ring[kPF][NF];
#pragma unroll
  for (int r = 0; r < kPF; r++)
#pragma unroll
    for (int f = 0; f < NF; f++) ring[r][f] = load(f, imin(k0 + r, k1));
  int kb = k0;
  for (; kb + kPF - 1 <= k1; kb += kPF) {
#pragma unroll
    for (int r = 0; r < kPF; r++) {
      T cur[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        cur[f] = ring[r][f];
        ring[r][f] = load(f, imin(kb + r + kPF, k1));
      }
      body(kb + r, cur);
    }
  }
#pragma unroll
  for (int r = 0; r < kPF - 1; r++)
    if (kb + r <= k1) body(kb + r, ring[r]);
}
// The same, visiting k = k1 down to k0.
template <int NF, typename T, int kPF = kPFdef, typename Load, typename Body>
DEVI void column_march_down(int k0, int k1, Load &&load, Body &&body) {
  T ring[kPF][NF];
#pragma unroll
  for (int r = 0; r < kPF; r++)
#pragma unroll
    for (int f = 0; f < NF; f++) ring[r][f] = load(f, imax(k1 - r, k0));
  int kb = k1;
  for (; kb - kPF + 1 >= k0; kb -= kPF) {
#pragma unroll
    for (int r = 0; r < kPF; r++) {
      T cur[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        cur[f] = ring[r][f];
        ring[r][f] = load(f, imax(kb - r - kPF, k0));
      }
      body(kb - r, cur);
    }
  }
#pragma unroll
  for (int r = 0; r < kPF - 1; r++)
    if (kb - r >= k0) body(kb - r, ring[r]);
}

}  // namespace wrfmi
