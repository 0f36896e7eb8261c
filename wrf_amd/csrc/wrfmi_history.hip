// History output in the reference's wrfout layout (include/wrfmi_history.h): a netCDF
// classic file in the 64-bit offset format, written without the netCDF library.
//
// Every variable has Time as its first (record) dimension, as wrfout's do, so the file is
// a header followed by records; a record holds each variable's values in header order,
// each padded to 4 bytes.  The GPU packs a record: per variable, the interior points of
// the field's staggering, reordered from the state's (j, k, i) memory layout to the file's
// (k, j, i) and byte-swapped to big-endian, into one device buffer that is copied to pinned
// host memory once and appended.  The header's record count is rewritten after each record.
//
// Names, dimensions and attributes follow the reference: dimension names from the Registry
// (west_east, south_north, bottom_top, *_stag, Time, DateStrLen = 19,
// external/io_netcdf/wrf_io.F90:1379-1410), variable attributes FieldType / MemoryOrder /
// description / units / stagger (wrf_io.F90:2693-2702), the Registry 'h' stream names and
// descriptions (Registry/Registry.EM_COMMON), and the global attributes of
// share/output_wrf.F:337-700 for the options of this step.
#include <errno.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>
#include "host_internal.h"
#include "wrfmi_history.h"

using namespace wrfmi;
using namespace wrfmi::host;

namespace wrfmi {

// one variable of a record: the interior (nx x ny x nk points, the field's staggering) of
// a 3-D (j, k, i), 2-D (j, i) or 1-D (k) array
struct PackVar {
  const float *src;
  const float *qv;      // mode 1: the QV array of the dry-theta diagnostic
  unsigned long long dst;  // element offset of the variable in the record buffer
  int nx, ny, nk, rank, mode;
};
struct PackArgs {
  PackVar v[64];
  int nvar;
  Dm d;
};

// grid-stride over the points of variable blockIdx.y
__global__ void __launch_bounds__(256) k_hist_pack(const PackArgs *pa, unsigned *out) {
  const PackArgs &a = *pa;
  const PackVar &v = a.v[blockIdx.y];
  const Dm &d = a.d;
  const long long n = (long long)v.nx * v.ny * v.nk;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % v.nx);
    const int j = (int)((t / v.nx) % v.ny);
    const int k = (int)(t / ((long long)v.nx * v.ny));
    size_t o;
    if (v.rank == 3)
      o = ((size_t)(j + d.jds - d.jms) * d.nk + (size_t)(k + d.kds - d.kms)) * d.ni + (size_t)(i + d.ids - d.ims);
    else if (v.rank == 2)
      o = (size_t)(j + d.jds - d.jms) * d.ni + (size_t)(i + d.ids - d.ims);
    else
      o = (size_t)(k + d.kds - d.kms);
    float x = v.src[o];
    if (v.mode == 1) {
      // dry theta perturbation from the moist one (module_big_step_utilities_em.F:5739)
      const float t0 = 300.0f, rvovrd = 461.6f / 287.0f;
      x = (x + t0) / (1.0f + rvovrd * v.qv[o]) - t0;
    }
    out[v.dst + (unsigned long long)t] = __builtin_bswap32(__float_as_uint(x));
  }
}

}  // namespace wrfmi

namespace {

enum { NC_CHAR = 2, NC_INT = 4, NC_FLOAT = 5 };
enum { NC_DIMENSION = 0x0A, NC_VARIABLE = 0x0B, NC_ATTRIBUTE = 0x0C };
enum { D_TIME, D_DATESTR, D_WE, D_SN, D_BT, D_BT_STAG, D_WE_STAG, D_SN_STAG, D_COUNT };

struct Attr {
  std::string name;
  int type;
  std::string text;  // NC_CHAR
  double num;        // NC_INT, NC_FLOAT (one value)
};
struct Var {
  std::string name;
  int type;
  std::vector<int> dims;  // after Time
  std::vector<Attr> atts;
  size_t nbytes = 0;      // unpadded bytes per record
  size_t vsize = 0;       // padded
  uint64_t begin = 0;
  // source: a state array (offset of its pointer in wrfmi_state), or a special value
  long field = -1;        // offsetof(wrfmi_state, <array>) or -1
  int species = -1;       // q[species]
  int rank = 0, mode = 0; // rank 3, 2, 1; mode 1 = dry theta from t_2 and q[0]
  int special = 0;        // 1 Times, 2 XTIME, 3 P_TOP
};

struct Out {
  std::vector<unsigned char> b;
  void u32(uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
  }
  void u64(uint64_t v) {
    u32((uint32_t)(v >> 32));
    u32((uint32_t)v);
  }
  void pad() {
    while (b.size() % 4) b.push_back(0);
  }
  void name(const std::string &s) {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    pad();
  }
  void f32(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    u32(u);
  }
  void atts(const std::vector<Attr> &as) {
    if (as.empty()) {
      u32(0);
      u32(0);
      return;
    }
    u32(NC_ATTRIBUTE);
    u32((uint32_t)as.size());
    for (const Attr &a : as) {
      name(a.name);
      u32((uint32_t)a.type);
      if (a.type == NC_CHAR) {
        u32((uint32_t)a.text.size());
        b.insert(b.end(), a.text.begin(), a.text.end());
        pad();
      } else {
        u32(1);
        if (a.type == NC_INT) u32((uint32_t)(int32_t)a.num);
        else f32((float)a.num);
      }
    }
  }
};

Attr A(const char *n, const std::string &s) { return Attr{n, NC_CHAR, s, 0.0}; }
Attr I(const char *n, int v) { return Attr{n, NC_INT, "", (double)v}; }
Attr F(const char *n, float v) { return Attr{n, NC_FLOAT, "", (double)v}; }

}  // namespace

struct wrfmi_history {
  FILE *f = nullptr;
  Dm d;
  int dimlen[D_COUNT];
  std::vector<Attr> gatts;
  std::vector<Var> vars;
  size_t header_bytes = 0, recsize = 0;
  int nrec = 0;
  int n_species = 0;
  int theta_m = 0;  // use_theta_m: T is diagnosed from the moist theta
  unsigned *dbuf = nullptr;       // device record buffer
  unsigned char *hbuf = nullptr;  // pinned host record buffer
  PackArgs *dargs = nullptr;      // device copy of the pack table
};

namespace {

std::vector<unsigned char> header(const wrfmi_history &h) {
  static const char *dname[D_COUNT] = {"Time", "DateStrLen", "west_east", "south_north", "bottom_top",
                                       "bottom_top_stag", "west_east_stag", "south_north_stag"};
  Out o;
  o.b.insert(o.b.end(), {'C', 'D', 'F', 2});
  o.u32((uint32_t)h.nrec);
  o.u32(NC_DIMENSION);
  o.u32(D_COUNT);
  for (int n = 0; n < D_COUNT; n++) {
    o.name(dname[n]);
    o.u32(n == D_TIME ? 0u : (uint32_t)h.dimlen[n]);
  }
  o.atts(h.gatts);
  o.u32(NC_VARIABLE);
  o.u32((uint32_t)h.vars.size());
  for (const Var &v : h.vars) {
    o.name(v.name);
    o.u32((uint32_t)(v.dims.size() + 1));
    o.u32(D_TIME);
    for (int dd : v.dims) o.u32((uint32_t)dd);
    o.atts(v.atts);
    o.u32((uint32_t)v.type);
    o.u32((uint32_t)v.vsize);
    o.u64(v.begin);
  }
  return o.b;
}

#define FIELD(n) ((long)offsetof(wrfmi_state, n))

// wrfout's variables for this state, in the file's order
void build_vars(wrfmi_history &h, const wrfmi_state *st, const char *start_date, const char *const *names) {
  auto var = [&](const std::string &name, std::vector<int> dims, const char *order, const char *desc,
                 const char *units, const char *stag) {
    Var v;
    v.name = name;
    v.type = NC_FLOAT;
    v.dims = dims;
    v.atts = {I("FieldType", 104), A("MemoryOrder", order), A("description", desc), A("units", units),
              A("stagger", stag)};
    v.rank = (int)dims.size();
    h.vars.push_back(v);
    return &h.vars.back();
  };
  Var times;
  times.name = "Times";
  times.type = NC_CHAR;
  times.dims = {D_DATESTR};
  times.special = 1;
  h.vars.push_back(times);
  Var *x = var("XTIME", {}, "0  ", "minutes since simulation start", "", "");
  x->atts[3] = A("units", std::string("minutes since ") + start_date);
  x->special = 2;
  var("P_TOP", {}, "0  ", "PRESSURE TOP OF THE MODEL", "Pa", "")->special = 3;
  struct L1 { const char *n; long f; bool stag; const char *desc; };
  const L1 one[] = {
      {"ZNU", FIELD(znu), false, "eta values on half (mass) levels"},
      {"ZNW", FIELD(znw), true, "eta values on full (w) levels"},
      {"DNW", FIELD(dnw), false, "d(eta) values between full (w) levels"},
      {"RDNW", FIELD(rdnw), false, "inverse d(eta) values between full (w) levels"},
      {"DN", FIELD(dn), false, "d(eta) values between half (mass) levels"},
      {"RDN", FIELD(rdn), false, "inverse d(eta) values between half (mass) levels"},
      {"FNM", FIELD(fnm), false, "upper weight for vertical stretching"},
      {"FNP", FIELD(fnp), false, "lower weight for vertical stretching"},
      {"C1H", FIELD(c1h), false, "half levels, c1h = d bf / d eta, using znw"},
      {"C2H", FIELD(c2h), false, "half levels, c2h = (1-c1h)*(p0-pt)"},
      {"C3H", FIELD(c3h), false, "half levels, c3h = bh"},
      {"C4H", FIELD(c4h), false, "half levels, c4h = (eta-bh)*(p0-pt), using znu"},
      {"C1F", FIELD(c1f), true, "full levels, c1f = d bf / d eta, using znu"},
      {"C2F", FIELD(c2f), true, "full levels, c2f = (1-c1f)*(p0-pt)"},
      {"C3F", FIELD(c3f), true, "full levels, c3f = bf"},
      {"C4F", FIELD(c4f), true, "full levels, c4f = (eta-bf)*(p0-pt), using znw"},
  };
  for (const L1 &l : one) {
    Var *v = var(l.n, {l.stag ? D_BT_STAG : D_BT}, "Z  ", l.desc, l.n[0] == 'C' && (l.n[1] == '2' || l.n[1] == '4') ? "Pa" : "",
                 l.stag ? "Z" : "");
    v->field = l.f;
  }
  struct L3 { const char *n; long f; int dz, dy, dx; const char *desc, *units, *stag; int mode; };
  const L3 three[] = {
      {"U", FIELD(u_2), D_BT, D_SN, D_WE_STAG, "x-wind component", "m s-1", "X", 0},
      {"V", FIELD(v_2), D_BT, D_SN_STAG, D_WE, "y-wind component", "m s-1", "Y", 0},
      {"W", FIELD(w_2), D_BT_STAG, D_SN, D_WE, "z-wind component", "m s-1", "Z", 0},
      {"PH", FIELD(ph_2), D_BT_STAG, D_SN, D_WE, "perturbation geopotential", "m2 s-2", "Z", 0},
      {"PHB", FIELD(phb), D_BT_STAG, D_SN, D_WE, "base-state geopotential", "m2 s-2", "Z", 0},
      {"T", FIELD(t_2), D_BT, D_SN, D_WE, "perturbation potential temperature theta-t0", "K", "", 1},
      {"THM", FIELD(t_2), D_BT, D_SN, D_WE,
       "either 1) pert moist pot temp=(1+Rv/Rd Qv)*(theta)-T0, or 2) pert dry pot temp=theta-T0; based on use_theta_m setting",
       "K", "", 0},
      {"P", FIELD(p), D_BT, D_SN, D_WE, "perturbation pressure", "Pa", "", 0},
      {"PB", FIELD(pb), D_BT, D_SN, D_WE, "BASE STATE PRESSURE", "Pa", "", 0},
  };
  for (const L3 &l : three) {
    Var *v = var(l.n, {l.dz, l.dy, l.dx}, "XYZ", l.desc, l.units, l.stag);
    v->field = l.f;
    v->mode = l.mode;
  }
  struct L2 { const char *n; long f; int dy, dx; const char *desc, *units, *stag; };
  const L2 two[] = {
      {"MU", FIELD(mu_2), D_SN, D_WE, "perturbation dry air mass in column", "Pa", ""},
      {"MUB", FIELD(mub), D_SN, D_WE, "base state dry air mass in column", "Pa", ""},
      {"HGT", FIELD(ht), D_SN, D_WE, "Terrain Height", "m", ""},
      {"MAPFAC_MX", FIELD(msftx), D_SN, D_WE, "Map scale factor on mass grid, x direction", "", ""},
      {"MAPFAC_MY", FIELD(msfty), D_SN, D_WE, "Map scale factor on mass grid, y direction", "", ""},
      {"MAPFAC_UX", FIELD(msfux), D_SN, D_WE_STAG, "Map scale factor on u-grid, x direction", "", "X"},
      {"MAPFAC_UY", FIELD(msfuy), D_SN, D_WE_STAG, "Map scale factor on u-grid, y direction", "", "X"},
      {"MAPFAC_VX", FIELD(msfvx), D_SN_STAG, D_WE, "Map scale factor on v-grid, x direction", "", "Y"},
      {"MAPFAC_VY", FIELD(msfvy), D_SN_STAG, D_WE, "Map scale factor on v-grid, y direction", "", "Y"},
  };
  for (const L2 &l : two) var(l.n, {l.dy, l.dx}, "XY ", l.desc, l.units, l.stag)->field = l.f;
  for (int m = 0; m < st->n_species; m++) {
    std::string n;
    if (names && names[m] && *names[m]) n = names[m];
    else if (m == 0 && st->n_moist > 0) n = "QVAPOR";
    else if (m < st->n_moist) n = "Q" + std::to_string(m + 1);
    else n = "SCALAR" + std::to_string(m - st->n_moist + 1);
    const bool moist = m < st->n_moist;
    Var *v = var(n, {D_BT, D_SN, D_WE}, "XYZ", m == 0 && moist ? "Water vapor mixing ratio" : moist ? "mixing ratio"
                                                                                                 : "passive scalar",
                 moist ? "kg kg-1" : "", "");
    v->species = m;
  }
}

int fail_io(const char *what) { return fail(WRFMI_E_IO, "history file: %s failed (%s)", what, strerror(errno)); }

void free_history(wrfmi_history *h) {
  if (!h) return;
  if (h->f) fclose(h->f);
  if (h->dbuf) (void)hipFree(h->dbuf);
  if (h->dargs) (void)hipFree(h->dargs);
  if (h->hbuf) (void)hipHostFree(h->hbuf);
  delete h;
}

}  // namespace

#define HCHK(x)                                                                                  \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess) {                                                                      \
      free_history(h);                                                                           \
      return fail(WRFMI_E_DEVICE, "%s: %s", #x, hipGetErrorString(e_));                          \
    }                                                                                            \
  } while (0)

extern "C" int wrfmi_history_open(const char *path, const wrfmi_state *st, const wrfmi_config *cfg,
                                  const wrfmi_dims *dims, const char *start_date, const char *const *species_names,
                                  wrfmi_history **out) {
  if (!path || !st || !cfg || !dims || !start_date || !out) return fail(WRFMI_E_POINTER, "null argument");
  *out = nullptr;
  if (strlen(start_date) != 19) return fail(WRFMI_E_BADOPT, "start_date must be YYYY-MM-DD_hh:mm:ss (19 characters)");
  if (st->n_moist < 0 || st->n_species < st->n_moist || st->n_species > WRFMI_MAX_SPECIES)
    return fail(WRFMI_E_BADOPT, "n_moist=%d n_species=%d out of range", st->n_moist, st->n_species);
  std::lock_guard<std::mutex> lk(lock());
  wrfmi_history *h = new wrfmi_history();
  if (int rc = make_dm(dims, &h->d)) {
    delete h;
    return rc;
  }
  const Dm &d = h->d;
  const int nx = d.ide - d.ids, ny = d.jde - d.jds, nz = d.kde - d.kds;
  const int len[D_COUNT] = {0, 19, nx, ny, nz, nz + 1, nx + 1, ny + 1};
  memcpy(h->dimlen, len, sizeof len);
  h->n_species = st->n_species;
  h->theta_m = cfg->use_theta_m;
  h->gatts = {A("TITLE", " OUTPUT FROM WRF V4 MODEL (MI355X DYNAMICS STEP)"), A("START_DATE", start_date),
              A("SIMULATION_START_DATE", start_date), I("WEST-EAST_GRID_DIMENSION", nx + 1),
              I("SOUTH-NORTH_GRID_DIMENSION", ny + 1), I("BOTTOM-TOP_GRID_DIMENSION", nz + 1), F("DX", st->dx),
              F("DY", st->dy), F("DT", st->dt), A("GRIDTYPE", "C"), I("DIFF_OPT", cfg->diff_opt), I("KM_OPT", cfg->km_opt),
              I("DAMP_OPT", cfg->damp_opt), F("DAMPCOEF", cfg->dampcoef), F("KHDIF", cfg->khdif), F("KVDIF", cfg->kvdif),
              I("DIFF_6TH_OPT", cfg->diff_6th_opt), F("DIFF_6TH_FACTOR", cfg->diff_6th_factor),
              I("DIFF_6TH_SLOPEOPT", cfg->diff_6th_slopeopt), F("DIFF_6TH_THRESH", cfg->diff_6th_thresh),
              I("MP_PHYSICS", 0), I("BL_PBL_PHYSICS", cfg->bl_pbl_physics), I("CU_PHYSICS", 0),
              I("MOIST_ADV_OPT", cfg->moist_adv_opt), I("SCALAR_ADV_OPT", cfg->scalar_adv_opt),
              I("HYPSOMETRIC_OPT", cfg->hypsometric_opt), I("USE_THETA_M", cfg->use_theta_m),
              I("W_DAMPING", cfg->w_damping), I("MAP_PROJ", cfg->map_proj)};
  build_vars(*h, st, start_date, species_names);
  if (h->vars.size() > 3 + 64) {
    delete h;
    return fail(WRFMI_E_BADOPT, "too many history variables");
  }
  // record layout: each variable's bytes per record, padded to 4
  size_t off = 0;
  for (Var &v : h->vars) {
    size_t n = v.type == NC_CHAR ? 1 : 4;
    for (int dd : v.dims) n *= (size_t)h->dimlen[dd];
    v.nbytes = n;
    v.vsize = (n + 3) & ~(size_t)3;
    v.begin = off;  // relative until the header size is known
    off += v.vsize;
  }
  h->recsize = off;
  h->header_bytes = header(*h).size();
  for (Var &v : h->vars) v.begin += h->header_bytes;
  h->f = fopen(path, "wb+");
  if (!h->f) {
    delete h;
    return fail(WRFMI_E_IO, "history file: cannot create %s (%s)", path, strerror(errno));
  }
  const std::vector<unsigned char> hd = header(*h);
  if (fwrite(hd.data(), 1, hd.size(), h->f) != hd.size() || fflush(h->f) != 0) {
    int rc = fail_io("header write");
    free_history(h);
    return rc;
  }
  HCHK(hipMalloc(&h->dbuf, h->recsize));
  HCHK(hipHostMalloc(&h->hbuf, h->recsize, hipHostMallocDefault));
  HCHK(hipMalloc(&h->dargs, sizeof(PackArgs)));
  *out = h;
  return 0;
}

extern "C" int wrfmi_history_write(wrfmi_history *h, const wrfmi_state *st, const char *date, float xtime,
                                   void *stream) {
  if (!h || !st || !date) return fail(WRFMI_E_POINTER, "null argument");
  if (strlen(date) != 19) return fail(WRFMI_E_BADOPT, "date must be YYYY-MM-DD_hh:mm:ss (19 characters)");
  if (st->n_species != h->n_species) return fail(WRFMI_E_BADOPT, "n_species differs from the one given at open");
  std::lock_guard<std::mutex> lk(lock());
  hipStream_t s = (hipStream_t)stream;
  PackArgs pa;
  memset(&pa, 0, sizeof pa);
  pa.d = h->d;
  long long maxn = 1;
  for (const Var &v : h->vars) {
    if (v.special) continue;
    const float *src;
    if (v.species >= 0) src = st->q[v.species];
    else src = *(float *const *)((const char *)st + v.field);
    const float *dev = (const float *)translate((void *)src);
    if (!dev) return fail(WRFMI_E_POINTER, "history variable %s: array is null", v.name.c_str());
    if (int rc = device_array(dev, v.name.c_str())) return rc;
    PackVar &p = pa.v[pa.nvar++];
    p.src = dev;
    p.rank = v.rank;
    p.mode = v.mode && h->theta_m && st->n_moist > 0 ? 1 : 0;
    if (p.mode) {
      p.qv = (const float *)translate((void *)st->q[0]);
      if (!p.qv) return fail(WRFMI_E_POINTER, "history variable T: QV array is null");
    }
    p.dst = (v.begin - h->header_bytes) / 4;
    p.nk = v.rank == 2 ? 1 : h->dimlen[v.dims[0]];
    p.ny = v.rank == 1 ? 1 : h->dimlen[v.dims[v.rank - 2]];
    p.nx = v.rank == 1 ? 1 : h->dimlen[v.dims[v.rank - 1]];
    const long long n = (long long)p.nx * p.ny * p.nk;
    if (n > maxn) maxn = n;
  }
  if (hipMemcpyAsync(h->dargs, &pa, sizeof pa, hipMemcpyHostToDevice, s) != hipSuccess)
    return fail(WRFMI_E_DEVICE, "history: copy of the pack table failed");
  // grid-stride: at most 4096 workgroups per variable
  const unsigned gx = (unsigned)((maxn + 255) / 256 > 4096 ? 4096 : (maxn + 255) / 256);
  hipLaunchKernelGGL(k_hist_pack, dim3(gx, (unsigned)pa.nvar), dim3(256), 0, s, (const PackArgs *)h->dargs, h->dbuf);
  if (hipGetLastError() != hipSuccess) return fail(WRFMI_E_DEVICE, "history: pack kernel launch failed");
  if (hipMemcpyAsync(h->hbuf, h->dbuf, h->recsize, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(WRFMI_E_DEVICE, "history: copy of the record failed");
  // the host-side values: Times, XTIME, P_TOP
  for (const Var &v : h->vars) {
    unsigned char *dst = h->hbuf + (v.begin - h->header_bytes);
    if (v.special == 1) {
      memset(dst, 0, v.vsize);
      memcpy(dst, date, 19);
    } else if (v.special == 2 || v.special == 3) {
      const float x = v.special == 2 ? xtime : st->p_top;
      uint32_t u;
      memcpy(&u, &x, 4);
      u = __builtin_bswap32(u);
      memcpy(dst, &u, 4);
    }
  }
  const long long at = (long long)h->header_bytes + (long long)h->nrec * (long long)h->recsize;
  if (fseeko(h->f, (off_t)at, SEEK_SET) != 0 || fwrite(h->hbuf, 1, h->recsize, h->f) != h->recsize)
    return fail_io("record write");
  h->nrec++;
  const uint32_t nr = __builtin_bswap32((uint32_t)h->nrec);
  if (fseeko(h->f, 4, SEEK_SET) != 0 || fwrite(&nr, 4, 1, h->f) != 1 || fflush(h->f) != 0)
    return fail_io("record count update");
  return 0;
}

extern "C" int wrfmi_history_records(const wrfmi_history *h) { return h ? h->nrec : -1; }

extern "C" int wrfmi_history_close(wrfmi_history *h) {
  if (!h) return fail(WRFMI_E_POINTER, "null argument");
  std::lock_guard<std::mutex> lk(lock());
  int rc = 0;
  if (h->f && fclose(h->f) != 0) rc = fail(WRFMI_E_IO, "history file: close failed (%s)", strerror(errno));
  h->f = nullptr;
  free_history(h);
  return rc;
}
