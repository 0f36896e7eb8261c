"""Domain set-up for the WRF-ARW dynamics step: index sextets, options, and the
analytic initial states of the benchmark configurations.

Arrays are numpy float32 in the reference's Fortran layout ``(ims:ime, kms:kme,
jms:jme)`` with i fastest; in C order that is shape ``(nj, nk, ni)`` indexed
``[j - jms, k - kms, i - ims]``.  The initial state follows the steps of the
reference's idealised initialisation (dyn_em/module_initialize_ideal.F:611-1130:
eta levels and their metrics, hybrid coefficients, dry base state from a dry
sounding, moist full state from the moist sounding, hydrostatic geopotential,
warm bubble) with an analytic Weisman-Klemp sounding in place of the
input_sounding file, computed in float64 and rounded once to float32.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np

from .fields import FIELDS_1D, FIELDS_2D, FIELDS_3D, MAX_SPECIES, Config, Dims

# share/module_model_constants.F:17-47
G = 9.81
R_D = 287.0
CP = 7.0 * R_D / 2.0
R_V = 461.6
CV = CP - R_D
CVPM = -CV / CP
RCP = R_D / CP
P1000 = 1.0e5
T0 = 300.0
RVOVRD = R_V / R_D
BDYZONE = 4  # share/module_bc.F:33


@dataclasses.dataclass
class DomainSpec:
    nx: int = 40              # mass points in x (e_we = nx + 1)
    ny: int = 40
    nz: int = 40              # mass levels (e_vert = nz + 1)
    dx: float = 2000.0
    dy: float = 2000.0
    dt: float = 12.0
    ztop: float = 20000.0
    time_step_sound: int = 6
    bc: str = "periodic"      # "periodic" | "open" | "specified" (both directions unless bc_x/bc_y
                              # say otherwise; "specified" in y with x "specified" or "periodic")
    bc_x: str = ""            # per-direction override of bc
    bc_y: str = ""
    hybrid_opt: int = 0
    etac: float = 0.2
    n_passive: int = 0        # passive scalars after QV
    n_bubbles: int = 1
    bubble_amp: float = 3.0
    noise_amp: float = 0.1
    noise_levels: int = 10
    wind_scale: float = 1.0   # multiplies the hodograph winds (0 gives a resting atmosphere)
    seed: int = 1001
    pad_i: int = 32           # pad the memory i-extent to a multiple of this (alignment)
    precision_mode: int = 0
    # sounding "wk" (Weisman-Klemp, hodograph winds) or "stable" (constant Brunt-Vaisala
    # frequency, uniform westerly); the stable sounding's parameters follow
    sounding: str = "wk"
    theta0: float = 288.0     # surface potential temperature (K)
    brunt: float = 0.01       # N (1/s): theta = theta0 * exp(N^2 z / g)
    qv0: float = 0.008        # surface vapour mixing ratio (kg/kg), e-folding height qv_scale
    qv_scale: float = 2500.0
    u0: float = 10.0          # uniform westerly (m/s)
    # terrain: n_ridges Witch-of-Agnesi hills h_m / (1 + (r/a)^2) at seeded positions
    n_ridges: int = 0
    ridge_a: tuple = (5000.0, 20000.0)
    ridge_h: tuple = (200.0, 1500.0)

    @property
    def bcx(self):
        return self.bc_x or self.bc

    @property
    def bcy(self):
        return self.bc_y or self.bc

    @property
    def e_we(self):
        return self.nx + 1

    @property
    def e_sn(self):
        return self.ny + 1

    @property
    def e_vert(self):
        return self.nz + 1


def make_dims(spec: DomainSpec) -> Dims:
    """Serial WRF domain/memory/tile sextets (frame/module_domain.F:686-687: halo = bdyzone).
    The tile is the whole patch, as in a single-tile run."""
    d = Dims()
    d.ids, d.ide = 1, spec.e_we
    d.jds, d.jde = 1, spec.e_sn
    d.kds, d.kde = 1, spec.e_vert
    d.ims, d.jms, d.kms = d.ids - BDYZONE, d.jds - BDYZONE, d.kds
    ime = d.ide + BDYZONE
    ni = ime - d.ims + 1
    if spec.pad_i > 1:
        ni = -(-ni // spec.pad_i) * spec.pad_i
    d.ime = d.ims + ni - 1
    d.jme = d.jde + BDYZONE
    d.kme = d.kde
    d.its, d.ite, d.jts, d.jte, d.kts, d.kte = d.ids, d.ide, d.jds, d.jde, d.kds, d.kde
    return d


def make_config(spec: DomainSpec) -> Config:
    """Namelist/Registry defaults (Registry/Registry.EM_COMMON:2840-2902) with physics,
    diffusion and damping off, as the benchmark configurations state."""
    c = Config()
    for b in (spec.bcx, spec.bcy):
        if b not in ("periodic", "open", "specified"):
            raise ValueError(f"lateral boundary {b!r}: periodic, open or specified")
    if "specified" in (spec.bcx, spec.bcy) and not (spec.bcy == "specified" and spec.bcx in ("specified", "periodic")):
        raise ValueError("specified boundaries: on all four sides, or in y with a periodic x")
    c.specified = int(spec.bcy == "specified")
    # namelist defaults (Registry/Registry.EM_COMMON: spec_bdy_width 5, spec_zone 1, relax_zone 4)
    c.spec_bdy_width, c.spec_zone, c.relax_zone = 5, 1, 4
    c.have_bcs_moist = c.have_bcs_scalar = 0
    c.periodic_x = int(spec.bcx == "periodic")
    c.periodic_y = int(spec.bcy == "periodic")
    c.open_xs = c.open_xe = int(spec.bcx == "open")
    c.open_ys = c.open_ye = int(spec.bcy == "open")
    c.rk_ord = 3
    c.h_mom_adv_order, c.v_mom_adv_order = 5, 3
    c.h_sca_adv_order, c.v_sca_adv_order = 5, 3
    c.moist_adv_opt = c.scalar_adv_opt = 0
    c.momentum_adv_opt = 1
    c.non_hydrostatic = 1
    c.top_lid = 0
    c.phi_adv_z = 1
    c.hypsometric_opt = 2
    c.use_theta_m = 1
    c.damp_opt = 0
    c.w_damping = 0
    c.diff_opt = 0
    c.diff_6th_opt = 0
    c.zadvect_implicit = 0
    c.pert_coriolis = 0
    c.map_proj = 0
    c.time_step_sound = spec.time_step_sound
    c.time_step = int(round(spec.dt))
    c.epssm, c.smdiv, c.emdiv = 0.1, 0.1, 0.01
    c.dampcoef, c.zdamp, c.w_crit_cfl = 0.2, 5000.0, 1.2
    c.diff_6th_factor, c.diff_6th_thresh, c.diff_6th_slopeopt = 0.12, 0.10, 0
    c.km_opt, c.bl_pbl_physics, c.khdif, c.kvdif = 1, 0, 0.0, 0.0
    c.precision_mode = spec.precision_mode
    return c


def shape3(d: Dims):
    return (d.jme - d.jms + 1, d.kme - d.kms + 1, d.ime - d.ims + 1)


def shape2(d: Dims):
    return (d.jme - d.jms + 1, d.ime - d.ims + 1)


def shape1(d: Dims):
    return (d.kme - d.kms + 1,)


# ----------------------------------------------------------------------------- sounding
def wk_theta(z):
    """Weisman-Klemp (1982) potential temperature: 300 K at the ground, 343 K at 12 km,
    isothermal (213 K) stratosphere above."""
    z = np.asarray(z, dtype=np.float64)
    ztr, th0, thtr, ttr = 12000.0, 300.0, 343.0, 213.0
    lo = th0 + (thtr - th0) * (np.clip(z, 0, ztr) / ztr) ** 1.25
    hi = thtr * np.exp(G * (z - ztr) / (CP * ttr))
    return np.where(z <= ztr, lo, hi)


def wk_rh(z):
    z = np.asarray(z, dtype=np.float64)
    ztr = 12000.0
    return np.where(z <= ztr, 1.0 - 0.75 * (np.clip(z, 0, ztr) / ztr) ** 1.25, 0.25)


def _qvs(t, p):
    es = 611.2 * np.exp(17.67 * (t - 273.15) / (t - 29.65))
    return 0.622 * es / np.maximum(p - es, 1.0)


def sounding(moist: bool, zmax=30000.0, dz=5.0, psfc=1.0e5, qv_max=0.014):
    """Hydrostatic integration of the analytic sounding on a fine z grid.
    Returns z, p (full), pd (dry hydrostatic), theta, qv."""
    z = np.arange(0.0, zmax + dz, dz)
    th = wk_theta(z)
    rh = wk_rh(z) if moist else np.zeros_like(z)
    n = z.size
    p = np.empty(n)
    qv = np.zeros(n)
    rho = np.empty(n)
    p[0] = psfc
    for it in range(n):
        if it > 0:
            # trapezoidal step with a fixed-point iteration (cf. get_sounding's 10 iterations)
            pg = p[it - 1] - G * rho[it - 1] * dz
            for _ in range(10):
                t = th[it] * (pg / P1000) ** RCP
                q = min(rh[it] * _qvs(t, pg), qv_max) if moist else 0.0
                rh_ = pg / (R_D * t * (1.0 + RVOVRD * q) / (1.0 + q))
                pg = p[it - 1] - 0.5 * G * (rho[it - 1] + rh_) * dz
            p[it] = pg
        t = th[it] * (p[it] / P1000) ** RCP
        qv[it] = min(rh[it] * _qvs(t, p[it]), qv_max) if moist else 0.0
        rho[it] = p[it] / (R_D * t * (1.0 + RVOVRD * qv[it]) / (1.0 + qv[it]))
    # dry hydrostatic pressure: remove the weight of the vapour above each level
    wv = G * rho * qv / (1.0 + qv)
    col = np.concatenate([np.cumsum((0.5 * (wv[1:] + wv[:-1]) * dz)[::-1])[::-1], [0.0]])
    pd = p - col
    return z, p, pd, th, qv


def stable_sounding(spec: DomainSpec, moist: bool, zmax=30000.0, dz=5.0, psfc=1.0e5):
    """Hydrostatic integration of a constant-N sounding, theta = theta0 * exp(N^2 z / g),
    with vapour qv0 * exp(-z / qv_scale) when `moist`.  Same integration and returns as
    `sounding` (z, p, pd, theta, qv)."""
    z = np.arange(0.0, zmax + dz, dz)
    th = spec.theta0 * np.exp(spec.brunt ** 2 * z / G)
    qv = spec.qv0 * np.exp(-z / spec.qv_scale) if moist else np.zeros_like(z)
    n = z.size
    p = np.empty(n)
    rho = np.empty(n)
    p[0] = psfc
    for it in range(n):
        if it > 0:
            pg = p[it - 1] - G * rho[it - 1] * dz
            for _ in range(10):
                t = th[it] * (pg / P1000) ** RCP
                rh_ = pg / (R_D * t * (1.0 + RVOVRD * qv[it]) / (1.0 + qv[it]))
                pg = p[it - 1] - 0.5 * G * (rho[it - 1] + rh_) * dz
            p[it] = pg
        t = th[it] * (p[it] / P1000) ** RCP
        rho[it] = p[it] / (R_D * t * (1.0 + RVOVRD * qv[it]) / (1.0 + qv[it]))
    wv = G * rho * qv / (1.0 + qv)
    col = np.concatenate([np.cumsum((0.5 * (wv[1:] + wv[:-1]) * dz)[::-1])[::-1], [0.0]])
    return z, p, p - col, th, qv


def ridge_field(spec: DomainSpec, rng) -> np.ndarray:
    """Terrain height (ny, nx) at mass points: the sum of `n_ridges` Witch-of-Agnesi hills
    h_m / (1 + (r/a)^2), centres uniform over the domain, a and h_m uniform in ridge_a and
    ridge_h; distances wrap in a periodic direction."""
    nx, ny = spec.nx, spec.ny
    Lx, Ly = nx * spec.dx, ny * spec.dy
    xs = (np.arange(nx) + 0.5) * spec.dx
    ys = (np.arange(ny) + 0.5) * spec.dy
    ht = np.zeros((ny, nx))
    for _ in range(spec.n_ridges):
        xc, yc = rng.uniform(0.0, Lx), rng.uniform(0.0, Ly)
        a = rng.uniform(*spec.ridge_a)
        hm = rng.uniform(*spec.ridge_h)
        dxr, dyr = xs - xc, ys - yc
        if spec.bcx == "periodic":
            dxr = (dxr + 0.5 * Lx) % Lx - 0.5 * Lx
        if spec.bcy == "periodic":
            dyr = (dyr + 0.5 * Ly) % Ly - 0.5 * Ly
        ht += hm / (1.0 + (dxr[None, :] ** 2 + dyr[:, None] ** 2) / (a * a))
    return ht


def hodograph(z):
    """Quarter-circle hodograph: radius 8 m/s to 2 km, then straight to 32 m/s at 6 km."""
    z = np.asarray(z, dtype=np.float64)
    ang = 0.5 * np.pi * np.clip(z, 0, 2000.0) / 2000.0
    u = 8.0 * (1.0 - np.cos(ang))
    v = 8.0 * np.sin(ang)
    up = 8.0 + (32.0 - 8.0) * np.clip((z - 2000.0) / 4000.0, 0.0, 1.0)
    u = np.where(z <= 2000.0, u, up)
    v = np.where(z <= 2000.0, v, 8.0)
    return u, v


def _interp_p(var, pin, plev):
    """interp_0 by pressure (pressure decreases with height)."""
    return np.interp(-np.asarray(plev), -pin, var)


# ----------------------------------------------------------------------------- state
class HostState:
    """Host-side arrays of one domain (numpy float32, Fortran layout)."""

    def __init__(self, spec: DomainSpec):
        self.spec = spec
        self.dims = make_dims(spec)
        self.cfg = make_config(spec)
        d = self.dims
        self.a3 = {n: np.zeros(shape3(d), np.float32) for n in FIELDS_3D}
        self.a2 = {n: np.zeros(shape2(d), np.float32) for n in FIELDS_2D}
        self.a1 = {n: np.zeros(shape1(d), np.float32) for n in FIELDS_1D}
        self.scal = {}
        self.n_moist = 1
        self.n_species = 1 + spec.n_passive
        if self.n_species > MAX_SPECIES:
            raise ValueError("too many species")
        self.q = [np.zeros(shape3(d), np.float32) for _ in range(self.n_species)]
        self.q_old = [np.zeros(shape3(d), np.float32) for _ in range(self.n_species)]
        self.q_tend = [np.zeros(shape3(d), np.float32) for _ in range(self.n_species)]
        self.coef_dtype = np.float64 if spec.precision_mode == 1 else np.float32
        self.lbc = None  # lateral boundary arrays of a specified domain (build_lbc)
        self.a = np.zeros(shape3(d), self.coef_dtype)
        self.alpha = np.zeros(shape3(d), self.coef_dtype)
        self.gamma = np.zeros(shape3(d), self.coef_dtype)

    # all arrays, name -> ndarray (species as q0.., q_old0.., q_tend0..)
    def arrays(self):
        out = {}
        out.update(self.a3)
        out.update(self.a2)
        out.update(self.a1)
        for m in range(self.n_species):
            out[f"q{m}"] = self.q[m]
            out[f"q_old{m}"] = self.q_old[m]
            out[f"q_tend{m}"] = self.q_tend[m]
        out["a"], out["alpha"], out["gamma"] = self.a, self.alpha, self.gamma
        if self.lbc is not None:
            out.update(lbc_arrays(self.lbc))
        return out

    def copy(self):
        other = HostState.__new__(HostState)
        other.spec, other.dims, other.cfg = self.spec, self.dims, self.cfg
        other.a3 = {k: v.copy() for k, v in self.a3.items()}
        other.a2 = {k: v.copy() for k, v in self.a2.items()}
        other.a1 = {k: v.copy() for k, v in self.a1.items()}
        other.scal = dict(self.scal)
        other.n_moist, other.n_species = self.n_moist, self.n_species
        other.q = [x.copy() for x in self.q]
        other.q_old = [x.copy() for x in self.q_old]
        other.q_tend = [x.copy() for x in self.q_tend]
        other.coef_dtype = self.coef_dtype
        other.a, other.alpha, other.gamma = self.a.copy(), self.alpha.copy(), self.gamma.copy()
        other.lbc = None
        if self.lbc is not None:
            other.lbc = {k: ([None if sides is None else [a.copy() for a in sides] for sides in v]
                             if k in ("q_b", "q_bt")
                             else [a.copy() for a in v] if isinstance(v, list) else np.array(v, copy=True))
                         for k, v in self.lbc.items()}
        return other


def _halo_fill_np(a: np.ndarray, d: Dims, periodic, stag_i: bool, stag_j: bool):
    """Initial halo fill (periodic wrap or zero-gradient copy) of a (nj, [nk,] ni) array.
    `periodic` is one flag for both directions or a pair (periodic in x, periodic in y)."""
    per_x, per_y = (periodic, periodic) if isinstance(periodic, (bool, int, np.bool_)) else periodic
    nx, ny = d.ide - d.ids, d.jde - d.jds
    i_idx = np.arange(d.ims, d.ime + 1)
    j_idx = np.arange(d.jms, d.jme + 1)
    if per_x:
        si = np.where(i_idx < d.ids, i_idx + nx, np.where(i_idx > d.ide - 1 + (1 if stag_i else 0), i_idx - nx, i_idx))
        if stag_i:
            si = np.where(i_idx == d.ide, d.ids, si)
    else:
        si = np.clip(i_idx, d.ids, d.ide if stag_i else d.ide - 1)
    if per_y:
        sj = np.where(j_idx < d.jds, j_idx + ny, np.where(j_idx > d.jde - 1 + (1 if stag_j else 0), j_idx - ny, j_idx))
        if stag_j:
            sj = np.where(j_idx == d.jde, d.jds, sj)
    else:
        sj = np.clip(j_idx, d.jds, d.jde if stag_j else d.jde - 1)
    si = np.clip(si, d.ims, d.ime) - d.ims
    sj = np.clip(sj, d.jms, d.jme) - d.jms
    # copy only the halo strips: i first over every row, then whole rows in j (corners
    # take a[sj, si]); every source index is a non-halo one
    ih = np.nonzero(si != np.arange(si.size))[0]
    jh = np.nonzero(sj != np.arange(sj.size))[0]
    a[..., ih] = a[..., si[ih]]
    a[jh] = a[sj[jh]]


def _vertical_coords(hs: HostState, spec: DomainSpec, p_top: float):
    """Eta levels, their metrics and the hybrid coefficients (module_initialize_ideal.F:
    617-877), in float64; stores the 1-D arrays and the scalars of `hs` and returns the
    float64 arrays by name."""
    kde = hs.dims.kde
    # eta levels and metrics (module_initialize_ideal.F:617-730), float64
    znw = 1.0 - np.arange(kde, dtype=np.float64) / (kde - 1)  # index k-1
    dnw = znw[1:] - znw[:-1]
    rdnw = 1.0 / dnw
    znu = 0.5 * (znw[1:] + znw[:-1])
    dn = np.zeros(kde)
    rdn = np.zeros(kde)
    fnp = np.zeros(kde)
    fnm = np.zeros(kde)
    for k in range(2, kde):  # Fortran k = 2..kde-1 -> index k-1
        dn[k - 1] = 0.5 * (dnw[k - 1] + dnw[k - 2])
        rdn[k - 1] = 1.0 / dn[k - 1]
        fnp[k - 1] = 0.5 * dnw[k - 1] / dn[k - 1]
        fnm[k - 1] = 0.5 * dnw[k - 2] / dn[k - 1]
    cof1 = (2.0 * dn[1] + dn[2]) / (dn[1] + dn[2]) * dnw[0] / dn[1]
    cof2 = dn[1] / (dn[1] + dn[2]) * dnw[0] / dn[2]
    cf1 = fnp[1] + cof1
    cf2 = fnm[1] - cof1 - cof2
    cf3 = cof2
    cfn = (0.5 * dnw[kde - 2] + dn[kde - 2]) / dn[kde - 2]
    cfn1 = -0.5 * dnw[kde - 2] / dn[kde - 2]

    # hybrid coefficients (module_initialize_ideal.F:791-877)
    if spec.hybrid_opt == 0:
        c3f = znw.copy()
    elif spec.hybrid_opt == 2:
        e = spec.etac
        b1 = 2.0 * e ** 2 * (1.0 - e)
        b2 = -e * (4.0 - 3.0 * e - e ** 3)
        b3 = 2.0 * (1.0 - e ** 3)
        b4 = -(1.0 - e ** 2)
        b5 = (1.0 - e) ** 4
        c3f = (b1 + b2 * znw + b3 * znw ** 2 + b4 * znw ** 3) / b5
        c3f = np.where(znw < e, 0.0, c3f)
        c3f[0], c3f[-1] = 1.0, 0.0
    else:
        raise ValueError("hybrid_opt must be 0 or 2")
    c4f = (znw - c3f) * (P1000 - p_top)
    c3h = (c3f[1:] + c3f[:-1]) * 0.5
    c4h = (znu - c3h) * (P1000 - p_top)
    c1f = np.zeros(kde)
    for k in range(2, kde):
        c1f[k - 1] = (c3h[k - 1] - c3h[k - 2]) / (znu[k - 1] - znu[k - 2])
    c1f[0] = 1.0
    c1f[kde - 1] = 1.0 if spec.hybrid_opt in (0, 1) else 0.0
    c2f = (1.0 - c1f) * (P1000 - p_top)
    c1h = (c3f[1:] - c3f[:-1]) / (znw[1:] - znw[:-1])
    c2h = (1.0 - c1h) * (P1000 - p_top)
    if spec.hybrid_opt == 0:
        c3h = znu.copy()
        c4h[:] = 0.0
        c1f[:] = 1.0
        c2f[:] = 0.0
        c1h[:] = 1.0
        c2h[:] = 0.0

    def put1(name, vals):  # vals indexed k-1 for k = 1..len
        arr = hs.a1[name]
        arr[: len(vals)] = np.asarray(vals, dtype=np.float64).astype(np.float32)

    put1("znw", znw); put1("dnw", dnw); put1("rdnw", rdnw); put1("znu", znu)
    put1("dn", dn); put1("rdn", rdn); put1("fnp", fnp); put1("fnm", fnm)
    put1("c1h", c1h); put1("c2h", c2h); put1("c3h", c3h); put1("c4h", c4h)
    put1("c1f", c1f); put1("c2f", c2f); put1("c3f", c3f); put1("c4f", c4f)
    hs.scal.update(dict(dt=spec.dt, dx=spec.dx, dy=spec.dy, rdx=1.0 / spec.dx, rdy=1.0 / spec.dy,
                        cf1=cf1, cf2=cf2, cf3=cf3, cfn=cfn, cfn1=cfn1, p_top=p_top))

    return dict(znw=znw, dnw=dnw, rdnw=rdnw, znu=znu, dn=dn, rdn=rdn, fnp=fnp, fnm=fnm,
                c1h=c1h, c2h=c2h, c3h=c3h, c4h=c4h, c1f=c1f, c2f=c2f, c3f=c3f, c4f=c4f)


def _base_profiles(hs: HostState):
    """The base-state profiles of the upper-level Rayleigh damper from the column (ids, jds)
    of the initial state, in REAL arithmetic (module_initialize_ideal.F:1579-1585): u, v,
    qv and the theta perturbation of that column, and the height of its mass levels."""
    d = hs.dims
    i, j = d.ids - d.ims, d.jds - d.jms
    nz = d.kde - d.kds
    f32, g = np.float32, np.float32(9.81)
    A1, A3 = hs.a1, hs.a3
    for n, src in (("t_base", "t_1"), ("u_base", "u_1"), ("v_base", "v_1")):
        A1[n][:nz] = A3[src][j, :nz, i]
    A1["qv_base"][:nz] = hs.q[0][j, :nz, i] if hs.n_species > 0 else 0.0
    phb, ph1 = A3["phb"][j, :, i], A3["ph_1"][j, :, i]
    for k in range(nz):
        A1["z_base"][k] = f32(0.5) * (((phb[k] + phb[k + 1]) + ph1[k]) + ph1[k + 1]) / g


def build_initial_state(spec: DomainSpec) -> HostState:
    """Analytic supercell-type initial state (see the module docstring); a spec with the
    stable sounding goes to build_terrain_state."""
    if spec.sounding == "stable":
        return build_terrain_state(spec)
    if spec.sounding != "wk":
        raise ValueError(f"sounding {spec.sounding!r}: wk or stable")
    hs = HostState(spec)
    d = hs.dims
    kde = d.kde
    nz = spec.nz
    rng = np.random.Generator(np.random.Philox(spec.seed))

    # soundings
    zs, pdry, _, thdry, _ = sounding(moist=False)
    zm, pm, pdm, thm, qvm = sounding(moist=True)
    p_top = float(np.interp(spec.ztop, zs, pdry))

    v = _vertical_coords(hs, spec, p_top)
    znw, dnw, rdnw, znu, dn, rdn = v["znw"], v["dnw"], v["rdnw"], v["znu"], v["dn"], v["rdn"]
    c1h, c2h, c3h, c4h, c1f, c2f, c3f, c4f = (v[n] for n in ("c1h", "c2h", "c3h", "c4h", "c1f", "c2f", "c3f", "c4f"))

    # base state: flat terrain (ht = 0, phb(1) = 0), dry sounding
    p_surf = float(np.interp(0.0, zs, pdry))
    mub = p_surf - p_top
    pb = c3h * (p_surf - p_top) + c4h + p_top
    t_init = _interp_p(thdry, pdry, pb) - T0
    alb = (R_D / P1000) * (t_init + T0) * (pb / P1000) ** CVPM
    phb = np.zeros(kde)
    for kk in range(2, kde + 1):
        k = kk - 1
        phb[kk - 1] = phb[kk - 2] - dnw[kk - 2] * (c1h[k - 1] * mub + c2h[k - 1]) * alb[kk - 2]

    # moist full state, horizontally uniform column
    pd_surf = float(np.interp(0.0, zm, pdm))
    mu1 = pd_surf - p_top - mub
    plev = c3h * (pd_surf - p_top) + c4h + p_top
    qv = _interp_p(qvm, pdm, plev)
    th = _interp_p(thm, pdm, plev) - T0

    ny_m, nx_m = spec.ny, spec.nx
    # 3-D working arrays over mass points, shape (ny, nz, nx), float64
    T = np.broadcast_to(th[None, :, None], (ny_m, nz, nx_m)).copy()
    QV = np.broadcast_to(qv[None, :, None], (ny_m, nz, nx_m)).copy()

    def hydro_column(T, QV):
        """p, alt, al, ph_1 from theta and qv (module_initialize_ideal.F:1000-1085).
        The vertical recurrences run on level-major copies (contiguous level slices)."""
        T = np.ascontiguousarray(T.transpose(1, 0, 2))
        QV = np.ascontiguousarray(QV.transpose(1, 0, 2))
        p = np.zeros_like(T)
        kk = nz - 1
        k = kk + 1  # Fortran kk = kte-1 (top mass level), k = kk + 1
        # top layer: the same qv average as pg_buoy_w's cqw(kde-1) so that the model's
        # top-level vertical pressure gradient balances exactly (the reference's ideal
        # init uses qv(kde-1) alone, module_initialize_ideal.F:1033)
        qvf1 = 0.5 * (QV[kk] + QV[kk - 1])
        qvf2 = 1.0 / (1.0 + qvf1)
        qvf1 = qvf1 * qvf2
        p[kk] = -0.5 * ((c1f[k] * mu1) + qvf1 * (c1f[k] * mub + c2f[k])) / rdnw[kk] / qvf2
        for kk in range(nz - 2, -1, -1):
            k = kk + 1
            qvf1 = 0.5 * (QV[kk] + QV[kk + 1])
            qvf2 = 1.0 / (1.0 + qvf1)
            qvf1 = qvf1 * qvf2
            p[kk] = p[kk + 1] - ((c1f[k] * mu1) + qvf1 * (c1f[k] * mub + c2f[k])) / qvf2 / rdn[kk + 1]
        qvf = 1.0 + RVOVRD * QV
        alt = (R_D / P1000) * (T + T0) * qvf * ((p + pb[:, None, None]) / P1000) ** CVPM
        al = alt - alb[:, None, None]
        # geopotential from the log-pressure hypsometric relation that the model's
        # equation of state inverts (calc_p_rho_phi, hypsometric_opt = 2), so that the
        # start-of-run diagnosis of al and p (start_em.F:794-830) returns this state
        muts = mub + mu1
        pf = c3f * muts + c4f + p_top
        phm = c3h * muts + c4h + p_top
        dlog = np.log(pf[:-1] / pf[1:])
        ph_tot = np.zeros((nz + 1, ny_m, nx_m))
        for k in range(nz):
            ph_tot[k + 1] = ph_tot[k] + alt[k] * phm[k] * dlog[k]
        ph = ph_tot - phb[:, None, None]
        al = (ph_tot[1:] - ph_tot[:-1]) / phm[:, None, None] / dlog[:, None, None] - alb[:, None, None]
        alt = al + alb[:, None, None]
        p = P1000 * ((R_D * (T0 + T * (1.0 + RVOVRD * QV) + T0 * RVOVRD * QV)) / (P1000 * alt)) ** (CP / CV) - pb[:, None, None]
        tr = lambda a: np.ascontiguousarray(a.transpose(1, 0, 2))
        return tr(p), tr(alt), tr(al), tr(ph)

    p, alt, al, ph = hydro_column(T, QV)

    # warm bubble(s): delt*cos^2(pi/2*rad), 10 km / 1.5 km radii, centred at z = 1.5 km
    zmid = 0.5 * (ph[:, :-1, :] + ph[:, 1:, :] + phb[None, :-1, None] + phb[None, 1:, None]) / G
    xs = (np.arange(nx_m) + 0.5) * spec.dx
    ys = (np.arange(ny_m) + 0.5) * spec.dy
    if spec.n_bubbles == 1:
        centres = [(0.5 * nx_m * spec.dx, 0.5 * ny_m * spec.dy, spec.bubble_amp)]
    else:
        centres = [(rng.uniform(0, nx_m * spec.dx), rng.uniform(0, ny_m * spec.dy), spec.bubble_amp)
                   for _ in range(spec.n_bubbles)]
    Lx, Ly = nx_m * spec.dx, ny_m * spec.dy
    for (xc, yc, amp) in centres:
        dxr = xs - xc
        dyr = ys - yc
        if spec.bcx == "periodic":
            dxr = (dxr + 0.5 * Lx) % Lx - 0.5 * Lx
        if spec.bcy == "periodic":
            dyr = (dyr + 0.5 * Ly) % Ly - 0.5 * Ly
        xr = dxr / 10000.0
        yr = dyr / 10000.0
        zr = (zmid - 1500.0) / 1500.0
        rad = np.sqrt(xr[None, None, :] ** 2 + yr[:, None, None] ** 2 + zr ** 2)
        T = T + np.where(rad <= 1.0, amp * np.cos(0.5 * np.pi * rad) ** 2, 0.0)
    # seeded +/- noise on the lowest levels
    if spec.noise_amp > 0:
        nl = min(spec.noise_levels, nz)
        T[:, :nl, :] += rng.uniform(-spec.noise_amp, spec.noise_amp, size=(ny_m, nl, nx_m))
    p, alt, al, ph = hydro_column(T, QV)
    # moist potential temperature (use_theta_m = 1)
    Tm = (T + T0) * (1.0 + RVOVRD * QV) - T0

    # winds from the hodograph at base-state mass-level heights
    zmass = 0.5 * (phb[:-1] + phb[1:]) / G
    uz, vz = hodograph(zmass)
    uz, vz = uz * spec.wind_scale, vz * spec.wind_scale

    # place into memory arrays
    js, is_ = slice(d.jds - d.jms, d.jds - d.jms + ny_m), slice(d.ids - d.ims, d.ids - d.ims + nx_m)
    ks = slice(0, nz)
    A3, A2 = hs.a3, hs.a2

    def put3_mass(name, v, kslice=ks):
        A3[name][js, kslice, is_] = v.astype(np.float32)

    put3_mass("t_2", Tm); put3_mass("t_1", Tm)
    put3_mass("p", p); put3_mass("al", al); put3_mass("alt", alt)
    put3_mass("ph_2", ph, slice(0, nz + 1)); put3_mass("ph_1", ph, slice(0, nz + 1))
    A3["phb"][:, : nz + 1, :] = phb.astype(np.float32)[None, :, None]
    A3["pb"][:, :nz, :] = pb.astype(np.float32)[None, :, None]
    A3["alb"][:, :nz, :] = alb.astype(np.float32)[None, :, None]
    A3["t_init"][:, :nz, :] = t_init.astype(np.float32)[None, :, None]
    A3["u_2"][:, :nz, :] = uz.astype(np.float32)[None, :, None]
    A3["v_2"][:, :nz, :] = vz.astype(np.float32)[None, :, None]
    A3["u_1"][...] = A3["u_2"]
    A3["v_1"][...] = A3["v_2"]
    A2["mub"][...] = np.float32(mub)
    A2["mu_2"][...] = np.float32(mu1)
    A2["mu_1"][...] = np.float32(mu1)
    for n in ("msfux", "msfuy", "msfvx", "msfvx_inv", "msfvy", "msftx", "msfty", "cosa"):
        A2[n][...] = 1.0
    hs.q[0][js, ks, is_] = QV.astype(np.float32)
    # passive scalars: log-normal per point, median 1e-5 kg/kg, sigma 1.0, clipped at 0
    for m in range(1, hs.n_species):
        vals = np.exp(np.log(1.0e-5) + 1.0 * rng.standard_normal((ny_m, nz, nx_m)))
        hs.q[m][js, ks, is_] = np.maximum(vals, 0.0).astype(np.float32)

    # halo fills of everything that carries a halo
    per = (spec.bcx == "periodic", spec.bcy == "periodic")
    stag = {"u_1": (1, 0), "u_2": (1, 0), "v_1": (0, 1), "v_2": (0, 1)}
    for n in ("u_1", "u_2", "v_1", "v_2", "w_1", "w_2", "t_1", "t_2", "ph_1", "ph_2", "phb", "pb", "alb",
              "p", "al", "alt"):
        si, sj = stag.get(n, (0, 0))
        _halo_fill_np(A3[n], d, per, bool(si), bool(sj))
    for m in range(hs.n_species):
        _halo_fill_np(hs.q[m], d, per, False, False)
        hs.q_old[m][...] = hs.q[m]
    for n in ("mub", "mu_1", "mu_2"):
        _halo_fill_np(A2[n], d, per, False, False)
    _base_profiles(hs)
    return hs


def build_terrain_state(spec: DomainSpec) -> HostState:
    """Stably stratified flow over terrain: the stable sounding, a uniform westerly u0 and
    ridge_field terrain, on the hybrid levels of `spec`.  The steps are those of the
    reference's idealised initialisation for its hill cases (module_initialize_ideal.F:
    611-1130): dry base state from the dry sounding with p_surf(ht) and phb(1) = g*ht,
    moist full state from the moist sounding, hydrostatic geopotential.  Every quantity
    varies per column through ht, so the columns are built in blocks of rows, level by
    level in float64, on a thread pool (numpy releases the GIL), and rounded once into the
    float32 arrays.  Each block draws its scalars from its own generator, so the state
    does not depend on the number of threads."""
    import concurrent.futures
    import os

    hs = HostState(spec)
    d = hs.dims
    nz = spec.nz
    rng = np.random.Generator(np.random.Philox(spec.seed))

    zs, pdry, _, thdry, _ = stable_sounding(spec, moist=False)
    zm, _, pdm, thm, qvm = stable_sounding(spec, moist=True)
    p_top = float(np.interp(spec.ztop, zs, pdry))
    v = _vertical_coords(hs, spec, p_top)
    dnw, rdnw, rdn = v["dnw"], v["rdnw"], v["rdn"]
    c1h, c2h, c3h, c4h, c1f, c2f, c3f, c4f = (v[n] for n in ("c1h", "c2h", "c3h", "c4h", "c1f", "c2f", "c3f", "c4f"))
    ht_all = ridge_field(spec, rng)
    A3, A2 = hs.a3, hs.a2
    f32 = np.float32
    i0 = d.ids - d.ims
    is_ = slice(i0, i0 + spec.nx)
    rows = 16
    blocks = [(j, min(j + rows, spec.ny)) for j in range(0, spec.ny, rows)]
    seeds = np.random.SeedSequence([spec.seed, 1]).spawn(len(blocks))

    def build_rows(blk):
        jb0, jb1 = blocks[blk]
        js = slice(d.jds - d.jms + jb0, d.jds - d.jms + jb1)
        ht = ht_all[jb0:jb1]
        # dry base state (module_initialize_ideal.F:905-960, p_surf from the sounding at ht)
        p_surf = np.interp(ht, zs, pdry)
        mub = p_surf - p_top
        pb = [c3h[k] * mub + c4h[k] + p_top for k in range(nz)]
        t_init = [_interp_p(thdry, pdry, pb[k]) - T0 for k in range(nz)]
        alb = [(R_D / P1000) * (t_init[k] + T0) * (pb[k] / P1000) ** CVPM for k in range(nz)]
        phb = [G * ht]
        for k in range(nz):
            phb.append(phb[k] - dnw[k] * (c1h[k] * mub + c2h[k]) * alb[k])
        # moist full state (module_initialize_ideal.F:965-1085)
        pd_surf = np.interp(ht, zm, pdm)
        mu1 = pd_surf - p_top - mub
        plev = [c3h[k] * (pd_surf - p_top) + c4h[k] + p_top for k in range(nz)]
        QV = [_interp_p(qvm, pdm, plev[k]) for k in range(nz)]
        T = [_interp_p(thm, pdm, plev[k]) - T0 for k in range(nz)]
        # first guess of p from the hydrostatic dry-mass relation, top down, giving alt
        alt = [None] * nz
        p_k = None
        for kk in range(nz - 1, -1, -1):
            k = kk + 1
            lo = kk - 1 if kk == nz - 1 else kk + 1
            qvf1 = 0.5 * (QV[kk] + QV[lo])
            qvf2 = 1.0 / (1.0 + qvf1)
            qvf1 = qvf1 * qvf2
            if kk == nz - 1:
                p_k = -0.5 * ((c1f[k] * mu1) + qvf1 * (c1f[k] * mub + c2f[k])) / rdnw[kk] / qvf2
            else:
                p_k = p_k - ((c1f[k] * mu1) + qvf1 * (c1f[k] * mub + c2f[k])) / qvf2 / rdn[kk + 1]
            alt[kk] = (R_D / P1000) * (T[kk] + T0) * (1.0 + RVOVRD * QV[kk]) * ((p_k + pb[kk]) / P1000) ** CVPM
        # geopotential from the hypsometric relation that calc_p_rho_phi inverts, then
        # al, alt and p from it
        muts = mub + mu1
        ph_lo = G * ht
        A3["ph_2"][js, 0, is_] = (ph_lo - phb[0]).astype(f32)
        for k in range(nz):
            pf_lo = c3f[k] * muts + c4f[k] + p_top
            pf_hi = c3f[k + 1] * muts + c4f[k + 1] + p_top
            phm = c3h[k] * muts + c4h[k] + p_top
            dlog = np.log(pf_lo / pf_hi)
            ph_hi = ph_lo + alt[k] * phm * dlog
            al = (ph_hi - ph_lo) / phm / dlog - alb[k]
            altk = al + alb[k]
            tv = T0 + T[k] * (1.0 + RVOVRD * QV[k]) + T0 * RVOVRD * QV[k]
            pk = P1000 * ((R_D * tv) / (P1000 * altk)) ** (CP / CV) - pb[k]
            A3["ph_2"][js, k + 1, is_] = (ph_hi - phb[k + 1]).astype(f32)
            A3["al"][js, k, is_] = al.astype(f32)
            A3["alt"][js, k, is_] = altk.astype(f32)
            A3["p"][js, k, is_] = pk.astype(f32)
            A3["t_2"][js, k, is_] = ((T[k] + T0) * (1.0 + RVOVRD * QV[k]) - T0).astype(f32)
            A3["pb"][js, k, is_] = pb[k].astype(f32)
            A3["alb"][js, k, is_] = alb[k].astype(f32)
            A3["t_init"][js, k, is_] = t_init[k].astype(f32)
            A3["phb"][js, k, is_] = phb[k].astype(f32)
            hs.q[0][js, k, is_] = QV[k].astype(f32)
            ph_lo = ph_hi
        A3["phb"][js, nz, is_] = phb[nz].astype(f32)
        A2["ht"][js, is_] = ht.astype(f32)
        A2["mub"][js, is_] = mub.astype(f32)
        A2["mu_2"][js, is_] = mu1.astype(f32)
        # passive scalars: log-normal per point, median 1e-5 kg/kg, sigma 1.0 (positive,
        # so the clip at 0 is implicit), drawn in float32
        g = np.random.Generator(np.random.Philox(seeds[blk]))
        for m in range(1, hs.n_species):
            z = g.standard_normal((jb1 - jb0, nz, spec.nx), dtype=np.float32)
            np.exp(z, out=z)
            z *= f32(1.0e-5)
            hs.q[m][js, :nz, is_] = z

    per = (spec.bcx == "periodic", spec.bcy == "periodic")
    stag = {"u_1": (1, 0), "u_2": (1, 0), "v_1": (0, 1), "v_2": (0, 1)}

    def finish(name):
        """Copies of the two time levels, then the halo fill of one array."""
        if name in ("t_1", "ph_1"):
            A3[name][...] = A3[name.replace("_1", "_2")]
        elif name in ("u_1", "u_2"):
            A3[name][:, :nz, :] = f32(spec.u0)
        elif name == "mu_1":
            A2["mu_1"][...] = A2["mu_2"]
        elif name.startswith("msf") or name == "cosa":
            A2[name][...] = 1.0
            return
        if name.startswith("q") and name[1:].isdigit():
            m = int(name[1:])
            _halo_fill_np(hs.q[m], d, per, False, False)
            hs.q_old[m][...] = hs.q[m]
        elif name in A2:
            _halo_fill_np(A2[name], d, per, False, False)
        else:
            si, sj = stag.get(name, (0, 0))
            _halo_fill_np(A3[name], d, per, bool(si), bool(sj))

    workers = max(1, min(16, os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        list(pool.map(build_rows, range(len(blocks))))
        names = ["u_1", "u_2", "v_1", "v_2", "t_1", "t_2", "ph_1", "ph_2", "phb", "pb", "alb", "t_init", "p", "al", "alt",
                 "mub", "mu_1", "mu_2", "ht", "msfux", "msfuy", "msfvx", "msfvx_inv", "msfvy", "msftx", "msfty",
                 "cosa"] + [f"q{m}" for m in range(hs.n_species)]
        # the time-level copies read halo-free sources: run them before their sources' fills
        first = ["t_1", "ph_1", "mu_1"]
        list(pool.map(finish, first))
        list(pool.map(finish, [n for n in names if n not in first]))
    # start_em.F:1480-1531: w of both time levels from the kinematic surface condition
    for w, u, v in (("w_1", "u_1", "v_1"), ("w_2", "u_2", "v_2")):
        set_w_surface_np(hs, A3[w], A3[u], A3[v], fill=True)
        _halo_fill_np(A3[w], d, per, False, False)
    _base_profiles(hs)
    return hs


def set_w_surface_np(hs: HostState, w, u, v, fill: bool):
    """set_w_surface (dyn_em/module_bc_em.F:1196-1295) in float32, as start_em calls it
    (start_em.F:1520-1531): w(1) = msfty*.5*rdy*(dht_n*v_n + dht_s*v_s) + msftx*.5*rdx*(...)
    with u, v extrapolated to the surface by cf1..cf3, a neighbour across a non-periodic edge
    being the point itself; with `fill`, w(k) = w(1)*znw(k)*znw(k) above.  Interior only."""
    d, c = hs.dims, hs.cfg
    f32 = np.float32
    cf1, cf2, cf3 = (f32(hs.scal[n]) for n in ("cf1", "cf2", "cf3"))
    rdx, rdy = f32(hs.scal["rdx"]), f32(hs.scal["rdy"])
    i = np.arange(d.ids, d.ide)                 # i = ids..ide-1
    j = np.arange(d.jds, d.jde)
    im1 = np.maximum(i - 1, d.ids - 1 if c.periodic_x else d.ids) - d.ims
    ip1 = np.minimum(i + 1, d.ide if c.periodic_x else d.ide - 1) - d.ims
    jm1 = np.maximum(j - 1, d.jds - 1 if c.periodic_y else d.jds) - d.jms
    jp1 = np.minimum(j + 1, d.jde if c.periodic_y else d.jde - 1) - d.jms
    ii, jj = i - d.ims, j - d.jms
    ht, msftx, msfty = hs.a2["ht"], hs.a2["msftx"], hs.a2["msfty"]
    h0 = ht[np.ix_(jj, ii)]
    srf = lambda a, jsel, isel: cf1 * a[jsel, 0][:, isel] + cf2 * a[jsel, 1][:, isel] + cf3 * a[jsel, 2][:, isel]
    vs_n = srf(v, jj + 1, ii)
    vs_s = srf(v, jj, ii)
    us_e = srf(u, jj, ii + 1)
    us_w = srf(u, jj, ii)
    w1 = (msfty[np.ix_(jj, ii)] * f32(.5) * rdy * ((ht[np.ix_(jp1, ii)] - h0) * vs_n + (h0 - ht[np.ix_(jm1, ii)]) * vs_s) +
          msftx[np.ix_(jj, ii)] * f32(.5) * rdx * ((ht[np.ix_(jj, ip1)] - h0) * us_e + (h0 - ht[np.ix_(jj, im1)]) * us_w))
    w[jj[0]:jj[-1] + 1, 0, ii[0]:ii[-1] + 1] = w1
    if fill:
        znw = hs.a1["znw"]
        for k in range(1, d.kde - d.kms + 1):   # Fortran k = kts+1..kte
            w[jj[0]:jj[-1] + 1, k, ii[0]:ii[-1] + 1] = w1 * znw[k] * znw[k]


# ----------------------------------------------------------------------------- presets
def supercell_c0(**kw) -> DomainSpec:
    """41x41x40 mass points, dx=2 km, dt=12 s, 6 sound steps, open boundaries (seed 1001)."""
    base = dict(nx=41, ny=41, nz=40, dx=2000.0, dy=2000.0, dt=12.0, ztop=20000.0, time_step_sound=6,
                bc="open", n_bubbles=1, bubble_amp=3.0, noise_amp=0.1, seed=1001)
    base.update(kw)
    return DomainSpec(**base)


def large_c1(**kw) -> DomainSpec:
    """400x400x80, dx=1 km, dt=6 s, doubly periodic, 6 bubbles, qv + 6 passive scalars (seed 1002)."""
    base = dict(nx=400, ny=400, nz=80, dx=1000.0, dy=1000.0, dt=6.0, ztop=20000.0, time_step_sound=6,
                bc="periodic", n_bubbles=6, bubble_amp=3.0, noise_amp=0.1, n_passive=6, seed=1002)
    base.update(kw)
    return DomainSpec(**base)


ENSEMBLE_SIZES = (64, 96, 128, 192, 256)


def ensemble_c3(n_members: int = 64, seed: int = 1004, **kw) -> list:
    """Uneven ensemble: member m has nx, ny drawn uniformly from ENSEMBLE_SIZES (generator
    seeded with `seed`), nz = 50, dx = 2 km, dt = 12 s, the 41x41x40 case's sounding and open
    boundaries, a bubble amplitude drawn from U[1, 4] K and +/-0.2 K theta noise, with its
    own seed seed*1000 + m.  Returns the members' DomainSpecs."""
    rng = np.random.Generator(np.random.Philox(seed))
    out = []
    for m in range(n_members):
        nx, ny = (int(v) for v in rng.choice(ENSEMBLE_SIZES, size=2))
        amp = float(rng.uniform(1.0, 4.0))
        base = dict(nx=nx, ny=ny, nz=50, dx=2000.0, dy=2000.0, dt=12.0, ztop=20000.0, time_step_sound=6,
                    bc="open", n_bubbles=1, bubble_amp=amp, noise_amp=0.2, seed=seed * 1000 + m)
        base.update(kw)
        out.append(DomainSpec(**base))
    return out


def terrain_c4(**kw) -> DomainSpec:
    """1200x1200x80 on hybrid levels (hybrid_opt 2), dx=500 m, dt=3 s, 6 sound steps,
    periodic in y and open in x, 12 Witch-of-Agnesi ridges (half-width 5-20 km, height
    200-1500 m), uniform 10 m/s westerly, N = 0.01 1/s, qv + 10 passive scalars (seed 1005)."""
    base = dict(nx=1200, ny=1200, nz=80, dx=500.0, dy=500.0, dt=3.0, ztop=20000.0, time_step_sound=6,
                bc="open", bc_x="open", bc_y="periodic", hybrid_opt=2, sounding="stable", n_ridges=12,
                ridge_a=(5000.0, 20000.0), ridge_h=(200.0, 1500.0), u0=10.0, brunt=0.01, n_passive=10,
                n_bubbles=0, noise_amp=0.0, seed=1005)
    base.update(kw)
    return DomainSpec(**base)


# ----------------------------------------------------------------------------- lateral boundaries
LBC_FIELDS = ("u", "v", "ph", "t", "w", "mu")
_SIDES = ("xs", "xe", "ys", "ye")


def lbc_fcx_gcx(spec_bdy_width, spec_zone, relax_zone, dt, spec_exp=0.0):
    """The relaxation weights of lbc_fcx_gcx (dyn_em/module_bc_em.F:1297-1341, specified),
    in REAL arithmetic; entries 0..spec_zone-1 (the spec zone) stay zero."""
    f32 = np.float32
    fcx = np.zeros(16, f32)
    gcx = np.zeros(16, f32)
    for loop in range(spec_zone + 1, spec_zone + relax_zone + 1):
        n, m = f32(spec_zone + relax_zone - loop), f32(relax_zone - 1)
        w = f32(np.exp(f32(-(loop - (spec_zone + 1))) * f32(spec_exp)))
        fcx[loop - 1] = (f32(0.1) / f32(dt) * n / m) * w
        gcx[loop - 1] = (f32(1.0) / f32(dt) / f32(50.) * n / m) * w
    return fcx, gcx


def build_lbc(hs, seed: int = 0, value_noise: float = 1e-3, tend_scale: float = 2e-4, dtype=np.float32):
    """Synthetic lateral boundary arrays for a specified domain: per field the coupled
    values of the initial state's boundary rows (u, v times c1 mu + c2, ph, t and w times
    c1 mu + c2, mu) perturbed by a relative `value_noise`, and tendencies of a relative
    `tend_scale` per second (seeded).  The species with boundary arrays (QV, and the others
    under have_bcs_moist / have_bcs_scalar) take theirs the same way.  The products are
    formed in `dtype` (float64 for an FP64 oracle, where they then equal its own mass
    weighting exactly).  Sets hs.lbc."""
    d, cfg = hs.dims, hs.cfg
    w = int(cfg.spec_bdy_width)
    rng = np.random.default_rng(seed)
    f32 = np.float32
    A3 = {k: v.astype(dtype) for k, v in hs.a3.items() if k in ("u_2", "v_2", "ph_2", "w_2", "t_2")}
    A2 = {k: hs.a2[k].astype(dtype) for k in ("mub", "mu_2")}
    A1 = {k: hs.a1[k].astype(dtype) for k in ("c1h", "c2h", "c1f", "c2f")}
    nkb = d.kde - d.kds + 1
    ci, cj = d.ids - d.ims, d.jds - d.jms
    mut = A2["mub"] + A2["mu_2"]
    c1h, c2h, c1f, c2f = (A1[n] for n in ("c1h", "c2h", "c1f", "c2f"))

    def coupled(name):
        if name == "u":
            return A3["u_2"] * (c1h[None, :, None] * mut[:, None, :] + c2h[None, :, None])
        if name == "v":
            return A3["v_2"] * (c1h[None, :, None] * mut[:, None, :] + c2h[None, :, None])
        if name in ("ph", "w"):
            return A3[name + "_2"] * (c1f[None, :, None] * mut[:, None, :] + c2f[None, :, None])
        if name == "t":
            return A3["t_2"] * (c1h[None, :, None] * mut[:, None, :] + c2h[None, :, None])
        if name == "mu":
            return A2["mu_2"][:, None, :]
        m = int(name[1:])
        return hs.q[m].astype(dtype) * (c1h[None, :, None] * mut[:, None, :] + c2h[None, :, None])

    def sides(name):
        full = coupled(name)  # (nj, nk, ni)
        nk = 1 if name == "mu" else nkb
        ibe = d.ide - d.ims if name == "u" else d.ide - 1 - d.ims
        jbe = d.jde - d.jms if name == "v" else d.jde - 1 - d.jms
        vals, tends = [], []
        for s_ in range(4):
            if s_ < 2:
                cols = [ci + b if s_ == 0 else ibe - b for b in range(w)]
                a = np.stack([full[:, :nk, c] for c in cols])  # (w, nj, nk)
                a = a.transpose(0, 2, 1)                       # (w, nk, nj)
            else:
                rows = [cj + b if s_ == 2 else jbe - b for b in range(w)]
                a = np.stack([full[r, :nk, :] for r in rows])  # (w, nk, ni)
            scale = np.abs(a).max() or 1.0
            if value_noise:
                a = a * (1.0 + value_noise * rng.standard_normal(a.shape))
            vals.append(np.ascontiguousarray(a.astype(dtype)))
            tends.append(np.ascontiguousarray((tend_scale * scale * rng.standard_normal(a.shape)).astype(dtype)))
        return vals, tends

    lbc = {"dtbc": np.float32(0.0)}
    lbc["fcx"], lbc["gcx"] = lbc_fcx_gcx(w, int(cfg.spec_zone), int(cfg.relax_zone), hs.scal["dt"])
    for n in LBC_FIELDS:
        lbc[n + "_b"], lbc[n + "_bt"] = sides(n)
    lbc["q_b"] = [None] * hs.n_species
    lbc["q_bt"] = [None] * hs.n_species
    for m in range(hs.n_species):
        moist = m < hs.n_moist
        if cfg.nested or (m == 0 and moist) or (cfg.have_bcs_moist if moist else cfg.have_bcs_scalar):
            lbc["q_b"][m], lbc["q_bt"][m] = sides(f"q{m}")
    hs.lbc = lbc
    return lbc


def lbc_arrays(lbc) -> dict:
    """The boundary arrays of `lbc` by flat name (lbc_<field>_b<side>, lbc_q<m>_b<side>)."""
    out = {}
    for n in LBC_FIELDS:
        for s_ in range(4):
            out[f"lbc_{n}_b{s_}"] = lbc[n + "_b"][s_]
            out[f"lbc_{n}_bt{s_}"] = lbc[n + "_bt"][s_]
    for m, sides in enumerate(lbc["q_b"]):
        if sides is not None:
            for s_ in range(4):
                out[f"lbc_q{m}_b{s_}"] = sides[s_]
                out[f"lbc_q{m}_bt{s_}"] = lbc["q_bt"][m][s_]
    return out


def lbc_struct(lbc, struct_type, ptr, ctype=ctypes.c_float):
    """A ctypes wrfmi_lbc (or oracle olbc) for `lbc` whose arrays are ptr(flat name)."""
    import ctypes as ct
    L = struct_type()
    L.dtbc = float(lbc["dtbc"])
    for k in range(16):
        L.fcx[k] = float(lbc["fcx"][k])
        L.gcx[k] = float(lbc["gcx"][k])
    P = ct.POINTER(ctype)
    for n in LBC_FIELDS:
        for s_ in range(4):
            getattr(L, n + "_b")[s_] = ct.cast(ptr(f"lbc_{n}_b{s_}"), P)
            getattr(L, n + "_bt")[s_] = ct.cast(ptr(f"lbc_{n}_bt{s_}"), P)
    for m, sides in enumerate(lbc["q_b"]):
        if sides is not None:
            for s_ in range(4):
                L.q_b[m][s_] = ct.cast(ptr(f"lbc_q{m}_b{s_}"), P)
                L.q_bt[m][s_] = ct.cast(ptr(f"lbc_q{m}_bt{s_}"), P)
    return L

