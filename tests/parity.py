"""Shared comparison helpers of the parity tests (test infrastructure)."""
from __future__ import annotations

import numpy as np

# The reference evaluates LOG and ** (vspow) in FP32 libm; the GPU evaluates them in
# double and rounds once.  Against the libm oracle, outputs of calc_p_rho_phi may differ by
# a few ulp of the full (base + perturbation) value at isolated points: 1 ulp of LOG moves
# alpha by up to ~1 ulp, which ** (exponent cp/cv = 1.4) and the product with p0 carry into
# p at up to ~8 ulp.  Against the once-rounded oracle the bound is the rare case where a
# double result rounds differently to float on the two sides.
EOS_FIELDS = {"p": "pb", "al": "alb"}
EOS_ULPS = 8


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sqrt(np.sum(b * b))
    num = np.sqrt(np.sum((a - b) ** 2))
    return float(num / den) if den > 0 else float(num)


def interior(a, d, kind="mass", nk=None):
    """Domain part of a (nj, nk, ni) or (nj, ni) array: i = ids..ide(-1), j = jds..jde(-1)."""
    i0, j0 = d.ids - d.ims, d.jds - d.jms
    ni = d.ide - d.ids + (1 if kind == "u" else 0)
    nj = d.jde - d.jds + (1 if kind == "v" else 0)
    if a.ndim == 3:
        nkk = nk if nk is not None else (d.kde if kind == "w" else d.kde - 1)
        return a[j0:j0 + nj, :nkk, i0:i0 + ni]
    return a[j0:j0 + nj, i0:i0 + ni]


def compare_exact(got: dict, exp: dict, names, eos_tol: bool = False, base: dict | None = None, ulps: int = EOS_ULPS):
    """Bitwise comparison; returns a list of failure descriptions."""
    fails = []
    for k in names:
        g, e = got[k], exp[k]
        if g.shape != e.shape:
            fails.append(f"{k}: shape {g.shape} vs {e.shape}")
            continue
        bad = g.view(np.uint8).reshape(g.shape + (-1,)) != e.view(np.uint8).reshape(e.shape + (-1,))
        bad = bad.any(axis=-1)
        # +0 and -0 compare equal in value; keep bitwise strictness except for signed zeros
        bad &= ~((g == 0) & (e == 0))
        if eos_tol and k in EOS_FIELDS and base is not None:
            full = np.abs(e.astype(np.float64) + base[EOS_FIELDS[k]].astype(np.float64))
            ulp = np.spacing(full.astype(np.float32)).astype(np.float64)
            bad &= np.abs(g.astype(np.float64) - e.astype(np.float64)) > ulps * ulp
        n = int(bad.sum())
        if n:
            idx = tuple(int(x[0]) for x in np.nonzero(bad))
            fails.append(f"{k}: {n} values differ, first at [j,k,i]={idx} got {g[idx]!r} expected {e[idx]!r}, "
                         f"max|diff|={float(np.max(np.abs(g.astype(np.float64) - e.astype(np.float64))[bad])):.3e}")
    return fails


def phase_parity(backend, session, phases, nsteps: int = 1, stop_at_first: bool = True, eos_ulps: int = EOS_ULPS):
    """Run the per-routine calls on the backend and on the oracle from identical inputs
    (the oracle's state is uploaded before every call) and compare every array bitwise.
    Returns a list of failure descriptions."""
    names = list(session.arrays)
    fails = []
    for step in range(nsteps):
        for n, (name, args) in enumerate(phases):
            backend.upload(session.arrays)
            backend.phase(name, *args)
            session.phase(name, *args)
            got = backend.download()
            f = compare_exact(got, session.arrays, names, eos_tol=(name == "calc_p_rho_phi"), base=session.arrays,
                              ulps=eos_ulps)
            if f:
                fails.append(f"step {step + 1} call {n} {name}{args}: " + "; ".join(f))
                if stop_at_first:
                    return fails
    return fails


def domain_totals_numpy(arr: dict, hs) -> np.ndarray:
    """The domain totals of wrfmi_domain_totals (dry mass, theta budget, kinetic, internal,
    potential and total energy) computed in float64 with numpy from the host arrays `arr`
    (the state's fields by name) of a domain with dims and scalars of `hs`."""
    d = hs.dims
    g, rd = 9.81, 287.0
    cp = 7.0 * rd / 2.0
    cv = cp - rd
    i0, j0 = d.ids - d.ims, d.jds - d.jms
    nx, ny, nk = d.ide - d.ids, d.jde - d.jds, d.kde - 1
    f = {k: np.asarray(v, np.float64) for k, v in arr.items()}
    col = lambda a: a[j0:j0 + ny, i0:i0 + nx]  # noqa: E731
    lev = lambda a, dk=0, di=0, dj=0: a[j0 + dj:j0 + dj + ny, dk:dk + nk, i0 + di:i0 + di + nx]  # noqa: E731
    mut = col(f["mub"]) + col(f["mu_2"])
    area = hs.scal["dx"] * hs.scal["dy"] / (col(f["msftx"]) * col(f["msfty"]))
    c1h, c2h, dnw = f["c1h"][:nk], f["c2h"][:nk], f["dnw"][:nk]
    dm = (c1h[None, :, None] * mut[:, None, :] + c2h[None, :, None]) * (-dnw[None, :, None]) / g
    theta = lev(f["t_2"]) + 300.0
    thd = theta / (1.0 + (461.6 / 287.0) * lev(f["q0"])) if (hs.cfg.use_theta_m and hs.n_moist >= 1) else theta
    temp = thd * ((lev(f["p"]) + lev(f["pb"])) / 1.0e5) ** (rd / cp)
    ub = 0.5 * (lev(f["u_2"]) + lev(f["u_2"], di=1))
    vb = 0.5 * (lev(f["v_2"]) + lev(f["v_2"], dj=1))
    wb = 0.5 * (lev(f["w_2"]) + lev(f["w_2"], dk=1))
    phi = f["ph_2"] + f["phb"]
    pe = dm * 0.5 * (lev(phi) + lev(phi, dk=1))
    a3 = area[:, None, :]
    out = np.array([np.sum(dm * a3), np.sum(dm * theta * a3), np.sum(dm * 0.5 * (ub ** 2 + vb ** 2 + wb ** 2) * a3),
                    np.sum(dm * cv * temp * a3), np.sum(pe * a3), 0.0])
    out[5] = out[2] + out[3] + out[4]
    return out


MSF6 = ("msfux", "msfuy", "msfvx", "msfvy", "msftx", "msfty")


def pd_advect_inputs(bc: str, nx: int, ny: int, nz: int, seed: int, same_old: bool = False):
    """Seeded inputs of advect_scalar_pd on a small domain: a non-negative tracer in blobs
    with many zero cells (so that the limiter engages at the blob edges), mass fluxes at
    Courant numbers of about a tenth, masses of a real column.  All halos are filled as the
    model fills them (periodic wrap or zero-gradient copy), so a periodic case conserves.
    Returns (dims, config, arrays) with float32 arrays and the float32 scalars rdx, rdy, dt."""
    from wrf_amd import grid
    spec = grid.DomainSpec(nx=nx, ny=ny, nz=nz, bc=bc, pad_i=1, seed=seed)
    d, c = grid.make_dims(spec), grid.make_config(spec)
    per = bc == "periodic"
    s3, s2, s1 = grid.shape3(d), grid.shape2(d), grid.shape1(d)
    rng = np.random.default_rng(seed)
    u = lambda sh, lo, hi: rng.uniform(lo, hi, sh).astype(np.float32)  # noqa: E731

    def blobs():
        q = np.maximum(0.0, rng.normal(0.0, 1.0, s3)) * (rng.uniform(0, 1, s3) < 0.45)
        return (1e-3 * q).astype(np.float32)

    # ru = mu u / msf with u up to 17 m/s; rom = mu deta/dt with w of a few m/s
    a = dict(field=blobs(), field_old=blobs(), ru=u(s3, -1.5e6, 1.5e6), rv=u(s3, -1.5e6, 1.5e6), rom=u(s3, -40.0, 40.0),
             mut=u(s2, 8.8e4, 9.2e4), mub=u(s2, 8.6e4, 9.0e4), mu_old=u(s2, 1.5e3, 2.5e3), c1=u(s1, 0.6, 1.0),
             c2=u(s1, 0.0, 5e2), fzm=u(s1, 0.4, 0.6), rdzw=u(s1, -32.0, -24.0))
    if same_old:
        a["field_old"] = a["field"].copy()
    a["fzp"] = (1 - a["fzm"]).astype(np.float32)
    for m in MSF6:
        a[m] = u(s2, 0.97, 1.03)
    for k, (si, sj) in dict(field=(0, 0), field_old=(0, 0), ru=(1, 0), rv=(0, 1), rom=(0, 0), mut=(0, 0), mub=(0, 0),
                            mu_old=(0, 0), msfux=(1, 0), msfuy=(1, 0), msfvx=(0, 1), msfvy=(0, 1), msftx=(0, 0),
                            msfty=(0, 0)).items():
        grid._halo_fill_np(a[k], d, per, bool(si), bool(sj))
    a["rdx"], a["rdy"], a["dt"] = np.float32(1 / 1000.), np.float32(1 / 1250.), np.float32(6.0)
    return d, c, a
