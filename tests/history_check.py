"""The history file (wrfmi_history_*) read back with scipy's independent netCDF reader and
compared with the state it was written from -- test infrastructure, shared by the
CPU-emulated library's check and the GPU test.

Dimension and variable names, staggered extents, the per-variable attributes and the global
attributes are those of the reference's wrfout files (Registry/Registry.EM_COMMON names,
external/io_netcdf/wrf_io.F90 attributes, share/output_wrf.F globals); every value must equal
the state's interior bit for bit (T: the dry theta diagnosed from the moist one in float32).
"""
from __future__ import annotations

import numpy as np

DIMS = {"DateStrLen": 19}
FIELD_VARS = {  # wrfout name: (state array, stagger) -- 3-D
    "U": ("u_2", "X"), "V": ("v_2", "Y"), "W": ("w_2", "Z"), "PH": ("ph_2", "Z"), "PHB": ("phb", "Z"),
    "THM": ("t_2", ""), "P": ("p", ""), "PB": ("pb", ""),
}
FIELD_VARS_2D = {"MU": ("mu_2", ""), "MUB": ("mub", ""), "HGT": ("ht", ""), "MAPFAC_MX": ("msftx", ""),
                 "MAPFAC_MY": ("msfty", ""), "MAPFAC_UX": ("msfux", "X"), "MAPFAC_UY": ("msfuy", "X"),
                 "MAPFAC_VX": ("msfvx", "Y"), "MAPFAC_VY": ("msfvy", "Y")}
LEVEL_VARS = {"ZNU": ("znu", False), "ZNW": ("znw", True), "C1H": ("c1h", False), "C4F": ("c4f", True),
              "RDNW": ("rdnw", False)}


def _interior(a, d, stag, rank):
    nx, ny, nz = d.ide - d.ids, d.jde - d.jds, d.kde - d.kds
    j0, i0, k0 = d.jds - d.jms, d.ids - d.ims, d.kds - d.kms
    ex, ey, ez = (1 if stag == "X" else 0), (1 if stag == "Y" else 0), (1 if stag == "Z" else 0)
    if rank == 3:
        return np.ascontiguousarray(a[j0:j0 + ny + ey, k0:k0 + nz + ez, i0:i0 + nx + ex].transpose(1, 0, 2))
    return np.ascontiguousarray(a[j0:j0 + ny + ey, i0:i0 + nx + ex])


def history_failures(path, records, d, cfg, scal, n_moist, n_species, start_date):
    """`records`: [(date, xtime, arrays)] as written, arrays the state's host copy at the
    time (keys as DeviceDomain.download's)."""
    from scipy.io import netcdf_file
    fails = []
    nx, ny, nz = d.ide - d.ids, d.jde - d.jds, d.kde - d.kds
    with netcdf_file(path, "r", mmap=False) as f:
        want = {"Time": None, "DateStrLen": 19, "west_east": nx, "south_north": ny, "bottom_top": nz,
                "bottom_top_stag": nz + 1, "west_east_stag": nx + 1, "south_north_stag": ny + 1}
        for k, v in want.items():
            if f.dimensions.get(k, "missing") != v:
                fails.append(f"dimension {k}: {f.dimensions.get(k, 'missing')} (expected {v})")
        ga = f._attributes
        for k, v in (("WEST-EAST_GRID_DIMENSION", nx + 1), ("SOUTH-NORTH_GRID_DIMENSION", ny + 1),
                     ("BOTTOM-TOP_GRID_DIMENSION", nz + 1), ("DIFF_OPT", cfg.diff_opt), ("DAMP_OPT", cfg.damp_opt),
                     ("DIFF_6TH_OPT", cfg.diff_6th_opt), ("MOIST_ADV_OPT", cfg.moist_adv_opt),
                     ("USE_THETA_M", cfg.use_theta_m)):
            if int(np.asarray(ga.get(k, -999)).ravel()[0]) != v:
                fails.append(f"global attribute {k}: {ga.get(k)} (expected {v})")
        for k, v in (("DX", scal["dx"]), ("DY", scal["dy"]), ("DT", scal["dt"])):
            if np.float32(np.asarray(ga.get(k, np.nan)).ravel()[0]) != np.float32(v):
                fails.append(f"global attribute {k}: {ga.get(k)} (expected {v})")
        if ga.get("START_DATE", b"").decode() != start_date:
            fails.append(f"START_DATE {ga.get('START_DATE')}")
        nrec = len(records)
        t = f.variables["Times"]
        if t.dimensions != ("Time", "DateStrLen") or t.shape[0] != nrec:
            fails.append(f"Times: {t.dimensions} {t.shape}")
        for r, (date, xtime, arr) in enumerate(records):
            if b"".join(t[r]).decode() != date:
                fails.append(f"Times[{r}] = {b''.join(t[r])!r}")
            if f.variables["XTIME"][r] != np.float32(xtime):
                fails.append(f"XTIME[{r}] = {f.variables['XTIME'][r]}")
            if f.variables["P_TOP"][r] != np.float32(scal["p_top"]):
                fails.append(f"P_TOP[{r}] = {f.variables['P_TOP'][r]}")
            for name, (src, stag) in FIELD_VARS.items():
                v = f.variables[name]
                exp = _interior(arr[src], d, stag, 3)
                dims = ("Time", "bottom_top_stag" if stag == "Z" else "bottom_top",
                        "south_north_stag" if stag == "Y" else "south_north",
                        "west_east_stag" if stag == "X" else "west_east")
                if v.dimensions != dims:
                    fails.append(f"{name}: dimensions {v.dimensions}")
                    continue
                got = np.asarray(v[r], dtype=np.float32)
                if not np.array_equal(got.view(np.int32), exp.view(np.int32)):
                    fails.append(f"{name}[{r}]: {int(np.count_nonzero(got != exp))} values differ")
                if r == 0:
                    a = v._attributes
                    bad = int(a.get("FieldType", 0)) != 104 or a.get("MemoryOrder") != b"XYZ"
                    if bad or a.get("stagger", b"").decode() != stag:
                        fails.append(f"{name}: attributes {a}")
            for name, (src, stag) in FIELD_VARS_2D.items():
                exp = _interior(arr[src], d, stag, 2)
                got = np.asarray(f.variables[name][r], dtype=np.float32)
                if got.shape != exp.shape or not np.array_equal(got.view(np.int32), exp.view(np.int32)):
                    fails.append(f"{name}[{r}] differs")
            for name, (src, stag) in LEVEL_VARS.items():
                exp = np.asarray(arr[src], dtype=np.float32)[d.kds - d.kms:d.kds - d.kms + nz + (1 if stag else 0)]
                got = np.asarray(f.variables[name][r], dtype=np.float32)
                if not np.array_equal(got, exp):
                    fails.append(f"{name}[{r}] differs")
            th = _interior(arr["t_2"], d, "", 3)
            if cfg.use_theta_m and n_moist > 0:
                qv = _interior(arr["q0"], d, "", 3)
                t0, rvovrd = np.float32(300.0), np.float32(461.6) / np.float32(287.0)
                th = (th + t0) / (np.float32(1.0) + rvovrd * qv) - t0
            got = np.asarray(f.variables["T"][r], dtype=np.float32)
            if not np.array_equal(got, th.astype(np.float32)):
                fails.append(f"T[{r}]: {int(np.count_nonzero(got != th))} values differ")
            for m in range(n_species):
                name = "QVAPOR" if (m == 0 and n_moist > 0) else (f"Q{m + 1}" if m < n_moist
                                                                   else f"SCALAR{m - n_moist + 1}")
                got = np.asarray(f.variables[name][r], dtype=np.float32)
                exp = _interior(arr[f"q{m}"], d, "", 3)
                if not np.array_equal(got.view(np.int32), exp.view(np.int32)):
                    fails.append(f"{name}[{r}] differs")
    return fails
