"""Python host of the MI355X dynamics step: loads ``wrf_amd/libwrfmi.so`` and drives it
through its C ABI (include/wrfmi.h) with device memory owned by torch.

The device-resident domain mirrors the reference's ``grid`` (frame/module_domain.F): one
device array per grid%<name> field in the Fortran layout (ims:ime, kms:kme, jms:jme),
i fastest.  ``DeviceDomain.step()`` is the counterpart of one call of ``solve_em``
(dyn_em/solve_em.F:1-4540) for the implemented options.

There is no CPU fallback: if the library is missing or the device is absent, the
calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .fields import FIELDS_1D, FIELDS_2D, FIELDS_3D, SCALARS_R, Config, Dims, Lbc32, State32

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwrfmi.so")

WRFMI_OK = 0
WRFMI_E_BADOPT = 1
WRFMI_E_DIMS = 2
WRFMI_E_DEVICE = 3
WRFMI_E_POINTER = 4
WRFMI_E_NONFINITE = 5
WRFMI_E_IO = 6


class WrfmiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"wrfmi error {code}: {msg}")
        self.code = code


# per-routine entry points: name -> types of the arguments between dims and stream
PHASE_ARGS = {
    "rk_step_prep": "i",          # rk_step
    "rk_tendency": "i",           # rk_step
    "small_step_prep": "if",      # rk_step, dts
    "advance_uv_state": "f",      # dts
    "advance_mu_t_w": "fii",      # dts, iteration, n_small
    "small_step_finish": "iif",   # rk_step, n_small, dts
    "rk_scalar_step": "iif",      # ispec, rk_step, dt_rk
    "rk_scalar_tend": "iif",      # ispec, rk_step, dt_rk
    "rk_scalar_pd_prep": "f",     # dt_rk
    "calc_p_rho_phi": "",
    "phys_bc_dry_2": "i",         # kind
    "set_w_surface": "",
}


def num_sound_steps(cfg, dt: float, dx: float, dy: float) -> int:
    """Acoustic steps per time step (dyn_em/solve_em.F:446-463)."""
    if cfg.time_step_sound != 0:
        return cfg.time_step_sound
    f32 = np.float32
    n = 2 * (int(f32(300.0) * f32(dt) / f32(min(dx, dy)) - f32(0.01)) + 1)
    return max(n, 4)


def pd_species(cfg, n_moist: int, m: int) -> bool:
    """Species m takes the positive-definite transport on the last substep: moist species
    (m < n_moist) follow moist_adv_opt, the others scalar_adv_opt (1 = positive definite)."""
    return (cfg.moist_adv_opt if m < n_moist else cfg.scalar_adv_opt) == 1


def step_phases(cfg, scal: dict, n_species: int, n_moist: int = 1):
    """The per-routine calls of one RK3 time step, in solve_em order, as (name, args) with
    FP32 substep lengths computed as the reference computes them (solve_em.F:601-615)."""
    f32 = np.float32
    dt = f32(scal["dt"])
    ns = num_sound_steps(cfg, scal["dt"], scal["dx"], scal["dy"])
    out = []
    for rk in (1, 2, 3):
        if rk == 1:
            dt_rk = dt / f32(3.0); dts = dt_rk; nsmall = 1
        elif rk == 2:
            dt_rk = f32(0.5) * dt; dts = dt / f32(ns); nsmall = ns // 2
        else:
            dt_rk = dt; dts = dt / f32(ns); nsmall = ns
        out.append(("rk_step_prep", (rk,)))
        out.append(("rk_tendency", (rk,)))
        out.append(("small_step_prep", (rk, float(dts))))
        for it in range(1, nsmall + 1):
            out.append(("advance_uv_state", (float(dts),)))
            out.append(("advance_mu_t_w", (float(dts), it, nsmall)))
        out.append(("small_step_finish", (rk, nsmall, float(dts))))
        if rk == 3 and any(pd_species(cfg, n_moist, m) for m in range(n_species)):
            out.append(("rk_scalar_pd_prep", (float(dt_rk),)))
        for m in range(n_species):
            out.append(("rk_scalar_step", (m, rk, float(dt_rk))))
        out.append(("calc_p_rho_phi", ()))
        if rk < 3:
            out.append(("phys_bc_dry_2", (0,)))
    out.append(("calc_p_rho_phi", ()))
    out.append(("phys_bc_dry_2", (1,)))
    out.append(("set_w_surface", ()))
    return out


_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the HIP library (built by ``make -C wrf_amd/csrc``); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build() or `make -C wrf_amd/csrc`")
    lib = ctypes.CDLL(path)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    lib.wrfmi_init.restype = ctypes.c_int
    lib.wrfmi_finalize.restype = ctypes.c_int
    lib.wrfmi_last_error.restype = ctypes.c_char_p
    lib.wrfmi_version.restype = ctypes.c_char_p
    lib.wrfmi_state_field_count.restype = ctypes.c_int
    lib.wrfmi_state_sizeof.restype = sz
    lib.wrfmi_register_field.argtypes = [vp, sz]
    lib.wrfmi_unregister_field.argtypes = [vp]
    lib.wrfmi_sync_to_device.argtypes = [vp]
    lib.wrfmi_sync_to_host.argtypes = [vp]
    for fn in (lib.wrfmi_solve_em_step, lib.wrfmi_solve_em_step_graph):
        fn.argtypes = [vp, vp, vp, vp, ctypes.c_int]
        fn.restype = ctypes.c_int
    ci, cf = ctypes.c_int, ctypes.c_float
    for name, extra in PHASE_ARGS.items():
        fn = getattr(lib, "wrfmi_" + name)
        fn.argtypes = [vp, vp, vp] + [ci if t == "i" else cf for t in extra] + [vp]
        fn.restype = ctypes.c_int
    for fn in (lib.wrfmi_set_physical_bc3d, lib.wrfmi_set_physical_bc2d):
        fn.argtypes = [vp, ctypes.c_char, vp, vp, vp]
        fn.restype = ctypes.c_int
    lib.wrfmi_ensemble_create.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.POINTER(vp)]
    for fn in (lib.wrfmi_ensemble_step, lib.wrfmi_ensemble_step_graph):
        fn.argtypes = [vp, vp, ctypes.c_int]
    lib.wrfmi_ensemble_destroy.argtypes = [vp]
    lib.wrfmi_check_finite.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_longlong)]
    lib.wrfmi_domain_release.argtypes = [vp]
    lib.wrfmi_domain_totals.argtypes = [vp, vp, vp, vp, ctypes.POINTER(ctypes.c_double)]
    lib.wrfmi_kernel_count.restype = ctypes.c_int
    lib.wrfmi_kernel_name.restype = ctypes.c_char_p
    lib.wrfmi_kernel_name.argtypes = [ctypes.c_int]
    lib.wrfmi_timing_enable.argtypes = [ctypes.c_uint]
    lib.wrfmi_tuning_set.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 6
    lib.wrfmi_tuning_load.argtypes = [ctypes.c_char_p]
    lib.wrfmi_tuning_get.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_int)] * 3
    lib.wrfmi_timing_get.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_double)]
    if lib.wrfmi_state_sizeof() != ctypes.sizeof(State32):
        raise RuntimeError("wrfmi_state layout of the library does not match wrf_amd.fields")
    _lib = lib
    return lib


def check(rc: int, lib=None):
    """Raises WrfmiError with the library's message unless rc is WRFMI_OK (lib: the library
    that returned rc, the GPU library by default)."""
    if rc != WRFMI_OK:
        lib = lib or load_library()
        lib.wrfmi_last_error.restype = ctypes.c_char_p
        raise WrfmiError(rc, lib.wrfmi_last_error().decode())


class KernelTimer:
    """Per-kernel HIP-event timing of the eager entry points (wrfmi_timing_*)."""

    def __init__(self, kernels=None):
        """kernels: names to time (None = all)."""
        self.lib = load_library()
        names = [self.lib.wrfmi_kernel_name(k).decode() for k in range(self.lib.wrfmi_kernel_count())]
        if kernels is None:
            self.mask = (1 << len(names)) - 1
        else:
            unknown = set(kernels) - set(names)
            if unknown:
                raise ValueError(f"unknown kernels {sorted(unknown)}")
            self.mask = sum(1 << names.index(k) for k in kernels)

    def __enter__(self):
        check(self.lib.wrfmi_timing_enable(self.mask))
        return self

    def __exit__(self, *exc):
        self.lib.wrfmi_timing_enable(0)
        return False

    def collect(self) -> dict:
        """name -> (launches, total_ms) of the events recorded so far."""
        check(self.lib.wrfmi_timing_collect())
        out = {}
        for kid in range(self.lib.wrfmi_kernel_count()):
            n, ms = ctypes.c_longlong(0), ctypes.c_double(0.0)
            check(self.lib.wrfmi_timing_get(kid, ctypes.byref(n), ctypes.byref(ms)))
            out[self.lib.wrfmi_kernel_name(kid).decode()] = (n.value, ms.value)
        return out


class DeviceDomain:
    """A domain resident in HBM: every field of a ``wrf_amd.grid.HostState`` as a torch
    tensor on ``device``, plus the ``wrfmi_state`` struct of their device addresses."""

    def __init__(self, hs, device="cuda:0"):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("DeviceDomain needs a GPU (torch.cuda.is_available() is False)")
        self.lib = load_library()
        check(self.lib.wrfmi_init())
        self.torch = torch
        self.device = torch.device(device)
        self.dims = Dims.from_buffer_copy(hs.dims)
        self.cfg = Config.from_buffer_copy(hs.cfg)
        self.n_moist, self.n_species = hs.n_moist, hs.n_species
        self.t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(self.device) for k, v in hs.arrays().items()}
        st = State32()
        P = ctypes.POINTER(ctypes.c_float)
        for n in FIELDS_3D + FIELDS_2D + FIELDS_1D:
            setattr(st, n, ctypes.cast(self.t[n].data_ptr(), P))
        for n in SCALARS_R:
            setattr(st, n, hs.scal[n])
        st.a, st.alpha, st.gamma = self.t["a"].data_ptr(), self.t["alpha"].data_ptr(), self.t["gamma"].data_ptr()
        st.n_moist, st.n_species = hs.n_moist, hs.n_species
        for m in range(hs.n_species):
            st.q[m] = ctypes.cast(self.t[f"q{m}"].data_ptr(), P)
            st.q_old[m] = ctypes.cast(self.t[f"q_old{m}"].data_ptr(), P)
            st.q_tend[m] = ctypes.cast(self.t[f"q_tend{m}"].data_ptr(), P)
        if hs.lbc is not None:
            # the lateral boundary arrays of a specified domain (the struct lives in host memory)
            from .grid import lbc_struct
            self.lbc = lbc_struct(hs.lbc, Lbc32, lambda n: self.t[n].data_ptr())
            st.lbc = ctypes.addressof(self.lbc)
        self.state = st
        self.scal = dict(hs.scal)

    def stream_handle(self, stream=None) -> int:
        s = stream if stream is not None else self.torch.cuda.current_stream(self.device)
        return s.cuda_stream

    def step(self, nsteps: int = 1, graph: bool = False, stream=None, sync: bool = True):
        """Advance the domain by `nsteps` RK3 time steps on `stream` (torch's current
        stream by default).  ``graph=True`` replays a captured hipGraph of the step."""
        fn = self.lib.wrfmi_solve_em_step_graph if graph else self.lib.wrfmi_solve_em_step
        h = self.stream_handle(stream)
        for n in range(nsteps):
            check(fn(ctypes.byref(self.state), ctypes.byref(self.cfg), ctypes.byref(self.dims), ctypes.c_void_p(h),
                     int(sync and n == nsteps - 1)))

    def phase(self, name: str, *args, stream=None):
        """Enqueue one per-routine entry point (``wrfmi_<name>``) on `stream`."""
        fn = getattr(self.lib, "wrfmi_" + name)
        h = self.stream_handle(stream)
        check(fn(ctypes.byref(self.state), ctypes.byref(self.cfg), ctypes.byref(self.dims), *args, ctypes.c_void_p(h)))

    def addr(self, name: str) -> int:
        """Device address of the field `name`."""
        return self.t[name].data_ptr()

    def upload(self, arrays: dict):
        """Overwrite device fields with host arrays of the same names and shapes."""
        torch = self.torch
        for k, v in arrays.items():
            t = self.t[k]
            src = torch.from_numpy(np.ascontiguousarray(v, dtype=np.dtype(str(t.dtype).replace("torch.", ""))))
            if tuple(src.shape) != tuple(t.shape):
                raise ValueError(f"{k}: shape {tuple(src.shape)} does not match {tuple(t.shape)}")
            t.copy_(src.to(self.device))
        torch.cuda.synchronize(self.device)

    def release(self):
        """Free the library's work arrays and cached graphs of this domain (it may be
        stepped again afterwards: they are allocated again on the next call)."""
        check(self.lib.wrfmi_domain_release(ctypes.byref(self.state)))

    TOTALS = ("dry_mass_kg", "theta_budget_kg_K", "kinetic_J", "internal_J", "potential_J", "total_energy_J")

    def domain_totals(self, stream=None) -> dict:
        """FP64 domain totals of the current state (wrfmi_domain_totals)."""
        out = (ctypes.c_double * 6)()
        check(self.lib.wrfmi_domain_totals(ctypes.byref(self.state), ctypes.byref(self.cfg), ctypes.byref(self.dims),
                                           ctypes.c_void_p(self.stream_handle(stream)), out))
        return dict(zip(self.TOTALS, out[:]))

    def check_finite(self) -> int:
        nbad = ctypes.c_longlong(0)
        rc = self.lib.wrfmi_check_finite(ctypes.byref(self.state), ctypes.byref(self.dims), ctypes.byref(nbad))
        if rc not in (WRFMI_OK, WRFMI_E_NONFINITE):
            check(rc)
        return nbad.value

    def to_host(self, names=None) -> dict:
        self.torch.cuda.synchronize(self.device)
        keys = names if names is not None else list(self.t)
        return {k: self.t[k].cpu().numpy() for k in keys}

    def download(self) -> dict:
        return self.to_host()

    def hbm_bytes(self) -> int:
        return sum(v.numel() * v.element_size() for v in self.t.values())



class History:
    """A history file in the reference's wrfout layout (wrfmi_history_open/_write/_close,
    include/wrfmi_history.h): `lib`, `state`, `cfg`, `dims` are a domain's library handle
    and ctypes structs (DeviceDomain's, or a registered-host-array domain's)."""

    def __init__(self, lib, state, cfg, dims, path, start_date="2000-01-01_00:00:00", species_names=None):
        vp = ctypes.c_void_p
        lib.wrfmi_history_open.argtypes = [ctypes.c_char_p, vp, vp, vp, ctypes.c_char_p, vp, vp]
        lib.wrfmi_history_write.argtypes = [vp, vp, ctypes.c_char_p, ctypes.c_float, vp]
        lib.wrfmi_history_close.argtypes = [vp]
        lib.wrfmi_history_records.argtypes = [vp]
        self.lib, self.state = lib, state
        names = None
        if species_names is not None:
            self._names = [n.encode() if n else None for n in species_names]
            names = (ctypes.c_char_p * len(self._names))(*self._names)
        self.h = ctypes.c_void_p()
        check(lib.wrfmi_history_open(str(path).encode(), ctypes.byref(state), ctypes.byref(cfg), ctypes.byref(dims),
                                     start_date.encode(), names, ctypes.byref(self.h)), lib)

    def write(self, date, xtime, stream=None):
        check(self.lib.wrfmi_history_write(self.h, ctypes.byref(self.state), date.encode(), float(xtime),
                                           ctypes.c_void_p(stream or 0)), self.lib)

    @property
    def records(self):
        return self.lib.wrfmi_history_records(self.h)

    def close(self):
        if self.h:
            check(self.lib.wrfmi_history_close(self.h), self.lib)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class DeviceEnsemble:
    """Independent domains (members, of any sizes) advanced together by the batched entry
    points wrfmi_ensemble_step / _graph: one launch per kernel of the step over the tiles
    of all members.  ``members`` are DeviceDomains (their HBM fields are used in place)."""

    def __init__(self, members):
        self.lib = load_library()
        self.members = list(members)
        n = len(self.members)
        self._states = (State32 * n)(*[m.state for m in self.members])
        self._cfgs = (Config * n)(*[m.cfg for m in self.members])
        self._dims = (Dims * n)(*[m.dims for m in self.members])
        self.handle = ctypes.c_void_p()
        check(self.lib.wrfmi_ensemble_create(n, ctypes.byref(self._states), ctypes.byref(self._cfgs),
                                             ctypes.byref(self._dims), ctypes.byref(self.handle)))

    def step(self, nsteps: int = 1, graph: bool = False, stream=None, sync: bool = True):
        """Advance every member by `nsteps` RK3 steps."""
        fn = self.lib.wrfmi_ensemble_step_graph if graph else self.lib.wrfmi_ensemble_step
        h = self.members[0].stream_handle(stream)
        for n in range(nsteps):
            check(fn(self.handle, ctypes.c_void_p(h), int(sync and n == nsteps - 1)))

    def close(self):
        if self.handle:
            check(self.lib.wrfmi_ensemble_destroy(self.handle))
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def points(self) -> int:
        """Mass points over all members."""
        return sum((m.dims.ide - m.dims.ids) * (m.dims.jde - m.dims.jds) * (m.dims.kde - m.dims.kds)
                   for m in self.members)
