"""Drives tests/cpu_emu/libwrfmi_emu.so (the kernel sources built for the CPU under
AddressSanitizer) through the same C ABI as the GPU library -- test infrastructure.
Must run in a process that has libasan preloaded (see test_cpu_emu.py)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from wrf_amd.fields import FIELDS_1D, FIELDS_2D, FIELDS_3D, SCALARS_R, Lbc32, State32
from wrf_amd.runtime import PHASE_ARGS

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_LIB = os.path.join(HERE, "cpu_emu", "libwrfmi_emu.so")


class EmuDomain:
    """Host arrays registered with the emulated library (each gets an exact-size
    malloc'd mirror, so out-of-bounds accesses hit AddressSanitizer redzones)."""

    def __init__(self, hs, order: int = 0):
        lib = ctypes.CDLL(EMU_LIB)
        vp = ctypes.c_void_p
        lib.wrfmi_register_field.argtypes = [vp, ctypes.c_size_t]
        lib.wrfmi_unregister_field.argtypes = [vp]
        lib.wrfmi_sync_to_device.argtypes = [vp]
        lib.wrfmi_sync_to_host.argtypes = [vp]
        lib.wrfmi_last_error.restype = ctypes.c_char_p
        for name, extra in PHASE_ARGS.items():
            fn = getattr(lib, "wrfmi_" + name)
            fn.argtypes = [vp, vp, vp] + [ctypes.c_int if t == "i" else ctypes.c_float for t in extra] + [vp]
        lib.wrfmi_solve_em_step.argtypes = [vp, vp, vp, vp, ctypes.c_int]
        lib.emu_set_launch_order(order)
        self.lib = lib
        self.cfg, self.dims = hs.cfg, hs.dims
        self.arrays = {k: np.array(v, copy=True) for k, v in hs.arrays().items()}
        for v in self.arrays.values():
            self._chk(lib.wrfmi_register_field(v.ctypes.data, v.nbytes))
        st = State32()
        P = ctypes.POINTER(ctypes.c_float)
        for n in FIELDS_3D + FIELDS_2D + FIELDS_1D:
            setattr(st, n, self.arrays[n].ctypes.data_as(P))
        for n in SCALARS_R:
            setattr(st, n, hs.scal[n])
        st.a, st.alpha, st.gamma = (self.arrays[k].ctypes.data for k in ("a", "alpha", "gamma"))
        st.n_moist, st.n_species = hs.n_moist, hs.n_species
        for m in range(hs.n_species):
            st.q[m] = self.arrays[f"q{m}"].ctypes.data_as(P)
            st.q_old[m] = self.arrays[f"q_old{m}"].ctypes.data_as(P)
            st.q_tend[m] = self.arrays[f"q_tend{m}"].ctypes.data_as(P)
        if hs.lbc is not None:
            from wrf_amd.grid import lbc_struct
            self.lbc = lbc_struct(hs.lbc, Lbc32, lambda n: self.arrays[n].ctypes.data)
            st.lbc = ctypes.addressof(self.lbc)
        self.state = st

    def __del__(self):
        arrays = getattr(self, "arrays", None)
        if arrays:
            for v in arrays.values():
                self.lib.wrfmi_unregister_field(ctypes.c_void_p(v.ctypes.data))

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(f"emulated wrfmi call failed ({rc}): {self.lib.wrfmi_last_error().decode()}")

    def upload(self, arrays: dict):
        for k, v in arrays.items():
            self.arrays[k][...] = v
            self._chk(self.lib.wrfmi_sync_to_device(self.arrays[k].ctypes.data))

    def addr(self, name: str) -> int:
        """Address of the registered host array `name` (the library uses its mirror)."""
        return self.arrays[name].ctypes.data

    def download(self) -> dict:
        for v in self.arrays.values():
            self._chk(self.lib.wrfmi_sync_to_host(v.ctypes.data))
        return {k: v.copy() for k, v in self.arrays.items()}

    def phase(self, name, *args):
        fn = getattr(self.lib, "wrfmi_" + name)
        self._chk(fn(ctypes.byref(self.state), ctypes.byref(self.cfg), ctypes.byref(self.dims), *args, None))

    def step(self, nsteps=1):
        for _ in range(nsteps):
            self._chk(self.lib.wrfmi_solve_em_step(ctypes.byref(self.state), ctypes.byref(self.cfg),
                                                   ctypes.byref(self.dims), None, 1))
