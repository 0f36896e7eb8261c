"""Phase-by-phase comparison of the CPU-emulated kernel sources with the FP32 oracle.
Run by test_cpu_emu.py in a subprocess with libasan preloaded:
    python tests/emu_phase_check.py <case> <launch-order> <nsteps>
Exit status 0 = every array bitwise equal after every call (EOS outputs within 4 ulp)."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(HERE), os.path.join(os.path.dirname(HERE), "oracle"), HERE]

from emu_backend import EmuDomain  # noqa: E402
from parity import phase_parity  # noqa: E402
from pyoracle import Oracle  # noqa: E402
from wrf_amd import grid, runtime  # noqa: E402

CASES = {
    "periodic": dict(nx=14, ny=11, nz=10, bc="periodic", n_passive=1, n_bubbles=2, bubble_amp=2.0, seed=11),
    "open": dict(nx=13, ny=12, nz=9, bc="open", n_bubbles=1, seed=12),
    "periodic_pm1": dict(nx=12, ny=10, nz=9, bc="periodic", n_bubbles=1, seed=13, precision_mode=1),
}


def main():
    case, order, nsteps = sys.argv[1], int(sys.argv[2]), int(sys.argv[3])
    spec = grid.DomainSpec(**CASES[case])
    hs = grid.build_initial_state(spec)
    emu = EmuDomain(hs, order=order)
    oracle = Oracle("f32")
    sess = oracle.session(hs)
    phases = runtime.step_phases(hs.cfg, hs.scal, hs.n_species)
    fails = phase_parity(emu, sess, phases, nsteps=nsteps)
    for f in fails:
        print("MISMATCH", f)
    print("OK" if not fails else "FAILED", case, "order", order, flush=True)
    sys.exit(1 if fails else 0)


if __name__ == "__main__":
    main()
