"""TEST INFRASTRUCTURE ONLY: speed of the oracle (the C restatement) beside
the reference's own compiled routine, on the CPU of this machine.

Times total_energy (the reference's bw_hamiltonian.f90:58-81 through
oracle/_ref, and the oracle's restatement) on the small config (fcc 8^3,
2 shells) and the quaternary bcc config (bcc 32^3, 4 shells): N site
energies, i.e. the routine that dominates the reference's Metropolis trial
(pair_energy = two site energies, twice per trial).  Needs the reference
build (python -c "from oracle import ref; ref.build()").

    python3 oracle/time_vs_reference.py
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from oracle import oracle as orc  # noqa: E402
from oracle import ref  # noqa: E402


@ref._isolated
def _time_ref(lattice, n, v_ex, cfg, reps):
    v = np.asfortranarray(v_ex, dtype=np.float64)
    S, _, R = v.shape
    c = np.asfortranarray(cfg, dtype=np.int8)
    f = ref.lib().ref_total_energy
    args = (lattice, *n, S, R, v.ctypes.data_as(C.POINTER(C.c_double)),
            c.ctypes.data_as(C.POINTER(C.c_int8)))
    f(*args)
    t0 = time.perf_counter()
    for _ in range(reps):
        e = f(*args)
    return (time.perf_counter() - t0) / reps, e


def main():
    from brawl_amd import workloads
    orc.build()
    if not ref.available():
        print("reference build absent (oracle/_ref)")
        return
    for name, ch, reps in (("config 1: fcc 8^3, R=2", workloads.config1(), 200),
                           ("config 2: bcc 32^3, R=4", workloads.config2(), 5)):
        o = orc.OracleLattice(ch.lattice, ch.n, ch.v_ex)
        cfg = o.initial_setup(orc.Generator(110179), species_numbers=ch.counts)
        o.total_energy(cfg)
        t0 = time.perf_counter()
        for _ in range(reps):
            e_o = o.total_energy(cfg)
        t_o = (time.perf_counter() - t0) / reps
        t_r, e_r = _time_ref(ch.lattice, ch.n, ch.v_ex, cfg, reps)
        print(f"{name}: total_energy reference {t_r * 1e3:.3f} ms, oracle {t_o * 1e3:.3f} ms, "
              f"reference/oracle {t_r / t_o:.2f}, energies equal: {e_o == e_r}")


if __name__ == "__main__":
    main()
