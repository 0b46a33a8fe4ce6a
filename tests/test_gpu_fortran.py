"""The Fortran drop-in: BraWl host code (tests/fortran/drop_in.f90, built
with the reference's own modules by oracle/ref.mk, called in-process) rebinds its
`hamiltonian`, `neighbour` and `monte_carlo` procedure pointers to the GPU
adaptors (brawl_amd/fortran/brawl_hip_adaptors.f90) and evaluates the same
configuration both ways.  Checks: total energy within 1e-12 relative of the
reference's total_energy, site energies bit-identical to the reference's
*_energy_Rshells, the device initial_setup identical to the reference's,
MC steps through the rebound pointer identical to the oracle's trial
sequence, and the parallel sampler's anneal called from Fortran through
the C-ABI (composition kept, reported energy that of the configuration it
returns)."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import ref as R

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(R.DROP_IN),
                                 reason="drop-in binary absent (built with the reference's modules)")]


def _run(tmp_path, lattice, n, S, Rs, trials):
    out = tmp_path / f"drop_in_{lattice}_{n}_{S}_{Rs}.txt"
    print("drop_in", lattice, n, S, Rs, trials, flush=True)
    # the library runs with the reference's stop-on-error convention turned
    # off (errors come back as the status), and its inputs are valid for
    # the reference's routines, so nothing in it stops the process
    lib = C.CDLL(R.DROP_IN)
    f = lib.drop_in_run
    f.restype = C.c_int
    f.argtypes = [C.c_int] * 5 + [C.c_char_p]
    rc = f(1 if lattice == "fcc" else 2, n, S, Rs, trials, str(out).encode())
    assert rc == 0, f"GPU library error {rc}"
    vals = {}
    for line in out.read_text().splitlines():
        k, _, v = line.strip().partition(" ")
        vals[k] = v
    return vals


@pytest.mark.parametrize("lattice,n,S,Rs", [("fcc", 4, 3, 4), ("bcc", 4, 2, 8), ("fcc", 3, 5, 6),
                                            ("bcc", 3, 4, 10)])
def test_fortran_host_rebinds_to_gpu(oracle, tmp_path, lattice, n, S, Rs):
    trials = 300
    v = _run(tmp_path, lattice, n, S, Rs, trials)
    e_ref, e_gpu = float(v["e_ref"]), float(v["e_gpu"])
    assert abs(e_gpu - e_ref) <= 1e-12 * max(1.0, abs(e_ref)) + 1e-15
    assert int(v["site_energy_checked"]) > 0
    assert int(v["site_energy_mismatches"]) == 0
    assert int(v["initial_setup_mismatches"]) == 0
    # replay the same stream on the oracle: initial_setup, then the trials
    lat = oracle.FCC if lattice == "fcc" else oracle.BCC
    o = oracle.OracleLattice(lat, (n, n, n), np.zeros((S, S, Rs)))
    N = o.n_atoms
    counts = np.full(S, N // S, dtype=np.int64)
    counts[0] += N - S * (N // S)
    vex = np.zeros((S, S, Rs), order="F")
    for s in range(1, Rs + 1):
        for b in range(1, S + 1):
            for a in range(1, S + 1):
                vex[a - 1, b - 1, s - 1] = 1.0e-3 * np.sin(float(a * b + 3 * s)) / float(s)
    o = oracle.OracleLattice(lat, (n, n, n), vex)
    gen = oracle.Generator(rank=0)
    cfg = o.initial_setup(gen, species_numbers=counts)
    beta = 1.0 / (600.0 * 8.167333262e-5 / 13.605693122990)
    acc = sum(o.mc_step(cfg, beta, gen) for _ in range(trials))
    assert int(v["mc_accepted"]) == acc
    got = np.array([int(x) for x in v["mc_config"].split()], dtype=np.int8)
    assert np.array_equal(got, cfg.ravel(order="F"))
    # the anneal called from Fortran: composition kept, and the energy it
    # reports is the energy of the configuration it returned
    if int(v["anneal_status"]) != 0:
        # only a box too small for the sampler's rounds may be refused
        assert (lattice, n) != ("fcc", 4) and int(v["anneal_status"]) == -2
        return
    fin = np.array([int(x) for x in v["anneal_config"].split()], dtype=np.int8)
    assert np.array_equal(np.bincount(fin, minlength=S + 1), np.bincount(got, minlength=S + 1))
    assert int(v["anneal_trials"]) >= 4 * N
    fin_cfg = fin.reshape((2 * n, 2 * n, 2 * n), order="F")
    e_fin = o.total_energy(np.asfortranarray(fin_cfg))
    assert abs(float(v["anneal_E"]) * N - e_fin) <= 1e-11 * max(1.0, abs(e_fin))
