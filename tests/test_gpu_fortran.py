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


def _setup_both(lattice, n, S, rank, numbers=None, conc=None):
    lib = C.CDLL(R.DROP_IN)
    f = lib.drop_in_setup
    f.restype = C.c_int
    P = C.POINTER
    f.argtypes = [C.c_int] * 5 + [P(C.c_int64), P(C.c_double), P(C.c_int8), P(C.c_int8),
                                  P(C.c_double), P(C.c_double)]
    m = 2 * n
    sn = np.zeros(S, dtype=np.int64) if numbers is None else np.asarray(numbers, dtype=np.int64)
    cc = np.zeros(S) if conc is None else np.asarray(conc, dtype=np.float64)
    cg = np.zeros(m ** 3, dtype=np.int8)
    cr = np.zeros(m ** 3, dtype=np.int8)
    kg, kr = np.zeros(S), np.zeros(S)
    print("drop_in_setup", lattice, n, S, rank, "numbers" if numbers is not None else "conc", flush=True)
    rc = f(lattice, n, S, rank, int(numbers is not None), sn.ctypes.data_as(P(C.c_int64)),
           cc.ctypes.data_as(P(C.c_double)), cg.ctypes.data_as(P(C.c_int8)),
           cr.ctypes.data_as(P(C.c_int8)), kg.ctypes.data_as(P(C.c_double)),
           kr.ctypes.data_as(P(C.c_double)))
    assert rc == 0, f"GPU library error {rc}"
    shape = (m, m, m)
    return (cg.reshape(shape, order="F"), cr.reshape(shape, order="F"), kg, kr)


@pytest.mark.parametrize("rank", [0, 1, 2, 3])
def test_fortran_initial_setup_concentrations_path_reproduces_reference_test_02(rank):
    """The reference's test 02 (bcc 4^3, 4 species at 0.25, static seeds of
    ranks 0-3) sets up from concentrations held in a 0:S array
    (io.f90:242): the rebound adaptor reproduces the golden initial
    configuration of each rank, as does the reference's own routine."""
    from conftest import load_golden
    g = load_golden("ref02_parallel_metropolis.npz")
    n = int(g["n"][0])
    cg, cr, kg, kr = _setup_both(int(g["lattice"]), n, 4, rank, conc=g["concentrations"])
    assert np.array_equal(cr, g[f"initial_config{rank}"])
    assert np.array_equal(cg, g[f"initial_config{rank}"])
    assert np.array_equal(kg, kr)


@pytest.mark.parametrize("lattice,n,S,conc", [(1, 3, 5, [0.2, 0.2, 0.2, 0.2, 0.2]),
                                              (2, 3, 3, [0.1, 0.35, 0.55])])
def test_fortran_initial_setup_concentrations_rounding(lattice, n, S, conc):
    cg, cr, kg, kr = _setup_both(lattice, n, S, 1, conc=conc)
    assert np.array_equal(cg, cr)
    assert np.array_equal(kg, kr)


@pytest.mark.parametrize("lattice,n,numbers", [(1, 3, [30, 40, 38]), (2, 4, [33, 31, 32, 32])])
def test_fortran_initial_setup_numbers_path_sets_concentrations(lattice, n, numbers):
    """On the species-numbers path the reference also stores the REAL(4)
    concentrations count/n_atoms in setup (initialise.F90:468-472)."""
    cg, cr, kg, kr = _setup_both(lattice, n, len(numbers), 0, numbers=numbers)
    assert np.array_equal(cg, cr)
    assert np.array_equal(kg, kr)
    assert kg.sum() > 0.99


@pytest.mark.parametrize("lattice,n", [(1, 3), (2, 4), (1, 5)])
def test_fortran_proposal_adaptors_match_reference_draws(lattice, n):
    """setup%rdm_site, setup%rdm_nbr and wl_params%mc_select (both modes)
    rebound to the device draw the same sites, from the same stream, as the
    reference's own procedures."""
    lib = C.CDLL(R.DROP_IN)
    f = lib.drop_in_draws
    f.restype = C.c_int
    f.argtypes = [C.c_int] * 4 + [C.POINTER(C.c_int)]
    mism = (C.c_int * 4)()
    print("drop_in_draws", lattice, n, flush=True)
    rc = f(lattice, n, 2, 300, mism)
    assert rc == 0, f"GPU library error {rc}"
    assert list(mism) == [0, 0, 0, 0]
