"""The oracle against the reference's own golden outputs (tests/99_ref,
committed as fixtures under tests/golden).  CPU only."""
import numpy as np
import pytest

from conftest import load_golden


def _traj_text(values):
    # energy_trajectory_writer format (I13,x,f20.10), metropolis_output.f90:49
    return ["%13d %20.10f" % (i, e) for i, e in enumerate(values)]


def test_serial_metropolis_fcc_bit_exact(oracle):
    """tests/01_serial_metropolis: initial config, 257-line energy trajectory,
    final config and radial densities reproduce exactly."""
    g = load_golden("ref01_serial_metropolis.npz")
    L = oracle.OracleLattice(int(g["lattice"]), g["n"], g["v_ex"])
    gen = oracle.Generator(int(g["seed"]))
    cfg = L.initial_setup(gen, species_numbers=g["species_numbers"])
    assert np.array_equal(cfg, g["initial_config"])
    res = L.anneal(cfg, gen, T=float(g["T"]), n_mc_steps=int(g["n_mc_steps"]),
                   n_sample_steps=int(g["n_sample_steps"]), trajectory=True,
                   calculate_asro=True, n_sample_steps_asro=int(g["n_mc_steps"]),
                   wc_range=int(g["wc_range"]))
    assert _traj_text(res["trajectory"]) == _traj_text(g["traj_energy"])
    assert np.array_equal(cfg, g["final_config"])
    assert np.array_equal(res["rho"], g["rho"])
    assert res["energies"][0] == g["U"][0]
    np.testing.assert_array_equal(L.lattice_shells(g["initial_config"], 3), g["shells"])


@pytest.mark.parametrize("rank", [0, 1, 2, 3])
def test_parallel_metropolis_bcc_ranks_bit_exact(oracle, rank):
    """tests/02_parallel_metropolis: each MPI rank is an independent chain
    seeded 110179 + 11*rank; everything it writes reproduces exactly."""
    g = load_golden("ref02_parallel_metropolis.npz")
    L = oracle.OracleLattice(int(g["lattice"]), g["n"], g["v_ex"])
    gen = oracle.Generator(int(g[f"seed{rank}"]))
    cfg = L.initial_setup(gen, concentrations=g["concentrations"])
    assert np.array_equal(cfg, g[f"initial_config{rank}"])
    res = L.anneal(cfg, gen, T=float(g["T"]), n_mc_steps=int(g["n_mc_steps"]),
                   n_sample_steps=1, trajectory=True, calculate_asro=True,
                   n_sample_steps_asro=int(g["n_mc_steps"]), wc_range=int(g["wc_range"]))
    assert _traj_text(res["trajectory"]) == _traj_text(g[f"traj_energy{rank}"])
    assert np.array_equal(cfg, g[f"final_config{rank}"])
    assert np.array_equal(res["rho"], g[f"rho{rank}"])
    # diagnostics_writer format (F8.1,2X,F24.15,2X,F24.15,2X,F6.4)
    d = g[f"diagnostics{rank}"]
    ours = "%8.1f  %24.15f  %24.15f  %6.4f" % (300.0, res["energies"][0], res["C"][0],
                                               res["acceptance"][0])
    gold = "%8.1f  %24.15f  %24.15f  %6.4f" % tuple(d)
    assert ours == gold


def test_rank_average_diagnostics(oracle):
    """av_energy_diagnostics.dat is the mean over the 4 ranks
    (comms.F90:122-160)."""
    g = load_golden("ref02_parallel_metropolis.npz")
    d = np.array([g[f"diagnostics{r}"] for r in range(4)])
    av = d.mean(axis=0)
    np.testing.assert_allclose(av[1:3], g["av_diagnostics"][1:3], rtol=0, atol=1e-15)
    assert abs(av[3] - g["av_diagnostics"][3]) < 1e-4


def test_nested_sampling_statistics(oracle):
    """tests/99_ref/03: the reference's NS golden is not reproducible bit for
    bit from the current sources (the upstream CI never compares it), so the
    restated NS loop is held to the golden statistically: the energy-limit
    curve E_lim(i) follows the same compression within a few percent of the
    golden's range."""
    g = load_golden("ref03_nested_sampling.npz")
    L = oracle.OracleLattice(int(g["lattice"]), g["n"], g["v_ex"])
    gen = oracle.Generator(int(g["seed"]))
    el = L.nested_sampling(gen, n_walkers=int(g["n_walkers"]), n_steps=int(g["n_steps"]),
                           n_iter=int(g["n_iter"]), species_numbers=g["species_numbers"])
    gold = g["e_limits"]
    assert np.all(np.diff(el) <= 0) and np.all(np.diff(gold) <= 0)
    span = gold[0] - gold[-1]
    late = slice(100, None)
    assert np.max(np.abs(el[late] - gold[late])) < 0.06 * span
    assert abs(np.mean(el[late] - gold[late])) < 0.03 * span
    # header: n_walkers, 1, 0, False, n1*n2*n3*n_basis*n_species (A11)
    assert list(g["header"]) == ["100", "1", "0", "False", "135"]
