"""The parallel sampler's equilibrium <E>/N against the sequential kernel,
which follows the reference's own Metropolis algorithm trial by trial
(random site pairs in the order of monte_carlo_step_lattice,
metropolis.F90:751-813; its trajectories are the reference's bit for bit,
test_gpu_parity.py), on fixed-pitch kernels that the other statistical
tests do not reach: bcc with 2 shells (the Wang-Landau workload's lattice)
and bcc and fcc with 3 shells (the batched workload's kernels).

Both samplers run 16 chains from independent random starts at the same
temperature, composition and V, with the same burn-in and sampled sweeps,
and the energy of every sampled sweep is averaged per chain; the heat
capacity per chain is the variance of those energies over k_B T^2 N, as
the reference forms it (metropolis.F90).  For <E>/N and for C the two
samplers' means agree within 2 sigma of their difference (sigma from the
chain-to-chain spread of each).  The seeds are fixed, so every run gives
the same numbers."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RY_TO_MEV = 13.605693122 * 1000.0

# lattice (1 fcc, 2 bcc), cells, species, sigma of V^(n) per shell (meV), T (K)
CASES = [
    (2, (8, 8, 8), 4, [30.0, 15.0], 2000.0),
    (2, (10, 10, 10), 3, [25.0, 12.5, 25.0 / 3], 1500.0),
    (1, (8, 8, 8), 4, [25.0, 12.5, 25.0 / 3], 1500.0),
]
CHAINS, BURN, SAMPLED = 16, 300, 1000


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("lat,n,S,sig,T", CASES)
def test_parallel_sampler_matches_reference_algorithm(lat, n, S, sig, T):
    from brawl_amd import Lattice, _lib, beta_of, workloads
    from brawl_amd.lattice import K_B_IN_RY, anneal_batch, mc_steps_batch
    _lib.load()
    v = workloads.random_v(S, sig, workloads.derive_seed(workloads.MASTER_SEED, 700 + lat * 10 + S))
    par, seq = [], []
    try:
        for k in range(CHAINS):
            for group, base in ((par, 1000), (seq, 2000)):
                L = Lattice(lat, n, v)
                L.seed(base + 11 * k)
                N = L.n_atoms
                L.initial_setup(species_numbers=workloads.equal_counts(N, S))
                group.append(L)
        res, _ = anneal_batch(par, [dict(T=T, n_sweeps=SAMPLED, burn_in_sweeps=BURN,
                                         seed=workloads.derive_seed(workloads.MASTER_SEED, 9000 + k))
                                    for k in range(CHAINS)])
        e_par = np.array([r.energies_of_T[0] for r in res]) * RY_TO_MEV
        c_par = np.array([r.C_of_T[0] for r in res])
        betas = [beta_of(T)] * CHAINS
        mc_steps_batch(seq, betas, BURN * N)
        _, E, _ = mc_steps_batch(seq, betas, SAMPLED * N, sample_every=N)
        e_seq = np.array([e.mean() for e in E]) / N * RY_TO_MEV
        c_seq = np.array([e.var() for e in E]) / (K_B_IN_RY * T * T) / N
    finally:
        for L in par + seq:
            L.close()
    name = f"{'fcc' if lat == 1 else 'bcc'} {n} S={S} R={len(sig)} {T:.0f} K"
    for what, a, b in (("<E>/N (meV/atom)", e_par, e_seq), ("C (Ry/K per atom)", c_par, c_seq)):
        d = a.mean() - b.mean()
        sd = np.sqrt(a.var(ddof=1) / CHAINS + b.var(ddof=1) / CHAINS)
        _say(f"{name} {what}: parallel {a.mean():.6g}, sequential {b.mean():.6g}, difference "
             f"{d:+.3g} = {d / sd:+.2f} sigma")
        assert abs(d) <= 2.0 * sd, what
