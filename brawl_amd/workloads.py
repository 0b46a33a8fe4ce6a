"""Synthetic inputs of the benchmark workloads (no data files needed).

Interactions: for each shell n, the S(S+1)/2 entries V(a<=b) are drawn from
N(0, sigma_n) meV, the mean of those entries is subtracted (zero-mean per
shell), the matrix is symmetrised and converted to Rydberg with the
reference's constant (constants.f90:49).  Seeds derive from a master seed
with SplitMix64; normals come from numpy's PCG64 generator.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .lattice import BCC, FCC, RY_TO_EV

MASTER_SEED = 20261202


def splitmix64(x: int) -> int:
    """One SplitMix64 output for state x (Steele, Lea, Flood 2014)."""
    z = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def derive_seed(master: int, index: int) -> int:
    return splitmix64((master + index * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF)


def random_v(S: int, sigmas_mev, seed: int) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    R = len(sigmas_mev)
    v = np.zeros((S, S, R), order="F")
    iu = np.triu_indices(S)
    for n, sig in enumerate(sigmas_mev):
        vals = rng.normal(0.0, sig, size=len(iu[0]))
        vals -= vals.mean()
        m = np.zeros((S, S))
        m[iu] = vals
        m = m + np.triu(m, 1).T
        v[:, :, n] = m / (1000.0 * RY_TO_EV)
    return v


def equal_counts(n_atoms: int, S: int) -> np.ndarray:
    base = np.full(S, n_atoms // S, dtype=np.int64)
    base[: n_atoms - base.sum()] += 1
    return base


def dirichlet_counts(n_atoms: int, S: int, rng: np.random.Generator) -> np.ndarray:
    """Largest-remainder rounding of Dirichlet(1) * N, every species >= 1."""
    w = rng.dirichlet(np.ones(S)) * (n_atoms - S)
    c = np.floor(w).astype(np.int64)
    rem = w - c
    short = (n_atoms - S) - c.sum()
    c[np.argsort(-rem)[:short]] += 1
    return c + 1


def n_atoms_of(lattice: int, n) -> int:
    return (4 if lattice == FCC else 2) * n[0] * n[1] * n[2]


@dataclass
class Chain:
    lattice: int
    n: tuple
    v_ex: np.ndarray
    counts: np.ndarray
    T: float
    delta_T: float
    T_steps: int
    n_sweeps: int
    seed: int
    name: str = ""

    @property
    def S(self):
        return self.v_ex.shape[0]

    @property
    def R(self):
        return self.v_ex.shape[2]

    @property
    def n_atoms(self):
        return n_atoms_of(self.lattice, self.n)


def config1(seed: int = MASTER_SEED) -> Chain:
    """fcc 8^3 (2048 sites) binary 1024/1024, 2 shells (20, 5 meV), 600 K,
    1000 sweeps."""
    s = derive_seed(seed, 1)
    v = random_v(2, [20.0, 5.0], s)
    return Chain(FCC, (8, 8, 8), v, np.array([1024, 1024], dtype=np.int64), 600.0, 0.0, 1,
                 1000, s, "fcc8_S2_R2_T600")


def config2(seed: int = MASTER_SEED) -> Chain:
    """bcc 32^3 (65536 sites) equiatomic quaternary, 4 shells (30/n meV),
    anneal 3000 K -> 100 K in 50 steps of 200 sweeps."""
    s = derive_seed(seed, 2)
    v = random_v(4, [30.0 / n for n in range(1, 5)], s)
    return Chain(BCC, (32, 32, 32), v, equal_counts(65536, 4), 3000.0, -2900.0 / 49.0, 50, 200,
                 s, "bcc32_S4_R4_anneal")


def config3(n_chains: int = 256, seed: int = MASTER_SEED, n_sweeps: int = 500):
    """Uneven replicas: L ~ U{8..24}, fcc or bcc, S ~ U{2..5}, Dirichlet(1)
    counts, 3 shells (25/n meV), T ~ U[200, 2000] K."""
    out = []
    for c in range(n_chains):
        s = derive_seed(seed, 1000 + c)
        rng = np.random.Generator(np.random.PCG64(s))
        L = int(rng.integers(8, 25))
        lat = FCC if rng.integers(0, 2) == 0 else BCC
        S = int(rng.integers(2, 6))
        v = random_v(S, [25.0 / n for n in range(1, 4)], s ^ 0x5A5A)
        N = n_atoms_of(lat, (L, L, L))
        counts = dirichlet_counts(N, S, rng)
        T = float(rng.uniform(200.0, 2000.0))
        out.append(Chain(lat, (L, L, L), v, counts, T, 0.0, 1, n_sweeps, s, f"chain{c}"))
    return out


def random_config(lattice: int, n, counts, rng: np.random.Generator) -> np.ndarray:
    """A uniformly random occupation with exact species counts, in the
    reference's layout (2n1, 2n2, 2n3) int8, Fortran order."""
    m = tuple(2 * x for x in n)
    x, y, z = np.meshgrid(*[np.arange(k) for k in m], indexing="ij")
    occ = ((x + y + z) % 2 == 0) if lattice == FCC else ((x % 2 == y % 2) & (y % 2 == z % 2))
    species = np.repeat(np.arange(1, len(counts) + 1, dtype=np.int8), counts)
    rng.shuffle(species)
    cfg = np.zeros(m, dtype=np.int8, order="F")
    cfg[occ] = species
    return np.asfortranarray(cfg)


@dataclass
class ServedRequest:
    kind: str            # "anneal", "ns", "wl"
    arrival_s: float
    lattice: int
    n: tuple
    v_ex: np.ndarray
    seed: int
    params: dict
    counts: np.ndarray = None
    config: np.ndarray = None

    @property
    def n_atoms(self):
        return n_atoms_of(self.lattice, self.n)


def config5(n_requests: int = 240, rate_hz: float = 400.0, seed: int = MASTER_SEED):
    """Served mixed stream, Poisson arrivals at rate_hz requests/s:
    60 % Metropolis (fcc or bcc; bcc L ~ U{8..32}, fcc L ~ U{8..28}, the
    largest fcc box whose 3-shell lattice fits one workgroup's LDS;
    S ~ U{2..5}, Dirichlet(1) counts, 3 shells of 25/n meV, T ~ U[200,
    2000] K, U{100..500} sweeps), 25 % nested sampling (quinary fcc 6^3,
    4 shells of 30/n meV, U{128..640} walkers, walk length 100, 200
    iterations), 15 % short Wang-Landau runs (bcc L ~ U{6..10}, S = 4, 2
    shells of 30/n meV, one window of 50 bins over [-10, 10] meV/atom,
    U{4..8} walkers, 20 sweeps at ln f = 1)."""
    rng = np.random.Generator(np.random.PCG64(derive_seed(seed, 5)))
    t = 0.0
    out = []
    for r in range(n_requests):
        t += float(rng.exponential(1.0 / rate_hz))
        s = derive_seed(seed, 50000 + r)
        u = float(rng.uniform())
        if u < 0.60:
            lat = FCC if rng.integers(0, 2) == 0 else BCC
            L = int(rng.integers(8, 29 if lat == FCC else 33))
            S = int(rng.integers(2, 6))
            v = random_v(S, [25.0 / k for k in range(1, 4)], s ^ 0x5A5A)
            N = n_atoms_of(lat, (L, L, L))
            counts = dirichlet_counts(N, S, rng)
            cfg = random_config(lat, (L, L, L), counts, rng)
            out.append(ServedRequest("anneal", t, lat, (L, L, L), v, s,
                                     dict(T=float(rng.uniform(200.0, 2000.0)),
                                          n_sweeps=int(rng.integers(100, 501))), counts, cfg))
        elif u < 0.85:
            S = 5
            v = random_v(S, [30.0 / k for k in range(1, 5)], s ^ 0x5A5A)
            out.append(ServedRequest("ns", t, FCC, (6, 6, 6), v, s & 0xFFFFFFFF,
                                     dict(n_walkers=int(rng.integers(128, 641)), n_steps=100,
                                          n_iter=200), equal_counts(864, S)))
        else:
            L = int(rng.integers(6, 11))
            v = random_v(4, [30.0, 15.0], s ^ 0x5A5A)
            out.append(ServedRequest("wl", t, BCC, (L, L, L), v, s,
                                     dict(bins=50, num_windows=1,
                                          walkers_per_window=int(rng.integers(4, 9)),
                                          bin_overlap=0.1, energy_min=-10.0, energy_max=10.0,
                                          mc_sweeps=20, wl_f=1.0, rank_offset=int(s % 1000)),
                                     equal_counts(n_atoms_of(BCC, (L, L, L)), 4)))
    return out


def config4(seed: int = MASTER_SEED):
    """Wang-Landau, quaternary bcc 16^3 (8192 sites), equiatomic, 2 shells
    with V^(n) ~ N(0, 30/n meV); 16 windows of 50 bins (800 bins over
    [-40, 10] meV/atom, 10 % overlap), 8 walkers per window, ln f from 1,
    flatness 0.8."""
    s = derive_seed(seed, 4)
    v = random_v(4, [30.0, 15.0], s)
    return dict(lattice=BCC, n=(16, 16, 16), v_ex=v, counts=equal_counts(8192, 4), bins=800,
                num_windows=16, walkers_per_window=8, bin_overlap=0.1, energy_min=-40.0,
                energy_max=10.0, wl_f=1.0, flatness=0.8, tolerance=1e-6)
