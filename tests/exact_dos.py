"""The exact density of states of a small lattice, shared by the tests that
hold the samplers to it: 5 solutes on bcc 3^3 (54 sites) with one shell,
whose energy is E = E_0 + m J with m the number of solute-solute shell-1
bonds (0..6), counted over all C(54, 5) = 3,162,510 placements."""
import itertools

import numpy as np

# V^(1) of the binary: J = V(1,1) - 2 V(1,2) + V(2,2) = 2 mRy per solute bond
V1 = ((2e-3, 0.5e-3), (0.5e-3, 1e-3))
J = 2e-3
SPECIES_NUMBERS = (5, 49)


def v_ex():
    V = np.zeros((2, 2, 1), order="F")
    V[:, :, 0] = V1
    return V


def exact_dos():
    """Sites (doubled coordinates), all placements, their bond counts m, and
    the distinct m with their multiplicities (the reference shell list's
    multiplicities)."""
    from oracle import oracle as orc
    lat, n = 2, (3, 3, 3)
    grid = tuple(2 * x for x in n)
    sites = [(i, j, k) for k in range(grid[2]) for j in range(grid[1]) for i in range(grid[0])
             if i % 2 == j % 2 == k % 2]
    idx = {s: q for q, s in enumerate(sites)}
    offs = np.ctypeslib.as_array(orc.lib().bwo_shell_offsets(lat, 1), shape=(24,)).reshape(8, 3)
    N = len(sites)
    A = np.zeros((N, N), dtype=np.int64)
    for q, s in enumerate(sites):
        for d in offs:
            A[q, idx[tuple((s[x] + int(d[x])) % grid[x] for x in range(3))]] += 1
    combs = np.array(list(itertools.combinations(range(N), 5)), dtype=np.int16)
    m = np.zeros(len(combs), dtype=np.int64)
    for x in range(5):
        for y in range(x + 1, 5):
            m += A[combs[:, x], combs[:, y]]
    levels, counts = np.unique(m, return_counts=True)
    return sites, combs, m, levels, counts


def level_energies(ref, sites, combs, m, levels):
    """Total energy of one placement of each level, by the oracle."""
    E = []
    for lv in levels:
        cfg = np.zeros((6, 6, 6), dtype=np.int8, order="F")
        for s in sites:
            cfg[s] = 2
        for p in combs[np.nonzero(m == lv)[0][0]]:
            cfg[sites[p]] = 1
        E.append(ref.total_energy(cfg))
    return np.array(E)
