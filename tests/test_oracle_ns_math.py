"""The oracle's nested-sampling loop (the restatement of
nested_sampling.f90:78-200) against what the algorithm must produce on a
lattice whose density of states is known exactly (tests/exact_dos.py:
5 solutes on bcc 3^3, one shell, E = E_0 + m J).

With walks long enough to mix, a clone is a uniform draw from the
configurations below the current limit, so the number of iterations spent
at each energy level depends only on the densities of states, the walkers'
starting levels and their tie-breaking offsets (the reference adds
genrand()*1e-8 to each starting energy, nested_sampling.f90:94-95, and a
clone inherits its parent's offset, so a clone may stay on the limit's own
level).  An independent simulation of the algorithm on the levels alone --
maxloc of the walkers' energies, a clone of a uniformly chosen walker, a
new level drawn with probability proportional to the exact counts among
the levels strictly below the limit for the clone's offset -- started from
the same walkers (the reference's initial_setup on the same streams) must
give the same mean iterations per level.  A loop that dropped the offsets
(every level then ends after a binomial number of iterations) is rejected
by the same comparison."""
import numpy as np

from exact_dos import J, SPECIES_NUMBERS, exact_dos, level_energies, v_ex

K, STEPS, NITER, RUNS = 64, 3000, 200, 80


def _ideal_counts(rng, w, eps, G, with_offsets=True):
    c = np.zeros(len(G))
    w, eps = w.astype(float).copy(), eps.copy()
    if not with_offsets:
        eps[:] = 0.0
    top = len(G) - 1
    for _ in range(NITER):
        e = w * J + eps
        i = int(np.argmax(e))   # the first maximum, as maxloc
        lim = e[i]
        c[int(w[i])] += 1
        src = int(rng.integers(0, len(w)))
        es = eps[src]
        mmax = top
        while mmax >= 0 and mmax * J + es >= lim:
            mmax -= 1
        if mmax < 0:            # no level below the limit: the clone stays
            w[i], eps[i] = w[src], es
            continue
        w[i] = np.searchsorted(G, rng.random() * G[mmax], side="right")
        eps[i] = es
    return c


def test_nested_sampling_levels_match_ideal_algorithm(oracle):
    sites, combs, m, levels, counts = exact_dos()
    assert list(levels) == list(range(7)) and counts.sum() == 3162510
    G = np.cumsum(counts)
    lat, n = 2, (3, 3, 3)
    ref = oracle.OracleLattice(lat, n, v_ex())
    E = level_energies(ref, sites, combs, m, levels)
    assert np.allclose(np.diff(E), J, rtol=1e-12, atol=0)
    rng = np.random.default_rng(20261203)
    obs, ideal, plain = [], [], []
    for r in range(RUNS):
        # the walkers' starting levels and offsets: initial_setup then one
        # genrand per walker, on the stream the run itself uses
        g = oracle.Generator(rank=r)
        w, eps = [], []
        for _ in range(K):
            cfg = ref.initial_setup(g, species_numbers=SPECIES_NUMBERS)
            eps.append(g.genrand() * 1e-8)
            w.append(int(np.rint((ref.total_energy(cfg) - E[0]) / J)))
        w, eps = np.array(w), np.array(eps)
        lim = ref.nested_sampling(oracle.Generator(rank=r), n_walkers=K, n_steps=STEPS, n_iter=NITER,
                                  species_numbers=SPECIES_NUMBERS)
        lv = np.rint((lim - E[0]) / J).astype(int)
        assert np.all(np.abs(lim - (E[0] + lv * J)) < 2e-8)
        obs.append(np.bincount(lv, minlength=7))
        ideal.append(_ideal_counts(rng, w, eps, G))
        plain.append(_ideal_counts(rng, w, eps, G, with_offsets=False))
    obs, ideal, plain = np.array(obs), np.array(ideal), np.array(plain)

    def z(a, b):
        return (a.mean(0) - b.mean(0)) / np.sqrt(a.var(0, ddof=1) / RUNS + b.var(0, ddof=1) / RUNS + 1e-30)

    zi, zp = z(obs, ideal), z(obs, plain)
    print("iterations per level (m = 6..0): oracle", np.round(obs.mean(0)[::-1], 2),
          "ideal", np.round(ideal.mean(0)[::-1], 2), "z", np.round(zi[::-1], 2))
    # levels 1..5 carry the information (level 6 is nearly empty, level 0
    # takes the remaining iterations)
    assert np.all(np.abs(zi[1:6]) <= 3.0), zi
    # the comparison has the power to see a loop without the offsets
    assert np.max(np.abs(zp[1:6])) > 5.0, zp
