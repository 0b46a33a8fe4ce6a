"""Canonical averages from a binned density of states (host logic, no
GPU), checked against a case the mathematics fixes: for a Gaussian density
of states exp(-(E - E0)^2 / (2 s^2)) the Boltzmann-weighted distribution at
inverse temperature beta is again Gaussian, with mean E0 - beta s^2."""
import math

import numpy as np
import pytest
from scipy.special import log_ndtr, ndtr

from brawl_amd.lattice import beta_of
from brawl_amd.wang_landau import canonical_energy


def _log_mass(a, b):
    """ln(Phi(b) - Phi(a)) for a < b, in the tails too."""
    if a >= 0.0:
        a, b = -b, -a
    if b <= 0.0:
        la, lb = log_ndtr(a), log_ndtr(b)
        return lb + math.log1p(-math.exp(la - lb))
    return math.log(ndtr(b) - ndtr(a))


def _gaussian_bins(E0, s, lo, hi, bins):
    edges = np.linspace(lo, hi, bins + 1)
    # ln of the exact number of states in each bin (up to a constant)
    z = (edges - E0) / s
    return edges, np.array([_log_mass(z[i], z[i + 1]) for i in range(bins)])


@pytest.mark.parametrize("T", [800.0, 1500.0, 3000.0])
def test_gaussian_density_of_states_gives_the_shifted_mean(T):
    n_atoms = 8192
    E0, s = -0.5, 0.2                       # Ry, whole lattice
    beta = beta_of(T)
    shift = beta * s * s
    lo, hi = E0 - shift - 12 * s, E0 + 12 * s
    # 400 bins: 7, 2.6 and 1 kT wide at the three temperatures (the
    # config-4 window set's bins are about 6 kT wide at 1000 K)
    edges, lng = _gaussian_bins(E0, s, lo, hi, 400)
    mean, outer = canonical_energy(lng, edges, n_atoms, [T])
    assert abs(mean[0] - (E0 - shift) / n_atoms) < 0.1 * s / n_atoms
    assert outer[0] < 1e-10


def test_binning_error_vanishes_as_the_bins_narrow():
    n_atoms, T = 8192, 1000.0
    E0, s = 0.0, 0.2
    beta = beta_of(T)
    exact = (E0 - beta * s * s) / n_atoms
    lo, hi = E0 - beta * s * s - 10 * s, E0 + 10 * s
    errs = []
    for bins in (100, 400, 1600, 6400):
        edges, lng = _gaussian_bins(E0, s, lo, hi, bins)
        mean, _ = canonical_energy(lng, edges, n_atoms, [T])
        errs.append(abs(mean[0] - exact) * n_atoms / s)
    assert errs == sorted(errs, reverse=True)
    assert errs[-1] < 0.005


def test_a_range_that_cuts_the_distribution_off_is_reported():
    n_atoms, T = 1000, 2000.0
    E0, s = 0.0, 0.1
    edges, lng = _gaussian_bins(E0, s, -0.1, 0.5, 300)    # cuts below E0 - s
    _, outer = canonical_energy(lng, edges, n_atoms, [T])
    assert outer[0] > 0.05
