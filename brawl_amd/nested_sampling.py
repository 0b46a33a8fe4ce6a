"""Nested sampling over the HIP C-ABI (nested_sampling.f90:45-206).

`nested_sampling` runs the reference's serial loop for one chain on the
device; `nested_sampling_batch` runs many independent chains (requests) in
one launch, one wavefront each.  Each run consumes its lattice's MT19937
stream exactly as the reference does, so its energy limits are the
reference's for the same stream, alone or in a batch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import NsParams, RunStats, check
from .lattice import Lattice


def _sn(species_numbers):
    return None if species_numbers is None else np.ascontiguousarray(species_numbers, dtype=np.int64)


def _cc(concentrations):
    return None if concentrations is None else np.ascontiguousarray(concentrations, dtype=np.float64)


def nested_sampling(lat: Lattice, *, n_walkers: int, n_steps: int, n_iter: int,
                    species_numbers=None, concentrations=None):
    """Returns (e_limits[n_iter], RunStats) for one run on `lat`'s stream."""
    res, _ = nested_sampling_batch([lat], [dict(n_walkers=n_walkers, n_steps=n_steps, n_iter=n_iter,
                                                species_numbers=species_numbers,
                                                concentrations=concentrations)])
    return res[0]


def nested_sampling_batch(lattices, params):
    """params: one dict per lattice with n_walkers, n_steps, n_iter and
    species_numbers / concentrations.  Returns ([(e_limits, stats)], makespan_ms)."""
    n = len(lattices)
    if n == 0 or n != len(params):
        raise ValueError("one parameter set per lattice")
    L = _lib.load()
    P = (NsParams * n)()
    outs, sns, ccs = [], [], []
    for i, p in enumerate(params):
        P[i] = NsParams(int(p["n_walkers"]), int(p["n_steps"]), int(p["n_iter"]), 0)
        outs.append(np.zeros(int(p["n_iter"])))
        sns.append(_sn(p.get("species_numbers")))
        ccs.append(_cc(p.get("concentrations")))
    H = (C.c_void_p * n)(*[lat._h for lat in lattices])
    I64P, DP = C.POINTER(C.c_int64), C.POINTER(C.c_double)
    SN = (I64P * n)(*[I64P() if a is None else a.ctypes.data_as(I64P) for a in sns])
    CC = (DP * n)(*[DP() if a is None else a.ctypes.data_as(DP) for a in ccs])
    E = (DP * n)(*[o.ctypes.data_as(DP) for o in outs])
    st = (RunStats * n)()
    mk = C.c_double()
    check(L.bwl_nested_sampling_batch(n, H, P, SN, CC, E, st, C.byref(mk)))
    return [(o, s) for o, s in zip(outs, st)], float(mk.value)


__all__ = ["nested_sampling", "nested_sampling_batch"]
