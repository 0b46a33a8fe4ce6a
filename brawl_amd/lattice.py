"""Host-side mirror of the reference's energy and sampler interface, over the
HIP C-ABI.

`Lattice` plays the role of one `run_params` plus its module-global
`config` and `V_ex` (derived_types.f90:56-100, shared_data.f90:30-37).  Its
methods carry the reference's procedure names:

    full_energy / total_energy   setup%full_energy   (bw_hamiltonian.f90:58-81)
    nbr_energy                    setup%nbr_energy    (bw_hamiltonian.f90:898-1910)
    pair_energy                   pair_energy         (bw_hamiltonian.f90:99-114)
    initial_setup                 initial_setup       (initialise.F90:434-617)
    rdm_site / rdm_nbr            setup%rdm_site, setup%rdm_nbr (random_site.f90:127-225)
    select_sites                  wl_params%mc_select (wang-landau.F90:1527-1539)
    mc_step / mc_steps            setup%mc_step       (metropolis.F90:751-891)
    metropolis_simulated_annealing  the temperature loop (metropolis.F90:193-437)
    lattice_shells / radial_densities  analytics.f90:205-404

Arrays use the reference's Fortran layouts (config (2n1,2n2,2n3) int8,
V_ex (S,S,R) float64, both order='F').  Errors raise BrawlError with the
library's status code; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import warnings
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import AnnealParams, BrawlError, LatticeDesc, RunStats, check

FCC, BCC = _lib.FCC, _lib.BCC
K_B_IN_RY = 8.167333262e-5 / 13.605693122990   # constants.f90:45-46 (typo kept)
RY_TO_EV = 13.605693122                        # constants.f90:49
LATTICE_NAMES = {"fcc": FCC, "bcc": BCC}


def beta_of(temp: float) -> float:
    """beta = 1/(T k_B) as metropolis.F90:204-206 computes it."""
    return 1.0 / (temp * K_B_IN_RY)


def _p(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class AnnealResult:
    energies_of_T: np.ndarray   # <E>/N per T (Ry/atom)
    C_of_T: np.ndarray          # heat capacity per atom
    acceptance_of_T: np.ndarray
    trials: int
    accepted: int
    kernel_ms: float
    rho_of_T: np.ndarray = None  # (S, S, R+1, T_steps) radial densities, if sampled
    e_drift_abs: float = 0.0     # carried energy vs FP64 recounts: largest |difference| (Ry)
    e_drift_rel: float = 0.0     # ... relative to the recounted energy


def warren_cowley(rho: np.ndarray, shell_sizes, concentrations) -> np.ndarray:
    """alpha(a, b, l) = 1 - rho(a, b, l) / (z_l c_b) for shells l >= 1 of a
    radial-density array rho(S, S, wc); shell 0 (on-site) is left as 0."""
    rho = np.asarray(rho)
    S = rho.shape[0]
    a = np.zeros_like(rho)
    for l in range(1, rho.shape[2]):
        for b in range(S):
            a[:, b, l, ...] = 1.0 - rho[:, b, l, ...] / (shell_sizes[l - 1] * concentrations[b])
    return a


class Lattice:
    def __init__(self, lattice, n, v_ex):
        if isinstance(lattice, str):
            lattice = LATTICE_NAMES[lattice]
        v = np.asfortranarray(np.asarray(v_ex, dtype=np.float64))
        if v.ndim != 3 or v.shape[0] != v.shape[1]:
            raise ValueError("v_ex must have shape (S, S, R)")
        self.v_ex = v
        self.S, _, self.R = v.shape
        self.n = tuple(int(x) for x in n)
        self.lattice = int(lattice)
        self.grid = tuple(2 * x for x in self.n)
        L = _lib.load()
        d = LatticeDesc(self.lattice, self.n[0], self.n[1], self.n[2], self.S, self.R)
        h = C.c_void_p()
        check(L.bwl_create(C.byref(d), _p(self.v_ex, C.c_double), C.byref(h)))
        self._h = h
        self._L = L
        na = C.c_int32()
        check(L.bwl_n_atoms(h, C.byref(na)))
        self.n_atoms = int(na.value)
        self.layout = "auto"
        self.threads = 0

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._L.bwl_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- configuration ----------------------------------------------------
    def _cfg(self, cfg):
        cfg = np.asfortranarray(cfg, dtype=np.int8)
        if cfg.shape != self.grid:
            raise ValueError(f"config must have shape {self.grid}")
        return cfg

    def set_config(self, cfg):
        cfg = self._cfg(cfg)
        check(self._L.bwl_set_config(self._h, _p(cfg, C.c_int8)))

    def get_config(self) -> np.ndarray:
        out = np.zeros(self.grid, dtype=np.int8, order="F")
        check(self._L.bwl_get_config(self._h, _p(out, C.c_int8)))
        return out

    # -- Hamiltonian ------------------------------------------------------
    def pair_histogram(self) -> np.ndarray:
        h = np.zeros((self.S, self.S, self.R), dtype=np.int64, order="F")
        check(self._L.bwl_pair_histogram(self._h, _p(h, C.c_int64)))
        return h

    def total_energy(self) -> float:
        e = C.c_double()
        check(self._L.bwl_total_energy(self._h, C.byref(e)))
        return e.value

    full_energy = total_energy

    def nbr_energy(self, site_b, site_i, site_j, site_k) -> float:
        e = C.c_double()
        check(self._L.bwl_site_energy(self._h, site_b, site_i, site_j, site_k, C.byref(e)))
        return e.value

    def pair_energy(self, idx1, idx2) -> float:
        a = np.ascontiguousarray(idx1, dtype=np.int32)
        b = np.ascontiguousarray(idx2, dtype=np.int32)
        e = C.c_double()
        check(self._L.bwl_pair_energy(self._h, _p(a, C.c_int32), _p(b, C.c_int32), C.byref(e)))
        return e.value

    # -- reference random stream ------------------------------------------
    def seed(self, seed: int) -> None:
        check(self._L.bwl_genrand_seed(self._h, seed))

    def genrand_state(self):
        mt = np.zeros(624, dtype=np.uint32)
        mti = C.c_int32()
        check(self._L.bwl_genrand_get_state(self._h, _p(mt, C.c_uint32), C.byref(mti)))
        return mt, int(mti.value)

    def set_genrand_state(self, mt, mti) -> None:
        mt = np.ascontiguousarray(mt, dtype=np.uint32)
        check(self._L.bwl_genrand_set_state(self._h, _p(mt, C.c_uint32), int(mti)))

    def initial_setup(self, species_numbers=None, concentrations=None) -> np.ndarray:
        sn = None if species_numbers is None else np.ascontiguousarray(species_numbers, dtype=np.int64)
        cc = None if concentrations is None else np.ascontiguousarray(concentrations, dtype=np.float64)
        check(self._L.bwl_initial_setup(self._h, None if sn is None else _p(sn, C.c_int64),
                                        None if cc is None else _p(cc, C.c_double)))
        return self.get_config()

    def rdm_site(self, n: int = 1) -> np.ndarray:
        """n draws of setup%rdm_site from the context's stream: int32 (n, 4)
        rows (b, i, j, k), 1-based."""
        out = np.zeros((int(n), 4), dtype=np.int32)
        check(self._L.bwl_random_site(self._h, int(n), _p(out, C.c_int32)))
        return out

    def rdm_nbr(self, site) -> np.ndarray:
        """setup%rdm_nbr(site): a random shell-1 neighbour (one draw)."""
        s = np.ascontiguousarray(site, dtype=np.int32)
        out = np.zeros(4, dtype=np.int32)
        check(self._L.bwl_random_nbr(self._h, _p(s, C.c_int32), _p(out, C.c_int32)))
        return out

    def select_sites(self, n: int = 1, nbr_swap: bool = False) -> np.ndarray:
        """n calls of wl_params%mc_select: int32 (n, 2, 4), (rdm1, rdm2)."""
        out = np.zeros((int(n), 2, 4), dtype=np.int32)
        check(self._L.bwl_select_sites(self._h, int(bool(nbr_swap)), int(n), _p(out, C.c_int32)))
        return out

    # -- sequential (bit-exact) sampler -------------------------------------
    def mc_steps(self, beta: float, n_trials: int, *, nbr_swap=False, sample_every=0):
        """n_trials reference trials; returns (accepted, energies or None)."""
        acc = C.c_int64()
        energies = None
        ep = None
        if sample_every:
            energies = np.zeros(n_trials // sample_every)
            ep = _p(energies, C.c_double)
        check(self._L.bwl_mc_steps(self._h, beta, int(nbr_swap), n_trials, sample_every,
                                   ep, C.byref(acc)))
        return int(acc.value), energies

    def mc_step(self, beta: float, nbr_swap=False) -> int:
        return self.mc_steps(beta, 1, nbr_swap=nbr_swap)[0]

    # -- analytics ----------------------------------------------------------
    def lattice_shells(self, wc_range: int) -> np.ndarray:
        out = np.zeros(int(wc_range))
        check(self._L.bwl_lattice_shells(self._h, int(wc_range), _p(out, C.c_double)))
        return out

    def radial_densities(self, wc_range: int) -> np.ndarray:
        """rho(S, S, wc_range) of the current configuration (Fortran order)."""
        rho = np.zeros((self.S, self.S, int(wc_range)), order="F")
        check(self._L.bwl_radial_densities(self._h, int(wc_range), _p(rho, C.c_double)))
        return rho

    # -- parallel sampler -------------------------------------------------
    def set_layout(self, layout) -> None:
        """The parallel sampler's LDS layout for later anneals: "auto" (by
        rule), "compact", "natural" or "fixed" (see bwl_set_layout)."""
        code = _lib.LAYOUTS[layout] if isinstance(layout, str) else int(layout)
        check(self._L.bwl_set_layout(self._h, code))
        self.layout = layout if isinstance(layout, str) else {v: k for k, v in _lib.LAYOUTS.items()}[code]

    def set_threads(self, threads: int) -> None:
        """Threads per chain of the parallel sampler for later anneals: 0 by
        rule, else a multiple of 64 up to 1024 (see bwl_set_threads)."""
        check(self._L.bwl_set_threads(self._h, int(threads)))
        self.threads = int(threads)

    def layouts(self) -> list:
        """The layouts available to this lattice, in the rule's order of
        preference (fixed, natural, compact)."""
        m = C.c_int32()
        check(self._L.bwl_layouts(self._h, C.byref(m)))
        return [name for name in ("fixed", "natural", "compact")
                if m.value & (1 << (_lib.LAYOUTS[name] - 1))]

    def layout_info(self, layout="auto"):
        """(threads per chain, LDS bytes) of the sampler with that layout."""
        code = _lib.LAYOUTS[layout] if isinstance(layout, str) else int(layout)
        t, b = C.c_int32(), C.c_int64()
        check(self._L.bwl_layout_info(self._h, code, C.byref(t), C.byref(b)))
        return int(t.value), int(b.value)

    def metropolis_simulated_annealing(self, *, T, delta_T=0.0, T_steps=1, n_sweeps,
                                       burn_in_sweeps=0, seed=20261202,
                                       asro_every=0) -> AnnealResult:
        """With asro_every > 0, radial densities are averaged over samples
        taken every asro_every sampled sweeps (result.rho_of_T)."""
        p = AnnealParams(float(T), float(delta_T), int(T_steps), int(n_sweeps),
                         int(burn_in_sweeps), int(asro_every), int(seed) & (2**64 - 1))
        e = np.zeros(T_steps); c = np.zeros(T_steps); a = np.zeros(T_steps)
        st = RunStats()
        rho = None
        if asro_every:
            rho = np.zeros((self.S, self.S, self.R + 1, T_steps), order="F")
            check(self._L.bwl_metropolis_anneal_rho(self._h, C.byref(p), _p(e, C.c_double),
                                                    _p(c, C.c_double), _p(a, C.c_double),
                                                    _p(rho, C.c_double), C.byref(st)))
        else:
            check(self._L.bwl_metropolis_anneal(self._h, C.byref(p), _p(e, C.c_double),
                                                _p(c, C.c_double), _p(a, C.c_double), C.byref(st)))
        return AnnealResult(e, c, a, int(st.trials), int(st.accepted), float(st.kernel_ms), rho,
                            float(st.e_drift_abs), float(st.e_drift_rel))


def mc_steps_batch(lattices, betas, n_trials: int, *, nbr_swap=False, sample_every=0):
    """Batched mc_steps: every lattice makes n_trials reference trials on its
    own MT19937 stream, all in one launch.  Returns (accepted per lattice,
    energies per lattice or None, device ms)."""
    n = len(lattices)
    if n == 0 or len(betas) != n:
        raise ValueError("one beta per lattice")
    L = _lib.load()
    H = (C.c_void_p * n)(*[lat._h for lat in lattices])
    b = np.ascontiguousarray(betas, dtype=np.float64)
    acc = np.zeros(n, dtype=np.int64)
    outs = None
    DP = C.POINTER(C.c_double)
    E = None
    if sample_every:
        outs = [np.zeros(n_trials // sample_every) for _ in range(n)]
        E = (DP * n)(*[_p(o, C.c_double) for o in outs])
    ms = C.c_double()
    check(L.bwl_mc_steps_batch(n, H, _p(b, C.c_double), int(bool(nbr_swap)), int(n_trials),
                               int(sample_every), E, _p(acc, C.c_int64), C.byref(ms)))
    return acc, outs, float(ms.value)


def anneal_batch(lattices, params):
    """Batched form of metropolis_simulated_annealing: one workgroup per
    chain, all chains in as few launches as possible.  `params` is a list of
    dicts with the keyword arguments of metropolis_simulated_annealing.
    Returns (list of AnnealResult, makespan_ms).

    A chain's random numbers depend only on its seed and its position in
    that stream, so chains given equal seeds from equal stream positions
    draw the same numbers (the same pairs and acceptance uniforms) and their
    statistics are not independent.  Chain i without a "seed" key gets its
    own, derived from the master seed and i by SplitMix64; equal explicit
    seeds draw a warning."""
    from .workloads import MASTER_SEED, derive_seed
    n = len(lattices)
    if n != len(params) or n == 0:
        raise ValueError("one parameter set per lattice")
    seeds = [int(p["seed"]) if "seed" in p else derive_seed(MASTER_SEED, i)
             for i, p in enumerate(params)]
    if len(set(seeds)) != n:
        warnings.warn("anneal_batch: chains share a seed; their random streams are identical",
                      stacklevel=2)
    L = _lib.load()
    P = (AnnealParams * n)()
    outs = []
    for i, p in enumerate(params):
        T_steps = int(p.get("T_steps", 1))
        P[i] = AnnealParams(float(p["T"]), float(p.get("delta_T", 0.0)), T_steps,
                            int(p["n_sweeps"]), int(p.get("burn_in_sweeps", 0)), 0,
                            seeds[i] & (2**64 - 1))
        outs.append((np.zeros(T_steps), np.zeros(T_steps), np.zeros(T_steps)))
    H = (C.c_void_p * n)(*[lat._h for lat in lattices])
    DP = C.POINTER(C.c_double)
    E = (DP * n)(*[_p(o[0], C.c_double) for o in outs])
    Cv = (DP * n)(*[_p(o[1], C.c_double) for o in outs])
    A = (DP * n)(*[_p(o[2], C.c_double) for o in outs])
    st = (RunStats * n)()
    mk = C.c_double()
    check(L.bwl_metropolis_anneal_batch(n, H, P, E, Cv, A, st, C.byref(mk)))
    res = [AnnealResult(o[0], o[1], o[2], int(s.trials), int(s.accepted), float(s.kernel_ms), None,
                        float(s.e_drift_abs), float(s.e_drift_rel))
           for o, s in zip(outs, st)]
    return res, float(mk.value)


__all__ = ["Lattice", "BrawlError", "AnnealResult", "anneal_batch", "mc_steps_batch", "beta_of",
           "warren_cowley",
           "FCC", "BCC", "K_B_IN_RY", "RY_TO_EV"]
