"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by
the brawl_amd package.

Arrays follow the reference's Fortran layouts:
  config  int8 (2*n1, 2*n2, 2*n3) Fortran order  -> numpy order='F'
  v_ex    float64 (S, S, R) Fortran order
  hist    int64  (S, S, R) Fortran order
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

FCC, BCC = 1, 2
K_B_IN_RY = 8.167333262e-5 / 13.605693122990   # constants.f90:45-46 (typo kept)
RY_TO_EV = 13.605693122                        # constants.f90:49


class Lattice(C.Structure):
    _fields_ = [("lattice", C.c_int32), ("n1", C.c_int32), ("n2", C.c_int32),
                ("n3", C.c_int32), ("n_species", C.c_int32), ("n_shells", C.c_int32),
                ("v_ex", C.POINTER(C.c_double))]


class MT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("mti", C.c_int32)]


class MetropolisParams(C.Structure):
    _fields_ = [("T", C.c_double), ("delta_T", C.c_double), ("T_steps", C.c_int32),
                ("n_mc_steps", C.c_int32), ("n_sample_steps", C.c_int32),
                ("burn_in_start", C.c_int32), ("burn_in", C.c_int32),
                ("n_burn_in_steps", C.c_int32), ("nbr_swap", C.c_int32),
                ("calculate_asro", C.c_int32), ("n_sample_steps_asro", C.c_int32),
                ("wc_range", C.c_int32)]


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        i8p, i64p, dp, i32p = P(C.c_int8), P(C.c_int64), P(C.c_double), P(C.c_int32)
        sig = {
            "bwo_mt_init": (None, [P(MT), C.c_uint32]),
            "bwo_mt_u32": (C.c_uint32, [P(MT)]),
            "bwo_genrand": (C.c_double, [P(MT)]),
            "bwo_static_seed": (C.c_uint32, [C.c_int32]),
            "bwo_n_atoms": (C.c_int32, [P(Lattice)]),
            "bwo_shell_size": (C.c_int32, [C.c_int32, C.c_int32]),
            "bwo_shell_offsets": (P(C.c_int8), [C.c_int32, C.c_int32]),
            "bwo_site_energy": (C.c_double, [P(Lattice), i8p, C.c_int32, C.c_int32, C.c_int32]),
            "bwo_total_energy": (C.c_double, [P(Lattice), i8p]),
            "bwo_pair_energy": (C.c_double, [P(Lattice), i8p, i32p, i32p]),
            "bwo_pair_histogram": (None, [P(Lattice), i8p, i64p]),
            "bwo_energy_from_histogram": (C.c_double, [P(Lattice), i64p]),
            "bwo_random_site": (None, [P(Lattice), P(MT), i32p]),
            "bwo_random_nbr": (None, [P(Lattice), P(MT), i32p, i32p]),
            "bwo_pair_swap": (None, [P(Lattice), i8p, i32p, i32p]),
            "bwo_initial_setup": (C.c_int32, [P(Lattice), i64p, dp, P(MT), i8p]),
            "bwo_mc_step": (C.c_int32, [P(Lattice), i8p, C.c_double, C.c_int32, P(MT)]),
            "bwo_metropolis_anneal": (C.c_int32, [P(Lattice), i8p, P(MT), P(MetropolisParams),
                                                  dp, dp, dp, dp, dp]),
            "bwo_lattice_shells": (None, [P(Lattice), i8p, C.c_int32, dp]),
            "bwo_radial_densities": (None, [P(Lattice), i8p, C.c_int32, dp, dp]),
            "bwo_nested_sampling": (C.c_int32, [P(Lattice), i64p, dp, P(MT), C.c_int32,
                                                C.c_int32, C.c_int32, dp]),
            "bwo_wl_bin_index": (C.c_int32, [C.c_double, dp, C.c_int32]),
            "bwo_wl_sweeps": (C.c_int64, [P(Lattice), i8p, P(MT), C.c_int32, C.c_int32, dp,
                                          C.c_int32, C.c_int32, C.c_int32, C.c_double, dp, dp]),
            "bwo_wl_enter_window": (C.c_int64, [P(Lattice), i8p, P(MT), C.c_double, C.c_double,
                                                C.c_float, C.c_float, i64p, dp, C.c_int64]),
            "bwo_wl_setup": (None, [P(Lattice), C.c_int32, C.c_int32, C.c_float, C.c_float,
                                    C.c_float, i32p, dp]),
            "bwo_wl_replica_exchange": (C.c_int32, [P(Lattice), P(i8p), P(P(MT)), dp, C.c_int32,
                                                    dp, i32p, C.c_int32, C.c_int32, i32p, i32p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Generator:
    """The reference's global MT19937 stream (mt19937ar.c), one object per chain."""

    def __init__(self, seed: int | None = None, rank: int = 0):
        self.state = MT()
        if seed is None:
            seed = lib().bwo_static_seed(rank)
        lib().bwo_mt_init(C.byref(self.state), seed)

    def u32(self) -> int:
        return lib().bwo_mt_u32(C.byref(self.state))

    def genrand(self) -> float:
        return lib().bwo_genrand(C.byref(self.state))

    def copy(self) -> "Generator":
        g = Generator.__new__(Generator)
        g.state = MT()
        C.memmove(C.byref(g.state), C.byref(self.state), C.sizeof(MT))
        return g

    @classmethod
    def from_state(cls, mt, mti) -> "Generator":
        """A stream at a given position: the 624 words and the index."""
        g = cls.__new__(cls)
        g.state = MT()
        g.state.mt[:] = [int(x) for x in np.asarray(mt, dtype=np.uint32)]
        g.state.mti = int(mti)
        return g

    def state_arrays(self):
        return np.frombuffer(bytes(self.state.mt), dtype=np.uint32).copy(), int(self.state.mti)


class OracleLattice:
    """A lattice description plus interaction table, as the oracle sees it."""

    def __init__(self, lattice: int, n: tuple[int, int, int], v_ex: np.ndarray):
        v = np.asfortranarray(np.asarray(v_ex, dtype=np.float64))
        assert v.ndim == 3 and v.shape[0] == v.shape[1]
        self.v_ex = v
        self.S, _, self.R = v.shape
        self.n = tuple(int(x) for x in n)
        self.lattice = lattice
        self.c = Lattice(lattice, self.n[0], self.n[1], self.n[2], self.S, self.R,
                         _ptr(self.v_ex, C.c_double))

    @property
    def ref(self):
        return C.byref(self.c)

    @property
    def grid(self):
        return (2 * self.n[0], 2 * self.n[1], 2 * self.n[2])

    @property
    def n_atoms(self) -> int:
        return lib().bwo_n_atoms(self.ref)

    def empty_config(self) -> np.ndarray:
        return np.zeros(self.grid, dtype=np.int8, order="F")

    def _cfg(self, cfg):
        assert cfg.flags.f_contiguous and cfg.dtype == np.int8 and cfg.shape == self.grid
        return _ptr(cfg, C.c_int8)

    def initial_setup(self, gen: Generator, species_numbers=None, concentrations=None):
        cfg = self.empty_config()
        sn = None if species_numbers is None else np.ascontiguousarray(species_numbers, dtype=np.int64)
        cc = None if concentrations is None else np.ascontiguousarray(concentrations, dtype=np.float64)
        rc = lib().bwo_initial_setup(self.ref, None if sn is None else _ptr(sn, C.c_int64),
                                     None if cc is None else _ptr(cc, C.c_double),
                                     C.byref(gen.state), self._cfg(cfg))
        if rc:
            raise ValueError(f"initial_setup failed ({rc})")
        return cfg

    def site_energy(self, cfg, i, j, k) -> float:
        return lib().bwo_site_energy(self.ref, self._cfg(cfg), i, j, k)

    def total_energy(self, cfg) -> float:
        return lib().bwo_total_energy(self.ref, self._cfg(cfg))

    def pair_energy(self, cfg, s1, s2) -> float:
        a = np.asarray(s1, dtype=np.int32); b = np.asarray(s2, dtype=np.int32)
        return lib().bwo_pair_energy(self.ref, self._cfg(cfg), _ptr(a, C.c_int32), _ptr(b, C.c_int32))

    def histogram(self, cfg) -> np.ndarray:
        h = np.zeros((self.S, self.S, self.R), dtype=np.int64, order="F")
        lib().bwo_pair_histogram(self.ref, self._cfg(cfg), _ptr(h, C.c_int64))
        return h

    def energy_from_histogram(self, h) -> float:
        h = np.asfortranarray(h, dtype=np.int64)
        return lib().bwo_energy_from_histogram(self.ref, _ptr(h, C.c_int64))

    def random_site(self, gen):
        s = np.zeros(3, dtype=np.int32)
        lib().bwo_random_site(self.ref, C.byref(gen.state), _ptr(s, C.c_int32))
        return tuple(int(x) for x in s)

    def random_nbr(self, gen, site):
        """(i, j, k) of a random shell-1 neighbour of site (i, j, k)."""
        s = np.ascontiguousarray(site, dtype=np.int32)
        out = np.zeros(3, dtype=np.int32)
        lib().bwo_random_nbr(self.ref, C.byref(gen.state), _ptr(s, C.c_int32), _ptr(out, C.c_int32))
        return tuple(int(x) for x in out)

    def mc_step(self, cfg, beta, gen, nbr_swap=False) -> int:
        return lib().bwo_mc_step(self.ref, self._cfg(cfg), beta, int(nbr_swap), C.byref(gen.state))

    def anneal(self, cfg, gen, *, T, delta_T=0.0, T_steps=1, n_mc_steps, n_sample_steps,
               burn_in_start=False, burn_in=False, n_burn_in_steps=0, nbr_swap=False,
               calculate_asro=False, n_sample_steps_asro=1, wc_range=1, trajectory=False):
        p = MetropolisParams(T, delta_T, T_steps, n_mc_steps, n_sample_steps, int(burn_in_start),
                             int(burn_in), n_burn_in_steps, int(nbr_swap), int(calculate_asro),
                             n_sample_steps_asro, wc_range)
        n_sweeps = n_mc_steps // n_sample_steps
        traj = np.zeros((n_sweeps + 1) * T_steps) if trajectory else None
        e = np.zeros(T_steps); c = np.zeros(T_steps); a = np.zeros(T_steps)
        rho = np.zeros((self.S, self.S, wc_range, T_steps), order="F") if calculate_asro else None
        rc = lib().bwo_metropolis_anneal(
            self.ref, self._cfg(cfg), C.byref(gen.state), C.byref(p),
            None if traj is None else _ptr(traj, C.c_double), _ptr(e, C.c_double),
            _ptr(c, C.c_double), _ptr(a, C.c_double),
            None if rho is None else _ptr(rho, C.c_double))
        if rc:
            raise ValueError(f"anneal failed ({rc})")
        return dict(trajectory=traj, energies=e, C=c, acceptance=a, rho=rho)

    def lattice_shells(self, cfg, wc_range):
        s = np.zeros(wc_range)
        lib().bwo_lattice_shells(self.ref, self._cfg(cfg), wc_range, _ptr(s, C.c_double))
        return s

    def radial_densities(self, cfg, wc_range, shells):
        shells = np.ascontiguousarray(shells, dtype=np.float64)
        rho = np.zeros((self.S, self.S, wc_range), order="F")
        lib().bwo_radial_densities(self.ref, self._cfg(cfg), wc_range, _ptr(shells, C.c_double),
                                   _ptr(rho, C.c_double))
        return rho

    def nested_sampling(self, gen, *, n_walkers, n_steps, n_iter, species_numbers=None,
                        concentrations=None):
        sn = None if species_numbers is None else np.ascontiguousarray(species_numbers, dtype=np.int64)
        cc = None if concentrations is None else np.ascontiguousarray(concentrations, dtype=np.float64)
        out = np.zeros(n_iter)
        rc = lib().bwo_nested_sampling(self.ref, None if sn is None else _ptr(sn, C.c_int64),
                                       None if cc is None else _ptr(cc, C.c_double),
                                       C.byref(gen.state), n_walkers, n_steps, n_iter,
                                       _ptr(out, C.c_double))
        if rc:
            raise ValueError(f"nested sampling failed ({rc})")
        return out

    def wl_setup(self, bins, num_windows, bin_overlap, energy_min, energy_max):
        w = np.zeros(2 * num_windows, dtype=np.int32)
        edges = np.zeros(bins + 1)
        lib().bwo_wl_setup(self.ref, bins, num_windows, bin_overlap, energy_min, energy_max,
                           _ptr(w, C.c_int32), _ptr(edges, C.c_double))
        return w.reshape(num_windows, 2), edges

    def wl_sweeps(self, cfg, gen, *, mc_sweeps, bin_edges, start_idx, end_idx, wl_f, lng, hist,
                  nbr_swap=False):
        bins = len(bin_edges) - 1
        assert lng.dtype == np.float64 and hist.dtype == np.float64
        return lib().bwo_wl_sweeps(self.ref, self._cfg(cfg), C.byref(gen.state), int(nbr_swap),
                                   mc_sweeps, _ptr(bin_edges, C.c_double), bins, start_idx,
                                   end_idx, wl_f, _ptr(lng, C.c_double), _ptr(hist, C.c_double))

    def wl_enter_window(self, cfg, gen, min_e, max_e, energy_min, energy_max, *,
                        species_numbers=None, concentrations=None, max_trials=10**9):
        sn = None if species_numbers is None else np.ascontiguousarray(species_numbers, dtype=np.int64)
        cc = None if concentrations is None else np.ascontiguousarray(concentrations, dtype=np.float64)
        return lib().bwo_wl_enter_window(self.ref, self._cfg(cfg), C.byref(gen.state), min_e, max_e,
                                         energy_min, energy_max,
                                         None if sn is None else _ptr(sn, C.c_int64),
                                         None if cc is None else _ptr(cc, C.c_double), max_trials)


def wl_replica_exchange(lat: "OracleLattice", cfgs, gens, lng, bin_edges, windows, wpw):
    """replica_exchange (wang-landau.F90:1392-1501) over walkers whose
    configurations (order='F' int8 grids) and Generators are listed in walker
    order; swaps configurations in place.  Returns (exchanged, partner,
    accepted)."""
    W = len(cfgs)
    windows = np.ascontiguousarray(np.asarray(windows, dtype=np.int32).reshape(-1))
    lng = np.ascontiguousarray(lng, dtype=np.float64)
    edges = np.ascontiguousarray(bin_edges, dtype=np.float64)
    assert lng.shape == (W, len(edges) - 1) and W == (len(windows) // 2) * wpw
    for c in cfgs:
        assert c.dtype == np.int8 and c.flags.f_contiguous
    cp = (C.POINTER(C.c_int8) * W)(*[_ptr(c, C.c_int8) for c in cfgs])
    gp = (C.POINTER(MT) * W)(*[C.pointer(g.state) for g in gens])
    partner = np.zeros(W, dtype=np.int32)
    acc = np.zeros(W, dtype=np.int32)
    n = lib().bwo_wl_replica_exchange(lat.ref, cp, gp, _ptr(lng, C.c_double), len(edges) - 1,
                                      _ptr(edges, C.c_double), _ptr(windows, C.c_int32),
                                      len(windows) // 2, int(wpw), _ptr(partner, C.c_int32),
                                      _ptr(acc, C.c_int32))
    return int(n), partner, acc


def read_vij(path: str, n_species: int, n_shells: int) -> np.ndarray:
    """read_exchange (io.f90:399-402): list-directed read filling V_ex(S,S,R)
    in column-major order."""
    vals = np.array(open(path).read().split(), dtype=np.float64)
    need = n_species * n_species * n_shells
    return vals[:need].reshape((n_species, n_species, n_shells), order="F")
