"""Wang-Landau walkers over the HIP C-ABI (wang-landau.F90).

`WangLandau` holds num_windows x walkers_per_window walkers on one device
(walker k is the reference's MPI rank k, in window k // walkers_per_window,
drawing from MT19937 seed 110179 + 11*k).  Its methods are the reference's
per-rank routines run for every walker at once:

    enter_windows     enter_energy_window      (wang-landau.F90:643-741)
    sweeps            sweeps + window average  (wang-landau.F90:539-632)
    replica_exchange  replica_exchange         (wang-landau.F90:1392-1501)
    pre_sampling      pre_sampling's loop      (wang-landau.F90:757-838)
    dos_average       dos_average              (wang-landau.F90:1127-1133)
    dos_combine       dos_combine              (wang-landau.F90:1147-1194)

`run` is wl_main (wang-landau.F90:160-306): window entry, pre-sampling,
then the main loop -- sweeps and a replica exchange until every walker's
window histogram is flat, then halve ln f, average and stitch ln g, until
ln f falls below the tolerance.  Not carried over: window load balancing
(mpi_window_optimise, :1211-1328, driven by MPI timings), rho_of_E
sampling, and the per-bin configuration stores with the reloads they feed
(merge_configs, load_window_config: Fortran's random_number picks them) --
the parts that exist to balance MPI ranks or write auxiliary output.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import LatticeDesc, WlParams, check
from .lattice import LATTICE_NAMES


def _p(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class WlRunResult:
    lng: np.ndarray          # combined ln g over all bins (min subtracted)
    bin_edges: np.ndarray
    iterations: int          # ln f halvings done
    sweeps_calls: int        # main loop and pre-sampling
    trials: int
    kernel_ms: float
    pre_sampling_calls: int = 0
    exchanges: int = 0       # configurations exchanged (pairs)
    exchange_pairs: int = 0  # pairs that drew for an exchange


class WangLandau:
    def __init__(self, lattice, n, v_ex, *, bins, num_windows, walkers_per_window,
                 bin_overlap, energy_min, energy_max, mc_sweeps=1, nbr_swap=False,
                 rank_offset=0):
        if isinstance(lattice, str):
            lattice = LATTICE_NAMES[lattice]
        v = np.asfortranarray(np.asarray(v_ex, dtype=np.float64))
        self.S, _, self.R = v.shape
        self.n = tuple(int(x) for x in n)
        self.grid = tuple(2 * x for x in self.n)
        self.lattice = int(lattice)
        self.n_atoms = (4 if self.lattice == _lib.FCC else 2) * self.n[0] * self.n[1] * self.n[2]
        self.v_ex = v
        self.params = WlParams(int(bins), int(num_windows), int(walkers_per_window),
                               float(bin_overlap), float(energy_min), float(energy_max),
                               int(bool(nbr_swap)), int(mc_sweeps), int(rank_offset))
        L = _lib.load()
        d = LatticeDesc(int(lattice), self.n[0], self.n[1], self.n[2], self.S, self.R)
        h = C.c_void_p()
        check(L.bwl_wl_create(C.byref(d), _p(v, C.c_double), C.byref(self.params), C.byref(h)))
        self._h, self._L = h, L
        nw, hs = C.c_int32(), C.c_int32()
        check(L.bwl_wl_n_walkers(h, C.byref(nw)))
        check(L.bwl_wl_hist_stride(h, C.byref(hs)))
        self.n_walkers, self.hist_stride = int(nw.value), int(hs.value)
        self.bins, self.num_windows = int(bins), int(num_windows)
        self.walkers_per_window = int(walkers_per_window)
        w = np.zeros(2 * self.num_windows, dtype=np.int32)
        e = np.zeros(self.bins + 1)
        check(L.bwl_wl_windows(h, _p(w, C.c_int32), _p(e, C.c_double)))
        self.windows = w.reshape(self.num_windows, 2)   # 1-based (start, end) bins
        self.bin_edges = e

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._L.bwl_wl_destroy(self._h)
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

    # -- walkers ------------------------------------------------------------
    def window_of(self, k: int) -> int:
        return k // self.walkers_per_window

    def initial_setup(self, species_numbers=None, concentrations=None):
        sn = None if species_numbers is None else np.ascontiguousarray(species_numbers, dtype=np.int64)
        cc = None if concentrations is None else np.ascontiguousarray(concentrations, dtype=np.float64)
        check(self._L.bwl_wl_initial_setup(self._h, None if sn is None else _p(sn, C.c_int64),
                                           None if cc is None else _p(cc, C.c_double)))

    def set_walker(self, k: int, cfg):
        cfg = np.asfortranarray(cfg, dtype=np.int8)
        check(self._L.bwl_wl_set_walker(self._h, int(k), _p(cfg, C.c_int8)))

    def get_walker(self, k: int) -> np.ndarray:
        out = np.zeros(self.grid, dtype=np.int8, order="F")
        check(self._L.bwl_wl_get_walker(self._h, int(k), _p(out, C.c_int8)))
        return out

    def genrand_state(self, k: int):
        mt = np.zeros(624, dtype=np.uint32)
        mti = C.c_int32()
        check(self._L.bwl_wl_genrand_get_state(self._h, int(k), _p(mt, C.c_uint32), C.byref(mti)))
        return mt, int(mti.value)

    def set_genrand_state(self, k: int, mt, mti: int):
        mt = np.ascontiguousarray(mt, dtype=np.uint32)
        check(self._L.bwl_wl_genrand_set_state(self._h, int(k), _p(mt, C.c_uint32), int(mti)))

    def lng_hist(self):
        lng = np.zeros((self.n_walkers, self.bins))
        hist = np.zeros((self.n_walkers, self.hist_stride))
        check(self._L.bwl_wl_get_lng(self._h, _p(lng, C.c_double), _p(hist, C.c_double)))
        return lng, hist

    def set_lng_hist(self, lng=None, hist=None):
        lp = hp = None
        if lng is not None:
            lng = np.ascontiguousarray(lng, dtype=np.float64)
            assert lng.shape == (self.n_walkers, self.bins)
            lp = _p(lng, C.c_double)
        if hist is not None:
            hist = np.ascontiguousarray(hist, dtype=np.float64)
            assert hist.shape == (self.n_walkers, self.hist_stride)
            hp = _p(hist, C.c_double)
        check(self._L.bwl_wl_set_lng(self._h, lp, hp))

    # -- the reference's routines ------------------------------------------
    def enter_windows(self, max_trials: int = 10**9):
        tr = np.zeros(self.n_walkers, dtype=np.int64)
        en = np.zeros(self.n_walkers, dtype=np.int32)
        check(self._L.bwl_wl_enter_windows(self._h, int(max_trials), _p(tr, C.c_int64),
                                           _p(en, C.c_int32)))
        return tr, en.astype(bool)

    def sweeps(self, wl_f: float, average: bool = True):
        acc = np.zeros(self.n_walkers, dtype=np.int64)
        ms = C.c_double()
        check(self._L.bwl_wl_sweeps(self._h, float(wl_f), int(bool(average)), _p(acc, C.c_int64),
                                    C.byref(ms)))
        return acc, float(ms.value)

    def replica_exchange(self):
        """One replica_exchange over all walkers (wang-landau.F90:1392-1501);
        returns (partner, exchanged, kernel_ms): partner[k] is the walker
        matched with k (-1 if none), exchanged[k] whether that pair swapped
        configurations."""
        partner = np.zeros(self.n_walkers, dtype=np.int32)
        exch = np.zeros(self.n_walkers, dtype=np.int32)
        ms = C.c_double()
        check(self._L.bwl_wl_replica_exchange(self._h, _p(partner, C.c_int32), _p(exch, C.c_int32),
                                              C.byref(ms)))
        return partner, exch.astype(bool), float(ms.value)

    def pre_sampling(self, wl_f: float, flatness: float, *, exchange: bool = True,
                     max_sweeps_calls=None):
        """pre_sampling's loop (wang-landau.F90:757-785): sweeps with window
        averages, a replica exchange after every 10th call (the counter then
        restarts at 1, so every 9th after the first), until every window has
        been flat once -- histogram minimum above 1000/walkers_per_window and
        flatness above the threshold, judged per window (the MPI_MAX of its
        walkers).  Returns (sweeps calls, device ms, accepted trials,
        exchanged pairs, pairs that drew, whether every window got flat)."""
        done = np.zeros(self.num_windows, dtype=bool)
        wpw = self.walkers_per_window
        i_sweeps = calls = acc_total = exch = pairs = 0
        ms = 0.0
        while not done.all():
            if max_sweeps_calls is not None and calls >= max_sweeps_calls:
                break
            acc, k_ms = self.sweeps(wl_f, average=True)
            calls += 1
            ms += k_ms
            acc_total += int(acc.sum())
            i_sweeps += 1
            if exchange and i_sweeps % 10 == 0:
                i_sweeps = 1
                partner, ex, x_ms = self.replica_exchange()
                ms += x_ms
                exch += int(ex.sum()) // 2
                pairs += int((partner >= 0).sum()) // 2
            _, hist = self.lng_hist()
            flat = self.flatness(hist)
            for w in range(self.num_windows):
                if done[w]:
                    continue
                nb = int(self.windows[w, 1] - self.windows[w, 0] + 1)
                for k in range(w * wpw, (w + 1) * wpw):
                    if hist[k, :nb].min() > 1000.0 / wpw and flat[k] > flatness:
                        done[w] = True
        return calls, ms, acc_total, exch, pairs, bool(done.all())

    def dos_average(self, lng: np.ndarray) -> np.ndarray:
        """Mean of each window's ln g over its walkers, summed in walker order."""
        out = lng.copy()
        wpw = self.walkers_per_window
        for w in range(self.num_windows):
            rows = lng[w * wpw:(w + 1) * wpw]
            s = rows[0].copy()
            for r in rows[1:]:
                s = s + r
            out[w * wpw:(w + 1) * wpw] = s / float(wpw)
        return out

    def dos_combine(self, lng: np.ndarray) -> np.ndarray:
        """Stitch the windows' ln g where their slopes agree best
        (wang-landau.F90:1147-1194); returns the combined ln g (min 0)."""
        wpw = self.walkers_per_window
        comb = lng[0].copy()
        beta_index = 0
        for i in range(2, self.num_windows + 1):
            buf = lng[(i - 1) * wpw]
            start, end = int(self.windows[i - 1, 0]), int(self.windows[i - 1, 1])
            beta_diff = np.finfo(np.float64).max
            for j in range(0, int(self.windows[i - 2, 1]) - start):
                bo = comb[start + j] - comb[start + j - 1]      # 1-based (s+j+1) - (s+j)
                bm = buf[start + j] - buf[start + j - 1]
                if abs(bo - bm) < beta_diff:
                    beta_diff = abs(bo - bm)
                    beta_index = start + j + 1
            for j in range(beta_index, end + 1):
                comb[j - 1] = buf[j - 1] + comb[beta_index - 1] - buf[beta_index - 1]
        return comb - comb.min()

    def flatness(self, hist: np.ndarray) -> np.ndarray:
        """min(hist)/(sum(hist)/window bins) per walker (wang-landau.F90:222)."""
        out = np.zeros(self.n_walkers)
        for k in range(self.n_walkers):
            w = self.window_of(k)
            nb = int(self.windows[w, 1] - self.windows[w, 0] + 1)
            h = hist[k, :nb]
            s = h.sum()
            out[k] = h.min() / (s / nb) if s > 0 else 0.0
        return out

    def _halve_setup(self, lng, hist):
        """dos_average then dos_combine, the stitched ln g written to every
        walker and the histograms cleared."""
        lng = self.dos_average(lng)
        comb = self.dos_combine(lng)
        self.set_lng_hist(np.tile(comb, (self.n_walkers, 1)), np.zeros_like(hist))

    def run(self, *, wl_f=1.0, tolerance=1e-6, flatness=0.8, max_sweeps_calls=None,
            enter_max_trials=10**9, pre_sampling=True, replica_exchange=True,
            max_pre_sampling_calls=None, progress=None):
        """wl_main (see module docstring for what is left out).  With
        pre_sampling, ln f stays at wl_f through the pre-sampling loop, whose
        ln g (averaged and stitched) starts the main loop; with
        replica_exchange, the main loop exchanges after every sweeps call
        (both as the reference runs by default, performance = 0).
        progress, if given, is called as progress(halvings, ln_f, calls)
        after every halving of ln f and every 2000 sweeps calls."""
        _, entered = self.enter_windows(enter_max_trials)
        if not entered.all():
            raise RuntimeError("some walkers did not reach their energy window")
        lng, hist = self.lng_hist()
        self.set_lng_hist(np.zeros_like(lng), np.zeros_like(hist))
        trials_per_call = self.params.mc_sweeps * self.n_atoms * self.n_walkers
        calls = iters = pre_calls = exch = pairs = 0
        ms = 0.0
        if pre_sampling:
            pre_calls, p_ms, _, exch, pairs, flat = self.pre_sampling(
                wl_f, flatness, exchange=replica_exchange, max_sweeps_calls=max_pre_sampling_calls)
            if not flat:
                raise RuntimeError("pre-sampling: not every window became flat")
            ms += p_ms
            lng, hist = self.lng_hist()
            self._halve_setup(lng, hist)
        converged = np.zeros(self.n_walkers, dtype=bool)
        acc_total = 0
        while wl_f > tolerance:
            if max_sweeps_calls is not None and calls >= max_sweeps_calls:
                break
            acc, k_ms = self.sweeps(wl_f, average=True)
            calls += 1
            ms += k_ms
            acc_total += int(acc.sum())
            if progress is not None and calls % 2000 == 0:
                progress(iters, wl_f, calls)
            if replica_exchange:
                partner, ex, x_ms = self.replica_exchange()
                ms += x_ms
                exch += int(ex.sum()) // 2
                pairs += int((partner >= 0).sum()) // 2
            lng, hist = self.lng_hist()
            converged |= self.flatness(hist) > flatness
            if converged.all():
                converged[:] = False
                wl_f = wl_f * 0.5
                self._halve_setup(lng, hist)
                iters += 1
                if progress is not None:
                    progress(iters, wl_f, calls)
        lng, _ = self.lng_hist()
        return WlRunResult(lng[0].copy(), self.bin_edges.copy(), iters, calls + pre_calls,
                           (calls + pre_calls) * trials_per_call, ms, pre_calls, exch, pairs)


def canonical_energy(lng, bin_edges, n_atoms, temps):
    """Canonical averages from a binned density of states: for each
    temperature (K) the mean energy per atom <E>/N (Ry) and the share of the
    Boltzmann weight that falls in the outer 5 % of the energy range (large
    when the range cuts the distribution off, and the mean is then biased).
    lng holds ln g per bin up to a constant and bin_edges the bins' total
    energies (Ry), as WlRunResult gives them.  Each bin's states are placed
    at its centre; the error this makes shrinks with the bins' width
    relative to kT (for a Gaussian density of states about 0.06 of the
    distribution's spread with bins 7 kT wide, 0.007 with bins 1 kT wide)."""
    from .lattice import beta_of
    edges = np.asarray(bin_edges, dtype=np.float64)
    lng = np.asarray(lng, dtype=np.float64)
    centres = 0.5 * (edges[:-1] + edges[1:])
    span = edges[-1] - edges[0]
    outer = (centres < edges[0] + 0.05 * span) | (centres > edges[-1] - 0.05 * span)
    means, edge_share = [], []
    for T in temps:
        w = lng - beta_of(T) * centres
        p = np.exp(w - w.max())
        means.append(float((p * centres).sum() / p.sum()) / n_atoms)
        edge_share.append(float(p[outer].sum() / p.sum()))
    return np.array(means), np.array(edge_share)


__all__ = ["WangLandau", "WlRunResult", "canonical_energy"]
