"""TEST INFRASTRUCTURE ONLY: the reference's own routines, compiled from the
reference's Fortran/C sources by oracle/ref.mk into oracle/_ref/ (never
committed).  Used by tests/ to check the C restatement (oracle.py) against
the reference itself on inputs its golden files do not cover.

available() is False when the reference build is absent (the reference
sources are not present, or the build was not run); callers skip then.
"""
from __future__ import annotations

import ctypes as C
import multiprocessing as mp
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(_HERE, "_ref")
LIB_PATH = os.path.join(REF_DIR, "libbrawl_ref.so")
DROP_IN = os.path.join(REF_DIR, "libdrop_in.so")
REF_SRC = "/root/reference/src"
_lib = None


def build() -> bool:
    """Build oracle/_ref when the reference sources are present; returns
    whether the build is available afterwards."""
    if os.path.isdir(REF_SRC):
        subprocess.check_call(["make", "-s", "-f", os.path.join(_HERE, "ref.mk")])
    return available()


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        i32, i8p, dp, i64p, i32p = C.c_int32, P(C.c_int8), P(C.c_double), P(C.c_int64), P(C.c_int32)
        sig = {
            "ref_total_energy": (C.c_double, [i32] * 6 + [dp, i8p]),
            "ref_site_energies": (None, [i32] * 6 + [dp, i8p, dp]),
            "ref_initial_setup": (None, [i32] * 6 + [i64p, i8p]),
            "ref_initial_setup_conc": (None, [i32] * 6 + [dp, i8p]),
            "ref_random_sites": (None, [i32] * 6 + [i32p]),
            "ref_random_nbrs": (None, [i32] * 6 + [i32p]),
            "ref_radial_densities": (None, [i32] * 6 + [i8p, dp, dp]),
            "ref_metropolis_sweeps": (C.c_int64, [i32] * 6 + [dp, i8p, C.c_double, i32, i32,
                                                             dp, dp]),
            "ref_wl_sweeps": (C.c_int64, [i32] * 6 + [dp, i8p] + [i32] * 5 + [dp, i32, i32,
                                                                         C.c_double, dp, dp]),
            "ref_wl_enter": (C.c_int64, [i32] * 6 + [dp, i8p, i32, C.c_double, C.c_double,
                                                     C.c_float, C.c_float, C.c_int64]),
            "ref_nested_sampling": (C.c_int32, [i32] * 6 + [dp, i32, i64p, i32, i32, i32, dp, i64p,
                                                            dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _isolated(fn):
    """Run fn in a forked child: the reference stops the whole process (a
    Fortran STOP exits with status 0) on inputs it rejects, which must not
    end the caller."""
    def wrapper(*args, **kw):
        ctx = mp.get_context("fork")
        rd, wr = ctx.Pipe(duplex=False)

        def child():
            try:
                wr.send(("ok", fn(*args, **kw)))
            except BaseException as e:  # report, never hang the parent
                wr.send(("err", repr(e)))
            wr.close()

        p = ctx.Process(target=child)
        p.start()
        wr.close()
        try:
            status, val = rd.recv()
        except EOFError:
            p.join()
            raise RuntimeError(f"reference routine {fn.__name__} ended the process "
                               f"(exit code {p.exitcode}): it rejected its input") from None
        p.join()
        if status != "ok":
            raise RuntimeError(f"reference routine {fn.__name__} failed: {val}")
        return val
    wrapper.__name__ = fn.__name__
    wrapper.__doc__ = fn.__doc__
    return wrapper


def _p(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _cfg(cfg):
    cfg = np.asfortranarray(cfg, dtype=np.int8)
    return cfg, _p(cfg, C.c_int8)


@_isolated
def total_energy(lattice, n, v_ex, cfg) -> float:
    v = np.asfortranarray(v_ex, dtype=np.float64)
    S, _, R = v.shape
    c, cp = _cfg(cfg)
    return lib().ref_total_energy(lattice, *n, S, R, _p(v, C.c_double), cp)


@_isolated
def site_energies(lattice, n, v_ex, cfg, n_atoms) -> np.ndarray:
    """setup%nbr_energy of every occupied site, sites in z, y, x order."""
    v = np.asfortranarray(v_ex, dtype=np.float64)
    S, _, R = v.shape
    c, cp = _cfg(cfg)
    out = np.zeros(n_atoms)
    lib().ref_site_energies(lattice, *n, S, R, _p(v, C.c_double), cp, _p(out, C.c_double))
    return out


@_isolated
def initial_setup(lattice, n, species_numbers, rank=0) -> np.ndarray:
    sn = np.ascontiguousarray(species_numbers, dtype=np.int64)
    cfg = np.zeros(tuple(2 * x for x in n), dtype=np.int8, order="F")
    lib().ref_initial_setup(lattice, *n, len(sn), rank, _p(sn, C.c_int64), _p(cfg, C.c_int8))
    return cfg


@_isolated
def initial_setup_conc(lattice, n, concentrations, rank=0) -> np.ndarray:
    """initial_setup on the concentrations path (species numbers 0)."""
    cc = np.ascontiguousarray(concentrations, dtype=np.float64)
    cfg = np.zeros(tuple(2 * x for x in n), dtype=np.int8, order="F")
    lib().ref_initial_setup_conc(lattice, *n, len(cc), rank, _p(cc, C.c_double), _p(cfg, C.c_int8))
    return cfg


@_isolated
def random_sites(lattice, n, count, rank=0) -> np.ndarray:
    out = np.zeros((count, 4), dtype=np.int32)
    lib().ref_random_sites(lattice, *n, rank, count, _p(out, C.c_int32))
    return out


@_isolated
def random_nbrs(lattice, n, count, rank=0) -> np.ndarray:
    """count (rdm_site, rdm_nbr) pairs: int32 (count, 2, 4)."""
    out = np.zeros((count, 2, 4), dtype=np.int32)
    lib().ref_random_nbrs(lattice, *n, rank, count, _p(out, C.c_int32))
    return out


@_isolated
def radial_densities(lattice, n, S, wc, cfg):
    c, cp = _cfg(cfg)
    shells = np.zeros(wc)
    rho = np.zeros((S, S, wc), order="F")
    lib().ref_radial_densities(lattice, *n, S, wc, cp, _p(shells, C.c_double), _p(rho, C.c_double))
    return shells, rho


def _metropolis_sweeps(lattice, n, v_ex, cfg, beta, n_sweeps, rank=0):
    """The reference's serial Metropolis path (ref_metropolis_sweeps):
    n_sweeps sweeps of n_atoms trials at beta after seeding for `rank`.
    Returns (accepted, last full energy, seconds of the trial loops, final
    configuration)."""
    v = np.asfortranarray(v_ex, dtype=np.float64)
    S, _, R = v.shape
    c, cp = _cfg(np.array(cfg, dtype=np.int8, order="F", copy=True))
    e, secs = C.c_double(0.0), C.c_double(0.0)
    acc = lib().ref_metropolis_sweeps(lattice, *n, S, R, _p(v, C.c_double), cp, float(beta),
                                      int(n_sweeps), int(rank), C.byref(e), C.byref(secs))
    return int(acc), float(e.value), float(secs.value), np.array(c)


metropolis_sweeps = _isolated(_metropolis_sweeps)


def _wl_sweeps(lattice, n, v_ex, cfg, *, rank, nbr_swap, mc_sweeps, n_calls, bin_edges, start_idx,
               end_idx, wl_f, lng, hist):
    """n_calls `sweeps` calls of one walker from the stream of `rank`
    (ref_wl_sweeps); returns (accepted, configuration, ln g, histogram)."""
    v = np.asfortranarray(v_ex, dtype=np.float64)
    S, _, R = v.shape
    c, cp = _cfg(np.array(cfg, dtype=np.int8, order="F", copy=True))
    e = np.ascontiguousarray(bin_edges, dtype=np.float64)
    lg = np.array(lng, dtype=np.float64, copy=True)
    h = np.array(hist, dtype=np.float64, copy=True)
    acc = lib().ref_wl_sweeps(lattice, *n, S, R, _p(v, C.c_double), cp, int(rank), int(nbr_swap),
                              int(mc_sweeps), int(n_calls), len(e) - 1, _p(e, C.c_double),
                              int(start_idx), int(end_idx), float(wl_f), _p(lg, C.c_double),
                              _p(h, C.c_double))
    return int(acc), np.array(c), lg, h


wl_sweeps = _isolated(_wl_sweeps)


@_isolated
def wl_enter(lattice, n, v_ex, cfg, *, rank, min_e, max_e, energy_min, energy_max, max_trials):
    """enter_energy_window of one walker from the stream of `rank`
    (ref_wl_enter); returns (trials, configuration)."""
    v = np.asfortranarray(v_ex, dtype=np.float64)
    S, _, R = v.shape
    c, cp = _cfg(np.array(cfg, dtype=np.int8, order="F", copy=True))
    t = lib().ref_wl_enter(lattice, *n, S, R, _p(v, C.c_double), cp, int(rank), float(min_e),
                           float(max_e), float(energy_min), float(energy_max), int(max_trials))
    return int(t), np.array(c)


def _nested_sampling(lattice, n, v_ex, species_numbers, *, rank, n_walkers, n_steps, n_iter):
    """The nested-sampling loop from the stream of `rank`
    (ref_nested_sampling); returns (the n_iter energy limits, walk trials,
    seconds of the iteration loop)."""
    v = np.asfortranarray(v_ex, dtype=np.float64)
    S, _, R = v.shape
    sn = np.ascontiguousarray(species_numbers, dtype=np.int64)
    out = np.zeros(n_iter)
    tr, secs = C.c_int64(0), C.c_double(0.0)
    rc = lib().ref_nested_sampling(lattice, *n, S, R, _p(v, C.c_double), int(rank),
                                   _p(sn, C.c_int64), int(n_walkers), int(n_steps), int(n_iter),
                                   _p(out, C.c_double), C.byref(tr), C.byref(secs))
    if rc:
        raise RuntimeError(f"ref_nested_sampling returned {rc}")
    return out, int(tr.value), float(secs.value)


@_isolated
def nested_sampling(lattice, n, v_ex, species_numbers, *, rank, n_walkers, n_steps, n_iter):
    """The n_iter energy limits of the reference's nested-sampling loop."""
    return _nested_sampling(lattice, n, v_ex, species_numbers, rank=rank, n_walkers=n_walkers,
                            n_steps=n_steps, n_iter=n_iter)[0]


nested_sampling_timed = _isolated(_nested_sampling)


def metropolis_sweeps_replicas(lattice, n, v_ex, cfgs, beta, n_sweeps):
    """One reference chain per configuration, each in its own forked process
    (rank k for cfgs[k]), all running at once; returns their results as
    metropolis_sweeps does, in order."""
    ctx = mp.get_context("fork")
    procs = []
    for k, cfg in enumerate(cfgs):
        rd, wr = ctx.Pipe(duplex=False)

        def child(wr=wr, k=k, cfg=cfg):
            try:
                wr.send(("ok", _metropolis_sweeps(lattice, n, v_ex, cfg, beta, n_sweeps, rank=k)))
            except BaseException as e:  # report, never hang the parent
                wr.send(("err", repr(e)))
            wr.close()

        p = ctx.Process(target=child)
        p.start()
        wr.close()
        procs.append((p, rd))
    out = []
    for p, rd in procs:
        try:
            status, val = rd.recv()
        except EOFError:
            status, val = "err", f"exit code {p.exitcode}"
        p.join()
        if status != "ok":
            raise RuntimeError(f"reference replica failed: {val}")
        out.append(val)
    return out


def run_parallel(fn, arg_list):
    """fn(*args) for every args in arg_list, each in its own forked process,
    all at once; returns the results in order (fn must return picklable
    values)."""
    ctx = mp.get_context("fork")
    procs = []
    for args in arg_list:
        rd, wr = ctx.Pipe(duplex=False)

        def child(wr=wr, args=args):
            try:
                wr.send(("ok", fn(*args)))
            except BaseException as e:  # report, never hang the parent
                wr.send(("err", repr(e)))
            wr.close()

        p = ctx.Process(target=child)
        p.start()
        wr.close()
        procs.append((p, rd))
    out = []
    for p, rd in procs:
        try:
            status, val = rd.recv()
        except EOFError:
            status, val = "err", f"exit code {p.exitcode}"
        p.join()
        if status != "ok":
            raise RuntimeError(f"reference worker failed: {val}")
        out.append(val)
    return out
