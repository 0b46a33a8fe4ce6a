"""Request server over the HIP C-ABI (bwl_server_*): continuous batching of
sampling runs on one device.

    srv = Server(device_bytes=256 << 20, max_launches=8)
    rid = srv.submit_anneal(lattice, n, v_ex, config, T=..., n_sweeps=..., seed=...)
    rid = srv.submit_ns(lattice, n, v_ex, seed, n_walkers=..., n_steps=..., n_iter=..., ...)
    rid = srv.submit_wl(lattice, n, v_ex, bins=..., ..., wl_f=...)
    while srv.progress() ...: ...          # non-blocking; or srv.drain()
    srv.result(rid)                        # outputs + timing

Each request keeps its own random stream, so its result equals the same
run made alone through the plain entry points.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import AnnealParams, LatticeDesc, NsParams, RequestInfo, ServerParams, WlParams, check
from .lattice import LATTICE_NAMES

KINDS = {1: "anneal", 2: "ns", 3: "wl"}
STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}


def _p(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _desc(lattice, n, v):
    if isinstance(lattice, str):
        lattice = LATTICE_NAMES[lattice]
    S, _, R = v.shape
    return LatticeDesc(int(lattice), int(n[0]), int(n[1]), int(n[2]), S, R)


@dataclass
class Request:
    id: int
    kind: str
    outputs: dict
    keep: list = field(default_factory=list)   # host buffers the library writes into


class Server:
    def __init__(self, device_bytes: int = 256 << 20, max_launches: int = 8, batch_wait_us: int = 0):
        self._L = _lib.load()
        h = C.c_void_p()
        p = ServerParams(int(device_bytes), int(max_launches), int(batch_wait_us))
        check(self._L.bwl_server_create(C.byref(p), C.byref(h)))
        self._h = h
        self.requests: dict[int, Request] = {}

    def close(self):
        if getattr(self, "_h", None):
            self._L.bwl_server_destroy(self._h)
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

    # -- submission -----------------------------------------------------------
    def submit_anneal(self, lattice, n, v_ex, config, *, T, n_sweeps, delta_T=0.0, T_steps=1,
                      burn_in_sweeps=0, seed=20261202, want_config=False) -> int:
        v = np.asfortranarray(v_ex, dtype=np.float64)
        cfg = np.asfortranarray(config, dtype=np.int8)
        p = AnnealParams(float(T), float(delta_T), int(T_steps), int(n_sweeps), int(burn_in_sweeps),
                         0, int(seed) & (2**64 - 1))
        E, Cv, A = np.zeros(T_steps), np.zeros(T_steps), np.zeros(T_steps)
        out_cfg = np.zeros_like(cfg) if want_config else None
        rid = C.c_int64()
        check(self._L.bwl_server_submit_anneal(
            self._h, C.byref(_desc(lattice, n, v)), _p(v, C.c_double), _p(cfg, C.c_int8), C.byref(p),
            _p(E, C.c_double), _p(Cv, C.c_double), _p(A, C.c_double),
            None if out_cfg is None else _p(out_cfg, C.c_int8), C.byref(rid)))
        outs = dict(energies_of_T=E, C_of_T=Cv, acceptance_of_T=A)
        if out_cfg is not None:
            outs["config"] = out_cfg
        self.requests[rid.value] = Request(rid.value, "anneal", outs, [v, cfg])
        return rid.value

    def submit_ns(self, lattice, n, v_ex, seed, *, n_walkers, n_steps, n_iter, species_numbers=None,
                  concentrations=None) -> int:
        v = np.asfortranarray(v_ex, dtype=np.float64)
        sn = None if species_numbers is None else np.ascontiguousarray(species_numbers, dtype=np.int64)
        cc = None if concentrations is None else np.ascontiguousarray(concentrations, dtype=np.float64)
        lim = np.zeros(int(n_iter))
        rid = C.c_int64()
        check(self._L.bwl_server_submit_ns(
            self._h, C.byref(_desc(lattice, n, v)), _p(v, C.c_double), int(seed) & 0xFFFFFFFF,
            C.byref(NsParams(int(n_walkers), int(n_steps), int(n_iter), 0)),
            None if sn is None else _p(sn, C.c_int64), None if cc is None else _p(cc, C.c_double),
            _p(lim, C.c_double), C.byref(rid)))
        self.requests[rid.value] = Request(rid.value, "ns", dict(e_limits=lim), [v])
        return rid.value

    def submit_wl(self, lattice, n, v_ex, *, bins, num_windows, walkers_per_window, bin_overlap,
                  energy_min, energy_max, mc_sweeps, wl_f, nbr_swap=False, rank_offset=0,
                  species_numbers=None, concentrations=None) -> int:
        v = np.asfortranarray(v_ex, dtype=np.float64)
        sn = None if species_numbers is None else np.ascontiguousarray(species_numbers, dtype=np.int64)
        cc = None if concentrations is None else np.ascontiguousarray(concentrations, dtype=np.float64)
        p = WlParams(int(bins), int(num_windows), int(walkers_per_window), float(bin_overlap),
                     float(energy_min), float(energy_max), int(bool(nbr_swap)), int(mc_sweeps),
                     int(rank_offset))
        lng = np.zeros((int(num_windows), int(bins)))
        hist = np.zeros((int(num_windows), int(bins)))   # stride <= bins
        rid = C.c_int64()
        check(self._L.bwl_server_submit_wl(
            self._h, C.byref(_desc(lattice, n, v)), _p(v, C.c_double), C.byref(p),
            None if sn is None else _p(sn, C.c_int64), None if cc is None else _p(cc, C.c_double),
            float(wl_f), _p(lng, C.c_double), _p(hist, C.c_double), C.byref(rid)))
        self.requests[rid.value] = Request(rid.value, "wl", dict(lng=lng, hist=hist), [v])
        return rid.value

    # -- progress -------------------------------------------------------------
    def progress(self):
        """Non-blocking; returns (running, queued)."""
        nr, nq = C.c_int32(), C.c_int32()
        check(self._L.bwl_server_progress(self._h, C.byref(nr), C.byref(nq)))
        return int(nr.value), int(nq.value)

    def poll(self, max_ids: int = 4096):
        ids = np.zeros(max_ids, dtype=np.int64)
        n = C.c_int32()
        check(self._L.bwl_server_poll(self._h, _p(ids, C.c_int64), int(max_ids), C.byref(n)))
        return [int(x) for x in ids[: n.value]]

    def drain(self):
        check(self._L.bwl_server_drain(self._h))

    def info(self, rid: int) -> dict:
        i = RequestInfo()
        check(self._L.bwl_server_info(self._h, int(rid), C.byref(i)))
        return dict(kind=KINDS.get(i.kind, "?"), state=STATES.get(i.state, "?"), error=i.error,
                    trials=int(i.trials), accepted=int(i.accepted), t_submit_ms=i.t_submit_ms,
                    t_admit_ms=i.t_admit_ms, t_done_ms=i.t_done_ms, kernel_ms=i.kernel_ms,
                    latency_ms=i.t_done_ms - i.t_submit_ms)

    def result(self, rid: int) -> dict:
        r = self.requests[rid]
        return dict(self.info(rid), **r.outputs)

    def memory(self):
        u, t = C.c_int64(), C.c_int64()
        check(self._L.bwl_server_memory(self._h, C.byref(u), C.byref(t)))
        return int(u.value), int(t.value)


__all__ = ["Server"]
