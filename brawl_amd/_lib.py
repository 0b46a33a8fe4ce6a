"""ctypes binding of libbrawl_hip.so (the C-ABI declared in include/brawl_hip.h).

The shared library is built in-tree (brawl_amd/libbrawl_hip.so) by
__graft_entry__.build() / `make -C brawl_amd/csrc`.  There is no fallback:
if the library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbrawl_hip.so")

BWL_OK = 0
BWL_ERR_ARG = -1
BWL_ERR_UNSUPPORTED = -2
BWL_ERR_DEVICE = -3
BWL_ERR_NOMEM = -4
BWL_ERR_STATE = -5
FCC, BCC = 1, 2
# parallel-sampler layouts (bwl_set_layout)
LAYOUTS = {"auto": 0, "compact": 1, "natural": 2, "fixed": 3}

# Every symbol include/brawl_hip.h declares.
EXPORTS = (
    "bwl_last_error", "bwl_device_count", "bwl_create", "bwl_destroy", "bwl_n_atoms",
    "bwl_set_config", "bwl_get_config", "bwl_pair_histogram", "bwl_total_energy",
    "bwl_site_energy", "bwl_pair_energy", "bwl_genrand_seed", "bwl_genrand_get_state",
    "bwl_genrand_set_state", "bwl_initial_setup", "bwl_mc_steps", "bwl_metropolis_anneal",
    "bwl_metropolis_anneal_batch", "bwl_metropolis_anneal_rho", "bwl_lattice_shells", "bwl_radial_densities",
    "bwl_nested_sampling", "bwl_nested_sampling_batch", "bwl_wl_create", "bwl_wl_destroy",
    "bwl_wl_n_walkers", "bwl_wl_windows", "bwl_wl_hist_stride", "bwl_wl_initial_setup",
    "bwl_wl_set_walker", "bwl_wl_get_walker", "bwl_wl_genrand_get_state",
    "bwl_wl_genrand_set_state", "bwl_wl_get_lng", "bwl_wl_set_lng", "bwl_wl_enter_windows",
    "bwl_wl_sweeps", "bwl_wl_replica_exchange", "bwl_server_create", "bwl_server_destroy",
    "bwl_server_submit_anneal",
    "bwl_server_submit_ns", "bwl_server_submit_wl", "bwl_server_progress", "bwl_server_poll",
    "bwl_server_info", "bwl_server_drain", "bwl_server_memory", "bwl_lds_bandwidth",
    "bwl_set_layout", "bwl_layouts", "bwl_layout_info", "bwl_set_threads", "bwl_mc_steps_batch",
    "bwl_random_site", "bwl_random_nbr", "bwl_select_sites", "bwl_metropolis_anneal_noalloc",
)


class LatticeDesc(C.Structure):
    _fields_ = [("lattice", C.c_int32), ("n1", C.c_int32), ("n2", C.c_int32),
                ("n3", C.c_int32), ("n_species", C.c_int32), ("n_shells", C.c_int32)]


class AnnealParams(C.Structure):
    _fields_ = [("T", C.c_double), ("delta_T", C.c_double), ("T_steps", C.c_int32),
                ("n_sweeps", C.c_int32), ("burn_in_sweeps", C.c_int32),
                ("asro_every", C.c_int32), ("seed", C.c_uint64)]


class RunStats(C.Structure):
    _fields_ = [("trials", C.c_int64), ("accepted", C.c_int64), ("kernel_ms", C.c_double),
                ("e_drift_abs", C.c_double), ("e_drift_rel", C.c_double)]


class NsParams(C.Structure):
    _fields_ = [("n_walkers", C.c_int32), ("n_steps", C.c_int32), ("n_iter", C.c_int32),
                ("reserved", C.c_int32)]


class WlParams(C.Structure):
    _fields_ = [("bins", C.c_int32), ("num_windows", C.c_int32),
                ("walkers_per_window", C.c_int32), ("bin_overlap", C.c_float),
                ("energy_min", C.c_float), ("energy_max", C.c_float), ("nbr_swap", C.c_int32),
                ("mc_sweeps", C.c_int32), ("rank_offset", C.c_uint32)]


class ServerParams(C.Structure):
    _fields_ = [("device_bytes", C.c_int64), ("max_launches", C.c_int32),
                ("batch_wait_us", C.c_int32)]


class RequestInfo(C.Structure):
    _fields_ = [("kind", C.c_int32), ("state", C.c_int32), ("error", C.c_int32),
                ("reserved", C.c_int32), ("trials", C.c_int64), ("accepted", C.c_int64),
                ("t_submit_ms", C.c_double), ("t_admit_ms", C.c_double), ("t_done_ms", C.c_double),
                ("kernel_ms", C.c_double)]


class BrawlError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "or `make -C brawl_amd/csrc`")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp, i8p, i32p, i64p, u32p, dp = (C.c_void_p, P(C.c_int8), P(C.c_int32), P(C.c_int64),
                                     P(C.c_uint32), P(C.c_double))
    sig = {
        "bwl_last_error": (C.c_char_p, []),
        "bwl_device_count": (C.c_int, [i32p]),
        "bwl_create": (C.c_int, [P(LatticeDesc), dp, P(vp)]),
        "bwl_destroy": (C.c_int, [vp]),
        "bwl_n_atoms": (C.c_int, [vp, i32p]),
        "bwl_set_config": (C.c_int, [vp, i8p]),
        "bwl_get_config": (C.c_int, [vp, i8p]),
        "bwl_pair_histogram": (C.c_int, [vp, i64p]),
        "bwl_total_energy": (C.c_int, [vp, dp]),
        "bwl_site_energy": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, dp]),
        "bwl_pair_energy": (C.c_int, [vp, i32p, i32p, dp]),
        "bwl_genrand_seed": (C.c_int, [vp, C.c_uint32]),
        "bwl_genrand_get_state": (C.c_int, [vp, u32p, i32p]),
        "bwl_genrand_set_state": (C.c_int, [vp, u32p, C.c_int32]),
        "bwl_initial_setup": (C.c_int, [vp, i64p, dp]),
        "bwl_random_site": (C.c_int, [vp, C.c_int32, i32p]),
        "bwl_random_nbr": (C.c_int, [vp, i32p, i32p]),
        "bwl_select_sites": (C.c_int, [vp, C.c_int32, C.c_int32, i32p]),
        "bwl_mc_steps": (C.c_int, [vp, C.c_double, C.c_int32, C.c_int64, C.c_int64, dp, i64p]),
        "bwl_metropolis_anneal": (C.c_int, [vp, P(AnnealParams), dp, dp, dp, P(RunStats)]),
        "bwl_metropolis_anneal_noalloc": (C.c_int, [vp, P(AnnealParams), dp, dp, dp, P(RunStats)]),
        "bwl_metropolis_anneal_batch": (C.c_int, [C.c_int32, P(vp), P(AnnealParams), P(dp), P(dp),
                                                  P(dp), P(RunStats), dp]),
        "bwl_metropolis_anneal_rho": (C.c_int, [vp, P(AnnealParams), dp, dp, dp, dp, P(RunStats)]),
        "bwl_lattice_shells": (C.c_int, [vp, C.c_int32, dp]),
        "bwl_radial_densities": (C.c_int, [vp, C.c_int32, dp]),
        "bwl_nested_sampling": (C.c_int, [vp, P(NsParams), i64p, dp, dp, P(RunStats)]),
        "bwl_nested_sampling_batch": (C.c_int, [C.c_int32, P(vp), P(NsParams), P(i64p), P(dp),
                                                P(dp), P(RunStats), dp]),
        "bwl_wl_create": (C.c_int, [P(LatticeDesc), dp, P(WlParams), P(vp)]),
        "bwl_wl_destroy": (C.c_int, [vp]),
        "bwl_wl_n_walkers": (C.c_int, [vp, i32p]),
        "bwl_wl_windows": (C.c_int, [vp, i32p, dp]),
        "bwl_wl_hist_stride": (C.c_int, [vp, i32p]),
        "bwl_wl_initial_setup": (C.c_int, [vp, i64p, dp]),
        "bwl_wl_set_walker": (C.c_int, [vp, C.c_int32, i8p]),
        "bwl_wl_get_walker": (C.c_int, [vp, C.c_int32, i8p]),
        "bwl_wl_genrand_get_state": (C.c_int, [vp, C.c_int32, u32p, i32p]),
        "bwl_wl_genrand_set_state": (C.c_int, [vp, C.c_int32, u32p, C.c_int32]),
        "bwl_wl_get_lng": (C.c_int, [vp, dp, dp]),
        "bwl_wl_set_lng": (C.c_int, [vp, dp, dp]),
        "bwl_wl_enter_windows": (C.c_int, [vp, C.c_int64, i64p, i32p]),
        "bwl_wl_sweeps": (C.c_int, [vp, C.c_double, C.c_int32, i64p, dp]),
        "bwl_wl_replica_exchange": (C.c_int, [vp, i32p, i32p, dp]),
        "bwl_server_create": (C.c_int, [P(ServerParams), P(vp)]),
        "bwl_server_destroy": (C.c_int, [vp]),
        "bwl_server_submit_anneal": (C.c_int, [vp, P(LatticeDesc), dp, i8p, P(AnnealParams), dp, dp,
                                               dp, i8p, i64p]),
        "bwl_server_submit_ns": (C.c_int, [vp, P(LatticeDesc), dp, C.c_uint32, P(NsParams), i64p, dp,
                                           dp, i64p]),
        "bwl_server_submit_wl": (C.c_int, [vp, P(LatticeDesc), dp, P(WlParams), i64p, dp, C.c_double,
                                           dp, dp, i64p]),
        "bwl_server_progress": (C.c_int, [vp, i32p, i32p]),
        "bwl_server_poll": (C.c_int, [vp, i64p, C.c_int32, i32p]),
        "bwl_server_info": (C.c_int, [vp, C.c_int64, P(RequestInfo)]),
        "bwl_server_drain": (C.c_int, [vp]),
        "bwl_server_memory": (C.c_int, [vp, i64p, i64p]),
        "bwl_lds_bandwidth": (C.c_int, [C.c_int32, dp, i32p]),
        "bwl_set_layout": (C.c_int, [vp, C.c_int32]),
        "bwl_set_threads": (C.c_int, [vp, C.c_int32]),
        "bwl_layouts": (C.c_int, [vp, i32p]),
        "bwl_layout_info": (C.c_int, [vp, C.c_int32, i32p, i64p]),
        "bwl_mc_steps_batch": (C.c_int, [C.c_int32, P(vp), dp, C.c_int32, C.c_int64, C.c_int64,
                                         P(dp), i64p, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != BWL_OK:
        msg = load().bwl_last_error()
        raise BrawlError(rc, msg.decode() if msg else "")


def device_count() -> int:
    n = C.c_int32(0)
    load().bwl_device_count(C.byref(n))
    return int(n.value)


def lds_bandwidth(iters: int = 20000):
    """(aggregate LDS read bytes/s over all CUs, number of CUs) measured by
    the library's probe kernel."""
    bps, cus = C.c_double(0.0), C.c_int32(0)
    check(load().bwl_lds_bandwidth(int(iters), C.byref(bps), C.byref(cus)))
    return float(bps.value), int(cus.value)
