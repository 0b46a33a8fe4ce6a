"""The C-ABI library loads, exports every symbol include/brawl_hip.h declares,
and rejects bad descriptors before touching a device.  CPU only: no compute
call is made here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "brawl_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bwl_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    from brawl_amd import _lib
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from brawl_amd import _lib
    L = _lib.load()
    for name in declared_symbols():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (bwl_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_library_is_gfx950():
    from brawl_amd import _lib
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def _create(desc_vals, v):
    from brawl_amd import _lib
    L = _lib.load()
    d = _lib.LatticeDesc(*desc_vals)
    h = C.c_void_p()
    rc = L.bwl_create(C.byref(d), v.ctypes.data_as(C.POINTER(C.c_double)), C.byref(h))
    if rc == 0:
        L.bwl_destroy(h)
    return rc, L.bwl_last_error().decode()


@pytest.mark.parametrize("desc,code", [
    ((3, 4, 4, 4, 2, 2), -2),    # simple cubic / unknown lattice: unsupported
    ((1, 0, 4, 4, 2, 2), -1),    # zero cells
    ((1, 4, 4, 4, 9, 2), -2),    # 9 species
    ((1, 4, 4, 4, 2, 7), -2),    # fcc has 6 shells
    ((2, 4, 4, 4, 2, 11), -2),   # bcc has 10 shells
])
def test_invalid_descriptors_rejected_before_device(desc, code):
    v = np.zeros(8 * 8 * 11)
    rc, msg = _create(desc, v)
    assert rc == code and msg


def test_null_arguments():
    from brawl_amd import _lib
    L = _lib.load()
    assert L.bwl_create(None, None, None) == _lib.BWL_ERR_ARG
    assert L.bwl_destroy(None) == 0
    assert L.bwl_total_energy(None, None) == _lib.BWL_ERR_ARG
    assert L.bwl_set_layout(None, 0) == _lib.BWL_ERR_ARG
    assert L.bwl_layouts(None, None) == _lib.BWL_ERR_ARG
    assert L.bwl_layout_info(None, 0, None, None) == _lib.BWL_ERR_ARG
    assert L.bwl_wl_replica_exchange(None, None, None, None) == _lib.BWL_ERR_ARG
    assert L.bwl_lds_bandwidth(0, None, None) == _lib.BWL_ERR_ARG
    assert L.bwl_mc_steps_batch(0, None, None, 0, 1, 0, None, None, None) == _lib.BWL_ERR_ARG


def test_no_cpu_fallback_without_gpu():
    """On a machine without a GPU the library reports a device error instead
    of computing anything on the host."""
    from brawl_amd import _lib
    if _lib.device_count() > 0:
        pytest.skip("a GPU is present")
    rc, msg = _create((1, 4, 4, 4, 2, 2), np.zeros(8))
    assert rc == _lib.BWL_ERR_DEVICE


def test_anneal_batch_seeds_are_distinct_by_default(monkeypatch):
    """Chains without a seed key get distinct derived seeds; equal explicit
    seeds warn (their streams would be identical).  The library call is
    replaced by a recorder: no device is touched."""
    import warnings

    from brawl_amd import _lib, lattice

    seen = {}

    class Fake:
        def bwl_metropolis_anneal_batch(self, n, H, P, *rest):
            seen["seeds"] = [P[i].seed for i in range(n)]
            return 0

    class Lat:
        _h = None

    monkeypatch.setattr(_lib, "load", lambda: Fake())
    lats = [Lat(), Lat(), Lat()]
    lattice.anneal_batch(lats, [dict(T=500.0, n_sweeps=1)] * 3)
    assert len(set(seen["seeds"])) == 3
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        lattice.anneal_batch(lats, [dict(T=500.0, n_sweeps=1, seed=7)] * 3)
    assert any("share a seed" in str(x.message) for x in w)
