"""bwl_metropolis_anneal_noalloc (the anneal whose scratch is a file-scope
device array, for callers that cannot allocate) against
bwl_metropolis_anneal on twin contexts: same configuration, seed and
stream position give bit-identical outputs, final configurations and
counts, on two calls in a row, the second with a smaller input; an input
whose scratch exceeds the array is refused before anything runs."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    from brawl_amd import _lib
    L = _lib.load()
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to the HIP library")
    return L


def _anneal(lib, L, params, noalloc):
    from brawl_amd import _lib
    T = params.T_steps
    e, c, a = np.zeros(T), np.zeros(T), np.zeros(T)
    st = _lib.RunStats()
    dp = C.POINTER(C.c_double)
    fn = lib.bwl_metropolis_anneal_noalloc if noalloc else lib.bwl_metropolis_anneal
    rc = fn(L._h, C.byref(params), e.ctypes.data_as(dp), c.ctypes.data_as(dp), a.ctypes.data_as(dp),
            C.byref(st))
    return rc, e, c, a, st


def _case(lib, lat, n, S, R, T_steps, n_sweeps, seed):
    from brawl_amd import Lattice, _lib, workloads
    v = workloads.random_v(S, [25.0 / k for k in range(1, R + 1)], seed)
    with Lattice(lat, n, v) as A, Lattice(lat, n, v) as B:
        N = A.n_atoms
        counts = workloads.dirichlet_counts(N, S, np.random.default_rng(seed))
        A.seed(110179)
        cfg = A.initial_setup(species_numbers=counts)
        B.set_config(cfg)
        p = _lib.AnnealParams(1800.0, -300.0, T_steps, n_sweeps, 2, 0, seed)
        _say("noalloc case", lat, n, S, R, T_steps, n_sweeps)
        rc_a, *out_a = _anneal(lib, A, p, False)
        rc_b, *out_b = _anneal(lib, B, p, True)
        assert rc_a == 0 and rc_b == 0, (rc_a, rc_b, _lib.load().bwl_last_error())
        for x, y in zip(out_a[:3], out_b[:3]):
            assert np.array_equal(x, y)
        assert out_a[3].trials == out_b[3].trials > 0
        assert out_a[3].accepted == out_b[3].accepted
        assert np.array_equal(A.get_config(), B.get_config())
        # a second call continues each context's stream the same way
        rc_a, *out_a = _anneal(lib, A, p, False)
        rc_b, *out_b = _anneal(lib, B, p, True)
        assert rc_a == 0 and rc_b == 0
        for x, y in zip(out_a[:3], out_b[:3]):
            assert np.array_equal(x, y)
        assert np.array_equal(A.get_config(), B.get_config())


def test_noalloc_anneal_equals_allocating_anneal(lib):
    # the larger input first, then a smaller one: the shared array keeps the
    # first call's contents, which must not leak into the second
    _case(lib, 2, (8, 8, 8), 4, 4, 6, 5, 11)
    _case(lib, 1, (4, 4, 4), 3, 2, 3, 4, 12)


def test_noalloc_anneal_refuses_oversized_scratch(lib):
    from brawl_amd import Lattice, _lib, workloads
    v = workloads.random_v(2, [20.0], 3)
    with Lattice(2, (4, 4, 4), v) as L:
        L.initial_setup(species_numbers=[64, 64])
        p = _lib.AnnealParams(1000.0, 0.0, 200000, 1, 0, 0, 5)   # 8 MB of per-T outputs
        rc, *_ = _anneal(lib, L, p, True)
        assert rc == _lib.BWL_ERR_ARG
        assert b"scratch" in _lib.load().bwl_last_error()
