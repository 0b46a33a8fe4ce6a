"""The request server: a mixed stream of anneals, nested-sampling runs and
short Wang-Landau runs, admitted under a small device-memory budget with few
launch slots (so requests queue, skip ahead and overlap), gives every
request exactly the result of the same run made alone through the plain
entry points."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    import brawl_amd
    from brawl_amd import _lib
    _lib.load()
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to the HIP library")
    return brawl_amd


def _mix(hip, oracle, n_anneal=14, n_ns=5, n_wl=4, seed=99):
    from brawl_amd import workloads
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n_anneal):
        lat = int(rng.integers(1, 3))
        n = int(rng.integers(4, 13))
        S = int(rng.integers(2, 6))
        v = workloads.random_v(S, [25.0 / k for k in range(1, 4)], 1000 + i)
        o = oracle.OracleLattice(lat, (n, n, n), v)
        counts = workloads.dirichlet_counts(o.n_atoms, S, rng)
        cfg = o.initial_setup(oracle.Generator(110179 + i), species_numbers=counts)
        reqs.append(dict(kind="anneal", lat=lat, n=(n, n, n), v=v, cfg=cfg,
                         T=float(rng.uniform(200, 2000)), n_sweeps=int(rng.integers(5, 40)),
                         seed=int(rng.integers(1, 2**62))))
    for i in range(n_ns):
        S = 5
        v = workloads.random_v(S, [30.0 / k for k in range(1, 5)], 2000 + i)
        reqs.append(dict(kind="ns", lat=1, n=(3, 3, 3), v=v, seed=int(rng.integers(1, 2**31)),
                         n_walkers=int(rng.integers(16, 64)), n_steps=20, n_iter=40,
                         conc=np.full(S, 1.0 / S)))
    for i in range(n_wl):
        S = 4
        v = workloads.random_v(S, [30.0, 15.0], 3000 + i)
        reqs.append(dict(kind="wl", lat=2, n=(4, 4, 4), v=v, conc=np.full(S, 0.25), bins=30,
                         num_windows=int(rng.integers(1, 3)), walkers_per_window=int(rng.integers(2, 5)),
                         bin_overlap=0.1, energy_min=-15.0, energy_max=15.0, mc_sweeps=3,
                         wl_f=1.0, rank_offset=int(rng.integers(0, 50))))
    order = rng.permutation(len(reqs))
    return [reqs[k] for k in order]


def _solo(hip, r):
    from brawl_amd import Lattice, nested_sampling, wang_landau
    if r["kind"] == "anneal":
        with Lattice(r["lat"], r["n"], r["v"]) as L:
            L.set_config(r["cfg"])
            res = L.metropolis_simulated_annealing(T=r["T"], n_sweeps=r["n_sweeps"], seed=r["seed"])
            return dict(energies_of_T=res.energies_of_T, C_of_T=res.C_of_T,
                        acceptance_of_T=res.acceptance_of_T, config=L.get_config(),
                        trials=res.trials)
    if r["kind"] == "ns":
        with Lattice(r["lat"], r["n"], r["v"]) as L:
            L.seed(r["seed"])
            lim, st = nested_sampling.nested_sampling(L, n_walkers=r["n_walkers"], n_steps=r["n_steps"],
                                                      n_iter=r["n_iter"], concentrations=r.get("conc"),
                                                      species_numbers=r.get("sn"))
            return dict(e_limits=lim, trials=st.trials)
    wl = wang_landau.WangLandau(r["lat"], r["n"], r["v"], bins=r["bins"], num_windows=r["num_windows"],
                                walkers_per_window=r["walkers_per_window"], bin_overlap=r["bin_overlap"],
                                energy_min=r["energy_min"], energy_max=r["energy_max"],
                                mc_sweeps=r["mc_sweeps"], rank_offset=r["rank_offset"])
    wl.initial_setup(concentrations=r.get("conc"), species_numbers=r.get("sn"))
    wl.sweeps(r["wl_f"], average=True)
    lng, hist = wl.lng_hist()
    out = dict(lng=lng[:: r["walkers_per_window"]].copy(), hist=hist[:: r["walkers_per_window"]].copy(),
               windows=wl.windows.copy())
    wl.close()
    return out


def _submit(srv, r):
    if r["kind"] == "anneal":
        return srv.submit_anneal(r["lat"], r["n"], r["v"], r["cfg"], T=r["T"], n_sweeps=r["n_sweeps"],
                                 seed=r["seed"], want_config=True)
    if r["kind"] == "ns":
        return srv.submit_ns(r["lat"], r["n"], r["v"], r["seed"], n_walkers=r["n_walkers"],
                             n_steps=r["n_steps"], n_iter=r["n_iter"], concentrations=r.get("conc"),
                             species_numbers=r.get("sn"))
    return srv.submit_wl(r["lat"], r["n"], r["v"], bins=r["bins"], num_windows=r["num_windows"],
                         walkers_per_window=r["walkers_per_window"], bin_overlap=r["bin_overlap"],
                         energy_min=r["energy_min"], energy_max=r["energy_max"],
                         mc_sweeps=r["mc_sweeps"], wl_f=r["wl_f"], rank_offset=r["rank_offset"],
                         concentrations=r.get("conc"), species_numbers=r.get("sn"))


def _check_same(r, got, solo):
    if r["kind"] == "anneal":
        for k in ("energies_of_T", "C_of_T", "acceptance_of_T", "config"):
            assert np.array_equal(got[k], solo[k]), k
        assert got["trials"] == solo["trials"]
    elif r["kind"] == "ns":
        assert np.array_equal(got["e_limits"], solo["e_limits"])
        assert got["trials"] == solo["trials"]
    else:
        assert np.array_equal(got["lng"], solo["lng"])
        for w in range(r["num_windows"]):
            s, e = solo["windows"][w]
            assert np.array_equal(got["hist"][w, : e - s + 1], solo["hist"][w, : e - s + 1])


def test_server_results_equal_solo_runs(hip, oracle):
    from brawl_amd.serve import Server
    reqs = _mix(hip, oracle)
    _say("requests", len(reqs))
    solo = [_solo(hip, r) for r in reqs]
    # a budget that holds only a few requests at a time, three launch slots
    with Server(device_bytes=3 << 20, max_launches=3) as srv:
        ids = []
        for k, r in enumerate(reqs):
            ids.append(_submit(srv, r))
            if k % 3 == 2:
                srv.progress()          # requests arrive while others run
        srv.drain()
        done = srv.poll()
        assert sorted(done) == sorted(ids)
        assert srv.memory()[0] == 0
        for r, rid, s in zip(reqs, ids, solo):
            got = srv.result(rid)
            assert got["state"] == "done", got
            assert got["t_submit_ms"] <= got["t_admit_ms"] <= got["t_done_ms"]
            _check_same(r, got, s)


def test_server_rejects_oversized_requests(hip):
    """A nested-sampling run whose 640 walkers need more than the whole
    device budget is refused at submission (BWL_ERR_NOMEM)."""
    from brawl_amd import BrawlError, workloads
    from brawl_amd.serve import Server
    v = workloads.random_v(5, [30.0, 15.0, 10.0, 7.5], 1)
    with Server(device_bytes=1 << 20, max_launches=1) as srv:
        with pytest.raises(BrawlError) as e:
            srv.submit_ns(1, (6, 6, 6), v, 7, n_walkers=640, n_steps=100, n_iter=10,
                          concentrations=np.full(5, 0.2))
        assert e.value.code == -4
        assert srv.memory() == (0, 1 << 20)


def _from_served(q):
    """A request of the served-stream workload as the helpers above take it."""
    r = dict(kind=q.kind, lat=q.lattice, n=q.n, v=q.v_ex, seed=q.seed)
    if q.kind == "anneal":
        r.update(cfg=q.config, T=q.params["T"], n_sweeps=q.params["n_sweeps"])
    elif q.kind == "ns":
        r.update(sn=q.counts, **q.params)
    else:
        r.update(sn=q.counts, **q.params)
    return r


def test_served_mix_at_stated_sizes(hip, oracle):
    """Requests of the served mixed stream at the workload's own sizes
    (anneals of fcc 8^3-28^3 / bcc 8^3-32^3 with 2-5 species and 100-500
    sweeps; quinary fcc 6^3 nested sampling with 128-640 walkers, walk
    length 100, 200 iterations; Wang-Landau runs of bcc 6^3-10^3 with 4-8
    walkers over 20 sweeps), admitted as they arrive under a budget that
    holds a few of them at a time: each result equals the same run made
    alone, and every nested-sampling run's 200 energy limits equal the
    oracle's run from the same MT19937 seed."""
    from brawl_amd import workloads
    from brawl_amd.serve import Server
    stream = workloads.config5(n_requests=36)
    kinds = [q.kind for q in stream]
    _say("served mix", {k: kinds.count(k) for k in set(kinds)})
    assert {"anneal", "ns", "wl"} <= set(kinds)
    reqs = [_from_served(q) for q in stream]
    solo = [_solo(hip, r) for r in reqs]
    with Server(device_bytes=48 << 20, max_launches=4, batch_wait_us=2000) as srv:
        ids = []
        for k, r in enumerate(reqs):
            ids.append(_submit(srv, r))
            srv.progress()
        srv.drain()
        assert sorted(srv.poll()) == sorted(ids)
        assert srv.memory()[0] == 0
        for r, rid, s in zip(reqs, ids, solo):
            got = srv.result(rid)
            assert got["state"] == "done", got
            _check_same(r, got, s)
    for r, rid, s in zip(reqs, ids, solo):
        if r["kind"] != "ns":
            continue
        ref = oracle.OracleLattice(r["lat"], r["n"], r["v"])
        _say("ns vs oracle", r["n_walkers"], "walkers")
        lim = ref.nested_sampling(oracle.Generator(r["seed"]), n_walkers=r["n_walkers"],
                                  n_steps=r["n_steps"], n_iter=r["n_iter"], species_numbers=r["sn"])
        assert np.array_equal(s["e_limits"], lim)
