"""Walker kernels vs the oracle, through the C-ABI: radial densities
(integer pair counts, bit-exact), nested-sampling runs (energy limits
bit-exact for the same MT19937 stream, alone and batched), Wang-Landau
`sweeps` (per-walker ln g, histogram, configuration and stream bit-exact;
window average equal to the walker-order mean) and enter_energy_window
(bit-exact, re-set-up path included)."""
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


ASRO_CASES = [
    # lattice, n, S, wc
    (1, (4, 4, 4), 3, 3),     # n < 5: the window wraps, offsets counted twice
    (1, (6, 6, 6), 5, 4),
    (2, (8, 8, 8), 4, 5),
    (2, (3, 5, 4), 2, 6),
    (1, (2, 2, 2), 2, 2),
    (2, (16, 16, 16), 4, 3),
    (1, (7, 5, 9), 8, 16),
]


@pytest.mark.parametrize("lat,n,S,wc", ASRO_CASES)
def test_radial_densities_bit_exact(hip, oracle, lat, n, S, wc):
    from brawl_amd import workloads
    _say("asro case", lat, n, S, wc)
    v = workloads.random_v(S, [10.0], 5)
    ref = oracle.OracleLattice(lat, n, v)
    N = ref.n_atoms
    cfg = ref.initial_setup(oracle.Generator(110179 + N),
                            species_numbers=workloads.dirichlet_counts(N, S, np.random.default_rng(N)))
    shells_ref = ref.lattice_shells(cfg, wc)
    rho_ref = ref.radial_densities(cfg, wc, shells_ref)
    with hip.Lattice(lat, n, v) as L:
        L.set_config(cfg)
        assert np.array_equal(L.lattice_shells(wc), shells_ref)
        rho = L.radial_densities(wc)
    ok = np.isfinite(rho_ref)
    assert np.array_equal(np.isfinite(rho), ok)
    assert np.array_equal(rho[ok], rho_ref[ok])
    # on-site shell: rho(a, a, 1) = 1 for present species
    present = [a for a in range(S) if np.isfinite(rho[a, a, 0])]
    assert all(rho[a, a, 0] == 1.0 for a in present)


NS_CASES = [
    # lattice, n, S, R, walkers, steps, iters
    (1, (3, 3, 3), 3, 4, 16, 20, 60),
    (2, (3, 3, 3), 2, 2, 9, 10, 40),
    (1, (2, 2, 2), 5, 6, 32, 30, 50),
]


@pytest.mark.parametrize("lat,n,S,R,nw,ns,ni", NS_CASES)
def test_nested_sampling_bit_exact(hip, oracle, lat, n, S, R, nw, ns, ni):
    from brawl_amd import nested_sampling, workloads
    _say("ns case", lat, n, S, R, nw, ns, ni)
    v = workloads.random_v(S, [30.0 / k for k in range(1, R + 1)], 3 + S)
    ref = oracle.OracleLattice(lat, n, v)
    conc = np.full(S, 1.0 / S)
    gen = oracle.Generator(110179)
    lim_ref = ref.nested_sampling(gen, n_walkers=nw, n_steps=ns, n_iter=ni, concentrations=conc)
    with hip.Lattice(lat, n, v) as L:
        L.seed(110179)
        lim, st = nested_sampling.nested_sampling(L, n_walkers=nw, n_steps=ns, n_iter=ni,
                                                  concentrations=conc)
        mt, mti = L.genrand_state()
    assert np.array_equal(lim, lim_ref)
    mt_ref, mti_ref = gen.state_arrays()
    assert mti == mti_ref and np.array_equal(mt, mt_ref)
    assert st.trials >= ns * ni
    assert np.all(np.diff(lim) <= 0)          # limits never rise


def test_nested_sampling_batch_matches_single_runs(hip, oracle):
    from brawl_amd import nested_sampling, workloads
    specs = [(1, (3, 3, 3), 3, 2, 12, 15, 30, 11), (2, (4, 4, 4), 4, 3, 20, 10, 25, 22),
             (1, (2, 2, 2), 5, 4, 8, 25, 40, 33), (2, (2, 3, 2), 2, 2, 6, 5, 20, 44)]
    lats, params, solo = [], [], []
    for lat, n, S, R, nw, ns, ni, seed in specs:
        v = workloads.random_v(S, [30.0 / k for k in range(1, R + 1)], seed)
        p = dict(n_walkers=nw, n_steps=ns, n_iter=ni, concentrations=np.full(S, 1.0 / S))
        with hip.Lattice(lat, n, v) as L:
            L.seed(seed)
            solo.append(nested_sampling.nested_sampling(L, **p)[0])
        L = hip.Lattice(lat, n, v)
        L.seed(seed)
        lats.append(L)
        params.append(p)
    res, mk = nested_sampling.nested_sampling_batch(lats, params)
    for (lim, _), s in zip(res, solo):
        assert np.array_equal(lim, s)
    for L in lats:
        L.close()


def _wl_case(hip, oracle, nbr_swap, seed):
    from brawl_amd import wang_landau, workloads
    lat, n, S, R = 2, (4, 4, 4), 3, 2
    v = workloads.random_v(S, [30.0, 15.0], seed)
    ref = oracle.OracleLattice(lat, n, v)
    # an energy range spanning random configurations: +-4 meV/atom around 0
    wl = wang_landau.WangLandau(lat, n, v, bins=40, num_windows=3, walkers_per_window=2,
                                bin_overlap=0.1, energy_min=-4.0, energy_max=4.0, mc_sweeps=2,
                                nbr_swap=nbr_swap)
    return wl, ref, S


@pytest.mark.parametrize("nbr_swap", [False, True])
def test_wang_landau_sweeps_bit_exact(hip, oracle, nbr_swap):
    wl, ref, S = _wl_case(hip, oracle, nbr_swap, 7)
    _say("wl sweeps, nbr_swap", nbr_swap)
    windows_ref, edges_ref = ref.wl_setup(40, 3, 0.1, -4.0, 4.0)
    assert np.array_equal(wl.windows, windows_ref)
    assert np.array_equal(wl.bin_edges, edges_ref)
    conc = np.full(S, 1.0 / S)
    wl.initial_setup(concentrations=conc)
    gens, cfgs = [], []
    for k in range(wl.n_walkers):
        g = oracle.Generator(rank=k)
        cfgs.append(ref.initial_setup(g, concentrations=conc))
        gens.append(g)
        assert np.array_equal(wl.get_walker(k), cfgs[k])
    rng = np.random.default_rng(3)
    lng0 = rng.normal(0.0, 0.5, (wl.n_walkers, wl.bins))
    wl.set_lng_hist(lng0, np.zeros((wl.n_walkers, wl.hist_stride)))
    lng_ref = lng0.copy()
    hist_ref = np.zeros((wl.n_walkers, wl.hist_stride))
    for call in range(2):
        acc, _ = wl.sweeps(0.5 / (call + 1), average=False)
        lng, hist = wl.lng_hist()
        for k in range(wl.n_walkers):
            s, e = wl.windows[wl.window_of(k)]
            h = np.zeros(e - s + 1)
            h[:] = hist_ref[k, :e - s + 1]
            lk = lng_ref[k].copy()
            a = ref.wl_sweeps(cfgs[k], gens[k], mc_sweeps=2, bin_edges=edges_ref, start_idx=int(s),
                              end_idx=int(e), wl_f=0.5 / (call + 1), lng=lk, hist=h,
                              nbr_swap=nbr_swap)
            lng_ref[k] = lk
            hist_ref[k, :e - s + 1] = h
            assert acc[k] == a
            assert np.array_equal(wl.get_walker(k), cfgs[k])
            mt, mti = wl.genrand_state(k)
            mt_ref, mti_ref = gens[k].state_arrays()
            assert mti == mti_ref and np.array_equal(mt, mt_ref)
        assert np.array_equal(lng, lng_ref)
        assert np.array_equal(hist, hist_ref)
    # the window average: walker-order mean, written to every walker
    wl.set_lng_hist(lng_ref, hist_ref)
    wl.sweeps(0.0, average=True)        # ln f = 0 leaves ln g as it is before the average
    lng, _ = wl.lng_hist()
    lng_noavg = np.zeros_like(lng)
    for k in range(wl.n_walkers):
        lng_noavg[k] = lng_ref[k]
    expect = wl.dos_average(lng_noavg)
    assert np.array_equal(lng, expect)
    wl.close()


def test_wang_landau_enter_window_bit_exact(hip, oracle):
    """Walkers walk into their windows exactly as the oracle's
    enter_energy_window; one unreachable window forces the reference's
    periodic re-set-up, compared after a fixed number of trials."""
    from brawl_amd import wang_landau, workloads
    lat, n, S, R = 1, (2, 2, 2), 2, 1
    v = workloads.random_v(S, [20.0], 1)
    ref = oracle.OracleLattice(lat, n, v)
    conc = np.full(S, 1.0 / S)
    # one window at an energy no configuration reaches (far above the range
    # of 32-atom binary configurations): every walker keeps re-setting up
    for emin, emax, max_trials in [(-3.0, 3.0, 10**7), (5000.0, 6000.0, 20000)]:
        _say("wl enter", emin, emax, max_trials)
        wl = wang_landau.WangLandau(lat, n, v, bins=20, num_windows=1, walkers_per_window=3,
                                    bin_overlap=0.1, energy_min=emin, energy_max=emax)
        wl.initial_setup(concentrations=conc)
        trials, entered = wl.enter_windows(max_trials)
        _, edges = ref.wl_setup(20, 1, 0.1, emin, emax)
        for k in range(wl.n_walkers):
            g = oracle.Generator(rank=k)
            cfg = ref.initial_setup(g, concentrations=conc)
            t = ref.wl_enter_window(cfg, g, edges[0], edges[-1], emin, emax, concentrations=conc,
                                    max_trials=max_trials)
            assert trials[k] == t
            assert np.array_equal(wl.get_walker(k), cfg)
            mt, mti = wl.genrand_state(k)
            mt_ref, mti_ref = g.state_arrays()
            assert mti == mti_ref and np.array_equal(mt, mt_ref)
            assert entered[k] == (t < max_trials)
        wl.close()


def _wl_rx_case(hip, oracle, bins, wpw, emin, emax, mc_sweeps, overlap=0.5):
    """bcc 4^3, 3 species, 3 windows: device walkers and oracle walkers from
    the same streams, both walked into their windows."""
    from brawl_amd import wang_landau, workloads
    lat, n, S = 2, (4, 4, 4), 3
    v = workloads.random_v(S, [30.0, 15.0], 7)
    ref = oracle.OracleLattice(lat, n, v)
    wl = wang_landau.WangLandau(lat, n, v, bins=bins, num_windows=3, walkers_per_window=wpw,
                                bin_overlap=overlap, energy_min=emin, energy_max=emax,
                                mc_sweeps=mc_sweeps)
    windows, edges = ref.wl_setup(bins, 3, overlap, emin, emax)
    assert np.array_equal(wl.windows, windows) and np.array_equal(wl.bin_edges, edges)
    conc = np.full(S, 1.0 / S)
    wl.initial_setup(concentrations=conc)
    _, entered = wl.enter_windows(10**7)
    assert entered.all()
    gens, cfgs = [], []
    for k in range(wl.n_walkers):
        g = oracle.Generator(rank=k)
        c = ref.initial_setup(g, concentrations=conc)
        s, e = windows[k // wpw]
        ref.wl_enter_window(c, g, edges[s - 1], edges[e], emin, emax, concentrations=conc)
        gens.append(g)
        cfgs.append(c)
    return wl, ref, windows, edges, gens, cfgs


def _wl_same_state(wl, gens, cfgs):
    for k in range(wl.n_walkers):
        assert np.array_equal(wl.get_walker(k), cfgs[k]), f"walker {k} configuration"
        mt, mti = wl.genrand_state(k)
        mt_ref, mti_ref = gens[k].state_arrays()
        assert mti == mti_ref and np.array_equal(mt, mt_ref), f"walker {k} stream"


def test_wang_landau_replica_exchange_bit_exact(hip, oracle):
    """Sweeps then replica_exchange, 12 times: matches, decisions, exchanged
    configurations and every walker's stream equal the oracle's."""
    wpw, bins = 4, 40
    wl, ref, windows, edges, gens, cfgs = _wl_rx_case(hip, oracle, bins, wpw, 8.0, 18.0, 1)
    _say("wl replica exchange: walkers", wl.n_walkers)
    _wl_same_state(wl, gens, cfgs)
    lng0 = np.random.default_rng(11).normal(0.0, 1.0, (wl.n_walkers, bins))
    wl.set_lng_hist(lng0, np.zeros((wl.n_walkers, wl.hist_stride)))
    pairs = exchanged = 0
    for it in range(12):
        wl.sweeps(0.0, average=False)          # ln f = 0: ln g stays lng0
        for k in range(wl.n_walkers):
            s, e = windows[k // wpw]
            ref.wl_sweeps(cfgs[k], gens[k], mc_sweeps=1, bin_edges=edges, start_idx=int(s),
                          end_idx=int(e), wl_f=0.0, lng=lng0[k].copy(), hist=np.zeros(e - s + 1))
        partner, ex, _ = wl.replica_exchange()
        n, p_ref, a_ref = oracle.wl_replica_exchange(ref, cfgs, gens, lng0, edges, windows, wpw)
        assert np.array_equal(partner, p_ref), f"round {it}"
        assert np.array_equal(ex, a_ref.astype(bool)), f"round {it}"
        assert int(ex.sum()) // 2 == n
        _wl_same_state(wl, gens, cfgs)
        pairs += int((partner >= 0).sum()) // 2
        exchanged += n
    _say("pairs", pairs, "exchanged", exchanged)
    assert pairs > exchanged > 0          # both outcomes occurred
    wl.close()


def test_wang_landau_pre_sampling_bit_exact(hip, oracle):
    """The pre-sampling loop (sweeps with window averages at ln f = 1, an
    exchange after every 10th call and then every 9th, until every window
    has been flat once) equals the same loop over the oracle's walkers:
    number of calls, ln g, histograms, configurations and streams."""
    wpw, bins = 2, 30
    wl, ref, windows, edges, gens, cfgs = _wl_rx_case(hip, oracle, bins, wpw, 9.0, 16.0, 5)
    calls, _, _, exch, pairs, flat = wl.pre_sampling(1.0, 0.8, max_sweeps_calls=2000)
    _say("wl pre-sampling: calls", calls, "pairs", pairs, "exchanged", exch)
    assert flat and pairs > 0
    W = wl.n_walkers
    lng = np.zeros((W, bins))
    hist = np.zeros((W, wl.hist_stride))
    i_sweeps = 0
    for call in range(calls):
        for k in range(W):
            s, e = windows[k // wpw]
            h = hist[k, :e - s + 1].copy()
            lk = lng[k].copy()
            ref.wl_sweeps(cfgs[k], gens[k], mc_sweeps=5, bin_edges=edges, start_idx=int(s),
                          end_idx=int(e), wl_f=1.0, lng=lk, hist=h)
            lng[k] = lk
            hist[k, :e - s + 1] = h
        lng = wl.dos_average(lng)
        hist = wl.dos_average(hist)
        i_sweeps += 1
        if i_sweeps % 10 == 0:
            i_sweeps = 1
            oracle.wl_replica_exchange(ref, cfgs, gens, lng, edges, windows, wpw)
    lng_d, hist_d = wl.lng_hist()
    assert np.array_equal(lng_d, lng)
    assert np.array_equal(hist_d, hist)
    _wl_same_state(wl, gens, cfgs)
    wl.close()
