"""Wang-Landau at full size and against an exact density of states.

Config 4 (bcc 16^3, 4 species, 2 shells; 16 windows x 50 bins with 10 %
overlap, 8 walkers per window = 128 walkers on one device): window entry,
then sweeps calls with replica exchange.  After each call every walker's
configuration keeps its composition and its recounted energy (the
reference's summation order) lies in its window; the window's ln g grew by
wl_f per trial and its histogram by one count every INT(0.02 N) trials
(wang-landau.F90:539-632); two walkers are followed bit for bit by the
oracle's `sweeps`.

Exact DOS: binary bcc 3^3 (54 sites) with 5 solutes and one shell has
energy E = E_0 + m J, where m is the number of solute-solute shell-1 bonds
(0..6); the number of configurations at each m is counted by enumerating
all C(54, 5) placements.  Whole Wang-Landau runs on the device (window
entry, pre-sampling, main loop to ln f < 1e-4) are bit-identical to the
same runs of the oracle's restatement of the reference's routines, and
their ln g(m) - ln g(0) lies within 2 standard deviations (of one run,
over independent runs) of the exact value.  The reference's algorithm
judges only flatness, so its error saturates instead of shrinking with
ln f (at 1e-4 and at 1e-6 alike, about 0.03 at m = 1 and 0.2 at m = 6
here): the spread of single runs, not the standard error of their mean,
is the scale its result is held to."""
import sys

import numpy as np
import pytest

from exact_dos import exact_dos as _exact_dos

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


def _bin_of(E, edges):
    """bin_index (wang-landau.F90:515-523): truncation toward zero, 1-based."""
    bins = len(edges) - 1
    return int((E - edges[0]) / (edges[-1] - edges[0]) * bins) + 1


def test_config4_full_size_entry_sweeps_and_exchange(hip, oracle):
    from brawl_amd import wang_landau, workloads
    c = workloads.config4()
    wl = wang_landau.WangLandau(c["lattice"], c["n"], c["v_ex"], bins=c["bins"],
                                num_windows=c["num_windows"],
                                walkers_per_window=c["walkers_per_window"],
                                bin_overlap=c["bin_overlap"], energy_min=c["energy_min"],
                                energy_max=c["energy_max"])
    assert wl.n_walkers == 128
    ref = oracle.OracleLattice(c["lattice"], c["n"], c["v_ex"])
    windows_ref, edges_ref = ref.wl_setup(c["bins"], c["num_windows"], c["bin_overlap"],
                                          c["energy_min"], c["energy_max"])
    assert np.array_equal(wl.windows, windows_ref) and np.array_equal(wl.bin_edges, edges_ref)
    wl.initial_setup(species_numbers=c["counts"])
    _say("config4: window entry of 128 walkers")
    trials, entered = wl.enter_windows(10**8)
    assert entered.all(), f"{(~entered).sum()} walkers outside their windows"
    N, wpw = wl.n_atoms, wl.walkers_per_window
    cadence = int(0.02 * N)

    def recount_all():
        cfgs = [wl.get_walker(k) for k in range(wl.n_walkers)]
        for k, cfg in enumerate(cfgs):
            assert np.array_equal(np.bincount(cfg[cfg > 0], minlength=5)[1:], c["counts"])
            s, e = wl.windows[wl.window_of(k)]
            b = _bin_of(ref.total_energy(cfg), wl.bin_edges)
            assert s <= b <= e, (k, b, s, e)
        return cfgs

    recount_all()
    lng0, hist0 = wl.lng_hist()
    assert not lng0.any() and not hist0.any()
    # call 1 without the window average: two walkers followed by the oracle
    spot = [0, 77]
    before = {k: (wl.get_walker(k), wl.genrand_state(k)) for k in spot}
    wl_f = 1.0
    acc, ms = wl.sweeps(wl_f, average=False)
    _say("config4 sweeps call (128 walkers x 8192 trials):", ms, "ms")
    lng, hist = wl.lng_hist()
    for k in spot:
        cfg, (mt, mti) = before[k]
        gen = oracle.Generator.from_state(mt, mti)
        s, e = wl.windows[wl.window_of(k)]
        lk = np.zeros(wl.bins)
        hk = np.zeros(e - s + 1)
        a = ref.wl_sweeps(cfg, gen, mc_sweeps=1, bin_edges=wl.bin_edges, start_idx=int(s),
                          end_idx=int(e), wl_f=wl_f, lng=lk, hist=hk)
        assert acc[k] == a
        assert np.array_equal(wl.get_walker(k), cfg)
        assert np.array_equal(lng[k], lk) and np.array_equal(hist[k, :e - s + 1], hk)
        mt2, mti2 = wl.genrand_state(k)
        mt_ref, mti_ref = gen.state_arrays()
        assert mti2 == mti_ref and np.array_equal(mt2, mt_ref)
    # every walker: ln g grew by wl_f per trial, the histogram by the cadence
    assert np.array_equal(lng.sum(axis=1), np.full(wl.n_walkers, wl_f * N))
    assert np.array_equal(hist.sum(axis=1), np.full(wl.n_walkers, float(N // cadence)))
    # window averages back in, then two more calls with the average and a
    # replica exchange after each (the reference's main loop)
    wl.set_lng_hist(wl.dos_average(lng), wl.dos_average(hist))
    exchanged = 0
    for call in range(2):
        lng_b, hist_b = wl.lng_hist()
        wl.sweeps(wl_f, average=True)
        lng, hist = wl.lng_hist()
        for w in range(wl.num_windows):
            rows = slice(w * wpw, (w + 1) * wpw)
            assert (lng[rows] == lng[rows][0]).all()           # averaged over the window
            grow = (lng[rows][0] - lng_b[rows][0]).sum()
            assert abs(grow - wl_f * N) <= 1e-9 * N
            assert abs((hist[rows][0] - hist_b[rows][0]).sum() - N // cadence) <= 1e-9 * N
        partner, ex, _ = wl.replica_exchange()
        exchanged += int(ex.sum()) // 2
        for k in range(wl.n_walkers):
            if partner[k] >= 0:
                assert partner[partner[k]] == k
                assert abs(wl.window_of(int(partner[k])) - wl.window_of(k)) == 1
        recount_all()
    _say("config4: pairs exchanged over 2 calls:", exchanged)
    wl.close()


def _oracle_wl_run(oracle, ref, species_numbers, edges, window, emin, emax, ranks, wl_f,
                   tolerance, flatness):
    """WangLandau.run for one window on the oracle: enter_energy_window per
    walker, pre-sampling (sweeps with the window average until the averaged
    histogram is flat with every bin above 1000 / walkers), then the main
    loop, halving ln f whenever the averaged histogram is flat."""
    s, e = int(window[0]), int(window[1])
    nb = e - s + 1
    W = len(ranks)
    gens = [oracle.Generator(rank=r) for r in ranks]
    cfgs = [ref.initial_setup(g, species_numbers=species_numbers) for g in gens]
    for k in range(W):
        ref.wl_enter_window(cfgs[k], gens[k], edges[s - 1], edges[e], emin, emax,
                            species_numbers=species_numbers)
    lng, hist = np.zeros((W, len(edges) - 1)), np.zeros((W, nb))

    def call(f):
        for k in range(W):
            ref.wl_sweeps(cfgs[k], gens[k], mc_sweeps=1, bin_edges=edges, start_idx=s, end_idx=e,
                          wl_f=f, lng=lng[k], hist=hist[k])
        for a in (lng, hist):               # the window average, walker order
            m = a[0].copy()
            for k in range(1, W):
                m = m + a[k]
            a[:] = m / float(W)

    calls = 0
    while True:
        call(wl_f)
        calls += 1
        h = hist[0]
        if h.min() > 1000.0 / W and h.min() / (h.sum() / nb) > flatness:
            break
    lng[:] = lng[0] - lng[0].min()
    hist[:] = 0.0
    halvings = 0
    while wl_f > tolerance:
        call(wl_f)
        calls += 1
        h = hist[0]
        if h.sum() > 0 and h.min() / (h.sum() / nb) > flatness:
            wl_f *= 0.5
            halvings += 1
            lng[:] = lng[0] - lng[0].min()
            hist[:] = 0.0
    return lng[0].copy(), calls, halvings


def test_wang_landau_reproduces_exact_dos(hip, oracle):
    from brawl_amd import wang_landau
    lat, n, N = 2, (3, 3, 3), 54
    V = np.zeros((2, 2, 1), order="F")
    V[0, 0, 0], V[0, 1, 0], V[1, 0, 0], V[1, 1, 0] = 2e-3, 0.5e-3, 0.5e-3, 1e-3   # J = 2 mRy
    sites, combs, m, levels, counts = _exact_dos()
    assert list(levels) == list(range(7)) and counts.sum() == 3162510
    # the energies of the levels by the oracle (pinned to the reference):
    # E = E_0 + m J exactly spaced
    ref = oracle.OracleLattice(lat, n, V)
    E = []
    for lv in levels:
        cfg = np.zeros((6, 6, 6), dtype=np.int8, order="F")
        for s in sites:
            cfg[s] = 2
        for p in combs[np.nonzero(m == lv)[0][0]]:
            cfg[sites[p]] = 1
        E.append(ref.total_energy(cfg))
    E = np.array(E)
    J = 2e-3
    assert np.allclose(np.diff(E), J, rtol=1e-12, atol=0)
    # 7 bins centred on the 7 levels, in meV/atom
    to_mev = 13.605693122 * 1000.0 / N
    emin, emax = (E[0] - J / 2) * to_mev, (E[-1] + J / 2) * to_mev
    exact = np.log(counts / counts[0])
    sn = np.array([5, 49])
    runs = []
    for r in range(6):
        wl = wang_landau.WangLandau(lat, n, V, bins=7, num_windows=1, walkers_per_window=4,
                                    bin_overlap=0.0, energy_min=emin, energy_max=emax,
                                    rank_offset=4 * r)
        assert [_bin_of(e, wl.bin_edges) for e in E] == list(range(1, 8))
        wl.initial_setup(species_numbers=sn)
        res = wl.run(wl_f=1.0, tolerance=1e-4, flatness=0.8)
        _say(f"run {r}: {res.iterations} halvings, {res.sweeps_calls} calls, ln g - ln g0 =",
             np.round(res.lng - res.lng[0], 3))
        runs.append(res.lng - res.lng[0])
        if r < 2:
            # the whole run, step for step, on the oracle
            lng_o, calls_o, halv_o = _oracle_wl_run(
                oracle, ref, sn, wl.bin_edges, wl.windows[0], float(np.float32(emin)),
                float(np.float32(emax)), range(4 * r, 4 * r + 4), 1.0, 1e-4, 0.8)
            assert (calls_o, halv_o) == (res.sweeps_calls, res.iterations)
            assert np.array_equal(lng_o, res.lng)
        wl.close()
    runs = np.array(runs)
    mean, sd = runs.mean(axis=0), runs.std(axis=0, ddof=1)
    _say("exact", np.round(exact, 3), "mean", np.round(mean, 3), "sd of one run", np.round(sd, 3))
    dev = (mean[1:] - exact[1:]) / sd[1:]
    _say("deviation / sd", np.round(dev, 2))
    assert np.all(np.abs(dev) <= 2.0), dev
