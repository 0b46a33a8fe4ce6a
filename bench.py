"""Benchmark: swap trials/s of the lattice-sampling hot path on one MI355X.

Workload (one "step"): the quaternary bcc anneal -- bcc 32^3 conventional
cells (65536 sites), 4 equiatomic species, 4 shells with V^(n) ~ N(0, 30/n
meV) (zero-mean per shell), T from 3000 K to 100 K in 50 linear steps of 200
sweeps each, lattice-wide swaps; the energy is recounted from the lattice
before the first T step and at the end of each (the carried value's largest
difference from the recounts is reported) and carried through the sweeps as
the sum of the accepted dE, sampled after every sweep.  One step = the whole anneal
(>= 655.36 M swap trials) from the same synthetic start, on an HBM-resident
lattice, through the C-ABI.

--calibrate-copy runs only a plain 256 MiB device-to-device copy (three
times), the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters
into bytes.

The served form (config 5): a Poisson stream of mixed requests (Metropolis
anneals, nested-sampling runs, short Wang-Landau runs) submitted at their
arrival times to the request server (continuous batching under a device
memory budget), with per-class p50/p99 latency from arrival to completion;
the same requests are then run one at a time through the plain entry points
for the plain rate and to check that every served result equals its solo
run.

Prints ONE JSON line.  With --gpus N > 1 it exits: the library runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
# hardware queues per process (HIP's default is 4), set before HIP starts:
# with 8, every concurrent launch (batch groups, server slots) has a queue
# of its own (measured on the 256-chain batch: 380 ms with 4 queues, 271 ms
# with 8 or 16, 468 ms with 32)
HW_QUEUES = next((a.split("=", 1)[1] for a in sys.argv if a.startswith("--hw-queues=")), "8")
os.environ["GPU_MAX_HW_QUEUES"] = HW_QUEUES


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_trial(z: int) -> int:
    # int8 occupations of both sites and their z neighbours, plus two writes
    return 2 * (z + 1) + 2


# HBM bytes per launch of the plain anneal kernel, from rocprofv3 --pmc
# FETCH_SIZE (106.125 KiB, x2.0 copy-kernel scale) and WRITE_SIZE
# (221.25 KiB, x1.0): the lattice is read once into LDS and written back
# once, plus the per-T-step outputs and the register spills of the
# radial-density reduction.
TRAFFIC_BYTES = int(round((106.125 * 2.0 + 221.25) * 1024))
PLAIN_KERNEL = "k_metro_nat<bcc, R=4, 32-bit counts, 32 cells>"
# radial densities (alpha^(1..4)) sampled every 10 sweeps of each T step
ASRO_EVERY = 10


def _alpha_summary(res, chain):
    """Warren-Cowley alpha^(n)(a,b) at the last temperature, shells 1..R."""
    from brawl_amd.lattice import warren_cowley
    z = {1: [12, 6, 24, 12, 24, 8], 2: [8, 6, 12, 24, 8, 6, 24, 24, 24, 32]}[chain.lattice][: chain.R]
    conc = chain.counts / chain.counts.sum()
    a = warren_cowley(res.rho_of_T[:, :, :, -1], z, conc)
    return [[[round(float(a[i, j, n]), 6) for j in range(chain.S)] for i in range(chain.S)]
            for n in range(1, chain.R + 1)]


def calibrate_copy():
    import torch
    n = 256 << 20
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "copies": 3}))


def cpu_baseline(chain, budget_s: float):
    """The reference's serial path on a bounded sample of the same anneal
    (three of its temperature steps, as many sweeps as fit the time budget):
    when the reference build is present (oracle/_ref, compiled from the
    reference's sources: its own rdm_site, pair_energy, pair_swap, genrand
    in the order of monte_carlo_step_lattice, full_energy after every sweep,
    timed around the trial loops), kind "reference"; the oracle (the C
    restatement, "port") is timed beside it.  Also the same sample as
    independent chains, one per available core (the reference's MPI replica
    mode), for the aggregate rate.  Runs before the GPU is touched: the
    reference runs in forked processes."""
    from concurrent.futures import ThreadPoolExecutor
    from brawl_amd.lattice import beta_of
    from oracle import oracle as orc
    from oracle import ref as oref
    orc.build()
    lat = orc.OracleLattice(chain.lattice, chain.n, chain.v_ex)
    gen = orc.Generator(110179)
    cfg = lat.initial_setup(gen, species_numbers=chain.counts)
    N = lat.n_atoms
    cores = min(len(os.sched_getaffinity(0)), 16)
    # the oracle: calibrate on one sweep, then a sample of ~budget/2 seconds
    c = np.array(cfg, order="F", copy=True)
    t0 = time.perf_counter()
    lat.anneal(c, gen, T=chain.T, n_mc_steps=N, n_sample_steps=N)
    one = time.perf_counter() - t0
    sweeps = max(1, int(0.5 * budget_s / max(one, 1e-6)))
    t0 = time.perf_counter()
    lat.anneal(c, gen, T=chain.T, n_mc_steps=sweeps * N, n_sample_steps=N)
    dt = time.perf_counter() - t0
    port = {"value": sweeps * N / dt, "unit": "swap trials/s", "cores": 1, "kind": "port",
            "sample": f"first T step (3000 K) of the same anneal, {sweeps} sweeps = {sweeps * N} "
                      f"trials with the reference's per-sweep energy recount, {dt:.1f} s"}
    if not oref.available():
        rsweeps = max(1, sweeps // 2)

        def replica(k):
            lk = orc.OracleLattice(chain.lattice, chain.n, chain.v_ex)
            g = orc.Generator(110179 + 11 * k)
            ck = lk.initial_setup(g, species_numbers=chain.counts)
            lk.anneal(ck, g, T=chain.T, n_mc_steps=rsweeps * N, n_sample_steps=N)
            return rsweeps * N

        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=cores) as ex:
            rtrials = sum(ex.map(replica, range(cores)))
        rdt = time.perf_counter() - t0
        port["replicas"] = {"value": rtrials / rdt, "cores": cores,
                            "sample": f"{cores} oracle chains (one per core), {rsweeps} sweeps "
                                      f"each, {rdt:.1f} s"}
        return port
    # three temperatures of the anneal (its first step, the heat-capacity
    # peak, its last step), a third of the budget each, from the same start
    temps = [chain.T, chain.T + chain.delta_T * 33, chain.T + chain.delta_T * (chain.T_steps - 1)]
    betas = [beta_of(t) for t in temps]
    _, _, one, _ = oref.metropolis_sweeps(chain.lattice, chain.n, chain.v_ex, cfg, betas[0], 1)
    rs = max(1, int(budget_s / 3 / max(one, 1e-6)))
    per_T, secs = [], 0.0
    for t, b in zip(temps, betas):
        _, _, st, _ = oref.metropolis_sweeps(chain.lattice, chain.n, chain.v_ex, cfg, b, rs)
        per_T.append({"T": round(t, 1), "value": rs * N / st})
        secs += st
    rsweeps = max(1, rs // 2)
    t0 = time.perf_counter()
    reps = oref.run_parallel(_ref_chain, [(chain, cfg, k, betas[k % 3], rsweeps) for k in range(cores)])
    rdt = time.perf_counter() - t0
    return {"value": 3 * rs * N / secs, "unit": "swap trials/s", "cores": 1, "kind": "reference",
            "sample": f"three T steps of the same anneal ({temps[0]:.0f} K, {temps[1]:.0f} K, "
                      f"{temps[2]:.0f} K) from the same random start through the reference's own "
                      f"compiled routines, {rs} sweeps = {rs * N} trials each with the per-sweep "
                      f"energy recount, {secs:.1f} s in the trial loops",
            "per_T": per_T,
            "replicas": {"value": cores * rsweeps * N / rdt, "cores": cores,
                         "sample": f"{cores} reference chains in parallel processes (ranks "
                                   f"0..{cores - 1}, the three temperatures in turn), {rsweeps} "
                                   f"sweeps each, {rdt:.1f} s wall (longest chain "
                                   f"{max(r[2] for r in reps):.1f} s in its loops); 16 cores is the "
                                   f"host share of one GPU"},
            "port": port}


def _ref_chain(chain, cfg, k, beta, n_sweeps):
    """One reference chain (rank k's stream) at beta for n_sweeps sweeps."""
    from oracle import ref as oref
    return oref._metropolis_sweeps(chain.lattice, chain.n, chain.v_ex, cfg, beta, n_sweeps, rank=k)


def _wl_walker_sample(k: int, n_calls: int):
    """One config-4 walker of the middle window through the reference's own
    compiled `sweeps` loop (ref_wl_sweeps: rdm_site, pair_energy, pair_swap,
    bin_index, log(genrand()) as wang-landau.F90:539-632, full_energy at the
    start of every call): entered into its window on the oracle first
    (untimed), then n_calls calls of one sweep timed around the call.
    Returns (trials, seconds)."""
    from brawl_amd import workloads
    from oracle import oracle as orc
    from oracle import ref as oref
    c = workloads.config4()
    lat = orc.OracleLattice(c["lattice"], c["n"], c["v_ex"])
    windows, edges = lat.wl_setup(c["bins"], c["num_windows"], c["bin_overlap"], c["energy_min"],
                                  c["energy_max"])
    s, e = (int(x) for x in windows[c["num_windows"] // 2])
    gen = orc.Generator(rank=k)
    cfg = lat.initial_setup(gen, species_numbers=c["counts"])
    lat.wl_enter_window(cfg, gen, edges[s - 1], edges[e], c["energy_min"], c["energy_max"],
                        species_numbers=c["counts"])
    lng, hist = np.zeros(c["bins"]), np.zeros(e - s + 1)
    t0 = time.perf_counter()
    oref._wl_sweeps(c["lattice"], c["n"], c["v_ex"], cfg, rank=k, nbr_swap=0, mc_sweeps=1,
                    n_calls=n_calls, bin_edges=edges, start_idx=s, end_idx=e, wl_f=1.0, lng=lng,
                    hist=hist)
    return n_calls * lat.n_atoms, time.perf_counter() - t0


def ns_request_shape():
    """The nested-sampling class of the served stream: quinary fcc 6^3 (864
    sites), 4 shells of 30/n meV, walk length 100; 384 walkers (the middle of
    its 128-640)."""
    from brawl_amd import workloads
    q = next(r for r in workloads.config5(n_requests=64) if r.kind == "ns")
    return q, 384


def _ns_sample(k: int, n_iter: int):
    """One nested-sampling run of that shape through the reference's own
    compiled routines (ref_nested_sampling: initial_setup, full_energy,
    rdm_site, pair_energy, pair_swap, genrand in the order of
    nested_sampling.f90:78-200), its iteration loop timed.  Returns
    (walk trials, seconds)."""
    from oracle import ref as oref
    q, walkers = ns_request_shape()
    _, trials, secs = oref._nested_sampling(q.lattice, q.n, q.v_ex, q.counts, rank=k,
                                            n_walkers=walkers, n_steps=100, n_iter=n_iter)
    return trials, secs


def cpu_walkers(budget_s: float):
    """The reference's walker loops on the host cores, one chain per core
    (the reference runs one walker per MPI rank): Wang-Landau `sweeps` on a
    config-4 walker and the nested-sampling loop on the served stream's NS
    shape, each on 1 core and on every available core at once (up to 16,
    the host share of one GPU).  None without the reference build."""
    from oracle import ref as oref
    if not oref.available():
        return None
    cores = min(len(os.sched_getaffinity(0)), 16)
    out = {}
    # Wang-Landau: calibrate with one call, then ~budget/4 seconds per leg
    tr, dt = oref.run_parallel(_wl_walker_sample, [(0, 2)])[0]
    calls = max(2, int(0.25 * budget_s / max(dt / 2, 1e-6)))
    tr1, dt1 = oref.run_parallel(_wl_walker_sample, [(0, calls)])[0]
    many = oref.run_parallel(_wl_walker_sample, [(k, calls) for k in range(cores)])
    out["wang_landau"] = {
        "value": tr1 / dt1, "unit": "swap trials/s", "cores": 1, "kind": "reference",
        "all_cores": {"value": sum(t for t, _ in many) / max(d for _, d in many), "cores": cores},
        "sample": f"one config-4 walker (bcc 16^3, 2 shells, middle window of 16) through the "
                  f"reference's compiled sweeps loop, {calls} calls of 8192 trials; all cores: "
                  f"{cores} walkers (ranks 0..{cores - 1}) at once"}
    # nested sampling
    tr, dt = oref.run_parallel(_ns_sample, [(0, 20)])[0]
    iters = max(20, int(0.25 * budget_s / max(dt / 20, 1e-6)))
    tr1, dt1 = oref.run_parallel(_ns_sample, [(0, iters)])[0]
    many = oref.run_parallel(_ns_sample, [(k, iters) for k in range(cores)])
    out["nested_sampling"] = {
        "value": tr1 / dt1, "unit": "swap trials/s", "cores": 1, "kind": "reference",
        "all_cores": {"value": sum(t for t, _ in many) / max(d for _, d in many), "cores": cores},
        "sample": f"quinary fcc 6^3, 4 shells, 384 walkers, walk length 100, {iters} iterations "
                  f"through the reference's compiled routines; all cores: {cores} runs at once"}
    return out


def nested_sampling_runs(n_runs: int = 64, n_iter: int = 200, fill_runs: int = 256):
    """Nested-sampling runs of the served stream's NS shape on the device:
    one run alone (one wavefront), then n_runs runs (their own seeds) in
    one batched launch, then fill_runs runs (one per CU of the MI355X)."""
    from brawl_amd import Lattice, nested_sampling
    q, walkers = ns_request_shape()
    p = dict(n_walkers=walkers, n_steps=100, n_iter=n_iter, species_numbers=q.counts)
    with Lattice(q.lattice, q.n, q.v_ex) as L:
        L.seed(110179)
        nested_sampling.nested_sampling(L, **p)                 # warm-up
        L.seed(110190)
        _, st = nested_sampling.nested_sampling(L, **p)
    one = st.trials / (st.kernel_ms / 1e3)
    def batch(runs):
        lats = []
        for k in range(runs):
            Lk = Lattice(q.lattice, q.n, q.v_ex)
            Lk.seed(110179 + 11 * k)
            lats.append(Lk)
        res, mk = nested_sampling.nested_sampling_batch(lats, [p] * runs)
        for Lk in lats:
            Lk.close()
        trials = sum(s.trials for _, s in res)
        log(f"[bench] nested sampling: {runs} runs {trials} trials in {mk:.1f} ms")
        return {"runs": runs, "trials": trials, "makespan_ms": mk, "value": trials / (mk / 1e3),
                "unit": "swap trials/s"}

    log(f"[bench] nested sampling: one run {one:.3e} trials/s")
    return {"workload": f"quinary fcc 6^3 (864 sites), 4 shells, {walkers} walkers, walk length 100, "
                        f"{n_iter} iterations",
            "one_run": {"value": one, "unit": "swap trials/s", "us_per_trial": 1e6 / one},
            "batched": batch(n_runs), "batched_filled": batch(fill_runs)}


def batched(steps: int, warmup: int, tune: bool = True):
    """The batched form: 256 uneven Metropolis chains (fcc/bcc, 8^3..24^3
    cells, 2-5 species with Dirichlet(1) counts, 3 shells, T in [200,2000] K,
    500 sweeps each) in one batched call; first with the layouts the library
    picks by rule, then (tune) with the layouts the autotuner picks per
    shape (tuning time reported, outside the timed steps)."""
    from brawl_amd import Lattice, anneal_batch, autotune, workloads
    chains = workloads.config3()
    lats, params = [], []
    for ch in chains:
        L = Lattice(ch.lattice, ch.n, ch.v_ex)
        L.seed(ch.seed & 0xFFFFFFFF)
        L.initial_setup(species_numbers=ch.counts)
        lats.append(L)
        params.append(dict(T=ch.T, n_sweeps=ch.n_sweeps, seed=ch.seed))
    starts = [L.get_config() for L in lats]

    def run(tag):
        rates, mks, trials, alg_bytes = [], [], 0, 0
        for k in range(warmup + steps):
            for L, s0 in zip(lats, starts):
                L.set_config(s0)
            t0 = time.perf_counter()
            res, mk = anneal_batch(lats, params)
            dt = time.perf_counter() - t0
            tr = sum(r.trials for r in res)
            # algorithmic bytes: B_trial of each chain's shells (3: fcc z=42, bcc z=26)
            alg = sum(r.trials * bytes_per_trial(42 if ch.lattice == 1 else 26)
                      for r, ch in zip(res, chains))
            log(f"[bench] batched ({tag}) {'warmup' if k < warmup else 'step'}: {tr} trials, "
                f"makespan {mk:.1f} ms, wall {dt * 1e3:.1f} ms")
            if k >= warmup:
                rates.append(tr / dt)
                mks.append(mk)
                trials, alg_bytes = tr, alg
        return rates, mks, trials, alg_bytes

    rates, mks, trials, alg_bytes = run("layouts by rule")
    out = {"workload": "256 chains, L~U{8..24} fcc/bcc, S~U{2..5}, 3 shells, T~U[200,2000] K, "
                       "500 sweeps", "chains": len(chains), "trials": trials,
           "value": float(np.mean(rates)), "unit": "swap trials/s",
           "device_rate": trials / (np.mean(mks) / 1e3), "makespan_ms": float(np.mean(mks)),
           "bytes_per_step": int(alg_bytes)}
    # the CUs' busy share over the makespan: each chain run alone from its
    # start (the same trajectory as in the batch, bit for bit) for its own
    # device time, summed, against every CU for the batch's makespan
    from brawl_amd import _lib
    _, n_cus = _lib.lds_bandwidth()
    solo_ms = []
    for L, s0, p in zip(lats, starts, params):
        L.set_config(s0)
        solo_ms.append(L.metropolis_simulated_annealing(**p).kernel_ms)
    out["cu_busy_share"] = float(sum(solo_ms) / (n_cus * out["makespan_ms"]))
    out["cu_busy_note"] = (f"sum of the {len(solo_ms)} chains' device times alone "
                           f"({sum(solo_ms):.0f} ms; longest {max(solo_ms):.1f} ms) over {n_cus} CUs "
                           f"x the makespan")
    log(f"[bench] batched: CUs busy {out['cu_busy_share']:.2f} of the makespan "
        f"(longest chain alone {max(solo_ms):.1f} ms)")
    if tune:
        for L, s0 in zip(lats, starts):
            L.set_config(s0)
        t0 = time.perf_counter()
        table = autotune.tune(lats)
        tune_s = time.perf_counter() - t0
        table.apply(lats)
        picks = list(table.choices.values())
        # shapes whose tuned thread count differs from the rule's for the tuned layout
        off_rule, seen = 0, set()
        for L in lats:
            k = autotune.shape_key(L)
            if k in seen or k not in table.choices:
                continue
            seen.add(k)
            off_rule += table.threads[k] != L.layout_info(table.choices[k])[0]
        t_rates, t_mks, t_trials, _ = run("autotuned layouts and threads")
        out["autotuned"] = {"value": float(np.mean(t_rates)),
                            "device_rate": t_trials / (np.mean(t_mks) / 1e3),
                            "makespan_ms": float(np.mean(t_mks)), "tuning_s": tune_s,
                            "shapes": len(picks), "natural": picks.count("natural"),
                            "fixed": picks.count("fixed"), "threads_not_rule": off_rule}
    for L in lats:
        L.close()
    return out


def lds_roofline(plain_gbs: float, batched_out=None):
    """The samplers' lattices live in LDS: their algorithmic bytes against
    the LDS read bandwidth measured by the library's probe (all CUs, and one
    CU's share for the one-chain plain form)."""
    from brawl_amd import _lib
    bps, cus = _lib.lds_bandwidth()
    peak = bps / 1e9
    out = {"bound": "lds", "unit": "GB/s", "peak": peak, "cus": cus, "peak_per_cu": peak / cus,
           "probe": "conflict-free 16-byte LDS reads, 8 blocks x 1024 threads per CU",
           "plain_achieved": plain_gbs, "plain_frac": plain_gbs / peak,
           "plain_frac_of_one_cu": plain_gbs / (peak / cus)}
    if batched_out:
        a = batched_out["bytes_per_step"] / (batched_out["makespan_ms"] / 1e3) / 1e9
        out["batched_achieved"] = a
        out["batched_frac"] = a / peak
    return out


def _served_submit(srv, r):
    p = r.params
    if r.kind == "anneal":
        return srv.submit_anneal(r.lattice, r.n, r.v_ex, r.config, T=p["T"], n_sweeps=p["n_sweeps"],
                                 seed=r.seed)
    if r.kind == "ns":
        return srv.submit_ns(r.lattice, r.n, r.v_ex, r.seed, n_walkers=p["n_walkers"],
                             n_steps=p["n_steps"], n_iter=p["n_iter"], species_numbers=r.counts)
    return srv.submit_wl(r.lattice, r.n, r.v_ex, bins=p["bins"], num_windows=p["num_windows"],
                         walkers_per_window=p["walkers_per_window"], bin_overlap=p["bin_overlap"],
                         energy_min=p["energy_min"], energy_max=p["energy_max"],
                         mc_sweeps=p["mc_sweeps"], wl_f=p["wl_f"], rank_offset=p["rank_offset"],
                         species_numbers=r.counts)


def _plain_run(r):
    """The same request alone through the plain entry points; returns
    (outputs, trials, seconds of the run call, seconds of device time)."""
    from brawl_amd import Lattice, nested_sampling, wang_landau
    p = r.params
    if r.kind == "anneal":
        with Lattice(r.lattice, r.n, r.v_ex) as L:
            L.set_config(r.config)
            t0 = time.perf_counter()
            res = L.metropolis_simulated_annealing(T=p["T"], n_sweeps=p["n_sweeps"], seed=r.seed)
            dt = time.perf_counter() - t0
            return (dict(energies_of_T=res.energies_of_T, C_of_T=res.C_of_T,
                         acceptance_of_T=res.acceptance_of_T), res.trials, dt, res.kernel_ms / 1e3)
    if r.kind == "ns":
        with Lattice(r.lattice, r.n, r.v_ex) as L:
            L.seed(r.seed)
            t0 = time.perf_counter()
            lim, st = nested_sampling.nested_sampling(L, n_walkers=p["n_walkers"], n_steps=p["n_steps"],
                                                      n_iter=p["n_iter"], species_numbers=r.counts)
            return dict(e_limits=lim), st.trials, time.perf_counter() - t0, st.kernel_ms / 1e3
    wl = wang_landau.WangLandau(r.lattice, r.n, r.v_ex, bins=p["bins"], num_windows=p["num_windows"],
                                walkers_per_window=p["walkers_per_window"], bin_overlap=p["bin_overlap"],
                                energy_min=p["energy_min"], energy_max=p["energy_max"],
                                mc_sweeps=p["mc_sweeps"], rank_offset=p["rank_offset"])
    t0 = time.perf_counter()
    wl.initial_setup(species_numbers=r.counts)
    _, k_ms = wl.sweeps(p["wl_f"], average=True)
    dt = time.perf_counter() - t0
    lng, _ = wl.lng_hist()
    out = dict(lng=lng[:: p["walkers_per_window"]].copy())
    wl.close()
    trials = p["num_windows"] * p["walkers_per_window"] * p["mc_sweeps"] * r.n_atoms
    return out, trials, dt, k_ms / 1e3


def _serve_stream(args, reqs):
    """Submit reqs at their arrival times to a fresh server, drain it;
    returns (results, submission delays after arrival, wall s, peak bytes)."""
    from brawl_amd.serve import Server
    with Server(device_bytes=args.served_budget_mb << 20, max_launches=args.served_slots,
                batch_wait_us=args.served_wait_us) as srv:
        ids, late = [], []
        t0 = time.perf_counter()
        for r in reqs:
            while time.perf_counter() - t0 < r.arrival_s:
                srv.progress()
            late.append(time.perf_counter() - t0 - r.arrival_s)   # submission after arrival
            ids.append(_served_submit(srv, r))
            srv.progress()
        srv.drain()
        wall = time.perf_counter() - t0
        res = [srv.result(i) for i in ids]
        peak_mem = srv.memory()[1]
    return res, late, wall, peak_mem


def _latency_by_class(reqs, res, late):
    classes = {}
    for r, x, lt in zip(reqs, res, late):
        lat_ms = x["t_done_ms"] - x["t_submit_ms"] + lt * 1e3
        c = classes.setdefault(r.kind, dict(lat=[], trials=0, n=0, k=[]))
        c["lat"].append(lat_ms)
        c["k"].append(x["kernel_ms"])
        c["trials"] += x["trials"]
        c["n"] += 1
    return classes


def served_sweep(args, rates=(200.0, 400.0, 800.0, 1600.0, 3200.0, 6400.0)):
    """The served stream at increasing Poisson arrival rates (the same
    request mix, args.served_requests requests each): aggregate swap
    trials/s over the makespan and per-class p50/p99 latency from arrival.
    A rate is sustained while the server keeps up: the requests arriving in
    the second half wait no longer (p99) than twice those of the first half;
    past that, a backlog builds and latency grows with the length of the
    stream."""
    from brawl_amd import workloads
    out = []
    for lam in rates:
        reqs = workloads.config5(args.served_requests, lam)
        res, late, wall, _ = _serve_stream(args, reqs)
        trials = sum(x["trials"] for x in res)
        first = min(x["t_submit_ms"] for x in res)
        last = max(x["t_done_ms"] for x in res)
        lat = np.array([x["t_done_ms"] - x["t_submit_ms"] + lt * 1e3 for x, lt in zip(res, late)])
        half = len(lat) // 2
        p99_a, p99_b = float(np.percentile(lat[:half], 99)), float(np.percentile(lat[half:], 99))
        classes = _latency_by_class(reqs, res, late)
        row = {"rate_hz": lam, "offered_trials_per_s": trials / reqs[-1].arrival_s,
               "value": trials / ((last - first) / 1e3), "makespan_ms": last - first,
               "p99_first_half_ms": p99_a, "p99_second_half_ms": p99_b,
               "sustained": p99_b <= 2.0 * p99_a,
               "per_class": {k: {"p50_ms": float(np.percentile(c["lat"], 50)),
                                 "p99_ms": float(np.percentile(c["lat"], 99))}
                             for k, c in classes.items()}}
        log(f"[bench] served sweep {lam:g}/s: {row['value']:.3e} trials/s, p99 {p99_a:.0f} / "
            f"{p99_b:.0f} ms, sustained {row['sustained']}")
        out.append(row)
    sustained = [r["rate_hz"] for r in out if r["sustained"]]
    return {"requests_per_rate": args.served_requests, "rates": out,
            "max_sustained_rate_hz": max(sustained) if sustained else None,
            "peak_value": max(r["value"] for r in out)}


def served(args):
    """Config 5 through the request server, then the same requests plain."""
    from brawl_amd import workloads
    reqs = workloads.config5(args.served_requests, args.served_rate)
    log(f"[bench] served: {len(reqs)} requests at {args.served_rate:g}/s")
    res, late, wall, peak_mem = _serve_stream(args, reqs)
    trials = sum(x["trials"] for x in res)
    first = min(x["t_submit_ms"] for x in res)
    last = max(x["t_done_ms"] for x in res)
    makespan_s = (last - first) / 1e3
    classes = _latency_by_class(reqs, res, late)
    log(f"[bench] served: {trials} trials, makespan {makespan_s * 1e3:.1f} ms, wall {wall * 1e3:.1f} ms")
    # plain: one request at a time, and the served results checked against it
    plain_s, plain_dev_s, plain_trials, mismatch = 0.0, 0.0, 0, 0
    subset = list(zip(reqs, res))[:: args.served_plain_every]
    for r, x in subset:
        out, tr, dt, dev = _plain_run(r)
        plain_s += dt
        plain_dev_s += dev
        plain_trials += tr
        same = tr == x["trials"] and all(np.array_equal(v, x[k]) for k, v in out.items())
        mismatch += 0 if same else 1
    log(f"[bench] plain over {len(subset)} of the requests: {plain_trials} trials in {plain_s:.2f} s "
        f"({plain_dev_s:.2f} s device), {mismatch} results differ from the served ones")
    per_class = {k: dict(requests=c["n"], trials=c["trials"],
                         p50_ms=float(np.percentile(c["lat"], 50)),
                         p99_ms=float(np.percentile(c["lat"], 99)),
                         launch_device_ms_mean=float(np.mean(c["k"]))) for k, c in classes.items()}
    value = trials / makespan_s
    # the plain rate counts device time only (the most favourable to it)
    plain_rate = plain_trials / plain_dev_s
    return {"workload": f"{len(reqs)} requests, Poisson arrivals at {args.served_rate:g}/s: 60% Metropolis "
                        "anneals (fcc L 8-28 / bcc L 8-32, S 2-5, 3 shells, 100-500 sweeps), 25% nested "
                        "sampling (quinary fcc 6^3, 128-640 walkers, walk length 100, 200 iterations), "
                        "15% short Wang-Landau runs (bcc L 6-10, 4-8 walkers, 20 sweeps)",
            "value": value, "unit": "swap trials/s", "trials": trials,
            "makespan_ms": makespan_s * 1e3, "wall_ms": wall * 1e3,
            "latency_from_arrival": per_class,
            "plain_same_requests": plain_rate,
            "plain_same_requests_wall": plain_trials / plain_s, "vs_plain": value / plain_rate,
            "plain_sample": f"every {args.served_plain_every}th request ({len(subset)} of {len(reqs)}), "
                            "run alone one after another through the plain entry points; the plain "
                            "rate counts device time only",
            "results_checked_against_solo": len(subset),
            "results_differing_from_solo": mismatch,
            "server": {"device_bytes": peak_mem, "launch_slots": args.served_slots,
                       "batch_wait_us": args.served_wait_us,
                       "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))}}


def wang_landau_walkers(calls: int = 10):
    """Config 4's walker set (16 windows x 8 walkers on bcc 16^3): window
    entry (bounded), then `calls` sweeps calls of one sweep each, every
    walker in one launch; returns the walkers' aggregate trial rate."""
    from brawl_amd import wang_landau, workloads
    c = workloads.config4()
    wl = wang_landau.WangLandau(c["lattice"], c["n"], c["v_ex"], bins=c["bins"],
                                num_windows=c["num_windows"], walkers_per_window=c["walkers_per_window"],
                                bin_overlap=c["bin_overlap"], energy_min=c["energy_min"],
                                energy_max=c["energy_max"], mc_sweeps=1)
    wl.initial_setup(species_numbers=c["counts"])
    _, entered = wl.enter_windows(20 * 8192)
    ms = x_ms = 0.0
    pairs = exchanged = 0
    for _ in range(calls):
        _, k = wl.sweeps(c["wl_f"], average=True)
        ms += k
        # the reference's main loop exchanges after every sweeps call
        partner, ex, xk = wl.replica_exchange()
        x_ms += xk
        pairs += int((partner >= 0).sum()) // 2
        exchanged += int(ex.sum()) // 2
    trials = calls * wl.n_walkers * 8192
    wl.close()
    log(f"[bench] Wang-Landau walkers: {trials} trials in {ms:.1f} ms device; "
        f"{calls} replica exchanges in {x_ms:.2f} ms")
    return {"workload": "16 windows x 8 walkers, bcc 16^3 (8192 sites), S=4, 2 shells, 800 bins; "
                        f"{calls} sweeps calls of 1 sweep after a bounded window entry",
            "walkers": wl.n_walkers, "entered": int(entered.sum()), "trials": trials,
            "value": trials / (ms / 1e3), "unit": "swap trials/s", "device_ms": ms,
            "replica_exchange": {"calls": calls, "ms_per_call": x_ms / calls, "pairs": pairs,
                                 "exchanged": exchanged}}


def wang_landau_run(tolerance: float):
    """Config 4 end to end through WangLandau.run: every walker walked into
    its window, pre-sampling, then sweeps calls of one sweep each with a
    replica exchange after each, halving ln f from 1 at flatness 0.8 until
    it falls below `tolerance`."""
    from brawl_amd import wang_landau, workloads
    c = workloads.config4()
    wl = wang_landau.WangLandau(c["lattice"], c["n"], c["v_ex"], bins=c["bins"],
                                num_windows=c["num_windows"], walkers_per_window=c["walkers_per_window"],
                                bin_overlap=c["bin_overlap"], energy_min=c["energy_min"],
                                energy_max=c["energy_max"], mc_sweeps=1)
    wl.initial_setup(species_numbers=c["counts"])
    t0 = time.perf_counter()
    tr, entered = wl.enter_windows(10**8)
    t1 = time.perf_counter()
    log(f"[bench] Wang-Landau run: window entry {t1 - t0:.1f} s, {int(entered.sum())} walkers in")
    def progress(halvings, ln_f, calls):
        log(f"[bench] Wang-Landau run: {halvings} halvings, ln f {ln_f:.3g}, {calls} sweeps calls, "
            f"{time.perf_counter() - t1:.1f} s")

    res = wl.run(wl_f=c["wl_f"], tolerance=tolerance, flatness=c["flatness"], max_sweeps_calls=200000,
                 max_pre_sampling_calls=50000, enter_max_trials=10**8, progress=progress)
    t2 = time.perf_counter()
    n_atoms = wl.n_atoms
    wl.close()
    log(f"[bench] Wang-Landau run: {res.iterations} halvings, {res.sweeps_calls} sweeps calls, "
        f"{t2 - t1:.1f} s")
    return {"workload": "16 windows x 8 walkers, bcc 16^3 (8192 sites), S=4, 2 shells, 800 bins "
                        "over [-40, 10] meV/atom, ln f from 1, flatness 0.8",
            "tolerance": tolerance, "entry_s": t1 - t0, "entry_max_trials": int(tr.max()),
            "run_s": t2 - t1, "halvings": res.iterations, "sweeps_calls": res.sweeps_calls,
            "pre_sampling_calls": res.pre_sampling_calls, "trials": res.trials,
            "device_ms": res.kernel_ms, "value": res.trials / (t2 - t1), "unit": "swap trials/s",
            "exchanges": res.exchanges, "exchange_pairs": res.exchange_pairs,
            "lng_span": float(res.lng.max() - res.lng.min()),
            "canonical_vs_metropolis": wl_canonical_check(c, res, n_atoms)}


def wl_canonical_check(c, res, n_atoms, chains: int = 8, sweeps: int = 2000):
    """The run's ln g(E) against an independent sampler on the same lattice:
    the canonical <E>/N from ln g at 1800 -> 1000 K in steps of 100 K (where
    the canonical distribution lies inside the windows' energy range; the
    random alloy's energy lies above it), and the parallel Metropolis
    sampler's <E>/N over `chains` independent anneals through the same
    temperatures (`sweeps` burn-in and `sweeps` sampled sweeps at each), with
    the standard error over the chains."""
    from brawl_amd import Lattice, wang_landau, workloads
    from brawl_amd.lattice import RY_TO_EV, anneal_batch
    temps = [1800.0 - 100.0 * k for k in range(9)]
    e_wl, outer = wang_landau.canonical_energy(res.lng, res.bin_edges, n_atoms, temps)
    lats = []
    try:
        for k in range(chains):
            L = Lattice(c["lattice"], c["n"], c["v_ex"])
            L.seed(110179 + 11 * k)
            L.initial_setup(species_numbers=c["counts"])
            lats.append(L)
        params = [dict(T=temps[0], delta_T=-100.0, T_steps=len(temps), n_sweeps=sweeps,
                       burn_in_sweeps=sweeps, seed=workloads.derive_seed(workloads.MASTER_SEED, 4000 + k))
                  for k in range(chains)]
        out, mk = anneal_batch(lats, params)
    finally:
        for L in lats:
            L.close()
    e = np.array([r.energies_of_T for r in out])
    mean, err = e.mean(axis=0), e.std(axis=0, ddof=1) / np.sqrt(chains)
    to_mev = RY_TO_EV * 1000.0           # the bins' conversion (constants.f90:49)
    rows = []
    for i, T in enumerate(temps):
        rows.append({"T": T, "E_wl_meV_per_atom": e_wl[i] * to_mev,
                     "E_mc_meV_per_atom": float(mean[i]) * to_mev,
                     "E_mc_err_meV_per_atom": float(err[i]) * to_mev,
                     "difference_meV_per_atom": (e_wl[i] - float(mean[i])) * to_mev,
                     "wl_weight_in_outer_5pct": float(outer[i])})
        log(f"[bench] {T:.0f} K: <E>/N from ln g {e_wl[i] * to_mev:.4f}, Metropolis "
            f"{mean[i] * to_mev:.4f} +- {err[i] * to_mev:.4f} meV/atom (outer weight {outer[i]:.2g})")
    diffs = np.array([r["difference_meV_per_atom"] for r in rows])
    return {"sampler": f"{chains} parallel-sampler anneals, {sweeps} burn-in + {sweeps} sampled "
                       f"sweeps per temperature, makespan {mk:.0f} ms",
            "bin_width_meV_per_atom": 50.0 / 800, "rows": rows,
            "max_abs_difference_meV_per_atom": float(np.abs(diffs).max())}


def sequential_bit_exact(sweeps: int = 100):
    """Config 1 (fcc 8^3, binary, 2 shells, 600 K) through the sequential
    mode that consumes the reference's MT19937 stream (bwl_mc_steps; its
    trajectory is the reference's bit for bit), `sweeps` sweeps with the
    energy recounted after each, as the reference samples it."""
    from brawl_amd import Lattice, beta_of, workloads
    c = workloads.config1()
    with Lattice(c.lattice, c.n, c.v_ex) as L:
        L.seed(110179)
        L.initial_setup(species_numbers=c.counts)
        N = L.n_atoms
        L.mc_steps(beta_of(c.T), N, sample_every=N)          # warm-up sweep
        t0 = time.perf_counter()
        acc, _ = L.mc_steps(beta_of(c.T), sweeps * N, sample_every=N)
        dt = time.perf_counter() - t0
    log(f"[bench] sequential bit-exact (config 1): {sweeps * N} trials in {dt * 1e3:.1f} ms")
    # the reference's replica ensemble, bit for bit: one chain per rank
    # (static seed 110179 + 11*rank), all ranks in one launch
    from brawl_amd.lattice import mc_steps_batch
    reps, rsweeps = 256, 20
    lats = []
    for k in range(reps):
        Lk = Lattice(c.lattice, c.n, c.v_ex)
        Lk.seed(110179 + 11 * k)
        Lk.initial_setup(species_numbers=c.counts)
        lats.append(Lk)
    betas = [beta_of(c.T)] * reps
    mc_steps_batch(lats, betas, N, sample_every=N)              # warm-up sweep
    t0 = time.perf_counter()
    racc, _, rms = mc_steps_batch(lats, betas, rsweeps * N, sample_every=N)
    rdt = time.perf_counter() - t0
    for Lk in lats:
        Lk.close()
    rtrials = reps * rsweeps * N
    log(f"[bench] sequential bit-exact replicas: {rtrials} trials in {rdt * 1e3:.1f} ms "
        f"({rms:.1f} ms device)")
    return {"workload": f"fcc 8^3 (2048 sites), S=2, 2 shells, 600 K, {sweeps} sweeps, energy "
                        "recounted after every sweep; one wavefront, the reference's MT19937 stream",
            "trials": sweeps * N, "value": sweeps * N / dt, "unit": "swap trials/s",
            "us_per_trial": dt / (sweeps * N) * 1e6, "accepted": acc,
            "replicas": {"chains": reps, "sweeps": rsweeps, "trials": rtrials,
                         "value": rtrials / rdt, "device_rate": rtrials / (rms / 1e3),
                         "unit": "swap trials/s",
                         "note": "one chain per rank, each the reference's trajectory bit for bit"}}


def largest_batch_chains(k: int, sweeps: int = 100):
    """The k chains of the 256-chain batch with the most trials, each run
    alone for `sweeps` sweeps at its own temperature: the chains that set
    the batch's makespan, for timing and profiling them apart."""
    from brawl_amd import Lattice, workloads
    chains = sorted(workloads.config3(), key=lambda c: -workloads.n_atoms_of(c.lattice, c.n))[:k]
    out = []
    for c in chains:
        with Lattice(c.lattice, c.n, c.v_ex) as L:
            L.seed(c.seed & 0xFFFFFFFF)
            L.initial_setup(species_numbers=c.counts)
            r = L.metropolis_simulated_annealing(T=c.T, n_sweeps=sweeps, seed=c.seed)
            out.append({"lattice": "fcc" if c.lattice == 1 else "bcc", "n": list(c.n),
                        "species": len(c.counts), "T": c.T, "threads": L.layout_info("auto")[0],
                        "trials": r.trials, "kernel_ms": r.kernel_ms,
                        "value": r.trials / (r.kernel_ms / 1e3)})
            log(f"[bench] chain {out[-1]['lattice']} {c.n} S={len(c.counts)}: {r.trials} trials in "
                f"{r.kernel_ms:.2f} ms")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--wl-run", type=float, default=None, metavar="TOL",
                    help="only: config 4's whole Wang-Landau run (window entry, pre-sampling, "
                         "sweeps with replica exchanges) down to ln f < TOL")
    ap.add_argument("--only-served", action="store_true",
                    help="run only the served workload (for profiling it)")
    ap.add_argument("--served-requests", type=int, default=240)
    ap.add_argument("--served-rate", type=float, default=400.0)
    ap.add_argument("--served-slots", type=int, default=8)
    ap.add_argument("--hw-queues", type=int, default=8,
                    help="GPU_MAX_HW_QUEUES for this process (give as --hw-queues=N)")
    ap.add_argument("--served-budget-mb", type=int, default=1024)
    ap.add_argument("--served-wait-us", type=int, default=5000)
    ap.add_argument("--served-plain-every", type=int, default=4)
    ap.add_argument("--largest-batch-chains", type=int, default=0, metavar="K",
                    help="only: the K chains of the batched workload with the most trials, each "
                         "alone for 100 sweeps (for profiling them)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.largest_batch_chains > 0:
        print(json.dumps({"largest_batch_chains": largest_batch_chains(args.largest_batch_chains)}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "brawl_amd runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    from brawl_amd import Lattice, workloads
    from brawl_amd import _lib

    if args.only_served:
        sv = served(args)
        sv["saturation"] = served_sweep(args)
        print(json.dumps({"served": sv}))
        return
    if args.wl_run is not None:
        print(json.dumps({"wang_landau_run": wang_landau_run(args.wl_run)}))
        return
    chain = workloads.config2()
    log(f"[bench] workload {chain.name}: N={chain.n_atoms} S={chain.S} R={chain.R}")
    cpu = None
    if not args.no_cpu:
        # before the GPU is touched (the reference baseline forks)
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(chain, args.cpu_budget)
        log(f"[bench] cpu baseline: {cpu['value']:.3e} trials/s ({cpu['kind']}), "
            f"{cpu['replicas']['value']:.3e} with {cpu['replicas']['cores']} chains")
        walkers_cpu = cpu_walkers(args.cpu_budget)
        if walkers_cpu:
            cpu["walkers"] = walkers_cpu
            log(f"[bench] cpu walkers: WL {walkers_cpu['wang_landau']['value']:.3e} / "
                f"{walkers_cpu['wang_landau']['all_cores']['value']:.3e}, NS "
                f"{walkers_cpu['nested_sampling']['value']:.3e} / "
                f"{walkers_cpu['nested_sampling']['all_cores']['value']:.3e} trials/s")
    # start configuration: equiatomic random occupation (reference initial_setup)
    lat = Lattice(chain.lattice, chain.n, chain.v_ex)
    lat.seed(110179)
    start = lat.initial_setup(species_numbers=chain.counts)
    z = {1: [12, 6, 24, 12, 24, 8], 2: [8, 6, 12, 24, 8, 6, 24, 24, 24, 32]}[chain.lattice]
    z = sum(z[: chain.R])
    bpt = bytes_per_trial(z)

    def step(seed):
        lat.set_config(start)
        return lat.metropolis_simulated_annealing(T=chain.T, delta_T=chain.delta_T,
                                                  T_steps=chain.T_steps, n_sweeps=chain.n_sweeps,
                                                  seed=seed, asro_every=ASRO_EVERY)

    for w in range(args.warmup):
        r = step(1000 + w)
        log(f"[bench] warmup {w}: {r.trials} trials, kernel {r.kernel_ms:.1f} ms")
    total_trials = 0
    kernel_ms = []
    results = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        r = step(k + 1)
        total_trials += r.trials
        kernel_ms.append(r.kernel_ms)
        results.append(r)
        log(f"[bench] step {k}: {r.trials} trials, kernel {r.kernel_ms:.1f} ms")
    dt = time.perf_counter() - t0
    value = total_trials / dt
    avg_kernel_s = np.mean(kernel_ms) / 1e3
    trials_per_launch = results[0].trials
    achieved = trials_per_launch * bpt / avg_kernel_s / 1e9
    peak = 8000.0
    last = results[-1]
    out = {
        "metric": "swap trials/s on 1 MI355X (plain vs batched-served); p50/p99 run latency",
        "value": value,
        "unit": "swap trials/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain: quaternary bcc anneal, bcc 32^3 cells (65536 sites), "
                               "S=4 equiatomic, 4 shells, 3000->100 K in 50 steps x 200 sweeps, "
                               "radial densities sampled every 10 sweeps",
                   "form": "plain (one chain per call)", "trials_per_step": trials_per_launch,
                   "bytes_per_trial": bpt},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": peak, "unit": "GB/s",
                     "frac": achieved / peak, "traffic": TRAFFIC_BYTES,
                     "kernel": PLAIN_KERNEL, "kernel_ms": float(np.mean(kernel_ms)),
                     "bytes_per_launch": trials_per_launch * bpt,
                     "limited_by": "VALU and LDS instruction issue of the one CU that holds the "
                                   "chain (the lattice is resident in LDS; HBM carries only "
                                   "the load and the write-back, 'traffic')",
                     "frac_of_one_cu_lds": None},
        "results": {"E_per_atom_first_T": float(last.energies_of_T[0]),
                    "E_per_atom_last_T": float(last.energies_of_T[-1]),
                    "C_max": float(np.max(last.C_of_T)),
                    "acceptance_first_last": [float(last.acceptance_of_T[0]),
                                              float(last.acceptance_of_T[-1])],
                    "alpha_last_T_shells_1_to_4": _alpha_summary(last, chain),
                    "carried_energy_vs_recount": {
                        "max_abs_Ry": max(r.e_drift_abs for r in results),
                        "max_rel": max(r.e_drift_rel for r in results),
                        "note": "the carried energy (recount + accepted dE) against an FP64 "
                                "recount at the end of every T step, worst over the timed steps"}},
    }
    lat.close()
    if not args.no_batched:
        out["batched"] = batched(max(1, args.steps), min(1, args.warmup))
        out["batched"]["vs_plain"] = out["batched"]["value"] / value
    if not args.no_served:
        out["served"] = served(args)
        out["served"]["saturation"] = served_sweep(args)
    if not args.no_batched:
        out["wang_landau_walkers"] = wang_landau_walkers()
        out["nested_sampling_runs"] = nested_sampling_runs()
        out["sequential_bit_exact"] = sequential_bit_exact()
    out["roofline_lds"] = lds_roofline(achieved, out.get("batched"))
    out["roofline"]["frac_of_one_cu_lds"] = out["roofline_lds"]["plain_frac_of_one_cu"]
    out["cpu_baseline"] = cpu
    print(json.dumps(out))


if __name__ == "__main__":
    main()
