"""Autotuner for the parallel sampler's lattice layout in LDS and its
threads per chain (the tile and domain shape of the batched anneal).

The layout candidates for a lattice shape are the two layouts in lattice
coordinates: "natural" (this size's own row pitch, or the kernel compiled
for this exact shape, at most 512 threads per chain unless the shape has a
kernel of its own) and "fixed" (a row pitch shared by every size of the
lattice type and shell count, neighbour offsets compiled in, up to 1024
threads).  The per-sublattice "compact" layout is never faster and is not
a candidate.  For each layout the thread counts tried are the rule's (the
fewest threads that need the fewest passes over a round's pairs) and the
counts of THREAD_CANDIDATES up to twice the rule's that the layout can run:
fewer passes per round against more wavefronts to hide each other's
latency.  Each candidate anneals a scratch copy of the lattice for a few
sweeps; the smallest device time per sweep wins.

A chain's random numbers are keyed by its thread and step, so the two
layouts give the same chain when they run it with the same thread count and
different, equally exact chains otherwise: the table of choices is part of
a run's definition.  Choices are kept per shape (lattice type, cells,
species, shells) in a LayoutTable, which saves and loads as JSON, so a run
that loads a table makes no timing decisions and gives the same chains
every time, alone or batched.
"""
from __future__ import annotations

import json

from ._lib import BrawlError
from .lattice import Lattice

TUNED_LAYOUTS = ("natural", "fixed")
THREAD_CANDIDATES = (256, 512, 768, 1024)


def shape_key(lat: Lattice) -> str:
    return f"{lat.lattice}:{lat.n[0]}x{lat.n[1]}x{lat.n[2]}:S{lat.S}:R{lat.R}"


def candidate_key(layout: str, threads: int) -> str:
    return f"{layout}/{threads}"


class LayoutTable:
    """Layout and thread-count choice per lattice shape, with the timings
    behind each (device ms per `sweeps` sweeps, by candidate_key)."""

    def __init__(self, choices=None, threads=None):
        self.choices = dict(choices or {})
        self.threads = dict(threads or {})
        self.timings = {}

    def get(self, lat: Lattice):
        """The layout chosen for this lattice's shape (None if untuned)."""
        return self.choices.get(shape_key(lat))

    def get_threads(self, lat: Lattice) -> int:
        """The threads per chain chosen for this shape (0: the rule's)."""
        return int(self.threads.get(shape_key(lat), 0))

    def apply(self, lattices) -> int:
        """Set the tuned layout and thread count on every lattice whose shape
        is in the table; returns how many were set."""
        k = 0
        for lat in lattices:
            c = self.get(lat)
            if c is not None:
                lat.set_layout(c)
                lat.set_threads(self.get_threads(lat))
                k += 1
        return k

    def to_json(self) -> str:
        return json.dumps({"choices": self.choices, "threads": self.threads,
                           "timings_ms": self.timings}, sort_keys=True)

    @classmethod
    def from_json(cls, text: str) -> "LayoutTable":
        d = json.loads(text)
        t = cls(d.get("choices", {}), d.get("threads", {}))
        t.timings = d.get("timings_ms", {})
        return t


def candidates(lat: Lattice):
    """(layout, threads) pairs to time for this lattice: per tuned layout
    the rule's thread count and the THREAD_CANDIDATES up to twice it."""
    out = []
    for layout in lat.layouts():
        if layout not in TUNED_LAYOUTS:
            continue
        rule, _ = lat.layout_info(layout)
        ts = sorted({rule} | {t for t in THREAD_CANDIDATES if t <= 2 * rule})
        out.extend((layout, t) for t in ts)
    return out


def tune(lattices, *, T=1000.0, sweeps=16, base_sweeps=4, repeats=2, seed=1,
         table=None) -> LayoutTable:
    """Time the candidates (layout, threads) for every distinct shape among
    `lattices` (each with a configuration uploaded) and record the fastest.
    The cost of a candidate is the device time of base_sweeps + sweeps
    sweeps less that of base_sweeps (best of `repeats` after a warm-up
    each): the sweeps' own cost, without the staging and recount every
    launch pays once, as in a long anneal.  A thread count that the layout
    cannot run is skipped.  Shapes already in `table` are skipped; a shape
    with one candidate records it untimed."""
    table = table if table is not None else LayoutTable()
    seen = set()
    for lat in lattices:
        key = shape_key(lat)
        if key in table.choices or key in seen:
            continue
        seen.add(key)
        cands = candidates(lat)
        if len(cands) < 2:
            if cands:
                table.choices[key], table.threads[key] = cands[0]
            continue
        cfg = lat.get_config()
        best, costs = None, {}
        for layout, threads in cands:
            with Lattice(lat.lattice, lat.n, lat.v_ex) as s:
                s.set_layout(layout)
                s.set_threads(threads)
                t = {}
                try:
                    for n_sw in (base_sweeps, base_sweeps + sweeps):
                        ms = []
                        for _ in range(repeats + 1):          # the first run warms up
                            s.set_config(cfg)
                            r = s.metropolis_simulated_annealing(T=T, n_sweeps=n_sw, seed=seed)
                            ms.append(r.kernel_ms)
                        t[n_sw] = min(ms[1:])
                except BrawlError:
                    continue                                  # the layout cannot run that many
            c = candidate_key(layout, threads)
            costs[c] = t[base_sweeps + sweeps] - t[base_sweeps]
            if best is None or costs[c] < costs[candidate_key(*best)]:
                best = (layout, threads)
        if best is None:
            continue
        table.choices[key], table.threads[key] = best
        table.timings[key] = costs
    return table


__all__ = ["LayoutTable", "tune", "shape_key", "candidates", "candidate_key", "TUNED_LAYOUTS",
           "THREAD_CANDIDATES"]
