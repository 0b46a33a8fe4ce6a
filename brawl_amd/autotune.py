"""Autotuner for the parallel sampler's lattice layout in LDS (the tile and
domain shape of the batched anneal).

The candidates for a lattice shape are the two layouts in lattice
coordinates: "natural" (this size's own row pitch, or the kernel compiled
for this exact shape, at most 512 threads per chain) and "fixed" (a row
pitch shared by every size of the lattice type and shell count, neighbour
offsets compiled in, up to 768 threads).  The per-sublattice "compact"
layout is never faster and is not a candidate.  Each candidate anneals a
scratch copy of the lattice for a few sweeps; the smallest device time
per sweep wins.  A chain's random numbers are keyed by its thread and step, so the
two layouts give the same chain when they run it with the same thread
count (rounds of at most 512 pairs) and different, equally valid chains
otherwise: the table of choices is part of a run's definition.  Choices are
kept per shape (lattice type, cells, species, shells) in a LayoutTable,
which saves and loads as JSON, so a run that loads a table makes no timing
decisions and gives the same chains every time, alone or batched.
"""
from __future__ import annotations

import json

from .lattice import Lattice

TUNED_LAYOUTS = ("natural", "fixed")


def shape_key(lat: Lattice) -> str:
    return f"{lat.lattice}:{lat.n[0]}x{lat.n[1]}x{lat.n[2]}:S{lat.S}:R{lat.R}"


class LayoutTable:
    """Layout choice per lattice shape, with the timings behind each."""

    def __init__(self, choices=None):
        self.choices = dict(choices or {})
        self.timings = {}

    def get(self, lat: Lattice):
        return self.choices.get(shape_key(lat))

    def apply(self, lattices) -> int:
        """Set the tuned layout on every lattice whose shape is in the table;
        returns how many were set."""
        k = 0
        for lat in lattices:
            c = self.get(lat)
            if c is not None:
                lat.set_layout(c)
                k += 1
        return k

    def to_json(self) -> str:
        return json.dumps({"choices": self.choices, "timings_ms": self.timings}, sort_keys=True)

    @classmethod
    def from_json(cls, text: str) -> "LayoutTable":
        d = json.loads(text)
        t = cls(d.get("choices", {}))
        t.timings = d.get("timings_ms", {})
        return t


def tune(lattices, *, T=1000.0, sweeps=16, base_sweeps=4, repeats=2, seed=1,
         table=None) -> LayoutTable:
    """Time the candidate layouts for every distinct shape among `lattices`
    (each with a configuration uploaded) and record the fastest.  The cost
    of a layout is the device time of base_sweeps + sweeps sweeps less that
    of base_sweeps (best of `repeats` after a warm-up each): the sweeps'
    own cost, without the staging and recount every launch pays once, as in
    a long anneal.  Shapes already in `table` are skipped; a shape with one
    candidate records it untimed."""
    table = table if table is not None else LayoutTable()
    seen = set()
    for lat in lattices:
        key = shape_key(lat)
        if key in table.choices or key in seen:
            continue
        seen.add(key)
        cands = [c for c in lat.layouts() if c in TUNED_LAYOUTS]
        if len(cands) < 2:
            if cands:
                table.choices[key] = cands[0]
            continue
        cfg = lat.get_config()
        best, costs = None, {}
        for c in cands:
            with Lattice(lat.lattice, lat.n, lat.v_ex) as s:
                s.set_layout(c)
                t = {}
                for n_sw in (base_sweeps, base_sweeps + sweeps):
                    ms = []
                    for _ in range(repeats + 1):          # the first run warms up
                        s.set_config(cfg)
                        r = s.metropolis_simulated_annealing(T=T, n_sweeps=n_sw, seed=seed)
                        ms.append(r.kernel_ms)
                    t[n_sw] = min(ms[1:])
                costs[c] = t[base_sweeps + sweeps] - t[base_sweeps]
            if best is None or costs[c] < costs[best]:
                best = c
        table.choices[key] = best
        table.timings[key] = costs
    return table


__all__ = ["LayoutTable", "tune", "shape_key", "TUNED_LAYOUTS"]
