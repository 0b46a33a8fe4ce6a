"""Host logic of the layout and thread-count autotuner (no GPU): the
candidates it times per shape, the choice of the cheapest, a thread count a
layout cannot run being skipped, and the table's JSON round trip."""
import json

import pytest

from brawl_amd import autotune
from brawl_amd._lib import BrawlError


class FakeResult:
    def __init__(self, ms):
        self.kernel_ms = ms


class FakeLattice:
    """Stands in for brawl_amd.Lattice: layouts by rule order, the rule's
    thread count per layout, a device time per (layout, threads) that grows
    with the sweeps, and a thread cap per layout."""

    lattice, n, S, R = 1, (8, 8, 8), 3, 3
    v_ex = None

    def __init__(self, rule, cost, cap):
        self.rule, self.cost, self.cap = rule, cost, cap
        self.layout, self.threads = "auto", 0

    def layouts(self):
        return ["fixed", "natural", "compact"]

    def layout_info(self, layout="auto"):
        return self.rule[layout], 1024

    def set_layout(self, layout):
        self.layout = layout

    def set_threads(self, threads):
        self.threads = threads

    def get_config(self):
        return None

    def set_config(self, cfg):
        pass

    def metropolis_simulated_annealing(self, *, T, n_sweeps, seed):
        if self.threads > self.cap[self.layout]:
            raise BrawlError(-3, "no layout of this lattice runs the requested threads per chain")
        return FakeResult(1.0 + n_sweeps * self.cost[(self.layout, self.threads)])

    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


def test_candidates_are_rule_and_listed_counts_up_to_twice_the_rule():
    lat = FakeLattice({"fixed": 384, "natural": 128, "compact": 64}, {}, {})
    c = autotune.candidates(lat)
    assert c == [("fixed", 256), ("fixed", 384), ("fixed", 512), ("fixed", 768),
                 ("natural", 128), ("natural", 256)]


def test_tune_takes_the_cheapest_and_skips_counts_a_layout_cannot_run(monkeypatch):
    rule = {"fixed": 384, "natural": 128, "compact": 64}
    cost = {("fixed", 256): 3.0, ("fixed", 384): 2.0, ("fixed", 512): 1.5, ("fixed", 768): 0.5,
            ("natural", 128): 2.5, ("natural", 256): 1.0}
    cap = {"fixed": 512, "natural": 512, "compact": 512}   # 768 is refused
    proto = FakeLattice(rule, cost, cap)
    monkeypatch.setattr(autotune, "Lattice", lambda lattice, n, v_ex: FakeLattice(rule, cost, cap))
    table = autotune.tune([proto], sweeps=4, base_sweeps=1, repeats=1)
    key = autotune.shape_key(proto)
    assert table.choices[key] == "natural" and table.threads[key] == 256
    costs = table.timings[key]
    assert "fixed/768" not in costs and len(costs) == 5
    # base + sweeps = 5 sweeps less the 1-sweep base: 4 sweeps at 1.0 each
    assert costs["natural/256"] == pytest.approx(4.0)


def test_table_json_round_trip_and_apply():
    t = autotune.LayoutTable({"1:8x8x8:S3:R3": "fixed"}, {"1:8x8x8:S3:R3": 768})
    t.timings = {"1:8x8x8:S3:R3": {"fixed/768": 1.5}}
    u = autotune.LayoutTable.from_json(t.to_json())
    assert u.choices == t.choices and u.threads == t.threads and u.timings == t.timings
    lat = FakeLattice({"fixed": 384, "natural": 128, "compact": 64}, {}, {})
    assert u.apply([lat]) == 1 and lat.layout == "fixed" and lat.threads == 768
    # a table written before thread counts were tuned applies the rule's count
    old = autotune.LayoutTable.from_json(json.dumps({"choices": {"1:8x8x8:S3:R3": "natural"}}))
    lat2 = FakeLattice({"fixed": 384, "natural": 128, "compact": 64}, {}, {})
    assert old.apply([lat2]) == 1 and lat2.layout == "natural" and lat2.threads == 0
