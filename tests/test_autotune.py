"""The autotuner's search and cache (host logic, no device): coordinate descent finds the optimum
of a separable cost, ignores moves inside the noise margin, stops at its budget and caches its
winner by workload key."""
import json

from vecruntime_amd import autotune


def cost_of(c):
    # separable: rows_per_thread 4 and max_load 65 are best; partitioned -1 is 30% slower
    t = 1.0 + 0.1 * abs(c["rows_per_thread"] - 4) + 0.002 * abs(c["max_load_percent"] - 65)
    return t * (1.3 if c["partitioned"] == -1 else 1.0)


def test_descent_finds_the_optimum_and_counts_runs():
    calls = []

    def run(c):
        calls.append(dict(c))
        return cost_of(c)

    best, trials = autotune.descend(run, autotune.SPACES["partitioned"], reps=3)
    assert best["rows_per_thread"] == 4 and best["max_load_percent"] == 65 and best["partitioned"] == 0
    assert trials[0][0] == cost_of(best)
    # every configuration is measured once, `reps` runs each
    assert len(calls) == 3 * len(trials)
    assert len({json.dumps(c, sort_keys=True) for c in calls}) == len(trials)


def test_moves_inside_the_margin_are_ignored():
    def run(c):  # 1% better elsewhere: noise for a 2% margin
        return 0.99 if c["rows_per_thread"] == 2 else 1.0

    best, _ = autotune.descend(run, autotune.SPACES["partitioned"], reps=1)
    assert best == autotune.DEFAULT


def test_budget_stops_the_descent():
    import time

    def run(c):
        time.sleep(0.01)
        return cost_of(c)

    best, trials = autotune.descend(run, autotune.SPACES["persistent"], reps=1, budget_s=0.05)
    assert len(trials) < 10


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tune.json")
    runs = []

    def run(c):
        runs.append(1)
        return cost_of(c)

    best, trials = autotune.tuned("k", run, autotune.SPACES["partitioned"], reps=1, path=path)
    assert trials is not None
    n = len(runs)
    again, trials2 = autotune.tuned("k", run, autotune.SPACES["partitioned"], reps=1, path=path)
    assert again == best and trials2 is None and len(runs) == n


def test_launch_config_carries_every_knob():
    c = dict(autotune.DEFAULT, rows_per_thread=4, partitioned=-1, specialize=-1, lds=0, lds_entries=512, flush_tiles=2,
             blocks_per_cu=8)
    lc = autotune.to_launch_config(c)
    assert (lc.rows_per_thread, lc.blocks_per_cu, lc.lds_preaggregate, lc.lds_entries, lc.flush_tiles, lc.partitioned,
            lc.specialize) == (4, 8, 0, 512, 2, -1, -1)
