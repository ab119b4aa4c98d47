"""Autotuner for the group-by update: the table's load factor, the update path (partitioned
hot-set scatter + region merge, or the persistent per-tile kernel), the partitioned path's rows per
thread and specialization, and the persistent kernel's launch shape and LDS hot-group cache.

A tuning is a coordinate descent: from a starting configuration each knob in turn takes the value
that times best with the others held, for a few passes or until no knob moves; a candidate's time
is the median of several runs on a sample of the real workload, and a move must beat the current
best by more than the noise margin.  The winner is cached as JSON keyed by the device and the
workload's signature, so a later run of the same shape picks it up without tuning again.

    best, trials = descend(run, space, start)       # run(cfg) -> seconds
"""
import json
import os
import time

from . import abi

CACHE_PATH = os.environ.get("VR_AUTOTUNE_CACHE",
                            os.path.join(os.path.expanduser("~"), ".cache", "vecruntime_amd", "autotune.json"))

# The library's defaults (what a configuration of zeros means), the descent's starting point.
DEFAULT = dict(rows_per_thread=0, blocks_per_cu=0, lds=-1, lds_entries=0, flush_tiles=0, partitioned=0, specialize=0,
               max_load_percent=50)

# Knob values per workload kind.  "partitioned": keys of at most 64 bits, inputs large enough for
# the hot-set scatter (partitioned 0 = the library's size rule, -1 = the persistent kernel instead);
# "persistent": wider keys, which take the persistent per-tile kernel.
SPACES = {
    "partitioned": dict(partitioned=[0, -1], rows_per_thread=[0, 1, 2, 4], specialize=[0, -1],
                        max_load_percent=[35, 50, 65]),
    "persistent": dict(rows_per_thread=[0, 1, 2, 4, 8], blocks_per_cu=[0, 2, 8], lds=[-1, 0],
                       lds_entries=[0, 256, 512], flush_tiles=[0, 2, 16], max_load_percent=[35, 50, 65, 80]),
}


def to_launch_config(c):
    return abi.launch_config(c.get("rows_per_thread", 0), c.get("blocks_per_cu", 0), c.get("lds", -1),
                             c.get("lds_entries", 0), c.get("flush_tiles", 0), c.get("partitioned", 0),
                             c.get("specialize", 0))


def signature(device_name, spec, tag=""):
    """A key for the cache: device, key/value lanes, aggregate kinds and a workload tag."""
    return json.dumps([device_name, spec.key_types, spec.value_types, [a.kind for a in spec.aggs], tag])


def load_cache(path=CACHE_PATH):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(cache, path=CACHE_PATH):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def descend(run, space, start=None, reps=3, budget_s=None, passes=2, margin=0.02):
    """Coordinate descent over `space` (knob -> values) from `start` (DEFAULT).  run(cfg) returns
    seconds; a configuration's time is the median of `reps` runs, measured once.  A move needs a
    time below (1 - margin) x the best so far.  Returns (best, [(seconds, cfg)] sorted); `budget_s`
    ends the descent early, keeping the best found."""
    t0 = time.time()
    times = {}

    def measure(c):
        k = json.dumps(c, sort_keys=True)
        if k not in times:
            ts = sorted(run(c) for _ in range(reps))
            times[k] = ts[len(ts) // 2]
        return times[k]

    best = dict(DEFAULT if start is None else start)
    best_t = measure(best)
    for _ in range(passes):
        moved = False
        for knob, values in space.items():
            for v in values:
                if v == best.get(knob):
                    continue
                if budget_s is not None and time.time() - t0 > budget_s:
                    return best, sorted(((t, json.loads(k)) for k, t in times.items()), key=lambda x: x[0])
                c = dict(best)
                c[knob] = v
                t = measure(c)
                if t < best_t * (1.0 - margin):
                    best, best_t, moved = c, t, True
        if not moved:
            break
    return best, sorted(((t, json.loads(k)) for k, t in times.items()), key=lambda x: x[0])


def tuned(key, run, space, start=None, reps=3, budget_s=None, path=CACHE_PATH):
    """The cached winner for `key`, or a fresh descent (cached afterwards): (best, trials or None)."""
    cache = load_cache(path)
    if key in cache:
        return cache[key]["config"], None
    best, trials = descend(run, space, start, reps, budget_s)
    cache[key] = {"config": best, "seconds": trials[0][0], "tried": len(trials)}
    save_cache(cache, path)
    return best, trials
