"""Autotuner for the group-by update: launch shape (rows per thread, workgroups per CU), the LDS
hot-group cache (entries, flush interval, on/off) and the table's load factor.

A tuning is a sweep over a small candidate grid, each candidate timed on a sample of the real
workload; the winner is cached as JSON keyed by the device and the workload's signature, so a
later run with the same shape picks it up without sweeping again.

    best, table = tune(run, candidates(), reps=2)    # run(cfg) -> seconds
"""
import itertools
import json
import os
import time

from . import abi

CACHE_PATH = os.environ.get("VR_AUTOTUNE_CACHE",
                            os.path.join(os.path.expanduser("~"), ".cache", "vecruntime_amd", "autotune.json"))


def candidates(rows_per_thread=(1, 2, 4, 8), blocks_per_cu=(2, 4, 8), lds_entries=(256, 512, 1024),
               flush_tiles=(2, 4, 16), max_load_percent=(50,)):
    """The candidate grid: every combination, plus the cache switched off."""
    out = []
    for r, b, e, f, m in itertools.product(rows_per_thread, blocks_per_cu, lds_entries, flush_tiles, max_load_percent):
        out.append(dict(rows_per_thread=r, blocks_per_cu=b, lds=1, lds_entries=e, flush_tiles=f, max_load_percent=m))
    for r, b, m in itertools.product(rows_per_thread, blocks_per_cu, max_load_percent):
        out.append(dict(rows_per_thread=r, blocks_per_cu=b, lds=0, lds_entries=0, flush_tiles=0, max_load_percent=m))
    return out


def to_launch_config(c):
    return abi.launch_config(c["rows_per_thread"], c["blocks_per_cu"], c["lds"], c["lds_entries"], c["flush_tiles"])


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


def tune(run, cands, reps=2, budget_s=None):
    """Times run(candidate) (seconds, the best of `reps`) for each candidate; returns (best
    candidate, [(seconds, candidate)] sorted).  `budget_s` stops the sweep early."""
    t0 = time.time()
    results = []
    for c in cands:
        best = min(run(c) for _ in range(reps))
        results.append((best, c))
        if budget_s is not None and time.time() - t0 > budget_s:
            break
    results.sort(key=lambda x: x[0])
    return results[0][1], results


def tuned(key, run, cands, reps=2, budget_s=None, path=CACHE_PATH):
    """The cached winner for `key`, or a fresh sweep (cached afterwards)."""
    cache = load_cache(path)
    if key in cache:
        return cache[key]["config"], None
    best, results = tune(run, cands, reps, budget_s)
    cache[key] = {"config": best, "seconds": results[0][0], "swept": len(results)}
    save_cache(cache, path)
    return best, results
