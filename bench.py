"""Benchmark: group-by aggregation (and join probe) rows/s on one MI355X.

Headline workload (the large group-by): 200M rows in 65,536-row Arrow batches, int64 key Zipf(1.1)
over 10M ranks (scrambled), SUM and AVG of decimal(12,2) (unscaled int64 uniform [0, 1e8)), AVG of a
lognormal(0,1) double, MIN/MAX of a uniform int32, COUNT(*); each value column 5% null.  Inputs are
synthetic, generated on the device and resident in HBM before timing.  A step is one pass over the
whole input: the table cleared, every batch folded in, every group emitted.

`value` is the batched path (one persistent launch for all batches); the plain per-batch path (one
launch per batch, the drop-in) and the join probe are reported beside it.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--skip-join] [--skip-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BATCH = 65536
# the uneven batched workload's columns: filter int32, keys int32 + int64 + int32 dictionary code,
# decimal(15,2) as int64, double (every row's line is read at 30% selectivity)
UNEVEN_BYTES_PER_ROW = 4 + 4 + 8 + 4 + 8 + 8
# HBM bytes moved by one partitioned update of the 200M-row group-by (sampling, count pass, scans,
# scatter, hot reduce, merge), from rocprofv3 FETCH_SIZE / WRITE_SIZE runs scaled by the same
# counters on a 1 GiB device copy (profiles/partitioned_update_groupby200M_pmc_traffic.md).
# Measured for the kernels as committed with it.
PMC_TRAFFIC_BYTES = {200_000_000: 12.93e9}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pack_bits_torch(valid_bool):
    """LSB-first bitmap (Arrow) of a bool tensor whose length is a multiple of 8."""
    import torch
    w = torch.tensor([1, 2, 4, 8, 16, 32, 64, 128], dtype=torch.uint8, device=valid_bool.device)
    return (valid_bool.view(-1, 8).to(torch.uint8) * w).sum(dim=1, dtype=torch.uint8)


def zipf_ranks(gen, n, distinct, s, chunk=1 << 25):
    """Zipf(s) ranks in [0, distinct) by inverse CDF over the exact probabilities."""
    import torch
    r = torch.arange(1, distinct + 1, dtype=torch.float64, device="cuda")
    cdf = torch.cumsum(r.pow(-s), 0)
    cdf /= cdf[-1].clone()
    out = torch.empty(n, dtype=torch.int64, device="cuda")
    for s0 in range(0, n, chunk):
        m = min(chunk, n - s0)
        u = torch.rand(m, dtype=torch.float64, device="cuda", generator=gen)
        out[s0:s0 + m] = torch.searchsorted(cdf, u).clamp_(max=distinct - 1)
    return out


def scramble(x):
    """A bijection of int64 (odd multiplier, then an add): hot ranks do not cluster."""
    return x * -7046029254386353131 + 0x5851F42D4C957F2D


def gen_groupby(n, seed=102):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    keys = scramble(zipf_ranks(g, n, 10_000_000, 1.1))
    dec = torch.randint(0, 10 ** 8, (n,), dtype=torch.int64, device="cuda", generator=g)
    dbl = torch.empty(n, dtype=torch.float64, device="cuda").log_normal_(0.0, 1.0, generator=g)
    i32 = torch.randint(-2 ** 31, 2 ** 31, (n,), dtype=torch.int32, device="cuda", generator=g)
    valid = [pack_bits_torch(torch.rand(n, device="cuda", generator=g) >= 0.05) for _ in range(3)]
    torch.cuda.synchronize()
    return keys, dec, dbl, i32, valid


def groupby_spec(capacity_hint, max_load_percent=0):
    from vecruntime_amd import abi
    return abi.AggSpec(
        [abi.INT64], [abi.INT64, abi.FLOAT64, abi.INT32],
        [abi.agg_desc(abi.SUM_WIDE, 0, buffer_precision=22, buffer_scale=2),
         abi.agg_desc(abi.AVG_WIDE, 0, buffer_precision=22, buffer_scale=2),
         abi.agg_desc(abi.AVG_DOUBLE, 1), abi.agg_desc(abi.MIN, 2), abi.agg_desc(abi.MAX, 2),
         abi.agg_desc(abi.COUNT_STAR)],
        key_nullable=[False], capacity_hint=capacity_hint, max_load_percent=max_load_percent)


def groupby_batches(keys, dec, dbl, i32, valid, n):
    from vecruntime_amd import abi, vr
    batches = []
    for s0 in range(0, n, BATCH):
        m = min(BATCH, n - s0)
        vb = [v[s0 // 8:] for v in valid]
        batches.append(([vr.device_column(abi.INT64, keys[s0:s0 + m])],
                        [vr.device_column(abi.INT64, dec[s0:s0 + m], vb[0], length=m),
                         vr.device_column(abi.FLOAT64, dbl[s0:s0 + m], vb[1], length=m),
                         vr.device_column(abi.INT32, i32[s0:s0 + m], vb[2], length=m)], m))
    return batches


def emit_buffers(agg, groups):
    """Device output columns for every group (keys + the aggregate buffers)."""
    import torch
    from vecruntime_amd import abi, vr
    dt = {abi.INT64: torch.int64, abi.FLOAT64: torch.float64, abi.INT32: torch.int32}
    nb = (groups + 63) // 64 * 8
    keys = [vr.device_out_column(abi.INT64, groups, torch.empty(groups, dtype=torch.int64, device="cuda"),
                                 torch.empty(nb, dtype=torch.uint8, device="cuda"))]
    outs = []
    for t in agg.output_types():
        if t == abi.DECIMAL128:
            data = torch.empty(2 * groups, dtype=torch.int64, device="cuda")
        elif t == abi.BOOL:
            data = torch.empty(nb, dtype=torch.uint8, device="cuda")
        else:
            data = torch.empty(groups, dtype=dt[t], device="cuda")
        outs.append(vr.device_out_column(t, groups, data, torch.empty(nb, dtype=torch.uint8, device="cuda")))
    return keys, outs


def unpack_bits_torch(bm, n):
    import torch
    bits = (bm.view(-1, 1) >> torch.arange(8, device=bm.device, dtype=torch.uint8)) & 1
    return bits.view(-1)[:n].bool()


def verify_groupby(keys, dec, dbl, i32, valid, n, groups, ek, eo):
    """Size-independent properties of the full-size result against the input, computed by torch:
    distinct keys, total rows, exact decimal total and non-null count, global MIN/MAX, the double
    total.  Outputs: [SUM dec (d128), isEmpty, AVG dec sum (d128), AVG dec count, AVG f64 sum,
    AVG f64 count, MIN i32, MAX i32, COUNT(*)]."""
    import torch
    v = [unpack_bits_torch(b, n) for b in valid]
    out = {}
    out["distinct_keys"] = int(torch.unique(keys).numel()) == groups
    cnt = eo[8].data[:groups]
    out["count_star_total"] = int(cnt.sum().item()) == n
    dsum = eo[0].data[:2 * groups].view(groups, 2)
    dok = unpack_bits_torch(eo[0].validity, groups)
    # every per-group sum fits 64 bits here (< 2^63): compare the low limbs' total exactly
    exact_total = int(dec[v[0]].sum().item())
    out["decimal_sum_total"] = bool((dsum[:, 1][dok] == 0).all().item()) and int(dsum[:, 0][dok].sum().item()) == exact_total
    out["decimal_count_total"] = int(eo[3].data[:groups].sum().item()) == int(v[0].sum().item())
    out["f64_count_total"] = int(eo[5].data[:groups].sum().item()) == int(v[1].sum().item())
    fs = eo[4].data[:groups].sum().item()
    ft = dbl[v[1]].sum().item()
    out["f64_sum_rel_err"] = abs(fs - ft) / abs(ft)
    mn_ok = unpack_bits_torch(eo[6].validity, groups)
    mx_ok = unpack_bits_torch(eo[7].validity, groups)
    out["min_i32"] = int(eo[6].data[:groups][mn_ok].min().item()) == int(i32[v[2]].min().item())
    out["max_i32"] = int(eo[7].data[:groups][mx_ok].max().item()) == int(i32[v[2]].max().item())
    out["ok"] = all(x for k, x in out.items() if k != "f64_sum_rel_err") and out["f64_sum_rel_err"] < 1e-9
    return out


def tune_groupby(dev, args, batches, stream):
    """The autotuner's descent over the partitioned path's knobs and the load factor, each
    configuration timed (median of 3) on what a step does to the table -- its reset and the update
    of every batch (the load factor sets the table's size, which the reset pays for) -- outside the
    timed steps; cached per device and workload."""
    import torch
    from vecruntime_amd import autotune, vr
    sample = vr.GroupedAggregate.prepare_batches(batches)
    srows = sum(b[2] for b in batches)
    handles = {}

    def run(c):
        m = c["max_load_percent"]
        if m not in handles:
            handles[m] = vr.GroupedAggregate(dev, groupby_spec(0, m))
            handles[m].update_batches(sample, stream=stream)  # grows the table to the workload's groups
        h = handles[m]
        torch.cuda.synchronize()
        t = time.perf_counter()
        h.reset()
        h.update_batches(sample, config=autotune.to_launch_config(c), stream=stream)
        torch.cuda.synchronize()
        return time.perf_counter() - t

    key = autotune.signature(torch.cuda.get_device_properties(0).name, groupby_spec(0), "groupby-zipf1.1-10M-step")
    t0 = time.time()
    best, trials = autotune.tuned(key, run, autotune.SPACES["partitioned"], reps=3, budget_s=args.tune_budget)
    for h in handles.values():
        h.close()
    info = {"sample_rows": srows, "seconds": time.time() - t0, "configurations": len(trials) if trials else None}
    if trials:
        info["sample_ms_best"] = trials[0][0] * 1e3
        info["sample_ms_default"] = next((t * 1e3 for t, c in trials if c == autotune.DEFAULT), None)
    log("autotune (group-by): %s, %s" % (best, json.dumps(info)))
    return best, info


def run_groupby(dev, args, n):
    import torch
    from vecruntime_amd import abi, autotune, vr
    log("generating %d rows of the group-by input on the device" % n)
    t0 = time.time()
    keys, dec, dbl, i32, valid = gen_groupby(n)
    batches = groupby_batches(keys, dec, dbl, i32, valid, n)
    log("generated in %.1f s: %d batches" % (time.time() - t0, len(batches)))
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    best, tuning = tune_groupby(dev, args, batches, stream)
    tuned_cfg = autotune.to_launch_config(best)
    agg = vr.GroupedAggregate(dev, groupby_spec(0, best["max_load_percent"]))

    batch_set = vr.GroupedAggregate.prepare_batches(batches)  # as the plain path's prepare(), once

    def step_batched(emit_cols, h=agg, cfg=tuned_cfg):
        h.reset()
        h.update_batches(batch_set, config=cfg, stream=stream)
        h.emit(0, emit_cols[2], emit_cols[0], emit_cols[1], stream=stream)

    prepared = [(vr.GroupedAggregate.prepare(k, v), m) for k, v, m in batches]

    def step_plain(emit_cols):
        agg.reset()
        for prep, m in prepared:
            agg.update_prepared(prep, m, stream=stream)
        agg.emit(0, emit_cols[2], emit_cols[0], emit_cols[1], stream=stream)

    agg.update_batches(batch_set, config=tuned_cfg, stream=stream)
    groups = agg.num_groups()
    ek, eo = emit_buffers(agg, groups)
    cols = (ek, eo, groups)
    log("groups: %d" % groups)
    agg.emit(0, groups, ek, eo, stream=stream)
    checks = verify_groupby(keys, dec, dbl, i32, valid, n, groups, ek, eo)
    log("checks: %s" % json.dumps(checks))
    res = {}
    legs = [("batched", step_batched, args.steps)]
    if not args.skip_plain:
        legs.append(("plain", step_plain, max(1, args.steps // 2)))
    for name, fn, steps in legs:
        for _ in range(args.warmup):
            fn(cols)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn(cols)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        res[name] = dict(ms_per_step=dt * 1e3, rows_per_s=n / dt, steps=steps)
        log("%s: %.2f ms/step, %.3e rows/s" % (name, dt * 1e3, n / dt))
    # the library's defaults beside the tuned configuration (a handle of the default load factor)
    res["tuning"] = dict(tuning, tuned=best)
    if best != autotune.DEFAULT:
        dh = agg if best["max_load_percent"] == 50 else vr.GroupedAggregate(dev, groupby_spec(0))
        dcfg = abi.launch_config()
        step_batched(cols, dh, dcfg)  # warm (sizes the table)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step_batched(cols, dh, dcfg)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        res["tuning"]["default_ms_per_step"] = dt * 1e3
        if dh is not agg:
            dh.close()
        log("default configuration: %.2f ms/step" % (dt * 1e3))
    res["tuning"]["tuned_ms_per_step"] = res["batched"]["ms_per_step"]
    # a cold table: a fresh handle with no capacity hint (2,048 slots at first, grown as groups
    # arrive), created, updated and emitted inside the timed region
    cold = []
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fresh = vr.GroupedAggregate(dev, groupby_spec(0))
        fresh.update_batches(batch_set, stream=stream)
        fresh.emit(0, groups, ek, eo, stream=stream)
        torch.cuda.synchronize()
        cold.append(time.perf_counter() - t0)
        fresh.close()
    res["cold"] = {"ms": min(cold) * 1e3, "rows_per_s": n / min(cold)}
    log("cold table: %.2f ms/step" % (min(cold) * 1e3))
    # the dominant kernel, timed with HIP events the library records around each of its launches
    reps = max(1, args.steps)
    agg.set_timing(True)
    for _ in range(reps):
        step_batched(cols)
    torch.cuda.synchronize()
    kms, kl = agg.kernel_time()
    agg.set_timing(False)
    res["update_kernel_ms"] = kms / reps
    res["update_launches_per_step"] = kl / reps
    res["update_kernel_avg_launch_ms"] = kms / max(1, kl)
    log("update kernel: %.3f ms/step over %.1f launches/step" % (kms / reps, kl / reps))
    res["groups"] = groups
    res["checks"] = checks
    # a bounded sample copied to the host for the CPU baseline
    res["_sample"] = None
    res["_host"] = None
    if not args.skip_cpu or not args.skip_pcie:
        host = dict(keys=keys.cpu().numpy(), dec=dec.cpu().numpy(), dbl=dbl.cpu().numpy(),
                    i32=i32.cpu().numpy(), valid=[v.cpu().numpy() for v in valid])
        res["_host"] = host
        if not args.skip_cpu:  # whole batches
            res["_sample"] = (host, min(n // BATCH, args.cpu_batches))
    agg.close()
    del batches, keys, dec, dbl, i32, valid
    torch.cuda.empty_cache()
    return res


def host_groupby_batches(host, pinned=False):
    """The group-by input as host-resident Arrow batches (the reference's off-heap buffers):
    pageable numpy arrays, or page-locked copies when `pinned`."""
    import torch
    from vecruntime_amd import abi
    from vecruntime_amd.columns import Column
    src = host
    if pinned:
        def pin(a):
            t = torch.empty(a.shape, dtype=getattr(torch, a.dtype.name), pin_memory=True)
            t.numpy()[...] = a
            return t.numpy()
        src = dict(keys=pin(host["keys"]), dec=pin(host["dec"]), dbl=pin(host["dbl"]), i32=pin(host["i32"]),
                   valid=[pin(v) for v in host["valid"]])
    out = []
    n = len(src["keys"])
    for s0 in range(0, n, BATCH):
        m = min(BATCH, n - s0)
        vb = [v[s0 // 8:] for v in src["valid"]]
        out.append(([Column(abi.INT64, src["keys"][s0:s0 + m])],
                    [Column(abi.INT64, src["dec"][s0:s0 + m], vb[0], length=m),
                     Column(abi.FLOAT64, src["dbl"][s0:s0 + m], vb[1], length=m),
                     Column(abi.INT32, src["i32"][s0:s0 + m], vb[2], length=m)], m))
    return out, src


def run_pcie(dev, args, host, groups):
    """End to end from host memory: the same batches fed from host-resident buffers, each staged
    to the device by the library (pinned staging buffers, copies on its side stream) and folded by
    the per-batch update.  Reports rows/s and the PCIe bytes/s this sustains."""
    import torch
    from vecruntime_amd import vr
    n = len(host["keys"])
    bytes_in = n * (8 + 8 + 8 + 4 + 3 / 8.0)
    res = {"rows": n, "host_bytes": bytes_in}
    agg = vr.GroupedAggregate(dev, groupby_spec(0))
    for kind in ("pageable", "pinned"):
        batches, keep = host_groupby_batches(host, pinned=kind == "pinned")
        agg.reset()
        agg.update_batches(batches[:64])  # warm-up: staging buffers sized, table touched
        torch.cuda.synchronize()
        times = []
        for _ in range(2):
            agg.reset()
            t0 = time.perf_counter()
            agg.update_batches(batches)
            g = agg.num_groups()
            times.append(time.perf_counter() - t0)
        dt = min(times)
        res[kind] = {"ms": dt * 1e3, "rows_per_s": n / dt, "host_GB_per_s": bytes_in / dt / 1e9,
                     "groups_match_device_resident": g == groups}
        log("pcie %s: %.1f ms, %.3e rows/s, %.1f GB/s from host" % (kind, dt * 1e3, n / dt, bytes_in / dt / 1e9))
        del batches, keep
    agg.close()
    return res


def run_cfg1(dev, args):
    """The reference's CPU-path case: one Arrow batch of 1,000,000 rows, int64 key uniform over
    1,000 values, COUNT(*), SUM(int64 uniform [-1e6, 1e6]), SUM(float64 uniform [0, 1)), no nulls.
    Timed device-resident and from host memory; checked against per-key sums computed by torch;
    the CPU port of the reference kernels timed on the same batch."""
    import torch
    from vecruntime_amd import abi, vr
    from vecruntime_amd.columns import Column
    n = 1_000_000
    rng = np.random.default_rng(101)
    keys = rng.integers(0, 1000, n).astype(np.int64)
    lv = rng.integers(-10 ** 6, 10 ** 6 + 1, n).astype(np.int64)
    dv = rng.random(n)
    spec = abi.AggSpec([abi.INT64], [abi.INT64, abi.FLOAT64],
                       [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_LONG, 0), abi.agg_desc(abi.SUM_DOUBLE, 1)],
                       key_nullable=[False])
    tk, tl, td = (torch.from_numpy(a).cuda() for a in (keys, lv, dv))
    dcols = ([vr.device_column(abi.INT64, tk)], [vr.device_column(abi.INT64, tl), vr.device_column(abi.FLOAT64, td)])
    hcols = ([Column.longs(keys)], [Column.longs(lv), Column.doubles(dv)])
    agg = vr.GroupedAggregate(dev, spec)
    res = {"rows": n, "groups": 1000}
    for kind, (kc, vc) in (("device_resident", dcols), ("from_host", hcols)):
        for _ in range(3):
            agg.reset()
            agg.update(kc, vc, n)
            agg.num_groups()
        reps = 20
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            agg.reset()
            agg.update(kc, vc, n)
            g = agg.num_groups()
        dt = (time.perf_counter() - t0) / reps
        res[kind] = {"ms": dt * 1e3, "rows_per_s": n / dt}
    # exact check of every group against torch
    ek, eo = agg.emit()
    G = agg.num_groups()
    k = torch.from_numpy(ek[0].data[:G].copy())
    cnt = torch.bincount(torch.from_numpy(keys), minlength=1000)
    ls = torch.zeros(1000, dtype=torch.int64).index_add_(0, torch.from_numpy(keys), torch.from_numpy(lv))
    ds = torch.zeros(1000, dtype=torch.float64).index_add_(0, torch.from_numpy(keys), torch.from_numpy(dv))
    ok = G == 1000 and bool((torch.from_numpy(eo[0].data[:G].copy()) == cnt[k]).all())
    ok = ok and bool((torch.from_numpy(eo[1].data[:G].copy()) == ls[k]).all())
    rel = float(((torch.from_numpy(eo[2].data[:G].copy()) - ds[k]).abs() / ds[k].abs()).max())
    res["checks"] = {"count_and_long_sum_exact": ok, "double_sum_max_rel_err": rel, "ok": ok and rel < 1e-12}
    agg.close()
    if not args.skip_cpu:
        from oracle.oracle import OracleAgg  # the CPU port: the baseline, timed on the same batch
        reps, t0 = 0, time.perf_counter()
        while reps < 3 or time.perf_counter() - t0 < 2.0:
            o = OracleAgg(spec)
            o.update(hcols[0], hcols[1], n)
            del o
            reps += 1
        dt = (time.perf_counter() - t0) / reps
        res["cpu_baseline"] = {"value": n / dt, "unit": "rows/s", "cores": 1, "kind": "port", **host_info(),
                               "sample": "the whole 1,000,000-row batch, one thread, %d runs" % reps}
    log("cfg1: %s" % json.dumps(res))
    return res


def gen_uneven(seed=104, scale=1.0):
    """64 partitions x U{1..200} batches of round(log-uniform[1e3, 1e6]) rows, interleaved across
    partitions; columns f (filter int32), k1 int32 1..1000, k2 int64 Zipf(1.2) over 1M (scrambled),
    k3 dictionary codes over 10,000 strings of 8-24 ASCII chars, dec(15,2) unscaled
    [-1e13, 1e13), f64 uniform [-1e6, 1e6)."""
    import torch
    from vecruntime_amd import abi
    from vecruntime_amd.columns import Column
    rng = np.random.default_rng(seed)
    per_part = rng.integers(1, 201, 64)
    sizes = [[int(round(np.exp(rng.uniform(np.log(1e3), np.log(1e6)) ))) for _ in range(nb)] for nb in per_part]
    if scale != 1.0:
        sizes = [[max(1, int(x * scale)) for x in p] for p in sizes]
    order = []
    for j in range(int(per_part.max())):
        for p in range(64):
            if j < len(sizes[p]):
                order.append(sizes[p][j])
    total = int(sum(order))
    words = ["".join(chr(c) for c in rng.integers(33, 127, int(rng.integers(8, 25)))) for _ in range(10_000)]
    dictionary = Column.strings(words)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    f = torch.randint(0, 2 ** 31, (total,), dtype=torch.int32, device="cuda", generator=g)
    k1 = torch.randint(1, 1001, (total,), dtype=torch.int32, device="cuda", generator=g)
    k2 = scramble(zipf_ranks(g, total, 1_000_000, 1.2))
    k3 = torch.randint(0, 10_000, (total,), dtype=torch.int32, device="cuda", generator=g)
    dec = torch.randint(-10 ** 13, 10 ** 13, (total,), dtype=torch.int64, device="cuda", generator=g)
    mx = torch.rand(total, dtype=torch.float64, device="cuda", generator=g) * 2e6 - 1e6
    torch.cuda.synchronize()
    return order, total, dictionary, (f, k1, k2, k3, dec, mx)


def uneven_spec(capacity_hint, max_load_percent=0):
    from vecruntime_amd import abi
    return abi.AggSpec([abi.INT32, abi.INT64, abi.UTF8], [abi.INT32, abi.INT64, abi.FLOAT64],
                       [abi.agg_desc(abi.SUM_WIDE, 1, buffer_precision=25, buffer_scale=2),
                        abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.MAX, 2)],
                       key_nullable=[False] * 3, capacity_hint=capacity_hint, max_load_percent=max_load_percent)


def uneven_batches(order, dictionary, cols, lo=0, hi=None):
    from vecruntime_amd import abi, vr
    f, k1, k2, k3, dec, mx = cols
    out = []
    off = 0
    for bi, n in enumerate(order):
        if hi is not None and bi >= hi:
            break
        if bi >= lo:
            sl = slice(off, off + n)
            out.append(([vr.device_column(abi.INT32, k1[sl]), vr.device_column(abi.INT64, k2[sl]),
                         vr.device_column(abi.UTF8, k3[sl], dictionary=dictionary, length=n)],
                        [vr.device_column(abi.INT32, f[sl]), vr.device_column(abi.INT64, dec[sl]),
                         vr.device_column(abi.FLOAT64, mx[sl])], n))
        off += n
    return out


def run_uneven(dev, args):
    """Batched uneven workload: one persistent launch sequence vs one launch per batch, with the
    autotuner's launch shape for the batched path."""
    import torch
    from vecruntime_amd import abi, autotune, vr
    log("generating the uneven batched input")
    t0 = time.time()
    order, total, dictionary, cols = gen_uneven(scale=args.uneven_scale)
    batches = uneven_batches(order, dictionary, cols)
    thr = int(0.3 * 2 ** 31)
    passing = int((cols[0] < thr).sum().item())
    log("generated %d rows in %d batches (%d pass the filter) in %.1f s" % (total, len(batches), passing, time.time() - t0))
    stream = torch.cuda.current_stream()
    # autotune the batched launch on what a step does: the table's reset (its size follows the load
    # factor) and the update of every batch, on a handle sized as the timed ones; one handle at a
    # time (at 35% load the table takes 96 GB)
    sample = batches
    srows = sum(b[2] for b in sample)
    handles = {}

    def run(c):
        m = c["max_load_percent"]
        if m not in handles:
            for h in handles.values():
                h.close()
            handles.clear()
            torch.cuda.synchronize()
            handles[m] = vr.GroupedAggregate(dev, uneven_spec(int(passing * 0.8), m))
        h = handles[m]
        torch.cuda.synchronize()
        t = time.perf_counter()
        h.reset()
        h.update_batches(sample, filter_col=0, threshold=thr, config=autotune.to_launch_config(c), stream=stream)
        torch.cuda.synchronize()
        return time.perf_counter() - t

    props = torch.cuda.get_device_properties(0)
    key = autotune.signature(props.name, uneven_spec(0), "uneven-filter30-step")
    t0 = time.time()
    best, results = autotune.tuned(key, run, autotune.SPACES["persistent"], reps=3, budget_s=args.tune_budget)
    for h in handles.values():
        h.close()
    log("autotune: %s (%.1f s, %s configurations%s)" % (best, time.time() - t0, len(results) if results else "-",
                                                         "" if results else ", cached"))
    default_cfg = abi.launch_config()
    tuned_cfg = autotune.to_launch_config(best)
    agg = vr.GroupedAggregate(dev, uneven_spec(int(passing * 0.8), best["max_load_percent"]))
    agg.set_launch_config(tuned_cfg)  # the per-batch launches use the tuned shape too
    prepared = [(vr.GroupedAggregate.prepare(k, v), n) for k, v, n in batches]
    res = {"rows": total, "batches": len(batches), "passing_rows": passing, "tuned": best,
           "tuning": {"sample_rows": srows, "configurations": len(results) if results else None,
                      "sample_ms_best": results[0][0] * 1e3 if results else None,
                      "sample_ms_default": next((t * 1e3 for t, c in results if c == autotune.DEFAULT), None)
                      if results else None}}

    # the default configuration on a handle of the default load factor
    agg_def = agg if best["max_load_percent"] == 50 else vr.GroupedAggregate(dev, uneven_spec(int(passing * 0.8), 50))

    def batched(cfg, h=agg):
        h.reset()
        h.update_batches(batches, filter_col=0, threshold=thr, config=cfg, stream=stream)
        torch.cuda.synchronize()

    def plain():
        agg.reset()
        for prep, n in prepared:
            agg.update_prepared(prep, n, 0, thr, stream=stream)
        torch.cuda.synchronize()

    graph = [None]

    def graphed():  # the captured step: one HIP graph of the table's reset and the update
        if graph[0] is None:
            graph[0] = agg.capture(batches, max_groups=int(passing * 0.8), filter_col=0, threshold=thr,
                                   config=tuned_cfg)
        graph[0].launch(stream)
        torch.cuda.synchronize()

    groups = {}
    for name, fn, h in (("batched_tuned", lambda: batched(tuned_cfg), agg),
                        ("batched_default", lambda: batched(default_cfg, agg_def), agg_def),
                        ("plain", plain, agg), ("graph", graphed, agg)):
        fn()  # warm (and sizes the table)
        ts = []
        for _ in range(max(1, args.uneven_steps)):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        dt = min(ts)
        groups[name] = h.num_groups()
        res[name] = dict(ms=dt * 1e3, rows_per_s=total / dt)
        log("uneven %s: %.1f ms, %.3e rows/s, %d groups" % (name, dt * 1e3, total / dt, groups[name]))
    # the captured step's device time, with HIP events on its stream
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(1, args.uneven_steps)
    e0.record(stream)
    for _ in range(reps):
        graph[0].launch(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    dev_ms = e0.elapsed_time(e1) / reps
    alg = total * UNEVEN_BYTES_PER_ROW
    achieved = alg / (dev_ms * 1e-3) / 1e9
    res["graph"].update({"device_ms": dev_ms, "captured": True})
    res["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                       "kernel": "agg_update_kernel<true> in the captured step (table reset + update)",
                       "algorithmic_bytes": alg, "bytes_per_row": UNEVEN_BYTES_PER_ROW}
    log("uneven graph step: %.1f ms device, %.1f GB/s = %.1f%% of HBM" % (dev_ms, achieved, 100 * achieved / HBM_PEAK_GBS))
    res["graph_over_plain"] = res["graph"]["rows_per_s"] / res["plain"]["rows_per_s"]
    res["batched_over_plain"] = res["batched_tuned"]["rows_per_s"] / res["plain"]["rows_per_s"]
    res["groups"] = groups["plain"]
    res["paths_agree"] = len(set(groups.values())) == 1
    graph[0].close()
    if agg_def is not agg:
        agg_def.close()
    agg.close()
    del batches, prepared, cols
    torch.cuda.empty_cache()
    return res


def cpu_baseline(sample, threads):
    from oracle import oracle
    from vecruntime_amd import abi
    from vecruntime_amd.columns import Column
    host, nb = sample
    bk, bv, ln = [], [], []
    for b in range(nb):
        s0, s1 = b * BATCH, (b + 1) * BATCH
        vb = [v[s0 // 8:s1 // 8] for v in host["valid"]]
        bk.append([Column(abi.INT64, host["keys"][s0:s1])])
        bv.append([Column(abi.INT64, host["dec"][s0:s1], vb[0]), Column(abi.FLOAT64, host["dbl"][s0:s1], vb[1]),
                   Column(abi.INT32, host["i32"][s0:s1], vb[2])])
        ln.append(BATCH)
    spec = groupby_spec(0)
    groups, secs = oracle.partial_threads(spec, bk, bv, ln, threads)
    return nb * BATCH / secs, secs, groups


def cpu_shuffle_baseline(keys, parts):
    """The shuffle's map side on the host: the oracle's restatement of PartitionKernels (scalar C++,
    one thread) over a sample of the same int64 keys -- hash partition ids, then the grouping."""
    import time
    from oracle import oracle
    from vecruntime_amd import abi
    from vecruntime_amd.columns import Column
    col = [Column(abi.INT64, keys)]
    t0 = time.perf_counter()
    ids, _ = oracle.hash_partition_ids(col, [abi.HASH_LONG], len(keys), parts)
    t1 = time.perf_counter()
    oracle.partition_rows(ids, parts)
    t2 = time.perf_counter()
    return {"hash_ids_rows_per_s": len(keys) / (t1 - t0), "group_rows_rows_per_s": len(keys) / (t2 - t1), "cores": 1,
            "kind": "port", "sample": "%d rows of the same keys, one thread, %.1f s" % (len(keys), t2 - t0)}


def run_join(dev, args):
    """Join probe: 10M unique int64 build keys (int32 + decimal(7,2) payloads); 400M probe rows in
    65,536-row batches (6,104 batches, 90% hits), two int32 payloads.  Inner materialized (the three
    streamed and the three build columns), semi and anti (the three streamed columns), each as one
    batched probe over every batch (vr_join_probe_batches) and as per-batch probes (vr_join_probe,
    the drop-in), with the C calls' arguments marshalled before the timed region."""
    import ctypes as C
    import torch
    from vecruntime_amd import abi, vr
    from vecruntime_amd.columns import c_array
    nb, npr = args.join_build, args.join_probe
    g = torch.Generator(device="cuda")
    g.manual_seed(103)
    bkeys = scramble(torch.arange(nb, dtype=torch.int64, device="cuda"))
    bpay = torch.randint(-2 ** 31, 2 ** 31, (nb,), dtype=torch.int32, device="cuda", generator=g)
    bdec = torch.randint(0, 10 ** 7, (nb,), dtype=torch.int64, device="cuda", generator=g)
    pkeys = torch.empty(npr, dtype=torch.int64, device="cuda")
    for s0 in range(0, npr, 1 << 26):  # in chunks: the temporaries of 400M rows at once are large
        m = min(1 << 26, npr - s0)
        hit = torch.rand(m, device="cuda", generator=g) < 0.9
        pick = torch.randint(0, nb, (m,), device="cuda", generator=g)
        miss = scramble(torch.arange(nb + s0, nb + s0 + m, dtype=torch.int64, device="cuda"))
        pkeys[s0:s0 + m] = torch.where(hit, bkeys[pick], miss)
        del hit, pick, miss
    p1 = torch.randint(-2 ** 31, 2 ** 31, (npr,), dtype=torch.int32, device="cuda", generator=g)
    p2 = torch.randint(-2 ** 31, 2 ** 31, (npr,), dtype=torch.int32, device="cuda", generator=g)
    spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32, abi.INT64], [0], capacity_hint=nb)
    bt = vr.BuildTable(dev, spec)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bt.append([vr.device_column(abi.INT64, bkeys), vr.device_column(abi.INT32, bpay),
               vr.device_column(abi.INT64, bdec)], nb)
    bt.finish()
    build_s = time.perf_counter() - t0
    expect_hits = int(torch.isin(pkeys[:4_000_000], bkeys).sum().item())  # a sample: the hit rate
    batches = []
    for s0 in range(0, npr, BATCH):
        m = min(BATCH, npr - s0)
        batches.append(([vr.device_column(abi.INT64, pkeys[s0:s0 + m]), vr.device_column(abi.INT32, p1[s0:s0 + m]),
                         vr.device_column(abi.INT32, p2[s0:s0 + m])], m))
    pb = vr.ProbeBatches(batches)
    cap = npr
    outs_all = [torch.empty(cap, dtype=torch.int64, device="cuda"), torch.empty(cap, dtype=torch.int32, device="cuda"),
                torch.empty(cap, dtype=torch.int32, device="cuda"), torch.empty(cap, dtype=torch.int64, device="cuda"),
                torch.empty(cap, dtype=torch.int32, device="cuda"), torch.empty(cap, dtype=torch.int64, device="cuda")]
    types = [abi.INT64, abi.INT32, abi.INT32, abi.INT64, abi.INT32, abi.INT64]
    stream = torch.cuda.current_stream()
    res = {"build_rows": nb, "probe_rows": npr, "batches": len(batches), "build_s": build_s}
    # bytes per probe row read (int64 key + two int32) and per output row written
    out_bytes = {"inner": 16 + 20, "semi": 16, "anti": 16}
    lib = vr.lib()
    for name, jt in (("inner", abi.INNER), ("semi", abi.LEFT_SEMI), ("anti", abi.LEFT_ANTI)):
        bo = [0, 1, 2] if jt == abi.INNER else []
        outs = [vr.device_out_column(t, cap, d) for t, d in zip(types[:3 + len(bo)], outs_all)]
        payload_before = dev.stats()["payload_join_probes"]
        bt.probe_batches(pb, [0], jt, outputs=outs, probe_out=[0, 1, 2], build_out=bo, capacity=cap)  # warm
        payload_table = dev.stats()["payload_join_probes"] > payload_before
        reps = 3
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(reps):
            total, per = bt.probe_batches(pb, [0], jt, outputs=outs, probe_out=[0, 1, 2], build_out=bo, capacity=cap)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / reps
        dev_ms = e0.elapsed_time(e1) / reps
        assert int(per.sum()) == total
        alg = npr * 16 + total * out_bytes[name]
        achieved = alg / (dev_ms * 1e-3) / 1e9
        # the same rows through per-batch calls (the drop-in), arguments marshalled beforehand
        ocs = c_array(outs, abi.VrOutColumn)
        po, boa = abi.int_array([0, 1, 2]), abi.int_array(bo)
        kc = abi.int_array([0])
        calls = []
        for cols, m in batches:
            ca = c_array(cols)
            pout = abi.VrProbeOutput(abi.MEM_DEVICE, cap, None, None, 3, C.cast(po, C.POINTER(C.c_int32)), len(bo),
                                     C.cast(boa, C.POINTER(C.c_int32)), C.cast(ocs, C.POINTER(abi.VrOutColumn)), 0)
            calls.append((ca, m, pout))
        sh = stream.cuda_stream
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ptotal = 0
        for ca, m, pout in calls:
            st = lib.vr_join_probe(bt.h, C.cast(ca, C.c_void_p), 3, C.cast(kc, C.c_void_p), m, None, jt, 0,
                                   C.byref(pout), sh)
            vr.check(st)
            ptotal += pout.rows
        torch.cuda.synchronize()
        plain_s = time.perf_counter() - t1
        assert ptotal == total, (ptotal, total)
        res[name] = {"out_rows": total, "payload_table": payload_table,
                     "batched": {"ms": wall * 1e3, "device_ms": dev_ms, "rows_per_s": npr / wall},
                     "plain": {"ms": plain_s * 1e3, "rows_per_s": npr / plain_s, "calls": len(calls)},
                     "batched_over_plain": plain_s / wall,
                     "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                                  "kernel": "join_probe_kernel<true> (+ tile map, state clear)",
                                  "algorithmic_bytes": alg,
                                  "bytes_per_probe_row": 16, "bytes_per_out_row": out_bytes[name]}}
        log("join %s: batched %.2f ms (device %.2f ms, %.3e probe rows/s, %.1f GB/s = %.1f%% of HBM), "
            "plain %.1f ms over %d calls, %d out rows" % (name, wall * 1e3, dev_ms, npr / wall, achieved,
                                                          100 * achieved / HBM_PEAK_GBS, plain_s * 1e3, len(calls),
                                                          total))
    res["hit_rate_sample"] = expect_hits / 4_000_000
    res["checks"] = {"inner_equals_semi": res["inner"]["out_rows"] == res["semi"]["out_rows"],
                     "semi_plus_anti": res["semi"]["out_rows"] + res["anti"]["out_rows"] == npr}
    bt.close()
    del batches, pb, outs_all
    if not args.skip_cpu:
        threads = cpu_threads()
        sp = min(npr, threads * 1_000_000)
        res["cpu_baseline"] = cpu_join_baseline(spec, [x.cpu().numpy() for x in (bkeys, bpay, bdec)],
                                                [x[:sp].cpu().numpy() for x in (pkeys, p1, p2)], threads)
        log("join cpu (%s): inner %.3g, semi %.3g, anti %.3g probe rows/s" % (
            res["cpu_baseline"]["sample"], res["cpu_baseline"]["inner_rows_per_s"],
            res["cpu_baseline"]["semi_rows_per_s"], res["cpu_baseline"]["anti_rows_per_s"]))
    del pkeys, p1, p2, bkeys, bpay, bdec
    torch.cuda.empty_cache()
    return res


def cpu_threads():
    """Host threads for the CPU baselines: the cores this process may run on, at most 16 (the host
    share of one GPU on the measurement box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = None
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cores": aff}


def cpu_join_baseline(spec, build, probe, threads):
    """The join leg on the host: the oracle's restatement of BuildTable.build (one thread) and the
    probe (scalar C++) over the same build side and a sample of the probe rows, split over `threads`
    threads that probe the one table at once (a broadcast build side shared by a task's threads)."""
    import time
    from concurrent.futures import ThreadPoolExecutor
    from oracle.oracle import OracleJoin
    from vecruntime_amd import abi
    from vecruntime_amd.columns import Column
    nb, npr = len(build[0]), len(probe[0])
    t0 = time.perf_counter()
    oj = OracleJoin(spec, [Column(abi.INT64, build[0]), Column(abi.INT32, build[1]), Column(abi.INT64, build[2])], nb)
    res = {"build_s": time.perf_counter() - t0}
    step = (npr + threads - 1) // threads
    for name, jt in (("inner", abi.INNER), ("semi", abi.LEFT_SEMI), ("anti", abi.LEFT_ANTI)):
        bo = [0, 1, 2] if jt == abi.INNER else []
        parts = []
        for s0 in range(0, npr, step):
            m = min(step, npr - s0)
            parts.append(([Column(abi.INT64, probe[0][s0:s0 + m]), Column(abi.INT32, probe[1][s0:s0 + m]),
                           Column(abi.INT32, probe[2][s0:s0 + m])], m))

        def one(part):
            cols, m = part
            return oj.probe(cols, [0], m, jt, probe_out=[0, 1, 2], build_out=bo, capacity=m)[3]

        t1 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(one, parts))
        res[name + "_rows_per_s"] = npr / (time.perf_counter() - t1)
    res.update({"cores": threads, "kind": "port", **host_info(),
                "sample": "build %d rows (one thread), %d probe rows in %d threads probing the one table, %.1f s"
                % (nb, npr, threads, time.perf_counter() - t0)})
    return res


def run_filters(dev, args):
    """The filter operator's device kernels on 200M-row columns: an int32 < literal selection
    (vr_filter_compare), its compaction of three columns (int64, float64, int32; about 30% pass,
    vr_compact) and a gather of 40M rows of an int64 column by a random index array (vr_gather).
    Kernel time from HIP events on the stream the calls run on; bytes are the columns read whole
    plus what is written."""
    import torch
    from vecruntime_amd import abi, vr
    n = args.filter_rows
    g = torch.Generator(device="cuda")
    g.manual_seed(107)
    f = torch.randint(0, 1000, (n,), dtype=torch.int32, device="cuda", generator=g)
    a = torch.randint(-10 ** 12, 10 ** 12, (n,), dtype=torch.int64, device="cuda", generator=g)
    b = torch.rand(n, dtype=torch.float64, device="cuda", generator=g)
    stream = torch.cuda.current_stream()
    fcol = vr.device_column(abi.INT32, f)
    cols = [vr.device_column(abi.INT64, a), vr.device_column(abi.FLOAT64, b), fcol]
    sel = torch.zeros(((n + 63) // 64) * 8, dtype=torch.uint8, device="cuda")

    def timed(fn, reps=5):
        fn()  # warm-up (and the output sizes)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            r = fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps, r

    ms_cmp, (_, m) = timed(lambda: vr.filter_compare(dev, fcol, abi.CMP_LT, 300, n, out=sel, stream=stream))
    outs = [vr.device_out_column(t, m, torch.empty(m * w, dtype=torch.uint8, device="cuda"))
            for t, w in ((abi.INT64, 8), (abi.FLOAT64, 8), (abi.INT32, 4))]
    ms_cpt, rows = timed(lambda: vr.compact(dev, cols, n, sel, outs, stream=stream))
    assert rows == m
    gm = 40_000_000
    idx = torch.randint(0, 10_000_000, (gm,), dtype=torch.int32, device="cuda", generator=g)
    gout = vr.device_out_column(abi.INT64, gm, torch.empty(gm * 8, dtype=torch.uint8, device="cuda"))
    src = vr.device_column(abi.INT64, a[:10_000_000])
    ms_gat, _ = timed(lambda: vr.gather(dev, src, idx, gm, gout, stream=stream))
    res = {}
    for name, ms, nbytes in (("compare", ms_cmp, n * 4 + n / 8), ("compact", ms_cpt, n * 20 + n / 8 + m * 20),
                             ("gather", ms_gat, gm * (4 + 8 + 8))):
        gbs = nbytes / (ms * 1e-3) / 1e9
        res[name] = {"ms": ms, "GB_per_s": gbs, "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}
        log("filters %s: %.3f ms, %.0f GB/s (%.1f%% of %d)" % (name, ms, gbs, 100 * gbs / HBM_PEAK_GBS, HBM_PEAK_GBS))
    res.update({"rows": n, "selected": m, "gather_rows": gm})
    if not args.skip_cpu:
        res["cpu_baseline"] = cpu_filter_baseline(f[:20_000_000].cpu().numpy(), a[:20_000_000].cpu().numpy(),
                                                  b[:20_000_000].cpu().numpy(), a[:10_000_000].cpu().numpy(),
                                                  idx[:4_000_000].cpu().numpy())
        log("filters cpu (%s): compare %.3g, compact %.3g, gather %.3g rows/s" % (
            res["cpu_baseline"]["sample"], res["cpu_baseline"]["compare_rows_per_s"],
            res["cpu_baseline"]["compact_rows_per_s"], res["cpu_baseline"]["gather_rows_per_s"]))
    del f, a, b, sel, outs, idx, gout
    return res


def cpu_filter_baseline(f, a, b, src, idx):
    """The filter leg on the host: the oracle's restatement of CompareKernels / CompactKernels /
    GatherKernels (scalar C++, one thread) over samples of the same columns and indices."""
    import time
    from oracle import oracle
    from vecruntime_amd import abi
    from vecruntime_amd.columns import Column
    n = len(f)
    t0 = time.perf_counter()
    bits, m = oracle.filter_compare(Column(abi.INT32, f), abi.CMP_LT, 300, n)
    t1 = time.perf_counter()
    oracle.compact([Column(abi.INT64, a), Column(abi.FLOAT64, b), Column(abi.INT32, f)], n, bits,
                   [abi.INT64, abi.FLOAT64, abi.INT32], m)
    t2 = time.perf_counter()
    oracle.gather(Column(abi.INT64, src), idx)
    t3 = time.perf_counter()
    return {"compare_rows_per_s": n / (t1 - t0), "compact_rows_per_s": n / (t2 - t1),
            "gather_rows_per_s": len(idx) / (t3 - t2), "cores": 1, "kind": "port",
            "sample": "%d rows compared and compacted (three columns), %d rows gathered from %d, one thread, %.1f s"
                      % (n, len(idx), len(src), t3 - t0)}


def run_shuffle(dev, args):
    """The shuffle writer's map side on the device over 200M rows: Spark's hash partition ids of an
    int64 key into 200 partitions (vr_hash_partition_ids: 8 bytes read and 4 written per row), then
    the rows grouped by partition with their destinations (vr_partition_rows: the ids read, rows
    and destinations written, 12 bytes per row at the least).  HIP events on the calls' stream."""
    import torch
    from vecruntime_amd import abi, vr
    n, parts = args.filter_rows, 200
    g = torch.Generator(device="cuda")
    g.manual_seed(109)
    a = torch.randint(-2 ** 62, 2 ** 62, (n,), dtype=torch.int64, device="cuda", generator=g)
    stream = torch.cuda.current_stream()
    key = [vr.device_column(abi.INT64, a)]

    def timed(fn, reps=5):
        r = fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            r = fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps, r

    ms_hash, (ids, _) = timed(lambda: vr.hash_partition_ids(dev, key, [abi.HASH_LONG], n, parts, stream=stream))
    ms_group, (rows, dest, starts) = timed(lambda: vr.partition_rows(dev, ids, n, parts, stream=stream))
    assert int(starts[-1].item()) == n
    res = {"rows": n, "partitions": parts}
    for name, ms, nbytes in (("hash_ids", ms_hash, n * 12), ("group_rows", ms_group, n * 12)):
        gbs = nbytes / (ms * 1e-3) / 1e9
        res[name] = {"ms": ms, "rows_per_s": n / (ms * 1e-3), "GB_per_s": gbs, "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}
        log("shuffle %s: %.3f ms, %.0f GB/s (%.1f%% of %d)" % (name, ms, gbs, 100 * gbs / HBM_PEAK_GBS, HBM_PEAK_GBS))
    if not args.skip_cpu:
        res["cpu_baseline"] = cpu_shuffle_baseline(a[:20_000_000].cpu().numpy(), parts)
        log("shuffle cpu (%s): hash_ids %.3g rows/s, group_rows %.3g rows/s"
            % (res["cpu_baseline"]["sample"], res["cpu_baseline"]["hash_ids_rows_per_s"],
               res["cpu_baseline"]["group_rows_rows_per_s"]))
    del a, ids, rows, dest, starts
    return res


SERVE_VALUE_TYPES = ("int32", "int64", "decimal(12,2)", "float64")


def gen_serve_pools(rows, seed=105):
    """Device pools every serving query takes a window of: int64 and int32 keys, uniform and
    Zipf(1.1) over 1e3..1e7 groups; int32 / int64 / decimal(12,2) / float64 values; unique build
    keys scramble(0..rows) and probe keys scramble(j) with j log-uniform over [1, 1.1 * rows], so a
    probe of a build over keys 0..b hits with probability ln(b) / ln(1.1 * rows)."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    pools = {"keys": {}}
    for e in range(3, 8):
        G = 10 ** e
        for dist in ("uniform", "zipf"):
            r = (torch.randint(0, G, (rows,), dtype=torch.int64, device="cuda", generator=g) if dist == "uniform"
                 else zipf_ranks(g, rows, G, 1.1))
            k32 = ((r * 2654435761) & 0xFFFFFFFF).to(torch.int32)
            pools["keys"][(G, dist)] = (scramble(r), k32)
            del r
    pools["int32"] = torch.randint(-2 ** 31, 2 ** 31, (rows,), dtype=torch.int32, device="cuda", generator=g)
    pools["int64"] = torch.randint(-10 ** 6, 10 ** 6 + 1, (rows,), dtype=torch.int64, device="cuda", generator=g)
    pools["decimal(12,2)"] = torch.randint(0, 10 ** 8, (rows,), dtype=torch.int64, device="cuda", generator=g)
    pools["float64"] = torch.empty(rows, dtype=torch.float64, device="cuda").log_normal_(0.0, 1.0, generator=g)
    pools["build_keys"] = scramble(torch.arange(rows, dtype=torch.int64, device="cuda"))
    u = torch.rand(rows, dtype=torch.float64, device="cuda", generator=g)
    j = torch.exp(u * float(np.log(1.1 * rows))).to(torch.int64) - 1
    pools["probe_keys"] = scramble(j)
    del u, j
    pools["build_dec"] = torch.randint(0, 10 ** 7, (rows,), dtype=torch.int64, device="cuda", generator=g)
    torch.cuda.synchronize()
    return pools


def serve_queries(pools, rows, seconds, rate, seed=106):
    """The query stream: Poisson arrivals, 70% group-by / 30% join, sizes log-uniform over
    [1e4, 1e8] rows (capped by the pools)."""
    from vecruntime_amd import abi, vr
    rng = np.random.default_rng(seed)
    qs, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            break
        n = int(min(rows, np.exp(rng.uniform(np.log(1e4), np.log(1e8)))))
        off = int(rng.integers(0, rows - n + 1))
        sl = slice(off, off + n)
        if rng.random() < 0.7:
            G = 10 ** int(rng.integers(3, 8))
            dist = "uniform" if rng.random() < 0.5 else "zipf"
            k64, k32 = pools["keys"][(G, dist)]
            kt = abi.INT64 if rng.random() < 0.5 else abi.INT32
            key = vr.device_column(kt, (k64 if kt == abi.INT64 else k32)[sl])
            vt = SERVE_VALUE_TYPES[int(rng.integers(0, 4))]
            lane = {"int32": abi.INT32, "int64": abi.INT64, "decimal(12,2)": abi.INT64, "float64": abi.FLOAT64}[vt]
            val = vr.device_column(lane, pools[vt][sl])
            aggs = [abi.agg_desc(abi.COUNT_STAR)]
            if vt == "int32":
                aggs += [abi.agg_desc(abi.SUM_LONG, 0), abi.agg_desc(abi.MIN, 0), abi.agg_desc(abi.MAX, 0)]
            elif vt == "int64":
                aggs += [abi.agg_desc(abi.SUM_LONG, 0), abi.agg_desc(abi.MAX, 0)]
            elif vt == "decimal(12,2)":
                aggs += [abi.agg_desc(abi.SUM_WIDE, 0, buffer_precision=22, buffer_scale=2),
                         abi.agg_desc(abi.AVG_WIDE, 0, buffer_precision=22, buffer_scale=2)]
            else:
                aggs += [abi.agg_desc(abi.SUM_DOUBLE, 0), abi.agg_desc(abi.AVG_DOUBLE, 0)]
            spec = abi.AggSpec([kt], [lane], aggs, key_nullable=[False])
            qs.append(dict(at=t, kind="agg", n=n, groups=min(G, n), spec=spec, batches=[([key], [val], n)],
                           desc="%s %s G=%d %s" % ("int64" if kt == abi.INT64 else "int32", dist, G, vt)))
        else:
            nb = max(1000, n // 5)
            npr = max(1, n - nb)
            poff = int(rng.integers(0, rows - npr + 1))
            ps = slice(poff, poff + npr)
            jt = [abi.INNER, abi.INNER, abi.LEFT_SEMI, abi.LEFT_ANTI][int(rng.integers(0, 4))]
            spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32, abi.INT64], [0], capacity_hint=nb)
            build = [vr.device_column(abi.INT64, pools["build_keys"][:nb]),
                     vr.device_column(abi.INT32, pools["int32"][:nb]),
                     vr.device_column(abi.INT64, pools["build_dec"][:nb])]
            probe = [vr.device_column(abi.INT64, pools["probe_keys"][ps]),
                     vr.device_column(abi.INT32, pools["int32"][ps]),
                     vr.device_column(abi.INT32, pools["int32"][rows - npr - poff:rows - poff])]
            qs.append(dict(at=t, kind="join", n=nb + npr, nb=nb, npr=npr, jt=jt, spec=spec, build=build,
                           probe=probe, build_out=[1, 2] if jt == abi.INNER else [],
                           desc="join %s build=%d probe=%d" % (["inner", "semi", "anti"][jt], nb, npr)))
    return qs


def run_serve(dev, args):
    """The serving loop: a Poisson stream of group-bys and joins through the query server."""
    import threading
    import torch
    from vecruntime_amd import abi
    from vecruntime_amd.serve import Server, prewarm
    rows = args.serve_pool_rows
    log("serving: generating device pools of %d rows" % rows)
    t0 = time.time()
    pools = gen_serve_pools(rows)
    qs = serve_queries(pools, rows, args.serve_seconds, args.serve_rate)
    log("serving: pools in %.1f s; %d queries over %.1f s (%d group-by, %d join), %.3e input rows"
        % (time.time() - t0, len(qs), args.serve_seconds, sum(q["kind"] == "agg" for q in qs),
           sum(q["kind"] == "join" for q in qs), sum(q["n"] for q in qs)))
    # the kernels compiled per spec (hipRTC) before the server starts
    t0 = time.time()
    uniq = {}
    for q in qs:
        if q["kind"] == "agg":
            uniq.setdefault(q["desc"].split()[0] + " " + q["desc"].split()[-1], q["spec"])
    prewarm(dev, uniq.values())
    prewarm_s = time.time() - t0
    log("serving: %d group-by specs compiled in %.1f s" % (len(uniq), prewarm_s))
    srv = Server(dev, workers=args.serve_workers, hbm_budget=int(args.serve_budget_gb * 1e9),
                 admission=abi.ADMIT_FIFO if args.serve_fifo else abi.ADMIT_SIZE_AWARE)

    def submit(q):
        if q["kind"] == "agg":
            return srv.submit_agg(q["spec"], q["batches"], expected_groups=q["groups"])
        return srv.submit_join(q["spec"], q["build"], q["nb"], q["probe"], [0], q["npr"], q["jt"],
                               probe_out=[0, 1, 2], build_out=q["build_out"], expected_rows=q["npr"])

    # warm-up: a few queries of each kind, untimed
    for q in [x for x in qs if x["kind"] == "agg"][:3] + [x for x in qs if x["kind"] == "join"][:2]:
        t = submit(q)
        r = srv.wait(t)
        assert r.status == abi.VR_OK, (q["desc"], r.status)
        srv.release(t)
    tickets = [None] * len(qs)
    submitted_at = [0.0] * len(qs)
    ready = [threading.Event() for _ in qs]
    results = [None] * len(qs)
    count_ok = [True]

    def waiter():
        for i, q in enumerate(qs):
            ready[i].wait()
            r = srv.wait(tickets[i])
            done = time.perf_counter()
            results[i] = (r.status, r.latency_ms, r.queue_ms, r.run_ms, r.rows, done)
            if r.status == abi.VR_OK and q["kind"] == "agg" and r.rows <= 100_000:
                cnt = srv.fetch(tickets[i], r)[1]
                count_ok[0] &= int(cnt.data[:r.rows].sum()) == q["n"]
            srv.release(tickets[i])

    th = threading.Thread(target=waiter)
    th.start()
    torch.cuda.synchronize()
    start = time.perf_counter()
    for i, q in enumerate(qs):
        delay = start + q["at"] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        submitted_at[i] = time.perf_counter()
        tickets[i] = submit(q)
        ready[i].set()
    th.join()
    end = max(r[5] for r in results)
    st = srv.stats()
    srv.close()
    failed = sum(1 for r in results if r[0] != abi.VR_OK)
    # latency from the scheduled arrival: any lag of the submitting thread counts against it
    lat = np.array([(submitted_at[i] - (start + q["at"])) * 1e3 + results[i][1] for i, q in enumerate(qs)])
    total_rows = sum(q["n"] for q in qs)
    queue = np.array([r[2] for r in results])
    run = np.array([r[3] for r in results])
    sizes = np.array([q["n"] for q in qs])

    def pct(x, p):
        return float(np.percentile(x, p)) if len(x) else None

    res = {"queries": len(qs), "failed": failed, "seconds": end - start, "rows": total_rows,
           "rows_per_s": total_rows / (end - start), "offered_rows_per_s": total_rows / args.serve_seconds,
           "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
           "queue_p50_ms": pct(queue, 50), "queue_p99_ms": pct(queue, 99),
           "run_p50_ms": pct(run, 50), "run_p99_ms": pct(run, 99), "mean_queue_ms": float(np.mean(queue)),
           "p99_ms_by_rows": {"<1e6": pct(lat[sizes < 1e6], 99), "1e6-1e7": pct(lat[(sizes >= 1e6) & (sizes < 1e7)], 99),
                              ">=1e7": pct(lat[sizes >= 1e7], 99)},
           "agg_p50_ms": pct([lat[i] for i, q in enumerate(qs) if q["kind"] == "agg"], 50),
           "join_p50_ms": pct([lat[i] for i, q in enumerate(qs) if q["kind"] == "join"], 50),
           "peak_inflight": st["peak_inflight"], "peak_reserved_gb": st["peak_reserved_bytes"] / 1e9,
           "workers": args.serve_workers, "rate_qps": args.serve_rate,
           "admission": "fifo" if args.serve_fifo else "size-aware (aging 100 ms)", "prewarm_s": prewarm_s,
           "count_checks_ok": count_ok[0]}
    log("serving: %s" % json.dumps(res))
    del pools, qs
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rows", type=int, default=200_000_000)
    ap.add_argument("--skip-join", action="store_true")
    ap.add_argument("--skip-groupby", action="store_true", help="run only the other legs (no headline line)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-batches", type=int, default=3052)
    ap.add_argument("--join-build", type=int, default=10_000_000)
    ap.add_argument("--join-probe", type=int, default=400_000_000)
    ap.add_argument("--skip-uneven", action="store_true")
    ap.add_argument("--uneven-scale", type=float, default=1.0)
    ap.add_argument("--uneven-steps", type=int, default=2)
    ap.add_argument("--tune-budget", type=float, default=45.0)
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--skip-pcie", action="store_true")
    ap.add_argument("--skip-plain", action="store_true", help="no per-batch leg (for profiling the batched update)")
    ap.add_argument("--skip-filters", action="store_true")
    ap.add_argument("--filter-rows", type=int, default=200_000_000)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-fifo", action="store_true", help="arrival-order admission")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-workers", type=int, default=4)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-pool-rows", type=int, default=100_000_000)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only copy this many MiB device to device three times (scales the PMC byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(args.copy_calibration << 17, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 8, "copies": 3}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs no multi-GPU path (--gpus %d refused)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    from vecruntime_amd import vr
    assert torch.cuda.is_available(), "bench.py needs an MI355X"
    # every leg on one side stream: the library's calls and the HIP events that time them share it
    # (torch's legacy default stream is not ordered with the library's non-blocking streams)
    torch.cuda.set_stream(torch.cuda.Stream())
    dev = vr.Device(0)
    n = args.rows
    if args.skip_groupby:  # the other legs alone, for profiling them; no headline line
        out = {}
        if not args.skip_join:
            out["join"] = run_join(dev, args)
        if not args.skip_uneven:
            out["batched_uneven"] = run_uneven(dev, args)
        if not args.skip_serve:
            out["serving"] = run_serve(dev, args)
        dev.close()
        print(json.dumps(out), flush=True)
        return
    cfg1 = run_cfg1(dev, args)
    gb = run_groupby(dev, args, n)
    pcie = None
    if gb["_host"] is not None and not args.skip_pcie:
        pcie = run_pcie(dev, args, gb["_host"], gb["groups"])
    join = None if args.skip_join else run_join(dev, args)
    uneven = None if args.skip_uneven else run_uneven(dev, args)
    serve = None if args.skip_serve else run_serve(dev, args)
    filters = None if args.skip_filters else run_filters(dev, args)
    shuffle = None if args.skip_filters else run_shuffle(dev, args)
    # roofline of the update: algorithmic bytes (the input columns, read once) / the duration of its
    # kernels (hot-set sampling, scatter, hot reduce, merge), timed with HIP events on their stream
    bytes_per_row = 8 + 8 + 8 + 4 + 3 / 8.0
    alg_bytes = n * bytes_per_row
    achieved = alg_bytes / (gb["update_kernel_ms"] * 1e-3) / 1e9
    cpu = None
    if gb["_sample"] is not None:
        threads = cpu_threads()
        rps, secs, _ = cpu_baseline(gb["_sample"], threads)
        cpu = {"value": rps, "unit": "rows/s", "cores": threads, "kind": "port", **host_info(),
               "sample": "%d batches x 65,536 rows of the same input, %d threads, one partial table per thread, %.1f s"
                         % (gb["_sample"][1], threads, secs)}
    line = {
        "metric": "rows/s group-by agg + join probe, batched vs plain on 1 MI355X; p99 ms",
        "value": gb["batched"]["rows_per_s"],
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": gb["batched"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64/decimal128/f64",
        "data": "synthetic (generated on device)",
        "config": {"workload": "group-by: %d rows, 65,536-row batches, int64 key Zipf(1.1) over 10M, SUM/AVG "
                               "decimal(12,2), AVG f64, MIN/MAX int32, COUNT, 5%% nulls; batched persistent launch"
                               % n, "global_batch": n, "seq_len": BATCH, "parallelism": "single GPU",
                   "groups": gb["groups"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": PMC_TRAFFIC_BYTES.get(n),
                     "kernel": "partitioned update: hot_sample_kernel, part_count_kernel, part_col_scan_kernel, "
                               "part_region_scan_kernel, part_scatter_kernel, hot_reduce_kernel, part_merge_kernel",
                     "kernel_ms_per_step": gb["update_kernel_ms"],
                     "launches_per_step": gb["update_launches_per_step"],
                     "avg_launch_ms": gb["update_kernel_avg_launch_ms"], "algorithmic_bytes_per_step": alg_bytes},
        "cpu_baseline": cpu,
        "plain": {"rows_per_s": gb["plain"]["rows_per_s"], "ms_per_step": gb["plain"]["ms_per_step"],
                  "batched_over_plain": gb["batched"]["rows_per_s"] / gb["plain"]["rows_per_s"]} if "plain" in gb else None,
        "cold_table": gb["cold"],
        "from_host": pcie,
        "cfg1_1m_rows_1k_groups": cfg1,
        "join": join,
        "batched_uneven": uneven,
        "serving": serve,
        "filters": filters,
        "shuffle": shuffle,
    }
    dev.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
