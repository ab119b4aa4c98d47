"""The partitioned group-by update (hot-set scatter, then one merge workgroup per table region)
against the CPU oracle and against the atomic per-tile kernel (MI355X).

The path is forced with launch_config(partitioned=1) so that small inputs take it; the same inputs
through partitioned=-1 take the per-tile kernel.  Parity bar: keys, counts, integer and decimal
sums and MIN/MAX bit-exact; double sums within 1e-12 of the group's sum of magnitudes."""
import struct

import numpy as np
import pytest

from gpu_util import compare_groups, group_rows, out_roles
from oracle.oracle import OracleAgg
from vecruntime_amd import abi
from vecruntime_amd.columns import Column, pack_bits

pytestmark = pytest.mark.gpu

FORCE = dict(partitioned=1)
NEVER = dict(partitioned=-1)
GENERIC = dict(partitioned=1, specialize=-1)  # the generic kernels, not ones compiled for the spec


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    from vecruntime_amd import vr
    d = vr.Device(0)
    yield d
    d.close()


def make(dev, spec, mode=FORCE):
    from vecruntime_amd import vr
    g = vr.GroupedAggregate(dev, spec)
    g.set_launch_config(abi.launch_config(**mode))
    return g


class Took:
    """Counts the partitioned updates the block runs, and those that ran kernels compiled for their
    spec (device statistics)."""

    def __init__(self, dev):
        self.dev = dev

    def __enter__(self):
        st = self.dev.stats()
        self.before = st["partitioned_updates"]
        self.before_spec = st["specialized_updates"]
        return self

    def __exit__(self, *exc):
        st = self.dev.stats()
        self.count = st["partitioned_updates"] - self.before
        self.specialized = st["specialized_updates"] - self.before_spec


def rows_of(agg, spec):
    G = agg.num_groups()
    k, o = agg.emit()
    dicts = {c: agg.key_dictionary(c) for c, t in enumerate(spec.key_types) if t == abi.UTF8}
    return group_rows(k, o, G, spec, dicts)


def check(g, o, spec, rel=1e-12, mags=None):
    """mags: {key tuple: sum of |x| of the group's doubles}, the scale of the double-sum bar
    (default |expected|, which suits sums of one sign)."""
    assert g.num_groups() == o.num_groups()
    ref = rows_of(o, spec)
    scale = None if mags is None else {k: [mags[k]] * len(v) for k, v in ref.items()}
    return compare_groups(rows_of(g, spec), ref, out_roles(spec), rel, scale)


def magnitudes(keys, vals):
    """{(key,): sum of |val|} per int64 key."""
    u, inv = np.unique(keys, return_inverse=True)
    m = np.zeros(len(u))
    np.add.at(m, inv, np.abs(vals))
    return {(int(k),): float(x) for k, x in zip(u, m)}


def zipf_keys(rng, n, distinct, s):
    ranks = np.arange(1, distinct + 1, dtype=np.float64)
    p = ranks ** -s
    p /= p.sum()
    r = rng.choice(distinct, size=n, p=p)
    return (r.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15) ^ np.uint64(0x5851F42D4C957F2D)).astype(np.int64)


def cfg2_spec(**kw):
    return abi.AggSpec(
        [abi.INT64], [abi.INT64, abi.FLOAT64, abi.INT32],
        [abi.agg_desc(abi.SUM_WIDE, 0, buffer_precision=22, buffer_scale=2),
         abi.agg_desc(abi.AVG_WIDE, 0, buffer_precision=22, buffer_scale=2),
         abi.agg_desc(abi.AVG_DOUBLE, 1), abi.agg_desc(abi.MIN, 2), abi.agg_desc(abi.MAX, 2),
         abi.agg_desc(abi.COUNT_STAR)],
        key_nullable=[False], **kw)


def cfg2_data(rng, n, distinct, s=1.1):
    keys = zipf_keys(rng, n, distinct, s)
    dec = rng.integers(0, 10 ** 8, n)
    dbl = rng.lognormal(0, 1, n)
    i32 = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
    nulls = [rng.random(n) < 0.05 for _ in range(3)]
    return keys, dec, dbl, i32, nulls


def device_batches(keys, dec, dbl, i32, nulls, size):
    import torch
    from vecruntime_amd import vr
    n = len(keys)
    out, keep = [], []
    for s0 in range(0, n, size):
        sl = slice(s0, min(n, s0 + size))
        m = sl.stop - sl.start
        t = [torch.from_numpy(np.ascontiguousarray(a[sl])).cuda() for a in (keys, dec, dbl, i32)]
        v = [torch.from_numpy(pack_bits(~nl[sl])).cuda() for nl in nulls]
        keep.append((t, v))
        out.append(([vr.device_column(abi.INT64, t[0])],
                    [vr.device_column(abi.INT64, t[1], validity=v[0]), vr.device_column(abi.FLOAT64, t[2], validity=v[1]),
                     vr.device_column(abi.INT32, t[3], validity=v[2])], m))
    return out, keep


def oracle_cfg2(spec, keys, dec, dbl, i32, nulls):
    o = OracleAgg(spec)
    n = len(keys)
    o.update([Column.longs(keys)], [Column.longs(dec, nulls=nulls[0]), Column.doubles(dbl, nulls=nulls[1]),
                                    Column.ints(i32, nulls=nulls[2])], n)
    return o


@pytest.mark.parametrize("distinct,s", [(100_000, 1.1), (2_000, 1.1), (300_000, 0.6)])
def test_zipf_batches_match_oracle_and_atomic_kernel(dev, distinct, s):
    """The large group-by's shape (Zipf int64 keys, decimal SUM/AVG, double AVG, int MIN/MAX,
    COUNT(*), 5% nulls per value column) in 65,536-row device batches, one batched call."""
    rng = np.random.default_rng(distinct + 11)
    n = 1_500_000
    data = cfg2_data(rng, n, distinct, s)
    spec = cfg2_spec()
    batches, keep = device_batches(*data, 65536)
    gp = make(dev, spec, FORCE)
    ga = make(dev, spec, NEVER)
    gg = make(dev, spec, GENERIC)
    with Took(dev) as t:
        gp.update_batches(batches, config=abi.launch_config(**FORCE))
    assert t.count == 1 and t.specialized == 1
    with Took(dev) as t:
        ga.update_batches(batches, config=abi.launch_config(**NEVER))
    assert t.count == 0
    with Took(dev) as t:
        gg.update_batches(batches, config=abi.launch_config(**GENERIC))
    assert t.count == 1 and t.specialized == 0
    o = oracle_cfg2(spec, *data)
    assert check(gp, o, spec) < 1e-12
    check(ga, o, spec)
    check(gg, o, spec)
    exact = [r == "exact" for r in out_roles(spec)]
    rp, ra, rg = rows_of(gp, spec), rows_of(ga, spec), rows_of(gg, spec)
    assert set(rp) == set(ra) == set(rg)
    for k in rp:
        assert [x for x, e in zip(rp[k], exact) if e] == [x for x, e in zip(ra[k], exact) if e]
        assert [x for x, e in zip(rp[k], exact) if e] == [x for x, e in zip(rg[k], exact) if e]


@pytest.mark.parametrize("n", [1, 3, 63, 64, 65, 127, 129, 1023, 4097, 70_000])
def test_awkward_lengths_with_selection(dev, n):
    """Host columns through the per-call entry point (staged), a selection bitmap, 40 groups."""
    rng = np.random.default_rng(n + 3)
    keys = rng.integers(-20, 20, n)
    vals = rng.integers(-10 ** 12, 10 ** 12, n)
    sel = rng.random(n) < 0.7
    spec = abi.AggSpec([abi.INT64], [abi.INT64], [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_LONG, 0)],
                       key_nullable=[False])
    g = make(dev, spec)
    o = OracleAgg(spec)
    kv = ([Column.longs(keys)], [Column.longs(vals)], n, pack_bits(sel))
    with Took(dev) as t:
        g.update(*kv)
    assert t.count == 1
    o.update(*kv)
    check(g, o, spec)


def test_null_and_sentinel_keys(dev):
    """Null keys and a key equal to the table's empty marker take the reserved rows (device
    atomics inside the scatter); 0 stays apart from null."""
    rng = np.random.default_rng(5)
    marker = 0x7FF4DEADBEEF7FF4
    n = 200_000
    keys = rng.choice(np.array([0, marker, 5, -1, 7, 1 << 40], np.int64), n)
    nulls = rng.random(n) < 0.1
    vals = rng.integers(-1000, 1000, n)
    spec = abi.AggSpec([abi.INT64], [abi.INT64], [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_LONG, 0),
                                                  abi.agg_desc(abi.MIN, 0)])
    g = make(dev, spec)
    o = OracleAgg(spec)
    kv = ([Column.longs(keys, nulls=nulls)], [Column.longs(vals)], n)
    g.update(*kv)
    o.update(*kv)
    assert g.num_groups() == 7
    check(g, o, spec)


def test_int32_and_pair_keys_with_filter(dev):
    """A nullable INT32 key (null flag in the 64-bit cell) and a pair of INT32 keys, behind the
    fused filter value < threshold, in uneven device batches."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(9)
    thr = int(0.3 * 2 ** 31)
    lengths = [int(x) for x in np.exp(rng.uniform(np.log(1000), np.log(80000), 20))]
    spec1 = abi.AggSpec([abi.INT32], [abi.INT32, abi.INT64, abi.FLOAT64],
                        [abi.agg_desc(abi.SUM_WIDE, 1, buffer_precision=25, buffer_scale=2), abi.agg_desc(abi.COUNT_STAR),
                         abi.agg_desc(abi.MAX, 2), abi.agg_desc(abi.COUNT, 2)])
    spec2 = abi.AggSpec([abi.INT32, abi.INT32], [abi.INT32, abi.INT64, abi.FLOAT64],
                        [abi.agg_desc(abi.SUM_LONG, 1), abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.MIN, 2)],
                        key_nullable=[False, False])
    for spec, two in ((spec1, False), (spec2, True)):
        g = make(dev, spec)
        o = OracleAgg(spec)
        batches, keep = [], []
        for n in lengths:
            f = rng.integers(0, 2 ** 31, n).astype(np.int32)
            k1 = rng.integers(0, 3000, n).astype(np.int32)
            k2 = rng.integers(0, 50, n).astype(np.int32)
            knull = rng.random(n) < 0.05
            dec = rng.integers(-10 ** 13, 10 ** 13, n)
            dbl = rng.normal(0, 1e3, n)
            dnull = rng.random(n) < 0.05
            hk = [Column.ints(k1), Column.ints(k2)] if two else [Column.ints(k1, nulls=knull)]
            hv = [Column.ints(f), Column.longs(dec), Column.doubles(dbl, nulls=dnull)]
            o.update(hk, hv, n, pack_bits(f < thr))
            t = [torch.from_numpy(a).cuda() for a in (k1, k2, f, dec, dbl)]
            kv = torch.from_numpy(pack_bits(~knull)).cuda()
            dv = torch.from_numpy(pack_bits(~dnull)).cuda()
            keep.append((t, kv, dv))
            dk = ([vr.device_column(abi.INT32, t[0]), vr.device_column(abi.INT32, t[1])] if two
                  else [vr.device_column(abi.INT32, t[0], validity=kv)])
            dvals = [vr.device_column(abi.INT32, t[2]), vr.device_column(abi.INT64, t[3]),
                     vr.device_column(abi.FLOAT64, t[4], validity=dv)]
            batches.append((dk, dvals, n))
        g.update_batches(batches, filter_col=0, threshold=thr, config=abi.launch_config(**FORCE))
        check(g, o, spec)


def test_merge_stops_grows_and_resumes(dev):
    """2.5M distinct keys into a table sized for 1,000 groups: merge workgroups stop at their
    region's limit, the table grows (regions double) and they resume -- every row folded once."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(108)
    n, distinct = 3_000_000, 2_500_000
    ids = np.concatenate([rng.permutation(distinct), rng.integers(0, distinct, n - distinct)])
    rng.shuffle(ids)
    keys = (ids.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15) ^ np.uint64(0x5851F42D4C957F2D)).astype(np.int64)
    i64 = rng.integers(-10 ** 12, 10 ** 12, n)
    dbl = rng.normal(0, 1, n)
    spec = abi.AggSpec([abi.INT64], [abi.INT64, abi.FLOAT64],
                       [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_LONG, 0), abi.agg_desc(abi.SUM_DOUBLE, 1)],
                       key_nullable=[False], capacity_hint=1000)
    grows0 = dev.stats()["table_grows"]
    g = make(dev, spec)
    tk, ti, td = (torch.from_numpy(a).cuda() for a in (keys, i64, dbl))
    g.update_batches([([vr.device_column(abi.INT64, tk)], [vr.device_column(abi.INT64, ti),
                                                            vr.device_column(abi.FLOAT64, td)], n)])
    assert dev.stats()["table_grows"] - grows0 >= 2
    o = OracleAgg(spec)
    o.update([Column.longs(keys)], [Column.longs(i64), Column.doubles(dbl)], n)
    check(g, o, spec, mags=magnitudes(keys, dbl))


def test_double_minmax_specials_first_seen(dev):
    """Double MIN/MAX carry the row index in the record: NaN greatest with the first NaN's payload,
    -0.0 / +0.0 ties keep the first seen, infinities; bit-exact against the oracle."""
    rng = np.random.default_rng(77)
    n = 400_000
    keys = rng.integers(0, 2000, n).astype(np.int64)
    vals = rng.normal(0, 1, n)
    special = rng.random(n) < 0.02
    pool = [struct.unpack("<d", struct.pack("<Q", b))[0]
            for b in (0x7FF8000000000001, 0x7FF8000000000002, 0xFFF8000000000003)] + [0.0, -0.0, np.inf, -np.inf]
    vals[special] = rng.choice(pool, int(special.sum()))
    nulls = rng.random(n) < 0.03
    spec = abi.AggSpec([abi.INT64], [abi.FLOAT64], [abi.agg_desc(abi.MIN, 0), abi.agg_desc(abi.MAX, 0),
                                                    abi.agg_desc(abi.COUNT, 0)], key_nullable=[False])
    g = make(dev, spec)
    o = OracleAgg(spec)
    for s0 in range(0, n, 100_000):
        sl = slice(s0, s0 + 100_000)
        kv = ([Column.longs(keys[sl])], [Column.doubles(vals[sl], nulls=nulls[sl])], 100_000)
        g.update(*kv)
        o.update(*kv)
    check(g, o, spec)


def test_mixed_with_per_tile_updates_and_reset(dev):
    """Partitioned and per-tile updates fold into one table (same regions); after a reset the
    same input gives the same groups again."""
    rng = np.random.default_rng(31)
    spec = cfg2_spec()
    from vecruntime_amd import vr
    g = vr.GroupedAggregate(dev, spec)
    o = OracleAgg(spec)
    parts = []
    for mode, n in ((FORCE, 300_000), (NEVER, 65_536), (FORCE, 200_000), (NEVER, 1000)):
        keys, dec, dbl, i32, nulls = cfg2_data(rng, n, 50_000)
        kv = ([Column.longs(keys)], [Column.longs(dec, nulls=nulls[0]), Column.doubles(dbl, nulls=nulls[1]),
                                     Column.ints(i32, nulls=nulls[2])], n)
        parts.append((mode, kv))
        g.set_launch_config(abi.launch_config(**mode))
        g.update(*kv)
        o.update(*kv)
    first = rows_of(g, spec)
    check(g, o, spec)
    g.reset()
    for mode, kv in parts:
        g.set_launch_config(abi.launch_config(**mode))
        g.update(*kv)
    again = rows_of(g, spec)
    assert set(first) == set(again)
    compare_groups(again, rows_of(o, spec), out_roles(spec))


def test_few_groups_all_hot(dev):
    """300 groups: every row folds into the scatter workgroups' hot copies (no records); 9M rows in
    one call take the partitioned path by default."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(107)
    n = 9_000_000
    pool = rng.integers(-2 ** 63, 2 ** 63 - 1, 300, dtype=np.int64)
    keys = pool[rng.integers(0, 300, n)]
    i64 = rng.integers(-10 ** 12, 10 ** 12, n)
    spec = abi.AggSpec([abi.INT64], [abi.INT64], [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_LONG, 0),
                                                  abi.agg_desc(abi.MAX, 0)], key_nullable=[False], capacity_hint=300)
    g = vr.GroupedAggregate(dev, spec)
    tk, ti = torch.from_numpy(keys).cuda(), torch.from_numpy(i64).cuda()
    with Took(dev) as t:
        g.update([vr.device_column(abi.INT64, tk)], [vr.device_column(abi.INT64, ti)], n)
    assert t.count == 1
    o = OracleAgg(spec)
    o.update([Column.longs(keys)], [Column.longs(i64)], n)
    check(g, o, spec)


def test_large_regions_claim_by_cas(dev):
    """A table far larger than the input (4M-group hint: 8M slots in few regions, each above the
    merge's LDS claim bits) makes the merge claim its new slots with device CAS."""
    rng = np.random.default_rng(131)
    n = 600_000
    data = cfg2_data(rng, n, 200_000, 0.8)
    spec = cfg2_spec(capacity_hint=4_000_000)
    batches, keep = device_batches(*data, 65536)
    g = make(dev, spec, FORCE)
    with Took(dev) as t:
        g.update_batches(batches, config=abi.launch_config(**FORCE))
    assert t.count == 1
    o = oracle_cfg2(spec, *data)
    check(g, o, spec)


@pytest.mark.parametrize("mode", [FORCE, GENERIC])
def test_uniform_keys_fill_the_record_pool(dev, mode):
    """Every row a record (uniform keys, no hot set to speak of): 85M rows over 10M groups in
    65,536-row device batches, the shape whose record pool once ran out.  Same groups and states
    as the per-tile kernel, compared column by column after sorting by key."""
    import torch
    from vecruntime_amd import vr
    n, G, bs = 85_000_000, 10_000_000, 65536
    print("uniform record pool: %s" % mode, flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(12)
    keys = torch.randint(0, G, (n,), dtype=torch.int64, device="cuda", generator=g) * 2654435761
    vals = torch.randint(-2 ** 31, 2 ** 31, (n,), dtype=torch.int32, device="cuda", generator=g)
    spec = abi.AggSpec([abi.INT64], [abi.INT32],
                       [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_LONG, 0), abi.agg_desc(abi.MIN, 0),
                        abi.agg_desc(abi.MAX, 0)], key_nullable=[False], capacity_hint=G)
    batches = [([vr.device_column(abi.INT64, keys[s:s + bs])], [vr.device_column(abi.INT32, vals[s:s + bs])],
                min(bs, n - s)) for s in range(0, n, bs)]
    out = []
    for m in (mode, NEVER):
        a = vr.GroupedAggregate(dev, spec)
        with Took(dev) as t:
            a.update_batches(batches, config=abi.launch_config(**m))
        assert t.count == (1 if m is not NEVER else 0)
        k, o = a.emit()
        m_ = a.num_groups()
        kk = k[0].data[:m_].copy()
        order = np.argsort(kk, kind="stable")
        out.append([kk[order]] + [c.data[:m_][order].copy() for c in o])
        a.close()
    assert len(out[0][0]) == len(out[1][0])
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)
    assert int(out[0][1].sum()) == n
