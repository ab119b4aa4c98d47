"""Device group-by aggregation against the golden fixtures and the CPU oracle (MI355X)."""
import math
import struct

import numpy as np
import pytest

from golden_util import batch_slices, f64_bits, grouped_batches, grouped_spec, load
from gpu_util import compare_groups, group_rows, out_roles
from oracle.oracle import OracleAgg
from vecruntime_amd import abi
from vecruntime_amd.columns import Column, pack_bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    from vecruntime_amd import vr
    d = vr.Device(0)
    yield d
    d.close()


def run_both(dev, spec, batches):
    """batches: list of (keys, values, n, selection). Returns (dev agg, oracle agg)."""
    from vecruntime_amd import vr
    g = vr.GroupedAggregate(dev, spec)
    o = OracleAgg(spec)
    for keys, values, n, sel in batches:
        g.update(keys, values, n, sel)
        o.update(keys, values, n, sel)
    return g, o


def dicts_of(agg, spec):
    return {c: agg.key_dictionary(c) for c, t in enumerate(spec.key_types) if t == abi.UTF8}


def compare(g, o, spec, rel=1e-12, scale=None):
    G = o.num_groups()
    assert g.num_groups() == G
    gk, go = g.emit()
    ok, oo = o.emit()
    dr = group_rows(gk, go, G, spec, dicts_of(g, spec))
    orr = group_rows(ok, oo, G, spec, dicts_of(o, spec))
    return compare_groups(dr, orr, out_roles(spec), rel, scale)


@pytest.mark.parametrize("with_selection", [False, True])
@pytest.mark.parametrize("cardinality", [1, 4, 64, 65, 300, 5000])
@pytest.mark.parametrize("dict_encoded", [True, False])
def test_grouped_accumulators_match_golden(dev, cardinality, with_selection, dict_encoded):
    """GroupedAggregationTest.groupedAccumulatorsMatchReference / HonourSelection draws on the
    device, deterministic mode: groups in first-seen order, every output equal to the fixture
    (double sums bit for bit)."""
    from vecruntime_amd import vr
    z = load("grouped_agg_c%d_s%d.npz" % (cardinality, int(with_selection)))
    batches, strings = grouped_batches(z, dict_encoded)
    spec = grouped_spec(deterministic=True)
    g = vr.GroupedAggregate(dev, spec)
    for keys, values, n, sel in batches:
        g.update(keys, values, n, sel)
    G = g.num_groups()
    assert G == len(z["exp_rows"])
    keys, outs = g.emit()
    k1ok = keys[0].valid(G)
    assert (~k1ok == z["exp_k1_null"]).all()
    assert (keys[0].data[:G][k1ok] == z["exp_k1"][k1ok]).all()
    tdict = g.key_dictionary(1)
    k2ok = keys[1].valid(G)
    got_k2 = [tdict[keys[1].data[i]] if k2ok[i] else None for i in range(G)]
    assert got_k2 == [strings[c] if c >= 0 else None for c in z["exp_k2"]]
    dsum, lsum, rows, cnt, dmin, dmax = outs
    assert (~dsum.valid(G) == z["exp_dsum_null"]).all()
    ok = ~z["exp_dsum_null"]
    assert (f64_bits(dsum.data[:G])[ok] == z["exp_dsum_bits"][ok]).all()
    assert (~lsum.valid(G) == z["exp_lsum_null"]).all()
    assert (lsum.data[:G][~z["exp_lsum_null"]] == z["exp_lsum"][~z["exp_lsum_null"]]).all()
    assert (rows.data[:G] == z["exp_rows"]).all()
    assert (cnt.data[:G] == z["exp_count"]).all()
    has = z["exp_has"]
    assert (dmin.valid(G) == has).all() and (dmax.valid(G) == has).all()
    assert (f64_bits(dmin.data[:G])[has] == z["exp_min_bits"][has]).all()
    assert (f64_bits(dmax.data[:G])[has] == z["exp_max_bits"][has]).all()


@pytest.mark.parametrize("cardinality", [1, 4, 300, 5000])
def test_grouped_fast_mode_matches_oracle(dev, cardinality):
    """The default (atomic) mode on the same draws: groups as a multiset, double sums within the
    reference test's fast-mode tolerance |x| * 1e-12 + 1e-9 (GroupedAggregationTest.java:240)."""
    z = load("grouped_agg_c%d_s1.npz" % cardinality)
    batches, _ = grouped_batches(z, True)
    spec = grouped_spec(deterministic=False)
    g, o = run_both(dev, spec, batches)
    G = o.num_groups()
    gk, go = g.emit()
    ok, oo = o.emit()
    dr = group_rows(gk, go, G, spec, dicts_of(g, spec))
    orr = group_rows(ok, oo, G, spec, dicts_of(o, spec))
    scale = {k: [abs(v[0]) + 1e-9 / 1e-12 if isinstance(v, tuple) else 0 for v in r] for k, r in orr.items()}
    compare_groups(dr, orr, out_roles(spec), 1e-12, scale)


@pytest.mark.parametrize("cardinality", [1, 3, 64, 300, 5000])
@pytest.mark.parametrize("deterministic", [True, False])
def test_strict_double_sum(dev, cardinality, deterministic):
    """strictDoubleSumIsBitIdenticalToSequentialReference: bit-identical in deterministic mode;
    the atomic mode within 1e-12 of the group's sum of magnitudes."""
    from vecruntime_amd import vr
    z = load("strict_sum_c%d.npz" % cardinality)
    spec = abi.AggSpec([abi.INT32], [abi.FLOAT64], [abi.agg_desc(abi.SUM_DOUBLE, 0), abi.agg_desc(abi.COUNT, 0)],
                       key_nullable=[False], deterministic=deterministic)
    g = vr.GroupedAggregate(dev, spec)
    for sl, n in batch_slices(z["lengths"]):
        sel = z["sel"][sl]
        g.update([Column.ints(z["ids"][sl])], [Column.doubles(z["vals"][sl], nulls=z["val_null"][sl])], n,
                 None if sel.all() else pack_bits(sel))
    G = g.num_groups()
    keys, (s, c) = g.emit()
    order = np.argsort(keys[0].data[:G]) if not deterministic else np.arange(G)
    exp_order = np.argsort(z["exp_key"]) if not deterministic else np.arange(G)
    assert (keys[0].data[:G][order] == z["exp_key"][exp_order]).all()
    assert (c.data[:G][order] == z["exp_count"][exp_order]).all()
    got = s.data[:G][order]
    exp_bits = z["exp_sum_bits"][exp_order]
    ok = ~z["exp_sum_null"][exp_order]
    if deterministic:
        assert (f64_bits(got)[ok] == exp_bits[ok]).all()
    else:
        ids = z["ids"]
        live = (~z["val_null"]) & z["sel"]
        mag = np.zeros(max(ids) + 1)
        np.add.at(mag, ids[live], np.abs(z["vals"][live]))
        keyv = keys[0].data[:G][order]
        exp = exp_bits.view(np.float64)
        assert (np.abs(got - exp)[ok] <= 1e-12 * mag[keyv][ok]).all()


def test_minmax_special_values(dev):
    """NaN greatest, signed zeros equal with the first-seen kept, NaN payload of the first NaN."""
    from vecruntime_amd import vr
    z = load("minmax_special.npz")
    spec = abi.AggSpec([abi.INT32], [abi.FLOAT64],
                       [abi.agg_desc(abi.MIN, 0), abi.agg_desc(abi.MAX, 0), abi.agg_desc(abi.SUM_DOUBLE, 0)],
                       deterministic=True)
    g = vr.GroupedAggregate(dev, spec)
    for sl, n in batch_slices(z["lengths"]):
        g.update([Column.ints(z["keys"][sl])], [Column.doubles(z["vals"][sl], nulls=z["val_null"][sl])], n)
    G = g.num_groups()
    keys, (mn, mx, sm) = g.emit()
    assert (keys[0].data[:G] == z["exp_key"]).all()
    has = z["exp_has"]
    assert (mn.valid(G) == has).all()
    assert (f64_bits(mn.data[:G])[has] == z["exp_min_bits"][has]).all()
    assert (f64_bits(mx.data[:G])[has] == z["exp_max_bits"][has]).all()


def test_signed_zero_and_nan_payload_first_seen(dev):
    """-0.0 before +0.0 keeps -0.0 (and the reverse); the first NaN's payload wins."""
    from vecruntime_amd import vr
    nan_a = struct.unpack("<d", struct.pack("<q", 0x7FF8000000000001))[0]
    nan_b = struct.unpack("<d", struct.pack("<q", 0x7FF8000000000002))[0]
    keys = [0, 0, 1, 1, 2, 2, 3]
    vals = [-0.0, 0.0, 0.0, -0.0, nan_a, nan_b, -1.0]
    spec = abi.AggSpec([abi.INT32], [abi.FLOAT64], [abi.agg_desc(abi.MIN, 0), abi.agg_desc(abi.MAX, 0)],
                       deterministic=True)
    for split in (1, 3, 7):
        g = vr.GroupedAggregate(dev, spec)
        o = OracleAgg(spec)
        for s0 in range(0, 7, split):
            kk, vv = [Column.ints(keys[s0:s0 + split])], [Column.doubles(vals[s0:s0 + split])]
            g.update(kk, vv, len(keys[s0:s0 + split]))
            o.update(kk, vv, len(keys[s0:s0 + split]))
        _, (gmn, gmx) = g.emit()
        _, (omn, omx) = o.emit()
        assert (f64_bits(gmn.data[:4]) == f64_bits(omn.data[:4])).all()
        assert (f64_bits(gmx.data[:4]) == f64_bits(omx.data[:4])).all()


def test_wide_sums_and_decimal_average(dev):
    """Decimal SUM/AVG buffers and results over DECIMAL128 extremes and INT64 unscaled lanes."""
    from vecruntime_amd import vr
    z = load("wide_sum.npz")
    aggs = [abi.agg_desc(abi.SUM_WIDE, 0, buffer_precision=38, buffer_scale=2),
            abi.agg_desc(abi.SUM_WIDE, 1, buffer_precision=22, buffer_scale=2),
            abi.agg_desc(abi.AVG_WIDE, 1, buffer_precision=22, buffer_scale=2),
            abi.agg_desc(abi.AVG_WIDE, 1, flags=abi.RESULT, buffer_precision=22, buffer_scale=2,
                         result_precision=16, result_scale=6),
            abi.agg_desc(abi.SUM_WIDE, 1, flags=abi.RESULT, buffer_precision=22, buffer_scale=2)]
    spec = abi.AggSpec([abi.INT32], [abi.DECIMAL128, abi.INT64], aggs, deterministic=True)
    g = vr.GroupedAggregate(dev, spec)
    o = OracleAgg(spec)
    off = 0
    for n in z["batch"]:
        n = int(n)
        sl = slice(off, off + n)
        wide = Column(abi.DECIMAL128, np.ascontiguousarray(z["wide"][2 * off:2 * (off + n)]),
                      pack_bits(~z["wide_null"][sl]), length=n)
        kv = [Column.ints(z["keys"][sl])], [wide, Column.longs(z["narrow"][sl], nulls=z["narrow_null"][sl])]
        g.update(*kv, n)
        o.update(*kv, n)
        off += n
    G = o.num_groups()
    gk, go = g.emit()
    ok, oo = o.emit()
    assert (gk[0].data[:G] == ok[0].data[:G]).all()
    for a, b in zip(go, oo):
        assert a.values(G) == b.values(G)


def test_wide_state_rows_emit_from_global_memory(dev):
    """State rows wider than the emission stages in LDS (MIN and MAX over four double columns,
    four words each, two double sums with their null counts, COUNT(*): 37 words) are emitted from
    global memory; every output equals the oracle's (the doubles are whole numbers, so sums are
    exact in any order).  The last workgroup of the emission holds fewer than 256 groups."""
    rng = np.random.default_rng(77)
    spec = abi.AggSpec([abi.INT64], [abi.FLOAT64] * 4,
                       [abi.agg_desc(abi.MIN, c) for c in range(4)] + [abi.agg_desc(abi.MAX, c) for c in range(4)] +
                       [abi.agg_desc(abi.SUM_DOUBLE, 0), abi.agg_desc(abi.SUM_DOUBLE, 3), abi.agg_desc(abi.COUNT_STAR)],
                       key_nullable=[False])
    batches = []
    for n in (1000, 777, 3001):
        keys = rng.integers(0, 1499, n)
        vals = [Column.doubles(rng.integers(-1000, 1000, n).astype(np.float64), nulls=rng.random(n) < 0.1)
                for _ in range(4)]
        batches.append(([Column.longs(keys)], vals, n, None))
    g, o = run_both(dev, spec, batches)
    assert compare(g, o, spec, rel=0.0) == 0.0


def test_checked_long_sum_overflow(dev):
    from vecruntime_amd import vr
    big = (1 << 63) - 1
    spec = abi.AggSpec([abi.INT32], [abi.INT64], [abi.agg_desc(abi.SUM_LONG, 0, flags=abi.CHECKED)])
    g = vr.GroupedAggregate(dev, spec)
    g.update([Column.ints([1, 1, 2])], [Column.longs([big, 1, 5])], 3)
    with pytest.raises(vr.VrError) as e:
        g.emit()
    assert e.value.status == abi.VR_ERR_ARITHMETIC_OVERFLOW
    spec = abi.AggSpec([abi.INT32], [abi.INT64], [abi.agg_desc(abi.SUM_LONG, 0)])
    g = vr.GroupedAggregate(dev, spec)
    g.update([Column.ints([1, 1])], [Column.longs([big, 1])], 2)
    assert g.emit()[1][0].values(1) == [-(1 << 63)]


def test_checked_long_sum_first_prefix_in_deterministic_mode(dev):
    """The reference's checked LongSum adds with Math.addExact row by row (GroupedAccumulators.java:
    563-601), so [Long.MAX, 1, -1] throws although its total fits.  The deterministic mode folds
    each group's rows in row order and raises the same, also when the overflowing prefix spans two
    batches; the oracle agrees.  Prefixes that all fit give the sum."""
    from oracle.oracle import OracleError
    from vecruntime_amd import vr
    big = (1 << 63) - 1
    spec = abi.AggSpec([abi.INT32], [abi.INT64], [abi.agg_desc(abi.SUM_LONG, 0, flags=abi.CHECKED)],
                       deterministic=True)
    for parts in ([([1, 1, 1, 2], [big, 1, -1, 5])], [([1, 2], [big, 5]), ([3, 1, 1], [0, 1, -1])]):
        g, o = vr.GroupedAggregate(dev, spec), OracleAgg(spec)
        raised = []
        for k, v in parts:
            g.update([Column.ints(k)], [Column.longs(v)], len(k))
            try:
                o.update([Column.ints(k)], [Column.longs(v)], len(k))
            except OracleError as oe:  # the oracle throws where the reference does: in the update
                raised.append(oe.status)
        assert raised == [abi.VR_ERR_ARITHMETIC_OVERFLOW]
        with pytest.raises(vr.VrError) as e:  # the device reports it when the group is emitted
            g.emit()
        assert e.value.status == abi.VR_ERR_ARITHMETIC_OVERFLOW
        g.close()
    g = vr.GroupedAggregate(dev, spec)
    g.update([Column.ints([1, 1, 1, 1])], [Column.longs([big, -1, 1, -(1 << 63) + 1])], 4)
    g.update([Column.ints([1])], [Column.longs([-(1 << 63)])], 1)
    assert g.emit()[1][0].values(1) == [-(1 << 63)]  # every prefix fits; the total is Long.MIN_VALUE
    g.close()


@pytest.mark.parametrize("det", [True, False])
def test_null_and_sentinel_keys(dev, det):
    """Null keys group together apart from 0; a key equal to the table's empty marker groups too.
    Both live in reserved rows after the table's slots: in first-seen order when deterministic,
    otherwise emitted after the occupied slots (any order of groups, same groups and counts)."""
    from vecruntime_amd import vr
    marker = 0x7FF4DEADBEEF7FF4
    keys = np.array([0, 0, marker, 5, marker, 0, -1], np.int64)
    nulls = [False, True, False, False, False, True, False]
    spec = abi.AggSpec([abi.INT64], [], [abi.agg_desc(abi.COUNT_STAR)], deterministic=det)
    g = vr.GroupedAggregate(dev, spec)
    o = OracleAgg(spec)
    col = [Column.longs(keys, nulls=nulls)]
    g.update(col, [], 7)
    o.update(col, [], 7)
    assert g.num_groups() == o.num_groups() == 5
    gk, (gc,) = g.emit()
    ok, (oc,) = o.emit()
    if det:
        assert gk[0].values(5) == ok[0].values(5)
        assert gc.values(5) == oc.values(5)
    else:
        assert sorted(zip(gk[0].values(5), gc.values(5)), key=repr) == sorted(zip(ok[0].values(5), oc.values(5)), key=repr)
    # a second emission of the same groups, and one after more rows, agree as well
    g.update(col, [], 7)
    o.update(col, [], 7)
    gk, (gc,) = g.emit()
    ok, (oc,) = o.emit()
    assert sorted(zip(gk[0].values(5), gc.values(5)), key=repr) == sorted(zip(ok[0].values(5), oc.values(5)), key=repr)


@pytest.mark.parametrize("n", [0, 1, 3, 7, 8, 15, 16, 31, 63, 64, 65, 100, 127, 128, 129, 1000, 1023, 4097])
def test_awkward_lengths(dev, n):
    rng = np.random.default_rng(n + 7)
    keys = rng.integers(-20, 20, n)
    vals = rng.integers(-10 ** 6, 10 ** 6, n)
    sel = rng.random(n) < 0.7
    spec = abi.AggSpec([abi.INT64], [abi.INT64], [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_LONG, 0)],
                       key_nullable=[False], deterministic=True)
    batches = [([Column.longs(keys)], [Column.longs(vals)], n, pack_bits(sel) if n else None)]
    g, o = run_both(dev, spec, batches)
    assert g.num_groups() == o.num_groups()
    if o.num_groups():
        gk, go = g.emit()
        ok, oo = o.emit()
        G = o.num_groups()
        assert (gk[0].data[:G] == ok[0].data[:G]).all()
        for a, b in zip(go, oo):
            assert (a.data[:G] == b.data[:G]).all()


def test_reference_cpu_config_1m_rows(dev):
    """The reference's CPU-path config: 1M rows, int64 key over 1,000 values, COUNT(*),
    SUM(int64 in [-1e6, 1e6]) checked, SUM(float64 in [0,1)); device-resident columns."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(101)
    n = 1_000_000
    key_pool = rng.integers(-2 ** 63, 2 ** 63 - 1, 1000, dtype=np.int64)
    keys = key_pool[rng.integers(0, 1000, n)]
    v1 = rng.integers(-10 ** 6, 10 ** 6 + 1, n)
    v2 = (rng.integers(0, 2 ** 63, n, dtype=np.uint64) >> np.uint64(10)).astype(np.float64) * 2.0 ** -53
    spec = abi.AggSpec([abi.INT64], [abi.INT64, abi.FLOAT64],
                       [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_LONG, 0, flags=abi.CHECKED),
                        abi.agg_desc(abi.SUM_DOUBLE, 1)], key_nullable=[False])
    tk, t1, t2 = (torch.from_numpy(a).cuda() for a in (keys, v1, v2))
    g = vr.GroupedAggregate(dev, spec)
    g.update([vr.device_column(abi.INT64, tk)], [vr.device_column(abi.INT64, t1), vr.device_column(abi.FLOAT64, t2)], n)
    o = OracleAgg(spec)
    o.update([Column.longs(keys)], [Column.longs(v1), Column.doubles(v2)], n)
    worst = compare(g, o, spec)
    assert worst < 1e-12


def zipf_keys(rng, n, distinct, s):
    ranks = np.arange(1, distinct + 1, dtype=np.float64)
    p = ranks ** -s
    p /= p.sum()
    r = rng.choice(distinct, size=n, p=p)
    return (r.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15) ^ np.uint64(0x5851F42D4C957F2D)).astype(np.int64)


@pytest.mark.parametrize("deterministic", [False, True])
def test_zipf_high_cardinality_with_nulls(dev, deterministic):
    """The large group-by's shape at 1/100 scale: Zipf(1.1) int64 keys over 100k values in 65,536-row
    batches, SUM/AVG decimal(12,2), AVG double, MIN/MAX int32, COUNT(*), 5% nulls per value column;
    the table grows from its initial size several times."""
    from vecruntime_amd import vr
    rng = np.random.default_rng(102)
    n = 2_000_000
    keys = zipf_keys(rng, n, 100_000, 1.1)
    dec = rng.integers(0, 10 ** 8, n)
    dbl = rng.lognormal(0, 1, n)
    i32 = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
    nulls = [rng.random(n) < 0.05 for _ in range(3)]
    spec = abi.AggSpec(
        [abi.INT64], [abi.INT64, abi.FLOAT64, abi.INT32],
        [abi.agg_desc(abi.SUM_WIDE, 0, buffer_precision=22, buffer_scale=2),
         abi.agg_desc(abi.AVG_WIDE, 0, buffer_precision=22, buffer_scale=2),
         abi.agg_desc(abi.AVG_DOUBLE, 1), abi.agg_desc(abi.MIN, 2), abi.agg_desc(abi.MAX, 2),
         abi.agg_desc(abi.COUNT_STAR)],
        key_nullable=[False], deterministic=deterministic)
    batches = []
    for s0 in range(0, n, 65536):
        sl = slice(s0, min(n, s0 + 65536))
        m = sl.stop - sl.start
        batches.append(([Column.longs(keys[sl])],
                        [Column.longs(dec[sl], nulls=nulls[0][sl]), Column.doubles(dbl[sl], nulls=nulls[1][sl]),
                         Column.ints(i32[sl], nulls=nulls[2][sl])], m, None))
    g, o = run_both(dev, spec, batches)
    if deterministic:
        G = o.num_groups()
        gk, go = g.emit()
        ok, oo = o.emit()
        assert (gk[0].data[:G] == ok[0].data[:G]).all()
        for a, b in zip(go, oo):
            assert (a.data[:G] == b.data[:G]).all() and (a.valid(G) == b.valid(G)).all()
    else:
        worst = compare(g, o, spec)
        assert worst < 1e-12


def test_composite_key_filter_batched_matches_plain_and_oracle(dev):
    """The batched workload's shape, small: filter int32 < 30% threshold, key (int32, int64 Zipf,
    dictionary string), SUM decimal(15,2), COUNT(*), MAX double; uneven batches in one persistent
    launch (device-resident) against per-batch launches and the oracle."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(104)
    words = ["".join(chr(97 + x) for x in rng.integers(0, 26, rng.integers(8, 25))) for _ in range(500)]
    dictionary = Column.strings(words)
    lengths = [int(x) for x in np.exp(rng.uniform(np.log(1000), np.log(60000), 24))]
    thr = int(0.3 * 2 ** 31)
    spec = abi.AggSpec([abi.INT32, abi.INT64, abi.UTF8], [abi.INT32, abi.INT64, abi.FLOAT64],
                       [abi.agg_desc(abi.SUM_WIDE, 1, buffer_precision=25, buffer_scale=2),
                        abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.MAX, 2)], key_nullable=[False] * 3)
    gb = vr.GroupedAggregate(dev, spec)
    gp = vr.GroupedAggregate(dev, spec)
    o = OracleAgg(spec)
    dev_batches, keep = [], []
    for n in lengths:
        f = rng.integers(0, 2 ** 31, n).astype(np.int32)
        k1 = rng.integers(1, 101, n).astype(np.int32)
        k2 = zipf_keys(rng, n, 5000, 1.2)
        k3 = rng.integers(0, len(words), n).astype(np.int32)
        dec = rng.integers(-10 ** 13, 10 ** 13, n)
        mx = rng.uniform(-1e6, 1e6, n)
        sel = f < thr
        host_k = [Column.ints(k1), Column.longs(k2), Column.dictionary_strings(k3, dictionary)]
        host_v = [Column.ints(f), Column.longs(dec), Column.doubles(mx)]
        o.update(host_k, host_v, n, pack_bits(sel))
        gp.update_filtered(host_k, host_v, n, 0, thr)
        t = [torch.from_numpy(a).cuda() for a in (k1, k2, k3, f, dec, mx)]
        keep.append(t)
        dk = [vr.device_column(abi.INT32, t[0]), vr.device_column(abi.INT64, t[1]),
              vr.device_column(abi.UTF8, t[2], dictionary=dictionary, length=n)]
        dv = [vr.device_column(abi.INT32, t[3]), vr.device_column(abi.INT64, t[4]), vr.device_column(abi.FLOAT64, t[5])]
        dev_batches.append((dk, dv, n))
    gb.update_batches(dev_batches, filter_col=0, threshold=thr)
    for g in (gb, gp):
        compare(g, o, spec)


@pytest.mark.parametrize("batched", [False, True])
def test_long_batch_over_few_groups_is_split(dev, batched):
    """One 9M-row batch over 300 groups: the table stays sized by its groups and the batch is cut
    into several launches (double MIN/MAX with NaN, infinities and signed zeros, sums): results
    equal the oracle's as for one launch."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(107)
    n = 9_000_000
    pool = rng.integers(-2 ** 63, 2 ** 63 - 1, 300, dtype=np.int64)
    keys = pool[rng.integers(0, 300, n)]
    dbl = rng.normal(0, 1, n)
    special = rng.random(n) < 0.001
    dbl[special] = rng.choice([np.nan, 0.0, -0.0, np.inf, -np.inf], int(special.sum()))
    i64 = rng.integers(-10 ** 12, 10 ** 12, n)
    spec = abi.AggSpec([abi.INT64], [abi.FLOAT64, abi.INT64],
                       [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.MIN, 0), abi.agg_desc(abi.MAX, 0),
                        abi.agg_desc(abi.SUM_LONG, 1), abi.agg_desc(abi.SUM_DOUBLE, 0)],
                       key_nullable=[False], capacity_hint=300)
    tk, td, ti = (torch.from_numpy(a).cuda() for a in (keys, dbl, i64))
    g = vr.GroupedAggregate(dev, spec)
    per_tile = abi.launch_config(partitioned=-1)  # the per-tile kernel's launch splitting
    g.set_launch_config(per_tile)
    dk, dv = [vr.device_column(abi.INT64, tk)], [vr.device_column(abi.FLOAT64, td), vr.device_column(abi.INT64, ti)]
    if batched:
        g.update_batches([(dk, dv, n)], config=per_tile)
    else:
        g.update(dk, dv, n)
    o = OracleAgg(spec)
    o.update([Column.longs(keys)], [Column.doubles(dbl), Column.longs(i64)], n)
    # double sums within 1e-12 of the group's sum of magnitudes (the non-finite values are compared
    # as such before any tolerance)
    uk, inv = np.unique(keys, return_inverse=True)
    mag = np.bincount(inv, np.where(np.isfinite(dbl), np.abs(dbl), 0.0))
    scale = {(int(k),): [0.0, 0.0, 0.0, 0.0, float(m)] for k, m in zip(uk, mag)}
    assert compare(g, o, spec, rel=1e-12, scale=scale) < 1e-12
    assert g.memory() < 2 ** 30  # sized by its groups plus one launch's room, not by 9M rows


@pytest.mark.parametrize("batched", [False, True])
def test_guarded_launch_stops_grows_and_resumes(dev, batched):
    """6M rows over 4.5M distinct keys in one call: the launch passes the table's load limit part way,
    its workgroups stop, the table grows and the same grid resumes where each stopped -- every row
    folded once (COUNT(*) total, sums, double MIN/MAX with NaNs as the oracle)."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(108)
    n, distinct = 6_000_000, 4_500_000
    ids = np.concatenate([rng.permutation(distinct), rng.integers(0, distinct, n - distinct)])
    keys = (ids.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15) ^ np.uint64(0x5851F42D4C957F2D)).astype(np.int64)
    dbl = rng.normal(0, 1, n)
    dbl[rng.random(n) < 0.01] = np.nan
    i64 = rng.integers(-10 ** 12, 10 ** 12, n)
    spec = abi.AggSpec([abi.INT64], [abi.FLOAT64, abi.INT64],
                       [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.MIN, 0), abi.agg_desc(abi.MAX, 0),
                        abi.agg_desc(abi.SUM_LONG, 1)],
                       key_nullable=[False], capacity_hint=1000)
    tk, td, ti = (torch.from_numpy(a).cuda() for a in (keys, dbl, i64))
    grows0 = dev.stats()["table_grows"]
    g = vr.GroupedAggregate(dev, spec)
    per_tile = abi.launch_config(partitioned=-1)  # the guarded launch of the per-tile kernel
    g.set_launch_config(per_tile)
    dk, dv = [vr.device_column(abi.INT64, tk)], [vr.device_column(abi.FLOAT64, td), vr.device_column(abi.INT64, ti)]
    if batched:
        g.update_batches([(dk, dv, n)], config=per_tile)
    else:
        g.update(dk, dv, n)
    assert dev.stats()["table_grows"] - grows0 >= 2  # sized for the guard's room, then again mid-launch
    o = OracleAgg(spec)
    o.update([Column.longs(keys)], [Column.doubles(dbl), Column.longs(i64)], n)
    compare(g, o, spec)
