"""The oracle's merge modes (PartialMerge / Final over Spark's partial buffers) pinned by the
two-phase invariant and by closed-form cases.

* Final over the Partial stage's buffers, for any partitioning, equals Complete over all rows
  (VectorAggregates.compileMerge, VectorAggregates.scala:1141-1250); PartialMerge of a subset of
  buffers followed by Final equals Final over all of them.
* Known answers for the wide decimal merges: isEmpty rows skipped, a null sum poisoning the group,
  zero while empty, null past the buffer precision, ANSI errors (WideDecimalSumMergeAgg
  :688-762, WideDecimalAvgMergeAgg :562-609), and the exact ungrouped final (:583-604) beside the
  update mode's ungrouped average (WideDecimalAvgAgg.sumValue :455-467)."""
import numpy as np
import pytest

from gpu_util import compare_groups, group_rows, out_roles
from merge_cases import columns, partial_batches, specs, table_data
from oracle.oracle import OracleAgg, OracleError
from vecruntime_amd import abi
from vecruntime_amd.columns import Column


def rows(agg, spec):
    k, o = agg.emit()
    return group_rows(k, o, agg.num_groups(), spec, {})


def group_scale(k, knull, d, dnull, keyed):
    """Sum of |x| of the double lane per group key (the tolerance of re-associated sums)."""
    out = {}
    ok = ~dnull & ~np.isnan(d)
    for i in np.nonzero(ok)[0]:
        key = ((None if knull[i] else int(k[i])),) if keyed else ()
        out[key] = out.get(key, 0.0) + abs(float(d[i]))
    return out


def scales_for(spec, per_key, keys):
    n = len(out_roles(spec))
    return {key: [max(per_key.get(key, 0.0), 1e-300)] * n for key in keys}


@pytest.mark.parametrize("keyed", [True, False])
@pytest.mark.parametrize("bounds", [(0, 700, 2500, 2501, 5000), (0, 5000), (0, 1, 2, 3, 5000)])
def test_final_over_partials_equals_complete(keyed, bounds):
    rng = np.random.default_rng(11 + len(bounds) + keyed)
    k, knull, vals, nulls = table_data(rng, 5000)
    kt = [abi.INT64] if keyed else []
    partial, complete, merge = specs(kt)
    whole = OracleAgg(complete)
    keys, cols = columns(k, knull, vals, nulls, 0, 5000, keyed)
    whole.update(keys, cols, 5000)
    fin = OracleAgg(merge)
    for kin, bin_, G in partial_batches(OracleAgg, partial, k, knull, vals, nulls, list(bounds), keyed):
        fin.update(kin, bin_, G)
    ref = rows(whole, complete)
    got = rows(fin, merge)
    sc = scales_for(merge, group_scale(k, knull, vals[1], nulls[1], keyed), ref.keys())
    compare_groups(got, ref, out_roles(merge), scale=sc)


def test_partial_merge_then_final_equals_final():
    rng = np.random.default_rng(5)
    k, knull, vals, nulls = table_data(rng, 4000)
    partial, _, final = specs([abi.INT64])
    _, _, pmerge = specs([abi.INT64], result=False)
    parts = partial_batches(OracleAgg, partial, k, knull, vals, nulls, [0, 900, 2000, 3100, 4000])
    direct = OracleAgg(final)
    for kin, bin_, G in parts:
        direct.update(kin, bin_, G)
    # PartialMerge of the first three partitions' buffers, then Final over that and the fourth
    mid = OracleAgg(pmerge)
    for kin, bin_, G in parts[:3]:
        mid.update(kin, bin_, G)
    assert pmerge.output_types() == partial.output_types()
    from merge_cases import as_input
    G = mid.num_groups()
    ko, bo = mid.emit()
    two = OracleAgg(final)
    two.update([as_input(c, t, G) for c, t in zip(ko, pmerge.key_types)],
               [as_input(c, t, G) for c, t in zip(bo, pmerge.output_types())], G)
    two.update(*parts[3][:2], parts[3][2])
    ref = rows(direct, final)
    sc = scales_for(final, group_scale(k, knull, vals[1], nulls[1], True), ref.keys())
    compare_groups(rows(two, final), ref, out_roles(final), scale=sc)


# ---- closed forms for the wide decimal merges

WIDE_TYPES = [abi.DECIMAL128, abi.BOOL, abi.DECIMAL128, abi.INT64]


def wide_spec(result, ansi=False, keyed=True, bp=22, bs=2, rp=16, rs=6):
    fl = (abi.RESULT if result else 0) | (abi.CHECKED if ansi else 0)
    aggs = [abi.agg_desc(abi.MERGE_SUM_WIDE, 0, 1, fl, bp, bs),
            abi.agg_desc(abi.MERGE_AVG_WIDE, 2, 3, fl, bp, bs, rp, rs)]
    return abi.AggSpec([abi.INT64] if keyed else [], WIDE_TYPES, aggs)


def wide_batch(recs, keyed=True):
    """recs: (key, sum or None, isEmpty or None, avg sum or None, count)."""
    n = len(recs)
    keys = [Column.longs([r[0] for r in recs])] if keyed else []
    s = [0 if r[1] is None else r[1] for r in recs]
    e = [bool(r[2]) for r in recs]
    a = [0 if r[3] is None else r[3] for r in recs]
    vals = [Column.decimals(s, nulls=[r[1] is None for r in recs]),
            Column.bools(e, nulls=[r[2] is None for r in recs]),
            Column.decimals(a, nulls=[r[3] is None for r in recs]),
            Column.longs([r[4] for r in recs])]
    return keys, vals, n


E21 = 6 * 10 ** 21
CASES = [  # key, sum, isEmpty, avg sum, count
    (1, 100, False, 100, 2), (1, 250, False, 250, 3),                  # 350; 350/5
    (2, 999, True, 7, 0), (2, None, None, 3, 0), (2, 5, False, 0, 0),  # empty rows skipped; count 0
    (3, None, False, None, 4),                                         # poisoned
    (4, 0, True, 0, 0), (4, 12, True, 0, 0),                           # empty
    (5, E21, False, E21, 1), (5, E21, False, E21, 1),                  # past Decimal(22)
]


def test_wide_merge_known_answers():
    fin, pm = OracleAgg(wide_spec(True)), OracleAgg(wide_spec(False))
    for a in (fin, pm):
        a.update(*wide_batch(CASES))
    got = {k[0]: v for k, v in rows(fin, wide_spec(True)).items()}
    # Final: sum; avg = HALF_UP(total * 10^4 / count) at scale 6 (Decimal(16, 6) in an INT64 lane)
    assert got[1] == [350, 350 * 10 ** 4 // 5]
    assert got[2] == [5, None]      # isEmpty true / null rows skipped; count 0 nulls the average
    assert got[3] == [None, None]
    assert got[4] == [None, None]
    assert got[5] == [None, None]
    part = {k[0]: v for k, v in rows(pm, wide_spec(False)).items()}
    assert part[1] == [350, False, 350, 5]
    assert part[3] == [None, False, None, 4]
    assert part[4] == [0, True, 0, 0]
    assert part[5] == [None, False, None, 2]
    assert part[2] == [5, False, 10, 0]  # the average's sums add whatever the counts


@pytest.mark.parametrize("recs,status", [
    ([(3, None, False, 1, 1)], abi.VR_ERR_ARITHMETIC_OVERFLOW),  # a null partial sum: overflow in sum of decimals
    ([(5, E21, False, 1, 1), (5, E21, False, 1, 1)], abi.VR_ERR_DECIMAL_OVERFLOW),  # past Decimal(22)
])
def test_wide_sum_merge_ansi_errors(recs, status):
    a = OracleAgg(wide_spec(True, ansi=True))
    a.update(*wide_batch(recs))
    with pytest.raises(OracleError) as e:
        a.emit()
    assert e.value.status == status


@pytest.mark.parametrize("recs", [
    [(3, 1, False, None, 4)],                            # a null partial sum
    [(5, 1, False, E21, 1), (5, 1, False, E21, 1)],      # the merged total past Decimal(22)
])
def test_wide_avg_merge_ansi_overflow(recs):
    a = OracleAgg(wide_spec(True, ansi=True))
    a.update(*wide_batch(recs))
    with pytest.raises(OracleError) as e:
        a.emit()
    assert e.value.status == abi.VR_ERR_ARITHMETIC_OVERFLOW


def test_wide_sum_merge_overflow_is_sticky():
    """A signed 128-bit overflow of the merged total poisons the group for good, even if later
    rows bring the total back (WideDecimalSumMergeAgg.State.merge :703-706)."""
    big = 99 * 10 ** 36
    spec = wide_spec(True, bp=38, bs=0, rp=38, rs=4)
    a = OracleAgg(spec)
    a.update(*wide_batch([(1, big, False, 1, 1), (1, big, False, 1, 1), (1, -big, False, 1, 1)]))
    assert rows(a, spec)[(1,)][0] is None
    b = OracleAgg(spec)
    b.update(*wide_batch([(1, big, False, 1, 1), (1, -big, False, 1, 1), (1, big, False, 1, 1)]))
    assert rows(b, spec)[(1,)][0] == big


# ---- the exact ungrouped final

def ung_avg_spec(keyed, merge, ansi=False):
    fl = abi.RESULT | (abi.CHECKED if ansi else 0)
    if merge:
        return abi.AggSpec([abi.INT64] if keyed else [], [abi.DECIMAL128, abi.INT64],
                           [abi.agg_desc(abi.MERGE_AVG_WIDE, 0, 1, fl, 38, 0, 38, 4)])
    return abi.AggSpec([abi.INT64] if keyed else [], [abi.DECIMAL128],
                       [abi.agg_desc(abi.AVG_WIDE, 0, -1, fl, 38, 0, 38, 4)])


def one_value(agg, spec):
    r = rows(agg, spec)
    assert len(r) == 1
    return next(iter(r.values()))[0]


def test_ungrouped_average_divides_a_total_past_the_buffer():
    """15,000 rows of 9e33 in Decimal(38, 0): the total 1.35e38 is past the buffer's 10^38 but
    within 128 bits.  Ungrouped, the average is exact (9e33 at scale 4); grouped, the buffer is
    null (ANSI: the average of an overflowed sum)."""
    n, x = 15000, 9 * 10 ** 33
    cols = [Column.decimals([x] * n)]
    u = OracleAgg(ung_avg_spec(False, False))
    u.update([], cols, n)
    assert one_value(u, ung_avg_spec(False, False)) == x * 10 ** 4
    g = OracleAgg(ung_avg_spec(True, False))
    g.update([Column.longs(np.ones(n, np.int64))], cols, n)
    assert one_value(g, ung_avg_spec(True, False)) is None
    ga = OracleAgg(ung_avg_spec(True, False, ansi=True))
    ga.update([Column.longs(np.ones(n, np.int64))], cols, n)
    with pytest.raises(OracleError) as e:
        ga.emit()
    assert e.value.status == abi.VR_ERR_ARITHMETIC_OVERFLOW


def merge_rows(recs, keyed):
    keys = [Column.longs(np.ones(len(recs), np.int64))] if keyed else []
    return keys, [Column.decimals([r[0] for r in recs]), Column.longs([r[1] for r in recs])], len(recs)


@pytest.mark.parametrize("recs,expect", [
    ([(8 * 10 ** 37, 10 ** 4), (8 * 10 ** 37, 10 ** 4)], 8 * 10 ** 37),            # past 10^38, within 128 bits
    ([(9 * 10 ** 37, 10 ** 4), (9 * 10 ** 37, 10 ** 4), (-9 * 10 ** 37, 10 ** 4)], 3 * 10 ** 37),  # wraps and returns
])
def test_ungrouped_final_merge_is_exact(recs, expect):
    u = OracleAgg(ung_avg_spec(False, True))
    u.update(*merge_rows(recs, False))
    assert one_value(u, ung_avg_spec(False, True)) == expect
    g = OracleAgg(ung_avg_spec(True, True))
    g.update(*merge_rows(recs, True))
    assert one_value(g, ung_avg_spec(True, True)) is None  # grouped: null past the buffer / sticky


def test_ungrouped_final_merge_beyond_128_bits_is_left_to_the_cpu_kernels():
    u = OracleAgg(ung_avg_spec(False, True))
    u.update(*merge_rows([(9 * 10 ** 37, 10 ** 4), (9 * 10 ** 37, 10 ** 4)], False))
    with pytest.raises(OracleError) as e:
        u.emit()
    assert e.value.status == abi.VR_ERR_UNSUPPORTED
