"""Two-phase aggregation cases shared by the oracle's and the device's merge tests.

Spark splits an aggregate into a Partial stage per partition and a merging stage (PartialMerge /
Final) over the partial buffers (VectorAggregates.compileMerge, VectorAggregates.scala:1141-1250).
Whatever the partitioning, Final over the partial buffers equals Complete over all rows: the
invariant the reference's end-to-end suites rely on (VectorAggregateSuite.scala:255-296 runs
multi-stage plans against Spark's answers).  Exact for counts, integer and decimal sums and
MIN/MAX; double sums are re-associated by the split, so they agree within a tolerance."""
import numpy as np

from vecruntime_amd import abi
from vecruntime_amd.columns import Column

VALUE_TYPES = [abi.INT64, abi.FLOAT64, abi.INT32, abi.DECIMAL128]

# (update-mode aggregate, merge kind, number of buffer columns)
PAIRS = [
    (abi.agg_desc(abi.COUNT_STAR), abi.MERGE_COUNT, 1),
    (abi.agg_desc(abi.COUNT, 0), abi.MERGE_COUNT, 1),
    (abi.agg_desc(abi.SUM_LONG, 0), abi.MERGE_SUM_LONG, 1),
    (abi.agg_desc(abi.SUM_DOUBLE, 1), abi.MERGE_SUM_DOUBLE, 1),
    (abi.agg_desc(abi.AVG_DOUBLE, 1), abi.MERGE_AVG_DOUBLE, 2),
    (abi.agg_desc(abi.SUM_WIDE, 0, buffer_precision=22, buffer_scale=2), abi.MERGE_SUM_WIDE, 2),
    (abi.agg_desc(abi.AVG_WIDE, 0, buffer_precision=22, buffer_scale=2, result_precision=16, result_scale=6),
     abi.MERGE_AVG_WIDE, 2),
    (abi.agg_desc(abi.MIN, 2), abi.MERGE_MIN, 1),
    (abi.agg_desc(abi.MAX, 1), abi.MERGE_MAX, 1),
    (abi.agg_desc(abi.SUM_WIDE, 3, buffer_precision=38, buffer_scale=2), abi.MERGE_SUM_WIDE, 2),
    (abi.agg_desc(abi.AVG_WIDE, 3, buffer_precision=38, buffer_scale=2, result_precision=32, result_scale=6),
     abi.MERGE_AVG_WIDE, 2),
]


def with_flags(d, flags):
    return abi.agg_desc(d.kind, d.input, d.input2, flags, d.buffer_precision, d.buffer_scale,
                        d.result_precision, d.result_scale)


def specs(key_types, result=True, ansi=False, deterministic=False):
    """(partial spec, complete spec, merge spec) over VALUE_TYPES; the merge spec reads the
    partial spec's output columns in order."""
    fl = abi.CHECKED if ansi else 0
    partial = abi.AggSpec(key_types, VALUE_TYPES, [with_flags(d, fl) for d, _, _ in PAIRS], deterministic=deterministic)
    complete = abi.AggSpec(key_types, VALUE_TYPES, [with_flags(d, fl | abi.RESULT) for d, _, _ in PAIRS],
                           deterministic=deterministic)
    merge_aggs, c = [], 0
    for d, mk, nb in PAIRS:
        merge_aggs.append(abi.agg_desc(mk, c, c + 1 if nb == 2 else -1, fl | (abi.RESULT if result else 0),
                                       d.buffer_precision, d.buffer_scale, d.result_precision, d.result_scale))
        c += nb
    merge = abi.AggSpec(key_types, partial.output_types(), merge_aggs, deterministic=deterministic)
    return partial, complete, merge


def table_data(rng, n, groups=50):
    """One nullable INT64 key over `groups` values and the four value lanes, with nulls, doubles
    holding signed zeros and NaN, decimals near their precision."""
    k = rng.integers(0, groups, n).astype(np.int64)
    knull = rng.random(n) < 0.05
    v0 = rng.integers(-10 ** 12 + 1, 10 ** 12, n).astype(np.int64)
    d = rng.standard_normal(n)
    pick = rng.random(n)
    d[pick < 0.02] = 0.0
    d[(pick >= 0.02) & (pick < 0.04)] = -0.0
    d[(pick >= 0.04) & (pick < 0.045)] = np.nan
    m = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
    big = [int(x) * 10 ** 18 + int(y) for x, y in zip(rng.integers(-10 ** 11, 10 ** 11, n), rng.integers(0, 10 ** 18, n))]
    nulls = [rng.random(n) < f for f in (0.05, 0.1, 0.05, 0.05)]
    return k, knull, [v0, d, m, big], nulls


def columns(k, knull, vals, nulls, lo, hi, keyed=True):
    keys = [Column.longs(k[lo:hi], nulls=knull[lo:hi])] if keyed else []
    v0, d, m, big = vals
    cols = [Column.longs(v0[lo:hi], nulls=nulls[0][lo:hi]), Column.doubles(d[lo:hi], nulls=nulls[1][lo:hi]),
            Column.ints(m[lo:hi], nulls=nulls[2][lo:hi]), Column.decimals(big[lo:hi], nulls=nulls[3][lo:hi])]
    return keys, cols


def as_input(oc, t, n):
    """An emitted column (OutColumn of n rows) as an input column of the next stage, bits intact."""
    if t == abi.BOOL:
        data = oc.data[: (n + 7) // 8 + 8].copy()
    elif t == abi.DECIMAL128:
        data = oc.data[: 2 * n].copy()
    else:
        data = oc.data[:n].copy()
    return Column(t, data, oc.validity.copy(), length=n)


def partial_batches(agg_factory, partial, k, knull, vals, nulls, bounds, keyed=True):
    """Runs the Partial stage once per partition [bounds[i], bounds[i+1]) and returns its outputs as
    merge batches: (key columns, buffer columns, rows)."""
    out = []
    for lo, hi in zip(bounds[:-1], bounds[1:]):
        a = agg_factory(partial)
        keys, cols = columns(k, knull, vals, nulls, lo, hi, keyed)
        a.update(keys, cols, hi - lo)
        G = a.num_groups()
        ko, bo = a.emit()
        kin = [as_input(c, t, G) for c, t in zip(ko, partial.key_types)]
        bin_ = [as_input(c, t, G) for c, t in zip(bo, partial.output_types())]
        out.append((kin, bin_, G))
        if hasattr(a, "close"):
            a.close()
    return out
