"""Ungrouped aggregation (no grouping key: AggKernels' reductions, AggKernels.java:133-621) on the
device against the CPU oracle (MI355X).

The device folds every selected row into one group keyed by a constant column it supplies; the
handle holds that group from creation on, so an empty input emits its one row.  Parity bar: counts,
integer and decimal sums and MIN/MAX bit-exact; double sums within 1e-12 of the sum of magnitudes
(atomic mode) or bit-exact (deterministic mode)."""
import numpy as np
import pytest

from gpu_util import compare_groups, group_rows, out_roles
from oracle.oracle import OracleAgg
from vecruntime_amd import abi
from vecruntime_amd.columns import Column, pack_bits

pytestmark = pytest.mark.gpu

TYPES = [abi.INT64, abi.FLOAT64, abi.INT32]
AGGS = [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.COUNT, 0), abi.agg_desc(abi.SUM_LONG, 0),
        abi.agg_desc(abi.SUM_DOUBLE, 1), abi.agg_desc(abi.AVG_DOUBLE, 1), abi.agg_desc(abi.MIN, 2),
        abi.agg_desc(abi.MAX, 2), abi.agg_desc(abi.SUM_WIDE, 0, buffer_precision=22, buffer_scale=2),
        abi.agg_desc(abi.MIN, 1), abi.agg_desc(abi.MAX, 0)]


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    from vecruntime_amd import vr
    d = vr.Device(0)
    yield d
    d.close()


def columns(rng, n):
    v = rng.integers(-10 ** 9, 10 ** 9, n)
    d = rng.standard_normal(n)
    m = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
    nulls = [rng.random(n) < 0.1 for _ in range(3)]
    return [Column.longs(v, nulls=nulls[0]), Column.doubles(d, nulls=nulls[1]), Column.ints(m, nulls=nulls[2])], d, nulls


def rows(agg, spec):
    k, o = agg.emit()
    return group_rows(k, o, agg.num_groups(), spec, {})


def mags(d, nulls, sel):
    """The double columns' bar: sum of |x| over the rows folded."""
    m = float(np.abs(d[(~nulls[1]) & sel]).sum())
    return m


@pytest.mark.parametrize("deterministic", [False, True])
def test_batches_with_selection_match_oracle(dev, deterministic):
    from vecruntime_amd import vr
    rng = np.random.default_rng(3 + deterministic)
    spec = abi.AggSpec([], TYPES, AGGS, deterministic=deterministic)
    g, o = vr.GroupedAggregate(dev, spec), OracleAgg(spec)
    scale = 0.0
    for n in (1, 63, 1000, 65536, 70001):
        cols, d, nulls = columns(rng, n)
        selb = rng.random(n) < 0.7
        sel = pack_bits(selb)
        g.update([], cols, n, sel)
        o.update([], cols, n, sel)
        scale += mags(d, nulls, selb)
    assert g.num_groups() == 1 == o.num_groups()
    roles = out_roles(spec)
    sc = {(): [scale] * len(roles)}
    compare_groups(rows(g, spec), rows(o, spec), roles, rel=0.0 if deterministic else 1e-12,
                   scale=None if deterministic else sc)
    g.close()


def test_empty_input_and_reset_emit_one_row(dev):
    from vecruntime_amd import vr
    spec = abi.AggSpec([], TYPES, AGGS)
    g, o = vr.GroupedAggregate(dev, spec), OracleAgg(spec)
    assert g.num_groups() == 1
    compare_groups(rows(g, spec), rows(o, spec), out_roles(spec))
    rng = np.random.default_rng(8)
    cols, _, _ = columns(rng, 500)
    g.update([], cols, 500)
    g.reset()
    assert g.num_groups() == 1
    compare_groups(rows(g, spec), rows(o, spec), out_roles(spec))  # identity again: COUNT 0, nulls
    g.close()


def test_large_device_batches_and_filter(dev):
    """6M device rows in 65,536-row batches through one batched call (the partitioned update, every
    row in the hot set), then a filtered call, against the oracle."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(21)
    n, bs = 6_000_000, 65536
    v = rng.integers(-10 ** 6, 10 ** 6, n)
    d = rng.random(n)
    m = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
    nl = rng.random(n) < 0.05
    spec = abi.AggSpec([], TYPES, AGGS[:8])  # state rows that fit the partitioned update's LDS
    g, o = vr.GroupedAggregate(dev, spec), OracleAgg(spec)
    tv, td, tm = (torch.from_numpy(x).cuda() for x in (v, d, m))
    tn = torch.from_numpy(pack_bits(~nl)).cuda()
    batches = []
    for s0 in range(0, n, bs):
        k = min(bs, n - s0)
        batches.append(([], [vr.device_column(abi.INT64, tv[s0:s0 + k], tn[s0 // 8:], length=k),
                             vr.device_column(abi.FLOAT64, td[s0:s0 + k]), vr.device_column(abi.INT32, tm[s0:s0 + k])], k))
    st = dev.stats()["partitioned_updates"]
    g.update_batches(batches)
    assert dev.stats()["partitioned_updates"] == st + 1
    o.update([], [Column.longs(v, nulls=nl), Column.doubles(d), Column.ints(m)], n)
    # the filter: rows with the INT32 lane below a threshold
    thr = 0
    h = 200_000
    g.update_filtered([], [Column.longs(v[:h], nulls=nl[:h]), Column.doubles(d[:h]), Column.ints(m[:h])], h, 2, thr)
    o.update([], [Column.longs(v[:h], nulls=nl[:h]), Column.doubles(d[:h]), Column.ints(m[:h])], h,
             pack_bits(m[:h] < thr))
    roles = out_roles(spec)
    scale = float(np.abs(d).sum() + np.abs(d[:h]).sum())
    compare_groups(rows(g, spec), rows(o, spec), roles, scale={(): [scale] * len(roles)})
    g.close()
