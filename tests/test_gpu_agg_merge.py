"""Merge modes on the device -- PartialMerge / Final over Spark's partial buffers
(VectorAggregates.compileMerge, VectorAggregates.scala:1141-1250) -- against the CPU oracle (MI355X).

Parity bar: counts, integer and decimal sums, MIN/MAX (doubles bit for bit), isEmpty flags and
nulls bit-exact; double sums bit-exact in deterministic mode, within 1e-12 of the sum of
magnitudes in atomic mode.  Not compared: a merged wide total that overflows 128 bits and comes
back, where the reference's flag is sticky and the device checks the final total."""
import numpy as np
import pytest

from gpu_util import compare_groups, group_rows, out_roles
from merge_cases import as_input, columns, partial_batches, specs, table_data
from oracle.oracle import OracleAgg
from test_oracle_merge import CASES, E21, group_scale, merge_rows, scales_for, ung_avg_spec, wide_batch, wide_spec
from vecruntime_amd import abi
from vecruntime_amd.columns import Column

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    from vecruntime_amd import vr
    d = vr.Device(0)
    yield d
    d.close()


def rows(agg, spec):
    k, o = agg.emit()
    return group_rows(k, o, agg.num_groups(), spec, {})


def device_agg(dev, spec):
    from vecruntime_amd import vr
    return vr.GroupedAggregate(dev, spec)


@pytest.mark.parametrize("keyed", [True, False])
@pytest.mark.parametrize("deterministic", [False, True])
def test_final_over_partials_matches_oracle(dev, keyed, deterministic):
    rng = np.random.default_rng(40 + 2 * keyed + deterministic)
    n = 20000
    k, knull, vals, nulls = table_data(rng, n, groups=300)
    kt = [abi.INT64] if keyed else []
    partial, _, merge = specs(kt, deterministic=deterministic)
    parts = partial_batches(OracleAgg, partial, k, knull, vals, nulls, [0, 3000, 3001, 11000, 20000], keyed)
    print("final over oracle partials: keyed=%s det=%s" % (keyed, deterministic), flush=True)
    g, o = device_agg(dev, merge), OracleAgg(merge)
    for kin, bin_, G in parts:
        g.update(kin, bin_, G)
        o.update(kin, bin_, G)
    ref = rows(o, merge)
    sc = scales_for(merge, group_scale(k, knull, vals[1], nulls[1], keyed), ref.keys())
    compare_groups(rows(g, merge), ref, out_roles(merge), rel=0.0 if deterministic else 1e-12,
                   scale=None if deterministic else sc)
    # PartialMerge: the merged buffers themselves
    _, _, pm = specs(kt, result=False, deterministic=deterministic)
    gp, op = device_agg(dev, pm), OracleAgg(pm)
    for kin, bin_, G in parts:
        gp.update(kin, bin_, G)
        op.update(kin, bin_, G)
    ref = rows(op, pm)
    compare_groups(rows(gp, pm), ref, out_roles(pm), rel=0.0 if deterministic else 1e-12,
                   scale=None if deterministic else scales_for(pm, group_scale(k, knull, vals[1], nulls[1], keyed), ref.keys()))
    for a in (g, gp):
        a.close()


def test_device_partials_then_device_final_equal_complete(dev):
    """Both stages on the device: Partial per partition, its buffers emitted and merged by Final,
    against the oracle's Complete over all rows."""
    rng = np.random.default_rng(77)
    n = 30000
    k, knull, vals, nulls = table_data(rng, n, groups=500)
    partial, complete, merge = specs([abi.INT64])
    print("device partials -> device final", flush=True)
    parts = partial_batches(lambda s: device_agg(dev, s), partial, k, knull, vals, nulls, [0, 10000, 17000, 30000])
    g = device_agg(dev, merge)
    for kin, bin_, G in parts:
        g.update(kin, bin_, G)
    o = OracleAgg(complete)
    keys, cols = columns(k, knull, vals, nulls, 0, n)
    o.update(keys, cols, n)
    ref = rows(o, complete)
    sc = scales_for(merge, group_scale(k, knull, vals[1], nulls[1], True), ref.keys())
    compare_groups(rows(g, merge), ref, out_roles(merge), scale=sc)
    g.close()


def test_batched_device_buffers_match_oracle(dev):
    """The partial buffers as device columns through one batched call (the gated isEmpty reads of
    the merged decimal sum run in the persistent kernel)."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(91)
    n = 40000
    k, knull, vals, nulls = table_data(rng, n, groups=2000)
    partial, _, merge = specs([abi.INT64])
    bounds = [0, 6000, 13000, 21000, 30000, 40000]
    parts = partial_batches(OracleAgg, partial, k, knull, vals, nulls, bounds)
    keep, batches = [], []

    def on_device(c):
        data = torch.from_numpy(np.ascontiguousarray(c.data)).cuda()
        valid = torch.from_numpy(np.ascontiguousarray(c.validity)).cuda() if c.validity is not None else None
        keep.append((data, valid))
        return vr.device_column(c.type, data, valid, length=c.length)

    o = OracleAgg(merge)
    for kin, bin_, G in parts:
        batches.append(([on_device(c) for c in kin], [on_device(c) for c in bin_], G))
        o.update(kin, bin_, G)
    print("batched merge over device buffers: %d batches" % len(batches), flush=True)
    g = device_agg(dev, merge)
    g.update_batches(batches)
    ref = rows(o, merge)
    sc = scales_for(merge, group_scale(k, knull, vals[1], nulls[1], True), ref.keys())
    compare_groups(rows(g, merge), ref, out_roles(merge), scale=sc)
    g.close()


@pytest.mark.parametrize("result", [True, False])
def test_wide_merge_known_answers(dev, result):
    spec = wide_spec(result)
    g, o = device_agg(dev, spec), OracleAgg(spec)
    g.update(*wide_batch(CASES))
    o.update(*wide_batch(CASES))
    compare_groups(rows(g, spec), rows(o, spec), out_roles(spec), rel=0.0)
    g.close()


@pytest.mark.parametrize("recs,status", [
    ([(3, None, False, 1, 1)], abi.VR_ERR_ARITHMETIC_OVERFLOW),
    ([(5, E21, False, 1, 1), (5, E21, False, 1, 1)], abi.VR_ERR_DECIMAL_OVERFLOW),
    ([(3, 1, False, None, 4)], abi.VR_ERR_ARITHMETIC_OVERFLOW),
    ([(5, 1, False, E21, 1), (5, 1, False, E21, 1)], abi.VR_ERR_ARITHMETIC_OVERFLOW),
])
def test_wide_merge_ansi_errors(dev, recs, status):
    from vecruntime_amd import vr
    g = device_agg(dev, wide_spec(True, ansi=True))
    g.update(*wide_batch(recs))
    with pytest.raises(vr.VrError) as e:
        g.emit()
    assert e.value.status == status
    g.close()


def test_ungrouped_exact_final_and_average(dev):
    from vecruntime_amd import vr
    n, x = 15000, 9 * 10 ** 33
    spec = ung_avg_spec(False, False)
    u = device_agg(dev, spec)
    u.update([], [Column.decimals([x] * n)], n)
    assert next(iter(rows(u, spec).values()))[0] == x * 10 ** 4
    u.close()
    gspec = ung_avg_spec(True, False)
    g = device_agg(dev, gspec)
    g.update([Column.longs(np.ones(n, np.int64))], [Column.decimals([x] * n)], n)
    assert rows(g, gspec)[(1,)][0] is None
    g.close()
    for recs, expect in [([(8 * 10 ** 37, 10 ** 4)] * 2, 8 * 10 ** 37),
                         ([(9 * 10 ** 37, 10 ** 4), (9 * 10 ** 37, 10 ** 4), (-9 * 10 ** 37, 10 ** 4)], 3 * 10 ** 37)]:
        mspec = ung_avg_spec(False, True)
        m = device_agg(dev, mspec)
        m.update(*merge_rows(recs, False))
        assert next(iter(rows(m, mspec).values()))[0] == expect
        m.close()
    m = device_agg(dev, ung_avg_spec(False, True))
    m.update(*merge_rows([(9 * 10 ** 37, 10 ** 4)] * 2, False))
    with pytest.raises(vr.VrError) as e:
        m.emit()
    assert e.value.status == abi.VR_ERR_UNSUPPORTED
    m.close()


def test_merge_spec_validation(dev):
    from vecruntime_amd import vr
    bad = [
        abi.AggSpec([abi.INT64], [abi.DECIMAL128, abi.INT64], [abi.agg_desc(abi.MERGE_SUM_WIDE, 0, 1, 0, 22, 2)]),
        abi.AggSpec([abi.INT64], [abi.DECIMAL128], [abi.agg_desc(abi.MERGE_AVG_WIDE, 0, 1, 0, 22, 2)]),
        abi.AggSpec([abi.INT64], [abi.INT32], [abi.agg_desc(abi.MERGE_COUNT, 0)]),
    ]
    for spec in bad:
        with pytest.raises(vr.VrError) as e:
            device_agg(dev, spec)
        assert e.value.status == abi.VR_ERR_INVALID_ARGUMENT
