"""MIN/MAX over DECIMAL128 lanes on the device against the CPU oracle (MI355X).

Decimal128MinMax (kernels/src/main/java/io/vecruntime/kernels/GroupedAccumulators.java:1357-1462)
keeps the first value and replaces it only with a strictly better one by signed 128-bit order
(Decimal128.compare, Decimal128.java:145); a group with no value emits null.  The device has no
128-bit atomics: the update records each row's slot, sorts the rows by slot and folds each group's
rows in row order after the kernel.  Parity bar: bit-exact (the oracle itself is pinned against
exact Python integers in tests/test_oracle.py)."""
import numpy as np
import pytest

from gpu_util import compare_groups, group_rows, out_roles
from oracle.oracle import OracleAgg
from vecruntime_amd import abi
from vecruntime_amd.columns import Column, pack_bits

pytestmark = pytest.mark.gpu

INT64_MIN, INT64_MAX = -(1 << 63), (1 << 63) - 1


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    from vecruntime_amd import vr
    d = vr.Device(0)
    yield d
    d.close()


def decimal_limbs(rng, n):
    """Low and high limbs: random 128-bit values, with the extremes, a shared high limb whose low
    limbs straddle 2^63, negative values and zero mixed in."""
    lo = rng.integers(INT64_MIN, INT64_MAX, n, dtype=np.int64, endpoint=True)
    hi = rng.integers(INT64_MIN, INT64_MAX, n, dtype=np.int64, endpoint=True) >> rng.integers(0, 64, n)
    pick = rng.random(n)
    special = [(0, INT64_MIN), (-1, INT64_MAX), (0, 0), (-1, -1), (INT64_MIN + 5, 7), (3, 7), (INT64_MIN + 5, -7),
               (3, -7)]
    for j, (l, h) in enumerate(special):
        m = (pick >= 0.04 * j) & (pick < 0.04 * (j + 1))
        lo[m] = l
        hi[m] = h
    return lo, hi


def decimal_column(lo, hi, nulls=None):
    data = np.empty(2 * len(lo), np.int64)
    data[0::2] = lo
    data[1::2] = hi
    return Column(abi.DECIMAL128, data, None if nulls is None else pack_bits(~np.asarray(nulls, bool)), length=len(lo))


def spec_of(deterministic=False, ungrouped=False):
    # (no decimal SUM beside them: full-range values make 128-bit totals that overflow and come
    # back, where the device's exact final total and the reference's sticky flag differ by design)
    aggs = [abi.agg_desc(abi.MIN, 0), abi.agg_desc(abi.MAX, 0), abi.agg_desc(abi.COUNT, 0),
            abi.agg_desc(abi.SUM_DOUBLE, 1), abi.agg_desc(abi.MAX, 2), abi.agg_desc(abi.COUNT_STAR)]
    return abi.AggSpec([] if ungrouped else [abi.INT64], [abi.DECIMAL128, abi.FLOAT64, abi.INT32], aggs,
                       key_nullable=None if ungrouped else [True], deterministic=deterministic)


def make_batches(rng, lengths, distinct):
    out = []
    for bi, n in enumerate(lengths):
        keys = (rng.zipf(1.3, n) % distinct).astype(np.int64) * 1_000_003
        lo, hi = decimal_limbs(rng, n)
        nulls = rng.random(n) < 0.1
        nulls |= keys == 5 * 1_000_003  # a group of nulls only: null MIN/MAX
        knull = rng.random(n) < 0.01
        dbl = rng.standard_normal(n)
        i32 = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
        sel = pack_bits(rng.random(n) < 0.7) if bi % 2 else None
        out.append(([Column.longs(keys, nulls=knull)], [decimal_column(lo, hi, nulls), Column.doubles(dbl),
                                                       Column.ints(i32)], n, sel))
    return out


def check(g, o, spec):
    G = o.num_groups()
    assert g.num_groups() == G
    gk, go = g.emit()
    ok, oo = o.emit()
    worst = compare_groups(group_rows(gk, go, G, spec, {}), group_rows(ok, oo, G, spec, {}), out_roles(spec))
    assert worst < 1e-12
    return gk, go, ok, oo, G


@pytest.mark.parametrize("deterministic", [False, True])
def test_decimal_min_max_match_oracle(dev, deterministic):
    from vecruntime_amd import vr
    rng = np.random.default_rng(1357 + deterministic)
    spec = spec_of(deterministic)
    batches = make_batches(rng, [70_000, 1, 0, 130_001, 5], 3000)
    g, o = vr.GroupedAggregate(dev, spec), OracleAgg(spec)
    for keys, values, n, sel in batches:
        if n == 0:
            continue
        g.update(keys, values, n, sel)
        o.update(keys, values, n, sel)
    gk, go, ok, oo, G = check(g, o, spec)
    mins = go[0].values(G)
    assert any(v is None for v in mins) and any(v is not None and v < 0 for v in mins)
    if deterministic:  # first-seen group order and every buffer bit for bit
        assert (gk[0].data[:G] == ok[0].data[:G]).all()
        for a, b in zip(go, oo):
            w = 2 * G if a.type == abi.DECIMAL128 else G
            assert (a.data[:w] == b.data[:w]).all() and (a.valid(G) == b.valid(G)).all()
    g.close()


def test_decimal_min_max_batched_filtered_and_merge(dev):
    """Device-resident batches through vr_agg_update_batches (per-batch launches for this spec),
    the fused filter, and MERGE_MIN/MERGE_MAX over decimal buffer columns (Final over partials)."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(24)
    spec = spec_of()
    batches = make_batches(rng, [40_000, 9_999, 65_536], 500)

    def dev_col(c):
        v = None if c.validity is None else torch.from_numpy(c.validity).cuda()
        return vr.device_column(c.type, torch.from_numpy(c.data).cuda(), v, length=c.length)

    dbs = [([dev_col(c) for c in k], [dev_col(c) for c in v], n) for k, v, n, _ in batches]
    g, o = vr.GroupedAggregate(dev, spec), OracleAgg(spec)
    g.update_batches(dbs)
    for keys, values, n, _ in batches:
        o.update(keys, values, n)
    check(g, o, spec)
    g.close()
    # the filter: rows whose INT32 column is below the threshold
    g, o = vr.GroupedAggregate(dev, spec), OracleAgg(spec)
    for keys, values, n, _ in batches:
        g.update_filtered(keys, values, n, 2, 0)
        keep = np.asarray(values[2].data) < 0
        o.update(keys, values, n, pack_bits(keep))
    check(g, o, spec)
    g.close()
    # merge: the partial MIN/MAX buffers of two partitions folded by MERGE_MIN/MERGE_MAX
    mspec = abi.AggSpec([abi.INT64], [abi.DECIMAL128, abi.DECIMAL128],
                        [abi.agg_desc(abi.MERGE_MIN, 0), abi.agg_desc(abi.MERGE_MAX, 1)], key_nullable=[True])
    gm, om = vr.GroupedAggregate(dev, mspec), OracleAgg(mspec)
    for part in (batches[:2], batches[2:]):
        pspec = spec_of()
        po = OracleAgg(pspec)
        for keys, values, n, sel in part:
            po.update(keys, values, n, sel)
        G = po.num_groups()
        pk, pouts = po.emit()
        kcol = Column.longs(pk[0].data[:G], nulls=~pk[0].valid(G))
        mn = Column(abi.DECIMAL128, pouts[0].data[:2 * G].copy(), pouts[0].validity.copy(), length=G)
        mx = Column(abi.DECIMAL128, pouts[1].data[:2 * G].copy(), pouts[1].validity.copy(), length=G)
        gm.update([kcol], [mn, mx], G)
        om.update([kcol], [mn, mx], G)
    check(gm, om, mspec)
    gm.close()


def test_decimal_min_max_ungrouped(dev):
    from vecruntime_amd import vr
    rng = np.random.default_rng(7)
    spec = spec_of(ungrouped=True)
    g, o = vr.GroupedAggregate(dev, spec), OracleAgg(spec)
    gk, go = g.emit()  # the empty input's one row: null MIN/MAX
    assert go[0].values(1) == [None] and go[1].values(1) == [None]
    for _, values, n, sel in make_batches(rng, [10_000, 3], 1):
        g.update([], values, n, sel)
        o.update([], values, n, sel)
    gk, go = g.emit()
    ok, oo = o.emit()
    for a, b in zip(go, oo):
        assert a.values(1) == b.values(1) or a.type == abi.FLOAT64
    g.close()
