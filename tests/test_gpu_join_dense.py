"""The join's direct key index (DenseKeyIndex, kernels/src/main/java/io/vecruntime/kernels/
DenseKeyIndex.java:34-177) on the device (MI355X).

A build over one INT32/INT64 key whose non-null values span at most 2^22 and at most 10x the
distinct keys is probed through slots[key - min] instead of the hash table; everything the probe
returns -- pairs in order, semi, anti, the null-aware anti, materialized columns -- is the same.
Each case runs with the index (dense_index 0) and without it (-1) and against the oracle; the
device statistics show which builds took it."""
import numpy as np
import pytest

from oracle.oracle import OracleJoin
from vecruntime_amd import abi
from vecruntime_amd.columns import Column, pack_bits

pytestmark = pytest.mark.gpu

MARKER = 0x7FF4DEADBEEF7FF4  # the hash table's empty-cell marker, as a key value


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    from vecruntime_amd import vr
    d = vr.Device(0)
    yield d
    d.close()


def column(t, vals, nulls=None):
    return Column.ints(vals, nulls=nulls) if t == abi.INT32 else Column.longs(vals, nulls=nulls)


def run_all(dev, t, bkeys, bnull, pkeys, pnull, psel, expect_dense):
    """Every join type through the dense build, the hash build and the oracle; identical results."""
    from vecruntime_amd import vr
    nb, npr = len(bkeys), len(pkeys)
    bcols = [column(t, bkeys, bnull), Column.ints(np.arange(nb, dtype=np.int32))]
    pcols = [column(t, pkeys, pnull), Column.ints(np.arange(npr, dtype=np.int32) * 3)]
    sel = pack_bits(psel) if psel is not None else None
    results = []
    for mode in (0, -1):
        spec = abi.JoinSpec([t], [t, abi.INT32], [0], dense_index=mode)
        before = dev.stats()["dense_join_builds"]
        bt = vr.BuildTable(dev, spec)
        bt.append(bcols, nb)
        bt.finish()
        took = dev.stats()["dense_join_builds"] - before
        assert took == (1 if (mode == 0 and expect_dense) else 0), (mode, took)
        out = []
        for jt, dn in ((abi.INNER, False), (abi.LEFT_SEMI, False), (abi.LEFT_ANTI, False), (abi.LEFT_ANTI, True)):
            bo = [1] if jt == abi.INNER else []
            p, b, outs, r = bt.probe(pcols, [0], npr, jt, sel, drop_null_keys=dn, probe_out=[1], build_out=bo)
            out.append((p.tolist(), None if b is None else b.tolist(), [o.data[:r].tolist() for o in outs], r))
        results.append(out)
        bt.close()
    assert results[0] == results[1]
    spec = abi.JoinSpec([t], [t, abi.INT32], [0])
    oj = OracleJoin(spec, bcols, nb)
    for (p, b, outs, r), (jt, dn) in zip(results[0], ((abi.INNER, False), (abi.LEFT_SEMI, False), (abi.LEFT_ANTI, False),
                                                      (abi.LEFT_ANTI, True))):
        bo = [1] if jt == abi.INNER else []
        op, ob, oouts, orows = oj.probe(pcols, [0], npr, jt, sel, drop_null_keys=dn, probe_out=[1], build_out=bo,
                                        capacity=max(r, 1))
        assert r == orows, (jt, dn)
        assert p == op.tolist()
        if jt == abi.INNER:
            assert b == ob.tolist()
        for a, c in zip(outs, oouts):
            assert a == c.data[:r].tolist()


@pytest.mark.parametrize("t", [abi.INT32, abi.INT64])
def test_dense_small_range_with_duplicates_nulls_and_misses(dev, t):
    rng = np.random.default_rng(17 + t)
    lo = -20000
    nb, npr = 60000, 200000
    bkeys = rng.integers(lo, lo + 50000, nb)          # ~35K distinct keys over a 50K range
    bnull = rng.random(nb) < 0.02
    edge = np.array([lo - 1, lo, lo + 49999, lo + 50000, -2 ** 31, 2 ** 31 - 1], np.int64)
    pkeys = np.concatenate([rng.integers(lo - 1000, lo + 51000, npr - len(edge)), edge])
    if t == abi.INT64:
        pkeys[: 100] = rng.integers(-2 ** 63, 2 ** 63 - 1, 100, dtype=np.int64)  # offsets that wrap
    pnull = rng.random(npr) < 0.02
    psel = rng.random(npr) < 0.9
    run_all(dev, t, bkeys, bnull, pkeys, pnull, psel, expect_dense=True)


def test_dense_covers_the_empty_marker_key(dev):
    """INT64 keys around the table's empty marker: the marker's own entry sits in the range."""
    rng = np.random.default_rng(5)
    bkeys = MARKER + rng.integers(-500, 500, 3000)
    bkeys[:3] = MARKER
    pkeys = MARKER + rng.integers(-600, 600, 20000)
    pkeys[:10] = MARKER
    run_all(dev, abi.INT64, bkeys, None, pkeys, None, None, expect_dense=True)


@pytest.mark.parametrize("case", ["ten_times", "past_ten_times", "max_range", "past_max_range", "full_span"])
def test_dense_eligibility_bounds(dev, case):
    """Range <= 10 x distinct keys and <= 2^22 slots take the index; one more does not; a span
    past the long range (min and max of int64) must not overflow the check."""
    rng = np.random.default_rng(9)
    if case in ("ten_times", "past_ten_times"):
        g = 1000
        top = 10 * g - 1 if case == "ten_times" else 10 * g
        bkeys = np.concatenate([np.array([0, top]), rng.choice(np.arange(1, top), g - 2, replace=False)])
        expect = case == "ten_times"
    elif case in ("max_range", "past_max_range"):
        g = (1 << 22) // 10 + 1
        top = (1 << 22) - 1 if case == "max_range" else 1 << 22
        bkeys = np.concatenate([np.array([0, top]), rng.choice(np.arange(1, top), g - 2, replace=False)])
        expect = case == "max_range"
    else:
        bkeys = np.array([-2 ** 63, 2 ** 63 - 1, 0, 1, 2], np.int64)
        expect = False
    bkeys = bkeys.astype(np.int64)
    pkeys = np.concatenate([bkeys[: 5000], rng.integers(-10, min(int(bkeys.max()), 2 ** 40) + 10, 20000)]).astype(np.int64)
    run_all(dev, abi.INT64, bkeys, None, pkeys, None, None, expect_dense=expect)
