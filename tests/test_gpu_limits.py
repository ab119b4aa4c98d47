"""Run-time limits of the device path that the caller answers by running the CPU kernels (MI355X).

An exhausted HBM budget returns VR_ERR_OOM from the call that needed the memory (a handle's
creation, an update that grows the table, a join build), after which the device serves new work
as before.  An aggregate whose integer or decimal sums are kept as sums of 32-bit pieces refuses an
update that would bring 2^32 rows between resets (VR_ERR_UNSUPPORTED), before reading any row."""
import numpy as np
import pytest

from oracle.oracle import OracleAgg
from vecruntime_amd import abi
from vecruntime_amd.columns import Column

pytestmark = pytest.mark.gpu


def _spec(hint=0):
    return abi.AggSpec([abi.INT64], [abi.INT64],
                       [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_LONG, 0), abi.agg_desc(abi.MAX, 0)],
                       capacity_hint=hint)


def test_oom_from_a_small_budget():
    import torch
    from vecruntime_amd import vr
    dev = vr.Device(0, hbm_budget=1 << 20)
    try:
        with pytest.raises(vr.VrError) as e:  # the table for 10M groups does not fit 1 MiB
            vr.GroupedAggregate(dev, _spec(10_000_000))
        assert e.value.status == abi.VR_ERR_OOM
    finally:
        dev.close()
    dev = vr.Device(0, hbm_budget=48 << 20)
    try:
        n = 3_000_000
        keys = torch.arange(n, dtype=torch.int64, device="cuda") * 7919
        g = vr.GroupedAggregate(dev, _spec())
        with pytest.raises(vr.VrError) as e:  # 3M groups outgrow 48 MiB while the table grows
            g.update([vr.device_column(abi.INT64, keys)], [vr.device_column(abi.INT64, keys)], n)
        assert e.value.status == abi.VR_ERR_OOM
        g.close()
        jspec = abi.JoinSpec([abi.INT64], [abi.INT64], [0])
        bt = vr.BuildTable(dev, jspec)
        with pytest.raises(vr.VrError) as e:  # a 3M-row build side does not fit either
            bt.append([vr.device_column(abi.INT64, keys)], n)
            bt.finish()
        assert e.value.status == abi.VR_ERR_OOM
        bt.close()
        # the device serves work that fits, with correct results
        rng = np.random.default_rng(3)
        k = rng.integers(0, 1000, 50_000)
        v = rng.integers(-10 ** 9, 10 ** 9, 50_000)
        g, o = vr.GroupedAggregate(dev, _spec()), OracleAgg(_spec())
        g.update([Column.longs(k)], [Column.longs(v)], len(k))
        o.update([Column.longs(k)], [Column.longs(v)], len(k))
        G = o.num_groups()
        assert g.num_groups() == G
        gk, go = g.emit()
        ok, oo = o.emit()
        got = {a: (b, c, d) for a, b, c, d in zip(gk[0].values(G), go[0].values(G), go[1].values(G), go[2].values(G))}
        exp = {a: (b, c, d) for a, b, c, d in zip(ok[0].values(G), oo[0].values(G), oo[1].values(G), oo[2].values(G))}
        assert got == exp
        g.close()
    finally:
        dev.close()


def test_row_bound_of_32bit_piece_sums():
    import torch
    from vecruntime_amd import vr
    dev = vr.Device(0)
    try:
        g = vr.GroupedAggregate(dev, _spec())
        rng = np.random.default_rng(5)
        k, v = rng.integers(0, 100, 1000), rng.integers(-10 ** 6, 10 ** 6, 1000)
        g.update([Column.longs(k)], [Column.longs(v)], 1000)
        tiny = torch.zeros(16, dtype=torch.int64, device="cuda")
        col = vr.device_column(abi.INT64, tiny, length=16)
        # 1000 rows folded: 2^32 - 1000 more would reach 2^32 (refused before any row is read)
        with pytest.raises(vr.VrError) as e:
            g.update_batches([([col], [col], 2 ** 31), ([col], [col], 2 ** 31 - 1000)])
        assert e.value.status == abi.VR_ERR_UNSUPPORTED and "2^32" in str(e.value)
        with pytest.raises(vr.VrError) as e:
            g.update([col], [col], 2 ** 32 - 1000)
        assert e.value.status == abi.VR_ERR_UNSUPPORTED
        # the handle is intact: its groups are the 1000 rows', and after a reset it counts from zero
        o = OracleAgg(_spec())
        o.update([Column.longs(k)], [Column.longs(v)], 1000)
        assert g.num_groups() == o.num_groups()
        g.reset()
        g.update([Column.longs(k)], [Column.longs(v)], 1000)
        assert g.num_groups() == o.num_groups()
        g.close()
    finally:
        dev.close()
