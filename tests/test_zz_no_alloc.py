"""The allocation-free probe (vr_join_probe_noalloc) against vr_join_probe on the same inputs (MI355X).

Two calls in a row, the second smaller than the first (the shared scratch keeps the first call's
contents), each checked bit for bit against vr_join_probe: rows, probe/build indices and the
materialized columns with their validity.  The library's device-allocation counter does not move
across the allocation-free calls."""
import numpy as np
import pytest

from vecruntime_amd import abi
from vecruntime_amd.columns import Column, pack_bits, unpack_bits

pytestmark = pytest.mark.gpu


def _device(col):
    import torch
    from vecruntime_amd import vr
    v = None if col.validity is None else torch.from_numpy(np.ascontiguousarray(col.validity)).cuda()
    return vr.device_column(col.type, torch.from_numpy(np.ascontiguousarray(col.data)).cuda(), v, length=col.length)


def _outputs(types, cap):
    import torch
    from vecruntime_amd import vr
    keep, outs = [], []
    for t in types:
        w = {abi.INT32: torch.int32, abi.INT64: torch.int64}[t]
        d = torch.full((cap,), -3, dtype=w, device="cuda")
        v = torch.zeros((cap + 63) // 64 * 8, dtype=torch.uint8, device="cuda")
        keep.append((d, v))
        outs.append(vr.device_out_column(t, cap, d, v))
    return keep, outs


def _run(bt, fn, cols, n, sel, jt, bo, cap):
    import torch
    pidx = torch.full((cap,), -9, dtype=torch.int32, device="cuda")
    bidx = torch.full((cap,), -9, dtype=torch.int32, device="cuda")
    types = [abi.INT64, abi.INT32] + [bt.spec.column_types[c] for c in bo]
    keep, outs = _outputs(types, cap)
    pairs = jt in (abi.INNER, abi.OUTER)
    if fn == "probe":
        _, _, _, r = bt.probe(cols, [0], n, jt, selection=sel, probe_out=[0, 1], build_out=bo, outputs=outs,
                              probe_idx=pidx, build_idx=bidx if pairs else None, capacity=cap, mem=abi.MEM_DEVICE)
    else:
        r = bt.probe_noalloc(cols, [0], n, jt, outputs=outs, probe_out=[0, 1], build_out=bo, capacity=cap,
                             probe_idx=pidx, build_idx=bidx if pairs else None, selection=sel)
    torch.cuda.synchronize()
    res = [r, pidx[:r].cpu().numpy(), bidx[:r].cpu().numpy() if pairs else None]
    for d, v in keep:
        valid = unpack_bits(v.cpu().numpy(), r)
        res.append((valid, np.where(valid, d[:r].cpu().numpy(), 0)))
    return res


def test_noalloc_probe_matches_probe():
    import torch
    from vecruntime_amd import vr
    dev = vr.Device(0)
    try:
        rng = np.random.default_rng(77)
        nb = 40_000
        base = rng.integers(-2 ** 62, 2 ** 62, nb // 2)
        bkeys = base[rng.integers(0, len(base), nb)].astype(np.int64)
        bcols = [Column.longs(bkeys, nulls=rng.random(nb) < 0.03), Column.ints(rng.integers(-10 ** 9, 10 ** 9, nb))]
        spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32], [0])
        bt = vr.BuildTable(dev, spec)
        bt.append(bcols, nb)
        bt.finish()
        for n in (300_001, 5_003):  # the second call smaller than the first
            k = np.where(rng.random(n) < 0.8, bkeys[rng.integers(0, nb, n)], rng.integers(-2 ** 62, 2 ** 62, n))
            hc = [Column.longs(k, nulls=rng.random(n) < 0.02), Column.ints(rng.integers(-5, 5, n))]
            cols = [_device(c) for c in hc]
            sel = torch.from_numpy(pack_bits(rng.random(n) < 0.7)).cuda()
            for jt, bo in ((abi.INNER, [1, 0]), (abi.OUTER, [1]), (abi.LEFT_SEMI, []), (abi.LEFT_ANTI, [])):
                cap = n * 4
                exp = _run(bt, "probe", cols, n, sel, jt, bo, cap)
                before = dev.stats()["device_allocations"]
                got = _run(bt, "noalloc", cols, n, sel, jt, bo, cap)
                assert dev.stats()["device_allocations"] == before, "the allocation-free probe allocated"
                assert got[0] == exp[0]
                assert (got[1] == exp[1]).all()
                if exp[2] is not None:
                    assert (got[2] == exp[2]).all()
                for (gv, gd), (ev, ed) in zip(got[3:], exp[3:]):
                    assert (gv == ev).all() and (gd == ed).all()
        # refused: host columns, host outputs, a call past the scratch's size
        with pytest.raises(vr.VrError) as e:
            bt.probe_noalloc([Column.longs(np.arange(10))], [0], 10, abi.LEFT_SEMI, capacity=10,
                             probe_idx=torch.zeros(10, dtype=torch.int32, device="cuda"))
        assert e.value.status == abi.VR_ERR_UNSUPPORTED
        small = [_device(Column.longs(np.arange(10)))]
        with pytest.raises(vr.VrError) as e:
            bt.probe_noalloc(small, [0], 131_070 * 2048 + 1, abi.LEFT_SEMI, capacity=10,
                             probe_idx=torch.zeros(10, dtype=torch.int32, device="cuda"))
        assert e.value.status == abi.VR_ERR_UNSUPPORTED and "at most" in str(e.value)
        bt.close()
    finally:
        dev.close()
