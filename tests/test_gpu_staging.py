"""Host-resident input through the library's staging (MI355X): batches in pageable host memory
(copied into the library's pinned buffers) and in page-locked host memory (large buffers copied by
DMA straight from the caller's) give the oracle's groups and join rows; a caller that overwrites
its page-locked buffer right after a call does not change that call's result."""
import numpy as np
import pytest

from gpu_util import compare_groups, group_rows, out_roles
from oracle.oracle import OracleAgg, OracleJoin
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


def pinned(a):
    import torch
    t = torch.empty(a.shape, dtype=getattr(torch, a.dtype.name), pin_memory=True)
    t.numpy()[...] = a
    return t.numpy(), t


SPEC = abi.AggSpec([abi.INT64], [abi.INT64, abi.FLOAT64, abi.INT32],
                   [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_LONG, 0), abi.agg_desc(abi.SUM_DOUBLE, 1),
                    abi.agg_desc(abi.MIN, 2), abi.agg_desc(abi.MAX, 2)], key_nullable=[False])


def batch(rng, n):
    k = rng.integers(0, 5000, n).astype(np.int64) * 1_000_003
    v0 = rng.integers(-10 ** 9, 10 ** 9, n)
    v1 = rng.lognormal(0, 1, n)  # positive: the 1e-12 bound relative to the total is the bound relative to sum |x|
    v2 = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
    nulls = rng.random(n) < 0.05
    return k, v0, v1, v2, pack_bits(~nulls)


@pytest.mark.parametrize("memory", ["pageable", "pinned"])
def test_host_batches_match_oracle(dev, memory):
    from vecruntime_amd import vr
    rng = np.random.default_rng(77)
    g, o = vr.GroupedAggregate(dev, SPEC), OracleAgg(SPEC)
    keep = []
    for n in (65_536, 1, 40_000, 600_000, 65_536, 777):  # 600,000 rows: 4.8 MB lanes, copied by DMA as they are
        k, v0, v1, v2, vb = batch(rng, n)
        o.update([Column.longs(k)], [Column.longs(v0), Column.doubles(v1), Column(abi.INT32, v2, vb.copy(), length=n)], n)
        if memory == "pinned":
            (k, tk), (v0, t0), (v1, t1), (v2, t2), (vb, tb) = (pinned(a) for a in (k, v0, v1, v2, vb))
            keep += [tk, t0, t1, t2, tb]
        cols = ([Column.longs(k)], [Column.longs(v0), Column.doubles(v1), Column(abi.INT32, v2, vb, length=n)])
        g.update(cols[0], cols[1], n)
        if memory == "pinned":  # the caller reuses its buffers at once
            for a in (k, v0, v1, v2):
                a[...] = 0
    G = o.num_groups()
    assert g.num_groups() == G
    gk, go = g.emit()
    ok, oo = o.emit()
    worst = compare_groups(group_rows(gk, go, G, SPEC, {}), group_rows(ok, oo, G, SPEC, {}), out_roles(SPEC))
    assert worst < 1e-12
    g.close()


@pytest.mark.parametrize("memory", ["pageable", "pinned"])
def test_host_probe_matches_oracle(dev, memory):
    from vecruntime_amd import vr
    rng = np.random.default_rng(78)
    nb, npr = 30_000, 200_000
    bkeys = rng.permutation(nb).astype(np.int64) * 7919
    bpay = rng.integers(0, 1000, nb).astype(np.int32)
    pk = np.where(rng.random(npr) < 0.8, bkeys[rng.integers(0, nb, npr)], -1 - rng.integers(0, 10 ** 9, npr))
    pp = rng.integers(0, 10 ** 6, npr).astype(np.int64)
    spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32], [0], capacity_hint=nb)
    bt = vr.BuildTable(dev, spec)
    bt.append([Column.longs(bkeys), Column.ints(bpay)], nb)
    bt.finish()
    oj = OracleJoin(spec, [Column.longs(bkeys), Column.ints(bpay)], nb)
    keep = []
    if memory == "pinned":
        (pk, tk), (pp, tp) = pinned(pk), pinned(pp)
        keep += [tk, tp]
    probe = [Column.longs(pk), Column.longs(pp)]
    for jt in (abi.INNER, abi.LEFT_SEMI, abi.LEFT_ANTI):
        want_p, want_b, _, _ = oj.probe([Column.longs(np.array(pk)), Column.longs(np.array(pp))], [0], npr, jt)
        gp, gb, _, r = bt.probe(probe, [0], npr, jt)
        assert r == len(want_p) and list(gp[:r]) == list(want_p)
        if jt == abi.INNER:
            assert list(gb[:r]) == list(want_b)
    bt.close()


@pytest.mark.parametrize("memory", ["pageable", "pinned"])
def test_host_batch_set_in_chunks_matches_oracle(dev, memory):
    """vr_agg_update_batches over host-resident batches (about 400 MB, so the library stages them in
    more than one chunk) with the fused filter: the groups equal the oracle's."""
    from vecruntime_amd import vr
    rng = np.random.default_rng(79)
    n_batches, bn = 220, 65_536
    N = n_batches * bn + 777
    k = (rng.integers(0, 20_000, N) * 1_000_003).astype(np.int64)
    v0 = rng.integers(-10 ** 9, 10 ** 9, N)
    v1 = rng.lognormal(0, 1, N)
    v2 = rng.integers(-2 ** 31, 2 ** 31, N).astype(np.int32)
    nulls = rng.random(N) < 0.05
    vb = pack_bits(~nulls)
    keep = []
    if memory == "pinned":
        (k, tk), (v0, t0), (v1, t1), (v2, t2), (vb, tb) = (pinned(a) for a in (k, v0, v1, v2, vb))
        keep += [tk, t0, t1, t2, tb]
    batches = []
    for s0 in range(0, N, bn):
        m = min(bn, N - s0)
        batches.append(([Column.longs(k[s0:s0 + m])],
                        [Column.longs(v0[s0:s0 + m]), Column.doubles(v1[s0:s0 + m]),
                         Column(abi.INT32, v2[s0:s0 + m], vb[s0 // 8:], length=m)], m))
    thr = 0
    g, o = vr.GroupedAggregate(dev, SPEC), OracleAgg(SPEC)
    g.update_batches(batches, filter_col=2, threshold=thr)
    for kc, vc, m in batches:
        f = np.asarray(vc[2].data[:m])
        ok = f < thr
        valid = np.unpackbits(np.asarray(vc[2].validity[: (m + 7) // 8]), bitorder="little", count=m).astype(bool)
        o.update(kc, vc, m, pack_bits(ok & valid))
    G = o.num_groups()
    assert g.num_groups() == G
    gk, go = g.emit()
    ok_, oo = o.emit()
    worst = compare_groups(group_rows(gk, go, G, SPEC, {}), group_rows(ok_, oo, G, SPEC, {}), out_roles(SPEC))
    assert worst < 1e-12
    g.close()
