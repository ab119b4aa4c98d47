"""Outer joins on the device against the oracle (MI355X): the streamed side preserved
(VR_JOIN_OUTER: pairs, or (row, -1) with the build columns null -- VectorHashJoinExec.scala:
1160-1175) and the build side preserved (track_matched + vr_join_unmatched_build: the build rows no
probe paired, in build order -- :1004-1013,1457-1484).  Pairs, order and materialized columns
bit-exact, with the direct key index and without it."""
import numpy as np
import pytest

from oracle.oracle import OracleJoin
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


def build_both(dev, spec, bcols, nb):
    from vecruntime_amd import vr
    bt = vr.BuildTable(dev, spec)
    bt.append(bcols, nb)
    bt.finish()
    return bt, OracleJoin(spec, bcols, nb)


def same(dev_res, ora_res):
    p, b, outs, r = dev_res
    op, ob, oouts, orows = ora_res
    assert r == orows
    assert p.tolist() == op.tolist()
    assert b.tolist() == ob.tolist()
    for a, c in zip(outs, oouts):
        assert a.values(r) == c.values(r)


@pytest.mark.parametrize("dense", [0, -1])
@pytest.mark.parametrize("key_range", [3000, 10 ** 9])
def test_outer_probe_batches_then_unmatched_build(dev, dense, key_range):
    """Full outer: OUTER probes over several batches (nulls, misses, a selection), then the build
    rows never paired.  A small key range takes the direct index when allowed."""
    print("outer join: dense=%d key_range=%d" % (dense, key_range), flush=True)
    rng = np.random.default_rng(31 + dense)
    nb = 20000
    bkeys = rng.integers(0, key_range, nb)
    bnull = rng.random(nb) < 0.03
    bcols = [Column.longs(bkeys, nulls=bnull), Column.ints(rng.integers(-10 ** 6, 10 ** 6, nb)),
             Column.doubles(rng.standard_normal(nb), nulls=rng.random(nb) < 0.1)]
    spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32, abi.FLOAT64], [0], dense_index=dense, track_matched=True)
    bt, oj = build_both(dev, spec, bcols, nb)
    for batch in range(3):
        n = 7000 + 1000 * batch
        pk = np.where(rng.random(n) < 0.6, rng.choice(bkeys, n), rng.integers(0, 2 * key_range, n))
        pcols = [Column.longs(pk, nulls=rng.random(n) < 0.05), Column.ints(rng.integers(0, 100, n))]
        sel = pack_bits(rng.random(n) < 0.8)
        cap = n * 8
        same(bt.probe(pcols, [0], n, abi.OUTER, sel, probe_out=[1, 0], build_out=[2, 1], capacity=cap),
             oj.probe(pcols, [0], n, abi.OUTER, sel, probe_out=[1, 0], build_out=[2, 1], capacity=cap))
    same(bt.unmatched_build(probe_types=[abi.INT32, abi.INT64], build_out=[0, 2], capacity=nb),
         oj.unmatched_build(probe_types=[abi.INT32, abi.INT64], build_out=[0, 2], capacity=nb))
    bt.close()


def test_build_preserved_by_inner_probes(dev):
    """A join preserving only the build side probes with INNER and marks what it pairs; the
    unmatched rows come out at the end (and a too-small output reports the rows it needs)."""
    from vecruntime_amd import vr
    rng = np.random.default_rng(8)
    nb, n = 5000, 30000
    bkeys = rng.permutation(10 ** 6)[:nb].astype(np.int64)
    bcols = [Column.longs(bkeys), Column.ints(np.arange(nb, dtype=np.int32))]
    spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32], [0], track_matched=True)
    bt, oj = build_both(dev, spec, bcols, nb)
    pk = np.where(rng.random(n) < 0.5, rng.choice(bkeys[: nb // 2], n), -rng.integers(1, 10 ** 6, n))
    pcols = [Column.longs(pk)]
    same(bt.probe(pcols, [0], n, abi.INNER, probe_out=[0], build_out=[1], capacity=n),
         oj.probe(pcols, [0], n, abi.INNER, probe_out=[0], build_out=[1], capacity=n))
    with pytest.raises(vr.VrError) as e:
        bt.unmatched_build(build_out=[1], capacity=10)
    assert e.value.status == abi.VR_ERR_CAPACITY
    same(bt.unmatched_build(build_out=[1], capacity=nb), oj.unmatched_build(build_out=[1], capacity=nb))
    bt.close()


def test_outer_over_an_empty_build_and_untracked_handle(dev):
    from vecruntime_amd import vr
    spec = abi.JoinSpec([abi.INT32], [abi.INT32], [0])
    bt = vr.BuildTable(dev, spec)
    bt.finish()
    pk = np.arange(100, dtype=np.int32)
    p, b, outs, r = bt.probe([Column.ints(pk, nulls=pk % 7 == 0)], [0], 100, abi.OUTER, build_out=[0], capacity=100)
    assert r == 100 and p.tolist() == list(range(100)) and b.tolist() == [-1] * 100
    assert outs[0].values(r) == [None] * 100
    with pytest.raises(vr.VrError) as e:
        bt.unmatched_build(capacity=10)
    assert e.value.status == abi.VR_ERR_STATE
    bt.close()


@pytest.mark.parametrize("dense", [0, -1])
@pytest.mark.parametrize("lane", [abi.INT32, abi.INT64])
def test_residual_matches_oracle(dev, dense, lane):
    """A residual `left + lo <op> right + ro` over streamed and build columns (every side pairing,
    every operator, offsets that wrap the lane) on every join type, then the build rows no passing
    pair touched."""
    print("residual: dense=%d lane=%d" % (dense, lane), flush=True)
    rng = np.random.default_rng(50 + dense + lane)
    nb, n = 6000, 20000
    top = 2 ** 31 - 1 if lane == abi.INT32 else 2 ** 63 - 1
    mk = Column.ints if lane == abi.INT32 else Column.longs

    def vals(m):
        v = rng.integers(-50, 50, m)
        v[rng.random(m) < 0.05] = top  # + offset wraps
        return v
    bkeys = rng.integers(0, 2000, nb)
    bcols = [Column.longs(bkeys, nulls=rng.random(nb) < 0.02), mk(vals(nb), nulls=rng.random(nb) < 0.1),
             Column.ints(np.arange(nb, dtype=np.int32))]
    spec = abi.JoinSpec([abi.INT64], [abi.INT64, lane, abi.INT32], [0], dense_index=dense, track_matched=True)
    bt, oj = build_both(dev, spec, bcols, nb)
    ops = [abi.CMP_EQ, abi.CMP_NE, abi.CMP_LT, abi.CMP_LE, abi.CMP_GT, abi.CMP_GE]
    sides = [((0, 1), (1, 1)), ((1, 1), (0, 1)), ((0, 1), (0, 1)), ((1, 1), (1, 1))]
    for t in range(8):
        pk = np.where(rng.random(n) < 0.7, rng.choice(bkeys, n), rng.integers(2000, 4000, n))
        pcols = [Column.longs(pk, nulls=rng.random(n) < 0.03), mk(vals(n), nulls=rng.random(n) < 0.1)]
        (ls, rs) = sides[t % 4]
        res = abi.residual(ops[t % 6], ls, rs, int(rng.choice([0, 1, -7])), int(rng.choice([0, 3])))
        for jt in (abi.INNER, abi.OUTER, abi.LEFT_SEMI, abi.LEFT_ANTI):
            bo = [2] if jt in (abi.INNER, abi.OUTER) else []
            cap = n * 16
            d = bt.probe(pcols, [0], n, jt, probe_out=[1], build_out=bo, capacity=cap, residual=res)
            o = oj.probe(pcols, [0], n, jt, probe_out=[1], build_out=bo, capacity=cap, residual=res)
            assert d[3] == o[3], (t, jt)
            assert d[0].tolist() == o[0].tolist(), (t, jt)
            if jt in (abi.INNER, abi.OUTER):
                assert d[1].tolist() == o[1].tolist(), (t, jt)
            for a, c in zip(d[2], o[2]):
                assert a.values(d[3]) == c.values(o[3])
    same(bt.unmatched_build(build_out=[2], capacity=nb), oj.unmatched_build(build_out=[2], capacity=nb))
    bt.close()
