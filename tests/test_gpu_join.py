"""Device hash join build/probe against the golden fixture and the CPU oracle (MI355X)."""
import numpy as np
import pytest

from golden_util import load
from oracle import pyref
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


def build(dev, spec, cols, n, chunks=1):
    from vecruntime_amd import vr
    bt = vr.BuildTable(dev, spec)
    step = max(1, (n + chunks - 1) // chunks)
    for s0 in range(0, n, step):
        m = min(n, s0 + step) - s0
        part = []
        for c in cols:
            if c.type == abi.DECIMAL128:
                part.append(Column(c.type, np.ascontiguousarray(c.data[2 * s0:2 * (s0 + m)]), length=m))
            else:
                vn = None
                if c.validity is not None:
                    vn = ~np.unpackbits(c.validity, bitorder="little", count=n).astype(bool)[s0:s0 + m]
                part.append(Column(c.type, np.ascontiguousarray(c.data[s0:s0 + m]),
                                   None if vn is None else pack_bits(~vn), length=m))
        bt.append(part, m)
    bt.finish()
    return bt


@pytest.mark.parametrize("chunks", [1, 7])
def test_join_fixture_pairs_semi_anti(dev, chunks):
    """Duplicate and null build keys, hits / misses / null probe keys, a selection: pair order,
    semi, anti and the null-aware anti exactly as the nested-loop fixture."""
    z = load("join_dup.npz")
    nb, npr = len(z["bkeys"]), len(z["pkeys"])
    bcols = [Column.longs(z["bkeys"], nulls=z["bkeys_null"]), Column.ints(z["bpay"])]
    spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32], [0])
    bt = build(dev, spec, bcols, nb, chunks)
    probe = [Column.longs(z["pkeys"], nulls=z["pkeys_null"])]
    sel = pack_bits(z["psel"])
    p, b, outs, r = bt.probe(probe, [0], npr, abi.INNER, sel, build_out=[1], probe_out=[0])
    assert p.tolist() == z["inner_probe"].tolist()
    assert b.tolist() == z["inner_build"].tolist()
    assert outs[1].data[:r].tolist() == z["bpay"][z["inner_build"]].tolist()
    assert outs[0].data[:r].tolist() == z["pkeys"][z["inner_probe"]].tolist()
    p, _, _, _ = bt.probe(probe, [0], npr, abi.LEFT_SEMI, sel)
    assert p.tolist() == z["semi"].tolist()
    p, _, _, _ = bt.probe(probe, [0], npr, abi.LEFT_ANTI, sel)
    assert p.tolist() == z["anti"].tolist()
    p, _, _, _ = bt.probe(probe, [0], npr, abi.LEFT_ANTI, sel, drop_null_keys=True)
    assert p.tolist() == z["anti_null_aware"].tolist()


def test_join_unique_keys_materialized_matches_oracle(dev):
    """The join config's shape, small: unique int64 build keys with int32 + decimal(7,2) payloads,
    probe keys 90% hits, two int32 payloads; inner materialized, semi, anti."""
    rng = np.random.default_rng(103)
    nb, npr = 100_000, 500_000
    bkeys = (np.arange(nb, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)).astype(np.int64)
    hit = rng.random(npr) < 0.9
    pkeys = np.where(hit, bkeys[rng.integers(0, nb, npr)],
                     (np.arange(nb, nb + npr, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)).astype(np.int64))
    bcols = [Column.longs(bkeys), Column.ints(rng.integers(-2 ** 31, 2 ** 31, nb)),
             Column.longs(rng.integers(0, 10 ** 7, nb))]
    pcols = [Column.longs(pkeys), Column.ints(rng.integers(-2 ** 31, 2 ** 31, npr)),
             Column.ints(rng.integers(-2 ** 31, 2 ** 31, npr))]
    spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32, abi.INT64], [0], capacity_hint=nb)
    bt = build(dev, spec, bcols, nb, 3)
    oj = OracleJoin(spec, bcols, nb)
    assert bt.info() == (nb, nb, 1)
    for jt in (abi.INNER, abi.LEFT_SEMI, abi.LEFT_ANTI):
        bo = [0, 1, 2] if jt == abi.INNER else []
        p, b, outs, r = bt.probe(pcols, [0], npr, jt, probe_out=[0, 1, 2], build_out=bo)
        op, ob, oouts, orows = oj.probe(pcols, [0], npr, jt, probe_out=[0, 1, 2], build_out=bo, capacity=npr)
        assert r == orows
        assert (p == op).all()
        if jt == abi.INNER:
            assert (b == ob).all()
        for a, c in zip(outs, oouts):
            assert (a.data[:r] == c.data[:r]).all()
        assert r == (int(hit.sum()) if jt != abi.LEFT_ANTI else npr - int(hit.sum()))


def test_join_device_resident_outputs(dev):
    """Device-resident probe columns and outputs (the plain per-batch drop-in on device memory)."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(7)
    nb, npr = 5000, 65536
    bkeys = rng.permutation(10 ** 6)[:nb].astype(np.int64)
    pk = np.where(rng.random(npr) < 0.5, rng.choice(bkeys, npr), -1 - rng.integers(0, 10 ** 6, npr))
    spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32], [0])
    bcols = [Column.longs(bkeys), Column.ints(np.arange(nb))]
    bt = build(dev, spec, bcols, nb)
    oj = OracleJoin(spec, bcols, nb)
    tk = torch.from_numpy(pk).cuda()
    cap = npr
    pidx = torch.zeros(cap, dtype=torch.int32, device="cuda")
    bidx = torch.zeros(cap, dtype=torch.int32, device="cuda")
    odata = torch.zeros(cap, dtype=torch.int32, device="cuda")
    outs = [vr.device_out_column(abi.INT32, cap, odata)]
    _, _, _, r = bt.probe([vr.device_column(abi.INT64, tk)], [0], npr, abi.INNER, build_out=[1], outputs=outs,
                          probe_idx=pidx, build_idx=bidx, capacity=cap, mem=abi.MEM_DEVICE)
    op, ob, oouts, orows = oj.probe([Column.longs(pk)], [0], npr, abi.INNER, build_out=[1], capacity=cap)
    assert r == orows
    assert (pidx[:r].cpu().numpy() == op).all()
    assert (bidx[:r].cpu().numpy() == ob).all()
    assert (odata[:r].cpu().numpy() == oouts[0].data[:r]).all()


def test_join_sentinel_and_capacity(dev):
    """A key equal to the table's empty marker joins like any other; a too-small output reports
    the rows it needs."""
    from vecruntime_amd import vr
    marker = 0x7FF4DEADBEEF7FF4
    bkeys = np.array([marker, 3, marker, 4], np.int64)
    pkeys = np.array([4, marker, 9, 3, marker], np.int64)
    spec = abi.JoinSpec([abi.INT64], [abi.INT64], [0])
    bt = build(dev, spec, [Column.longs(bkeys)], 4)
    p, b, _, r = bt.probe([Column.longs(pkeys)], [0], 5, abi.INNER, capacity=20)
    exp = pyref.join([bkeys.tolist()], [pkeys.tolist()], 0)
    assert list(zip(p.tolist(), b.tolist())) == exp
    with pytest.raises(vr.VrError) as e:
        bt.probe([Column.longs(pkeys)], [0], 5, abi.INNER, capacity=2)
    assert e.value.status == abi.VR_ERR_CAPACITY


def test_join_two_int32_keys_and_empty_build(dev):
    from vecruntime_amd import vr
    rng = np.random.default_rng(3)
    nb, npr = 3000, 20000
    b1, b2 = rng.integers(0, 50, nb).astype(np.int32), rng.integers(0, 50, nb).astype(np.int32)
    p1, p2 = rng.integers(0, 60, npr).astype(np.int32), rng.integers(0, 60, npr).astype(np.int32)
    spec = abi.JoinSpec([abi.INT32, abi.INT32], [abi.INT32, abi.INT32], [0, 1])
    bcols = [Column.ints(b1), Column.ints(b2)]
    bt = build(dev, spec, bcols, nb)
    oj = OracleJoin(spec, bcols, nb)
    pcols = [Column.ints(p1), Column.ints(p2)]
    p, b, _, r = bt.probe(pcols, [0, 1], npr, abi.INNER)
    op, ob, _, orows = oj.probe(pcols, [0, 1], npr, abi.INNER, capacity=npr * 10)
    assert r == orows and (p == op).all() and (b == ob).all()
    empty = vr.BuildTable(dev, spec)
    empty.finish()
    p, _, _, r = empty.probe(pcols, [0, 1], npr, abi.LEFT_ANTI)
    assert r == npr
    p, _, _, r = empty.probe(pcols, [0, 1], npr, abi.INNER)
    assert r == 0
