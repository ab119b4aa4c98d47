"""Inner and outer joins that read the build side's payload lanes from the payload table's cells
(unique keys; fixed-width payload lanes without nulls, at most 20 bytes) against the CPU oracle
(MI355X).

The payload table only changes where a probe reads the build columns it writes: pairs, their
order and every output value must equal the oracle's per-batch probes
(VectorHashJoinExec.scala:1116-1182), for single-batch calls and for many batches in one launch,
across home buckets that overflowed into the next ones."""
import numpy as np
import pytest

from oracle.oracle import OracleJoin
from vecruntime_amd import abi
from vecruntime_amd.columns import Column, unpack_bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    from vecruntime_amd import vr
    d = vr.Device(0)
    yield d
    d.close()


def build_cols(rng, nb, dup=False, nulls=False):
    keys = rng.permutation(np.arange(nb, dtype=np.int64) * 7919 - 3 * 10 ** 9)
    if dup:
        keys[nb // 2:] = keys[: nb - nb // 2]
    i32 = rng.integers(-2 ** 31, 2 ** 31, nb)
    dec = np.stack([rng.integers(-2 ** 62, 2 ** 62, nb), rng.integers(-2 ** 20, 2 ** 20, nb)], axis=1).astype(np.int64)
    n32 = (rng.random(nb) < 0.1) if nulls else None
    return keys, [Column.longs(keys), Column.ints(i32, nulls=n32), Column(abi.DECIMAL128, dec.reshape(-1))]


def probe_cols(rng, keys, n, hit=0.85):
    h = rng.random(n) < hit
    k = np.where(h, keys[rng.integers(0, len(keys), n)], rng.integers(2 ** 40, 2 ** 41, n)).astype(np.int64)
    return [Column.longs(k), Column.ints(rng.integers(-2 ** 31, 2 ** 31, n)), Column.ints(rng.integers(0, 100, n))]


SPEC_TYPES = [abi.INT64, abi.INT32, abi.DECIMAL128]


def spec(nb):
    return abi.JoinSpec([abi.INT64], SPEC_TYPES, [0], capacity_hint=nb)


def make(dev, cols, nb):
    from vecruntime_amd import vr
    bt = vr.BuildTable(dev, spec(nb))
    bt.append(cols, nb)
    bt.finish()
    return bt


def out_data(cols, r):
    """(data, validity) of host output columns, decimals as (r, 2) int64."""
    res = []
    for c in cols:
        d = np.asarray(c.data)
        d = d[: 2 * r].reshape(r, 2) if c.type == abi.DECIMAL128 else d[:r]
        v = unpack_bits(c.validity, r) if c.validity is not None else np.ones(r, bool)
        res.append((d.copy(), v))
    return res


def assert_outputs(got, exp):
    for (gd, gv), (ed, ev) in zip(got, exp):
        assert (gv == ev).all()
        assert (gd[ev] == ed[ev]).all()


@pytest.mark.parametrize("jt", [abi.INNER, abi.OUTER])
@pytest.mark.parametrize("nb", [1, 37, 20_000])
def test_payload_probe_matches_oracle(dev, jt, nb):
    rng = np.random.default_rng(nb + jt)
    keys, cols = build_cols(rng, nb)
    bt = make(dev, cols, nb)
    oj = OracleJoin(spec(nb), cols, nb)
    before = dev.stats()["payload_join_probes"]
    for n in (1, 2047, 2048, 50_001):
        pc = probe_cols(rng, keys, n)
        build_out = (2, 0, 1)  # decimal, the key (from the streamed key), int32
        cap = max(1, n)
        outs = [abi_out(t, cap) for t in [abi.INT32, abi.INT64] + [SPEC_TYPES[c] for c in build_out]]
        p, b, o, r = bt.probe(pc, [0], n, jt, probe_out=(1, 0), build_out=build_out, capacity=cap, outputs=outs)
        op, ob, oo, orows = oj.probe(pc, [0], n, jt, probe_out=(1, 0), build_out=build_out, capacity=cap)
        assert r == orows
        assert (p == op[:r]).all() and (b == ob[:r]).all()
        assert_outputs(out_data(o, r), out_data(oo, r))
    assert dev.stats()["payload_join_probes"] - before == 4


def abi_out(t, cap):
    from vecruntime_amd.columns import OutColumn
    return OutColumn(t, cap)


def test_payload_table_many_batches(dev):
    """One launch over uneven device batches (tile and batch edges) against the oracle's per-batch
    probes concatenated."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(5)
    nb = 30_000
    keys, cols = build_cols(rng, nb)
    bt = make(dev, cols, nb)
    oj = OracleJoin(spec(nb), cols, nb)
    lengths = [1, 2048, 2049, 0, 65_536, 777, 4095]
    hb = [probe_cols(rng, keys, n) for n in lengths]
    dbatches, keep = [], []
    for pc, n in zip(hb, lengths):
        dc = []
        for c in pc:
            t = torch.from_numpy(np.ascontiguousarray(c.data)).cuda()
            keep.append(t)
            dc.append(vr.device_column(c.type, t, None, length=n))
        dbatches.append((dc, n, None))
    pb = vr.ProbeBatches(dbatches)
    build_out = (1, 2)
    total = sum(lengths)
    cap = max(64, total)
    pidx = torch.full((cap,), -7, dtype=torch.int32, device="cuda")
    bidx = torch.full((cap,), -7, dtype=torch.int32, device="cuda")
    o32 = torch.zeros(cap, dtype=torch.int32, device="cuda")
    odec = torch.zeros(2 * cap, dtype=torch.int64, device="cuda")
    v32 = torch.zeros((cap + 63) // 64 * 8, dtype=torch.uint8, device="cuda")
    vdec = torch.zeros((cap + 63) // 64 * 8, dtype=torch.uint8, device="cuda")
    outs = [vr.device_out_column(abi.INT32, cap, o32, v32), vr.device_out_column(abi.DECIMAL128, cap, odec, vdec)]
    before = dev.stats()["payload_join_probes"]
    rows, per = bt.probe_batches(pb, [0], abi.INNER, outputs=outs, build_out=build_out, capacity=cap, probe_idx=pidx,
                                 build_idx=bidx)
    torch.cuda.synchronize()
    assert dev.stats()["payload_join_probes"] - before == 1
    ps, bs, e32, edec, eper = [], [], [], [], []
    for pc, n in zip(hb, lengths):
        p, b, o, r = oj.probe(pc, [0], n, abi.INNER, build_out=build_out, capacity=max(64, n))
        ps.append(p[:r])
        bs.append(b[:r])
        e32.append(np.asarray(o[0].data)[:r])
        edec.append(np.asarray(o[1].data)[: 2 * r].reshape(r, 2))
        eper.append(r)
    assert rows == sum(eper) and per.tolist() == eper
    assert (pidx[:rows].cpu().numpy() == np.concatenate(ps)).all()
    assert (bidx[:rows].cpu().numpy() == np.concatenate(bs)).all()
    assert (o32[:rows].cpu().numpy() == np.concatenate(e32)).all()
    assert (odec[: 2 * rows].cpu().numpy().reshape(rows, 2) == np.concatenate(edec)).all()
    assert unpack_bits(v32.cpu().numpy(), rows).all() and unpack_bits(vdec.cpu().numpy(), rows).all()


@pytest.mark.parametrize("dup,nulls", [(True, False), (False, True)])
def test_no_payload_table_when_keys_repeat_or_payload_has_nulls(dev, dup, nulls):
    """Duplicate build keys or a nullable payload lane keep the gathers; results equal the oracle."""
    rng = np.random.default_rng(11 + dup)
    nb = 5_000
    keys, cols = build_cols(rng, nb, dup=dup, nulls=nulls)
    bt = make(dev, cols, nb)
    oj = OracleJoin(spec(nb), cols, nb)
    n = 9_000
    pc = probe_cols(rng, keys, n)
    before = dev.stats()["payload_join_probes"]
    cap = 4 * n
    p, b, o, r = bt.probe(pc, [0], n, abi.INNER, build_out=(1, 2), capacity=cap)
    op, ob, oo, orows = oj.probe(pc, [0], n, abi.INNER, build_out=(1, 2), capacity=cap)
    assert dev.stats()["payload_join_probes"] == before
    assert r == orows and (p == op[:r]).all() and (b == ob[:r]).all()
    assert_outputs(out_data(o, r), out_data(oo, r))
