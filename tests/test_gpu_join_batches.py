"""Many streamed batches probed with one launch (vr_join_probe_batches) against the CPU oracle's
per-batch probes, concatenated in batch order (MI355X).

The reference probes batch by batch (VectorHashJoinIterator.probe/emitMatches/emitSemiAnti,
VectorHashJoinExec.scala:1038-1182); one launch over all batches must give exactly the
concatenation of those per-batch results: pairs in streamed-row order, build rows in build order,
across tile and batch boundaries."""
import threading

import numpy as np
import pytest

from oracle.oracle import OracleJoin
from vecruntime_amd import abi
from vecruntime_amd.columns import Column, pack_bits, unpack_bits

pytestmark = pytest.mark.gpu

TILE = 2048  # rows of a probe tile


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    from vecruntime_amd import vr
    d = vr.Device(0)
    yield d
    d.close()


def to_device(col):
    import torch
    from vecruntime_amd import vr
    v = None if col.validity is None else torch.from_numpy(np.ascontiguousarray(col.validity)).cuda()
    return vr.device_column(col.type, torch.from_numpy(np.ascontiguousarray(col.data)).cuda(), v, length=col.length)


def make_build(dev, spec, cols, n):
    from vecruntime_amd import vr
    bt = vr.BuildTable(dev, spec)
    bt.append(cols, n)
    bt.finish()
    return bt


def batched(dev, bt, host_batches, key_cols, jt, probe_out=(), build_out=(), drop_null_keys=False, residual=None,
            validity=True, capacity=None):
    """Runs vr_join_probe_batches over host batches copied to the device; returns host arrays."""
    import torch
    from vecruntime_amd import vr
    dbatches = []
    for cols, n, sel in host_batches:
        dcols = [to_device(c) for c in cols]
        dsel = None if sel is None else torch.from_numpy(sel).cuda()
        dbatches.append((dcols, n, dsel))
    pb = vr.ProbeBatches(dbatches)
    pairs = jt in (abi.INNER, abi.OUTER)
    if capacity is None:
        capacity = max(64, sum(n for _, n, _ in host_batches) * (16 if pairs else 1))
    pidx = torch.full((capacity,), -7, dtype=torch.int32, device="cuda")
    bidx = torch.full((capacity,), -7, dtype=torch.int32, device="cuda") if pairs else None
    types = [host_batches[0][0][c].type for c in probe_out] + [bt.spec.column_types[c] for c in build_out]
    outs, keep = [], []
    for t in types:
        w = {abi.INT32: torch.int32, abi.INT64: torch.int64, abi.FLOAT64: torch.float64}[t]
        d = torch.zeros(capacity, dtype=w, device="cuda")
        v = torch.zeros((capacity + 63) // 64 * 8, dtype=torch.uint8, device="cuda") if validity else None
        keep.append((d, v))
        outs.append(vr.device_out_column(t, capacity, d, v))
    rows, per = bt.probe_batches(pb, key_cols, jt, outputs=outs, probe_out=probe_out, build_out=build_out,
                                 capacity=capacity, probe_idx=pidx, build_idx=bidx, drop_null_keys=drop_null_keys,
                                 residual=residual)
    torch.cuda.synchronize()
    r = min(rows, capacity)
    res = {"rows": rows, "per": per.tolist(), "p": pidx[:r].cpu().numpy(),
           "b": None if bidx is None else bidx[:r].cpu().numpy(),
           "outs": [(d[:r].cpu().numpy(), None if v is None else unpack_bits(v.cpu().numpy(), r)) for d, v in keep]}
    return res


def oracle_concat(oj, host_batches, key_cols, jt, probe_out=(), build_out=(), drop_null_keys=False, residual=None):
    ps, bs, per = [], [], []
    outs = None
    for cols, n, sel in host_batches:
        p, b, o, r = oj.probe(cols, key_cols, n, jt, sel, drop_null_keys, probe_out, build_out,
                              capacity=max(64, n * 32), residual=residual)
        ps.append(p[:r])
        if b is not None:
            bs.append(b[:r])
        per.append(r)
        cur = [(c.data[:r].copy(), unpack_bits(c.validity, r) if c.validity is not None else np.ones(r, bool)) for c in o]
        outs = cur if outs is None else [(np.concatenate([a[0], x[0]]), np.concatenate([a[1], x[1]])) for a, x in zip(outs, cur)]
    return {"rows": sum(per), "per": per, "p": np.concatenate(ps) if ps else np.zeros(0, np.int32),
            "b": np.concatenate(bs) if bs else None, "outs": outs or []}


def assert_same(got, exp):
    assert got["rows"] == exp["rows"]
    assert got["per"] == exp["per"]
    assert (got["p"] == exp["p"]).all()
    if exp["b"] is not None:
        assert (got["b"] == exp["b"]).all()
    for (gd, gv), (ed, ev) in zip(got["outs"], exp["outs"]):
        if gv is not None:
            assert (gv == ev).all()
            assert (gd[ev] == ed[ev]).all()  # data under null bits is not compared
        else:
            assert (gd == ed).all()


def probe_batches(rng, bkeys, lengths, hit=0.8, null_frac=0.0, sel_every=0, miss_base=None):
    """Host batches: key (int64, hits from bkeys, misses elsewhere), two int32 payloads."""
    out = []
    for bi, n in enumerate(lengths):
        h = rng.random(n) < hit
        miss = rng.integers(-2 ** 62, 2 ** 62, n) if miss_base is None else miss_base[rng.integers(0, len(miss_base), n)]
        k = np.where(h, bkeys[rng.integers(0, len(bkeys), n)], miss).astype(np.int64)
        nulls = rng.random(n) < null_frac if null_frac else None
        cols = [Column.longs(k, nulls=nulls), Column.ints(rng.integers(-2 ** 31, 2 ** 31, n)),
                Column.ints(rng.integers(-2 ** 31, 2 ** 31, n))]
        sel = pack_bits(rng.random(n) < 0.6) if sel_every and bi % sel_every == 0 else None
        out.append((cols, n, sel))
    return out


LENGTHS = [1, 0, TILE - 1, TILE, TILE + 1, 3 * TILE, 0, 65536, 7, 100_003, 2 * TILE + 5, 1]


@pytest.mark.parametrize("dups", [False, True])
def test_batches_equal_per_batch_oracle(dev, dups):
    """Uneven batches (empty, one row, tile edges, 65,536 and 100,003 rows), selections, null keys,
    misses; unique build keys or runs of duplicates (pairs crossing tile and batch boundaries)."""
    rng = np.random.default_rng(11 + dups)
    nb = 60_000
    base = rng.integers(-2 ** 62, 2 ** 62, nb // (4 if dups else 1))
    bkeys = (base[rng.integers(0, len(base), nb)] if dups else base).astype(np.int64)
    bnull = rng.random(nb) < 0.02
    bcols = [Column.longs(bkeys, nulls=bnull), Column.ints(rng.integers(-2 ** 31, 2 ** 31, nb)),
             Column.longs(rng.integers(0, 10 ** 7, nb))]
    spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32, abi.INT64], [0])
    bt = make_build(dev, spec, bcols, nb)
    oj = OracleJoin(spec, bcols, nb)
    hb = probe_batches(rng, bkeys, LENGTHS, null_frac=0.03, sel_every=3)
    for jt, po, bo, nak in ((abi.INNER, [0, 1, 2], [0, 1, 2], False), (abi.INNER, [], [], False),
                            (abi.OUTER, [0, 2], [2, 0], False), (abi.LEFT_SEMI, [0, 1, 2], [], False),
                            (abi.LEFT_ANTI, [0, 1], [], False), (abi.LEFT_ANTI, [0], [], True)):
        got = batched(dev, bt, hb, [0], jt, po, bo, drop_null_keys=nak)
        exp = oracle_concat(oj, hb, [0], jt, po, bo, drop_null_keys=nak)
        assert_same(got, exp)
    bt.close()


def test_batches_two_int32_keys_dense_and_residual(dev):
    """Two INT32 keys (hashed), one INT32 key over a dense range (the direct index), and a fused
    residual over a streamed and a build column."""
    from vecruntime_amd import vr
    rng = np.random.default_rng(5)
    nb = 20_000
    b1, b2 = rng.integers(0, 200, nb).astype(np.int32), rng.integers(0, 200, nb).astype(np.int32)
    spec2 = abi.JoinSpec([abi.INT32, abi.INT32], [abi.INT32, abi.INT32], [0, 1])
    bcols2 = [Column.ints(b1), Column.ints(b2)]
    bt2 = make_build(dev, spec2, bcols2, nb)
    oj2 = OracleJoin(spec2, bcols2, nb)
    hb2 = []
    for n in (5000, 0, TILE, 31_111):
        hb2.append(([Column.ints(rng.integers(0, 220, n)), Column.ints(rng.integers(0, 220, n))], n, None))
    for jt, bo in ((abi.INNER, [0, 1]), (abi.LEFT_SEMI, []), (abi.LEFT_ANTI, [])):
        assert_same(batched(dev, bt2, hb2, [0, 1], jt, [0, 1], bo), oracle_concat(oj2, hb2, [0, 1], jt, [0, 1], bo))
    bt2.close()
    # dense: keys 1000..40999 with duplicates, INT32
    dk = rng.integers(1000, 41000, nb).astype(np.int32)
    spec1 = abi.JoinSpec([abi.INT32], [abi.INT32, abi.INT32], [0])
    bcols1 = [Column.ints(dk), Column.ints(rng.integers(-100, 100, nb))]
    bt1 = make_build(dev, spec1, bcols1, nb)
    oj1 = OracleJoin(spec1, bcols1, nb)
    hb1 = [([Column.ints(rng.integers(900, 41100, n)), Column.ints(rng.integers(-100, 100, n))], n,
            pack_bits(rng.random(n) < 0.7)) for n in (TILE * 3 + 1, 17, 40_000)]
    res = abi.VrJoinResidual(abi.CMP_LT, 0, 1, 1, 1, 0, 5)
    for jt in (abi.INNER, abi.OUTER, abi.LEFT_SEMI, abi.LEFT_ANTI):
        bo = [1, 0] if jt in (abi.INNER, abi.OUTER) else []
        assert_same(batched(dev, bt1, hb1, [0], jt, [0, 1], bo), oracle_concat(oj1, hb1, [0], jt, [0, 1], bo))
        assert_same(batched(dev, bt1, hb1, [0], jt, [0, 1], bo, residual=res),
                    oracle_concat(oj1, hb1, [0], jt, [0, 1], bo, residual=res))
    bt1.close()


def fmix64(x):
    x = np.asarray(x, np.uint64)
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(33)
        x *= np.uint64(0xff51afd7ed558ccd)
        x ^= x >> np.uint64(33)
        x *= np.uint64(0xc4ceb9fe1a85ec53)
        x ^= x >> np.uint64(33)
    return x


def test_bucket_chains_wrap_and_stop(dev):
    """Forty build keys and a crowd of missing probe keys share the last bucket as their home: the
    build walks on past full buckets (wrapping to bucket 0) and marks them passed; hits are found
    down the chain and misses stop at the first bucket nothing walked past."""
    nb_rows = 4000
    cells = int(nb_rows / 0.50) + 1
    nbk = (cells + 7) // 8
    rng = np.random.default_rng(9)
    cand = rng.integers(-2 ** 62, 2 ** 62, 4_000_000).astype(np.int64)
    home = ((fmix64(cand.view(np.uint64)) >> np.uint64(32)) * np.uint64(nbk)) >> np.uint64(32)
    crowd = cand[home == nbk - 1]
    assert len(crowd) >= 80
    others = cand[home != nbk - 1][: nb_rows - 40]
    bkeys = np.concatenate([others[: nb_rows // 2], crowd[:40], others[nb_rows // 2:]]).astype(np.int64)
    spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32], [0], dense_index=-1)
    bcols = [Column.longs(bkeys), Column.ints(np.arange(len(bkeys)))]
    bt = make_build(dev, spec, bcols, len(bkeys))
    oj = OracleJoin(spec, bcols, len(bkeys))
    probe = np.concatenate([crowd, bkeys, rng.integers(-2 ** 62, 2 ** 62, 5000)]).astype(np.int64)
    rng.shuffle(probe)
    hb = [([Column.longs(probe), Column.ints(np.arange(len(probe)))], len(probe), None)]
    for jt in (abi.INNER, abi.LEFT_SEMI, abi.LEFT_ANTI):
        bo = [1] if jt == abi.INNER else []
        assert_same(batched(dev, bt, hb, [0], jt, [0], bo), oracle_concat(oj, hb, [0], jt, [0], bo))
    # the per-batch entry point takes the same lookups
    p, b, _, r = bt.probe([Column.longs(probe)], [0], len(probe), abi.INNER)
    op, ob, _, orows = oj.probe([Column.longs(probe)], [0], len(probe), abi.INNER, capacity=len(probe))
    assert r == orows and (p == op).all() and (b == ob).all()
    bt.close()


def test_batches_capacity_and_errors(dev):
    """A too-small output reports the rows it needs; host outputs and host batches are refused."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(2)
    bkeys = rng.integers(0, 1000, 500).astype(np.int64)
    spec = abi.JoinSpec([abi.INT64], [abi.INT64], [0])
    bt = make_build(dev, spec, [Column.longs(bkeys)], 500)
    hb = probe_batches(rng, bkeys, [3000, 5000], hit=1.0)
    exp = oracle_concat(OracleJoin(spec, [Column.longs(bkeys)], 500), hb, [0], abi.INNER)
    with pytest.raises(vr.VrError) as e:
        batched(dev, bt, hb, [0], abi.INNER, capacity=100)
    assert e.value.status == abi.VR_ERR_CAPACITY
    got = batched(dev, bt, hb, [0], abi.INNER, capacity=exp["rows"])
    assert_same(got, exp)
    pb = vr.ProbeBatches([([Column.longs(bkeys)], 500)])
    with pytest.raises(vr.VrError) as e:
        d = torch.zeros(1000, dtype=torch.int32, device="cuda")
        bt.probe_batches(pb, [0], abi.LEFT_SEMI, capacity=1000, probe_idx=d)
    assert e.value.status == abi.VR_ERR_UNSUPPORTED
    bt.close()


def test_concurrent_probes_of_one_table(dev):
    """A finished table probed from two threads at once, each on its own stream (the broadcast
    build side shared by a task's threads): each thread's results equal the oracle's."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(4)
    nb = 50_000
    bkeys = rng.integers(-2 ** 62, 2 ** 62, nb).astype(np.int64)
    spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32], [0])
    bcols = [Column.longs(bkeys), Column.ints(np.arange(nb))]
    bt = make_build(dev, spec, bcols, nb)
    oj = OracleJoin(spec, bcols, nb)
    work = [probe_batches(np.random.default_rng(40 + t), bkeys, [65536, 30_000, TILE + 3]) for t in range(2)]
    exps = [[oracle_concat(oj, w, [0], jt, [0, 1], [1] if jt == abi.INNER else []) for jt in (abi.INNER, abi.LEFT_SEMI)]
            for w in work]
    errors = []

    def run(t):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for rep in range(3):
                    for q, jt in enumerate((abi.INNER, abi.LEFT_SEMI)):
                        got = batched(dev, bt, work[t], [0], jt, [0, 1], [1] if jt == abi.INNER else [])
                        assert_same(got, exps[t][q])
                        # the per-batch entry point from the same thread, host outputs
                        cols, n, sel = work[t][0]
                        p, b, _, r = bt.probe(cols, [0], n, jt, sel)
                        assert r == exps[t][q]["per"][0] and (p == exps[t][q]["p"][:r]).all()
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    th = [threading.Thread(target=run, args=(t,)) for t in range(2)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errors, errors
    bt.close()
