"""Shuffle partitioning on the device against the oracle (MI355X): vr_hash_partition_ids
(Spark's Pmod(Murmur3Hash(keys, seed), n), PartitionKernels.hashPartitionIds), vr_round_robin_ids
and vr_partition_rows (PartitionKernels.partitionOrder / partitionDestinations).  Bit-exact: every
hash, id, row order, destination and partition start."""
import numpy as np
import pytest

from oracle import oracle
from test_gpu_filter import on_device
from test_oracle_filter import utf8_column
from test_oracle_partition import ALL_KINDS, make_key
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


@pytest.mark.parametrize("kind", ALL_KINDS)
def test_each_kind_matches_oracle(dev, kind):
    from vecruntime_amd import vr
    rng = np.random.default_rng(200 + kind)
    for n in (1, 63, 64, 65, 20_011):
        print("hash kind %d n %d" % (kind, n), flush=True)
        nulls = rng.random(n) < 0.15
        col, _ = make_key(kind, rng, n, nulls)
        keep = []
        dcol = on_device(col, keep)
        for parts, seed in ((200, 42), (1, 42), (2 ** 31 - 1, 42), (1000, -7)):
            want_ids, want_h = oracle.hash_partition_ids([col], [kind], n, parts, seed)
            ids, hs = vr.hash_partition_ids(dev, [dcol], [kind], n, parts, seed, want_hashes=True)
            assert np.array_equal(hs.cpu().numpy(), want_h), (kind, n, parts, seed)
            assert np.array_equal(ids.cpu().numpy(), want_ids), (kind, n, parts, seed)


def test_many_keys_dictionary_and_errors(dev):
    """Eight keys (every kind and a dictionary string key with its dictionary on the host), then the
    same eight twice and three times over (24 keys), with and without the hashes asked for; no
    keys; a code outside its dictionary; a kind that does not fit its lane."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(77)
    n = 50_001
    nulls = [rng.random(n) < 0.1 for _ in ALL_KINDS]
    cols = [make_key(k, rng, n, z)[0] for k, z in zip(ALL_KINDS, nulls)]
    words = ["", "a", "abcd", "été", "x" * 13, "фыва"]
    codes = rng.integers(0, len(words), n)
    cols.append(utf8_column(words, rng.random(n) < 0.1, codes))
    kinds = ALL_KINDS + [abi.HASH_UTF8]
    keep = []
    dcols = [on_device(c, keep) for c in cols]
    for reps in (1, 2, 3):
        print("keys %d" % (len(kinds) * reps), flush=True)
        want_ids, want_h = oracle.hash_partition_ids(cols * reps, kinds * reps, n, 200)
        for hashes in (True, False):
            ids, hs = vr.hash_partition_ids(dev, dcols * reps, kinds * reps, n, 200, want_hashes=hashes)
            assert np.array_equal(ids.cpu().numpy(), want_ids), (reps, hashes)
            if hashes:
                assert np.array_equal(hs.cpu().numpy(), want_h), reps
    ids, hs = vr.hash_partition_ids(dev, [], [], 10, 16, want_hashes=True)
    assert hs.cpu().tolist() == [42] * 10 and ids.cpu().tolist() == [10] * 10
    bad = on_device(utf8_column(words, np.zeros(4, bool), [0, 1, 6, 2]), keep)
    with pytest.raises(vr.VrError) as e:
        vr.hash_partition_ids(dev, [bad], [abi.HASH_UTF8], 4, 8)
    assert e.value.status == abi.VR_ERR_INVALID_ARGUMENT
    with pytest.raises(vr.VrError) as e:  # a kind that does not fit the lane
        vr.hash_partition_ids(dev, [dcols[0]], [abi.HASH_LONG], n, 8)
    assert e.value.status == abi.VR_ERR_INVALID_ARGUMENT
    torch.cuda.synchronize()


def test_round_robin_matches_oracle(dev):
    from vecruntime_amd import vr
    for n, parts, start in ((10, 4, 2), (0, 3, 1), (100_003, 200, 199), (5, 1, 0)):
        want, wnext = oracle.round_robin_ids(n, parts, start)
        ids, nxt = vr.round_robin_ids(dev, n, parts, start)
        assert ids.cpu().numpy().tolist() == want.tolist() and nxt == wnext


@pytest.mark.parametrize("parts", [1, 2, 200, 4096, 1 << 20])
def test_partition_rows_matches_oracle(dev, parts):
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(parts)
    for n in (0, 1, 65, 100_003, 1_000_000):
        print("partition rows parts %d n %d" % (parts, n), flush=True)
        ids = rng.integers(0, parts, n).astype(np.int32)
        if n > 10:  # skewed: half the rows in one partition
            ids[rng.random(n) < 0.5] = parts - 1
        rows, dest, starts = oracle.partition_rows(ids, parts)
        dids = torch.from_numpy(ids).cuda() if n else torch.zeros(1, dtype=torch.int32, device="cuda")
        r, d, s = vr.partition_rows(dev, dids, n, parts)
        assert np.array_equal(s.cpu().numpy(), starts)
        assert np.array_equal(r.cpu().numpy(), rows) and np.array_equal(d.cpu().numpy(), dest)
        r2, d2, s2 = vr.partition_rows(dev, dids, n, parts, want_rows=False, want_dest=False)
        assert r2 is None and d2 is None and np.array_equal(s2.cpu().numpy(), starts)


def test_partition_rows_rejects_an_id_outside(dev):
    import torch
    from vecruntime_amd import vr
    for parts in (5, 1000):  # the counting sort, then the radix sort
        for bad in (parts, -1, 1 << 30):
            ids = torch.tensor([0, 1, bad, 2], dtype=torch.int32, device="cuda")
            with pytest.raises(vr.VrError) as e:
                vr.partition_rows(dev, ids, 4, parts)
            assert e.value.status == abi.VR_ERR_INVALID_ARGUMENT
    r, d, s = vr.partition_rows(dev, torch.tensor([2, 0, 1, 0, 2], dtype=torch.int32, device="cuda"), 5, 4)
    assert r.cpu().tolist() == [1, 3, 2, 0, 4] and d.cpu().tolist() == [3, 0, 2, 1, 4]
    assert s.cpu().tolist() == [0, 2, 3, 5, 5]


def test_shuffle_map_side_at_size(dev):
    """The map side of a 200-partition shuffle over 20M rows with a long and a double key: ids
    from the hash, then the grouping; properties checked at full size on the device (the
    ids of the grouped rows are sorted, rows keep their order within a partition, starts are the
    cumulative counts, dest inverts rows) and the first 200,000 rows against the oracle."""
    import torch
    from vecruntime_amd import vr
    n, parts = 20_000_000, 200
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    a = torch.randint(-2 ** 62, 2 ** 62, (n,), dtype=torch.int64, device="cuda", generator=g)
    b = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    keys = [vr.device_column(abi.INT64, a), vr.device_column(abi.FLOAT64, b)]
    ids, _ = vr.hash_partition_ids(dev, keys, [abi.HASH_LONG, abi.HASH_DOUBLE], n, parts)
    rows, dest, starts = vr.partition_rows(dev, ids, n, parts)
    counts = torch.bincount(ids.long(), minlength=parts)
    assert starts[0].item() == 0 and torch.equal(starts[1:], torch.cumsum(counts, 0))
    sid = ids[rows.long()]
    assert bool((sid[1:] >= sid[:-1]).all())
    same = sid[1:] == sid[:-1]
    assert bool((rows[1:][same] > rows[:-1][same]).all())
    assert torch.equal(dest[rows.long()], torch.arange(n, dtype=torch.int32, device="cuda"))
    m = 200_000
    host = [Column.longs(a[:m].cpu().numpy()), Column.doubles(b[:m].cpu().numpy())]
    want, _ = oracle.hash_partition_ids(host, [abi.HASH_LONG, abi.HASH_DOUBLE], m, parts)
    assert np.array_equal(ids[:m].cpu().numpy(), want)
