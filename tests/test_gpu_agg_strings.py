"""UTF8 group keys through the device dictionary (MI355X) against the CPU oracle: plain columns in
host and in device memory, and dictionary-encoded ones (host and device codes), mixed in one
handle, alone and beside an INT64 key, with null keys; the batched entry point over device-resident
plain columns; the emitted ids resolved through vr_agg_key_dictionary.

The reference maps a UTF8 key to its table dictionary's id before it hashes the key
(GroupKeyTable.toIds, kernels/src/main/java/io/vecruntime/kernels/GroupKeyTable.java:1308-1430);
the device dictionary gives ids in its own order, so groups are compared by their decoded key
strings.  Parity bar: the same groups, exact integer outputs, double sums within 1e-12."""
import numpy as np
import pytest

from gpu_util import compare_groups, group_rows, out_roles
from oracle.oracle import OracleAgg
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


def words(rng, m):
    out = []
    for i in range(m):
        n = int(rng.integers(0, 30))
        out.append("".join(chr(int(c)) for c in rng.integers(97, 123, n)) + ("é" if i % 11 == 0 else ""))
    return sorted(set(out))


def to_device(col, keep):
    import torch

    def t(a):
        if a is None:
            return None
        x = torch.from_numpy(np.ascontiguousarray(a)).cuda()
        keep.append(x)
        return x
    return Column(col.type, t(col.data), t(col.validity), offsets=t(col.offsets), dictionary=col.dictionary,
                  length=col.length, mem=abi.MEM_DEVICE)


def string_batch(rng, vocab, n, form, keep):
    """(host Column for the oracle, the Column the device gets) of n keys drawn from vocab."""
    idx = rng.integers(0, len(vocab), n)
    nulls = rng.random(n) < 0.03
    vals = [vocab[i] for i in idx]
    plain = Column.strings(vals)
    host = Column(abi.UTF8, plain.data, pack_bits(~nulls), offsets=plain.offsets, length=n)
    if form in ("dict-host", "dict-device"):
        sub = sorted(set(vals) | set(vocab[: len(vocab) // 3]))  # the dictionary holds unused entries too
        d = Column.strings(sub)
        codes = np.array([sub.index(v) for v in vals], np.int32)
        host = Column.dictionary_strings(codes, d, nulls=nulls)
    col = host if form in ("plain-host", "dict-host") else to_device(host, keep)
    return host, col


@pytest.mark.parametrize("with_long", [False, True])
@pytest.mark.parametrize("deterministic", [False, True])
def test_string_keys_every_form(dev, with_long, deterministic):
    from vecruntime_amd import vr
    rng = np.random.default_rng(31 + with_long + 2 * deterministic)
    vocab = words(rng, 3000)
    kt = [abi.UTF8] + ([abi.INT64] if with_long else [])
    spec = abi.AggSpec(kt, [abi.INT64, abi.FLOAT64],
                       [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_LONG, 0), abi.agg_desc(abi.SUM_DOUBLE, 1),
                        abi.agg_desc(abi.MAX, 1)], key_nullable=[True] * len(kt), deterministic=deterministic)
    g, o = vr.GroupedAggregate(dev, spec), OracleAgg(spec)
    keep = []
    forms = ["plain-device", "plain-host", "dict-host", "dict-device", "plain-device", "plain-host"]
    for b, form in enumerate(forms):
        n = [50_000, 1, 30_000, 65_536, 7, 20_000][b]
        hk, dk = string_batch(rng, vocab, n, form, keep)
        v0 = rng.integers(-10 ** 9, 10 ** 9, n)
        v1 = rng.lognormal(0, 1, n)
        hks, dks = [hk], [dk]
        if with_long:
            k2 = rng.integers(0, 3, n) * 1_000_003
            hks.append(Column.longs(k2))
            dks.append(Column.longs(k2))
        vals = [Column.longs(v0), Column.doubles(v1)]
        g.update(dks, vals, n)
        o.update(hks, vals, n)
    G = o.num_groups()
    assert g.num_groups() == G
    gk, go = g.emit()
    ok, oo = o.emit()
    worst = compare_groups(group_rows(gk, go, G, spec, {0: g.key_dictionary(0)}),
                           group_rows(ok, oo, G, spec, {0: o.key_dictionary(0)}), out_roles(spec))
    assert worst < 1e-12
    g.close()


def test_batched_device_plain_strings(dev):
    """vr_agg_update_batches over device-resident plain UTF8 batches with the fused filter."""
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(77)
    vocab = words(rng, 20_000)
    spec = abi.AggSpec([abi.UTF8], [abi.INT32, abi.INT64],
                       [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_LONG, 1), abi.agg_desc(abi.MIN, 1)],
                       key_nullable=[True])
    g, o = vr.GroupedAggregate(dev, spec), OracleAgg(spec)
    keep, batches = [], []
    thr = 0
    for n in (40_000, 3, 65_536, 0, 12_345):
        hk, dk = string_batch(rng, vocab, n, "plain-device", keep)
        f = rng.integers(-1000, 1000, n).astype(np.int32)
        v = rng.integers(-10 ** 12, 10 ** 12, n)
        tf, tv = torch.from_numpy(f).cuda(), torch.from_numpy(v).cuda()
        keep += [tf, tv]
        batches.append(([dk], [vr.device_column(abi.INT32, tf), vr.device_column(abi.INT64, tv)], n))
        if n:
            o.update([hk], [Column.ints(f), Column.longs(v)], n, pack_bits(f < thr))
    g.update_batches(batches, filter_col=0, threshold=thr)
    G = o.num_groups()
    assert g.num_groups() == G
    gk, go = g.emit()
    ok, oo = o.emit()
    compare_groups(group_rows(gk, go, G, spec, {0: g.key_dictionary(0)}),
                   group_rows(ok, oo, G, spec, {0: o.key_dictionary(0)}), out_roles(spec))
    g.close()
