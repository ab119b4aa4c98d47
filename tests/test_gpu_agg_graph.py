"""A captured batched update (vr_agg_graph_*) replayed several times against the CPU oracle and the
eager batched update (MI355X).

The workload is the uneven batched group-by's shape, small: a filter (int32 < 30% threshold), a
composite key (int32, int64 Zipf, dictionary string), SUM decimal(15,2), COUNT(*), MAX double with
special values.  Between replays the program does eager work of its own and synchronizes, and one
replay runs over new contents of the same buffers; every replay's groups are checked."""
import numpy as np
import pytest

from gpu_util import compare_groups, group_rows, out_roles
from oracle.oracle import OracleAgg
from vecruntime_amd import abi
from vecruntime_amd.columns import Column, pack_bits

pytestmark = pytest.mark.gpu

THR = int(0.3 * 2 ** 31)


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    from vecruntime_amd import vr
    d = vr.Device(0)
    yield d
    d.close()


def _spec():
    return abi.AggSpec([abi.INT32, abi.INT64, abi.UTF8], [abi.INT32, abi.INT64, abi.FLOAT64],
                       [abi.agg_desc(abi.SUM_WIDE, 1, buffer_precision=25, buffer_scale=2),
                        abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.MAX, 2)], key_nullable=[False] * 3)


def _fill(rng, n, nwords):
    f = rng.integers(0, 2 ** 31, n).astype(np.int32)
    k1 = rng.integers(1, 101, n).astype(np.int32)
    k2 = (rng.zipf(1.2, n) % 5000).astype(np.int64) * 7919
    k3 = rng.integers(0, nwords, n).astype(np.int32)
    dec = rng.integers(-10 ** 13, 10 ** 13, n)
    mx = rng.uniform(-1e6, 1e6, n)
    special = rng.random(n) < 0.002
    mx[special] = rng.choice([np.nan, 0.0, -0.0, np.inf], int(special.sum()))
    return [k1, k2, k3, f, dec, mx]


class Batches:
    """Device buffers of the batches (refilled in place) and the host copies the oracle reads."""

    def __init__(self, rng, lengths, dictionary, nwords):
        import torch
        from vecruntime_amd import vr
        self.lengths, self.dictionary, self.nwords = lengths, dictionary, nwords
        self.host = [_fill(rng, n, nwords) for n in lengths]
        self.dev = [[torch.from_numpy(a).cuda() for a in h] for h in self.host]
        self.batches = []
        for t, n in zip(self.dev, lengths):
            dk = [vr.device_column(abi.INT32, t[0]), vr.device_column(abi.INT64, t[1]),
                  vr.device_column(abi.UTF8, t[2], dictionary=dictionary, length=n)]
            dv = [vr.device_column(abi.INT32, t[3]), vr.device_column(abi.INT64, t[4]),
                  vr.device_column(abi.FLOAT64, t[5])]
            self.batches.append((dk, dv, n))

    def refill(self, rng):
        import torch
        self.host = [_fill(rng, n, self.nwords) for n in self.lengths]
        for t, h in zip(self.dev, self.host):
            for d, a in zip(t, h):
                d.copy_(torch.from_numpy(a))
        torch.cuda.synchronize()

    def oracle(self, spec):
        o = OracleAgg(spec)
        for h, n in zip(self.host, self.lengths):
            k1, k2, k3, f, dec, mx = h
            o.update([Column.ints(k1), Column.longs(k2), Column.dictionary_strings(k3, self.dictionary)],
                     [Column.ints(f), Column.longs(dec), Column.doubles(mx)], n, pack_bits(f < THR))
        return o


def _check(g, o, spec):
    G = o.num_groups()
    assert g.num_groups() == G
    gk, go = g.emit()
    ok, oo = o.emit()
    dicts = {2: g.key_dictionary(2)}
    odicts = {2: o.key_dictionary(2)}
    worst = compare_groups(group_rows(gk, go, G, spec, dicts), group_rows(ok, oo, G, spec, odicts), out_roles(spec))
    assert worst < 1e-12
    return G


def test_graph_replays_match_oracle_and_eager(dev):
    import torch
    from vecruntime_amd import vr
    rng = np.random.default_rng(444)
    words = ["".join(chr(97 + x) for x in rng.integers(0, 26, rng.integers(8, 25))) for _ in range(300)]
    dictionary = Column.strings(words)
    lengths = [int(x) for x in np.exp(rng.uniform(np.log(1000), np.log(60000), 20))] + [0, 1]
    data = Batches(rng, lengths, dictionary, len(words))
    spec = _spec()
    eager = vr.GroupedAggregate(dev, spec)
    eager.update_batches(data.batches, filter_col=0, threshold=THR)
    o = data.oracle(spec)
    G = _check(eager, o, spec)
    g = vr.GroupedAggregate(dev, spec)
    graph = g.capture(data.batches, max_groups=sum(lengths), filter_col=0, threshold=THR)
    scratch = torch.zeros(1 << 20, device="cuda")
    for rep in range(4):
        if rep == 2:  # new contents of the same buffers: the replay folds them
            data.refill(rng)
            o = data.oracle(spec)
        graph.launch()
        scratch.add_(1.0)  # the program's own eager work between replays
        torch.cuda.synchronize()
        assert _check(g, o, spec) > 0
        if rep < 2:
            assert g.num_groups() == G
    graph.close()
    eager.close()
    g.close()


def test_graph_over_a_too_small_table_reports_oom_and_a_moved_table_is_refused(dev):
    from vecruntime_amd import vr
    rng = np.random.default_rng(9)
    words = ["w%d" % i for i in range(50)]
    dictionary = Column.strings(words)
    data = Batches(rng, [50_000, 40_000], dictionary, len(words))
    spec = _spec()
    g = vr.GroupedAggregate(dev, spec)
    graph = g.capture(data.batches, max_groups=16, filter_col=0, threshold=THR)  # far fewer than the groups
    graph.launch()
    with pytest.raises(vr.VrError) as e:
        g.num_groups()
    assert e.value.status == abi.VR_ERR_OOM
    graph.close()
    g.close()
    # an eager update that grows the table moves it: the graph captured before is refused
    g = vr.GroupedAggregate(dev, spec)
    graph = g.capture(data.batches[:1], max_groups=100, filter_col=0, threshold=THR)
    g.reset()
    g.update_batches(data.batches, filter_col=0, threshold=THR)  # about 20k groups: the table grows
    with pytest.raises(vr.VrError) as e:
        graph.launch()
    assert e.value.status == abi.VR_ERR_STATE
    graph.close()
    g.close()
    # a graph that outlives its aggregate refuses to launch and is still destroyed cleanly
    g = vr.GroupedAggregate(dev, spec)
    graph = g.capture(data.batches[:1], max_groups=60_000, filter_col=0, threshold=THR)
    graph.launch()
    g.close()
    with pytest.raises(vr.VrError) as e:
        graph.launch()
    assert e.value.status == abi.VR_ERR_STATE
    graph.close()
