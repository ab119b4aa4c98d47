"""The query server (include/vr_serve.h) on MI355X: every query's result in a concurrent mix equals
its result run alone, and both match the CPU oracle."""
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


def scramble(x):
    return (np.asarray(x, np.uint64) * np.uint64(0x9E3779B97F4A7C15) ^ np.uint64(0x5851F42D4C957F2D)).astype(np.int64)


class AggCase:
    """A group-by query: host columns for the oracle, device columns for the server."""

    def __init__(self, seed, groups, lengths, kind):
        import torch
        from vecruntime_amd import vr
        rng = np.random.default_rng(seed)
        self.words = None
        if kind == "long":
            self.spec = abi.AggSpec([abi.INT64], [abi.INT64, abi.FLOAT64, abi.INT32],
                                    [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_LONG, 0),
                                     abi.agg_desc(abi.SUM_DOUBLE, 1), abi.agg_desc(abi.MIN, 2),
                                     abi.agg_desc(abi.MAX, 2), abi.agg_desc(abi.AVG_DOUBLE, 1)],
                                    key_nullable=[False])
        elif kind == "composite":
            self.spec = abi.AggSpec([abi.INT32, abi.INT64], [abi.INT64, abi.FLOAT64, abi.INT32],
                                    [abi.agg_desc(abi.SUM_WIDE, 0, buffer_precision=22, buffer_scale=2),
                                     abi.agg_desc(abi.AVG_WIDE, 0, buffer_precision=22, buffer_scale=2),
                                     abi.agg_desc(abi.COUNT, 2)], key_nullable=[False, False])
        else:  # dictionary-encoded string key
            self.words = ["w%05d_%s" % (i, "x" * (i % 13)) for i in range(max(groups, 1))]
            self.dictionary = Column.strings(self.words)
            self.spec = abi.AggSpec([abi.UTF8], [abi.INT64, abi.FLOAT64, abi.INT32],
                                    [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_DOUBLE, 1),
                                     abi.agg_desc(abi.MAX, 1)], key_nullable=[False])
        self.groups = groups
        self.host, self.batches, self._t = [], [], []
        for n in lengths:
            g = rng.integers(0, groups, n)
            v0 = rng.integers(-10 ** 9, 10 ** 9, n)
            v1 = rng.lognormal(0, 1, n)
            v2 = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
            n2 = rng.random(n) < 0.05
            if kind == "long":
                hk = [Column.longs(scramble(g))]
                karr = [(abi.INT64, scramble(g))]
            elif kind == "composite":
                k1 = (g % 97).astype(np.int32)
                k2 = scramble(g // 97)
                hk = [Column.ints(k1), Column.longs(k2)]
                karr = [(abi.INT32, k1), (abi.INT64, k2)]
            else:
                k = g.astype(np.int32)
                hk = [Column.dictionary_strings(k, self.dictionary)]
                karr = [(abi.UTF8, k)]
            hv = [Column.longs(v0), Column.doubles(v1), Column.ints(v2, nulls=n2)]
            self.host.append((hk, hv, n))
            dk = []
            for t, a in karr:
                ta = torch.from_numpy(a).cuda()
                self._t.append(ta)
                if t == abi.UTF8:
                    dk.append(vr.device_column(abi.UTF8, ta, dictionary=self.dictionary, length=n))
                else:
                    dk.append(vr.device_column(t, ta))
            tv = [torch.from_numpy(a).cuda() for a in (v0, v1, v2)]
            tn = torch.from_numpy(pack_bits(~n2)).cuda()
            self._t += tv + [tn]
            dv = [vr.device_column(abi.INT64, tv[0]), vr.device_column(abi.FLOAT64, tv[1]),
                  vr.device_column(abi.INT32, tv[2], validity=tn)]
            self.batches.append((dk, dv, n))

    def submit(self, srv):
        return srv.submit_agg(self.spec, self.batches, expected_groups=self.groups)

    def rows_of(self, srv, ticket, res):
        cols = srv.fetch(ticket, res)
        nk = len(self.spec.key_types)
        dicts = {0: srv.key_dictionary(res, 0)} if self.words is not None else {}
        return group_rows(cols[:nk], cols[nk:], res.rows, self.spec, dicts)

    def oracle_rows(self):
        o = OracleAgg(self.spec)
        for hk, hv, n in self.host:
            o.update(hk, hv, n, None)
        G = o.num_groups()
        ok, oo = o.emit()
        dicts = {0: o.key_dictionary(0)} if self.words is not None else {}
        return group_rows(ok, oo, G, self.spec, dicts)


class JoinCase:
    def __init__(self, seed, nb, npr, join_type):
        import torch
        from vecruntime_amd import vr
        rng = np.random.default_rng(seed)
        self.join_type = join_type
        bkeys = scramble(rng.permutation(nb) + seed * 10 ** 7)
        dup = rng.random(nb) < 0.1  # some duplicate keys
        bkeys[dup] = bkeys[rng.integers(0, nb, int(dup.sum()))]
        bpay = rng.integers(-2 ** 31, 2 ** 31, nb).astype(np.int32)
        hit = rng.random(npr) < 0.7
        pkeys = np.where(hit, bkeys[rng.integers(0, nb, npr)], scramble(np.arange(npr) + 2 ** 40 + seed))
        ppay = rng.integers(0, 1000, npr).astype(np.int64)
        self.nb, self.npr = nb, npr
        self.spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32], [0], capacity_hint=nb)
        self.host_build = [Column.longs(bkeys), Column.ints(bpay)]
        self.host_probe = [Column.longs(pkeys), Column.longs(ppay)]
        self._t = [torch.from_numpy(a).cuda() for a in (bkeys, bpay, pkeys, ppay)]
        self.build = [vr.device_column(abi.INT64, self._t[0]), vr.device_column(abi.INT32, self._t[1])]
        self.probe = [vr.device_column(abi.INT64, self._t[2]), vr.device_column(abi.INT64, self._t[3])]
        self.build_out = [1] if join_type == abi.INNER else []

    def submit(self, srv):
        # a small output hint on purpose: the server grows the output and probes again
        return srv.submit_join(self.spec, self.build, self.nb, self.probe, [0], self.npr, self.join_type,
                               probe_out=[0, 1], build_out=self.build_out, expected_rows=self.npr // 4)

    def rows_of(self, srv, ticket, res):
        cols = srv.fetch(ticket, res)
        return [c.data[:res.rows].tolist() for c in cols]

    def oracle_rows(self):
        oj = OracleJoin(self.spec, self.host_build, self.nb)
        p, b, _, _ = oj.probe(self.host_probe, [0], self.npr, self.join_type)
        pk, pp = self.host_probe[0].data, self.host_probe[1].data
        out = [pk[p].tolist(), pp[p].tolist()]
        if self.join_type == abi.INNER:
            out.append(self.host_build[1].data[b].tolist())
        return out


def compare_case(case, got, want):
    if isinstance(case, AggCase):
        worst = compare_groups(got, want, out_roles(case.spec), rel=1e-12)
        assert worst < 1e-12
    else:
        assert got == want


def make_cases():
    return [AggCase(1, 1000, [50_000, 120_000, 7], "long"),
            AggCase(2, 200_000, [300_000, 250_000], "long"),
            AggCase(3, 30_000, [100_000, 100_000, 100_000], "composite"),
            AggCase(4, 800, [90_000, 10_000], "string"),
            JoinCase(5, 50_000, 300_000, abi.INNER),
            JoinCase(6, 40_000, 200_000, abi.LEFT_SEMI),
            JoinCase(7, 40_000, 200_000, abi.LEFT_ANTI)]


def test_mix_equals_alone_equals_oracle(dev):
    """Each query alone (one in flight), then every query three times over submitted at once on 4
    workers: the mix's results equal the lone results, and both equal the oracle's."""
    from vecruntime_amd.serve import Server
    cases = make_cases()
    srv = Server(dev, workers=4)
    try:
        alone = []
        for c in cases:
            t = c.submit(srv)
            res = srv.wait(t)
            assert res.status == abi.VR_OK
            alone.append(c.rows_of(srv, t, res))
            srv.release(t)
        for c, got in zip(cases, alone):
            compare_case(c, got, c.oracle_rows())
        tickets = [(c, i, c.submit(srv)) for _ in range(3) for i, c in enumerate(cases)]
        for c, i, t in tickets:
            res = srv.wait(t)
            assert res.status == abi.VR_OK
            assert res.latency_ms >= res.run_ms >= 0
            compare_case(c, c.rows_of(srv, t, res), alone[i])
            srv.release(t)
        st = srv.stats()
        assert st["completed"] == st["submitted"] == 4 * len(cases)
        assert st["failed"] == 0
        assert st["peak_inflight"] >= 2
        assert st["reserved_bytes"] == 0
    finally:
        srv.close()


def test_budget_admits_one_at_a_time(dev):
    """A budget smaller than any query: each query is admitted alone (never two in flight), and
    results are unchanged."""
    from vecruntime_amd.serve import Server
    cases = make_cases()[:2]
    srv = Server(dev, workers=4, hbm_budget=1 << 20)
    try:
        tickets = [(c, c.submit(srv)) for c in cases for _ in range(2)]
        for c, t in tickets:
            res = srv.wait(t)
            assert res.status == abi.VR_OK
            compare_case(c, c.rows_of(srv, t, res), c.oracle_rows())
            srv.release(t)
        assert srv.stats()["peak_inflight"] == 1
    finally:
        srv.close()


def test_failed_query_reports_its_status(dev):
    """An invalid query (an aggregate reading a value column that does not exist) fails alone; the
    server keeps serving."""
    from vecruntime_amd.serve import Server
    good = make_cases()[0]
    bad_spec = abi.AggSpec([abi.INT64], [abi.INT64], [abi.agg_desc(abi.SUM_LONG, 5)], key_nullable=[False])
    srv = Server(dev, workers=2)
    try:
        tb = srv.submit_agg(bad_spec, [(good.batches[0][0], good.batches[0][1][:1], good.batches[0][2])])
        tg = good.submit(srv)
        rb = srv.wait(tb)
        assert rb.status != abi.VR_OK
        rg = srv.wait(tg)
        assert rg.status == abi.VR_OK
        compare_case(good, good.rows_of(srv, tg, rg), good.oracle_rows())
        srv.release(tb)
        srv.release(tg)
        st = srv.stats()
        assert st["failed"] == 1 and st["reserved_bytes"] == 0
    finally:
        srv.close()


@pytest.mark.parametrize("admission", ["size-aware", "fifo"])
def test_admission_order(dev, admission):
    """One worker busy with a large group-by; a second large one is queued, then three small ones.
    Size-aware admission (aging long enough not to matter here) starts the small ones first;
    arrival-order admission starts the second large one first.  Every result matches the oracle."""
    import time
    from vecruntime_amd.serve import Server, prewarm
    big = AggCase(11, 100_000, [2_000_000] * 3, "long")
    smalls = [AggCase(12 + i, 500, [20_000], "long") for i in range(3)]
    prewarm(dev, [big.spec])
    mode = abi.ADMIT_SIZE_AWARE if admission == "size-aware" else abi.ADMIT_FIFO
    srv = Server(dev, workers=1, admission=mode, aging_ms=10 ** 9)
    try:
        order = [big, big] + smalls
        sub, tickets = [], []
        for c in order:
            sub.append(time.perf_counter())
            tickets.append(c.submit(srv))
        admitted = []
        for c, t, at in zip(order, tickets, sub):
            res = srv.wait(t)
            assert res.status == abi.VR_OK
            admitted.append(at * 1e3 + res.queue_ms)
            compare_case(c, c.rows_of(srv, t, res), c.oracle_rows())
            srv.release(t)
        if admission == "size-aware":
            assert max(admitted[2:]) < admitted[1]
        else:
            assert admitted[1] < min(admitted[2:])
        assert srv.stats()["peak_inflight"] == 1
    finally:
        srv.close()
