"""The oracle's outer join pinned by brute force: every 0-3 build x 0-3 probe combination over keys
{null, 1, 2}, probed in one or two batches, against a nested loop that follows the outer path of
VectorHashJoinExec.scala:1160-1175 (a streamed row's build rows in build order, or (row, -1) when
it pairs with nothing) and emitUnmatchedBuild (:1457-1484, the build rows never paired, in build
order)."""
import itertools

import numpy as np

from oracle.oracle import OracleJoin
from vecruntime_amd import abi
from vecruntime_amd.columns import Column

VALS = (None, 1, 2)


def col(keys):
    return Column.longs([0 if k is None else k for k in keys], nulls=[k is None for k in keys])


def brute(bkeys, probes):
    """probes: list of batches of keys.  Returns (per-batch outer pairs, unmatched build pairs)."""
    seen = set()
    out = []
    for pk in probes:
        pairs = []
        for i, k in enumerate(pk):
            hits = [r for r, b in enumerate(bkeys) if k is not None and b is not None and b == k]
            if hits:
                pairs += [(i, r) for r in hits]
                seen.update(hits)
            else:
                pairs.append((i, -1))
        out.append(pairs)
    return out, [(-1, r) for r in range(len(bkeys)) if r not in seen]


def test_outer_and_unmatched_build_exhaustive():
    cases = 0
    for nb in range(4):
        for bkeys in itertools.product(VALS, repeat=nb):
            for npr in range(4):
                for pkeys in itertools.product(VALS, repeat=npr):
                    for split in range(npr + 1):
                        probes = [list(pkeys[:split]), list(pkeys[split:])]
                        spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32], [0], track_matched=True)
                        bcols = [col(bkeys), Column.ints(np.arange(nb, dtype=np.int32) + 100)]
                        oj = OracleJoin(spec, bcols, nb)
                        want, want_un = brute(list(bkeys), probes)
                        for pk, exp in zip(probes, want):
                            n = len(pk)
                            p, b, outs, r = oj.probe([col(pk)], [0], n, abi.OUTER, probe_out=[0], build_out=[1],
                                                     capacity=max(1, 4 * n))
                            assert list(zip(p.tolist(), b.tolist())) == exp
                            # padded rows: build payload null; matched rows: the payload
                            bv = outs[1].values(r)
                            assert bv == [None if j < 0 else j + 100 for _, j in exp]
                            pv = outs[0].values(r)
                            assert pv == [pk[i] for i, _ in exp]
                        p, b, outs, r = oj.unmatched_build(probe_types=[abi.INT64], build_out=[1], capacity=max(1, nb))
                        assert list(zip(p.tolist(), b.tolist())) == want_un
                        assert outs[0].values(r) == [None] * r
                        assert outs[1].values(r) == [j + 100 for _, j in want_un]
                        cases += 1
    assert cases > 2000
