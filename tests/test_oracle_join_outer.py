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


def wrap32(x):
    return ((x + 2 ** 31) % 2 ** 32) - 2 ** 31


def test_residual_against_nested_loop():
    """Pairs by key equality and a residual `left + lo <op> right + ro` (PairPredicate.test,
    VectorHashJoinExec.scala:1287-1303: both operands non-null, the offset added in the lane's width,
    wrapping) for every join type, against a nested loop."""
    rng = np.random.default_rng(19)
    ops = [abi.CMP_EQ, abi.CMP_NE, abi.CMP_LT, abi.CMP_LE, abi.CMP_GT, abi.CMP_GE]
    cases = 0
    for trial in range(300):
        nb, npr = int(rng.integers(0, 7)), int(rng.integers(0, 7))
        bkeys = [None if rng.random() < 0.15 else int(rng.integers(0, 3)) for _ in range(nb)]
        pkeys = [None if rng.random() < 0.15 else int(rng.integers(0, 3)) for _ in range(npr)]
        edge = [2 ** 31 - 1, -2 ** 31, 0, 5, -5]
        bval = [None if rng.random() < 0.15 else edge[int(rng.integers(0, 5))] for _ in range(nb)]
        pval = [None if rng.random() < 0.15 else edge[int(rng.integers(0, 5))] for _ in range(npr)]
        op = ops[trial % 6]
        lo, ro = int(rng.choice([0, 1, -1])), int(rng.choice([0, 2, -3]))
        sides = [(0, 1), (1, 1), (0, 0), (1, 0)][trial % 4]  # (left side, right side)
        spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32], [0], track_matched=True)
        bcols = [col(bkeys), Column.ints([0 if v is None else v for v in bval], nulls=[v is None for v in bval])]
        pcols = [col(pkeys), Column.ints([0 if v is None else v for v in pval], nulls=[v is None for v in pval])]
        oj = OracleJoin(spec, bcols, nb)
        res = abi.residual(op, (sides[0], 1), (sides[1], 1), lo, ro)

        def val(side, i, r):
            return pval[i] if side == 0 else bval[r]

        def passes(i, r):
            x, y = val(sides[0], i, r), val(sides[1], i, r)
            if x is None or y is None:
                return False
            x, y = wrap32(x + lo), wrap32(y + ro)
            return [x == y, x != y, x < y, x <= y, x > y, x >= y][op]

        hits = {i: [r for r in range(nb) if pkeys[i] is not None and bkeys[r] == pkeys[i] and passes(i, r)]
                for i in range(npr)}
        for jt in (abi.INNER, abi.OUTER, abi.LEFT_SEMI, abi.LEFT_ANTI):
            p, b, _, r = oj.probe(pcols, [0], npr, jt, capacity=max(1, npr * max(1, nb)), residual=res)
            if jt == abi.INNER:
                want = [(i, x) for i in range(npr) for x in hits[i]]
            elif jt == abi.OUTER:
                want = [(i, x) for i in range(npr) for x in (hits[i] or [-1])]
            else:
                keep = [i for i in range(npr) if bool(hits[i]) == (jt == abi.LEFT_SEMI)]
                assert p.tolist() == keep
                continue
            assert list(zip(p.tolist(), b.tolist())) == want, (trial, jt)
        paired = {x for i in range(npr) for x in hits[i]}
        p, b, _, r = oj.unmatched_build(capacity=max(1, nb))
        assert b.tolist() == [x for x in range(nb) if x not in paired]
        cases += 1
    assert cases == 300
