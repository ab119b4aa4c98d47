"""Pins the CPU oracle (oracle/cpu_ref.cpp) against the golden fixtures, the pure-Python reference
(oracle/pyref.py), closed forms and planted mistakes.  CPU only."""
import math
import struct
from fractions import Fraction

import numpy as np
import pytest

from golden_util import GROUPED_AGGS, batch_slices, f64_bits, grouped_batches, grouped_spec, load
from oracle import pyref
from oracle.oracle import OracleAgg, OracleError, OracleJoin
from vecruntime_amd import abi
from vecruntime_amd.columns import Column, pack_bits

CARDS = [1, 4, 64, 65, 300, 5000]


def run_grouped(z, dict_encoded):
    batches, strings = grouped_batches(z, dict_encoded)
    agg = OracleAgg(grouped_spec())
    seen = {}
    for keys, values, n, sel in batches:
        ids = agg.update(keys, values, n, sel)
        # stable ids: a key tuple keeps its id across batches, unselected rows get -1
        k1 = keys[0].data
        k1n = ~np.unpackbits(keys[0].validity, bitorder="little", count=n).astype(bool)
        for i in range(n):
            if sel is not None and not (sel[i >> 3] >> (i & 7)) & 1:
                assert ids[i] == -1
                continue
            assert ids[i] >= 0
    return agg


@pytest.mark.parametrize("with_selection", [False, True])
@pytest.mark.parametrize("cardinality", CARDS)
@pytest.mark.parametrize("dict_encoded", [True, False])
def test_grouped_accumulators_match_golden(cardinality, with_selection, dict_encoded):
    """GroupedAggregationTest.groupedAccumulatorsMatchReference / HonourSelection
    (GroupedAggregationTest.java:59-193) on the same draws: every group in first-seen order,
    strict double sums bit for bit."""
    z = load("grouped_agg_c%d_s%d.npz" % (cardinality, int(with_selection)))
    agg = run_grouped(z, dict_encoded)
    G = agg.num_groups()
    assert G == len(z["exp_rows"])
    keys, outs = agg.emit()
    k1ok = keys[0].valid(G)
    assert (~k1ok == z["exp_k1_null"]).all()
    assert (keys[0].data[:G][k1ok] == z["exp_k1"][k1ok]).all()
    tdict = agg.key_dictionary(1)
    raw = z["dict_bytes"].tobytes()
    do = z["dict_offsets"]
    strings = [raw[do[i]:do[i + 1]].decode() for i in range(len(do) - 1)]
    k2ok = keys[1].valid(G)
    got_k2 = [tdict[keys[1].data[g]] if k2ok[g] else None for g in range(G)]
    exp_k2 = [strings[c] if c >= 0 else None for c in z["exp_k2"]]
    assert got_k2 == exp_k2
    dsum, lsum, rows, cnt, dmin, dmax = outs
    assert (~dsum.valid(G) == z["exp_dsum_null"]).all()
    ok = ~z["exp_dsum_null"]
    assert (f64_bits(dsum.data[:G])[ok] == z["exp_dsum_bits"][ok]).all()
    assert (~lsum.valid(G) == z["exp_lsum_null"]).all()
    assert (lsum.data[:G][~z["exp_lsum_null"]] == z["exp_lsum"][~z["exp_lsum_null"]]).all()
    assert (rows.data[:G] == z["exp_rows"]).all()
    assert (cnt.data[:G] == z["exp_count"]).all()
    has = z["exp_has"]
    assert (dmin.valid(G) == has).all() and (dmax.valid(G) == has).all()
    assert (f64_bits(dmin.data[:G])[has] == z["exp_min_bits"][has]).all()
    assert (f64_bits(dmax.data[:G])[has] == z["exp_max_bits"][has]).all()


@pytest.mark.parametrize("cardinality", [1, 3, 64, 300, 5000])
def test_strict_double_sum_bit_identical(cardinality):
    """strictDoubleSumIsBitIdenticalToSequentialReference (GroupedAggregationTest.java:201-244)."""
    z = load("strict_sum_c%d.npz" % cardinality)
    spec = abi.AggSpec([abi.INT32], [abi.FLOAT64], [abi.agg_desc(abi.SUM_DOUBLE, 0), abi.agg_desc(abi.COUNT, 0)])
    agg = OracleAgg(spec)
    for sl, n in batch_slices(z["lengths"]):
        sel = z["sel"][sl]
        agg.update([Column.ints(z["ids"][sl])], [Column.doubles(z["vals"][sl], nulls=z["val_null"][sl])], n,
                   None if sel.all() else pack_bits(sel))
    keys, (s, c) = agg.emit()
    G = agg.num_groups()
    assert (keys[0].data[:G] == z["exp_key"]).all()
    assert (c.data[:G] == z["exp_count"]).all()
    ok = ~z["exp_sum_null"]
    assert (f64_bits(s.data[:G])[ok] == z["exp_sum_bits"][ok]).all()


def test_minmax_special_values():
    """MIN/MAX with NaN greatest, -0.0 == 0.0 keeping the first-seen zero, infinities
    (GroupedAccumulators.DoubleMinMax.offer, GroupedAccumulators.java:1248-1258)."""
    z = load("minmax_special.npz")
    spec = abi.AggSpec([abi.INT32], [abi.FLOAT64],
                       [abi.agg_desc(abi.MIN, 0), abi.agg_desc(abi.MAX, 0), abi.agg_desc(abi.SUM_DOUBLE, 0)])
    agg = OracleAgg(spec)
    for sl, n in batch_slices(z["lengths"]):
        agg.update([Column.ints(z["keys"][sl])], [Column.doubles(z["vals"][sl], nulls=z["val_null"][sl])], n)
    G = agg.num_groups()
    keys, (mn, mx, sm) = agg.emit()
    assert (keys[0].data[:G] == z["exp_key"]).all()
    has = z["exp_has"]
    assert (mn.valid(G) == has).all()
    assert (f64_bits(mn.data[:G])[has] == z["exp_min_bits"][has]).all()
    assert (f64_bits(mx.data[:G])[has] == z["exp_max_bits"][has]).all()
    # sums: NaN payloads are unspecified by IEEE-754 (and canonical in Java), so any NaN matches a NaN
    exp_sum = z["exp_sum_bits"].view(np.float64)
    got_sum = sm.data[:G]
    nan = np.isnan(exp_sum)
    assert (np.isnan(got_sum) == nan)[has].all()
    assert (f64_bits(got_sum)[has & ~nan] == z["exp_sum_bits"][has & ~nan]).all()
    # the fixture does exercise signed zeros and NaN
    bits = set(z["exp_max_bits"][has].tolist()) | set(z["exp_min_bits"][has].tolist())
    assert struct.unpack("<q", struct.pack("<d", float("nan")))[0] in bits or any(
        math.isnan(struct.unpack("<d", struct.pack("<q", b))[0]) for b in bits)


def _limbs_to_ints(arr):
    out = []
    for i in range(len(arr) // 2):
        lo = int(arr[2 * i]) & ((1 << 64) - 1)
        out.append((int(arr[2 * i + 1]) << 64) | lo)
    return out


def test_wide_sums_and_decimal_average():
    """WideLongSum (128-bit with carry and sticky overflow) and the decimal SUM/AVG buffers and
    results (VectorAggregates.scala:318-518) against exact Python integers."""
    z = load("wide_sum.npz")
    aggs = [abi.agg_desc(abi.SUM_WIDE, 0, buffer_precision=38, buffer_scale=2),
            abi.agg_desc(abi.SUM_WIDE, 1, buffer_precision=22, buffer_scale=2),
            abi.agg_desc(abi.AVG_WIDE, 1, buffer_precision=22, buffer_scale=2),
            abi.agg_desc(abi.AVG_WIDE, 1, flags=abi.RESULT, buffer_precision=22, buffer_scale=2,
                         result_precision=16, result_scale=6),
            abi.agg_desc(abi.SUM_WIDE, 1, flags=abi.RESULT, buffer_precision=22, buffer_scale=2)]
    spec = abi.AggSpec([abi.INT32], [abi.DECIMAL128, abi.INT64], aggs)
    agg = OracleAgg(spec)
    off = 0
    for n in z["batch"]:
        n = int(n)
        sl = slice(off, off + n)
        wide = Column(abi.DECIMAL128, np.ascontiguousarray(z["wide"][2 * off:2 * (off + n)]),
                      pack_bits(~z["wide_null"][sl]), length=n)
        agg.update([Column.ints(z["keys"][sl])], [wide, Column.longs(z["narrow"][sl], nulls=z["narrow_null"][sl])], n)
        off += n
    G = agg.num_groups()
    keys, outs = agg.emit()
    assert (keys[0].data[:G] == z["exp_key"]).all()
    ws, wempty, ns, nempty, asum, acount, ares, sres = outs
    got = ws.values(G)
    exp = [None if nul else v for v, nul in zip(_limbs_to_ints(z["exp_wide_sum"]), z["exp_wide_sum_null"])]
    assert [None if v is None else v for v in got] == exp
    assert wempty.values(G) == z["exp_wide_empty"].tolist()
    assert ns.values(G) == [None if nul else v for v, nul in
                            zip(_limbs_to_ints(z["exp_narrow_sum"]), z["exp_narrow_sum_null"])]
    assert asum.values(G) == _limbs_to_ints(z["exp_avg_sum"])
    assert acount.values(G) == z["exp_avg_count"].tolist()
    assert ares.values(G) == [None if nul else v for v, nul in zip(z["exp_avg_result"].tolist(),
                                                                   z["exp_avg_result_null"])]
    assert sres.values(G) == [None if nul else v for v, nul in
                              zip(_limbs_to_ints(z["exp_sum_result"]), z["exp_sum_result_null"])]
    # the extremes do overflow some group's 38-digit sum
    assert any(v is None for v in got)


def test_decimal_average_known_answers():
    """Spark's AVG(decimal(12,2)) -> decimal(16,6), HALF_UP (ties away from zero), closed form."""
    cases = [([1], 32, None), ([-1], 32, None), ([2], 3, None), ([200], 3, None),
             ([10 ** 8 - 1] * 7, 7, None), ([5, -5, 3], 3, None)]
    aggs = [abi.agg_desc(abi.AVG_WIDE, 0, flags=abi.RESULT, buffer_precision=22, buffer_scale=2,
                         result_precision=16, result_scale=6)]
    for vals, _, _ in cases:
        spec = abi.AggSpec([abi.INT32], [abi.INT64], aggs)
        agg = OracleAgg(spec)
        n = len(vals)
        agg.update([Column.ints([0] * n)], [Column.longs(vals)], n)
        _, (res,) = agg.emit()
        expect = pyref.round_half_up(Fraction(sum(vals), 100) / n, 6)
        assert res.values(1) == [expect]
    # one value spread over 32 rows: 0.01/32 = 0.0003125 -> 0.000313 (HALF_EVEN would give 0.000312)
    spec = abi.AggSpec([abi.INT32], [abi.INT64], aggs)
    agg = OracleAgg(spec)
    agg.update([Column.ints([0] * 32)], [Column.longs([1] + [0] * 31)], 32)
    assert agg.emit()[1][0].values(1) == [313]
    agg = OracleAgg(abi.AggSpec([abi.INT32], [abi.INT64], aggs))
    agg.update([Column.ints([0] * 32)], [Column.longs([-1] + [0] * 31)], 32)
    assert agg.emit()[1][0].values(1) == [-313]


def test_checked_long_sum_overflow_raises():
    """SUM(bigint) in ANSI mode: ARITHMETIC_OVERFLOW (VectorAggregates.scala:142-170); legacy wraps."""
    big = (1 << 63) - 1
    spec = abi.AggSpec([abi.INT32], [abi.INT64], [abi.agg_desc(abi.SUM_LONG, 0, flags=abi.CHECKED)])
    agg = OracleAgg(spec)
    with pytest.raises(OracleError) as e:
        agg.update([Column.ints([1, 1])], [Column.longs([big, 1])], 2)
    assert e.value.status == abi.VR_ERR_ARITHMETIC_OVERFLOW
    spec = abi.AggSpec([abi.INT32], [abi.INT64], [abi.agg_desc(abi.SUM_LONG, 0)])
    agg = OracleAgg(spec)
    agg.update([Column.ints([1, 1])], [Column.longs([big, 1])], 2)
    assert agg.emit()[1][0].values(1) == [-(1 << 63)]


def test_null_key_is_its_own_group_and_not_zero():
    """Null keys form one group distinct from 0 (GroupedAggregationTest.java:781-799)."""
    for t, mk in [(abi.INT64, Column.longs), (abi.INT32, Column.ints)]:
        spec = abi.AggSpec([t], [], [abi.agg_desc(abi.COUNT_STAR)])
        agg = OracleAgg(spec)
        ids = agg.update([mk([0, 0, 5, 0], nulls=[False, True, False, True])], [], 4)
        assert ids.tolist() == [0, 1, 2, 1]
        keys, (c,) = agg.emit()
        assert keys[0].valid(3).tolist() == [True, False, True]
        assert c.values(3) == [1, 2, 1]


@pytest.mark.parametrize("n", [0, 1, 3, 7, 8, 15, 16, 31, 63, 64, 65, 100, 127, 128, 129, 1000, 1023, 4097])
def test_awkward_lengths_count_invariants(n):
    """TestData.LENGTHS: COUNT(*) over all groups equals the selected rows, SUM equals the total."""
    rng = np.random.default_rng(n)
    keys = rng.integers(-20, 20, n)
    vals = rng.integers(-10 ** 6, 10 ** 6, n)
    sel = rng.random(n) < 0.7
    spec = abi.AggSpec([abi.INT64], [abi.INT64], [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.SUM_LONG, 0)])
    agg = OracleAgg(spec)
    agg.update([Column.longs(keys)], [Column.longs(vals)], n, pack_bits(sel) if n else None)
    G = agg.num_groups()
    assert G == len(set(keys[sel].tolist()))
    if G:
        _, (c, s) = agg.emit()
        assert int(c.data[:G].sum()) == int(sel.sum())
        assert int(s.data[:G].sum()) == int(vals[sel].sum())


def test_join_multisets_match_nested_loop():
    """BuildTable + probe pair order (streamed rows, build order), semi, anti and the
    null-aware anti against the nested-loop fixture."""
    z = load("join_dup.npz")
    nb, npr = len(z["bkeys"]), len(z["pkeys"])
    build = [Column.longs(z["bkeys"], nulls=z["bkeys_null"]), Column.ints(z["bpay"])]
    spec = abi.JoinSpec([abi.INT64], [abi.INT64, abi.INT32], [0], capacity_hint=nb)
    j = OracleJoin(spec, build, nb)
    probe = [Column.longs(z["pkeys"], nulls=z["pkeys_null"])]
    sel = pack_bits(z["psel"])
    p, b, outs, r = j.probe(probe, [0], npr, abi.INNER, sel, build_out=[1], capacity=npr * 20)
    assert p.tolist() == z["inner_probe"].tolist()
    assert b.tolist() == z["inner_build"].tolist()
    assert outs[0].data[:r].tolist() == z["bpay"][z["inner_build"]].tolist()
    p, _, _, _ = j.probe(probe, [0], npr, abi.LEFT_SEMI, sel)
    assert p.tolist() == z["semi"].tolist()
    p, _, _, _ = j.probe(probe, [0], npr, abi.LEFT_ANTI, sel)
    assert p.tolist() == z["anti"].tolist()
    p, _, _, _ = j.probe(probe, [0], npr, abi.LEFT_ANTI, sel, drop_null_keys=True)
    assert p.tolist() == z["anti_null_aware"].tolist()


def test_join_brute_force_tiny():
    """Every combination of 0-3 build rows x 0-3 probe rows over keys {None, 1, 2}."""
    import itertools
    vals = [None, 1, 2]
    for nb in range(4):
        for bk in itertools.product(vals, repeat=nb):
            for pk in itertools.product(vals, repeat=2):
                build = [Column.longs([0 if v is None else v for v in bk] or [0],
                                      nulls=[v is None for v in bk] or [True])]
                spec = abi.JoinSpec([abi.INT64], [abi.INT64], [0])
                j = OracleJoin(spec, build, nb)
                probe = [Column.longs([0 if v is None else v for v in pk], nulls=[v is None for v in pk])]
                p, b, _, _ = j.probe(probe, [0], 2, abi.INNER)
                assert list(zip(p.tolist(), b.tolist())) == pyref.join([list(bk)], [list(pk)], 0)
                for jt in (abi.LEFT_SEMI, abi.LEFT_ANTI):
                    p, _, _, _ = j.probe(probe, [0], 2, jt)
                    assert p.tolist() == pyref.join([list(bk)], [list(pk)], jt)


def test_oracle_matches_pyref_on_random_multikey():
    """Random INT64 + FLOAT64 (raw-bit) + BOOL keys with nulls, every update-mode aggregate."""
    rng = np.random.default_rng(5)
    n = 3000
    k1 = rng.integers(0, 6, n)
    k2 = rng.choice([0.0, -0.0, 1.5, float("nan")], n)
    k3 = rng.random(n) < 0.5
    nul = rng.random((3, n)) < 0.1
    v = rng.integers(-10 ** 12, 10 ** 12, n)
    d = rng.standard_normal(n)
    i32 = rng.integers(-2 ** 31, 2 ** 31, n)
    vn = rng.random(n) < 0.2
    keys = [Column.longs(k1, nulls=nul[0]), Column.doubles(k2, nulls=nul[1]), Column.bools(k3, nulls=nul[2])]
    values = [Column.longs(v, nulls=vn), Column.doubles(d, nulls=vn), Column.ints(i32, nulls=vn)]
    aggs = [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.COUNT, 0), abi.agg_desc(abi.SUM_LONG, 0),
            abi.agg_desc(abi.SUM_DOUBLE, 1), abi.agg_desc(abi.AVG_DOUBLE, 1), abi.agg_desc(abi.MIN, 2),
            abi.agg_desc(abi.MAX, 2), abi.agg_desc(abi.MIN, 0), abi.agg_desc(abi.MAX, 1),
            abi.agg_desc(abi.SUM_WIDE, 0, buffer_precision=22, buffer_scale=2)]
    spec = abi.AggSpec([abi.INT64, abi.FLOAT64, abi.BOOL], [abi.INT64, abi.FLOAT64, abi.INT32], aggs)
    agg = OracleAgg(spec)
    agg.update(keys, values, n)
    pk = [[None if nul[0][i] else int(k1[i]) for i in range(n)],
          [None if nul[1][i] else float(k2[i]) for i in range(n)],
          [None if nul[2][i] else bool(k3[i]) for i in range(n)]]
    pv = [[None if vn[i] else int(v[i]) for i in range(n)], [None if vn[i] else float(d[i]) for i in range(n)],
          [None if vn[i] else int(i32[i]) for i in range(n)]]
    paggs = [dict(kind=a.kind, input=a.input, flags=a.flags, buffer_precision=a.buffer_precision,
                  buffer_scale=a.buffer_scale, in_type=t)
             for a, t in zip(aggs, [None, "int64", "int64", "double", "double", "int32", "int32", "int64",
                                    "double", "decimal"])]
    order, outs = pyref.aggregate([dict(keys=pk, values=pv, selection=None)], [False, True, False], paggs)
    G = agg.num_groups()
    assert G == len(order)
    kout, aouts = agg.emit()
    got_keys = list(zip(kout[0].values(G), kout[1].values(G), kout[2].values(G)))
    for g in range(G):
        ek = order[g]
        gk = got_keys[g]
        assert gk[0] == ek[0] and gk[2] == ek[2]
        if ek[1] is None:
            assert gk[1] is None
        else:
            assert struct.pack("<d", gk[1]) == struct.pack("<d", ek[1])
    col = 0
    flat = []
    for o in aouts:
        flat.append(o.values(G))
    for g in range(G):
        k = 0
        for a_i, exp in enumerate(outs[g]):
            for val in exp:
                got = flat[k][g]
                if isinstance(val, float) and math.isnan(val):
                    assert math.isnan(got)
                else:
                    assert got == val, (g, a_i, got, val)
                k += 1


# ------------------------------------------------------------------ planted mistakes (CPU only)

def test_planted_mistake_reversed_fold_is_caught():
    """A fold in the wrong order (a plausible parallel-sum slip) differs from the strict
    fixture's bits: the bit-exact check has teeth."""
    z = load("strict_sum_c1.npz")
    ok = (~z["val_null"]) & z["sel"]
    vals = z["vals"][ok]
    fwd = 0.0
    for x in vals:
        fwd += x
    rev = 0.0
    for x in vals[::-1]:
        rev += x
    assert f64_bits([fwd])[0] == z["exp_sum_bits"][0]
    assert f64_bits([rev])[0] != z["exp_sum_bits"][0]


def test_planted_mistake_half_even_is_caught():
    """Rounding ties to even instead of away from zero fails the known answer 0.000313."""
    fr = Fraction(1, 100) / 32
    assert pyref.round_half_up(fr, 6) == 313
    scaled = fr * 10 ** 6
    assert round(scaled) == 312  # Python's round is half-even: the mistake the test would catch


def test_planted_mistake_dropped_carry_is_caught():
    """A 128-bit add that drops the carry out of the low limb changes the wide-sum fixture."""
    z = load("wide_sum.npz")
    keys = z["keys"]
    exact, nocarry = {}, {}
    limbs = z["wide"]
    for i in range(len(keys)):
        if z["wide_null"][i]:
            continue
        lo = int(limbs[2 * i]) & ((1 << 64) - 1)
        hi = int(limbs[2 * i + 1])
        k = int(keys[i])
        exact[k] = exact.get(k, 0) + ((hi << 64) | lo)
        l0, h0 = nocarry.get(k, (0, 0))
        nocarry[k] = ((l0 + lo) & ((1 << 64) - 1), h0 + hi)  # carry dropped
    assert any(((h << 64) | l) != exact[k] for k, (l, h) in nocarry.items())


UNGROUPED_AGGS = [abi.agg_desc(abi.COUNT_STAR), abi.agg_desc(abi.COUNT, 0), abi.agg_desc(abi.SUM_LONG, 0),
                  abi.agg_desc(abi.SUM_DOUBLE, 1), abi.agg_desc(abi.AVG_DOUBLE, 1), abi.agg_desc(abi.MIN, 2),
                  abi.agg_desc(abi.MAX, 2), abi.agg_desc(abi.SUM_WIDE, 0, buffer_precision=22, buffer_scale=2)]


def ungrouped_input(rng, n):
    v = rng.integers(-10 ** 9, 10 ** 9, n)
    d = rng.standard_normal(n)
    m = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
    nulls = [rng.random(n) < 0.1 for _ in range(3)]
    return [Column.longs(v, nulls=nulls[0]), Column.doubles(d, nulls=nulls[1]), Column.ints(m, nulls=nulls[2])], (v, d, m, nulls)


@pytest.mark.parametrize("n", [1, 7, 64, 1000, 4097])
def test_ungrouped_equals_one_constant_key_group(n):
    """AggKernels' ungrouped reductions (AggKernels.java:133-621) are the grouped accumulators of a
    single group: the oracle without keys must emit the one row the grouped oracle gives for a
    constant key, with selection, over several batches."""
    rng = np.random.default_rng(n + 5)
    types = [abi.INT64, abi.FLOAT64, abi.INT32]
    ug = OracleAgg(abi.AggSpec([], types, UNGROUPED_AGGS))
    gr = OracleAgg(abi.AggSpec([abi.INT64], types, UNGROUPED_AGGS, key_nullable=[False]))
    for part in range(3):
        cols, _ = ungrouped_input(rng, n)
        sel = pack_bits(rng.random(n) < 0.8)
        ids = ug.update([], cols, n, sel)
        assert set(ids.tolist()) <= {0, -1}
        gr.update([Column.longs(np.full(n, 42))], cols, n, sel)
    assert ug.num_groups() == 1 and gr.num_groups() == 1
    _, uo = ug.emit()
    _, go = gr.emit()
    for a, b in zip(uo, go):
        assert a.values(1) == b.values(1)


def test_ungrouped_closed_forms_and_empty_input():
    """COUNT(*) = the selected rows, COUNT(x) = the non-null ones, SUM = the Python integer total,
    MIN/MAX = numpy's over the non-null values; with no input the one row is COUNT 0 and nulls
    (Spark's ungrouped aggregate over an empty relation)."""
    rng = np.random.default_rng(77)
    n = 3000
    cols, (v, d, m, nulls) = ungrouped_input(rng, n)
    spec = abi.AggSpec([], [abi.INT64, abi.FLOAT64, abi.INT32], UNGROUPED_AGGS)
    ug = OracleAgg(spec)
    ug.update([], cols, n)
    _, out = ug.emit()
    # columns: COUNT(*), COUNT, SUM, SUM double, AVG (sum, count), MIN, MAX, wide SUM (sum, isEmpty)
    by = dict(zip(range(len(out)), out))
    assert by[0].values(1) == [n]
    assert by[1].values(1) == [int((~nulls[0]).sum())]
    assert by[2].values(1) == [sum(int(x) for x in v[~nulls[0]])]
    assert by[5].values(1) == [int((~nulls[1]).sum())]
    assert by[6].values(1) == [int(m[~nulls[2]].min())] and by[7].values(1) == [int(m[~nulls[2]].max())]
    assert math.isclose(by[3].values(1)[0], math.fsum(d[~nulls[1]]), rel_tol=1e-12)
    empty = OracleAgg(spec)
    assert empty.num_groups() == 1
    _, eo = empty.emit()
    assert eo[0].values(1) == [0] and eo[1].values(1) == [0]
    assert eo[2].values(1) == [None] and eo[3].values(1) == [None]
    assert eo[6].values(1) == [None] and eo[7].values(1) == [None]


def _decimal_pool(rng, n):
    """Unscaled 128-bit values: the extremes, values that differ only in a low limb whose top bit
    is set or clear, negative and positive values sharing a high limb, and random ones."""
    lo_big = (1 << 63) + 5
    special = [pyref.INT128_MIN, pyref.INT128_MAX, 0, -1, 1, (7 << 64) + lo_big, (7 << 64) + 3,
               -(7 << 64) + lo_big, -(7 << 64) + 3, (1 << 64) - 1, -(1 << 64), lo_big, 3]
    out = []
    for _ in range(n):
        r = rng.random()
        if r < 0.4:
            out.append(special[int(rng.integers(0, len(special)))])
        elif r < 0.7:
            out.append(int(rng.integers(-10 ** 18, 10 ** 18)) * int(rng.integers(1, 10 ** 18)))
        else:
            out.append(int(rng.integers(-2 ** 62, 2 ** 62)) << int(rng.integers(0, 64)))
    return out


def test_decimal128_min_max_match_pyref():
    """Decimal128MinMax (GroupedAccumulators.java:1357-1462): the first value, then one strictly
    better by signed 128-bit order (Decimal128.compare, Decimal128.java:145); a group with no
    value emits null.  Against the Python restatement's exact integers over three batches with nulls
    and a selection."""
    rng = np.random.default_rng(128)
    batches, pb = [], []
    for n in (700, 1, 1300):
        keys = rng.integers(0, 40, n)
        vals = _decimal_pool(rng, n)
        for i in np.flatnonzero(keys == 38):  # one high limb: the low limb's order decides
            vals[i] = (7 << 64) + ((1 << 63) + 5 if rng.random() < 0.5 else 3)
        nul = (rng.random(n) < 0.15) | (keys == 39)
        sel = rng.random(n) < 0.8
        batches.append((keys, vals, nul, sel, n))
        pb.append(dict(keys=[[int(k) for k in keys]], values=[[None if nul[i] else vals[i] for i in range(n)]],
                       selection=[bool(x) for x in sel]))
    aggs = [abi.agg_desc(abi.MIN, 0), abi.agg_desc(abi.MAX, 0), abi.agg_desc(abi.COUNT, 0)]
    spec = abi.AggSpec([abi.INT32], [abi.DECIMAL128], aggs)
    agg = OracleAgg(spec)
    for keys, vals, nul, sel, n in batches:
        agg.update([Column.ints(keys)], [Column.decimals(vals, nulls=nul)], n, selection=pack_bits(sel))
    paggs = [dict(kind=a.kind, input=a.input, flags=a.flags, in_type="decimal") for a in aggs]
    order, outs = pyref.aggregate(pb, [False], paggs)
    G = agg.num_groups()
    assert G == len(order)
    kout, (mn, mx, cnt) = agg.emit()
    assert kout[0].values(G) == [o[0] for o in order]
    assert mn.values(G) == [o[0][0] for o in outs]
    assert mx.values(G) == [o[1][0] for o in outs]
    assert cnt.values(G) == [o[2][0] for o in outs]
    assert any(v is None for v in mn.values(G))  # key 39 holds only nulls
    # a compare that takes the low limb as signed (a plausible slip) picks differently on this data
    def signed_lo_key(v):
        lo = v & ((1 << 64) - 1)
        return (v >> 64, lo - (1 << 64) if lo >= (1 << 63) else lo)
    wrong = {}
    for b in pb:
        for i, (k, v) in enumerate(zip(b["keys"][0], b["values"][0])):
            if v is None or not b["selection"][i]:
                continue
            if k not in wrong or signed_lo_key(v) < signed_lo_key(wrong[k]):
                wrong[k] = v
    assert any(wrong.get(o[0]) != e[0][0] for o, e in zip(order, outs))


@pytest.mark.parametrize("kinds", [k for k in __import__("join_key_cases").KIND_SETS])
def test_join_wide_and_string_keys_match_nested_loop(kinds):
    """Key tuples over more than 64 bits (INT64 pairs, DECIMAL128, four mixed lanes), FLOAT64 keys
    compared bit for bit (-0.0 and 0.0 differ, NaN equals its own bits) and UTF8 keys (plain build,
    dictionary-encoded probe): the oracle's pairs, semi and anti rows equal the nested loop's."""
    from join_key_cases import make_case
    case = make_case(sum(map(ord, "".join(kinds))), kinds, 300, 500, 60, dict_probe=("utf8" in kinds))
    j = OracleJoin(case["spec"], case["build"], case["nb"])
    keys = list(range(case["nk"]))
    p, b, _, _ = j.probe(case["probe"], keys, case["npr"], abi.INNER, capacity=case["npr"] * case["nb"])
    want = pyref.join(case["pybuild"], case["pyprobe"], 0)
    assert len(want) > 100
    assert list(zip(p.tolist(), b.tolist())) == want
    for jt in (abi.LEFT_SEMI, abi.LEFT_ANTI):
        p, _, _, _ = j.probe(case["probe"], keys, case["npr"], jt)
        assert p.tolist() == pyref.join(case["pybuild"], case["pyprobe"], jt)
    p, _, _, _ = j.probe(case["probe"], keys, case["npr"], abi.LEFT_ANTI, drop_null_keys=True)
    assert p.tolist() == pyref.join(case["pybuild"], case["pyprobe"], 2, drop_null_keys=True)
