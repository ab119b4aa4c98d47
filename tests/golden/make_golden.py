"""Writes the golden fixtures under tests/golden/ (run: python tests/golden/make_golden.py).

Inputs are drawn with java.util.Random (restated in java_random.py) from the seeds and in the draw
order of the reference's own kernel tests, so each fixture holds the inputs those tests feed the
reference, and the expected outputs computed from them by oracle/pyref.py -- the reference's
in-test oracle style (a map per key tuple, exact longs, sequential double folds).

  grouped_agg_c{C}_s{S}.npz   GroupedAggregationTest.groupedAccumulatorsMatchReference /
                              groupedAccumulatorsHonourSelection (GroupedAggregationTest.java:59-193),
                              C = cardinality, S = with selection
  strict_sum_c{C}.npz         GroupedAggregationTest.strictDoubleSumIsBitIdenticalToSequentialReference
                              (GroupedAggregationTest.java:201-244)
  minmax_special.npz          MIN/MAX over TestData.doubles' special values (TestData.java:120-133)
  wide_sum.npz                SUM/AVG over TestData.decimal128s' extremes (TestData.java:97-118)
  join_dup.npz                join multisets with duplicate and null keys (nested-loop reference)
"""
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from java_random import JavaRandom  # noqa: E402
from oracle import pyref  # noqa: E402

GROUPED_CARDINALITIES = [1, 4, 64, 65, 300, 5000]
GROUPED_LENGTHS = [1000, 37, 2048, 64, 999]
STRICT_CARDINALITIES = [1, 3, 64, 300, 5000]
STRICT_LENGTHS = [1000, 37, 2048, 64, 999, 4097]


def dbits(x):
    return struct.unpack("<q", struct.pack("<d", x))[0]


def random_bitmap(rnd, n):
    """TestData.randomBitmap (TestData.java:135-143)."""
    return [rnd.next_boolean() for _ in range(n)]


def grouped_case(cardinality, with_selection):
    """The draws of GroupedAggregationTest.run (GroupedAggregationTest.java:75-110)."""
    rnd = JavaRandom(cardinality)
    batches = []
    for n in GROUPED_LENGTHS:
        k1, k2, vals, lvals, keynull, valnull = [], [], [], [], [], []
        for _ in range(n):
            g = rnd.next_int(cardinality)
            a = g % 7
            k1.append(a)
            k2.append(None if g % 13 == 0 else "k%d" % (g // 7))
            keynull.append(a == 3 and rnd.next_int(5) == 0)
            vals.append(float("nan") if rnd.next_int(6) == 0 else rnd.next_int_range(-100, 100) / 4.0)
            lvals.append(rnd.next_int_range(-1000, 1000))
            valnull.append(rnd.next_int(4) == 0)
        sel = random_bitmap(rnd, n) if with_selection else None
        batches.append(dict(k1=k1, k2=k2, vals=vals, lvals=lvals, keynull=keynull, valnull=valnull, sel=sel))
    return batches


GROUPED_AGGS = [
    dict(kind=pyref.SUM_DOUBLE, input=0, flags=0, in_type="double"),
    dict(kind=pyref.SUM_LONG, input=1, flags=0, in_type="int64"),
    dict(kind=pyref.COUNT_STAR, input=-1, flags=0, in_type=None),
    dict(kind=pyref.COUNT, input=0, flags=0, in_type="double"),
    dict(kind=pyref.MIN, input=0, flags=0, in_type="double"),
    dict(kind=pyref.MAX, input=0, flags=0, in_type="double"),
]


def write_grouped(cardinality, with_selection):
    batches = grouped_case(cardinality, with_selection)
    pb = []
    for b in batches:
        k1 = [None if kn else v for v, kn in zip(b["k1"], b["keynull"])]
        vals = [None if vn else v for v, vn in zip(b["vals"], b["valnull"])]
        lvals = [None if vn else v for v, vn in zip(b["lvals"], b["valnull"])]
        pb.append(dict(keys=[k1, b["k2"]], values=[vals, lvals], selection=b["sel"]))
    order, outs = pyref.aggregate(pb, [False, False], GROUPED_AGGS)
    strings = sorted({s for b in batches for s in b["k2"] if s is not None})
    code = {s: i for i, s in enumerate(strings)}
    dict_bytes = "".join(strings).encode()
    dict_offsets = np.cumsum([0] + [len(s) for s in strings]).astype(np.int32)
    cat = lambda f: np.concatenate([np.asarray(b[f]) for b in batches])  # noqa: E731
    arrays = dict(
        lengths=np.asarray(GROUPED_LENGTHS, np.int64),
        k1=cat("k1").astype(np.int32),
        k1_null=cat("keynull").astype(bool),
        k2_code=np.asarray([code[s] if s is not None else -1 for b in batches for s in b["k2"]], np.int32),
        dict_bytes=np.frombuffer(dict_bytes, np.uint8),
        dict_offsets=dict_offsets,
        vals=cat("vals").astype(np.float64),
        lvals=cat("lvals").astype(np.int64),
        val_null=cat("valnull").astype(bool),
        sel=(np.concatenate([np.asarray(b["sel"], bool) for b in batches]) if with_selection
             else np.ones(sum(GROUPED_LENGTHS), bool)),
        exp_k1=np.asarray([k[0] if k[0] is not None else 0 for k in order], np.int32),
        exp_k1_null=np.asarray([k[0] is None for k in order], bool),
        exp_k2=np.asarray([code[k[1]] if k[1] is not None else -1 for k in order], np.int32),
        exp_dsum_bits=np.asarray([dbits(o[0][0]) if o[0][0] is not None else 0 for o in outs], np.int64),
        exp_dsum_null=np.asarray([o[0][0] is None for o in outs], bool),
        exp_lsum=np.asarray([o[1][0] if o[1][0] is not None else 0 for o in outs], np.int64),
        exp_lsum_null=np.asarray([o[1][0] is None for o in outs], bool),
        exp_rows=np.asarray([o[2][0] for o in outs], np.int64),
        exp_count=np.asarray([o[3][0] for o in outs], np.int64),
        exp_min_bits=np.asarray([dbits(o[4][0]) if o[4][0] is not None else 0 for o in outs], np.int64),
        exp_max_bits=np.asarray([dbits(o[5][0]) if o[5][0] is not None else 0 for o in outs], np.int64),
        exp_has=np.asarray([o[4][0] is not None for o in outs], bool),
    )
    np.savez_compressed(os.path.join(HERE, "grouped_agg_c%d_s%d.npz" % (cardinality, int(with_selection))), **arrays)


def write_strict(cardinality):
    """The draws of strictDoubleSumIsBitIdenticalToSequentialReference (GroupedAggregationTest.java:201-230)."""
    rnd = JavaRandom(7 * cardinality + 1)
    ids, vals, nulls, sels, lengths = [], [], [], [], []
    for n in STRICT_LENGTHS:
        bi, bv, bn = [], [], []
        for _ in range(n):
            bi.append(rnd.next_int(cardinality))
            sign = 1 if rnd.next_boolean() else -1
            d = rnd.next_double()
            bv.append(sign * d * (10.0 ** rnd.next_int(16)))
            bn.append(rnd.next_int(5) == 0)
        bs = random_bitmap(rnd, n) if n % 2 == 0 else [True] * n
        ids += bi
        vals += bv
        nulls += bn
        sels += bs
        lengths.append(n)
    # Expected: the per-group sequential fold over selected, non-null rows, keyed by id.
    batches, off = [], 0
    for n in lengths:
        sl = slice(off, off + n)
        batches.append(dict(keys=[ids[sl]], values=[[None if nl else v for v, nl in zip(vals[sl], nulls[sl])]],
                            selection=sels[sl]))
        off += n
    aggs = [dict(kind=pyref.SUM_DOUBLE, input=0, flags=0, in_type="double"),
            dict(kind=pyref.COUNT, input=0, flags=0, in_type="double")]
    order, outs = pyref.aggregate(batches, [False], aggs)
    np.savez_compressed(
        os.path.join(HERE, "strict_sum_c%d.npz" % cardinality),
        lengths=np.asarray(lengths, np.int64), ids=np.asarray(ids, np.int32), vals=np.asarray(vals, np.float64),
        val_null=np.asarray(nulls, bool), sel=np.asarray(sels, bool),
        exp_key=np.asarray([k[0] for k in order], np.int32),
        exp_sum_bits=np.asarray([dbits(o[0][0]) if o[0][0] is not None else 0 for o in outs], np.int64),
        exp_sum_null=np.asarray([o[0][0] is None for o in outs], bool),
        exp_count=np.asarray([o[1][0] for o in outs], np.int64))


def test_doubles(rnd, n):
    """TestData.doubles (TestData.java:120-133): NaN, +-Inf, signed zeros and small values."""
    out = []
    for _ in range(n):
        c = rnd.next_int(12)
        if c == 0:
            out.append(float("nan"))
        elif c == 1:
            out.append(float("inf"))
        elif c == 2:
            out.append(float("-inf"))
        elif c == 3:
            out.append(-0.0)
        elif c == 4:
            out.append(0.0)
        elif c == 5:
            out.append(1.5)
        else:
            out.append(rnd.next_int_range(-5, 5) + (0.0 if rnd.next_boolean() else 0.5))
    return out


def write_minmax_special():
    rnd = JavaRandom(20260)
    lengths = [0, 1, 3, 7, 8, 15, 16, 31, 63, 64, 65, 100, 127, 128, 129, 1000, 1023, 4097]  # TestData.LENGTHS
    keys, vals, nulls = [], [], []
    for n in lengths:
        keys += [rnd.next_int(40) for _ in range(n)]
        vals += test_doubles(rnd, n)
        nulls += [rnd.next_double() < 0.1 for _ in range(n)]
    batches, off = [], 0
    for n in lengths:
        sl = slice(off, off + n)
        batches.append(dict(keys=[keys[sl]], values=[[None if nl else v for v, nl in zip(vals[sl], nulls[sl])]],
                            selection=None))
        off += n
    aggs = [dict(kind=pyref.MIN, input=0, flags=0, in_type="double"),
            dict(kind=pyref.MAX, input=0, flags=0, in_type="double"),
            dict(kind=pyref.SUM_DOUBLE, input=0, flags=0, in_type="double")]
    order, outs = pyref.aggregate(batches, [False], aggs)
    np.savez_compressed(
        os.path.join(HERE, "minmax_special.npz"),
        lengths=np.asarray(lengths, np.int64), keys=np.asarray(keys, np.int32), vals=np.asarray(vals, np.float64),
        val_null=np.asarray(nulls, bool), exp_key=np.asarray([k[0] for k in order], np.int32),
        exp_min_bits=np.asarray([dbits(o[0][0]) if o[0][0] is not None else 0 for o in outs], np.int64),
        exp_max_bits=np.asarray([dbits(o[1][0]) if o[1][0] is not None else 0 for o in outs], np.int64),
        exp_has=np.asarray([o[0][0] is not None for o in outs], bool),
        exp_sum_bits=np.asarray([dbits(o[2][0]) if o[2][0] is not None else 0 for o in outs], np.int64))


MAX_DECIMAL38 = 10 ** 38 - 1


def random_decimal128(rnd):
    """TestData.randomDecimal128 (TestData.java:103-118), the BigInteger(bits, rnd) draw restated as
    uniform bits from nextInt (the fixture keeps the values it drew)."""
    c = rnd.next_int(12)
    if c == 0:
        return MAX_DECIMAL38
    if c == 1:
        return -MAX_DECIMAL38
    if c == 2:
        return -(1 << 127)
    if c == 3:
        return -(1 << 63)
    if c == 4:
        return -(1 << 63) << 64
    if c == 5:
        return (1 << 63)
    if c == 6:
        return 0
    if c == 7:
        return rnd.next_int(2000) - 1000
    bits = rnd.next_int(127)
    v = 0
    for k in range(0, bits, 32):
        v |= (rnd.next_int() & 0xFFFFFFFF) << k
    v &= (1 << bits) - 1 if bits else 0
    return v if rnd.next_boolean() else -v


def to_limbs(vals):
    arr = np.zeros(2 * len(vals), np.int64)
    for i, x in enumerate(vals):
        x &= (1 << 128) - 1
        lo, hi = x & ((1 << 64) - 1), x >> 64
        arr[2 * i] = lo - (1 << 64) if lo >= (1 << 63) else lo
        arr[2 * i + 1] = hi - (1 << 64) if hi >= (1 << 63) else hi
    return arr


def write_wide_sum():
    """SUM / AVG over DECIMAL128 lanes (extremes included) and over INT64 unscaled lanes."""
    rnd = JavaRandom(12838)
    n = 3000
    keys = [rnd.next_int(25) for _ in range(n)]
    wide = [random_decimal128(rnd) for _ in range(n)]
    wide_null = [rnd.next_int(10) == 0 for _ in range(n)]
    narrow = [rnd.next_int_range(0, 100000000) for _ in range(n)]  # decimal(12,2) unscaled in [0, 1e8)
    narrow_null = [rnd.next_int(20) == 0 for _ in range(n)]
    batches = []
    for s in range(0, n, 700):
        sl = slice(s, min(n, s + 700))
        batches.append(dict(keys=[keys[sl]],
                            values=[[None if z else v for v, z in zip(wide[sl], wide_null[sl])],
                                    [None if z else v for v, z in zip(narrow[sl], narrow_null[sl])]],
                            selection=None))
    aggs = [dict(kind=pyref.SUM_WIDE, input=0, flags=0, buffer_precision=38, buffer_scale=2, in_type="decimal"),
            dict(kind=pyref.SUM_WIDE, input=1, flags=0, buffer_precision=22, buffer_scale=2, in_type="decimal"),
            dict(kind=pyref.AVG_WIDE, input=1, flags=0, buffer_precision=22, buffer_scale=2, in_type="decimal"),
            dict(kind=pyref.AVG_WIDE, input=1, flags=pyref.RESULT, buffer_precision=22, buffer_scale=2,
                 result_precision=16, result_scale=6, in_type="decimal"),
            dict(kind=pyref.SUM_WIDE, input=1, flags=pyref.RESULT, buffer_precision=22, buffer_scale=2,
                 in_type="decimal")]
    order, outs = pyref.aggregate(batches, [False], aggs)

    def limbs_or_zero(vs):
        return to_limbs([0 if v is None else v for v in vs])

    np.savez_compressed(
        os.path.join(HERE, "wide_sum.npz"),
        batch=np.asarray([700] * (n // 700) + ([n % 700] if n % 700 else []), np.int64),
        keys=np.asarray(keys, np.int32), wide=to_limbs(wide), wide_null=np.asarray(wide_null, bool),
        narrow=np.asarray(narrow, np.int64), narrow_null=np.asarray(narrow_null, bool),
        exp_key=np.asarray([k[0] for k in order], np.int32),
        exp_wide_sum=limbs_or_zero([o[0][0] for o in outs]),
        exp_wide_sum_null=np.asarray([o[0][0] is None for o in outs], bool),
        exp_wide_empty=np.asarray([o[0][1] for o in outs], bool),
        exp_narrow_sum=limbs_or_zero([o[1][0] for o in outs]),
        exp_narrow_sum_null=np.asarray([o[1][0] is None for o in outs], bool),
        exp_avg_sum=limbs_or_zero([o[2][0] for o in outs]),
        exp_avg_count=np.asarray([o[2][1] for o in outs], np.int64),
        exp_avg_result=np.asarray([0 if o[3][0] is None else o[3][0] for o in outs], np.int64),
        exp_avg_result_null=np.asarray([o[3][0] is None for o in outs], bool),
        exp_sum_result=limbs_or_zero([o[4][0] for o in outs]),
        exp_sum_result_null=np.asarray([o[4][0] is None for o in outs], bool))


def write_join_dup():
    """Build keys with duplicates and nulls, probe keys hitting, missing and null; pairs by
    nested loop (the order the reference emits: streamed rows, then build order)."""
    rnd = JavaRandom(97531)
    nb, npr = 600, 2500
    bkeys = [None if rnd.next_int(30) == 0 else rnd.next_int_range(-50, 150) for _ in range(nb)]
    bpay = [rnd.next_int() for _ in range(nb)]
    pkeys = [None if rnd.next_int(25) == 0 else rnd.next_int_range(-80, 200) for _ in range(npr)]
    psel = [rnd.next_int(7) != 0 for _ in range(npr)]
    inner = pyref.join([bkeys], [pkeys], 0, psel)
    semi = pyref.join([bkeys], [pkeys], 1, psel)
    anti = pyref.join([bkeys], [pkeys], 2, psel)
    anti_null_aware = pyref.join([bkeys], [pkeys], 2, psel, drop_null_keys=True)
    np.savez_compressed(
        os.path.join(HERE, "join_dup.npz"),
        bkeys=np.asarray([0 if k is None else k for k in bkeys], np.int64),
        bkeys_null=np.asarray([k is None for k in bkeys], bool), bpay=np.asarray(bpay, np.int32),
        pkeys=np.asarray([0 if k is None else k for k in pkeys], np.int64),
        pkeys_null=np.asarray([k is None for k in pkeys], bool), psel=np.asarray(psel, bool),
        inner_probe=np.asarray([p for p, _ in inner], np.int32), inner_build=np.asarray([b for _, b in inner], np.int32),
        semi=np.asarray(semi, np.int32), anti=np.asarray(anti, np.int32),
        anti_null_aware=np.asarray(anti_null_aware, np.int32))


def main():
    for c in GROUPED_CARDINALITIES:
        for s in (False, True):
            write_grouped(c, s)
    for c in STRICT_CARDINALITIES:
        write_strict(c)
    write_minmax_special()
    write_wide_sum()
    write_join_dup()
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
