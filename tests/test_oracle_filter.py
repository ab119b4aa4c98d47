"""The oracle's selection producer and compaction pinned by brute force and closed forms.

* vro_filter_compare (CompareKernels.compareScalar + VectorExpr.selection): against Python's own
  integer, float and big-integer comparisons, with the reference's NaN-literal truth table
  (CompareKernels.java:689-721: EQ/GE keep the NaN rows, NE/LT the others, LE every row, GT none),
  -0.0 == 0.0, the INT32 literal narrowed as Java's intValue, and decimals across the low limb's
  sign boundary (Decimal128.compare: signed high limb, unsigned low limb).
* vro_compact (CompactKernels.compactFixed / compactUtf8): the selected rows in order, values and
  validity, for every lane including BOOL bits, dictionary codes and plain UTF8 bytes."""
import math

import numpy as np
import pytest

from oracle import oracle
from vecruntime_amd import abi
from vecruntime_amd.columns import Column, pack_bits, unpack_bits

OPS = [abi.CMP_EQ, abi.CMP_NE, abi.CMP_LT, abi.CMP_LE, abi.CMP_GT, abi.CMP_GE]


def holds(op, c):
    return [c == 0, c != 0, c < 0, c <= 0, c > 0, c >= 0][op]


def cmp_py(x, s, t):
    if t == abi.FLOAT64:
        if (math.isnan(x) and math.isnan(s)) or x == s:
            return 0
        if math.isnan(x):
            return 1
        if math.isnan(s):
            return -1
        return 1 if x > s else -1
    if t == abi.INT32:
        s = ((int(s) + 2 ** 31) % 2 ** 32) - 2 ** 31
    return (x > s) - (x < s)


def brute(vals, nulls, sel, op, s, t):
    return [bool(sel[i]) and not nulls[i] and holds(op, cmp_py(vals[i], s, t)) for i in range(len(vals))]


def column(t, vals, nulls):
    if t == abi.INT32:
        return Column.ints(vals, nulls=nulls)
    if t == abi.INT64:
        return Column.longs(vals, nulls=nulls)
    if t == abi.FLOAT64:
        return Column.doubles(vals, nulls=nulls)
    return Column.decimals(vals, nulls=nulls)


def values(rng, t, n):
    if t == abi.INT32:
        v = [int(x) for x in rng.integers(-5, 5, n)]
        v[:4] = [-2 ** 31, 2 ** 31 - 1, 0, -1][:n]
        return v
    if t == abi.INT64:
        v = [int(x) for x in rng.integers(-5, 5, n)]
        v[:4] = [-2 ** 63, 2 ** 63 - 1, 0, -1][:n]
        return v
    if t == abi.FLOAT64:
        pool = [0.0, -0.0, 1.5, -2.5, float("nan"), float("inf"), -float("inf"), 3.0]
        return [pool[int(k)] for k in rng.integers(0, len(pool), n)]
    pool = [0, -1, 1, 2 ** 63, -2 ** 63, 2 ** 64, -2 ** 64, 2 ** 64 - 1, 5 * 10 ** 30, -5 * 10 ** 30]
    return [pool[int(k)] for k in rng.integers(0, len(pool), n)]


LITERALS = {
    abi.INT32: [0, -1, 3, 2 ** 31 - 1, -2 ** 31, 2 ** 32 + 3],  # the last narrows to 3
    abi.INT64: [0, -1, 3, 2 ** 63 - 1, -2 ** 63],
    abi.FLOAT64: [0.0, -0.0, 1.5, float("nan"), float("inf"), -float("inf")],
    abi.DECIMAL128: [0, -1, 2 ** 63, 2 ** 64, -2 ** 64, 2 ** 64 - 1, 5 * 10 ** 30],
}


@pytest.mark.parametrize("t", [abi.INT32, abi.INT64, abi.FLOAT64, abi.DECIMAL128])
def test_compare_matches_brute_force(t):
    rng = np.random.default_rng(t)
    for n in (1, 63, 64, 65, 500):
        vals = values(rng, t, n)
        nulls = rng.random(n) < 0.2
        sel = rng.random(n) < 0.8
        col = column(t, vals, nulls)
        for s in LITERALS[t]:
            for op in OPS:
                for selection in (None, pack_bits(sel)):
                    bits, cnt = oracle.filter_compare(col, op, s, n, selection)
                    want = brute(vals, nulls, sel if selection is not None else [True] * n, op, s, t)
                    got = unpack_bits(bits, n).tolist()
                    assert got == want, (t, n, s, op)
                    assert cnt == sum(want)
                    assert not unpack_bits(bits, len(bits) * 8)[n:].any()  # padding stays clear


def test_nan_literal_truth_table_and_signed_zero():
    x = [float("nan"), 1.0, -0.0, 0.0, float("inf"), float("nan")]
    col = Column.doubles(x)
    isnan = [math.isnan(v) for v in x]
    table = {abi.CMP_EQ: isnan, abi.CMP_GE: isnan, abi.CMP_NE: [not b for b in isnan],
             abi.CMP_LT: [not b for b in isnan], abi.CMP_LE: [True] * 6, abi.CMP_GT: [False] * 6}
    for op, want in table.items():
        bits, _ = oracle.filter_compare(col, op, float("nan"), 6)
        assert unpack_bits(bits, 6).tolist() == want
    bits, cnt = oracle.filter_compare(col, abi.CMP_EQ, -0.0, 6)
    assert unpack_bits(bits, 6).tolist() == [False, False, True, True, False, False] and cnt == 2


def utf8_column(strs, nulls, dict_codes=None):
    if dict_codes is not None:
        return Column.dictionary_strings(np.asarray(dict_codes, np.int32), Column.strings(strs), nulls=nulls)
    enc = [s.encode() for s in strs]
    off = np.zeros(len(enc) + 1, np.int32)
    off[1:] = np.cumsum([len(e) for e in enc])
    data = np.frombuffer(b"".join(enc) or b"\0", np.uint8).copy()
    return Column(abi.UTF8, data, pack_bits(~np.asarray(nulls, bool)), offsets=off)


def test_compact_every_lane():
    rng = np.random.default_rng(4)
    for n in (1, 64, 65, 300):
        sel = rng.random(n) < 0.4
        nulls = [rng.random(n) < 0.2 for _ in range(7)]
        i32 = rng.integers(-100, 100, n)
        i64 = rng.integers(-10 ** 12, 10 ** 12, n)
        f64 = rng.standard_normal(n)
        dec = [int(x) * 10 ** 20 + 7 for x in rng.integers(-10 ** 6, 10 ** 6, n)]
        bools = rng.random(n) < 0.5
        strs = ["".join(chr(97 + int(c)) for c in rng.integers(0, 26, int(k))) for k in rng.integers(0, 9, n)]
        dictionary = ["x", "yy", "zzz"]
        codes = rng.integers(0, 3, n)
        cols = [Column.ints(i32, nulls=nulls[0]), Column.longs(i64, nulls=nulls[1]), Column.doubles(f64, nulls=nulls[2]),
                Column.decimals(dec, nulls=nulls[3]), Column.bools(bools, nulls=nulls[4]), utf8_column(strs, nulls[5]),
                utf8_column(dictionary, nulls[6], codes)]
        types = [abi.INT32, abi.INT64, abi.FLOAT64, abi.DECIMAL128, abi.BOOL, abi.UTF8, abi.INT32]
        nbytes = sum(len(strs[i]) for i in range(n) if sel[i])
        outs, m = oracle.compact(cols, n, pack_bits(sel), types, n, utf8_bytes=nbytes)
        idx = np.nonzero(sel)[0]
        assert m == len(idx)
        srcs = [i32, i64, f64, dec, bools, strs, codes]
        for c, (o, src) in enumerate(zip(outs, srcs)):
            assert o.valid(m).tolist() == [not nulls[c][i] for i in idx], c
            if types[c] == abi.UTF8:
                got = [bytes(o.data[o.offsets[r]:o.offsets[r + 1]]).decode() for r in range(m)]
                assert got == [strs[i] for i in idx]
            elif types[c] == abi.BOOL:
                assert unpack_bits(o.data, m).tolist() == [bool(bools[i]) for i in idx]
            elif types[c] == abi.DECIMAL128:
                assert [v for v in o.values(m) if v is not None] == [dec[i] for i in idx if not nulls[c][i]]
            else:
                assert o.data[:m].tolist() == [src[i] for i in idx]


def test_gather_every_lane():
    """vro_gather against the rules of GatherKernels (GatherKernels.java:99-308): row idx[k], -1 a
    null whose fixed value and BOOL bit are zero, validity unless null or padded, UTF8 bytes of the
    valid real rows only."""
    rng = np.random.default_rng(12)
    n, m = 200, 333
    nulls = rng.random(n) < 0.2
    idx = rng.integers(-1, n, m).astype(np.int32)
    idx[idx < 0] = -1
    strs = ["".join(chr(97 + int(c)) for c in rng.integers(0, 26, int(k))) for k in rng.integers(0, 7, n)]
    i32, i64, f64 = rng.integers(-9, 9, n), rng.integers(-10 ** 15, 10 ** 15, n), rng.standard_normal(n)
    dec = [int(x) * 10 ** 25 - 3 for x in rng.integers(-10 ** 6, 10 ** 6, n)]
    bools = rng.random(n) < 0.5
    cases = [(Column.ints(i32, nulls=nulls), i32), (Column.longs(i64, nulls=nulls), i64),
             (Column.doubles(f64, nulls=nulls), f64), (Column.decimals(dec, nulls=nulls), dec),
             (Column.bools(bools, nulls=nulls), bools), (utf8_column(strs, nulls), strs)]
    for col, src in cases:
        nbytes = sum(len(strs[i]) for i in idx if i >= 0 and not nulls[i]) if col.type == abi.UTF8 else 0
        o = oracle.gather(col, idx, utf8_bytes=nbytes)
        assert o.valid(m).tolist() == [bool(i >= 0 and not nulls[i]) for i in idx]
        if col.type == abi.UTF8:
            got = [bytes(o.data[o.offsets[k]:o.offsets[k + 1]]).decode() for k in range(m)]
            assert got == [strs[i] if i >= 0 and not nulls[i] else "" for i in idx]
        elif col.type == abi.BOOL:
            assert unpack_bits(o.data, m).tolist() == [bool(i >= 0 and bools[i]) for i in idx]
        elif col.type == abi.DECIMAL128:
            raw = o.data[: 2 * m].reshape(m, 2)
            want = [(dec[i] if i >= 0 else 0) for i in idx]
            assert [(int(lo) & (2 ** 64 - 1)) | (int(hi) << 64) for lo, hi in raw] == want
        else:
            assert o.data[:m].tolist() == [src[i] if i >= 0 else 0 for i in idx]
