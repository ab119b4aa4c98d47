"""The oracle's shuffle partitioning pinned independently of its own code.

* A Python MurmurHash3_x86_32 written from the published algorithm is checked against the
  published test vectors (the standard tail) and against Spark's documented
  `hash('Spark', array(123), 2) = -1321691492` (the SQL function reference for `hash`), which goes
  through Spark's non-standard tail (each remaining byte sign-extended and mixed as a word,
  PartitionKernels.java:131-150) -- a standard tail gives another value.
* vro_hash_partition_ids against that Python hash for every key kind of PartitionKernels.KeyKind
  (PartitionKernels.java:55-86, mixColumn :199-250): ints as their 4 little-endian bytes, longs as
  8, doubles normalized (both zeros 0, one NaN), floats narrowed from the FLOAT64 lane, booleans
  as 1 / 0, strings with Spark's tail (bytes >= 0x80 included), wide decimals as the minimal
  big-endian two's-complement bytes of BigInteger.toByteArray, nulls leaving the hash, several
  keys mixed in order, another seed, and Pmod as Python's non-negative %.
* vro_partition_rows against numpy's stable argsort, bincount and the inverse permutation;
  vro_round_robin_ids against PartitionKernelsSuite.scala:230-235's own expectation."""
import math
import struct

import numpy as np
import pytest

from oracle import oracle
from test_oracle_filter import utf8_column
from vecruntime_amd import abi
from vecruntime_amd.columns import Column

M32 = 0xFFFFFFFF


def _rotl(x, r):
    return ((x << r) | (x >> (32 - r))) & M32


def _k1(k):
    return (_rotl((k * 0xCC9E2D51) & M32, 15) * 0x1B873593) & M32


def _h1(h, k):
    return (_rotl(h ^ k, 13) * 5 + 0xE6546B64) & M32


def _fmix(h, n):
    h ^= n
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M32
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & M32
    return h ^ (h >> 16)


def murmur3(data, seed, spark_tail=False):
    """MurmurHash3_x86_32 over bytes; spark_tail mixes each tail byte sign-extended as its own word."""
    h, n = seed & M32, len(data)
    a = n - n % 4
    for i in range(0, a, 4):
        h = _h1(h, _k1(int.from_bytes(data[i:i + 4], "little")))
    tail = data[a:]
    if spark_tail:
        for b in tail:
            h = _h1(h, _k1((b - 256 if b > 127 else b) & M32))
    elif tail:
        k = 0
        for j, b in enumerate(tail):
            k |= b << (8 * j)
        h ^= _k1(k)
    return _fmix(h, n)


def s32(x):
    return x - 2 ** 32 if x >= 2 ** 31 else x


PUBLISHED = [  # (bytes, seed, MurmurHash3_x86_32)
    (b"", 0, 0), (b"", 1, 0x514E28B7), (b"", 0xFFFFFFFF, 0x81F16F39), (b"\0\0\0\0", 0, 0x2362F9DE),
    (b"\xff\xff\xff\xff", 0, 0x76293B50), (b"\x21\x43\x65\x87", 0, 0xF55B516B),
    (b"\x21\x43\x65\x87", 0x5082EDEE, 0x2362F9DE), (b"\x21\x43\x65", 0, 0x7E4A8634), (b"\x21\x43", 0, 0xA0F7B07A),
    (b"\x21", 0, 0x72661CF4), (b"\0\0\0", 0, 0x85F0B427), (b"\0\0", 0, 0x30F4C306), (b"\0", 0, 0x514E28B7),
    (b"aaaa", 0x9747B28C, 0x5A97808A), (b"aaa", 0x9747B28C, 0x283E0130), (b"aa", 0x9747B28C, 0x5D211726),
    (b"a", 0x9747B28C, 0x7FA09EA6), (b"abcd", 0x9747B28C, 0xF0478627), (b"abc", 0x9747B28C, 0xC84A62DD),
    (b"ab", 0x9747B28C, 0x74875592), (b"Hello, world!", 0x9747B28C, 0x24884CBA),
    (b"The quick brown fox jumps over the lazy dog", 0x9747B28C, 0x2FA826CD),
]


def test_python_murmur3_against_published_vectors():
    for data, seed, want in PUBLISHED:
        assert murmur3(data, seed) == want, (data, seed)
    # the two tails agree exactly when there is no tail
    assert murmur3(b"abcd", 7, True) == murmur3(b"abcd", 7)
    assert murmur3(b"abcde", 7, True) != murmur3(b"abcde", 7)


def hash_int(v, seed):
    return murmur3(struct.pack("<i", v), seed)


def hash_long(v, seed):
    return murmur3(struct.pack("<q", v), seed)


def test_spark_documented_hash():
    """Spark's SQL reference: hash('Spark', array(123), 2) = -1321691492 (a string, an array whose
    element is hashed as an int into the running hash, an int)."""
    h = murmur3(b"Spark", 42, spark_tail=True)
    h = hash_int(123, h)
    h = hash_int(2, h)
    assert s32(h) == -1321691492


def java_bytes(v):
    """BigInteger.toByteArray: bitLength / 8 + 1 big-endian two's-complement bytes."""
    bl = (v if v >= 0 else ~v).bit_length()
    return v.to_bytes(bl // 8 + 1, "big", signed=True)


def key_hash(kind, v, seed):
    if kind == abi.HASH_INT:
        return hash_int(v, seed)
    if kind == abi.HASH_LONG:
        return hash_long(v, seed)
    if kind == abi.HASH_DOUBLE:
        bits = 0 if v == 0.0 else (0x7FF8000000000000 if math.isnan(v) else struct.unpack("<Q", struct.pack("<d", v))[0])
        return murmur3(struct.pack("<Q", bits), seed)
    if kind == abi.HASH_FLOAT:  # Java's (float) narrowing: round to nearest even, past the range infinite
        with np.errstate(over="ignore"):
            f = np.array([v], np.float64).astype(np.float32)
        bits = 0x7FC00000 if math.isnan(v) else (0 if f[0] == 0.0 else int(f.view(np.uint32)[0]))
        return murmur3(struct.pack("<I", bits), seed)
    if kind == abi.HASH_BOOL:
        return hash_int(1 if v else 0, seed)
    if kind == abi.HASH_UTF8:
        return murmur3(v, seed, spark_tail=True)
    return murmur3(java_bytes(v), seed, spark_tail=True)


def brute_ids(kinds, vals, nulls, n, parts, seed):
    hs = []
    for i in range(n):
        h = seed & M32
        for c, kind in enumerate(kinds):
            if not nulls[c][i]:
                h = key_hash(kind, vals[c][i], h)
        hs.append(s32(h))
    return [h % parts for h in hs], hs


DEC_EDGES = [0, -1, 1, 127, 128, -128, -129, 255, 256, 2 ** 31, -2 ** 31, 2 ** 63, -2 ** 63 - 1, 2 ** 64,
             10 ** 37, -10 ** 37, 2 ** 127 - 1, -2 ** 127, 2 ** 120, -(2 ** 120) - 1]
FLT_EDGES = [0.0, -0.0, float("nan"), float("inf"), -float("inf"), 1.5, -2.25, 0.1, 1e-40, 3.4e38, 1e300, -1e-300]
DBL_EDGES = [0.0, -0.0, float("nan"), float("inf"), -float("inf"), 1.5, -2.25, 0.1, 5e-324, 1.7976931348623157e308]


def make_key(kind, rng, n, nulls):
    """(Column, Python values) of one key kind."""
    if kind == abi.HASH_INT:
        v = [int(x) for x in rng.integers(-2 ** 31, 2 ** 31, n)]
        v[:3] = [-2 ** 31, 2 ** 31 - 1, 0][:n]
        return Column.ints(v, nulls=nulls), v
    if kind == abi.HASH_LONG:
        v = [int(x) for x in rng.integers(-2 ** 63, 2 ** 63 - 1, n, dtype=np.int64)]
        v[:3] = [-2 ** 63, 2 ** 63 - 1, -1][:n]
        return Column.longs(v, nulls=nulls), v
    if kind in (abi.HASH_DOUBLE, abi.HASH_FLOAT):
        edges = DBL_EDGES if kind == abi.HASH_DOUBLE else FLT_EDGES
        v = [edges[int(k)] if k < len(edges) else float(x) for k, x in
             zip(rng.integers(0, 2 * len(edges), n), rng.standard_normal(n) * 1e3)]
        col = Column.doubles(v, nulls=nulls)
        if n > 1:  # a NaN with other payload bits hashes as the canonical one
            col.data.view(np.uint64)[1] = np.uint64(0x7FF0000000000123)
            v[1] = float("nan")
        return col, v
    if kind == abi.HASH_BOOL:
        v = [bool(x) for x in rng.random(n) < 0.5]
        return Column.bools(v, nulls=nulls), v
    if kind == abi.HASH_UTF8:
        strs = ["".join(chr(int(c)) for c in rng.integers(32, 0x500, int(k))) for k in rng.integers(0, 10, n)]
        return utf8_column(strs, nulls), [s.encode() for s in strs]
    v = [DEC_EDGES[int(k)] if k < len(DEC_EDGES) else int(x) * 10 ** 20 + int(y)
         for k, x, y in zip(rng.integers(0, 2 * len(DEC_EDGES), n), rng.integers(-10 ** 9, 10 ** 9, n),
                            rng.integers(-10 ** 9, 10 ** 9, n))]
    return Column.decimals(v, nulls=nulls), v


ALL_KINDS = [abi.HASH_INT, abi.HASH_LONG, abi.HASH_DOUBLE, abi.HASH_BOOL, abi.HASH_UTF8, abi.HASH_DECIMAL128,
             abi.HASH_FLOAT]


@pytest.mark.parametrize("kind", ALL_KINDS)
def test_each_kind_against_python_murmur3(kind):
    rng = np.random.default_rng(100 + kind)
    n = 300
    nulls = rng.random(n) < 0.15
    col, vals = make_key(kind, rng, n, nulls)
    for parts, seed in ((200, 42), (7, 42), (1, 42), (2 ** 31 - 1, 42), (16, 12345)):
        ids, hs = oracle.hash_partition_ids([col], [kind], n, parts, seed)
        want_ids, want_h = brute_ids([kind], [vals], [nulls], n, parts, seed)
        assert hs.tolist() == want_h, (kind, parts, seed)
        assert ids.tolist() == want_ids, (kind, parts, seed)


def test_several_keys_dictionary_strings_and_no_keys():
    rng = np.random.default_rng(5)
    n = 200
    kinds = [abi.HASH_UTF8, abi.HASH_INT, abi.HASH_DECIMAL128, abi.HASH_DOUBLE]
    nulls = [rng.random(n) < 0.2 for _ in kinds]
    cols, vals = zip(*[make_key(k, rng, n, z) for k, z in zip(kinds, nulls)])
    ids, hs = oracle.hash_partition_ids(list(cols), kinds, n, 200)
    want_ids, want_h = brute_ids(kinds, vals, nulls, n, 200, 42)
    assert hs.tolist() == want_h and ids.tolist() == want_ids
    # the same strings dictionary-encoded hash the same
    words = ["", "a", "abcd", "été", "x" * 13]
    codes = rng.integers(0, len(words), n)
    dcol = utf8_column(words, nulls[0], codes)
    pcol = utf8_column([words[c] for c in codes], nulls[0])
    assert (oracle.hash_partition_ids([dcol], [abi.HASH_UTF8], n, 97)[1] ==
            oracle.hash_partition_ids([pcol], [abi.HASH_UTF8], n, 97)[1]).all()
    # no keys: every row the seed
    ids, hs = oracle.hash_partition_ids([], [], 5, 10)
    assert hs.tolist() == [42] * 5 and ids.tolist() == [2] * 5


def test_partition_rows_against_stable_sort():
    rng = np.random.default_rng(9)
    for n, parts in ((0, 3), (1, 1), (50, 7), (5000, 200), (3000, 10 ** 5)):
        ids = rng.integers(0, parts, n).astype(np.int32)
        rows, dest, starts = oracle.partition_rows(ids, parts)
        order = np.argsort(ids, kind="stable")
        assert rows.tolist() == order.tolist()
        assert dest[rows].tolist() == list(range(n))
        assert starts.tolist() == [0] + np.cumsum(np.bincount(ids, minlength=parts)).tolist()
    with pytest.raises(RuntimeError):
        oracle.partition_rows(np.array([0, 3], np.int32), 3)


def test_round_robin_reference_expectation():
    ids, nxt = oracle.round_robin_ids(10, 4, 2)
    assert ids.tolist() == [2, 3, 0, 1, 2, 3, 0, 1, 2, 3] and nxt == 0
