"""Join cases over key tuples wider than 64 bits and UTF8 keys, shared by the oracle's pin
(tests/test_oracle.py, nested loop in pure Python) and the device parity tests.

Each case gives the build and probe key columns three ways: as host Columns, and as the Python
values the nested loop compares (doubles as their raw bits -- the kernels compare key lanes bit for
bit -- decimals as ints, strings as str, nulls as None)."""
import struct

import numpy as np

from vecruntime_amd import abi
from vecruntime_amd.columns import Column, pack_bits

INT64_MIN, INT64_MAX = -(1 << 63), (1 << 63) - 1


def f64_bits(x):
    return struct.unpack("<q", struct.pack("<d", x))[0]


def _col(kind, vals, nulls, dictionary=None, codes=None):
    """A host Column of `vals` (Python values) with nulls."""
    n = len(vals)
    if kind == "int32":
        return Column.ints([0 if v is None else v for v in vals], nulls=nulls)
    if kind == "int64":
        return Column.longs([0 if v is None else v for v in vals], nulls=nulls)
    if kind == "float64":
        return Column.doubles([0.0 if v is None else v for v in vals], nulls=nulls)
    if kind == "decimal":
        c = Column.decimals([0 if v is None else v for v in vals], nulls=nulls)
        return c
    if kind == "utf8":
        if codes is not None:  # dictionary-encoded
            return Column.dictionary_strings(codes, dictionary, nulls=nulls)
        col = Column.strings(["" if v is None else v for v in vals])
        return Column(abi.UTF8, col.data, pack_bits(~np.asarray(nulls, bool)) if any(nulls) else None,
                      offsets=col.offsets, length=n)
    raise ValueError(kind)


LANE = {"int32": abi.INT32, "int64": abi.INT64, "float64": abi.FLOAT64, "decimal": abi.DECIMAL128, "utf8": abi.UTF8}


def make_case(seed, kinds, nb, npr, distinct, null_rate=0.05, hit=0.7, dict_probe=False):
    """Build and probe sides over key columns of `kinds`: `distinct` key tuples, build rows drawn
    with duplicates, probe rows hitting with probability `hit` (else fresh tuples), nulls mixed in.
    Returns dict(spec, build (Columns: keys then an INT32 payload), probe (Columns: keys then an
    INT64 payload), pybuild / pyprobe (key value lists for the nested loop))."""
    rng = np.random.default_rng(seed)
    words = ["k%d_%s" % (i, "xyz"[i % 3] * (i % 17)) for i in range(4 * distinct + 8)]
    specials = {"float64": [0.0, -0.0, float("nan"), float("inf"), -1.5],
                "decimal": [0, -1, INT64_MIN, INT64_MAX, (1 << 100) + 7, -(1 << 90)],
                "int64": [INT64_MIN, INT64_MAX, 0, -1], "int32": [-2 ** 31, 2 ** 31 - 1, 0, -1]}

    def draw(kind, m):
        if kind == "int32":
            v = rng.integers(-2 ** 31, 2 ** 31, m).tolist()
        elif kind == "int64":
            v = rng.integers(INT64_MIN, INT64_MAX, m, endpoint=True).tolist()
        elif kind == "float64":
            v = rng.standard_normal(m).tolist()
        elif kind == "decimal":
            v = [int(a) * (1 << 64) + int(b) for a, b in zip(rng.integers(-2 ** 40, 2 ** 40, m), rng.integers(0, 2 ** 63, m))]
        else:
            v = [words[int(x)] for x in rng.integers(0, len(words), m)]
        sp = specials.get(kind)
        if sp:
            for i in range(m):
                if rng.random() < 0.1:
                    v[i] = sp[int(rng.integers(0, len(sp)))]
        return v

    tuples = list(zip(*[draw(k, distinct) for k in kinds]))
    fresh = list(zip(*[draw(k, npr) for k in kinds]))
    bidx = rng.integers(0, distinct, nb)
    brows = [tuples[i] for i in bidx]
    prows = [tuples[int(rng.integers(0, distinct))] if rng.random() < hit else fresh[i] for i in range(npr)]
    bnull = rng.random((len(kinds), nb)) < null_rate
    pnull = rng.random((len(kinds), npr)) < null_rate

    def side(rows, nulls, dict_codes):
        cols, py = [], []
        for c, k in enumerate(kinds):
            vals = [r[c] for r in rows]
            nl = nulls[c].tolist()
            if k == "utf8" and dict_codes:
                uniq = sorted(set(vals))
                dictionary = Column.strings(uniq)
                codes = [uniq.index(v) for v in vals]
                cols.append(_col(k, vals, nl, dictionary, codes))
            else:
                cols.append(_col(k, vals, nl))
            pv = [None if nl[i] else (f64_bits(vals[i]) if k == "float64" else vals[i]) for i in range(len(vals))]
            py.append(pv)
        return cols, py

    bcols, pybuild = side(brows, bnull, False)
    pcols, pyprobe = side(prows, pnull, dict_probe)
    bpay = rng.integers(-2 ** 31, 2 ** 31, nb).astype(np.int32)
    ppay = rng.integers(-10 ** 12, 10 ** 12, npr)
    spec = abi.JoinSpec([LANE[k] for k in kinds], [LANE[k] for k in kinds] + [abi.INT32], list(range(len(kinds))),
                        capacity_hint=nb, dense_index=-1)
    return dict(spec=spec, build=bcols + [Column.ints(bpay)], probe=pcols + [Column.longs(ppay)], bpay=bpay, ppay=ppay,
                pybuild=pybuild, pyprobe=pyprobe, nb=nb, npr=npr, nk=len(kinds))


# key kinds of the cases: more than 64 bits of lanes (wide), or 64-bit keys of new lane types
KIND_SETS = [("int64", "int64"), ("int32", "int64", "decimal"), ("decimal",), ("float64",), ("utf8",),
             ("utf8", "int64"), ("int32", "utf8"), ("utf8", "int32", "int64", "decimal")]
