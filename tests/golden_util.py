"""Loads the golden fixtures (tests/golden/*.npz) into Arrow-layout batches."""
import os

import numpy as np

from vecruntime_amd import abi
from vecruntime_amd.columns import Column, pack_bits

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def batch_slices(lengths):
    off = 0
    for n in lengths:
        yield slice(off, off + int(n)), int(n)
        off += int(n)


def grouped_batches(z, dict_encoded=True):
    """GroupedAggregationTest batches: keys (k1 INT32 with nulls, k2 UTF8), values (vals FLOAT64,
    lvals INT64, shared nulls), selection bitmap or None."""
    dict_offsets = z["dict_offsets"]
    raw = z["dict_bytes"].tobytes()
    strings = [raw[dict_offsets[i]:dict_offsets[i + 1]].decode() for i in range(len(dict_offsets) - 1)]
    dictionary = Column.strings(strings)
    has_sel = not bool(z["sel"].all())
    out = []
    for sl, n in batch_slices(z["lengths"]):
        k1 = Column.ints(z["k1"][sl], nulls=z["k1_null"][sl])
        codes = z["k2_code"][sl]
        if dict_encoded:
            k2 = Column.dictionary_strings(np.where(codes < 0, 0, codes), dictionary, nulls=codes < 0)
        else:
            k2 = Column.strings([strings[c] if c >= 0 else None for c in codes])
        vals = Column.doubles(z["vals"][sl], nulls=z["val_null"][sl])
        lvals = Column.longs(z["lvals"][sl], nulls=z["val_null"][sl])
        sel = pack_bits(z["sel"][sl]) if has_sel else None
        out.append(([k1, k2], [vals, lvals], n, sel))
    return out, strings


GROUPED_AGGS = [
    abi.agg_desc(abi.SUM_DOUBLE, 0),
    abi.agg_desc(abi.SUM_LONG, 1),
    abi.agg_desc(abi.COUNT_STAR),
    abi.agg_desc(abi.COUNT, 0),
    abi.agg_desc(abi.MIN, 0),
    abi.agg_desc(abi.MAX, 0),
]


def grouped_spec(deterministic=False):
    return abi.AggSpec([abi.INT32, abi.UTF8], [abi.FLOAT64, abi.INT64], GROUPED_AGGS,
                       deterministic=deterministic)


def f64_bits(a):
    return np.asarray(a, np.float64).view(np.int64)
