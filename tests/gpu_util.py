"""Helpers for the device tests: running the same batches through the device library and the CPU
oracle, and comparing group multisets."""
import math
import struct

import numpy as np

from vecruntime_amd import abi

SUM_KINDS = (abi.SUM_DOUBLE, abi.AVG_DOUBLE)


def out_roles(spec):
    """Per output column: 'fsum' (double sum/avg: tolerance), 'exact' otherwise."""
    roles = []
    for a in spec.aggs:
        result = bool(a.flags & abi.RESULT)
        if a.kind in (abi.SUM_DOUBLE, abi.MERGE_SUM_DOUBLE):
            roles.append("fsum")
        elif a.kind in (abi.AVG_DOUBLE, abi.MERGE_AVG_DOUBLE):
            roles += ["fsum"] if result else ["fsum", "exact"]
        elif a.kind in (abi.SUM_WIDE, abi.AVG_WIDE, abi.MERGE_SUM_WIDE, abi.MERGE_AVG_WIDE):
            roles += ["exact"] if result else ["exact", "exact"]
        else:
            roles.append("exact")
    return roles


def group_rows(agg_keys, agg_outs, G, spec, key_dicts):
    """{key tuple: [output values]} with UTF8 ids resolved through the table dictionaries and
    doubles as (float, bits)."""
    kv = []
    for c, t in enumerate(spec.key_types):
        vals = agg_keys[c].values(G)
        if t == abi.UTF8:
            vals = [None if v is None else key_dicts[c][v] for v in vals]
        elif t == abi.FLOAT64:
            vals = [None if v is None else struct.pack("<d", v) for v in vals]
        kv.append(vals)
    cols = []
    for oc, t in zip(agg_outs, spec.output_types()):
        vals = oc.values(G)
        if t == abi.FLOAT64:
            vals = [None if v is None else (v, struct.unpack("<q", struct.pack("<d", v))[0]) for v in vals]
        cols.append(vals)
    out = {}
    for g in range(G):
        k = tuple(kv[c][g] for c in range(len(kv)))
        assert k not in out, "duplicate group %r" % (k,)
        out[k] = [col[g] for col in cols]
    return out


def compare_groups(dev_rows, ref_rows, roles, rel=1e-12, scale=None):
    """Same key set; exact outputs equal (doubles bit for bit); double sums within rel * scale
    (scale[key][col] = sum of |x| of the group, default |expected|) or both NaN."""
    assert set(dev_rows) == set(ref_rows), "group keys differ: %d vs %d (only dev: %r, only ref: %r)" % (
        len(dev_rows), len(ref_rows), list(set(dev_rows) - set(ref_rows))[:3], list(set(ref_rows) - set(dev_rows))[:3])
    worst = 0.0
    for k, ref in ref_rows.items():
        dev = dev_rows[k]
        for c, (d, r) in enumerate(zip(dev, ref)):
            if r is None or d is None:
                assert r is None and d is None, (k, c, d, r)
                continue
            if isinstance(r, tuple):
                if roles[c] == "fsum":
                    if math.isnan(r[0]) or math.isnan(d[0]):
                        assert math.isnan(r[0]) and math.isnan(d[0]), (k, c, d, r)
                        continue
                    if math.isinf(r[0]) or math.isinf(d[0]):
                        assert r[0] == d[0], (k, c, d, r)
                        continue
                    s = abs(r[0]) if scale is None else scale[k][c]
                    err = abs(d[0] - r[0])
                    assert err <= rel * s + 1e-300, (k, c, d[0], r[0], err, s)
                    if s:
                        worst = max(worst, err / s)
                else:
                    assert d[1] == r[1], (k, c, d, r)
            else:
                assert d == r, (k, c, d, r)
    return worst


def split_batches(arrays, lengths):
    off = 0
    for n in lengths:
        yield {k: v[off:off + n] for k, v in arrays.items()}, n
        off += n
