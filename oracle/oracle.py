"""ctypes front end of the CPU oracle (oracle/libvroracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, the package's smoke check and bench.py's cpu_baseline leg import this module, and
only as the checker or the timed CPU baseline; the product path never does.  The oracle takes the
same vr_column / vr_agg_spec / vr_join_spec structs as the device library (include/vr.h), host
memory only.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from vecruntime_amd import abi
from vecruntime_amd.columns import OutColumn, c_array

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvroracle.so")


def build():
    """Compiles the oracle (make in oracle/)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.vro_last_error.restype = C.c_char_p
        L.vro_agg_create.argtypes = [C.POINTER(abi.VrAggSpec), C.POINTER(P)]
        L.vro_agg_update.argtypes = [P, P, P, C.c_int64, P, P]
        L.vro_agg_num_groups.argtypes = [P]
        L.vro_agg_num_groups.restype = C.c_int64
        L.vro_agg_output_layout.argtypes = [P, P, C.c_int32]
        L.vro_agg_emit.argtypes = [P, C.c_int64, C.c_int64, P, P]
        L.vro_agg_key_dictionary.argtypes = [P, C.c_int32, P, P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.vro_agg_destroy.argtypes = [P]
        L.vro_join_build.argtypes = [C.POINTER(abi.VrJoinSpec), P, C.c_int64, C.POINTER(P)]
        L.vro_join_distinct_keys.argtypes = [P]
        L.vro_join_distinct_keys.restype = C.c_int64
        L.vro_join_probe.argtypes = [P, P, C.c_int32, P, C.c_int64, P, C.c_int32, C.c_int32,
                                     C.POINTER(abi.VrProbeOutput)]
        L.vro_join_unmatched_build.argtypes = [P, C.POINTER(abi.VrProbeOutput)]
        L.vro_filter_compare.argtypes = [C.POINTER(abi.VrColumn), C.c_int32, C.POINTER(abi.VrScalar), P, C.c_int64, P,
                                         C.POINTER(C.c_int64)]
        L.vro_compact.argtypes = [P, C.c_int32, C.c_int64, P, P, C.POINTER(C.c_int64)]
        L.vro_gather.argtypes = [C.POINTER(abi.VrColumn), P, C.c_int64, C.POINTER(abi.VrOutColumn)]
        L.vro_join_destroy.argtypes = [P]
        L.vro_hash_partition_ids.argtypes = [P, P, C.c_int32, C.c_int64, C.c_int32, C.c_int32, P, P]
        L.vro_round_robin_ids.argtypes = [C.c_int64, C.c_int32, C.c_int32, P, C.POINTER(C.c_int32)]
        L.vro_partition_rows.argtypes = [P, C.c_int64, C.c_int32, P, P, P]
        L.vro_agg_partial_threads.argtypes = [C.POINTER(abi.VrAggSpec), P, P, P, C.c_int32, C.c_int32,
                                              C.POINTER(C.c_double)]
        L.vro_agg_partial_threads.restype = C.c_int64
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (abi.STATUS_NAMES.get(status, status), msg))
        self.status = status


def _check(st):
    if st != abi.VR_OK:
        raise OracleError(st, lib().vro_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data


class OracleAgg:
    """GroupKeyTable + GroupedAccumulators restated on the CPU (group ids in first-seen order)."""

    def __init__(self, spec):
        self.spec = spec
        self.h = C.c_void_p()
        _check(lib().vro_agg_create(C.byref(spec.c), C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib().vro_agg_destroy(self.h)
            self.h = None

    def update(self, keys, values, n, selection=None):
        ka, va = c_array(keys), c_array(values)
        ids = np.zeros(max(1, n), np.int32)
        _check(lib().vro_agg_update(self.h, C.cast(ka, C.c_void_p), C.cast(va, C.c_void_p), n,
                                    _ptr(selection), ids.ctypes.data))
        return ids[:n]

    def num_groups(self):
        return lib().vro_agg_num_groups(self.h)

    def emit(self, start=0, stop=None):
        """Emits groups [start, stop): (key OutColumns, aggregate OutColumns)."""
        if stop is None:
            stop = self.num_groups()
        m = stop - start
        keys = [OutColumn(abi.INT32 if t == abi.UTF8 else t, m) for t in self.spec.key_types]
        outs = [OutColumn(t, m) for t in self.spec.output_types()]
        ka, oa = c_array(keys, abi.VrOutColumn), c_array(outs, abi.VrOutColumn)
        _check(lib().vro_agg_emit(self.h, start, stop, C.cast(ka, C.c_void_p), C.cast(oa, C.c_void_p)))
        return keys, outs

    def key_dictionary(self, c):
        count, nbytes = C.c_int64(), C.c_int64()
        _check(lib().vro_agg_key_dictionary(self.h, c, None, None, C.byref(count), C.byref(nbytes)))
        offs = np.zeros(count.value + 1, np.int32)
        data = np.zeros(max(1, nbytes.value), np.uint8)
        _check(lib().vro_agg_key_dictionary(self.h, c, offs.ctypes.data, data.ctypes.data,
                                            C.byref(count), C.byref(nbytes)))
        raw = data.tobytes()
        return [raw[offs[i]:offs[i + 1]].decode("utf-8") for i in range(count.value)]


class OracleJoin:
    """BuildTable.build + the probe iterator's pair / semi / anti emission, on the CPU."""

    def __init__(self, spec, columns, n):
        self.spec = spec
        self.n = n
        self.columns = columns  # borrowed by the oracle: keep alive
        self._ca = c_array(columns)
        self.h = C.c_void_p()
        _check(lib().vro_join_build(C.byref(spec.c), C.cast(self._ca, C.c_void_p), n, C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib().vro_join_destroy(self.h)
            self.h = None

    def distinct_keys(self):
        return lib().vro_join_distinct_keys(self.h)

    def probe(self, columns, key_columns, n, join_type, selection=None, drop_null_keys=False,
              probe_out=(), build_out=(), capacity=None, residual=None):
        """Returns (probe_idx, build_idx or None, output columns, rows)."""
        if capacity is None:
            capacity = max(1, n) * 4
        ca = c_array(columns)
        kc = abi.int_array(key_columns)
        pidx = np.zeros(capacity, np.int32)
        bidx = np.zeros(capacity, np.int32) if join_type in (abi.INNER, abi.OUTER) else None
        types = [columns[c].type for c in probe_out] + [self.spec.column_types[c] for c in build_out]
        outs = [OutColumn(t, capacity) for t in types]
        oa = c_array(outs, abi.VrOutColumn)
        po, bo = abi.int_array(list(probe_out)), abi.int_array(list(build_out))
        out = abi.VrProbeOutput(abi.MEM_HOST, capacity, pidx.ctypes.data,
                                bidx.ctypes.data if bidx is not None else None,
                                len(probe_out), C.cast(po, C.POINTER(C.c_int32)),
                                len(build_out), C.cast(bo, C.POINTER(C.c_int32)),
                                C.cast(oa, C.POINTER(abi.VrOutColumn)), 0)
        if residual is not None:
            out.residual = C.addressof(residual)
        _check(lib().vro_join_probe(self.h, C.cast(ca, C.c_void_p), len(columns), C.cast(kc, C.c_void_p), n,
                                    _ptr(selection), join_type, 1 if drop_null_keys else 0, C.byref(out)))
        r = out.rows
        return pidx[:r], (bidx[:r] if bidx is not None else None), outs, r

    def unmatched_build(self, probe_types=(), build_out=(), capacity=None):
        """The build rows no probe paired: (probe_idx (-1), build_idx, output columns, rows)."""
        if capacity is None:
            capacity = max(1, self.n)
        pidx = np.zeros(capacity, np.int32)
        bidx = np.zeros(capacity, np.int32)
        types = list(probe_types) + [self.spec.column_types[c] for c in build_out]
        outs = [OutColumn(t, capacity) for t in types]
        oa = c_array(outs, abi.VrOutColumn)
        bo = abi.int_array(list(build_out))
        out = abi.VrProbeOutput(abi.MEM_HOST, capacity, pidx.ctypes.data, bidx.ctypes.data, len(probe_types), None,
                                len(build_out), C.cast(bo, C.POINTER(C.c_int32)), C.cast(oa, C.POINTER(abi.VrOutColumn)), 0)
        _check(lib().vro_join_unmatched_build(self.h, C.byref(out)))
        r = out.rows
        return pidx[:r], bidx[:r], outs, r


def partial_threads(spec, batch_keys, batch_values, lengths, threads):
    """Aggregates batches round-robin over `threads` independent tables (Spark tasks); returns
    (total groups over the tables, wall seconds)."""
    nb = len(lengths)
    ka = c_array([c for b in batch_keys for c in b])
    va = c_array([c for b in batch_values for c in b])
    ln = np.asarray(lengths, np.int64)
    secs = C.c_double()
    g = lib().vro_agg_partial_threads(C.byref(spec.c), C.cast(ka, C.c_void_p), C.cast(va, C.c_void_p),
                                      ln.ctypes.data, nb, threads, C.byref(secs))
    return g, secs.value


def filter_compare(col, op, value, n, selection=None):
    """The selection of `col <op> value` over rows [0, n): (bitmap bytes padded to 8, count)."""
    out = np.zeros(max(8, ((n + 63) // 64) * 8), np.uint8)
    sc = abi.scalar(value, col.type)
    cnt = C.c_int64()
    _check(lib().vro_filter_compare(C.byref(col.c()), int(op), C.byref(sc), _ptr(selection), int(n), out.ctypes.data,
                                    C.byref(cnt)))
    return out, cnt.value


def compact(columns, n, selection, types, capacity, utf8_bytes=0):
    """The selected rows of each column: (OutColumns, rows).  `types`: the output lane per column."""
    outs = []
    for t in types:
        if t != abi.UTF8:
            oc = OutColumn(t, capacity)
        else:
            oc = OutColumn(t, capacity, data=np.zeros(max(1, utf8_bytes), np.uint8))
            oc.offsets = np.zeros(capacity + 1, np.int32)
            oc._c = abi.VrOutColumn(t, abi.MEM_HOST, capacity, _ptr(oc.validity), _ptr(oc.data), _ptr(oc.offsets))
        outs.append(oc)
    ca, oa = c_array(columns), c_array(outs, abi.VrOutColumn)
    rows = C.c_int64()
    _check(lib().vro_compact(C.cast(ca, C.c_void_p), len(columns), int(n), _ptr(selection), C.cast(oa, C.c_void_p),
                             C.byref(rows)))
    return outs, rows.value


def gather(col, idx, utf8_bytes=0):
    """Rows idx[k] of `col` (-1: null): an OutColumn of len(idx) rows."""
    idx = np.ascontiguousarray(idx, np.int32)
    m = len(idx)
    if col.type == abi.UTF8:
        oc = OutColumn(abi.UTF8, max(1, m), data=np.zeros(max(1, utf8_bytes), np.uint8))
        oc.offsets = np.zeros(max(1, m) + 1, np.int32)
        oc._c = abi.VrOutColumn(abi.UTF8, abi.MEM_HOST, max(1, m), _ptr(oc.validity), _ptr(oc.data), _ptr(oc.offsets))
    else:
        oc = OutColumn(col.type, max(1, m))
    _check(lib().vro_gather(C.byref(col.c()), idx.ctypes.data, m, C.byref(oc.c())))
    return oc


def hash_partition_ids(keys, kinds, n, num_partitions, seed=42):
    """Spark's Pmod(Murmur3Hash(keys, seed), num_partitions): (ids, hashes), int32 arrays of n."""
    ids = np.zeros(max(1, n), np.int32)
    hashes = np.zeros(max(1, n), np.int32)
    ka = c_array(keys) if keys else None
    kd = np.ascontiguousarray(kinds, np.int32)
    _check(lib().vro_hash_partition_ids(C.cast(ka, C.c_void_p) if ka is not None else None, kd.ctypes.data, len(keys),
                                        int(n), int(num_partitions), int(seed), ids.ctypes.data, hashes.ctypes.data))
    return ids[:n], hashes[:n]


def round_robin_ids(n, num_partitions, start):
    """(ids, next start) of Spark's round-robin partitioning."""
    ids = np.zeros(max(1, n), np.int32)
    nxt = C.c_int32()
    _check(lib().vro_round_robin_ids(int(n), int(num_partitions), int(start), ids.ctypes.data, C.byref(nxt)))
    return ids[:n], nxt.value


def partition_rows(ids, num_partitions):
    """The stable grouping of rows by partition: (rows, dest, starts)."""
    ids = np.ascontiguousarray(ids, np.int32)
    n = len(ids)
    rows = np.zeros(max(1, n), np.int32)
    dest = np.zeros(max(1, n), np.int32)
    starts = np.zeros(num_partitions + 1, np.int64)
    _check(lib().vro_partition_rows(ids.ctypes.data, n, int(num_partitions), rows.ctypes.data, dest.ctypes.data,
                                    starts.ctypes.data))
    return rows[:n], dest[:n], starts
