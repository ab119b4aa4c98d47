"""Python front end of libvrgpu.so (include/vr.h) through ctypes.

The classes mirror the reference's operator-level use of the kernels module
(kernels/src/main/java/io/vecruntime/kernels/GroupKeyTable.java, GroupedAccumulators.java and
spark/.../sql/vecruntime/VectorHashJoinExec.scala's BuildTable / probe iterator): a
GroupedAggregate is fed batches and emits keys plus Spark's partial buffers (or results); a
BuildTable is appended to, finished, then probed batch by batch.

The device library is required: importing this module on a machine where libvrgpu.so has not
been built raises, and there is no CPU fallback here.
"""
import ctypes as C
import os
import sys

import numpy as np

from . import abi
from .columns import OutColumn, c_array

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvrgpu.so")

if not os.path.exists(LIB_PATH):
    raise ImportError("libvrgpu.so is not built (run __graft_entry__.build() or python -m vecruntime_amd.build)")

try:
    # PyTorch ships its own HIP runtime and links to it by the unversioned name
    # (libamdhip64.so), which the dynamic loader does not match with an already loaded
    # libamdhip64.so.7.  Loaded after this library, torch would start a second runtime that
    # sees no GPU; loaded first, its runtime is the one this library binds to.
    import torch  # noqa: F401
except ImportError:
    pass

_P = C.c_void_p
_lib = C.CDLL(LIB_PATH)


def _sig(name, args, res=C.c_int32):
    f = getattr(_lib, name)
    f.argtypes = args
    f.restype = res
    return f


_sig("vr_last_error", [], C.c_char_p)
_sig("vr_device_open", [C.c_int32, C.c_int64, C.POINTER(_P)])
_sig("vr_device_close", [_P])
_sig("vr_device_memory", [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)])
_sig("vr_device_sync", [_P, _P])
_sig("vr_stats_get", [_P, C.POINTER(abi.VrStats)])
_sig("vr_agg_create", [_P, C.POINTER(abi.VrAggSpec), C.POINTER(_P)])
_sig("vr_agg_update", [_P, _P, _P, C.c_int64, _P, _P])
_sig("vr_agg_update_filtered", [_P, _P, _P, C.c_int64, C.c_int32, C.c_int32, _P])
_sig("vr_agg_update_batches", [_P, C.POINTER(abi.VrBatchSet), C.POINTER(abi.VrLaunchConfig), _P])
_sig("vr_agg_num_groups", [_P, C.POINTER(C.c_int64)])
_sig("vr_agg_set_launch_config", [_P, C.POINTER(abi.VrLaunchConfig)])
_sig("vr_agg_set_timing", [_P, C.c_int32])
_sig("vr_agg_graph_create", [_P, C.POINTER(abi.VrBatchSet), C.POINTER(abi.VrLaunchConfig), C.c_int64, C.POINTER(_P)])
_sig("vr_agg_graph_launch", [_P, _P])
_sig("vr_agg_graph_destroy", [_P])
_sig("vr_agg_kernel_time", [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)])
_sig("vr_agg_output_layout", [_P, C.POINTER(C.c_int32), _P, C.c_int32])
_sig("vr_agg_emit", [_P, C.c_int64, C.c_int64, _P, _P, _P])
_sig("vr_agg_key_dictionary", [_P, C.c_int32, _P, C.c_int64, _P, C.c_int64, C.POINTER(C.c_int64),
                               C.POINTER(C.c_int64)])
_sig("vr_agg_memory", [_P, C.POINTER(C.c_int64)])
_sig("vr_agg_prewarm", [_P, C.POINTER(abi.VrAggSpec)])
_sig("vr_agg_reset", [_P])
_sig("vr_agg_destroy", [_P])
_sig("vr_join_create", [_P, C.POINTER(abi.VrJoinSpec), C.POINTER(_P)])
_sig("vr_join_build_append", [_P, _P, C.c_int64, _P])
_sig("vr_join_build_finish", [_P, _P])
_sig("vr_join_info", [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)])
_sig("vr_join_probe", [_P, _P, C.c_int32, _P, C.c_int64, _P, C.c_int32, C.c_int32,
                       C.POINTER(abi.VrProbeOutput), _P])
_sig("vr_join_probe_noalloc", [_P, _P, C.c_int32, _P, C.c_int64, _P, C.c_int32, C.c_int32,
                               C.POINTER(abi.VrProbeOutput), _P])
_sig("vr_join_probe_batches", [_P, C.c_int32, _P, C.c_int32, _P, _P, _P, C.c_int32, C.c_int32,
                               C.POINTER(abi.VrProbeOutput), _P, _P])
_sig("vr_join_unmatched_build", [_P, C.POINTER(abi.VrProbeOutput), _P])
_sig("vr_join_memory", [_P, C.POINTER(C.c_int64)])
_sig("vr_filter_compare", [_P, C.POINTER(abi.VrColumn), C.c_int32, C.POINTER(abi.VrScalar), _P, C.c_int64, _P,
                           C.POINTER(C.c_int64), _P])
_sig("vr_compact_utf8_bytes", [_P, C.POINTER(abi.VrColumn), C.c_int64, _P, C.POINTER(C.c_int64), _P])
_sig("vr_compact", [_P, _P, C.c_int32, C.c_int64, _P, _P, C.POINTER(C.c_int64), _P])
_sig("vr_gather_utf8_bytes", [_P, C.POINTER(abi.VrColumn), _P, C.c_int64, C.POINTER(C.c_int64), _P])
_sig("vr_gather", [_P, C.POINTER(abi.VrColumn), _P, C.c_int64, C.POINTER(abi.VrOutColumn), _P])
_sig("vr_join_destroy", [_P])
_sig("vr_hash_partition_ids", [_P, _P, _P, C.c_int32, C.c_int64, C.c_int32, C.c_int32, _P, _P, _P])
_sig("vr_round_robin_ids", [_P, C.c_int64, C.c_int32, C.c_int32, _P, C.POINTER(C.c_int32), _P])
_sig("vr_partition_rows", [_P, _P, C.c_int64, C.c_int32, _P, _P, _P, _P])


def lib():
    return _lib


def lib_path():
    return LIB_PATH


class VrError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (abi.STATUS_NAMES.get(status, status), msg))
        self.status = status


def check(st):
    if st != abi.VR_OK:
        raise VrError(st, _lib.vr_last_error().decode())


def _stream(s):
    """A HIP stream handle: an int, a torch stream, or None.

    None is torch's current stream once torch has initialized the device, so that the call is
    ordered with the work that filled its input tensors and with whatever torch's caching allocator
    later gives their memory to.  torch's legacy default stream has handle 0, which the library
    takes for the handle's own (non-blocking) stream: that stream is not ordered with the legacy
    one, so the call first waits for torch's work and the wrapper synchronizes after a call that
    returns with work queued (_settle).  Without torch, None is the handle's own stream."""
    h = s if s is None or isinstance(s, int) else s.cuda_stream
    if not h:
        t = sys.modules.get("torch")
        if t is not None and t.cuda.is_initialized():
            cur = t.cuda.current_stream()
            if cur.cuda_stream:
                return cur.cuda_stream
            cur.synchronize()
        return None
    return h


def _settle(device_h, s):
    """After a call that may return with kernels queued: when it ran on the handle's own stream
    while torch uses its legacy default stream, wait for it (the caller may drop the inputs)."""
    h = s if s is None or isinstance(s, int) else s.cuda_stream
    if h:
        return
    t = sys.modules.get("torch")
    if t is not None and t.cuda.is_initialized() and not t.cuda.current_stream().cuda_stream:
        check(_lib.vr_device_sync(device_h, None))


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


class Device:
    def __init__(self, ordinal=0, hbm_budget=0):
        self.h = _P()
        check(_lib.vr_device_open(ordinal, int(hbm_budget), C.byref(self.h)))

    def close(self):
        if self.h:
            _lib.vr_device_close(self.h)
            self.h = None

    def memory(self):
        a, b = C.c_int64(), C.c_int64()
        check(_lib.vr_device_memory(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def sync(self, stream=None):
        check(_lib.vr_device_sync(self.h, _stream(stream)))

    def stats(self):
        s = abi.VrStats()
        check(_lib.vr_stats_get(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in abi.VrStats._fields_}


class PreparedBatches:
    """A vr_batch_set and the ctypes/numpy arrays it points to (kept alive with it)."""

    def __init__(self, bs, keep):
        self.bs = bs
        self.keep = keep


class GroupedAggregate:
    """Group-by aggregation handle: GroupKeyTable.assign + GroupedAccumulators updates per batch,
    writeKeys + the aggregate buffers (or results) at emission."""

    def __init__(self, device, spec):
        self.device = device
        self.spec = spec
        self.h = _P()
        check(_lib.vr_agg_create(device.h, C.byref(spec.c), C.byref(self.h)))
        self._keep = []

    def close(self):
        if self.h:
            _lib.vr_agg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, keys, values, n, selection=None, stream=None):
        ka, va = c_array(keys), c_array(values)
        check(_lib.vr_agg_update(self.h, C.cast(ka, _P), C.cast(va, _P), int(n), _ptr(selection), _stream(stream)))
        _settle(self.device.h, stream)

    @staticmethod
    def prepare(keys, values):
        """The ctypes arrays of one batch's columns, built once for repeated updates."""
        ka, va = c_array(keys), c_array(values)
        return (ka, va, C.cast(ka, _P), C.cast(va, _P), keys, values)

    def update_prepared(self, prep, n, filter_col=-1, threshold=0, stream=None):
        """vr_agg_update_filtered over prepare()d columns (filter_col -1: no filter)."""
        check(_lib.vr_agg_update_filtered(self.h, prep[2], prep[3], int(n), filter_col, threshold, _stream(stream)))
        _settle(self.device.h, stream)

    def update_filtered(self, keys, values, n, filter_col, threshold, stream=None):
        ka, va = c_array(keys), c_array(values)
        check(_lib.vr_agg_update_filtered(self.h, C.cast(ka, _P), C.cast(va, _P), int(n), filter_col, threshold,
                                          _stream(stream)))
        _settle(self.device.h, stream)

    @staticmethod
    def prepare_batches(batches, filter_col=-1, threshold=0):
        """The vr_batch_set of `batches` (list of (keys, values, n)) and the arrays it points to,
        built once for repeated vr_agg_update_batches calls over the same columns."""
        ka = c_array([c for b in batches for c in b[0]])
        va = c_array([c for b in batches for c in b[1]])
        ln = np.asarray([b[2] for b in batches], np.int64)
        bs = abi.VrBatchSet(len(batches), C.cast(ka, C.POINTER(abi.VrColumn)), C.cast(va, C.POINTER(abi.VrColumn)),
                            ln.ctypes.data_as(C.POINTER(C.c_int64)), filter_col, threshold)
        return PreparedBatches(bs, (ka, va, ln, batches))

    def capture(self, batches, max_groups, filter_col=-1, threshold=0, config=None):
        """vr_agg_graph_create: an AggGraph replaying one whole step (table cleared, every batch
        folded) over `batches` (list of (keys, values, n) with device columns, or a prepare_batches()
        result)."""
        if not isinstance(batches, PreparedBatches):
            batches = self.prepare_batches(batches, filter_col, threshold)
        cfg = config if config is not None else abi.launch_config()
        h = _P()
        check(_lib.vr_agg_graph_create(self.h, C.byref(batches.bs), C.byref(cfg), int(max_groups), C.byref(h)))
        return AggGraph(self, h, batches)

    def update_batches(self, batches, filter_col=-1, threshold=0, config=None, stream=None):
        """batches: list of (keys, values, n), or a prepare_batches() result (which carries its own
        filter).  One persistent launch when every column is device resident."""
        if not isinstance(batches, PreparedBatches):
            batches = self.prepare_batches(batches, filter_col, threshold)
        cfg = config if config is not None else abi.launch_config()
        check(_lib.vr_agg_update_batches(self.h, C.byref(batches.bs), C.byref(cfg), _stream(stream)))
        _settle(self.device.h, stream)
        self._keep = [batches]

    def set_timing(self, on):
        """Brackets every update-kernel launch with HIP events (a measurement aid)."""
        check(_lib.vr_agg_set_timing(self.h, 1 if on else 0))

    def kernel_time(self):
        """(device ms summed over the timed update-kernel launches, launches) since the last read."""
        ms, n = C.c_double(), C.c_int64()
        check(_lib.vr_agg_kernel_time(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def num_groups(self):
        g = C.c_int64()
        check(_lib.vr_agg_num_groups(self.h, C.byref(g)))
        return g.value

    def set_launch_config(self, cfg):
        check(_lib.vr_agg_set_launch_config(self.h, C.byref(cfg)))

    def output_types(self):
        n = C.c_int32()
        types = (C.c_int32 * 64)()
        check(_lib.vr_agg_output_layout(self.h, C.byref(n), types, 64))
        return list(types[: n.value])

    def emit(self, start=0, stop=None, keys=None, outs=None, stream=None):
        """Emits groups [start, stop) into host OutColumns (or the given ones); returns (keys, outs)."""
        if stop is None:
            stop = self.num_groups()
        m = stop - start
        if keys is None:
            keys = [OutColumn(abi.INT32 if t == abi.UTF8 else t, m) for t in self.spec.key_types]
        if outs is None:
            outs = [OutColumn(t, m) for t in self.output_types()]
        ka, oa = c_array(keys, abi.VrOutColumn), c_array(outs, abi.VrOutColumn)
        check(_lib.vr_agg_emit(self.h, start, stop, C.cast(ka, _P), C.cast(oa, _P), _stream(stream)))
        _settle(self.device.h, stream)
        return keys, outs

    def key_dictionary(self, c):
        return agg_key_dictionary(self.h, c)

    def memory(self):
        b = C.c_int64()
        check(_lib.vr_agg_memory(self.h, C.byref(b)))
        return b.value

    def reset(self):
        check(_lib.vr_agg_reset(self.h))


class BuildTable:
    """A join's build side (BuildTable): appended batches, then a finished, read-only table that
    streamed batches probe."""

    def __init__(self, device, spec):
        self.device = device
        self.spec = spec
        self.h = _P()
        check(_lib.vr_join_create(device.h, C.byref(spec.c), C.byref(self.h)))

    def close(self):
        if self.h:
            _lib.vr_join_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def append(self, columns, n, stream=None):
        ca = c_array(columns)
        check(_lib.vr_join_build_append(self.h, C.cast(ca, _P), int(n), _stream(stream)))

    def finish(self, stream=None):
        check(_lib.vr_join_build_finish(self.h, _stream(stream)))

    def info(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        check(_lib.vr_join_info(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def memory(self):
        b = C.c_int64()
        check(_lib.vr_join_memory(self.h, C.byref(b)))
        return b.value

    def probe(self, columns, key_columns, n, join_type, selection=None, drop_null_keys=False, probe_out=(),
              build_out=(), capacity=None, outputs=None, probe_idx=None, build_idx=None, mem=abi.MEM_HOST,
              stream=None, want_pairs=True, residual=None):
        """Returns (probe_idx, build_idx or None, output columns, rows).  Host outputs are
        allocated here unless given; device outputs (mem=MEM_DEVICE) must be given."""
        if capacity is None:
            _, _, maxc = self.info()
            capacity = max(1, n) * max(1, maxc if join_type == abi.INNER else 1)
        ca = c_array(columns)
        kc = abi.int_array(list(key_columns))
        if mem == abi.MEM_HOST:
            if want_pairs and probe_idx is None:
                probe_idx = np.zeros(capacity, np.int32)
            if want_pairs and build_idx is None and join_type in (abi.INNER, abi.OUTER):
                build_idx = np.zeros(capacity, np.int32)
            if outputs is None:
                types = [columns[c].type for c in probe_out] + [self.spec.column_types[c] for c in build_out]
                outputs = [OutColumn(t, capacity) for t in types]
        outputs = outputs or []
        oa = c_array(outputs, abi.VrOutColumn)
        po, bo = abi.int_array(list(probe_out)), abi.int_array(list(build_out))
        pairs = join_type in (abi.INNER, abi.OUTER)
        out = abi.VrProbeOutput(mem, int(capacity), _ptr(probe_idx), _ptr(build_idx) if pairs else None,
                                len(probe_out), C.cast(po, C.POINTER(C.c_int32)), len(build_out),
                                C.cast(bo, C.POINTER(C.c_int32)), C.cast(oa, C.POINTER(abi.VrOutColumn)), 0)
        if residual is not None:
            out.residual = C.addressof(residual)
        st = _lib.vr_join_probe(self.h, C.cast(ca, _P), len(columns), C.cast(kc, _P), int(n), _ptr(selection),
                                join_type, 1 if drop_null_keys else 0, C.byref(out), _stream(stream))
        check(st)
        r = out.rows
        if mem == abi.MEM_HOST:
            return (probe_idx[:r] if probe_idx is not None else None,
                    build_idx[:r] if build_idx is not None else None, outputs, r)
        return probe_idx, build_idx, outputs, r

    def probe_noalloc(self, columns, key_columns, n, join_type, outputs=(), probe_out=(), build_out=(), capacity=None,
                      probe_idx=None, build_idx=None, selection=None, drop_null_keys=False, residual=None, stream=None):
        """vr_join_probe_noalloc over device columns into device outputs (torch tensors): the rows
        produced."""
        ca = c_array(columns)
        kc = abi.int_array(list(key_columns))
        oa = c_array(list(outputs), abi.VrOutColumn)
        po, bo = abi.int_array(list(probe_out)), abi.int_array(list(build_out))
        pairs = join_type in (abi.INNER, abi.OUTER)
        out = abi.VrProbeOutput(abi.MEM_DEVICE, int(capacity), _ptr(probe_idx), _ptr(build_idx) if pairs else None,
                                len(probe_out), C.cast(po, C.POINTER(C.c_int32)), len(build_out),
                                C.cast(bo, C.POINTER(C.c_int32)), C.cast(oa, C.POINTER(abi.VrOutColumn)), 0)
        if residual is not None:
            out.residual = C.addressof(residual)
        check(_lib.vr_join_probe_noalloc(self.h, C.cast(ca, _P), len(columns), C.cast(kc, _P), int(n), _ptr(selection),
                                         join_type, 1 if drop_null_keys else 0, C.byref(out), _stream(stream)))
        return out.rows

    def probe_batches(self, batches, key_columns, join_type, outputs=(), probe_out=(), build_out=(), capacity=None,
                      probe_idx=None, build_idx=None, drop_null_keys=False, residual=None, stream=None):
        """vr_join_probe_batches over a ProbeBatches set into device outputs (torch tensors):
        returns (total rows, rows of each batch)."""
        oa = c_array(list(outputs), abi.VrOutColumn)
        po, bo = abi.int_array(list(probe_out)), abi.int_array(list(build_out))
        kc = abi.int_array(list(key_columns))
        pairs = join_type in (abi.INNER, abi.OUTER)
        out = abi.VrProbeOutput(abi.MEM_DEVICE, int(capacity), _ptr(probe_idx), _ptr(build_idx) if pairs else None,
                                len(probe_out), C.cast(po, C.POINTER(C.c_int32)), len(build_out),
                                C.cast(bo, C.POINTER(C.c_int32)), C.cast(oa, C.POINTER(abi.VrOutColumn)), 0)
        if residual is not None:
            out.residual = C.addressof(residual)
        rows = np.zeros(max(1, batches.count), np.int64)
        check(_lib.vr_join_probe_batches(self.h, batches.count, C.cast(batches.cols, _P), batches.ncols,
                                         batches.lengths.ctypes.data, C.cast(batches.sels, _P) if batches.has_sel else None,
                                         C.cast(kc, _P), join_type, 1 if drop_null_keys else 0, C.byref(out),
                                         rows.ctypes.data, _stream(stream)))
        return out.rows, rows[:batches.count]

    def unmatched_build(self, probe_types=(), build_out=(), capacity=None, stream=None):
        """vr_join_unmatched_build into host buffers: (probe_idx (-1), build_idx, output columns, rows)."""
        if capacity is None:
            capacity = max(1, self.info()[0])
        pidx = np.zeros(capacity, np.int32)
        bidx = np.zeros(capacity, np.int32)
        types = list(probe_types) + [self.spec.column_types[c] for c in build_out]
        outputs = [OutColumn(t, capacity) for t in types]
        oa = c_array(outputs, abi.VrOutColumn)
        bo = abi.int_array(list(build_out))
        out = abi.VrProbeOutput(abi.MEM_HOST, int(capacity), _ptr(pidx), _ptr(bidx), len(probe_types), None,
                                len(build_out), C.cast(bo, C.POINTER(C.c_int32)), C.cast(oa, C.POINTER(abi.VrOutColumn)), 0)
        check(_lib.vr_join_unmatched_build(self.h, C.byref(out), _stream(stream)))
        r = out.rows
        return pidx[:r], bidx[:r], outputs, r


class AggGraph:
    """A captured batched update of a GroupedAggregate (vr_agg_graph_*); keeps its batches alive."""

    def __init__(self, agg, h, batches):
        self.agg, self.h, self._batches = agg, h, batches

    def launch(self, stream=None):
        check(_lib.vr_agg_graph_launch(self.h, _stream(stream)))
        _settle(self.agg.device.h, stream)

    def close(self):
        if self.h:
            _lib.vr_agg_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ProbeBatches:
    """The streamed batches of one vr_join_probe_batches call, marshalled once: `batches` is a list
    of (columns, n) or (columns, n, selection) with device-resident columns, every batch with the
    same column lanes."""

    def __init__(self, batches):
        self.count = len(batches)
        self.ncols = len(batches[0][0]) if batches else 0
        self._keep = [b[0] for b in batches]
        flat = [c for b in batches for c in b[0]]
        self.cols = c_array(flat)
        self.lengths = np.array([int(b[1]) for b in batches] or [0], np.int64)
        sels = [b[2] if len(b) > 2 else None for b in batches]
        self.has_sel = any(x is not None for x in sels)
        self._sel_keep = sels
        self.sels = (C.c_void_p * max(1, self.count))(*[_ptr(x) for x in sels])


def agg_key_dictionary(handle, c):
    """The strings of UTF8 key column `c` of an aggregate handle, in id order."""
    count, nbytes = C.c_int64(), C.c_int64()
    check(_lib.vr_agg_key_dictionary(handle, c, None, 0, None, 0, C.byref(count), C.byref(nbytes)))
    offs = np.zeros(count.value + 1, np.int32)
    data = np.zeros(max(1, nbytes.value), np.uint8)
    check(_lib.vr_agg_key_dictionary(handle, c, offs.ctypes.data, len(offs), data.ctypes.data, len(data),
                                     C.byref(count), C.byref(nbytes)))
    raw = data.tobytes()
    return [raw[offs[i]:offs[i + 1]].decode("utf-8") for i in range(count.value)]


def device_column(type, data, validity=None, length=None, dictionary=None, offsets=None):
    """An input column over device tensors (torch); `dictionary` is a host Column; a plain UTF8
    column passes its bytes as `data` and its int32 `offsets`."""
    from .columns import Column
    return Column(type, data, validity, offsets=offsets, dictionary=dictionary, length=length, mem=abi.MEM_DEVICE)


def device_out_column(type, capacity, data, validity=None):
    """An output column over device tensors (torch)."""
    oc = OutColumn.__new__(OutColumn)
    oc.type = type
    oc.capacity = int(capacity)
    oc.mem = abi.MEM_DEVICE
    oc.validity = validity
    oc.data = data
    oc._c = abi.VrOutColumn(type, abi.MEM_DEVICE, int(capacity), _ptr(validity), _ptr(data), None)
    return oc


def filter_compare(device, col, op, value, n, selection=None, out=None, stream=None):
    """vr_filter_compare over a device column: (selection bitmap as a uint8 device tensor padded to
    8 bytes, selected rows)."""
    import torch
    if out is None:  # every word is written
        out = torch.empty(((n + 63) // 64) * 8, dtype=torch.uint8, device="cuda")
    sc = abi.scalar(value, col.type)
    cnt = C.c_int64()
    check(_lib.vr_filter_compare(device.h, C.byref(col.c()), int(op), C.byref(sc), _ptr(selection), int(n), _ptr(out),
                                 C.byref(cnt), _stream(stream)))
    _settle(device.h, stream)
    return out, cnt.value


def compact_utf8_bytes(device, col, n, selection, stream=None):
    b = C.c_int64()
    check(_lib.vr_compact_utf8_bytes(device.h, C.byref(col.c()), int(n), _ptr(selection), C.byref(b), _stream(stream)))
    _settle(device.h, stream)
    return b.value


def compact(device, columns, n, selection, outputs, stream=None):
    """vr_compact of device columns into device output columns (device_out_column); returns the rows."""
    ca, oa = c_array(columns), c_array(outputs, abi.VrOutColumn)
    rows = C.c_int64()
    check(_lib.vr_compact(device.h, C.cast(ca, _P), len(columns), int(n), _ptr(selection), C.cast(oa, _P), C.byref(rows),
                          _stream(stream)))
    _settle(device.h, stream)
    return rows.value


def gather_utf8_bytes(device, col, idx, m, stream=None):
    b = C.c_int64()
    check(_lib.vr_gather_utf8_bytes(device.h, C.byref(col.c()), _ptr(idx), int(m), C.byref(b), _stream(stream)))
    _settle(device.h, stream)
    return b.value


def gather(device, col, idx, m, out, stream=None):
    """vr_gather of a device column by a device int32 index tensor into a device output column."""
    check(_lib.vr_gather(device.h, C.byref(col.c()), _ptr(idx), int(m), C.byref(out.c()), _stream(stream)))
    _settle(device.h, stream)


def hash_partition_ids(device, keys, kinds, n, num_partitions, seed=abi.SPARK_SEED, want_hashes=False, stream=None):
    """vr_hash_partition_ids over device key columns: the partition ids (int32 device tensor) and,
    when asked, the hashes -- (ids, hashes or None)."""
    import torch
    ids = torch.empty(max(1, n), dtype=torch.int32, device="cuda")
    hashes = torch.empty(max(1, n), dtype=torch.int32, device="cuda") if want_hashes else None
    ka = c_array(keys) if keys else None
    kd = (C.c_int32 * max(1, len(kinds)))(*kinds)
    check(_lib.vr_hash_partition_ids(device.h, C.cast(ka, _P) if ka is not None else None, C.cast(kd, _P), len(keys),
                                     int(n), int(num_partitions), int(seed), _ptr(ids), _ptr(hashes), _stream(stream)))
    _settle(device.h, stream)
    return ids[:n], (hashes[:n] if hashes is not None else None)


def round_robin_ids(device, n, num_partitions, start, stream=None):
    """vr_round_robin_ids: (ids as an int32 device tensor, the next batch's start)."""
    import torch
    ids = torch.empty(max(1, n), dtype=torch.int32, device="cuda")
    nxt = C.c_int32()
    check(_lib.vr_round_robin_ids(device.h, int(n), int(num_partitions), int(start), _ptr(ids), C.byref(nxt),
                                  _stream(stream)))
    _settle(device.h, stream)
    return ids[:n], nxt.value


def partition_rows(device, ids, n, num_partitions, want_rows=True, want_dest=True, stream=None):
    """vr_partition_rows of a device int32 id tensor: (rows, dest, starts) device tensors, rows or
    dest None when not asked for."""
    import torch
    rows = torch.empty(max(1, n), dtype=torch.int32, device="cuda") if want_rows else None
    dest = torch.empty(max(1, n), dtype=torch.int32, device="cuda") if want_dest else None
    starts = torch.empty(num_partitions + 1, dtype=torch.int64, device="cuda")
    check(_lib.vr_partition_rows(device.h, _ptr(ids), int(n), int(num_partitions), _ptr(rows), _ptr(dest),
                                 _ptr(starts), _stream(stream)))
    _settle(device.h, stream)
    return (rows[:n] if rows is not None else None), (dest[:n] if dest is not None else None), starts
