"""ctypes mirror of include/vr.h: the structs and constants of the C ABI.

Only layout lives here; loading the device library is vecruntime_amd._lib's job.
"""
import ctypes as C

# status codes (vr_status)
VR_OK = 0
VR_ERR_INVALID_ARGUMENT = 1
VR_ERR_UNSUPPORTED = 2
VR_ERR_OOM = 3
VR_ERR_DEVICE = 4
VR_ERR_ARITHMETIC_OVERFLOW = 5
VR_ERR_DECIMAL_OVERFLOW = 6
VR_ERR_CAPACITY = 7
VR_ERR_STATE = 8

STATUS_NAMES = {
    VR_OK: "OK",
    VR_ERR_INVALID_ARGUMENT: "INVALID_ARGUMENT",
    VR_ERR_UNSUPPORTED: "UNSUPPORTED",
    VR_ERR_OOM: "OOM",
    VR_ERR_DEVICE: "DEVICE",
    VR_ERR_ARITHMETIC_OVERFLOW: "ARITHMETIC_OVERFLOW",
    VR_ERR_DECIMAL_OVERFLOW: "DECIMAL_OVERFLOW",
    VR_ERR_CAPACITY: "CAPACITY",
    VR_ERR_STATE: "STATE",
}

# vr_type: the reference's VecType order (BOOL, INT32, INT64, FLOAT64, UTF8, DECIMAL128)
BOOL, INT32, INT64, FLOAT64, UTF8, DECIMAL128 = range(6)
TYPE_NAMES = ["BOOL", "INT32", "INT64", "FLOAT64", "UTF8", "DECIMAL128"]
BYTE_WIDTH = {BOOL: 0, INT32: 4, INT64: 8, FLOAT64: 8, UTF8: 4, DECIMAL128: 16}

MEM_HOST, MEM_DEVICE = 0, 1

# vr_agg_kind
COUNT_STAR, COUNT, SUM_LONG, SUM_DOUBLE, SUM_WIDE, AVG_DOUBLE, AVG_WIDE, MIN, MAX = range(9)
MERGE_COUNT, MERGE_SUM_LONG, MERGE_SUM_DOUBLE, MERGE_SUM_WIDE = 16, 17, 18, 19
MERGE_AVG_DOUBLE, MERGE_AVG_WIDE, MERGE_MIN, MERGE_MAX = 20, 21, 22, 23

# vr_agg_flags
CHECKED = 1
RESULT = 2

# vr_join_type
INNER, LEFT_SEMI, LEFT_ANTI, OUTER = 0, 1, 2, 3
CMP_EQ, CMP_NE, CMP_LT, CMP_LE, CMP_GT, CMP_GE = 0, 1, 2, 3, 4, 5

# vr_hash_kind: how a shuffle key is hashed (PartitionKernels.KeyKind)
HASH_INT, HASH_LONG, HASH_DOUBLE, HASH_BOOL, HASH_UTF8, HASH_DECIMAL128, HASH_FLOAT = 0, 1, 2, 3, 4, 5, 6
SPARK_SEED = 42


class VrColumn(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("mem", C.c_int32),
        ("length", C.c_int64),
        ("validity", C.c_void_p),
        ("data", C.c_void_p),
        ("offsets", C.c_void_p),
        ("dictionary", C.c_void_p),
    ]


class VrOutColumn(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("mem", C.c_int32),
        ("capacity", C.c_int64),
        ("validity", C.c_void_p),
        ("data", C.c_void_p),
        ("offsets", C.c_void_p),
    ]


class VrAggDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("input", C.c_int32),
        ("input2", C.c_int32),
        ("flags", C.c_int32),
        ("buffer_precision", C.c_int32),
        ("buffer_scale", C.c_int32),
        ("result_precision", C.c_int32),
        ("result_scale", C.c_int32),
    ]


class VrAggSpec(C.Structure):
    _fields_ = [
        ("num_keys", C.c_int32),
        ("key_types", C.POINTER(C.c_int32)),
        ("key_nullable", C.POINTER(C.c_uint8)),
        ("num_values", C.c_int32),
        ("value_types", C.POINTER(C.c_int32)),
        ("num_aggs", C.c_int32),
        ("aggs", C.POINTER(VrAggDesc)),
        ("capacity_hint", C.c_int64),
        ("deterministic", C.c_int32),
        ("max_load_percent", C.c_int32),
    ]


class VrStats(C.Structure):
    _fields_ = [
        ("calls", C.c_int64),
        ("input_rows", C.c_int64),
        ("output_rows", C.c_int64),
        ("h2d_bytes", C.c_int64),
        ("d2h_bytes", C.c_int64),
        ("kernel_launches", C.c_int64),
        ("table_grows", C.c_int64),
        ("cpu_fallbacks_signalled", C.c_int64),
        ("partitioned_updates", C.c_int64),
        ("specialized_updates", C.c_int64),
        ("dense_join_builds", C.c_int64),
        ("device_allocations", C.c_int64),
        ("payload_join_probes", C.c_int64),
    ]


class VrBatchSet(C.Structure):
    _fields_ = [
        ("num_batches", C.c_int32),
        ("keys", C.POINTER(VrColumn)),
        ("values", C.POINTER(VrColumn)),
        ("lengths", C.POINTER(C.c_int64)),
        ("filter_col", C.c_int32),
        ("filter_threshold", C.c_int32),
    ]


class VrLaunchConfig(C.Structure):
    _fields_ = [
        ("block_threads", C.c_int32),
        ("rows_per_thread", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("lds_preaggregate", C.c_int32),
        ("lds_entries", C.c_int32),
        ("flush_tiles", C.c_int32),
        ("partitioned", C.c_int32),
        ("specialize", C.c_int32),
    ]


def launch_config(rows_per_thread=0, blocks_per_cu=0, lds=-1, lds_entries=0, flush_tiles=0, partitioned=0,
                  specialize=0):
    """partitioned: 1 takes the partitioned update at any size, -1 never, 0 from 4M rows per call.
    specialize: -1 runs the partitioned update's generic kernels instead of ones compiled for the spec."""
    return VrLaunchConfig(0, rows_per_thread, blocks_per_cu, lds, lds_entries, flush_tiles, partitioned, specialize)


class VrScalar(C.Structure):
    _fields_ = [("i64", C.c_int64), ("f64", C.c_double), ("lo", C.c_int64), ("hi", C.c_int64)]


def scalar(value, type):
    """A comparison literal for a lane of `type`: ints for INT32/INT64, float for FLOAT64, the
    unscaled Python int for DECIMAL128."""
    if type == DECIMAL128:
        x = int(value) & ((1 << 128) - 1)
        lo, hi = x & ((1 << 64) - 1), x >> 64
        return VrScalar(0, 0.0, lo - (1 << 64) if lo >= 1 << 63 else lo, hi - (1 << 64) if hi >= 1 << 63 else hi)
    if type == FLOAT64:
        return VrScalar(0, float(value), 0, 0)
    x = int(value) & ((1 << 64) - 1)
    return VrScalar(x - (1 << 64) if x >= 1 << 63 else x, 0.0, 0, 0)


class VrJoinSpec(C.Structure):
    _fields_ = [
        ("num_keys", C.c_int32),
        ("key_types", C.POINTER(C.c_int32)),
        ("num_columns", C.c_int32),
        ("column_types", C.POINTER(C.c_int32)),
        ("key_columns", C.POINTER(C.c_int32)),
        ("capacity_hint", C.c_int64),
        ("dense_index", C.c_int32),
        ("track_matched", C.c_int32),
    ]


class VrProbeOutput(C.Structure):
    _fields_ = [
        ("mem", C.c_int32),
        ("capacity", C.c_int64),
        ("probe_idx", C.c_void_p),
        ("build_idx", C.c_void_p),
        ("num_probe_out", C.c_int32),
        ("probe_out_columns", C.POINTER(C.c_int32)),
        ("num_build_out", C.c_int32),
        ("build_out_columns", C.POINTER(C.c_int32)),
        ("columns", C.POINTER(VrOutColumn)),
        ("rows", C.c_int64),
        ("residual", C.c_void_p),
    ]


class VrJoinResidual(C.Structure):
    _fields_ = [("op", C.c_int32), ("left_side", C.c_int32), ("left_column", C.c_int32), ("right_side", C.c_int32),
                ("right_column", C.c_int32), ("left_offset", C.c_int64), ("right_offset", C.c_int64)]


PROBE_SIDE, BUILD_SIDE = 0, 1


def residual(op, left, right, left_offset=0, right_offset=0):
    """left / right: (side, column) -- PROBE_SIDE with a streamed column index, BUILD_SIDE with a
    build column index."""
    return VrJoinResidual(op, left[0], left[1], right[0], right[1], left_offset, right_offset)


# vr_query_kind (include/vr_serve.h)
QUERY_AGG, QUERY_JOIN = 0, 1
MAX_RESULT_COLUMNS = 40


ADMIT_SIZE_AWARE, ADMIT_FIFO = 0, 1


class VrServerConfig(C.Structure):
    _fields_ = [
        ("workers", C.c_int32),
        ("admission", C.c_int32),
        ("hbm_budget_bytes", C.c_int64),
        ("aging_ms", C.c_int64),
    ]


class VrQuery(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("join_type", C.c_int32),
        ("agg", C.POINTER(VrAggSpec)),
        ("batches", C.POINTER(VrBatchSet)),
        ("expected_groups", C.c_int64),
        ("join", C.POINTER(VrJoinSpec)),
        ("build_columns", C.POINTER(VrColumn)),
        ("build_rows", C.c_int64),
        ("probe_columns", C.POINTER(VrColumn)),
        ("num_probe_columns", C.c_int32),
        ("num_probe_out", C.c_int32),
        ("probe_key_columns", C.POINTER(C.c_int32)),
        ("probe_out_columns", C.POINTER(C.c_int32)),
        ("probe_rows", C.c_int64),
        ("num_build_out", C.c_int32),
        ("reserved", C.c_int32),
        ("build_out_columns", C.POINTER(C.c_int32)),
        ("expected_rows", C.c_int64),
    ]


class VrQueryResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("num_columns", C.c_int32),
        ("rows", C.c_int64),
        ("columns", VrOutColumn * MAX_RESULT_COLUMNS),
        ("queue_ms", C.c_double),
        ("run_ms", C.c_double),
        ("latency_ms", C.c_double),
        ("device_bytes", C.c_int64),
        ("agg", C.c_void_p),
    ]


class VrServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64),
        ("completed", C.c_int64),
        ("failed", C.c_int64),
        ("peak_inflight", C.c_int64),
        ("peak_reserved_bytes", C.c_int64),
        ("reserved_bytes", C.c_int64),
    ]


def int_array(values):
    """A ctypes int32 array (kept alive by the caller)."""
    arr = (C.c_int32 * max(1, len(values)))(*values)
    return arr


def agg_desc(kind, input=-1, input2=-1, flags=0, buffer_precision=0, buffer_scale=0,
             result_precision=0, result_scale=0):
    return VrAggDesc(kind, input, input2, flags, buffer_precision, buffer_scale,
                     result_precision, result_scale)


class AggSpec:
    """Owns a VrAggSpec and the arrays it points at."""

    def __init__(self, key_types, value_types, aggs, key_nullable=None, capacity_hint=0,
                 deterministic=False, max_load_percent=0):
        self.key_types = list(key_types)
        self.value_types = list(value_types)
        self.aggs = list(aggs)
        self._kt = int_array(self.key_types)
        self._vt = int_array(self.value_types)
        self._aggs = (VrAggDesc * max(1, len(self.aggs)))(*self.aggs)
        if key_nullable is not None:
            self._kn = (C.c_uint8 * max(1, len(key_nullable)))(*[1 if x else 0 for x in key_nullable])
            kn = C.cast(self._kn, C.POINTER(C.c_uint8))
        else:
            kn = C.POINTER(C.c_uint8)()
        self.c = VrAggSpec(len(self.key_types), C.cast(self._kt, C.POINTER(C.c_int32)), kn,
                           len(self.value_types), C.cast(self._vt, C.POINTER(C.c_int32)),
                           len(self.aggs), C.cast(self._aggs, C.POINTER(VrAggDesc)),
                           int(capacity_hint), 1 if deterministic else 0, int(max_load_percent))

    def output_types(self):
        """The lanes an emission writes, aggregate by aggregate (Spark's buffer or result)."""
        out = []
        for a in self.aggs:
            result = bool(a.flags & RESULT)
            k = a.kind
            in_type = self.value_types[a.input] if a.input >= 0 else None
            if k in (COUNT_STAR, COUNT, MERGE_COUNT, SUM_LONG, MERGE_SUM_LONG):
                out.append(INT64)
            elif k in (SUM_DOUBLE, MERGE_SUM_DOUBLE):
                out.append(FLOAT64)
            elif k in (AVG_DOUBLE, MERGE_AVG_DOUBLE):
                out += [FLOAT64] if result else [FLOAT64, INT64]
            elif k in (SUM_WIDE, MERGE_SUM_WIDE):
                out += [DECIMAL128] if result else [DECIMAL128, BOOL]
            elif k in (AVG_WIDE, MERGE_AVG_WIDE):
                if result:
                    out.append(DECIMAL128 if a.result_precision > 18 else INT64)
                else:
                    out += [DECIMAL128, INT64]
            elif k in (MIN, MAX, MERGE_MIN, MERGE_MAX):
                out.append(in_type)
            else:
                raise ValueError("aggregate kind %d" % k)
        return out


class JoinSpec:
    """Owns a VrJoinSpec and the arrays it points at."""

    def __init__(self, key_types, column_types, key_columns, capacity_hint=0, dense_index=0, track_matched=False):
        self.key_types = list(key_types)
        self.column_types = list(column_types)
        self.key_columns = list(key_columns)
        self._kt = int_array(self.key_types)
        self._ct = int_array(self.column_types)
        self._kc = int_array(self.key_columns)
        self.c = VrJoinSpec(len(self.key_types), C.cast(self._kt, C.POINTER(C.c_int32)),
                            len(self.column_types), C.cast(self._ct, C.POINTER(C.c_int32)),
                            C.cast(self._kc, C.POINTER(C.c_int32)), int(capacity_hint), int(dense_index),
                            1 if track_matched else 0)
