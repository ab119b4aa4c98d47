/*
 * vr.h -- C ABI of libvrgpu.so, the MI355X (gfx950) hash-aggregation and
 * hash-join library that sits under VecRuntime's operator API.
 *
 * Every entry point returns a vr_status (0 = VR_OK); on failure the message is
 * available from vr_last_error() on the calling thread.  No torch or HIP type
 * appears in a signature: handles are opaque, buffers are plain pointers with
 * sizes, and a stream is passed as `void*` (a hipStream_t, or NULL for the
 * handle's own stream).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   vr_agg_*   GroupKeyTable.assign + GroupedAccumulators.*.update fused
 *              (kernels/src/main/java/io/vecruntime/kernels/GroupKeyTable.java:709,
 *               GroupedAccumulators.java:224,380,652,1017,1212,1287), driven by
 *              VectorGroupedAggregateIterator.consume
 *              (spark/src/main/scala/org/apache/spark/sql/vecruntime/VectorHashAggregateExec.scala:584-602);
 *              emission replaces GroupKeyTable.writeKeys (GroupKeyTable.java:2851)
 *              and GroupedAggState.writeBuffer (spark/.../agg/VectorAggregates.scala:63).
 *   vr_join_*  BuildTable.build (VectorHashJoinExec.scala:478-516), GroupKeyTable.lookup
 *              (GroupKeyTable.java:772), VectorHashJoinIterator.emitMatches /
 *              emitSemiAnti / flush (VectorHashJoinExec.scala:1116-1182,1605).
 */
#ifndef VR_H
#define VR_H

#ifndef __HIPCC_RTC__ /* hipRTC provides the fixed-width types itself (vr_types.h) */
#include <stddef.h>
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */

typedef int32_t vr_status;

enum {
    VR_OK = 0,
    /* A bad argument (IllegalArgumentException in the reference). */
    VR_ERR_INVALID_ARGUMENT = 1,
    /* A lane or aggregate the device path does not take (IllegalStateException,
     * GroupKeyTable.java:638).  The caller runs the CPU kernels instead; no new
     * planning-time fallback reason is introduced. */
    VR_ERR_UNSUPPORTED = 2,
    /* Device memory (or the HBM budget) exhausted: the caller re-runs the batch on the CPU kernels. */
    VR_ERR_OOM = 3,
    /* A HIP runtime error. */
    VR_ERR_DEVICE = 4,
    /* A checked (ANSI) bigint SUM left the long range: Spark's ARITHMETIC_OVERFLOW
     * (VectorAggregates.scala:142-170). */
    VR_ERR_ARITHMETIC_OVERFLOW = 5,
    /* An ANSI decimal result past its precision (NUMERIC_VALUE_OUT_OF_RANGE,
     * VectorAggregates.scala:329-349). */
    VR_ERR_DECIMAL_OVERFLOW = 6,
    /* A caller-allocated output is too small; the size needed is reported. */
    VR_ERR_CAPACITY = 7,
    /* A call out of order (e.g. a probe before vr_join_build_finish). */
    VR_ERR_STATE = 8
};

/* Message of the last failed call on this thread ("" when none). */
const char* vr_last_error(void);

/* ------------------------------------------------------------------ columns */

/* Physical lanes, in the order of the reference's VecType enum (VecType.java:23-91). */
enum vr_type {
    VR_BOOL = 0,       /* bitmap, one bit per row, LSB first */
    VR_INT32 = 1,
    VR_INT64 = 2,
    VR_FLOAT64 = 3,
    VR_UTF8 = 4,       /* int32 offsets[n+1] + bytes, or dictionary-encoded int32 codes */
    VR_DECIMAL128 = 5  /* 16 bytes per row: low limb at 16i, high limb at 16i+8, little endian */
};

/* Where a column's buffers live. */
enum vr_mem { VR_MEM_HOST = 0, VR_MEM_DEVICE = 1 };

/* An Arrow-layout input column (the reference's VectorBuffers, VectorBuffers.java:48-215).
 * `validity` is NULL when no row is null.  A plain UTF8 column has `offsets` (n + 1 int32) and its
 * bytes in `data`, both in the column's memory; for a dictionary-encoded UTF8 column `data` holds
 * int32 codes and `dictionary` points at a plain UTF8 column in host memory. */
typedef struct vr_column {
    int32_t type;
    int32_t mem;
    int64_t length;
    const uint8_t* validity;
    const void* data;
    const int32_t* offsets;
    const struct vr_column* dictionary;
} vr_column;

/* A caller-allocated Arrow-layout output column.  `validity` holds ceil(capacity/8) bytes padded
 * to a multiple of 8 (Arrow pads buffers to 64 bytes; bitmaps are written in whole 8-byte words);
 * fixed-width `data` holds capacity * width bytes (a bitmap for VR_BOOL).  A NULL `validity` means
 * the caller needs no validity for the column. */
typedef struct vr_out_column {
    int32_t type;
    int32_t mem;
    int64_t capacity;
    uint8_t* validity;
    void* data;
    int32_t* offsets;
} vr_out_column;

/* ------------------------------------------------------------------ device */

typedef struct vr_device vr_device;

/* Opens device `ordinal` (one handle per process and device is typical; the handle is
 * thread-safe).  `hbm_budget_bytes` = 0 means no limit beyond the device's memory. */
vr_status vr_device_open(int32_t ordinal, int64_t hbm_budget_bytes, vr_device** out);
vr_status vr_device_close(vr_device* dev);
/* Bytes currently allocated by handles of this device. */
vr_status vr_device_memory(vr_device* dev, int64_t* allocated, int64_t* budget);
/* Waits for all work the library queued on `stream` (NULL: every stream of the device). */
vr_status vr_device_sync(vr_device* dev, void* stream);

/* Per-device counters (the reference's SQLMetrics numInputBatches/numOutputRows/time,
 * VectorExec.scala:31-45, plus device-side figures). */
typedef struct vr_stats {
    int64_t calls;
    int64_t input_rows;
    int64_t output_rows;
    int64_t h2d_bytes;
    int64_t d2h_bytes;
    int64_t kernel_launches;
    int64_t table_grows;
    int64_t cpu_fallbacks_signalled;
    int64_t partitioned_updates;  /* aggregate calls that took the partitioned update */
    int64_t specialized_updates;  /* of those, calls that ran kernels compiled for their spec */
    int64_t dense_join_builds;    /* join builds that took the direct key index */
    int64_t device_allocations;   /* device blocks the library handed out (new or from its cache) */
    int64_t payload_join_probes;  /* probe calls that read build columns from the payload table's cells */
} vr_stats;
vr_status vr_stats_get(vr_device* dev, vr_stats* out);

/* ------------------------------------------------------------------ aggregation */

/* Aggregate functions in update mode (Partial / Complete), mapped from Spark's functions as
 * VectorAggregates.compileFunction does (VectorAggregates.scala:1327-1460). */
enum vr_agg_kind {
    VR_AGG_COUNT_STAR = 0,   /* COUNT(*): every selected row (GroupedAccumulators.Count.updateAll) */
    VR_AGG_COUNT = 1,        /* COUNT(x): non-null rows (Count.updateNonNull) */
    VR_AGG_SUM_LONG = 2,     /* SUM over INT32/INT64 into bigint (SumLongAgg); VR_AGG_CHECKED = ANSI */
    VR_AGG_SUM_DOUBLE = 3,   /* SUM over FLOAT64 (SumDoubleAgg) */
    VR_AGG_SUM_WIDE = 4,     /* SUM over a decimal lane into Decimal(p+10,s) (WideDecimalSumAgg) */
    VR_AGG_AVG_DOUBLE = 5,   /* AVG over FLOAT64, buffer (sum, count) (AverageAgg) */
    VR_AGG_AVG_WIDE = 6,     /* AVG over a decimal lane, buffer (sum, count) (WideDecimalAvgAgg) */
    VR_AGG_MIN = 7,          /* MIN over INT32/INT64/FLOAT64/DECIMAL128 (MinMaxAgg) */
    VR_AGG_MAX = 8,          /* MAX over INT32/INT64/FLOAT64/DECIMAL128 (MinMaxAgg) */
    /* Merge modes (PartialMerge / Final): inputs are partial buffer columns
     * (VectorAggregates.compileMerge, VectorAggregates.scala:1141). */
    VR_AGG_MERGE_COUNT = 16,      /* input: count bigint */
    VR_AGG_MERGE_SUM_LONG = 17,   /* input: sum bigint (nullable) */
    VR_AGG_MERGE_SUM_DOUBLE = 18, /* input: sum double (nullable) */
    VR_AGG_MERGE_SUM_WIDE = 19,   /* inputs: sum DECIMAL128 (null = overflowed), isEmpty BOOL */
    VR_AGG_MERGE_AVG_DOUBLE = 20, /* inputs: sum double, count bigint */
    VR_AGG_MERGE_AVG_WIDE = 21,   /* inputs: sum DECIMAL128 (null = overflowed), count bigint */
    VR_AGG_MERGE_MIN = 22,
    VR_AGG_MERGE_MAX = 23
};

enum vr_agg_flags {
    /* ANSI evaluation: overflow raises instead of wrapping / nulling. */
    VR_AGG_CHECKED = 1,
    /* Emit the final result in the first slot (Complete / Final) instead of the partial buffer. */
    VR_AGG_RESULT = 2
};

typedef struct vr_agg_desc {
    int32_t kind;          /* vr_agg_kind */
    int32_t input;         /* value column index into the update's `values`, -1 for COUNT(*) */
    int32_t input2;        /* second buffer column of a merge (isEmpty / count), else -1 */
    int32_t flags;         /* vr_agg_flags */
    int32_t buffer_precision; /* decimal sum buffer Decimal(p+10, s): precision; 0 otherwise */
    int32_t buffer_scale;
    int32_t result_precision; /* AVG over a decimal: result Decimal(p+4, s+4) */
    int32_t result_scale;
} vr_agg_desc;

typedef struct vr_agg_spec {
    int32_t num_keys;             /* 0: ungrouped aggregation, one group (one output row even for no
                                     input: AggKernels' reductions); keys are then passed as NULL */
    const int32_t* key_types;     /* vr_type per key; a UTF8 key, plain or dictionary encoded, in host
                                     or device memory, is mapped to the handle's dictionary ids on the
                                     device (ids in the dictionary's own order, not first-seen order) */
    const uint8_t* key_nullable;  /* may be NULL: every key nullable */
    int32_t num_values;
    const int32_t* value_types;   /* vr_type per value column */
    int32_t num_aggs;
    const vr_agg_desc* aggs;
    int64_t capacity_hint;        /* expected groups; 0 = start small and grow */
    /* 1: emit groups in first-seen order and fold double sums in row order, bit-identical to the
     * reference's strict mode (GroupedAccumulators.DoubleSum(strict), GroupedAccumulators.java:202-206).
     * 0: group order unspecified, double sums by device atomics (within 1e-12 relative). */
    int32_t deterministic;
    /* Largest fraction of the table's slots in use before it grows, in percent (0: 50). */
    int32_t max_load_percent;
} vr_agg_spec;

typedef struct vr_agg vr_agg;

vr_status vr_agg_create(vr_device* dev, const vr_agg_spec* spec, vr_agg** out);
/* One batch: n rows of `keys` (num_keys columns) and `values` (num_values columns), rows not set
 * in `selection` (a bitmap in the columns' memory, or NULL) ignored.  Asynchronous on `stream`. */
vr_status vr_agg_update(vr_agg* agg, const vr_column* keys, const vr_column* values, int64_t n,
                        const uint8_t* selection, void* stream);
/* Groups so far (synchronises with the handle's work). */
vr_status vr_agg_num_groups(vr_agg* agg, int64_t* groups);
/* Number of output columns an emission writes per aggregate, in order. */
vr_status vr_agg_output_layout(vr_agg* agg, int32_t* num_columns, int32_t* types, int32_t max_columns);
/* Writes groups [from, to) of the emission order: `keys` (num_keys columns, key lanes; a UTF8 key
 * is written as its table-dictionary id in an INT32 lane, see vr_agg_key_dictionary) and `out`
 * (the columns vr_agg_output_layout lists).  Synchronous with respect to `stream`. */
vr_status vr_agg_emit(vr_agg* agg, int64_t from, int64_t to, vr_out_column* keys,
                      vr_out_column* out, void* stream);
/* The table's dictionary of UTF8 key column `c` (GroupKeyTable.dictionary, GroupKeyTable.java:2804):
 * ids in first-seen order; offsets[count+1] and bytes, copied to host buffers. */
vr_status vr_agg_key_dictionary(vr_agg* agg, int32_t c, int32_t* offsets, int64_t offsets_capacity,
                                uint8_t* bytes, int64_t bytes_capacity, int64_t* count,
                                int64_t* num_bytes);
/* Device bytes held by the handle (GroupKeyTable.memoryBytes, GroupKeyTable.java:2596). */
vr_status vr_agg_memory(vr_agg* agg, int64_t* bytes);
/* Compiles ahead of time the kernels the partitioned update runs for `spec` (hipRTC, seconds per
 * spec, otherwise paid by the spec's first large update).  No counterpart in the reference (its
 * kernels are compiled with the JVM): a start-up aid for a server that knows the specs it will
 * see.  Safe from several threads at once; VR_OK also when the spec takes no compiled kernels. */
vr_status vr_agg_prewarm(vr_device* dev, const vr_agg_spec* spec);
/* Drops every group, keeping the allocation (emit-and-reset, VectorHashAggregateExec.scala:193-246). */
vr_status vr_agg_reset(vr_agg* agg);
vr_status vr_agg_destroy(vr_agg* agg);

/* ------------------------------------------------------------------ batched aggregation */

/* Many batches (typically from concurrent partitions) folded into one handle with a single
 * persistent launch.  Batch b's columns are keys[b*num_keys ...] and values[b*num_values ...],
 * its length lengths[b]; a batch's optional filter is `filter_col` (an INT32 column index into
 * its values) with predicate value < filter_threshold (filter_col < 0: none). */
typedef struct vr_batch_set {
    int32_t num_batches;
    const vr_column* keys;
    const vr_column* values;
    const int64_t* lengths;
    int32_t filter_col;
    int32_t filter_threshold;
} vr_batch_set;

/* Launch shape of the update kernel; zeros pick the defaults. */
typedef struct vr_launch_config {
    int32_t block_threads;      /* fixed at 256 (the kernel's launch bounds); reserved */
    int32_t rows_per_thread;    /* rows per thread per tile */
    int32_t blocks_per_cu;      /* workgroups per CU in the persistent grid */
    int32_t lds_preaggregate;   /* 1: the LDS hot-group cache, 0: off, -1: default (on) */
    int32_t lds_entries;        /* hot-group cache entries (a power of two, <= 2048) */
    int32_t flush_tiles;        /* tiles between flushes of the cache */
    int32_t partitioned;        /* keys of at most 64 bits: 1 takes the partitioned update (hot-set
                                   scatter, then one merge workgroup per table region) whatever the
                                   input size, -1 never, 0 from 4M rows per call */
    int32_t specialize;         /* partitioned update: 0 or 1 runs kernels compiled at run time for
                                   the spec's aggregates (hipRTC, once per spec and device, about
                                   4 s; the generic kernels if that fails), -1 the generic kernels */
} vr_launch_config;

/* Sets the handle's default launch shape (what vr_agg_update uses, and what
 * vr_agg_update_batches uses for fields its own config leaves at zero). */
vr_status vr_agg_set_launch_config(vr_agg* agg, const vr_launch_config* cfg);

vr_status vr_agg_update_batches(vr_agg* agg, const vr_batch_set* batches,
                                const vr_launch_config* cfg, void* stream);
/* A captured batched update (no counterpart in the reference, whose per-batch loop is
 * VectorGroupedAggregateIterator.consume, VectorHashAggregateExec.scala:584-602): one HIP graph
 * of a whole step over a fixed set of device-resident batches -- the table cleared, every batch
 * folded by one persistent launch (the set's filter fused), the NaN-payload pass -- replayed by
 * vr_agg_graph_launch with no host round trip and no allocation.  The caller refills the batches'
 * buffers between replays as it likes; after a replay, vr_agg_num_groups / vr_agg_emit read the
 * step's groups as after any update.  The table is sized for `max_groups` groups when the graph
 * is captured and nothing grows inside it: a step with more groups than the table holds sets the
 * table-full error, which the next read reports as VR_ERR_OOM.  A replay after the handle's table
 * moved (an update that grew it) returns VR_ERR_STATE: capture again.  A graph may outlive its
 * handle: its launches then return VR_ERR_STATE and it is still destroyed with
 * vr_agg_graph_destroy.  Not for deterministic specs, 128-bit MIN/MAX or ungrouped aggregates
 * (VR_ERR_UNSUPPORTED); UTF8 keys need one dictionary per key over the set. */
typedef struct vr_agg_graph vr_agg_graph;
vr_status vr_agg_graph_create(vr_agg* agg, const vr_batch_set* batches, const vr_launch_config* cfg,
                              int64_t max_groups, vr_agg_graph** out);
vr_status vr_agg_graph_launch(vr_agg_graph* graph, void* stream);
vr_status vr_agg_graph_destroy(vr_agg_graph* graph);
/* Measurement aid (no reference counterpart): while timing is on, every update-kernel launch of the
 * handle is bracketed by HIP events on its stream; vr_agg_kernel_time waits for them and returns
 * the summed device time and the launch count since timing was switched on or last read. */
vr_status vr_agg_set_timing(vr_agg* agg, int32_t enable);
vr_status vr_agg_kernel_time(vr_agg* agg, double* ms, int64_t* launches);
/* The same with a filter fused in front of a plain per-batch update (one launch per batch):
 * the baseline the persistent launch is measured against. */
vr_status vr_agg_update_filtered(vr_agg* agg, const vr_column* keys, const vr_column* values,
                                 int64_t n, int32_t filter_col, int32_t filter_threshold,
                                 void* stream);

/* ------------------------------------------------------------------ join */

enum vr_join_type {
    VR_JOIN_INNER = 0,
    VR_JOIN_LEFT_SEMI = 1,
    VR_JOIN_LEFT_ANTI = 2,
    /* The streamed side preserved (LeftOuter built on the right, RightOuter built on the left,
     * FullOuter): the inner pairs, and (row, -1) with the build columns null for every selected
     * streamed row that pairs with nothing (VectorHashJoinExec.scala:999-1009,1160-1175). */
    VR_JOIN_OUTER = 3
};

typedef struct vr_join_spec {
    int32_t num_keys;
    const int32_t* key_types;      /* 1-4 keys of INT32 / INT64 / FLOAT64 (bit for bit) / DECIMAL128 /
                                      UTF8 (plain or dictionary encoded: ids of the build side's device
                                      dictionary; a dictionary with null entries is refused); more than
                                      64 bits of key lanes are looked up by a 64-bit hash of the lanes
                                      and compared lane by lane.  A UTF8 column can only be a key. */
    int32_t num_columns;           /* build-side columns kept for the output (keys included) */
    const int32_t* column_types;
    const int32_t* key_columns;    /* index into the build columns of each key */
    int64_t capacity_hint;         /* expected build rows */
    /* 0: a direct index over the key range when the build allows it -- one INT32/INT64 key whose
     * non-null values span at most 2^22 and at most 10x the distinct keys (DenseKeyIndex.tryBuild,
     * kernels/src/main/java/io/vecruntime/kernels/DenseKeyIndex.java:70-110); -1: always the hash
     * table.  The probe's results are the same either way. */
    int32_t dense_index;
    /* 1: remember every build row an INNER or OUTER probe pairs, for vr_join_unmatched_build (an
     * outer join that preserves the build side, VectorHashJoinExec.scala:1004-1013). */
    int32_t track_matched;
} vr_join_spec;

typedef struct vr_join vr_join;

vr_status vr_join_create(vr_device* dev, const vr_join_spec* spec, vr_join** out);
/* Appends a build batch (ColumnBuilder.append in BuildTable.fromBatches, VectorHashJoinExec.scala:653-680). */
vr_status vr_join_build_append(vr_join* join, const vr_column* columns, int64_t n, void* stream);
/* Builds the table over every appended row (BuildTable.build): rows with a null key never match. */
vr_status vr_join_build_finish(vr_join* join, void* stream);
/* Build rows, distinct non-null keys, longest chain. */
vr_status vr_join_info(vr_join* join, int64_t* build_rows, int64_t* distinct_keys,
                       int64_t* max_rows_per_key);

/* Output of a probe.  For VR_JOIN_INNER either `pairs` (probe_idx/build_idx int32 arrays, the
 * reference's probeIdx/buildIdx) or materialized columns (probe_out_columns of the streamed batch
 * then build_out_columns of the build side, in that order), or both.  For semi/anti: `pairs`'
 * probe_idx lists the kept streamed rows and/or the streamed columns are compacted. */
/* Comparison operators (CompareOp, kernels/src/main/java/io/vecruntime/kernels/CompareOp.java). */
enum vr_compare_op { VR_CMP_EQ = 0, VR_CMP_NE = 1, VR_CMP_LT = 2, VR_CMP_LE = 3, VR_CMP_GT = 4, VR_CMP_GE = 5 };

/* A residual condition on a join's candidate pairs, evaluated in the probe (the fused
 * PairPredicate, VectorHashJoinExec.scala:1225-1334): `left + left_offset <op> right +
 * right_offset` over two INT32 or two INT64 columns, each a streamed column (side 0: an index into
 * the probe's `columns`) or a build column (side 1).  The offsets are added in the lane's width,
 * wrapping; a pair with either operand null fails.  Inner: the passing pairs; semi / anti: rows with
 * a passing pair / without one; outer: the passing pairs, or (row, -1) when none passes
 * (emitOuterConditional, :1415-1450).  vr_compare_op values for `op`. */
typedef struct vr_join_residual {
    int32_t op;
    int32_t left_side, left_column;
    int32_t right_side, right_column;
    int64_t left_offset, right_offset;
} vr_join_residual;

typedef struct vr_probe_output {
    int32_t mem;                   /* vr_mem of every output buffer */
    int64_t capacity;              /* rows the outputs hold */
    int32_t* probe_idx;            /* may be NULL */
    int32_t* build_idx;            /* may be NULL (always NULL for semi/anti) */
    int32_t num_probe_out;         /* streamed columns to materialize */
    const int32_t* probe_out_columns;
    int32_t num_build_out;         /* build columns to materialize (inner only) */
    const int32_t* build_out_columns;
    vr_out_column* columns;        /* num_probe_out + num_build_out columns */
    int64_t rows;                  /* out: rows produced (or needed, with VR_ERR_CAPACITY) */
    const struct vr_join_residual* residual;  /* NULL: pairs by key equality only */
} vr_probe_output;

/* Probes one streamed batch (`columns`: num_columns streamed columns; `key_columns`: the key
 * column index of each join key) against the built table.  Pair order is the reference's:
 * streamed rows in order, each row's build rows in build order.  `drop_null_keys` is the
 * null-aware anti join (VectorHashJoinExec.scala:994,1070). */
vr_status vr_join_probe(vr_join* join, const vr_column* columns, int32_t num_columns,
                        const int32_t* key_columns, int64_t n, const uint8_t* selection,
                        int32_t join_type, int32_t drop_null_keys, vr_probe_output* out,
                        void* stream);
/* vr_join_probe without any device allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, hosts that forbid hipMalloc).  It has no
 * counterpart in the reference.  Columns, selection and outputs must be device-resident; a call
 * takes at most (131,070 * 2,048) = 268,431,360 rows (VR_ERR_UNSUPPORTED past that).  Its kernels
 * keep their scratch (tile states, the row count) in a 1 MiB array of the library on each device,
 * held until the process ends and shared by every call on that device: calls that overlap on one
 * device race, so a caller issues them one at a time per device.  Results are those of
 * vr_join_probe on the same inputs. */
vr_status vr_join_probe_noalloc(vr_join* join, const vr_column* columns, int32_t num_columns,
                                const int32_t* key_columns, int64_t n, const uint8_t* selection,
                                int32_t join_type, int32_t drop_null_keys, vr_probe_output* out,
                                void* stream);
/* Probes many streamed batches with one launch: the probe iterator's per-batch loop
 * (VectorHashJoinIterator.hasNext/probe/emitMatches/emitSemiAnti, VectorHashJoinExec.scala:1017-1182)
 * over every batch a task holds, without a host round trip per batch.  Batch b's columns are
 * columns[b * num_columns ...] (device-resident, the same lanes in every batch), its length
 * lengths[b], its selection selections[b] (selections or an entry may be NULL).  The outputs of all
 * batches follow each other in `out` (device memory), in batch order, each batch's in the order
 * vr_join_probe gives it; probe_idx is the row within its batch.  batch_rows (NULL: not wanted)
 * receives each batch's output rows; out->rows the total (with VR_ERR_CAPACITY when it passes
 * out->capacity). */
vr_status vr_join_probe_batches(vr_join* join, int32_t num_batches, const vr_column* columns,
                                int32_t num_columns, const int64_t* lengths,
                                const uint8_t* const* selections, const int32_t* key_columns,
                                int32_t join_type, int32_t drop_null_keys, vr_probe_output* out,
                                int64_t* batch_rows, void* stream);
/* The build rows no probe has paired since the build, as (-1, row) pairs in build order:
 * probe_idx -1, build_idx the row, the streamed columns null (of the types `columns` gives) and the
 * build columns materialized (emitUnmatchedBuild, VectorHashJoinExec.scala:1457-1484).  Needs
 * track_matched; `out->probe_out_columns` is ignored. */
vr_status vr_join_unmatched_build(vr_join* join, vr_probe_output* out, void* stream);
vr_status vr_join_memory(vr_join* join, int64_t* bytes);
vr_status vr_join_destroy(vr_join* join);

/* ------------------------------------------------------------------ filters (selection producers) */

/* A comparison's literal: `i64` for INT32 (narrowed as Java's intValue) and INT64 lanes, `f64` for
 * FLOAT64, the unscaled value's two limbs for DECIMAL128 (already at the column's scale). */
typedef struct vr_scalar {
    int64_t i64;
    double f64;
    int64_t lo, hi;
} vr_scalar;

/* The selection of `col <op> scalar` over rows [0, n) (CompareKernels.compareScalar,
 * kernels/src/main/java/io/vecruntime/kernels/CompareKernels.java:57-160, turned into a selection by
 * VectorExpr.selection, spark/src/main/scala/io/vecruntime/spark/expr/VectorExpr.scala:438-443):
 * bit i of `out` = row i selected by `selection` (NULL: every row) AND non-null AND the comparison
 * true; FLOAT64 compares in Spark's NaN-safe order (NaN equals NaN and is above everything, -0.0
 * equals 0.0).  `out` holds ceil(n/8) bytes padded to 8; `count` (may be NULL) receives its set
 * bits.  INT32 / INT64 / FLOAT64 / DECIMAL128 lanes; the column and `out` in device memory. */
vr_status vr_filter_compare(vr_device* dev, const vr_column* col, int32_t op, const vr_scalar* scalar,
                            const uint8_t* selection, int64_t n, uint8_t* out, int64_t* count, void* stream);

/* Bytes of the plain UTF8 values of the rows set in `selection` (CompactKernels.selectedUtf8Bytes,
 * CompactKernels.java:379-408): what a compacted UTF8 column's `data` must hold. */
vr_status vr_compact_utf8_bytes(vr_device* dev, const vr_column* col, int64_t n, const uint8_t* selection,
                                int64_t* bytes, void* stream);

/* Compacts `num_columns` columns of n rows to the rows set in `selection`, in row order
 * (CompactKernels.compactFixed / compactUtf8, CompactKernels.java:108-471; the compacting branch of
 * VectorFilterExec.process, spark/.../sql/vecruntime/VectorFilterExec.scala:80-115): out[c] gets the
 * values (fixed-width lanes, BOOL bits, the int32 codes of a dictionary-encoded UTF8 column, or a
 * plain UTF8 column's offsets and bytes) and, when it has a validity buffer, the validity bits.
 * `rows` (may be NULL) receives the selected count; every out[c] needs that capacity.  Columns and
 * outputs in device memory. */
vr_status vr_compact(vr_device* dev, const vr_column* columns, int32_t num_columns, int64_t n,
                     const uint8_t* selection, vr_out_column* out, int64_t* rows, void* stream);

/* Gathers rows of `col` by an index array (GatherKernels.gatherFixed / gatherValidity / gatherUtf8,
 * kernels/src/main/java/io/vecruntime/kernels/GatherKernels.java:42-308; ArrowOutput.gather, the
 * join's materialization of probeIdx / buildIdx): out row k = row idx[k], and an index of -1 is a
 * null (an outer join's padding) whose fixed-width value is zero.  Validity is set unless the row
 * is null or padded; a plain UTF8 column gets the bytes of its valid, real rows only (size `data`
 * with vr_gather_utf8_bytes).  Fixed-width, BOOL and plain UTF8 lanes; `idx` (int32), the column
 * and `out` in device memory; `out->capacity` at least m. */
vr_status vr_gather_utf8_bytes(vr_device* dev, const vr_column* col, const int32_t* idx, int64_t m,
                               int64_t* bytes, void* stream);
vr_status vr_gather(vr_device* dev, const vr_column* col, const int32_t* idx, int64_t m, vr_out_column* out,
                    void* stream);

/* ------------------------------------------------------------------ shuffle partitioning */

/* How Spark hashes a key column (PartitionKernels.KeyKind,
 * kernels/src/main/java/io/vecruntime/kernels/PartitionKernels.java:55-86); the lane alone does not
 * say (a long or a decimal, a float or a double). */
enum vr_hash_kind {
    VR_HASH_INT = 0,        /* IntegerType, DateType on an INT32 lane: hashInt */
    VR_HASH_LONG = 1,       /* LongType, timestamps, decimals of at most 18 digits on an INT64 lane: hashLong */
    VR_HASH_DOUBLE = 2,     /* DoubleType: hashLong of the bits, -0.0 as 0.0, one NaN */
    VR_HASH_BOOL = 3,       /* BooleanType on a BOOL lane: hashInt(1 / 0) */
    VR_HASH_UTF8 = 4,       /* StringType on a plain UTF8 lane: hashUnsafeBytes (Spark's byte-wise tail) */
    VR_HASH_DECIMAL128 = 5, /* a decimal above 18 digits: hashUnsafeBytes of the unscaled value's
                               minimal big-endian two's-complement bytes */
    VR_HASH_FLOAT = 6       /* FloatType on a FLOAT64 lane: hashInt of the float's bits, both zeros 0,
                               one NaN */
};

/* Spark's HashPartitioning ids, Pmod(Murmur3Hash(keys, seed), num_partitions) with Spark's seed 42
 * (PartitionKernels.hashPartitionIds, kernels/src/main/java/io/vecruntime/kernels/PartitionKernels.java:
 * 375-401; VectorShuffleManager.partitionIds, shuffle/src/main/scala/org/apache/spark/sql/vecruntime/
 * shuffle/VectorShuffleManager.scala:332-350; another seed for the buckets of an aggregate spill,
 * AggregateSpill.scala:106): every key mixes into the running hash in key order, a null key leaving
 * it unchanged (PartitionKernels.mixColumn, :199-250).  `ids` gets each row's partition, `hashes`
 * (may be NULL) its hash.  Key data, validity and offsets, `ids` and `hashes` in device memory; the
 * dictionary of a dictionary-encoded UTF8 key may be in host or device memory.  A code outside its
 * dictionary fails with VR_ERR_INVALID_ARGUMENT.  Synchronous with respect to `stream`. */
vr_status vr_hash_partition_ids(vr_device* dev, const vr_column* keys, const int32_t* kinds, int32_t num_keys,
                                int64_t n, int32_t num_partitions, int32_t seed, int32_t* ids, int32_t* hashes,
                                void* stream);

/* Spark's RoundRobinPartitioning ids continuing from `start` (PartitionKernels.roundRobinIds,
 * PartitionKernels.java:411-422): ids[i] = (start + i) mod num_partitions; `next` (may be NULL) gets
 * the start of the following batch.  `ids` in device memory. */
vr_status vr_round_robin_ids(vr_device* dev, int64_t n, int32_t num_partitions, int32_t start, int32_t* ids,
                             int32_t* next, void* stream);

/* The rows grouped by partition, row order kept within each -- a stable counting sort
 * (PartitionKernels.partitionOrder / partitionDestinations, PartitionKernels.java:437-482, as the
 * shuffle writer uses them, shuffle/src/main/scala/io/vecruntime/shuffle/PartitionedIpcWriter.scala:
 * 1003,1150-1154).  `rows` (may be NULL, n int32) lists partition 0's rows, then partition 1's, ...;
 * `dest` (may be NULL, n int32) the slot each row takes in that order; `starts` (num_partitions + 1
 * int64) where each partition begins, starts[num_partitions] = n.  An id outside [0, num_partitions)
 * fails with VR_ERR_INVALID_ARGUMENT.  Device memory; synchronous with respect to `stream`. */
vr_status vr_partition_rows(vr_device* dev, const int32_t* ids, int64_t n, int32_t num_partitions, int32_t* rows,
                            int32_t* dest, int64_t* starts, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VR_H */
