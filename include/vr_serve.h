/*
 * vr_serve.h -- continuous-batching query server of libvrgpu.so.
 *
 * Queries (a group-by over device-resident batches, or a join build + probe) are submitted from
 * any thread; the server admits a queued query as soon as a worker is free and the query's
 * estimated device memory fits the HBM budget (smallest input first, with aging, or in arrival
 * order: vr_server_config.admission), runs each admitted query on its worker's own HIP
 * stream (so admitted queries overlap on the device), and keeps the query's hash table resident
 * across its batches.  A finished query's results stay in server-owned device memory, counted
 * against the budget, until vr_server_release.
 *
 * The reference has no serving layer: its queries are Spark tasks, one hash table per task
 * (spark/src/main/scala/org/apache/spark/sql/vecruntime/VectorHashAggregateExec.scala:478,1022-1026;
 * a broadcast build table shared by an executor's tasks, VectorHashJoinExec.scala:733-773).  This
 * header has no counterpart there; it is the caller-facing form of the same per-query operators
 * (vr_agg_*, vr_join_* in vr.h) for a process that serves many concurrent queries.
 */
#ifndef VR_SERVE_H
#define VR_SERVE_H

#include "vr.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vr_server vr_server;

enum vr_admission {
    /* A free worker takes the queued query with the least input rows divided by
     * (1 + its wait / aging_ms), among those whose memory fits the budget: small queries pass
     * large ones, and a large query's claim grows with its wait, so none waits for ever. */
    VR_ADMIT_SIZE_AWARE = 0,
    VR_ADMIT_FIFO = 1, /* arrival order; the head query waits until its memory fits */
};

typedef struct vr_server_config {
    int32_t workers;           /* concurrent queries (each on its own stream); 0: 4 */
    int32_t admission;         /* vr_admission */
    int64_t hbm_budget_bytes;  /* device bytes admitted queries may hold at once; 0: 200 GB */
    int64_t aging_ms;          /* VR_ADMIT_SIZE_AWARE: wait that halves a query's size; 0: 100 */
} vr_server_config;

enum vr_query_kind { VR_QUERY_AGG = 0, VR_QUERY_JOIN = 1 };

typedef struct vr_query {
    int32_t kind;
    int32_t join_type;                  /* join: vr_join_type */
    /* group-by: the spec and its batches (device-resident columns, valid until the query ends) */
    const vr_agg_spec* agg;
    const vr_batch_set* batches;
    int64_t expected_groups;            /* sizing and admission hint (0: the input rows) */
    /* join: build columns and probe columns, device resident */
    const vr_join_spec* join;
    const vr_column* build_columns;
    int64_t build_rows;
    const vr_column* probe_columns;
    int32_t num_probe_columns;
    int32_t num_probe_out;
    const int32_t* probe_key_columns;
    const int32_t* probe_out_columns;
    int64_t probe_rows;
    int32_t num_build_out;
    int32_t reserved;
    const int32_t* build_out_columns;
    int64_t expected_rows;              /* join output capacity (0: the probe rows) */
} vr_query;

#define VR_MAX_RESULT_COLUMNS 40

typedef struct vr_query_result {
    vr_status status;                   /* the query's own status */
    int32_t num_columns;
    int64_t rows;                       /* groups (group-by) or output rows (join) */
    vr_out_column columns[VR_MAX_RESULT_COLUMNS]; /* device results: keys then buffers, or probe then build columns */
    double queue_ms;                    /* submit -> admitted */
    double run_ms;                      /* admitted -> done */
    double latency_ms;                  /* submit -> done */
    int64_t device_bytes;               /* budget the query held while it ran */
    vr_agg* agg;                        /* group-by over UTF8 keys: its finished handle, for
                                           vr_agg_key_dictionary (the keys are dictionary ids);
                                           kept until release.  NULL otherwise. */
} vr_query_result;

typedef struct vr_server_stats {
    int64_t submitted;
    int64_t completed;
    int64_t failed;
    int64_t peak_inflight;              /* most queries running at once */
    int64_t peak_reserved_bytes;        /* most budget in use at once */
    int64_t reserved_bytes;             /* budget in use now (running queries and unreleased results) */
} vr_server_stats;

vr_status vr_server_create(vr_device* dev, const vr_server_config* cfg, vr_server** out);
/* Queues a query (the structs are copied; column buffers must stay valid until it completes). */
vr_status vr_server_submit(vr_server* srv, const vr_query* q, int64_t* ticket);
/* Blocks until the query finishes; fills its result (the call succeeds even if the query failed:
 * see result->status). */
vr_status vr_server_wait(vr_server* srv, int64_t ticket, vr_query_result* out);
/* Copies result column `column` of a finished query (its `rows` rows: data and validity) into a
 * caller's column, host or device (dst->mem), of at least that capacity. */
vr_status vr_server_copy_result(vr_server* srv, int64_t ticket, int32_t column, vr_out_column* dst);
/* Frees a finished query's result memory and returns its budget. */
vr_status vr_server_release(vr_server* srv, int64_t ticket);
vr_status vr_server_stats_get(vr_server* srv, vr_server_stats* out);
/* Waits for every queued query, then stops the workers. */
vr_status vr_server_destroy(vr_server* srv);

#ifdef __cplusplus
}
#endif

#endif /* VR_SERVE_H */
