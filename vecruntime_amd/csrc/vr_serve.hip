// Continuous-batching query server: the vr_server_* entry points of include/vr_serve.h.
//
// One queue.  A free worker admits the queued query whose estimated device bytes fit the budget
// and whose input rows, divided by (1 + its wait / aging), are least (or, in FIFO mode, the head
// query once it fits); a query larger than the whole budget is admitted alone.
// Each worker owns a HIP stream; an admitted query runs every launch on it, so several admitted
// queries overlap on the device.  A group-by keeps its table resident across its batches and
// emits every group into server-owned device columns; a join builds, probes once and keeps its
// materialized output.  Results hold their bytes of the budget until released.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

#include "../../include/vr_serve.h"
#include "vr_internal.h"

namespace vr {

using Clock = std::chrono::steady_clock;

static double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// A query with every struct it points at copied (column buffers stay the caller's).
struct QueryCopy {
    vr_query q{};
    std::vector<int32_t> key_types, value_types;
    std::vector<uint8_t> nullable;
    std::vector<vr_agg_desc> aggs;
    vr_agg_spec spec{};
    std::vector<vr_column> bkeys, bvals;
    std::vector<int64_t> lengths;
    vr_batch_set bs{};
    std::vector<int32_t> jkt, jct, jkc;
    vr_join_spec jspec{};
    std::vector<vr_column> build_cols, probe_cols;
    std::vector<int32_t> pkc, pout, bout;
};

struct Ticket {
    int64_t id = 0;
    std::unique_ptr<QueryCopy> qc;
    int64_t est = 0;       // bytes reserved while it runs
    double rows = 0;       // input rows: the size admission orders by
    int64_t kept = 0;      // bytes its results hold until released
    Clock::time_point submit, admit, done;
    bool finished = false;
    vr_query_result res{};
    std::vector<std::pair<void*, size_t>> bufs;  // result allocations
    vr_agg* agg = nullptr;                       // kept for its key dictionaries (UTF8 keys)
};

static int key_lane_words(int32_t t) { return t == VR_INT64 || t == VR_FLOAT64 ? 2 : t == VR_DECIMAL128 ? 4 : 1; }

static int64_t estimate_agg(const QueryCopy& c) {
    int64_t rows = 0;
    for (int64_t n : c.lengths) rows += n;
    int64_t g = c.q.expected_groups > 0 ? c.q.expected_groups : std::max<int64_t>(rows, 1);
    // the table leaves room for up to 3M new groups per launch beyond the groups it holds
    int64_t room = g + std::min<int64_t>(rows, 3ll << 20);
    uint64_t slots = 2048;
    while ((double)slots * 0.5 < (double)room + 1) slots <<= 1;
    int words = 0;
    for (int32_t t : c.key_types) words += key_lane_words(t);
    int width = words > 2 || c.key_types.size() > 2 ? words + 1 : 1;
    int stride = 2;
    for (auto& a : c.aggs) {  // state words per aggregate, at most
        const bool wide = a.kind == VR_AGG_SUM_WIDE || a.kind == VR_AGG_AVG_WIDE || a.kind == VR_AGG_MERGE_SUM_WIDE ||
                          a.kind == VR_AGG_MERGE_AVG_WIDE;
        const bool minmax = a.kind == VR_AGG_MIN || a.kind == VR_AGG_MAX || a.kind == VR_AGG_MERGE_MIN || a.kind == VR_AGG_MERGE_MAX;
        stride += wide ? 7 : minmax ? 4 : 3;
    }
    int64_t out_row = 8 * (int64_t)c.key_types.size() + 24 * (int64_t)c.aggs.size() + 8;
    return (int64_t)slots * 8 * (width + stride) + g * out_row + (64ll << 20);
}

static int64_t estimate_join(const QueryCopy& c) {
    int64_t row_bytes = 0;
    for (int32_t t : c.jct) row_bytes += (int64_t)lane_bytes(t, 1);
    int64_t cap = c.q.expected_rows > 0 ? c.q.expected_rows : c.q.probe_rows;
    int64_t out_row = 8 * (int64_t)(c.pout.size() + c.bout.size());
    return c.q.build_rows * (32 + row_bytes + 24) + cap * (out_row + 8) + (64ll << 20);
}

}  // namespace vr

using namespace vr;

struct vr_server {
    vr_device* dev = nullptr;
    DeviceRef ref;
    int64_t budget = 0;
    int32_t admission = VR_ADMIT_SIZE_AWARE;
    double aging_ms = 100;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Ticket*> queue;
    std::unordered_map<int64_t, std::unique_ptr<Ticket>> tickets;
    int64_t next_id = 1;
    int64_t reserved = 0;
    int64_t running = 0;
    bool stopping = false;
    vr_server_stats stats{};
    std::vector<std::thread> threads;
};

static void free_results(vr_server* srv, Ticket* t) {
    for (auto& b : t->bufs) srv->dev->d.release(b.first, b.second);
    t->bufs.clear();
    if (t->agg) vr_agg_destroy(t->agg);
    t->agg = nullptr;
    t->res.agg = nullptr;
}

static vr_status alloc_out(vr_server* srv, Ticket* t, int32_t type, int64_t rows, bool validity, vr_out_column* oc) {
    std::memset(oc, 0, sizeof(*oc));
    oc->type = type;
    oc->mem = VR_MEM_DEVICE;
    oc->capacity = rows;
    size_t db = align_up(lane_bytes(type == VR_UTF8 ? VR_INT32 : type, std::max<int64_t>(rows, 1)) + 8, 256);
    void* p = nullptr;
    VR_TRY(srv->dev->d.alloc(&p, db));
    t->bufs.push_back({p, db});
    t->kept += (int64_t)db;
    oc->data = p;
    if (validity) {
        size_t vb = align_up((size_t)(std::max<int64_t>(rows, 1) + 63) / 64 * 8, 256);
        VR_TRY(srv->dev->d.alloc(&p, vb));
        t->bufs.push_back({p, vb});
        t->kept += (int64_t)vb;
        oc->validity = (uint8_t*)p;
    }
    return VR_OK;
}

static vr_status run_agg(vr_server* srv, Ticket* t, hipStream_t s) {
    QueryCopy& c = *t->qc;
    vr_agg* agg = nullptr;
    vr_agg_spec spec = c.spec;
    if (spec.capacity_hint <= 0) spec.capacity_hint = c.q.expected_groups;
    VR_TRY(vr_agg_create(srv->dev, &spec, &agg));
    // on an early return the worker's stream may still run this query's kernels: wait for them
    // before the handle's blocks go back to the shared cache
    auto destroy = [s](vr_agg* a) {
        (void)hipStreamSynchronize(s);
        vr_agg_destroy(a);
    };
    std::unique_ptr<vr_agg, decltype(destroy)> guard(agg, destroy);
    VR_TRY(vr_agg_update_batches(agg, &c.bs, nullptr, s));
    int64_t g = 0;
    VR_TRY(vr_agg_num_groups(agg, &g));
    int32_t nout = 0;
    int32_t types[VR_MAX_RESULT_COLUMNS];
    VR_TRY(vr_agg_output_layout(agg, &nout, types, VR_MAX_RESULT_COLUMNS));
    const int nk = (int)c.key_types.size();
    if (nk + nout > VR_MAX_RESULT_COLUMNS) return fail(VR_ERR_UNSUPPORTED, "too many result columns");
    bool utf8 = false;
    for (int k = 0; k < nk; k++) {
        utf8 |= c.key_types[k] == VR_UTF8;
        VR_TRY(alloc_out(srv, t, c.key_types[k] == VR_UTF8 ? VR_INT32 : c.key_types[k], g, true, &t->res.columns[k]));
    }
    for (int k = 0; k < nout; k++) VR_TRY(alloc_out(srv, t, types[k], g, true, &t->res.columns[nk + k]));
    if (g) VR_TRY(vr_agg_emit(agg, 0, g, t->res.columns, t->res.columns + nk, s));
    VR_TRY_HIP(hipStreamSynchronize(s));
    t->res.num_columns = nk + nout;
    t->res.rows = g;
    if (utf8) {  // the ids mean nothing without the table's dictionaries: keep the handle
        int64_t bytes = 0;
        VR_TRY(vr_agg_memory(agg, &bytes));
        t->kept += bytes;
        t->agg = guard.release();
        t->res.agg = t->agg;
    }
    return VR_OK;
}

static vr_status run_join(vr_server* srv, Ticket* t, hipStream_t s) {
    QueryCopy& c = *t->qc;
    vr_join* j = nullptr;
    VR_TRY(vr_join_create(srv->dev, &c.jspec, &j));
    auto destroy = [s](vr_join* h) {  // as in run_agg: the stream's work ends before the blocks return
        (void)hipStreamSynchronize(s);
        vr_join_destroy(h);
    };
    std::unique_ptr<vr_join, decltype(destroy)> guard(j, destroy);
    VR_TRY(vr_join_build_append(j, c.build_cols.data(), c.q.build_rows, s));
    VR_TRY(vr_join_build_finish(j, s));
    int64_t cap = c.q.expected_rows > 0 ? c.q.expected_rows : c.q.probe_rows;
    const int np = (int)c.pout.size(), nb = (int)c.bout.size();
    if (np + nb > VR_MAX_RESULT_COLUMNS) return fail(VR_ERR_UNSUPPORTED, "too many result columns");
    vr_probe_output out{};
    for (int attempt = 0; attempt < 2; attempt++) {
        for (int k = 0; k < np; k++) VR_TRY(alloc_out(srv, t, c.probe_cols[c.pout[k]].type, cap, true, &t->res.columns[k]));
        for (int k = 0; k < nb; k++) VR_TRY(alloc_out(srv, t, c.jct[c.bout[k]], cap, true, &t->res.columns[np + k]));
        out.mem = VR_MEM_DEVICE;
        out.capacity = cap;
        out.num_probe_out = np;
        out.probe_out_columns = c.pout.data();
        out.num_build_out = nb;
        out.build_out_columns = c.bout.data();
        out.columns = t->res.columns;
        vr_status st = vr_join_probe(j, c.probe_cols.data(), (int32_t)c.probe_cols.size(), c.pkc.data(), c.q.probe_rows,
                                     nullptr, c.q.join_type, 0, &out, s);
        if (st != VR_ERR_CAPACITY || attempt) {
            VR_TRY(st);
            break;
        }
        // the output outgrew the hint: size it to the rows the probe reported and probe again
        (void)hipStreamSynchronize(s);
        free_results(srv, t);
        t->kept = 0;
        cap = out.rows;
    }
    VR_TRY_HIP(hipStreamSynchronize(s));
    t->res.num_columns = np + nb;
    t->res.rows = out.rows;
    return VR_OK;
}

// The queued query a free worker admits next (srv->mu held), or queue.end() when none may start:
// the budget must hold its bytes unless nothing runs.
static std::deque<Ticket*>::iterator pick(vr_server* srv) {
    auto fits = [&](const Ticket* t) { return srv->reserved + t->est <= srv->budget || srv->running == 0; };
    if (srv->queue.empty()) return srv->queue.end();
    if (srv->admission == VR_ADMIT_FIFO) return fits(srv->queue.front()) ? srv->queue.begin() : srv->queue.end();
    const Clock::time_point now = Clock::now();
    auto best = srv->queue.end();
    double best_score = 0;
    for (auto it = srv->queue.begin(); it != srv->queue.end(); ++it) {
        if (!fits(*it)) continue;
        const double score = (*it)->rows / (1.0 + ms_between((*it)->submit, now) / srv->aging_ms);
        if (best == srv->queue.end() || score < best_score) {
            best = it;
            best_score = score;
        }
    }
    if (best == srv->queue.end() && srv->running == 0) best = srv->queue.begin();
    return best;
}

static void worker(vr_server* srv) {
    (void)hipSetDevice(srv->dev->d.ordinal);
    hipStream_t s = nullptr;
    (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    while (true) {
        std::unique_lock<std::mutex> lk(srv->mu);
        std::deque<Ticket*>::iterator next;
        srv->cv_work.wait(lk, [&] {
            if (srv->queue.empty()) return srv->stopping;
            next = pick(srv);
            return next != srv->queue.end();
        });
        if (srv->queue.empty()) break;
        Ticket* t = *next;
        srv->queue.erase(next);
        srv->reserved += t->est;
        srv->running++;
        srv->stats.peak_inflight = std::max(srv->stats.peak_inflight, srv->running);
        srv->stats.peak_reserved_bytes = std::max(srv->stats.peak_reserved_bytes, srv->reserved);
        t->admit = Clock::now();
        lk.unlock();
        vr_status st = t->qc->q.kind == VR_QUERY_AGG ? run_agg(srv, t, s) : run_join(srv, t, s);
        if (st == VR_OK) {
            hipError_t e = hipStreamSynchronize(s);
            if (e != hipSuccess) st = hip_fail(e, "hipStreamSynchronize");
        }
        if (st != VR_OK) {  // kernels queued before the failure may still write the result blocks
            (void)hipStreamSynchronize(s);
            free_results(srv, t);
            t->kept = 0;
        }
        lk.lock();
        t->done = Clock::now();
        t->res.status = st;
        t->res.queue_ms = ms_between(t->submit, t->admit);
        t->res.run_ms = ms_between(t->admit, t->done);
        t->res.latency_ms = ms_between(t->submit, t->done);
        t->res.device_bytes = t->est;
        srv->reserved += t->kept - t->est;
        srv->running--;
        t->finished = true;
        srv->stats.completed++;
        if (st != VR_OK) srv->stats.failed++;
        srv->stats.reserved_bytes = srv->reserved;
        srv->cv_done.notify_all();
        srv->cv_work.notify_all();
    }
    (void)hipStreamDestroy(s);
}

extern "C" {

vr_status vr_server_create(vr_device* dev, const vr_server_config* cfg, vr_server** out) {
    if (!dev || !out) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    auto* srv = new vr_server();
    srv->dev = dev;
    srv->ref.hold(dev);
    int workers = cfg && cfg->workers > 0 ? cfg->workers : 4;
    if (workers > 16) workers = 16;
    srv->budget = cfg && cfg->hbm_budget_bytes > 0 ? cfg->hbm_budget_bytes : (200ll << 30);
    if (cfg && (cfg->admission < VR_ADMIT_SIZE_AWARE || cfg->admission > VR_ADMIT_FIFO || cfg->aging_ms < 0)) {
        delete srv;
        return fail(VR_ERR_INVALID_ARGUMENT, "admission mode or aging");
    }
    if (cfg) srv->admission = cfg->admission;
    if (cfg && cfg->aging_ms > 0) srv->aging_ms = (double)cfg->aging_ms;
    for (int w = 0; w < workers; w++) srv->threads.emplace_back(worker, srv);
    *out = srv;
    return VR_OK;
}

vr_status vr_server_submit(vr_server* srv, const vr_query* q, int64_t* ticket) {
    if (!srv || !q || !ticket) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    auto qc = std::make_unique<QueryCopy>();
    qc->q = *q;
    if (q->kind == VR_QUERY_AGG) {
        if (!q->agg || !q->batches) return fail(VR_ERR_INVALID_ARGUMENT, "a group-by query needs its spec and batches");
        const vr_agg_spec& a = *q->agg;
        qc->key_types.assign(a.key_types, a.key_types + a.num_keys);
        if (a.key_nullable) qc->nullable.assign(a.key_nullable, a.key_nullable + a.num_keys);
        if (a.num_values) qc->value_types.assign(a.value_types, a.value_types + a.num_values);
        if (a.num_aggs) qc->aggs.assign(a.aggs, a.aggs + a.num_aggs);
        qc->spec = a;
        qc->spec.key_types = qc->key_types.data();
        qc->spec.key_nullable = a.key_nullable ? qc->nullable.data() : nullptr;
        qc->spec.value_types = qc->value_types.data();
        qc->spec.aggs = qc->aggs.data();
        const vr_batch_set& b = *q->batches;
        qc->bkeys.assign(b.keys, b.keys + (size_t)b.num_batches * a.num_keys);
        qc->bvals.assign(b.values, b.values + (size_t)b.num_batches * a.num_values);
        qc->lengths.assign(b.lengths, b.lengths + b.num_batches);
        qc->bs = b;
        qc->bs.keys = qc->bkeys.data();
        qc->bs.values = qc->bvals.data();
        qc->bs.lengths = qc->lengths.data();
    } else if (q->kind == VR_QUERY_JOIN) {
        if (!q->join || !q->build_columns || !q->probe_columns || !q->probe_key_columns)
            return fail(VR_ERR_INVALID_ARGUMENT, "a join query needs its spec, build and probe columns");
        const vr_join_spec& js = *q->join;
        qc->jkt.assign(js.key_types, js.key_types + js.num_keys);
        qc->jct.assign(js.column_types, js.column_types + js.num_columns);
        qc->jkc.assign(js.key_columns, js.key_columns + js.num_keys);
        qc->jspec = js;
        qc->jspec.key_types = qc->jkt.data();
        qc->jspec.column_types = qc->jct.data();
        qc->jspec.key_columns = qc->jkc.data();
        qc->jspec.capacity_hint = q->build_rows;
        qc->build_cols.assign(q->build_columns, q->build_columns + js.num_columns);
        qc->probe_cols.assign(q->probe_columns, q->probe_columns + q->num_probe_columns);
        qc->pkc.assign(q->probe_key_columns, q->probe_key_columns + js.num_keys);
        if (q->num_probe_out) qc->pout.assign(q->probe_out_columns, q->probe_out_columns + q->num_probe_out);
        if (q->num_build_out) qc->bout.assign(q->build_out_columns, q->build_out_columns + q->num_build_out);
        for (int32_t c : qc->pout)
            if (c < 0 || c >= q->num_probe_columns) return fail(VR_ERR_INVALID_ARGUMENT, "probe output column out of range");
        for (int32_t c : qc->bout)
            if (c < 0 || c >= js.num_columns) return fail(VR_ERR_INVALID_ARGUMENT, "build output column out of range");
    } else {
        return fail(VR_ERR_INVALID_ARGUMENT, "query kind");
    }
    auto t = std::make_unique<Ticket>();
    t->est = q->kind == VR_QUERY_AGG ? estimate_agg(*qc) : estimate_join(*qc);
    if (q->kind == VR_QUERY_AGG)
        for (int64_t n : qc->lengths) t->rows += (double)n;
    else
        t->rows = (double)q->build_rows + (double)q->probe_rows;
    t->qc = std::move(qc);
    t->submit = Clock::now();
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stopping) return fail(VR_ERR_STATE, "the server is stopping");
    t->id = srv->next_id++;
    *ticket = t->id;
    srv->queue.push_back(t.get());
    srv->tickets[t->id] = std::move(t);
    srv->stats.submitted++;
    srv->cv_work.notify_all();
    return VR_OK;
}

vr_status vr_server_wait(vr_server* srv, int64_t ticket, vr_query_result* out) {
    if (!srv || !out) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return fail(VR_ERR_INVALID_ARGUMENT, "unknown ticket");
    Ticket* t = it->second.get();
    srv->cv_done.wait(lk, [&] { return t->finished; });
    *out = t->res;
    return VR_OK;
}

vr_status vr_server_copy_result(vr_server* srv, int64_t ticket, int32_t column, vr_out_column* dst) {
    if (!srv || !dst) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    vr_out_column src;
    int64_t rows = 0;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->tickets.find(ticket);
        if (it == srv->tickets.end()) return fail(VR_ERR_INVALID_ARGUMENT, "unknown ticket");
        Ticket* t = it->second.get();
        if (!t->finished) return fail(VR_ERR_STATE, "the query has not finished");
        if (t->res.status != VR_OK) return fail(VR_ERR_STATE, "the query failed");
        if (column < 0 || column >= t->res.num_columns) return fail(VR_ERR_INVALID_ARGUMENT, "result column out of range");
        src = t->res.columns[column];
        rows = t->res.rows;
    }
    if (dst->type != src.type) return fail(VR_ERR_INVALID_ARGUMENT, "destination lane differs from the result column");
    if (dst->capacity < rows) return fail(VR_ERR_CAPACITY, "destination holds fewer rows than the result");
    if (!dst->data) return fail(VR_ERR_INVALID_ARGUMENT, "destination has no data buffer");
    VR_TRY_HIP(hipSetDevice(srv->dev->d.ordinal));
    size_t db = lane_bytes(src.type, rows);
    if (db) VR_TRY_HIP(hipMemcpy(dst->data, src.data, db, hipMemcpyDefault));
    if (dst->validity) {
        size_t vb = bitmap_bytes(rows);
        if (src.validity && vb) {
            VR_TRY_HIP(hipMemcpy(dst->validity, src.validity, vb, hipMemcpyDefault));
        } else if (vb) {  // no nulls: every bit set
            std::vector<uint8_t> ones(vb, 0xFF);
            VR_TRY_HIP(hipMemcpy(dst->validity, ones.data(), vb, hipMemcpyDefault));
        }
    }
    return VR_OK;
}

vr_status vr_server_release(vr_server* srv, int64_t ticket) {
    if (!srv) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return fail(VR_ERR_INVALID_ARGUMENT, "unknown ticket");
    Ticket* t = it->second.get();
    if (!t->finished) return fail(VR_ERR_STATE, "the query has not finished");
    free_results(srv, t);
    srv->reserved -= t->kept;
    srv->stats.reserved_bytes = srv->reserved;
    srv->tickets.erase(it);
    srv->cv_work.notify_all();
    return VR_OK;
}

vr_status vr_server_stats_get(vr_server* srv, vr_server_stats* out) {
    if (!srv || !out) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = srv->stats;
    out->reserved_bytes = srv->reserved;
    return VR_OK;
}

vr_status vr_server_destroy(vr_server* srv) {
    if (!srv) return VR_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stopping = true;
        srv->cv_work.notify_all();
    }
    for (auto& th : srv->threads) th.join();
    for (auto& kv : srv->tickets) free_results(srv, kv.second.get());
    delete srv;
    return VR_OK;
}

}  // extern "C"
