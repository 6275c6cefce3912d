// serve.hip - continuous batching of filtered GROUP BY requests over one device-resident table
// (qk_serve_* in include/qikkdb_amd.h).  Host code only: the device work is the batch engine's.
//
// Requests are queued as they are submitted.  A dispatcher thread owns `engines` batch
// engines, each with its own stream.  Whenever an engine is free it takes the oldest requests
// (first come, first served) while the projected bytes of everything admitted and not yet
// released stay under the memory cap, and runs them as one batch.  A new batch is started
// while another is still on the device, so the two overlap there.  It is started when the
// device is idle, when a full batch is waiting, or when the oldest request has waited
// `wait_ms`.  A batch writes its groups into a result block of its own; a request is done when
// its batch's completion event has fired, and the engine is free again at that point.  The
// block stays in device memory until every request of the batch is released, so results are
// never held by an engine and callers may release requests in any order.
//
// Projected bytes of a request: its dense result table, key_mod groups x 32 B (count, sum, min,
// max; a dense table has load factor 1), admitted until the request is released, plus, for a
// request with more than lds_groups groups, its share of the two-pass pool, (estimated rows x
// 1.25 + 4096) x 12 B, admitted until its batch finishes.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace {

using Clock = std::chrono::steady_clock;

enum class TState { Pending, Running, Done, Failed };
enum class EState { Free, Running };

// Device result arrays of one batch, shared by its requests.
struct Block {
    void* mem = nullptr;    // count, sum, min, max: 4 x total int64
    int64_t total = 0;      // sum of key_mod over the batch
    int32_t refs = 0;       // requests not yet released
};

struct Ticket {
    qk_batch_query q{};
    int64_t bytes = 0;      // projected
    int64_t rbytes = 0;     // of which result table (held until release)
    double t_submit = 0.0;  // ms since the server started
    double t_due = 0.0;     // latency is measured from here (the submit time, or a scheduled arrival)
    double t_done = 0.0;
    TState state = TState::Pending;
    int err = QK_SUCCESS;
    std::string msg;
    Block* blk = nullptr;
    int64_t off = 0;        // its group 0 in the block's arrays
    bool auto_release = false;
    double* latency_out = nullptr;  // auto-release: where the latency goes
    bool admitted = false;          // dispatched (its projected bytes were admitted)
};

struct Engine {
    qk_batch b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t done = nullptr;
    EState state = EState::Free;
    std::vector<int64_t> tickets;
    int64_t pool_bytes = 0; // projected pool share of the batch (admitted until it finishes)
    double t_start = 0.0;
};

}  // namespace

struct qk_serve_s {
    qk_batch_columns cols{};
    qk_serve_config cfg{};
    int device = 0;
    Clock::time_point t0;
    std::vector<Engine> eng;
    std::mutex mu;
    std::condition_variable cv_work;  // the dispatcher waits here
    std::condition_variable cv_done;  // callers of qk_serve_wait wait here
    std::deque<int64_t> pending;
    std::unordered_map<int64_t, Ticket> tickets;
    int64_t next_ticket = 1;
    int64_t admitted_bytes = 0;       // projected bytes of dispatched, unreleased requests
    std::vector<Block*> blocks;       // live result blocks
    std::vector<void*> to_free;       // device memory of released blocks (freed by the dispatcher)
    int64_t outstanding_auto = 0;     // auto-release tickets not yet done
    bool stop = false;
    std::thread th;
    // statistics
    int64_t batches = 0, served = 0, failed = 0, max_batch_seen = 0, peak_bytes = 0;
    double busy_ms = 0.0, last_batch_ms = 0.0;
};

namespace {

double now_ms(const qk_serve_s* s) {
    return std::chrono::duration<double, std::milli>(Clock::now() - s->t0).count();
}

// Drops one reference to a result block (lock held); the last one frees it.
void unref(qk_serve_s* s, Block* b) {
    if (!b || --b->refs > 0) return;
    if (b->mem) s->to_free.push_back(b->mem);
    s->blocks.erase(std::find(s->blocks.begin(), s->blocks.end(), b));
    delete b;
}

// A request leaves the server (lock held): its result table is no longer admitted.
void drop_ticket(qk_serve_s* s, int64_t id) {
    Ticket& k = s->tickets[id];
    if (k.admitted) s->admitted_bytes -= k.rbytes;
    unref(s, k.blk);
    s->tickets.erase(id);
}

// Finishes ticket `id` of a completed batch (lock held).
void finish(qk_serve_s* s, int64_t id, double t, int err, const std::string& msg) {
    Ticket& k = s->tickets[id];
    k.t_done = t;
    k.state = err == QK_SUCCESS ? TState::Done : TState::Failed;
    k.err = err;
    k.msg = msg;
    s->admitted_bytes -= k.bytes - k.rbytes;  // its pool share
    if (err == QK_SUCCESS) s->served++;
    else s->failed++;
    if (k.auto_release) {
        if (k.latency_out) *k.latency_out = err == QK_SUCCESS ? t - k.t_due : -1.0;
        drop_ticket(s, id);
        s->outstanding_auto--;
    }
}

// Fails every request not yet dispatched (lock held): the server is stopping.
void drop_pending(qk_serve_s* s) {
    const double t = now_ms(s);
    for (int64_t id : s->pending) {
        Ticket& k = s->tickets[id];
        k.t_done = t;
        k.state = TState::Failed;
        k.err = QK_ERROR;
        k.msg = "qk_serve: the server stopped before the request ran";
        s->failed++;
        if (k.auto_release) {
            if (k.latency_out) *k.latency_out = -1.0;
            s->tickets.erase(id);
            s->outstanding_auto--;
        }
    }
    s->pending.clear();
}


// Marks the batches whose completion event fired (lock held).  Returns true if any did.
bool collect(qk_serve_s* s) {
    bool any = false;
    for (Engine& e : s->eng) {
        if (e.state != EState::Running) continue;
        const hipError_t q = hipEventQuery(e.done);
        if (q == hipErrorNotReady) continue;
        const double t = now_ms(s);
        int err = QK_SUCCESS;
        std::string msg;
        if (q != hipSuccess) {
            err = QK_RUNTIME_ERROR;
            msg = std::string("qk_serve: batch failed on the device: ") + hipGetErrorString(q);
            (void)hipGetLastError();
        }
        e.state = EState::Free;
        s->batches++;
        s->last_batch_ms = t - e.t_start;
        s->busy_ms += s->last_batch_ms;
        for (int64_t id : e.tickets) finish(s, id, t, err, msg);
        e.tickets.clear();
        any = true;
    }
    return any;
}

void dispatcher(qk_serve_s* s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    std::vector<qk_batch_query> qs;
    while (true) {
        if (collect(s)) s->cv_done.notify_all();
        for (void* m : s->to_free) (void)qk::pool_free(m);
        s->to_free.clear();
        int running = 0;
        Engine* free_e = nullptr;
        for (Engine& e : s->eng) {
            running += e.state == EState::Running;
            if (e.state == EState::Free && !free_e) free_e = &e;
        }
        if (s->stop) {
            if (!s->pending.empty()) {
                drop_pending(s);
                s->cv_done.notify_all();
            }
            if (running == 0) break;
        }
        bool launched = false;
        if (free_e && !s->pending.empty()) {
            const double t = now_ms(s);
            const double waited = t - s->tickets[s->pending.front()].t_submit;
            const double wait_ms = s->cfg.wait_ms > 0 ? s->cfg.wait_ms : 0.5 * s->last_batch_ms;
            if (running == 0 || (int64_t)s->pending.size() >= s->cfg.max_batch || waited >= wait_ms) {
                // admit the oldest requests while their projected bytes fit
                Engine& e = *free_e;
                int64_t bytes = 0, pool_bytes = 0, total = 0;
                e.tickets.clear();
                qs.clear();
                while (!s->pending.empty() && (int64_t)qs.size() < s->cfg.max_batch) {
                    Ticket& k = s->tickets[s->pending.front()];
                    if (s->admitted_bytes + bytes + k.bytes > s->cfg.memory_cap) break;
                    bytes += k.bytes;
                    pool_bytes += k.bytes - k.rbytes;
                    k.state = TState::Running;
                    k.admitted = true;
                    k.off = total;
                    total += k.q.key_mod;
                    qs.push_back(k.q);
                    e.tickets.push_back(s->pending.front());
                    s->pending.pop_front();
                }
                if (!qs.empty()) {
                    e.state = EState::Running;
                    e.pool_bytes = pool_bytes;
                    e.t_start = t;
                    s->admitted_bytes += bytes;
                    s->peak_bytes = std::max(s->peak_bytes, s->admitted_bytes);
                    s->max_batch_seen = std::max<int64_t>(s->max_batch_seen, (int64_t)qs.size());
                    Block* blk = new Block();
                    blk->total = total;
                    blk->refs = (int32_t)qs.size();
                    s->blocks.push_back(blk);
                    for (int64_t id : e.tickets) s->tickets[id].blk = blk;
                    lk.unlock();
                    int st = qk::pool_alloc(&blk->mem, (size_t)total * 32);
                    if (st == QK_SUCCESS) {
                        int64_t* r = (int64_t*)blk->mem;
                        st = qk_batch_run_into(e.b, (int32_t)qs.size(), qs.data(), r, r + total, r + 2 * total,
                                               r + 3 * total, e.s);
                    }
                    std::string msg;
                    if (st == QK_SUCCESS && hipEventRecord(e.done, e.s) != hipSuccess) {
                        (void)hipGetLastError();
                        st = QK_RUNTIME_ERROR;
                        msg = "qk_serve: hipEventRecord failed";
                    } else if (st != QK_SUCCESS) {
                        msg = qk_last_error_message();
                    }
                    lk.lock();
                    if (st != QK_SUCCESS) {
                        const double tf = now_ms(s);
                        e.state = EState::Free;
                        for (int64_t id : e.tickets) finish(s, id, tf, st, msg);
                        e.tickets.clear();
                        s->cv_done.notify_all();
                    }
                    launched = true;
                }
            }
        }
        if (launched) continue;
        // completion events are polled; arrivals, releases and stop wake the wait early
        if (running > 0 || !s->pending.empty()) s->cv_work.wait_for(lk, std::chrono::microseconds(100));
        else s->cv_work.wait(lk);
    }
}

int64_t projected_bytes(const qk_serve_s* s, const qk_batch_query& q, int64_t est_rows) {
    int64_t b = (int64_t)q.key_mod * 32;
    if (q.key_mod > s->cfg.tuning.lds_groups) b += (est_rows + est_rows / 4 + 4096) * 12;
    return b;
}

int submit(qk_serve_s* s, const qk_batch_query* q, double due_ms, bool auto_release, double* latency_out,
           int64_t* ticket) {
    int64_t est = 0;
    QK_TRY(qk_batch_estimate_rows(s->eng[0].b, 1, q, &est));  // validates the block range too
    if (q->key_mod < 1 || q->key_mod > (1 << 24)) return qk::fail(QK_ERROR, "qk_serve_submit: key_mod must be 1..2^24");
    const int64_t bytes = projected_bytes(s, *q, est);
    if (bytes > s->cfg.memory_cap)
        return qk::fail(QK_OUT_OF_MEMORY, "qk_serve_submit: projected %lld bytes exceed the cap of %lld",
                        (long long)bytes, (long long)s->cfg.memory_cap);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return qk::fail(QK_ERROR, "qk_serve_submit: server is stopping");
        const int64_t id = s->next_ticket++;
        Ticket& k = s->tickets[id];
        k.q = *q;
        k.bytes = bytes;
        k.rbytes = (int64_t)q->key_mod * 32;
        k.t_submit = now_ms(s);
        k.t_due = due_ms >= 0 ? due_ms : k.t_submit;
        k.auto_release = auto_release;
        k.latency_out = latency_out;
        s->pending.push_back(id);
        if (auto_release) s->outstanding_auto++;
        if (ticket) *ticket = id;
    }
    s->cv_work.notify_one();
    return QK_SUCCESS;
}

void destroy(qk_serve_s* s) {
    if (s->th.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        s->th.join();
    }
    (void)hipSetDevice(s->device);
    for (Engine& e : s->eng) {
        if (e.s) (void)hipStreamSynchronize(e.s);
        if (e.b) (void)qk_batch_destroy(e.b);
        if (e.done) (void)hipEventDestroy(e.done);
        if (e.s) (void)hipStreamDestroy(e.s);
    }
    // result blocks of requests never released
    for (Block* b : s->blocks) {
        if (b->mem) s->to_free.push_back(b->mem);
        delete b;
    }
    for (void* m : s->to_free) (void)qk::pool_free(m);
    delete s;
}

}  // namespace

extern "C" {

int qk_serve_create(qk_serve* out, const qk_batch_columns* cols, const qk_serve_config* cfg) {
    if (!out || !cols) return qk::fail(QK_ERROR, "qk_serve_create: null argument");
    *out = nullptr;
    QK_TRY(qk::ensure_init());
    qk_serve_config c{};
    if (cfg) c = *cfg;
    if (c.engines == 0) c.engines = 2;
    if (c.max_batch == 0) c.max_batch = 4096;
    if (c.memory_cap == 0) c.memory_cap = (int64_t)200 * 1000 * 1000 * 1000;
    if (c.tuning.tile_rows == 0) {
        c.tuning.tile_rows = 1 << 18;
        c.tuning.lds_groups = 4096;
        c.tuning.wgs_per_cu = 1;
        c.tuning.use_graph = 0;
        c.tuning.scratch_rows = 0;
    }
    if (c.engines < 1 || c.engines > 8 || c.max_batch < 1 || c.memory_cap < 1 || c.wait_ms < 0)
        return qk::fail(QK_ERROR, "qk_serve_create: engines 1..8, max_batch >= 1, memory_cap >= 1, wait_ms >= 0");
    {
        // Each engine keeps its row pool (12 bytes per pooled row, grown to the largest batch it
        // ran) beside the bytes the cap admits, so the engines share what the device has free
        // over the cap, less a tenth left to the runtime (at their default of 2^32 rows, four
        // engines' pools would not fit beside a 200 GB cap on 288 GB).
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            (void)hipGetLastError();
            return qk::fail(QK_RUNTIME_ERROR, "qk_serve_create: hipMemGetInfo failed");
        }
        const int64_t spare = (int64_t)free_b - c.memory_cap;
        const int64_t rows = spare > 0 ? spare / 10 * 9 / c.engines / 12 : 0;
        if (rows < ((int64_t)1 << 20))
            return qk::fail(QK_OUT_OF_MEMORY,
                            "qk_serve_create: a memory cap of %lld bytes leaves %lld of the %zu free bytes for "
                            "the row pools of %d engines", (long long)c.memory_cap, (long long)std::max<int64_t>(spare, 0),
                            free_b, c.engines);
        if (c.tuning.scratch_rows <= 0 || c.tuning.scratch_rows > rows) c.tuning.scratch_rows = rows;
    }
    qk_serve_s* s = new (std::nothrow) qk_serve_s();
    if (!s) return qk::fail(QK_OUT_OF_MEMORY, "qk_serve_create: host allocation");
    s->cols = *cols;
    s->cfg = c;
    s->t0 = Clock::now();
    (void)hipGetDevice(&s->device);
    s->eng.resize((size_t)c.engines);
    int st = QK_SUCCESS;
    for (Engine& e : s->eng) {
        st = qk_batch_create(&e.b, cols);
        if (st == QK_SUCCESS) st = qk_batch_set_tuning(e.b, &c.tuning);
        if (st == QK_SUCCESS && (hipStreamCreateWithFlags(&e.s, hipStreamNonBlocking) != hipSuccess ||
                                 hipEventCreateWithFlags(&e.done, hipEventDisableTiming) != hipSuccess)) {
            (void)hipGetLastError();
            st = qk::fail(QK_RUNTIME_ERROR, "qk_serve_create: stream or event creation failed");
        }
        if (st != QK_SUCCESS) break;
    }
    if (st != QK_SUCCESS) {
        destroy(s);
        return st;
    }
    s->th = std::thread(dispatcher, s);
    *out = s;
    return QK_SUCCESS;
}

int qk_serve_destroy(qk_serve s) {
    if (!s) return QK_SUCCESS;
    destroy(s);
    return QK_SUCCESS;
}

int qk_serve_submit(qk_serve s, const qk_batch_query* q, int64_t* ticket) {
    if (!s || !q || !ticket) return qk::fail(QK_ERROR, "qk_serve_submit: null argument");
    return submit(s, q, -1.0, false, nullptr, ticket);
}

int qk_serve_wait(qk_serve s, int64_t ticket, double timeout_ms) {
    if (!s) return qk::fail(QK_ERROR, "qk_serve_wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end() || it->second.auto_release) return qk::fail(QK_ERROR, "qk_serve_wait: unknown ticket");
    const auto ready = [&] {
        auto i = s->tickets.find(ticket);
        return i == s->tickets.end() || i->second.state == TState::Done || i->second.state == TState::Failed;
    };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready))
        return qk::fail(QK_ERROR, "qk_serve_wait: timed out");
    it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return qk::fail(QK_ERROR, "qk_serve_wait: ticket released while waiting");
    if (it->second.state == TState::Failed) return qk::fail(it->second.err, "%s", it->second.msg.c_str());
    return QK_SUCCESS;
}

int qk_serve_result(qk_serve s, int64_t ticket, const int64_t** count, const int64_t** sum, const int64_t** min,
                    const int64_t** max) {
    if (!s) return qk::fail(QK_ERROR, "qk_serve_result: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end() || it->second.state != TState::Done)
        return qk::fail(QK_ERROR, "qk_serve_result: ticket %lld is not done", (long long)ticket);
    const Ticket& k = it->second;
    const int64_t* r = (const int64_t*)k.blk->mem;
    const int64_t t = k.blk->total;
    if (count) *count = r + k.off;
    if (sum) *sum = r + t + k.off;
    if (min) *min = r + 2 * t + k.off;
    if (max) *max = r + 3 * t + k.off;
    return QK_SUCCESS;
}

int qk_serve_latency_ms(qk_serve s, int64_t ticket, double* out) {
    if (!s || !out) return qk::fail(QK_ERROR, "qk_serve_latency_ms: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end() || it->second.state != TState::Done)
        return qk::fail(QK_ERROR, "qk_serve_latency_ms: ticket %lld is not done", (long long)ticket);
    *out = it->second.t_done - it->second.t_due;
    return QK_SUCCESS;
}

int qk_serve_release(qk_serve s, int64_t ticket) {
    if (!s) return qk::fail(QK_ERROR, "qk_serve_release: null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->tickets.find(ticket);
        if (it == s->tickets.end() || it->second.auto_release)
            return qk::fail(QK_ERROR, "qk_serve_release: unknown ticket");
        Ticket& k = it->second;
        if (k.state == TState::Running) return qk::fail(QK_ERROR, "qk_serve_release: ticket is running");
        if (k.state == TState::Pending) {
            s->pending.erase(std::find(s->pending.begin(), s->pending.end(), ticket));
            s->tickets.erase(it);
        } else {
            drop_ticket(s, ticket);
        }
    }
    s->cv_work.notify_one();
    return QK_SUCCESS;
}

int qk_serve_replay(qk_serve s, int32_t n, const qk_batch_query* q, const double* arrival_ms, double* latency_ms) {
    if (!s || n < 0 || (n > 0 && (!q || !arrival_ms || !latency_ms)))
        return qk::fail(QK_ERROR, "qk_serve_replay: bad argument");
    for (int32_t i = 1; i < n; i++)
        if (arrival_ms[i] < arrival_ms[i - 1]) return qk::fail(QK_ERROR, "qk_serve_replay: arrivals must not decrease");
    for (int32_t i = 0; i < n; i++) latency_ms[i] = -1.0;
    const double start = now_ms(s);
    int st = QK_SUCCESS;
    for (int32_t i = 0; i < n && st == QK_SUCCESS; i++) {
        const double due = start + arrival_ms[i];
        const double wait = due - now_ms(s);
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double, std::milli>(wait));
        int64_t id = 0;
        st = submit(s, q + i, due, true, latency_ms + i, &id);
    }
    // the submitted requests write latency_ms when they finish: wait for all of them
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->outstanding_auto == 0; });
    if (st != QK_SUCCESS) return st;
    for (int32_t i = 0; i < n; i++)
        if (latency_ms[i] < 0) return qk::fail(QK_RUNTIME_ERROR, "qk_serve_replay: request %d failed", i);
    return QK_SUCCESS;
}

int qk_serve_replay_plain(const qk_batch_columns* cols, int32_t n, const qk_batch_query* q,
                          const double* arrival_ms, double* latency_ms, void* stream) {
    if (!cols || n < 0 || (n > 0 && (!q || !arrival_ms || !latency_ms)))
        return qk::fail(QK_ERROR, "qk_serve_replay_plain: bad argument");
    if (!cols->a || !cols->b || !cols->ts || !cols->value || cols->rows <= 0 || cols->block_rows <= 0)
        return qk::fail(QK_ERROR, "qk_serve_replay_plain: bad columns");
    for (int32_t i = 1; i < n; i++)
        if (arrival_ms[i] < arrival_ms[i - 1])
            return qk::fail(QK_ERROR, "qk_serve_replay_plain: arrivals must not decrease");
    const int64_t nblocks = (cols->rows + cols->block_rows - 1) / cols->block_rows;
    for (int32_t i = 0; i < n; i++) {
        latency_ms[i] = -1.0;
        if (q[i].key_mod < 1 || q[i].nblocks < 1 || q[i].first_block < 0 || q[i].first_block + (int64_t)q[i].nblocks > nblocks)
            return qk::fail(QK_ERROR, "qk_serve_replay_plain: request %d out of range", i);
    }
    QK_TRY(qk::ensure_init());
    void* keybuf = nullptr;
    QK_TRY(qk::pool_alloc(&keybuf, (size_t)cols->block_rows * 8));
    const int32_t key_type = QK_LONG;
    const int32_t aggs[4] = {QK_AGG_SUM, QK_AGG_MIN, QK_AGG_MAX, QK_AGG_COUNT};
    const auto t0 = Clock::now();
    auto ms_since = [&] { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); };
    int st = QK_SUCCESS;
    for (int32_t i = 0; i < n && st == QK_SUCCESS; i++) {
        const double wait = arrival_ms[i] - ms_since();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double, std::milli>(wait));
        const qk_batch_query& r = q[i];
        qk_groupby h = nullptr;
        st = qk_gb_create(&h, 1, &key_type, QK_LONG, 4, aggs, nullptr, r.key_mod, 1, QK_GB_KEYS_NOT_NULL);
        for (int64_t bi = r.first_block; st == QK_SUCCESS && bi < r.first_block + r.nblocks; bi++) {
            const int64_t off = bi * cols->block_rows;
            const int32_t rows = (int32_t)std::min<int64_t>(cols->block_rows, cols->rows - off);
            st = qk_batch_key_column(cols->a + off, cols->b + off, cols->b_range, r.key_mod, rows,
                                     (int64_t*)keybuf, stream);
            if (st != QK_SUCCESS) break;
            qk_predicate p[2];
            std::memset(p, 0, sizeof(p));
            for (int k = 0; k < 2; k++) {
                p[k].op = k == 0 ? QK_GREATER_EQUAL : QK_LESS;
                p[k].type = QK_LONG;
                p[k].column = cols->ts + off;
                p[k].rhs_kind = QK_CONST;
                p[k].rhs_type = QK_LONG;
                p[k].rhs_int = k == 0 ? r.ts_lo : r.ts_hi;
            }
            const void* kp = keybuf;
            st = qk_gb_process_block_where(h, &kp, nullptr, cols->value + off, nullptr, rows, 2, p, stream);
        }
        void* okeys[1] = {nullptr};
        void* ovals[4] = {nullptr, nullptr, nullptr, nullptr};
        int32_t groups = 0;
        if (st == QK_SUCCESS) st = qk_gb_get_results(h, okeys, nullptr, ovals, nullptr, &groups, stream);
        if (st == QK_SUCCESS) st = qk_stream_synchronize(stream);
        if (st == QK_SUCCESS) latency_ms[i] = ms_since() - arrival_ms[i];
        qk_free(okeys[0]);
        for (void* v : ovals) qk_free(v);
        if (h) qk_gb_destroy(h);
    }
    qk::pool_free(keybuf);
    return st;
}

int qk_serve_get_stats(qk_serve s, qk_serve_stats* out) {
    if (!s || !out) return qk::fail(QK_ERROR, "qk_serve_get_stats: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->batches = s->batches;
    out->served = s->served;
    out->failed = s->failed;
    out->pending = (int64_t)s->pending.size();
    out->max_batch = s->max_batch_seen;
    out->admitted_bytes = s->admitted_bytes;
    out->peak_admitted_bytes = s->peak_bytes;
    out->busy_ms = s->busy_ms;
    return QK_SUCCESS;
}

}  // extern "C"
