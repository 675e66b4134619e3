// olap_server.cpp -- continuous-batching request loop over resident datasets.
//
// Submitted requests queue in arrival order. Each worker thread (the GPU
// executor gives it its own context and non-blocking HIP stream) waits until
// admission can take a batch: the oldest pending request's dataset, further
// requests of that dataset in arrival order while their bytes fit the free
// budget (olap_server_plan). It runs the batch as one fused scan into its
// output staging, copies each request's slice to the request's host buffers
// and releases the budget. Requests arriving while batches run are taken by
// the next idle worker, so `streams` batches overlap on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "olap_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    int dataset = 0;
    olap_batch_query q{};
    int64_t *sum_host = nullptr;
    int64_t *count_host = nullptr;
    int64_t bytes = 0;
    Clock::time_point t_submit;
    double latency_us = 0.0;
    int status = OLAP_OK;
    bool admitted = false;
    bool done = false;
};

// A batch with more expected matches than max(kStagedMatches, kStagedPerRow
// x rows) uses the one-pass event-grouped scan with staged aggregation;
// fewer go to LDS tables plus global atomics. Measured on 1-64 serving
// templates over 6M, 60M and 600M fact rows: the one-pass scan wins from
// about 1e6 (60M rows) and 7e6 matches (600M rows); below about 1e6 its
// fixed cost (~0.24 ms) loses to the atomics layout. It beats two-pass
// staging everywhere (38 against 80 ms at 64 queries over 600M rows), and
// its record pool is smaller than a two-pass buffer, so workers keep one
// staging workspace. Two-pass staging remains the fallback when the
// one-pass layout does not fit (too many partitions for LDS, or one row).
constexpr double kStagedMatches = 2.0e6;
constexpr double kStagedPerRow = 0.015;
constexpr int kCfgAtomics = 0;   // 512 threads, 48 KiB of LDS tables
constexpr int kCfgOnePass = 33;  // 1024 threads, small dimension tables in LDS, one-pass staging, event segments
constexpr int kCfgOnePassEv = 30;  // the same with counted event lists (when segments do not fit)
constexpr int kCfgTwoPass = 9;   // 1024 threads, 48 KiB of LDS tables, two-pass staging

}  // namespace

struct olap_server_s {
    olap_ctx ctx = nullptr;                  // GPU executor: the creating context
    olap_server_exec_fn exec = nullptr;      // host executor
    void *exec_user = nullptr;
    olap_server_alloc_fn alloc_fn = nullptr;
    olap_server_free_fn free_fn = nullptr;
    void *alloc_user = nullptr;
    olap_server_config cfg{};
    std::vector<olap_star> datasets;
    std::mutex mu;
    std::condition_variable cv_work;   // pending requests, budget released, resume, stop
    std::condition_variable cv_done;   // a request finished
    std::deque<Request *> pending;
    std::map<uint64_t, Request *> tickets;
    int64_t bytes_in_use = 0, bytes_peak = 0;
    uint64_t next_ticket = 1;
    bool stop = false;
    bool paused = false;
    bool trace = false;                // OLAP_SERVER_TRACE=1: one stderr line per batch
    std::string err;
    std::vector<std::thread> workers;
    // statistics
    int64_t submitted = 0, completed = 0, failed = 0, rejected = 0, cancelled = 0, batches = 0;
    int64_t batched_requests = 0;
    std::vector<double> latency_us;
};

namespace {

bool gpu_mode(olap_server s) { return s->exec == nullptr; }

bool query_valid(const olap_star *star, const olap_batch_query *q, std::string *why)
{
    if (q->ndims < 1 || q->ndims > 4 || q->group_nums < 1 || q->group_nums > INT32_MAX) {
        *why = "ndims 1..4 and group_nums in [1, 2^31) required";
        return false;
    }
    uint32_t seen = 0;
    int64_t max_gid = 0;
    for (int s = 0; s < q->ndims; s++) {
        const int t = q->table[s];
        if (t < 0 || t > 3 || ((seen >> t) & 1u) || q->group_mod[s] < 0 || q->factor[s] < 0) {
            *why = "distinct tables 0..3, group_mod >= 0, factor >= 0";
            return false;
        }
        if (!star->filter[t] || !star->fk[t] || (q->group_mod[s] > 0 && !star->attr[t])) {
            *why = "the dataset lacks a probed table";
            return false;
        }
        seen |= 1u << t;
        if (q->group_mod[s] > 1) max_gid += (int64_t)(q->group_mod[s] - 1) * q->factor[s];
    }
    if (max_gid >= q->group_nums) {
        *why = "largest gid reaches group_nums";
        return false;
    }
    return true;
}

// One worker's output staging: sums then counts of up to `slots` groups in
// one allocation, grown on demand and kept across batches; device memory
// for the GPU executor, host memory for a host executor. The GPU executor
// also keeps two pinned host mirrors: a batch's results come back into one
// in two DMA copies while the worker's slicer still copies the previous
// batch's requests out of the other.
struct Staging {
    olap_server s = nullptr;
    hipStream_t stream = nullptr;   // GPU executor: the stream whose copies read buf
    int64_t *buf = nullptr;     // [2 * slots]: sums, then counts
    size_t slots = 0;
    int64_t *host[2] = {nullptr, nullptr};  // pinned mirrors (GPU executor)
    size_t host_slots[2] = {0, 0};

    int64_t *sum() const { return buf; }
    int64_t *count() const { return buf + slots; }
    void *alloc(size_t bytes)
    {
        if (s->alloc_fn) return s->alloc_fn(s->alloc_user, bytes);
        if (!gpu_mode(s)) return std::malloc(bytes);
        void *p = nullptr;
        return hipMalloc(&p, bytes) == hipSuccess ? p : nullptr;
    }
    void dealloc(void *p)
    {
        if (!p) return;
        if (s->free_fn) s->free_fn(s->alloc_user, p);
        else if (!gpu_mode(s)) std::free(p);
        else (void)hipFree(p);
    }
    // the device (or host-executor) buffer for `want` slots; a GPU buffer
    // is replaced only after the copies still reading it have finished
    bool fit(size_t want)
    {
        if (want > slots) {
            if (gpu_mode(s) && buf && stream && hipStreamSynchronize(stream) != hipSuccess) return false;
            dealloc(buf);
            slots = 0;
            buf = (int64_t *)alloc(16 * want);
            if (!buf) return false;
            slots = want;
        }
        return true;
    }
    // pinned mirror k for `want` slots (the caller owns mirror k: no copy and
    // no slicer uses it)
    bool fit_host(int k, size_t want)
    {
        if (want <= host_slots[k]) return true;
        if (host[k]) (void)hipHostFree(host[k]);
        host_slots[k] = 0;
        if (hipHostMalloc((void **)&host[k], 16 * want, hipHostMallocDefault) != hipSuccess) {
            host[k] = nullptr;
            return false;
        }
        host_slots[k] = want;
        return true;
    }
    void release()
    {
        dealloc(buf);
        buf = nullptr;
        slots = 0;
        for (int k = 0; k < 2; k++) {
            if (host[k]) (void)hipHostFree(host[k]);
            host[k] = nullptr;
            host_slots[k] = 0;
        }
    }
};

void finish(olap_server s, const std::vector<Request *> &batch, int status);

// A GPU worker's slicer thread. The worker hands it a batch whose results
// are being copied into pinned mirror k (an event marks the end of the
// copies) and goes on to the next batch; the slicer waits for the event,
// copies each request's slice to its host buffers and finishes the batch.
// Mirror k is the worker's again once the slicer is done with it.
struct SliceJob {
    std::vector<Request *> batch;
    std::vector<int64_t> off;
    int mirror = 0;
    int worker = 0, dataset = 0, cfg = -1;
    double create_ms = 0.0, setup_ms = 0.0, scan_ms = 0.0;
    std::chrono::steady_clock::time_point t_launched;
};

class Slicer {
  public:
    Slicer(olap_server s, Staging *out, int device) : s_(s), out_(out), device_(device)
    {
        for (int k = 0; k < 2; k++)
            if (hipEventCreateWithFlags(&ev_[k], hipEventDisableTiming) != hipSuccess) ev_[k] = nullptr;
        th_ = std::thread([this] { run(); });
    }
    ~Slicer()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        th_.join();
        for (int k = 0; k < 2; k++)
            if (ev_[k]) (void)hipEventDestroy(ev_[k]);
    }
    bool ok() const { return ev_[0] && ev_[1]; }
    hipEvent_t event(int k) const { return ev_[k]; }
    // block until mirror k is free (no job of the slicer reads it)
    void wait_free(int k)
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !busy_[k]; });
    }
    void submit(SliceJob &&job)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            busy_[job.mirror] = true;
            jobs_.push_back(std::move(job));
        }
        cv_.notify_all();
    }

  private:
    void run();
    olap_server s_;
    Staging *out_;
    int device_;
    hipEvent_t ev_[2] = {nullptr, nullptr};
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<SliceJob> jobs_;
    bool busy_[2] = {false, false};
    bool stop_ = false;
    std::thread th_;
};

struct Worker {
    int id = 0;
    olap_ctx ctx = nullptr;
    hipStream_t stream = nullptr;
    olap::BatchWorkspace *ws = nullptr;   // kept across batches: no steady-state allocation
    Staging out;
    int setup = OLAP_OK;
    Slicer *slicer = nullptr;             // GPU executor: copies results out, finishes requests
    int next_mirror = 0;
};

void set_err(olap_server s, const std::string &msg)
{
    std::lock_guard<std::mutex> lk(s->mu);
    s->err = msg;
}

// a GPU batch handed to the worker's slicer (it finishes the requests)
constexpr int kHandedOff = 1 << 30;

// run a batch on the GPU executor: the fused scan into the staging and the
// copies of its results into a free pinned mirror, then the batch goes to
// the slicer; returns kHandedOff, or an error status (the batch is then
// finished by the caller)
int run_gpu(olap_server s, Worker &w, const std::vector<Request *> &batch,
            const std::vector<olap_batch_query> &qs, const std::vector<int64_t> &off)
{
    const olap_star &star = s->datasets[(size_t)batch[0]->dataset];
    const Clock::time_point t0 = Clock::now();
    olap_batch b = nullptr;
    // the previous batch's kernels, which read the workspace this setup may
    // grow, have finished: the worker synchronised its stream before admission
    int st = olap::batch_create_in(w.ctx, &star, qs.data(), (int32_t)qs.size(), w.ws, w.stream, &b);
    const Clock::time_point tc = Clock::now();
    int cfg = -1;
    if (st == OLAP_OK) {
        const double matches = olap_batch_expected_events(b) * (double)star.rows;
        cfg = matches > std::max(kStagedMatches, kStagedPerRow * (double)star.rows) ? kCfgOnePass : kCfgAtomics;
        st = olap_batch_set_config(b, cfg);
        if (st == OLAP_E_CAPACITY && cfg == kCfgOnePass) st = olap_batch_set_config(b, cfg = kCfgOnePassEv);
        if (st == OLAP_E_CAPACITY && cfg == kCfgOnePassEv) st = olap_batch_set_config(b, cfg = kCfgTwoPass);
    }
    const size_t total = (size_t)off.back();
    if (st == OLAP_OK && !w.out.fit(std::max<size_t>(total, 1))) st = OLAP_E_NOMEM;
    const Clock::time_point t1 = Clock::now();
    if (st == OLAP_OK) st = olap_batch_run(b, w.out.sum(), w.out.count(), w.stream);
    const Clock::time_point t2 = Clock::now();
    // two DMA copies into a free mirror, an event after them, then the slicer
    const int k = w.next_mirror;
    if (st == OLAP_OK) {
        w.slicer->wait_free(k);
        if (!w.out.fit_host(k, std::max<size_t>(total, 1))) st = OLAP_E_NOMEM;
    }
    if (st == OLAP_OK &&
        (hipMemcpyAsync(w.out.host[k], w.out.sum(), 8 * total, hipMemcpyDeviceToHost, w.stream) != hipSuccess ||
         hipMemcpyAsync(w.out.host[k] + total, w.out.count(), 8 * total, hipMemcpyDeviceToHost, w.stream) !=
             hipSuccess ||
         hipEventRecord(w.slicer->event(k), w.stream) != hipSuccess))
        st = OLAP_E_HIP;
    if (b) olap_batch_destroy(b);
    if (st != OLAP_OK) {
        const char *m = olap_last_error(w.ctx);
        set_err(s, std::string("server batch: ") + (st == OLAP_E_NOMEM && (!m || !*m)
                                                        ? "output staging allocation failed"
                                                        : (m ? m : "")));
        return st;
    }
    auto ms = [](Clock::time_point a, Clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    SliceJob job;
    job.batch = batch;
    job.off = off;
    job.mirror = k;
    job.worker = w.id;
    job.dataset = batch[0]->dataset;
    job.cfg = cfg;
    job.create_ms = ms(t0, tc);
    job.setup_ms = ms(t0, t1);
    job.scan_ms = ms(t1, t2);
    job.t_launched = Clock::now();
    w.slicer->submit(std::move(job));
    w.next_mirror ^= 1;
    return kHandedOff;
}

void Slicer::run()
{
    (void)hipSetDevice(device_);
    for (;;) {
        SliceJob job;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
            if (jobs_.empty()) return;  // stopping, nothing left
            job = std::move(jobs_.front());
            jobs_.pop_front();
        }
        int st = hipEventSynchronize(ev_[job.mirror]) == hipSuccess ? OLAP_OK : OLAP_E_HIP;
        const Clock::time_point t3 = Clock::now();
        const int64_t *h = out_->host[job.mirror];
        const size_t total = (size_t)job.off.back();
        for (size_t i = 0; i < job.batch.size() && st == OLAP_OK; i++) {
            const size_t n = (size_t)(job.off[i + 1] - job.off[i]);
            std::memcpy(job.batch[i]->sum_host, h + job.off[i], 8 * n);
            std::memcpy(job.batch[i]->count_host, h + total + job.off[i], 8 * n);
        }
        if (st != OLAP_OK) set_err(s_, "server batch: the result copies failed");
        if (s_->trace) {
            auto ms = [](Clock::time_point a, Clock::time_point b) {
                return std::chrono::duration<double, std::milli>(b - a).count();
            };
            std::fprintf(stderr,
                         "olap_server: worker %d dataset %d queries %zu config %d slots %zu: setup %.2f ms "
                         "(create %.2f), scan %.2f ms, results %.2f ms, slices %.2f ms, status %d\n",
                         job.worker, job.dataset, job.batch.size(), job.cfg, total, job.setup_ms, job.create_ms,
                         job.scan_ms,
                         ms(job.t_launched, t3), ms(t3, Clock::now()), st);
        }
        finish(s_, job.batch, st);
        {
            std::lock_guard<std::mutex> lk(mu_);
            busy_[job.mirror] = false;
        }
        cv_.notify_all();
    }
}

// run a batch on a host executor
int run_host(olap_server s, Worker &w, const std::vector<Request *> &batch,
             const std::vector<olap_batch_query> &qs, const std::vector<int64_t> &off)
{
    if (!w.out.fit(std::max<size_t>((size_t)off.back(), 1))) {
        set_err(s, "server batch: output staging allocation failed");
        return OLAP_E_NOMEM;
    }
    const int st = s->exec(s->exec_user, batch[0]->dataset, qs.data(), (int32_t)qs.size(), off.data(),
                           w.out.sum(), w.out.count());
    if (st != OLAP_OK) {
        set_err(s, "server batch: the executor failed with status " + std::to_string(st));
        return st;
    }
    for (size_t i = 0; i < batch.size(); i++) {
        const size_t n = (size_t)(off[i + 1] - off[i]);
        std::memcpy(batch[i]->sum_host, w.out.sum() + off[i], 8 * n);
        std::memcpy(batch[i]->count_host, w.out.count() + off[i], 8 * n);
    }
    return OLAP_OK;
}

void finish(olap_server s, const std::vector<Request *> &batch, int status)
{
    const Clock::time_point now = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    for (Request *r : batch) {
        r->status = status;
        r->done = true;
        s->bytes_in_use -= r->bytes;
        r->latency_us = std::chrono::duration<double, std::micro>(now - r->t_submit).count();
        if (status == OLAP_OK) {
            s->latency_us.push_back(r->latency_us);
            s->completed++;
        } else {
            s->failed++;
        }
    }
    s->batches++;
    s->batched_requests += (int64_t)batch.size();
    s->cv_done.notify_all();
    s->cv_work.notify_all();
}

// under s->mu: move the next admissible batch out of the pending queue
void admit(olap_server s, std::vector<Request *> &batch)
{
    std::vector<int32_t> ds, chosen(s->pending.size());
    std::vector<int64_t> bytes;
    for (Request *r : s->pending) {
        ds.push_back(r->dataset);
        bytes.push_back(r->bytes);
    }
    const int32_t n = olap_server_plan(ds.data(), bytes.data(), (int32_t)ds.size(),
                                       s->cfg.memory_budget - s->bytes_in_use, s->cfg.max_batch,
                                       chosen.data());
    if (n <= 0) return;
    // one batch holds fewer than 2^31 output slots
    std::vector<bool> take(s->pending.size(), false);
    int64_t slots = 0;
    for (int32_t i = 0; i < n; i++) {
        const int64_t g = s->pending[(size_t)chosen[(size_t)i]]->q.group_nums;
        if (i > 0 && slots + g > INT32_MAX) break;
        slots += g;
        take[(size_t)chosen[(size_t)i]] = true;
    }
    std::deque<Request *> rest;
    for (size_t i = 0; i < s->pending.size(); i++) {
        if (take[i]) batch.push_back(s->pending[i]);
        else rest.push_back(s->pending[i]);
    }
    s->pending.swap(rest);
    for (Request *r : batch) {
        r->admitted = true;
        s->bytes_in_use += r->bytes;
    }
    s->bytes_peak = std::max(s->bytes_peak, s->bytes_in_use);
}

void worker_main(olap_server s, int id)
{
    Worker w;
    w.id = id;
    w.out.s = s;
    if (gpu_mode(s)) {
        w.ws = olap::batch_workspace_create();
        if (hipSetDevice(s->ctx->device) != hipSuccess || olap_ctx_create(s->ctx->device, &w.ctx) != OLAP_OK ||
            hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking) != hipSuccess) {
            w.setup = OLAP_E_HIP;
            set_err(s, "server worker: device context or stream creation failed");
        } else {
            w.out.stream = w.stream;
            w.slicer = new Slicer(s, &w.out, s->ctx->device);
            if (!w.slicer->ok()) {
                w.setup = OLAP_E_HIP;
                set_err(s, "server worker: event creation failed");
            }
        }
    }
    for (;;) {
        std::vector<Request *> batch;
        // a GPU worker takes new requests only once its previous batch has
        // left the device (its slicer may still copy the results out), so
        // requests wait in the shared queue while both workers are busy
        if (w.stream) (void)hipStreamSynchronize(w.stream);
        {
            std::unique_lock<std::mutex> lk(s->mu);
            while (!s->stop) {
                if (!s->paused && !s->pending.empty()) admit(s, batch);
                if (!batch.empty()) break;
                s->cv_work.wait(lk);
            }
            if (batch.empty()) break;  // stopping
        }
        std::vector<olap_batch_query> qs;
        std::vector<int64_t> off(1, 0);
        for (Request *r : batch) {
            qs.push_back(r->q);
            off.push_back(off.back() + r->q.group_nums);
        }
        int st = w.setup;
        if (st == OLAP_OK) st = gpu_mode(s) ? run_gpu(s, w, batch, qs, off) : run_host(s, w, batch, qs, off);
        if (st != kHandedOff) finish(s, batch, st);
    }
    delete w.slicer;  // finishes the batches handed to it first
    w.slicer = nullptr;
    if (w.stream) (void)hipStreamSynchronize(w.stream);
    w.out.release();
    if (w.ws) olap::batch_workspace_destroy(w.ws);
    if (w.stream) (void)hipStreamDestroy(w.stream);
    if (w.ctx) olap_ctx_destroy(w.ctx);
}

bool config_valid(const olap_server_config *cfg)
{
    return cfg && cfg->max_batch >= 1 && cfg->max_batch <= OLAP_BATCH_MAX && cfg->streams >= 1 &&
           cfg->streams <= 8 && cfg->memory_budget > 0;
}

olap_server start(olap_server s)
{
    const char *tr = std::getenv("OLAP_SERVER_TRACE");
    s->trace = tr && *tr && *tr != '0';
    for (int i = 0; i < s->cfg.streams; i++) s->workers.emplace_back(worker_main, s, i);
    return s;
}

}  // namespace

extern "C" {

int64_t olap_server_request_bytes(const olap_star *star, const olap_batch_query *q)
{
    if (!star || !q || q->ndims < 1 || q->ndims > 4 || q->group_nums < 0) return -1;
    int64_t b = 16 * q->group_nums;
    for (int s = 0; s < q->ndims; s++) {
        const int t = q->table[s];
        if (t < 0 || t > 3) return -1;
        b += 4 * star->dim_rows[t];
    }
    return b;
}

int32_t olap_server_plan(const int32_t *datasets, const int64_t *bytes, int32_t n,
                         int64_t budget_left, int32_t max_batch, int32_t *chosen)
{
    if (!datasets || !bytes || !chosen || n <= 0 || max_batch <= 0) return 0;
    if (bytes[0] > budget_left) return 0;  // the oldest waits for memory
    int32_t k = 0;
    int64_t left = budget_left;
    for (int32_t i = 0; i < n && k < max_batch; i++) {
        if (datasets[i] != datasets[0] || bytes[i] > left) continue;
        chosen[k++] = i;
        left -= bytes[i];
    }
    return k;
}

olap_status olap_server_create(olap_ctx ctx, const olap_server_config *cfg, olap_server *out)
{
    if (!ctx || !out) return OLAP_E_INVALID;
    if (!config_valid(cfg)) {
        ctx->err = "server_create: max_batch 1..64, streams 1..8, memory_budget > 0";
        return OLAP_E_INVALID;
    }
    olap_server s = new olap_server_s();
    s->ctx = ctx;
    s->cfg = *cfg;
    *out = start(s);
    return OLAP_OK;
}

olap_status olap_server_create_host(const olap_server_config *cfg, olap_server_exec_fn exec,
                                    void *exec_user, olap_server *out)
{
    if (!exec || !out || !config_valid(cfg)) return OLAP_E_INVALID;
    olap_server s = new olap_server_s();
    s->exec = exec;
    s->exec_user = exec_user;
    s->cfg = *cfg;
    *out = start(s);
    return OLAP_OK;
}

olap_status olap_server_set_allocator(olap_server s, olap_server_alloc_fn alloc_fn,
                                      olap_server_free_fn free_fn, void *user)
{
    if (!s || !alloc_fn != !free_fn) return OLAP_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->submitted) {
        s->err = "server_set_allocator: set it before the first submit";
        return OLAP_E_BUSY;
    }
    s->alloc_fn = alloc_fn;
    s->free_fn = free_fn;
    s->alloc_user = user;
    return OLAP_OK;
}

olap_status olap_server_add_dataset(olap_server s, const olap_star *star, int32_t *dataset_id)
{
    if (!s || !star || !dataset_id) return OLAP_E_INVALID;
    auto aligned = [](const void *p) { return ((uintptr_t)p & 15u) == 0; };
    bool ok = star->rows >= 0 && star->rows <= (int64_t)UINT32_MAX - 65536 && (star->m1 || !star->rows) &&
              aligned(star->m1) && aligned(star->m2);
    for (int t = 0; t < 4; t++)
        ok = ok && aligned(star->fk[t]) && star->dim_rows[t] >= 0 && star->dim_rows[t] <= INT32_MAX;
    std::lock_guard<std::mutex> lk(s->mu);
    if (!ok) {
        s->err = "server_add_dataset: rows < 2^32 - 65536, m1 required, 16-byte aligned fact columns";
        return OLAP_E_INVALID;
    }
    s->datasets.push_back(*star);
    *dataset_id = (int32_t)s->datasets.size() - 1;
    return OLAP_OK;
}

olap_status olap_server_submit(olap_server s, int32_t dataset, const olap_batch_query *q,
                               int64_t *sum_host, int64_t *count_host, uint64_t *ticket)
{
    if (!s || !q || !sum_host || !count_host || !ticket) return OLAP_E_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (dataset < 0 || dataset >= (int32_t)s->datasets.size()) {
            s->err = "server_submit: unknown dataset";
            return OLAP_E_INVALID;
        }
        std::string why;
        if (!query_valid(&s->datasets[(size_t)dataset], q, &why)) {
            s->err = "server_submit: " + why;
            s->rejected++;
            return OLAP_E_INVALID;
        }
        const int64_t bytes = olap_server_request_bytes(&s->datasets[(size_t)dataset], q);
        if (bytes > s->cfg.memory_budget) {
            s->err = "server_submit: the query alone exceeds the memory budget";
            s->rejected++;
            return OLAP_E_CAPACITY;
        }
        Request *r = new Request();
        r->ticket = s->next_ticket++;
        r->dataset = dataset;
        r->q = *q;
        r->sum_host = sum_host;
        r->count_host = count_host;
        r->bytes = bytes;
        r->t_submit = Clock::now();
        s->pending.push_back(r);
        s->tickets[r->ticket] = r;
        s->submitted++;
        *ticket = r->ticket;
    }
    s->cv_work.notify_one();
    return OLAP_OK;
}

olap_status olap_server_wait(olap_server s, uint64_t ticket, int64_t timeout_us, int32_t *status,
                             double *latency_us)
{
    if (!s) return OLAP_E_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return OLAP_E_INVALID;
    Request *r = it->second;
    auto pred = [&] { return r->done; };
    if (timeout_us < 0) s->cv_done.wait(lk, pred);
    else if (!s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), pred)) return OLAP_E_BUSY;
    if (status) *status = r->status;
    if (latency_us) *latency_us = r->latency_us;
    s->tickets.erase(r->ticket);
    delete r;
    return OLAP_OK;
}

olap_status olap_server_pause(olap_server s, int32_t paused)
{
    if (!s) return OLAP_E_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->paused = paused != 0;
    }
    s->cv_work.notify_all();
    return OLAP_OK;
}

olap_status olap_server_cancel(olap_server s, uint64_t ticket)
{
    if (!s) return OLAP_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return OLAP_E_INVALID;
    Request *r = it->second;
    if (r->admitted) {
        s->err = "server_cancel: the request is already in a batch";
        return OLAP_E_BUSY;
    }
    s->pending.erase(std::remove(s->pending.begin(), s->pending.end(), r), s->pending.end());
    s->tickets.erase(it);
    s->cancelled++;
    delete r;
    return OLAP_OK;
}

olap_status olap_server_get_stats(olap_server s, olap_server_stats *out)
{
    if (!s || !out) return OLAP_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    olap_server_stats st{};
    st.submitted = s->submitted;
    st.completed = s->completed;
    st.failed = s->failed;
    st.rejected = s->rejected;
    st.cancelled = s->cancelled;
    st.batches = s->batches;
    st.mean_batch = s->batches ? (double)s->batched_requests / (double)s->batches : 0.0;
    std::vector<double> l = s->latency_us;
    if (!l.empty()) {
        std::sort(l.begin(), l.end());
        // nearest rank
        auto pct = [&](double p) {
            const size_t i = (size_t)std::min<double>((double)l.size() - 1, p * (double)(l.size() - 1) + 0.5);
            return l[i];
        };
        st.p50_us = pct(0.50);
        st.p99_us = pct(0.99);
        st.max_us = l.back();
        double acc = 0;
        for (double v : l) acc += v;
        st.mean_us = acc / (double)l.size();
    }
    st.bytes_in_use = s->bytes_in_use;
    st.bytes_peak = s->bytes_peak;
    *out = st;
    return OLAP_OK;
}

olap_status olap_server_reset_stats(olap_server s)
{
    if (!s) return OLAP_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    s->submitted = s->completed = s->failed = s->rejected = s->cancelled = s->batches = 0;
    s->batched_requests = 0;
    s->latency_us.clear();
    s->bytes_peak = s->bytes_in_use;
    return OLAP_OK;
}

const char *olap_server_last_error(olap_server s)
{
    if (!s) return "null server";
    std::lock_guard<std::mutex> lk(s->mu);
    return s->err.c_str();
}

olap_status olap_server_destroy(olap_server s)
{
    if (!s) return OLAP_E_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (std::thread &t : s->workers) t.join();
    for (auto &kv : s->tickets) delete kv.second;
    delete s;
    return OLAP_OK;
}

}  // extern "C"
