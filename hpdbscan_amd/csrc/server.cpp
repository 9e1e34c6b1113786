// Continuous-batching clustering server (hpd_server_* of hpdbscan_capi.h).
//
// Requests (one point cloud each, own eps / minPts) are queued by
// hpd_server_submit.  A scheduler thread groups waiting requests of the same
// dtype and dimension into a batch and hands it to the first idle lane.  A
// lane owns an hpd_context (its own HIP stream and device workspace) and runs
// the batch through hpd_cluster_batch, so the batches of different lanes
// overlap on the device.  Batches are formed as lanes free up: a request that
// arrives while others run joins the next batch (continuous batching, not a
// queue of fixed batches).
//
// Admission: every request is charged an estimate of the device memory its
// share of a batch needs (rows * bytes_per_row); a batch is dispatched only
// while the charges of all running batches plus its own stay within the
// budget.  A request that cannot fit even alone fails with HPD_ENOMEM.
//
// Data movement: a lane keeps grow-only device buffers for a batch's rows
// and labels and pinned host buffers to stage them.  A host request's rows
// go host -> pinned -> device and its labels back the same way; a device
// request (hpd_server_submit_device) has its rows copied device to device
// into the batch and its labels copied device to device into the caller's
// buffer, so its data never leaves HBM.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hpdbscan_capi.h"

namespace {

using Clock = std::chrono::steady_clock;

size_t dtype_bytes(int dt) {
    switch (dt) {
        case HPD_DTYPE_F32: return 4;
        case HPD_DTYPE_F64: return 8;
        case HPD_DTYPE_I8: case HPD_DTYPE_U8: return 1;
        case HPD_DTYPE_I16: case HPD_DTYPE_U16: return 2;
        case HPD_DTYPE_I32: case HPD_DTYPE_U32: return 4;
        case HPD_DTYPE_I64: case HPD_DTYPE_U64: return 8;
        default: return 0;
    }
}

struct Request {
    uint64_t ticket = 0;
    int dtype = 0;
    int32_t d = 0;
    int64_t n = 0;
    float eps = 0.0f;
    uint64_t min_pts = 0;
    std::vector<char> rows;       // copy of the caller's rows (host requests)
    std::vector<int64_t> labels;  // result (host requests)
    const void* d_rows = nullptr; // device requests: the caller's rows ...
    int64_t* d_labels = nullptr;  // ... and labels, in device memory
    bool device = false;
    size_t charge = 0;            // device bytes reserved for it
    int status = HPD_OK;
    std::string error;
    bool done = false;
    Clock::time_point t_submit, t_start, t_done;
    int batch_jobs = 0;           // size of the batch it ran in
};

struct Lane {
    hpd_context* ctx = nullptr;
    std::thread thread;
    std::vector<std::shared_ptr<Request>> work;  // current batch
    bool busy = false;
    bool has_work = false;
    // grow-only batch buffers: device rows / labels, pinned staging
    void* d_rows = nullptr;
    int64_t* d_labels = nullptr;
    char* h_rows = nullptr;
    int64_t* h_labels = nullptr;
    size_t rows_cap = 0, labels_cap = 0, h_rows_cap = 0, h_labels_cap = 0;
};

// Grow a device (pinned: host) buffer to at least `bytes` (with a quarter
// of headroom, so a stream of growing batches reallocates rarely).
hipError_t grow(void** p, size_t* cap, size_t bytes, bool pinned) {
    if (bytes <= *cap) return hipSuccess;
    const size_t want = bytes + bytes / 4;
    if (*p) {
        if (pinned)
            (void)hipHostFree(*p);
        else
            (void)hipFree(*p);
        *p = nullptr;
        *cap = 0;
    }
    const hipError_t e = pinned ? hipHostMalloc(p, want, hipHostMallocDefault) : hipMalloc(p, want);
    if (e == hipSuccess) *cap = want;
    return e;
}

void release(Lane& l) {
    if (l.d_rows) (void)hipFree(l.d_rows);
    if (l.d_labels) (void)hipFree(l.d_labels);
    if (l.h_rows) (void)hipHostFree(l.h_rows);
    if (l.h_labels) (void)hipHostFree(l.h_labels);
    l.d_rows = nullptr;
    l.d_labels = nullptr;
    l.h_rows = nullptr;
    l.h_labels = nullptr;
}

}  // namespace

struct hpd_server {
    hpd_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_sched;   // scheduler: new request, lane free, stop
    std::condition_variable cv_lane;    // lanes: work assigned, stop
    std::condition_variable cv_done;    // waiters: a request finished
    std::deque<std::shared_ptr<Request>> queue;
    std::map<uint64_t, std::shared_ptr<Request>> table;
    std::vector<std::unique_ptr<Lane>> lanes;
    std::thread scheduler;
    bool stop = false;
    uint64_t next_ticket = 1;
    size_t charged = 0;
    // statistics
    std::vector<double> latencies_ms;  // completed requests
    uint64_t completed = 0, failed = 0, batches = 0, batched_jobs = 0;
    int64_t rows_done = 0;
    Clock::time_point t_first_submit, t_last_done;
    bool any_submit = false;
};

namespace {

thread_local std::string g_server_error;

int fail(int code, const std::string& msg) {
    g_server_error = msg;
    return code;
}

size_t request_charge(const hpd_server& s, const Request& r) {
    return static_cast<size_t>(r.n) * s.cfg.bytes_per_row_estimate +
           static_cast<size_t>(r.n) * static_cast<size_t>(r.d) * dtype_bytes(r.dtype);
}

// One batched call for the lane's batch: rows into the lane's device
// buffer (host requests through the pinned buffer, device requests device
// to device), hpd_cluster_batch_device, labels out (host requests into the
// pinned labels, device requests into the caller's buffers).
int run_batch(hpd_server* srv, Lane* lane, const std::vector<std::shared_ptr<Request>>& work,
              const std::vector<int64_t>& offsets, const std::vector<float>& eps,
              const std::vector<uint64_t>& mps, int dtype, int32_t d, size_t row_bytes,
              int64_t total, std::string* err) {
    auto hip_fail = [&](hipError_t e, const char* what) {
        *err = std::string(what) + ": " + hipGetErrorString(e);
        return HPD_EDEVICE;
    };
    hipError_t e = hipSetDevice(srv->cfg.device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    void* sp = nullptr;
    if (hpd_context_stream(lane->ctx, &sp) != HPD_OK) {
        *err = hpd_last_error();
        return HPD_EDEVICE;
    }
    hipStream_t stream = static_cast<hipStream_t>(sp);
    const size_t nrows = static_cast<size_t>(total);
    bool any_host = false;
    for (const auto& r : work) any_host = any_host || !r->device;
    if ((e = grow(&lane->d_rows, &lane->rows_cap, std::max<size_t>(1, nrows * row_bytes), false)) !=
            hipSuccess ||
        (e = grow(reinterpret_cast<void**>(&lane->d_labels), &lane->labels_cap,
                  std::max<size_t>(1, nrows) * sizeof(int64_t), false)) != hipSuccess)
        return hip_fail(e, "hipMalloc of the batch buffers");
    if (any_host &&
        ((e = grow(reinterpret_cast<void**>(&lane->h_rows), &lane->h_rows_cap,
                   std::max<size_t>(1, nrows * row_bytes), true)) != hipSuccess ||
         (e = grow(reinterpret_cast<void**>(&lane->h_labels), &lane->h_labels_cap,
                   std::max<size_t>(1, nrows) * sizeof(int64_t), true)) != hipSuccess))
        return hip_fail(e, "hipHostMalloc of the staging buffers");
    char* drows = static_cast<char*>(lane->d_rows);
    for (size_t j = 0; j < work.size(); ++j) {
        const Request& r = *work[j];
        const size_t off = static_cast<size_t>(offsets[j]) * row_bytes;
        const size_t bytes = static_cast<size_t>(r.n) * row_bytes;
        if (bytes == 0) continue;
        if (r.device) {
            e = hipMemcpyAsync(drows + off, r.d_rows, bytes, hipMemcpyDeviceToDevice, stream);
        } else {
            std::memcpy(lane->h_rows + off, r.rows.data(), bytes);
            e = hipMemcpyAsync(drows + off, lane->h_rows + off, bytes, hipMemcpyHostToDevice,
                               stream);
        }
        if (e != hipSuccess) return hip_fail(e, "copy of a request's rows");
    }
    const int rc = hpd_cluster_batch_device(lane->ctx, dtype, lane->d_rows, offsets.data(),
                                            static_cast<int32_t>(work.size()), d, eps.data(),
                                            mps.data(), lane->d_labels, stream);
    if (rc != HPD_OK) {
        *err = hpd_last_error();
        return rc;
    }
    for (size_t j = 0; j < work.size(); ++j) {
        const Request& r = *work[j];
        const size_t bytes = static_cast<size_t>(r.n) * sizeof(int64_t);
        if (bytes == 0) continue;
        const int64_t* src = lane->d_labels + offsets[j];
        e = r.device ? hipMemcpyAsync(r.d_labels, src, bytes, hipMemcpyDeviceToDevice, stream)
                     : hipMemcpyAsync(lane->h_labels + offsets[j], src, bytes,
                                      hipMemcpyDeviceToHost, stream);
        if (e != hipSuccess) return hip_fail(e, "copy of a request's labels");
    }
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    return HPD_OK;
}

void lane_main(hpd_server* srv, Lane* lane) {
    for (;;) {
        std::vector<std::shared_ptr<Request>> work;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_lane.wait(lk, [&] { return srv->stop || lane->has_work; });
            if (!lane->has_work && srv->stop) return;
            work.swap(lane->work);
            lane->has_work = false;
        }
        // one batched call for the whole group
        const int dtype = work.front()->dtype;
        const int32_t d = work.front()->d;
        const size_t row_bytes = dtype_bytes(dtype) * static_cast<size_t>(d);
        std::vector<int64_t> offsets(work.size() + 1, 0);
        std::vector<float> eps(work.size());
        std::vector<uint64_t> mps(work.size());
        for (size_t j = 0; j < work.size(); ++j) {
            offsets[j + 1] = offsets[j] + work[j]->n;
            eps[j] = work[j]->eps;
            mps[j] = work[j]->min_pts;
        }
        const int64_t total = offsets.back();
        const Clock::time_point t0 = Clock::now();
        std::string err;
        const int rc = run_batch(srv, lane, work, offsets, eps, mps, dtype, d, row_bytes, total,
                                 &err);
        const Clock::time_point t1 = Clock::now();
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            for (size_t j = 0; j < work.size(); ++j) {
                Request& r = *work[j];
                r.t_start = t0;
                r.t_done = t1;
                r.batch_jobs = static_cast<int>(work.size());
                r.status = rc;
                r.error = err;
                if (rc == HPD_OK && !r.device)
                    r.labels.assign(lane->h_labels + offsets[j], lane->h_labels + offsets[j + 1]);
                std::vector<char>().swap(r.rows);
                r.done = true;
                srv->charged -= r.charge;
                if (rc == HPD_OK) {
                    ++srv->completed;
                    srv->rows_done += r.n;
                    srv->latencies_ms.push_back(
                        std::chrono::duration<double, std::milli>(t1 - r.t_submit).count());
                } else {
                    ++srv->failed;
                }
            }
            ++srv->batches;
            srv->batched_jobs += work.size();
            srv->t_last_done = t1;
            lane->busy = false;
        }
        srv->cv_done.notify_all();
        srv->cv_sched.notify_one();
    }
}

// Take the next batch for an idle lane: the oldest waiting request and the
// following ones of the same dtype and dimension, within the job / row caps
// and the memory budget.  Called with the lock held.
std::vector<std::shared_ptr<Request>> form_batch(hpd_server* srv) {
    std::vector<std::shared_ptr<Request>> batch;
    if (srv->queue.empty()) return batch;
    const auto head = srv->queue.front();
    int64_t rows = 0;
    size_t charge = 0;
    for (auto it = srv->queue.begin(); it != srv->queue.end();) {
        const auto& r = *it;
        if (r->dtype != head->dtype || r->d != head->d) {
            ++it;
            continue;
        }
        if (static_cast<int32_t>(batch.size()) >= srv->cfg.max_batch_jobs) break;
        if (!batch.empty() && rows + r->n > srv->cfg.max_batch_rows) {
            ++it;
            continue;
        }
        if (srv->charged + charge + r->charge > srv->cfg.hbm_budget_bytes) {
            if (batch.empty()) return batch;  // wait for memory to free up
            ++it;
            continue;
        }
        rows += r->n;
        charge += r->charge;
        batch.push_back(r);
        it = srv->queue.erase(it);
    }
    srv->charged += charge;
    return batch;
}

void scheduler_main(hpd_server* srv) {
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
        srv->cv_sched.wait(lk, [&] {
            if (srv->stop) return true;
            if (srv->queue.empty()) return false;
            for (auto& l : srv->lanes)
                if (!l->busy) return true;
            return false;
        });
        if (srv->stop) return;
        bool dispatched = false;
        for (auto& l : srv->lanes) {
            if (l->busy) continue;
            auto batch = form_batch(srv);
            if (batch.empty()) break;
            l->work = std::move(batch);
            l->busy = true;
            l->has_work = true;
            dispatched = true;
        }
        if (dispatched) {
            srv->cv_lane.notify_all();
        } else {
            // the oldest request does not fit next to the running batches:
            // wait until one of them finishes and releases its memory
            const uint64_t seen = srv->batches;
            srv->cv_sched.wait(lk, [&] { return srv->stop || srv->batches != seen; });
        }
    }
}

double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = p / 100.0 * static_cast<double>(v.size() - 1);
    const size_t lo = static_cast<size_t>(pos);
    const size_t hi = std::min(lo + 1, v.size() - 1);
    const double f = pos - static_cast<double>(lo);
    return v[lo] * (1.0 - f) + v[hi] * f;
}

}  // namespace

extern "C" {

const char* hpd_server_last_error(void) { return g_server_error.c_str(); }

void hpd_server_default_config(hpd_server_config* cfg) {
    if (!cfg) return;
    cfg->device = 0;
    cfg->lanes = 3;
    cfg->hbm_budget_bytes = size_t(96) << 30;
    cfg->bytes_per_row_estimate = 200;
    cfg->max_batch_jobs = 4096;
    cfg->max_batch_rows = int64_t(32) << 20;
}

int hpd_server_create(const hpd_server_config* cfg, hpd_server** out) {
    if (!out) return fail(HPD_EINVAL, "out is NULL");
    *out = nullptr;
    hpd_server_config c;
    hpd_server_default_config(&c);
    if (cfg) c = *cfg;
    if (c.lanes < 1 || c.max_batch_jobs < 1 || c.max_batch_rows < 1)
        return fail(HPD_EINVAL, "lanes, max_batch_jobs and max_batch_rows must be positive");
    auto* srv = new hpd_server();
    srv->cfg = c;
    for (int i = 0; i < c.lanes; ++i) {
        auto lane = std::make_unique<Lane>();
        const int rc = hpd_context_create(c.device, &lane->ctx);
        if (rc != HPD_OK) {
            const std::string msg = hpd_last_error();
            for (auto& l : srv->lanes) hpd_context_destroy(l->ctx);
            delete srv;
            return fail(rc, msg);
        }
        srv->lanes.push_back(std::move(lane));
    }
    for (auto& l : srv->lanes) l->thread = std::thread(lane_main, srv, l.get());
    srv->scheduler = std::thread(scheduler_main, srv);
    *out = srv;
    return HPD_OK;
}

}  // extern "C"

namespace {

int enqueue(hpd_server* srv, std::shared_ptr<Request> r, uint64_t* ticket) {
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stop) return fail(HPD_EINVAL, "server is shutting down");
        r->charge = request_charge(*srv, *r);
        if (r->charge > srv->cfg.hbm_budget_bytes || r->n > srv->cfg.max_batch_rows)
            return fail(HPD_ENOMEM, "request of " + std::to_string(r->n) +
                                        " rows exceeds the server's memory budget");
        r->ticket = srv->next_ticket++;
        srv->table[r->ticket] = r;
        srv->queue.push_back(r);
        if (!srv->any_submit) {
            srv->any_submit = true;
            srv->t_first_submit = r->t_submit;
        }
        *ticket = r->ticket;
    }
    srv->cv_sched.notify_one();
    return HPD_OK;
}

int check_request(hpd_server* srv, int dtype, const void* data, int64_t n, int32_t d, float eps,
                  const uint64_t* ticket) {
    if (!srv || !ticket) return fail(HPD_EINVAL, "NULL argument");
    if (dtype_bytes(dtype) == 0) return fail(HPD_EINVAL, "unknown dtype code");
    if (!(eps > 0.0f)) return fail(HPD_EINVAL, "epsilon needs to be positive");
    if (n < 0 || d < 1 || (n > 0 && !data)) return fail(HPD_EINVAL, "bad shape or data");
    return HPD_OK;
}

}  // namespace

extern "C" {

int hpd_server_submit(hpd_server* srv, int dtype, const void* data, int64_t n, int32_t d,
                      float eps, uint64_t min_pts, uint64_t* ticket) {
    const int rc = check_request(srv, dtype, data, n, d, eps, ticket);
    if (rc != HPD_OK) return rc;
    auto r = std::make_shared<Request>();
    r->dtype = dtype;
    r->d = d;
    r->n = n;
    r->eps = eps;
    r->min_pts = min_pts;
    r->rows.assign(static_cast<const char*>(data),
                   static_cast<const char*>(data) + static_cast<size_t>(n) * d * dtype_bytes(dtype));
    return enqueue(srv, std::move(r), ticket);
}

int hpd_server_submit_device(hpd_server* srv, int dtype, const void* data, int64_t n, int32_t d,
                             float eps, uint64_t min_pts, int64_t* labels_out, uint64_t* ticket) {
    int rc = check_request(srv, dtype, data, n, d, eps, ticket);
    if (rc != HPD_OK) return rc;
    if (n > 0 && !labels_out) return fail(HPD_EINVAL, "labels_out is NULL");
    auto r = std::make_shared<Request>();
    r->dtype = dtype;
    r->d = d;
    r->n = n;
    r->eps = eps;
    r->min_pts = min_pts;
    r->device = true;
    r->d_rows = data;
    r->d_labels = labels_out;
    return enqueue(srv, std::move(r), ticket);
}

int hpd_server_wait(hpd_server* srv, uint64_t ticket, int64_t* labels_out, int64_t n,
                    double timeout_ms, hpd_request_info* info) {
    if (!srv) return fail(HPD_EINVAL, "server is NULL");
    std::shared_ptr<Request> r;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->table.find(ticket);
    if (it == srv->table.end()) return fail(HPD_EINVAL, "unknown ticket");
    r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) {
        srv->cv_done.wait(lk, ready);
    } else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms),
                                      ready)) {
        return fail(HPD_ETIMEOUT, "timed out");  // not done yet; ticket stays valid
    }
    if (info) {
        info->rows = r->n;
        info->batch_jobs = r->batch_jobs;
        info->queue_ms = std::chrono::duration<double, std::milli>(r->t_start - r->t_submit).count();
        info->latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    }
    const int status = r->status;
    if (status == HPD_OK && !r->device) {
        if (n != r->n || (n > 0 && !labels_out)) return fail(HPD_EINVAL, "labels buffer does not match the request");
        if (n) std::memcpy(labels_out, r->labels.data(), sizeof(int64_t) * static_cast<size_t>(n));
    }
    const std::string err = r->error;
    srv->table.erase(ticket);
    if (status != HPD_OK) return fail(status, err);
    return HPD_OK;
}

int hpd_server_get_stats(hpd_server* srv, hpd_server_stats* out) {
    if (!srv || !out) return fail(HPD_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    std::memset(out, 0, sizeof(*out));
    out->completed = srv->completed;
    out->failed = srv->failed;
    out->batches = srv->batches;
    out->mean_batch_jobs = srv->batches ? static_cast<double>(srv->batched_jobs) / srv->batches : 0.0;
    out->rows = srv->rows_done;
    out->p50_ms = percentile(srv->latencies_ms, 50.0);
    out->p99_ms = percentile(srv->latencies_ms, 99.0);
    out->max_ms = srv->latencies_ms.empty() ? 0.0
                                            : *std::max_element(srv->latencies_ms.begin(),
                                                                srv->latencies_ms.end());
    if (srv->any_submit && srv->completed) {
        const double s =
            std::chrono::duration<double>(srv->t_last_done - srv->t_first_submit).count();
        out->wall_s = s;
        out->jobs_per_s = s > 0 ? static_cast<double>(srv->completed) / s : 0.0;
        out->mrows_per_s = s > 0 ? static_cast<double>(srv->rows_done) / s / 1e6 : 0.0;
    }
    out->queued = srv->queue.size();
    out->charged_bytes = srv->charged;
    return HPD_OK;
}

int hpd_server_destroy(hpd_server* srv) {
    if (!srv) return HPD_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_sched.notify_all();
    srv->cv_lane.notify_all();
    srv->scheduler.join();
    for (auto& l : srv->lanes) l->thread.join();
    (void)hipSetDevice(srv->cfg.device);
    for (auto& l : srv->lanes) {
        release(*l);
        hpd_context_destroy(l->ctx);
    }
    delete srv;
    return HPD_OK;
}

}  // extern "C"
