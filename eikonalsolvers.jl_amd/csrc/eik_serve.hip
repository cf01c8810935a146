// eik_serve.hip -- request loop of libeik: continuous batching of traveltime and gradient
// requests on one GPU.  No counterpart in the reference (it evaluates one call at a time).
//
// A server owns `nslots` execution slots; each slot is a context (its own HIP stream and its
// own device pools, reused from request to request) driven by one host thread.  Submitted
// requests wait in a queue; a free slot admits the oldest request whose device working set fits
// in what the memory budget has left (smaller requests may pass a large one that does not fit,
// unless it has waited longer than kMaxBypassMs).  A traveltime request takes along the other
// queued traveltime requests of the same grid kind and parameters and of a similar size (working
// set within a factor kBatchSpread, so a small request does not wait for a large one) that fit,
// up to kMaxBatch requests and kMaxBatchSources sources: they run as one call whose launches
// hold all their sources (eik_traveltime_batch).  Admitted work runs concurrently on the slots' streams, so
// rounds of different slots overlap on the device.  Models are resident handles created by the
// caller; requests on one model only read it.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>

#include "eik_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Req {
    eik_request r{};
    int64_t bytes = 0;
    int status = EIK_OK;
    bool started = false, done = false;
    std::string err;
    Clock::time_point t_submit, t_start, t_end;
};

constexpr double kMaxBypassMs = 2000.0;
constexpr int kMaxBatch = 16;
constexpr double kBatchSpread = 2.0;
constexpr int64_t kMaxBatchSources = 512;

bool batchable(const eik_request& a, const eik_request& b) {
    return a.type == EIK_REQ_TRAVELTIME && b.type == EIK_REQ_TRAVELTIME && a.model && b.model &&
           a.model->g.kind == b.model->g.kind && a.model->g.ndim == b.model->g.ndim &&
           std::memcmp(&a.params, &b.params, sizeof(eik_params)) == 0;
}

}  // namespace

struct eik_server {
    int device = 0;
    int64_t budget = 0;
    std::vector<eik_ctx*> slots;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;
    std::map<int64_t, Req> reqs;
    int64_t inflight_bytes = 0;
    int64_t next_id = 1;
    int64_t nbatches = 0, nbatched = 0;
    bool stop = false;
};

namespace eik {

// Device working set of one request: traveltime grid + pop keys per source (coarse pools), the
// refined boxes, and for gradients the adjoint pools (codes, D_d t, lambda, rhs) and the image.
static int64_t request_bytes(const eik_request& r) {
    const eik_model* M = r.model;
    const int nd = M->g.ndim;
    const int64_t N = M->N;
    const int64_t esz = r.params.precision == EIK_PREC_FP32 ? 4 : 8;
    int64_t b = r.nsrc * N * (3 * esz + 1);
    const bool refine = r.params.refinearoundsrc || r.params.precision == EIK_PREC_FP32;
    if (refine) {
        const int64_t df = r.params.refinearoundsrc ? std::max(1, r.params.downscalefactor) : 1;
        const int64_t rad = r.params.refinearoundsrc ? std::max(0, r.params.noderadius) : 4;
        int64_t nf = 1;
        for (int d = 0; d < nd; d++) nf *= (2 * rad + 1) * df + 1;
        b += r.nsrc * nf * (4 * 8 + 1);
    }
    if (r.type == EIK_REQ_GRADIENT) b += r.nsrc * N * (2 + 8 * nd + 16) + 2 * N * 8;
    return b + (4 << 20);
}

static void run_request(eik_ctx* ctx, Req& q) {
    const eik_request& r = q.r;
    int rc;
    if (r.type == EIK_REQ_TRAVELTIME)
        rc = eik_traveltime(ctx, r.model, r.nsrc, r.coordsrc, r.rec_off, r.coordrec, &r.params, r.ttpicks, nullptr);
    else
        rc = eik_gradient(ctx, r.model, r.nsrc, r.coordsrc, r.rec_off, r.coordrec, r.pickobs, r.stdobs, r.whichgrad,
                          &r.params, r.gradvel, r.gradsrc, r.misf);
    q.status = rc;
    if (rc) q.err = eik_last_error(ctx);
}

static void worker(eik_server* S, int slot) {
    eik_ctx* ctx = S->slots[slot];
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        int64_t id = -1;
        S->cv_work.wait(lk, [&] {
            if (S->stop && S->queue.empty()) return true;
            const double waited =
                S->queue.empty()
                    ? 0.0
                    : std::chrono::duration<double, std::milli>(Clock::now() - S->reqs[S->queue.front()].t_submit).count();
            for (auto it = S->queue.begin(); it != S->queue.end(); ++it) {
                const Req& q = S->reqs[*it];
                if (S->inflight_bytes == 0 || S->inflight_bytes + q.bytes <= S->budget) {
                    id = *it;
                    S->queue.erase(it);
                    return true;
                }
                if (waited > kMaxBypassMs) break;  // the oldest request goes next
            }
            return false;
        });
        if (id < 0) return;  // stopping and nothing queued
        // the batch: this request and compatible queued traveltime requests that fit
        std::vector<int64_t> ids{id};
        int64_t bytes = S->reqs[id].bytes, nsrc = S->reqs[id].r.nsrc;
        for (auto it = S->queue.begin(); it != S->queue.end() && (int)ids.size() < kMaxBatch;) {
            const Req& q = S->reqs[*it];
            const double ratio = (double)q.bytes / (double)S->reqs[id].bytes;
            if (batchable(S->reqs[id].r, q.r) && ratio <= kBatchSpread && ratio >= 1.0 / kBatchSpread &&
                nsrc + q.r.nsrc <= kMaxBatchSources && S->inflight_bytes + bytes + q.bytes <= S->budget) {
                ids.push_back(*it);
                bytes += q.bytes;
                nsrc += q.r.nsrc;
                it = S->queue.erase(it);
            } else {
                ++it;
            }
        }
        S->inflight_bytes += bytes;
        const Clock::time_point t0 = Clock::now();
        for (int64_t k : ids) {
            S->reqs[k].started = true;
            S->reqs[k].t_start = t0;
        }
        if (ids.size() > 1) {
            S->nbatches++;
            S->nbatched += (int64_t)ids.size();
        }
        std::vector<Req*> qs;
        for (int64_t k : ids) qs.push_back(&S->reqs[k]);  // std::map nodes stay put while unlocked
        lk.unlock();
        if (qs.size() == 1) {
            run_request(ctx, *qs[0]);
        } else {
            std::vector<eik_request> rr;
            for (Req* q : qs) rr.push_back(q->r);
            const int rc = eik_traveltime_batch(ctx, (int32_t)rr.size(), rr.data());
            if (rc == EIK_OK) {
                for (Req* q : qs) q->status = EIK_OK;
            } else {
                for (Req* q : qs) run_request(ctx, *q);  // each request gets its own outcome
            }
        }
        lk.lock();
        const Clock::time_point t1 = Clock::now();
        for (Req* q : qs) {
            q->t_end = t1;
            q->done = true;
        }
        S->inflight_bytes -= bytes;
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
}

}  // namespace eik

using namespace eik;

extern "C" {

int eik_server_create(int32_t device, int32_t nslots, int64_t mem_budget, eik_server** out) {
    if (!out || nslots < 1 || nslots > 64 || mem_budget <= 0) return EIK_EINVAL;
    *out = nullptr;
    eik_server* S = new eik_server();
    S->device = device;
    S->budget = mem_budget;
    for (int i = 0; i < nslots; i++) {
        eik_ctx* c = nullptr;
        const int rc = eik_ctx_create(device, &c);
        if (rc) {
            for (eik_ctx* x : S->slots) eik_ctx_destroy(x);
            delete S;
            return rc;
        }
        S->slots.push_back(c);
    }
    for (int i = 0; i < nslots; i++) S->workers.emplace_back(worker, S, i);
    *out = S;
    return EIK_OK;
}

void eik_server_destroy(eik_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto& t : S->workers) t.join();
    for (eik_ctx* c : S->slots) eik_ctx_destroy(c);
    delete S;
}

int eik_submit(eik_server* S, const eik_request* r, int64_t* id) {
    if (!S || !r || !id || !r->model) return EIK_EINVAL;
    if (r->type != EIK_REQ_TRAVELTIME && r->type != EIK_REQ_GRADIENT) return EIK_EINVAL;
    Req q;
    q.r = *r;
    q.bytes = request_bytes(*r);
    q.t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) return EIK_EINVAL;
        *id = S->next_id++;
        S->reqs.emplace(*id, std::move(q));
        S->queue.push_back(*id);
    }
    S->cv_work.notify_one();
    return EIK_OK;
}

int eik_wait(eik_server* S, int64_t id, double* times_ms, char* errbuf, int64_t errlen) {
    if (!S) return EIK_EINVAL;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return EIK_EINVAL;
    S->cv_done.wait(lk, [&] { return it->second.done; });
    const Req& q = it->second;
    if (times_ms) {
        times_ms[0] = std::chrono::duration<double, std::milli>(q.t_end - q.t_submit).count();
        times_ms[1] = std::chrono::duration<double, std::milli>(q.t_start - q.t_submit).count();
        times_ms[2] = std::chrono::duration<double, std::milli>(q.t_end - q.t_start).count();
    }
    if (errbuf && errlen > 0) {
        std::strncpy(errbuf, q.err.c_str(), (size_t)errlen - 1);
        errbuf[errlen - 1] = '\0';
    }
    const int st = q.status;
    S->reqs.erase(it);
    return st;
}

int eik_poll(eik_server* S, int64_t id, int32_t* done) {
    if (!S || !done) return EIK_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return EIK_EINVAL;
    *done = it->second.done ? 1 : 0;
    return EIK_OK;
}

int eik_server_stats(eik_server* S, int64_t* out, int32_t nout) {
    if (!S || !out) return EIK_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    int64_t running = 0;
    for (auto& kv : S->reqs) running += kv.second.started && !kv.second.done;
    const int64_t v[6] = {(int64_t)S->queue.size(), running, S->inflight_bytes, S->budget, S->nbatches, S->nbatched};
    for (int i = 0; i < nout && i < 6; i++) out[i] = v[i];
    return EIK_OK;
}

}  // extern "C"
