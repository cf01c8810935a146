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
//
// Memory: a slot's pools grow to the largest call it has run and are kept for the next one, so
// what the server holds on the device is, per slot, the larger of its retained pools and the
// working set of the request it runs.  A request is admitted only if that sum, with the request
// placed on the free slot, stays within the budget: a fixed number of bytes, or by default what
// the device has free (hipMemGetInfo, read at every admission, plus what the slots already hold)
// less a reserve.  When the sum does not fit but would without the pools that idle slots keep,
// those pools are released first.  A request larger than the whole budget fails at once with
// EIK_ENOMEM, and a call that still runs out of device memory is retried once on released pools.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>

#include "eik_internal.h"

namespace {

using Clock = std::chrono::steady_clock;
// per-request statistics: the 13 eik_last_stats values of its call, then the requests in that call
constexpr int kReqStats = 14;

struct Req {
    eik_request r{};
    int64_t bytes = 0;
    int status = EIK_OK;
    bool started = false, done = false;
    std::string err;
    Clock::time_point t_submit, t_start, t_end;
    double stats[kReqStats] = {};  // the solver statistics of the call that ran it
};

// longest a blocked request lets smaller ones pass it (EIK_SERVER_BYPASS_MS overrides, read when
// the server is created: a test hook)
constexpr double kMaxBypassMs = 2000.0;
constexpr int kMaxBatch = 16;
constexpr double kBatchSpread = 2.0;
constexpr int64_t kMaxBatchSources = 512;
// device memory left out of the default budget: the larger of 2 GiB and 1/32 of the device
constexpr int64_t kReserveMin = int64_t(2) << 30;

bool batchable(const eik_request& a, const eik_request& b) {
    return a.type == EIK_REQ_TRAVELTIME && b.type == EIK_REQ_TRAVELTIME && a.model && b.model &&
           a.model->g.kind == b.model->g.kind && a.model->g.ndim == b.model->g.ndim &&
           std::memcmp(&a.params, &b.params, sizeof(eik_params)) == 0;
}

}  // namespace

struct eik_server {
    int device = 0;
    int64_t budget = 0;      // fixed budget in bytes; 0: free device memory less the reserve
    int64_t budget_now = 0;  // the budget the last admission used
    int64_t reserve = 0;
    double bypass_ms = kMaxBypassMs;
    std::vector<eik_ctx*> slots;
    std::vector<int64_t> held;  // per slot: device bytes its pools keep
    std::vector<int64_t> run;   // per slot: working set of the request it runs (0: idle)
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;
    std::map<int64_t, Req> reqs;
    int64_t inflight_bytes = 0;
    int64_t next_id = 1;
    int64_t nbatches = 0, nbatched = 0;
    int64_t rejected = 0, trims = 0, retries = 0;
    bool stop = false;
};

namespace eik {

// Device working set of one request: traveltime grid + pop keys per source (coarse pools), the
// refined boxes, and for gradients the adjoint pools (codes, D_d t, lambda, rhs, image) per
// source and node, the fine-box adjoint pools, and the two output grids.
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
        if (r.type == EIK_REQ_GRADIENT) b += r.nsrc * nf * (2 + 8 * nd + 5 * 8);
    }
    if (r.type == EIK_REQ_GRADIENT) b += r.nsrc * N * (2 + 8 * nd + 24) + 2 * N * 8;
    return b + (4 << 20);
}

static int call_request(eik_ctx* ctx, const eik_request& r) {
    if (r.type == EIK_REQ_TRAVELTIME)
        return eik_traveltime(ctx, r.model, r.nsrc, r.coordsrc, r.rec_off, r.coordrec, &r.params, r.ttpicks, nullptr);
    return eik_gradient(ctx, r.model, r.nsrc, r.coordsrc, r.rec_off, r.coordrec, r.pickobs, r.stdobs, r.whichgrad,
                        &r.params, r.gradvel, r.gradsrc, r.misf);
}

static void keep_stats(eik_ctx* ctx, Req& q, int64_t ncall) {
    for (int i = 0; i < kReqStats - 1; i++) q.stats[i] = ctx->stats[i];
    q.stats[kReqStats - 1] = (double)ncall;
}

// One request on its own (the slot is marked busy, the server lock is not held); a call that
// runs out of device memory runs once more on released pools.
static void run_request(eik_server* S, eik_ctx* ctx, Req& q) {
    int rc = call_request(ctx, q.r);
    if (rc == EIK_ENOMEM) {
        ctx_trim(ctx);
        rc = call_request(ctx, q.r);
        std::lock_guard<std::mutex> lk(S->mu);
        S->retries++;
    }
    q.status = rc;
    if (rc) q.err = eik_last_error(ctx);
    keep_stats(ctx, q, 1);
}

// What the slots would hold with a request of b bytes on slot `slot`: every other slot the
// larger of its retained pools and its running request's working set, this slot the same.
static int64_t footprint_with(const eik_server* S, int slot, int64_t b) {
    int64_t t = 0;
    for (size_t o = 0; o < S->slots.size(); o++)
        t += (int)o == slot ? std::max(S->held[o], b) : std::max(S->held[o], S->run[o]);
    return t;
}

// The budget now: the fixed one, or free device memory plus what the slots hold, less the
// reserve.  Called with the server lock held.
static int64_t budget_now(eik_server* S) {
    if (S->budget > 0) return S->budget_now = S->budget;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return S->budget_now;  // keep the last reading
    int64_t h = 0;
    for (int64_t x : S->held) h += x;
    return S->budget_now = std::max<int64_t>(0, (int64_t)fr + h - S->reserve);
}

// A request larger than the whole budget could never run: it fails on its own with EIK_ENOMEM.
// Called with the server lock held.
static void reject(eik_server* S, Req& q, int64_t B) {
    char b[256];
    snprintf(b, sizeof(b), "the request needs about %lld bytes of device memory; the server's budget is %lld",
             (long long)q.bytes, (long long)B);
    q.status = EIK_ENOMEM;
    q.err = b;
    q.started = q.done = true;
    q.t_start = q.t_end = Clock::now();
    S->rejected++;
    S->cv_done.notify_all();
}

// Whether a request of b bytes can run on idle slot `slot` within budget B; releases the pools
// of idle slots when that makes the room.  Called with the server lock held (an idle slot's
// context is then in no call).  A release waits for the device (hipFree synchronises), so the
// other slots' running calls finish first; it happens only when the budget binds.
static bool admit(eik_server* S, int slot, int64_t b, int64_t B) {
    if (footprint_with(S, slot, b) <= B) return true;
    int64_t floor = b;  // the footprint with every idle slot's pools released
    for (size_t o = 0; o < S->slots.size(); o++)
        if ((int)o != slot && S->run[o]) floor += std::max(S->held[o], S->run[o]);
    if (floor > B) return false;  // running work holds the room: wait for it
    for (size_t k = 0; k <= S->slots.size() && footprint_with(S, slot, b) > B; k++) {
        const size_t o = k == 0 ? (size_t)slot : k - 1;  // this slot's own pools first
        if (S->run[o] || S->held[o] == 0) continue;
        ctx_trim(S->slots[o]);
        S->held[o] = 0;
        S->trims++;
    }
    return footprint_with(S, slot, b) <= B;
}

static void worker(eik_server* S, int slot) {
    eik_ctx* ctx = S->slots[slot];
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        int64_t id = -1;
        S->cv_work.wait(lk, [&] {
            if (S->stop && S->queue.empty()) return true;
            const double waited =
                S->queue.empty()
                    ? 0.0
                    : std::chrono::duration<double, std::milli>(Clock::now() - S->reqs[S->queue.front()].t_submit).count();
            if (S->queue.empty()) return false;
            const int64_t B = budget_now(S);
            for (auto it = S->queue.begin(); it != S->queue.end();) {
                Req& q = S->reqs[*it];
                if (q.bytes > B) {
                    reject(S, q, B);
                    it = S->queue.erase(it);
                    continue;
                }
                if (admit(S, slot, q.bytes, B)) {
                    id = *it;
                    S->queue.erase(it);
                    return true;
                }
                if (waited > S->bypass_ms) break;  // the oldest request goes next
                ++it;
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
                nsrc + q.r.nsrc <= kMaxBatchSources && footprint_with(S, slot, bytes + q.bytes) <= S->budget_now) {
                ids.push_back(*it);
                bytes += q.bytes;
                nsrc += q.r.nsrc;
                it = S->queue.erase(it);
            } else {
                ++it;
            }
        }
        S->inflight_bytes += bytes;
        S->run[slot] = bytes;
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
            run_request(S, ctx, *qs[0]);
        } else {
            std::vector<eik_request> rr;
            for (Req* q : qs) rr.push_back(q->r);
            const int rc = eik_traveltime_batch(ctx, (int32_t)rr.size(), rr.data());
            if (rc == EIK_OK) {
                for (Req* q : qs) {
                    q->status = EIK_OK;
                    keep_stats(ctx, *q, (int64_t)qs.size());
                }
            } else {
                for (Req* q : qs) run_request(S, ctx, *q);  // each request gets its own outcome
            }
        }
        const int64_t now_held = ctx_pool_bytes(ctx);
        lk.lock();
        S->held[slot] = now_held;
        S->run[slot] = 0;
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
    if (!out || nslots < 1 || nslots > 64 || mem_budget < 0) return EIK_EINVAL;
    *out = nullptr;
    eik_server* S = new eik_server();
    S->device = device;
    S->budget = mem_budget;
    S->held.assign((size_t)nslots, 0);
    S->run.assign((size_t)nslots, 0);
    if (const char* e = getenv("EIK_SERVER_BYPASS_MS")) S->bypass_ms = std::max(0.0, atof(e));
    for (int i = 0; i < nslots; i++) {
        eik_ctx* c = nullptr;
        const int rc = eik_ctx_create(device, &c);
        if (rc) {
            for (eik_ctx* x : S->slots) eik_ctx_destroy(x);
            delete S;
            return rc;
        }
        c->split = false;  // the slots already run concurrently, one stream each
        S->slots.push_back(c);
    }
    size_t fr = 0, tot = 0;
    if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) {
        for (eik_ctx* x : S->slots) eik_ctx_destroy(x);
        delete S;
        return EIK_EDEVICE;
    }
    S->reserve = std::max<int64_t>(kReserveMin, (int64_t)tot / 32);
    S->budget_now = mem_budget > 0 ? mem_budget : std::max<int64_t>(0, (int64_t)fr - S->reserve);
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
        auto ins = S->reqs.emplace(*id, std::move(q)).first;
        // the budget of the last admission (a fixed budget, or the last free-memory reading)
        if (ins->second.bytes > S->budget_now) {
            reject(S, ins->second, S->budget_now);
            return EIK_OK;
        }
        S->queue.push_back(*id);
    }
    S->cv_work.notify_one();
    return EIK_OK;
}

int eik_request_bytes(const eik_request* r, int64_t* out) {
    if (!r || !out || !r->model || r->nsrc < 0) return EIK_EINVAL;
    if (r->type != EIK_REQ_TRAVELTIME && r->type != EIK_REQ_GRADIENT) return EIK_EINVAL;
    *out = request_bytes(*r);
    return EIK_OK;
}

int eik_wait(eik_server* S, int64_t id, double* times_ms, char* errbuf, int64_t errlen) {
    return eik_wait_stats(S, id, times_ms, nullptr, 0, errbuf, errlen);
}

int eik_wait_stats(eik_server* S, int64_t id, double* times_ms, double* stats, int32_t nstats, char* errbuf,
                   int64_t errlen) {
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
    if (stats)
        for (int i = 0; i < nstats && i < kReqStats; i++) stats[i] = q.stats[i];
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
    int64_t running = 0, held = 0;
    for (auto& kv : S->reqs) running += kv.second.started && !kv.second.done;
    for (int64_t x : S->held) held += x;
    const int64_t v[10] = {(int64_t)S->queue.size(), running, S->inflight_bytes, S->budget_now, S->nbatches,
                           S->nbatched, held, S->rejected, S->trims, S->retries};
    for (int i = 0; i < nout && i < 10; i++) out[i] = v[i];
    return EIK_OK;
}

}  // extern "C"
