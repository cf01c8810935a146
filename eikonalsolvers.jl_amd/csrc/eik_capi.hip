// eik_capi.hip -- C ABI of libeik (declared in include/eik.h) and its host-side geometry.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <system_error>
#include <thread>

#include "eik_internal.h"

namespace eik {

void make_inst(const eik_grid& g, InstDev& I) {
    memset(&I, 0, sizeof(I));
    for (int d = 0; d < 3; d++) {
        I.n[d] = d < g.ndim ? (int32_t)g.n[d] : 1;
        I.h[d] = g.d[d];
        I.o[d] = d < g.ndim ? g.o[d] : 0.0;
    }
    if (g.kind == EIK_GRID2D_CART || g.kind == EIK_GRID3D_CART) I.h[1] = I.h[2] = g.d[0];
    I.kind = g.kind;
    I.dthrad = deg2rad(g.d[1]);
    I.dphrad = deg2rad(g.d[2]);
    for (int d = 0; d < 3; d++) {
        I.xlo[d] = 1;  // no FP64 region
        I.xhi[d] = 0;
    }
    I.tstopK = INFINITY;
    I.tstopS = -INFINITY;
    I.tstopD = 0;
}

void set_tiles(InstDev& I, int nd) {
    const int fe = fwd_tile_edge(nd), ae = adj_tile_edge(nd);
    for (int d = 0; d < 3; d++) {
        I.t[d] = d < nd ? (I.n[d] + fe - 1) / fe : 1;
        I.at[d] = d < nd ? (I.n[d] + ae - 1) / ae : 1;
    }
}

static inline double gcoord(const eik_grid& g, int d, int64_t i) {
    const double h = (g.kind == EIK_GRID2D_CART || g.kind == EIK_GRID3D_CART) ? g.d[0] : g.d[d];
    return fma((double)i, h, g.o[d]);
}

// findenclosingbox (src/EikSolv/eikmainfwd.jl:1114-1186): 1 corner when the source sits on a node
// (remainder of every coordinate within 1e-8), else the 2^ndim corners of its cell in the
// reference's order, shifted back by one at the upper edge.
int find_enclosing_box(const eik_grid& g, const double* src, int corn[8][3]) {
    const int nd = g.ndim;
    int ijk[3] = {0, 0, 0};
    bool onnode = true;
    for (int d = 0; d < nd; d++) {
        const double h = (g.kind == EIK_GRID2D_CART || g.kind == EIK_GRID3D_CART) ? g.d[0] : g.d[d];
        const double x = src[d] - g.o[d];
        const double r = std::fmod(x, h);
        const double q = std::nearbyint((x - r) / h);
        ijk[d] = (int)std::floor(q);
        if (!(std::fabs(r) <= 1e-8)) onnode = false;
    }
    if (onnode) {
        for (int d = 0; d < 3; d++) corn[0][d] = d < nd ? ijk[d] : 0;
        return 1;
    }
    for (int d = 0; d < nd; d++)
        if (ijk[d] == g.n[d] - 1) ijk[d] -= 1;
    static const int off2[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
    static const int off3[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                   {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
    const int nc = 1 << nd;
    for (int c = 0; c < nc; c++)
        for (int d = 0; d < 3; d++) corn[c][d] = d < nd ? ijk[d] + (nd == 2 ? off2[c][d] : off3[c][d]) : 0;
    return nc;
}

// createfinegrid (src/EikSolv/eikmainfwd.jl:950-1104): window = enclosing box +- noderadius
// coarse nodes clipped to the model, fine spacing h/df, fine size (delta ijk)*df+1.
void build_fine(const eik_grid& g, const double* src, int df, int noderadius, FineDev& F, InstDev& I) {
    const int nd = g.ndim;
    memset(&F, 0, sizeof(F));
    int corn[8][3];
    const int nc = find_enclosing_box(g, src, corn);
    int i1[3] = {0, 0, 0}, i2[3] = {0, 0, 0};
    for (int d = 0; d < nd; d++) {
        int mn = corn[0][d], mx = corn[0][d];
        for (int c = 1; c < nc; c++) {
            mn = std::min(mn, corn[c][d]);
            mx = std::max(mx, corn[c][d]);
        }
        const int v1 = mn - noderadius, v2 = mx + noderadius;
        F.outmin[d] = v1 < 0;
        F.outmax[d] = v2 > g.n[d] - 1;
        i1[d] = F.outmin[d] ? 0 : v1;
        i2[d] = F.outmax[d] ? (int)g.n[d] - 1 : v2;
    }
    F.df = df;
    eik_grid gf{};
    gf.kind = g.kind;
    gf.ndim = nd;
    for (int d = 0; d < 3; d++) {
        F.i1c[d] = d < nd ? i1[d] : 0;
        F.nc[d] = d < nd ? (int32_t)g.n[d] : 1;
        gf.n[d] = d < nd ? (int64_t)(i2[d] - i1[d]) * df + 1 : 1;
    }
    const bool cart = (g.kind == EIK_GRID2D_CART || g.kind == EIK_GRID3D_CART);
    for (int d = 0; d < 3; d++) gf.d[d] = (cart ? g.d[0] : g.d[d]) / df;
    for (int d = 0; d < nd; d++) gf.o[d] = gcoord(g, d, i1[d]);
    make_inst(gf, I);
    set_tiles(I, nd);
    F.ncorn = find_enclosing_box(gf, src, F.corn);
}

// Receivers per call beyond which rec_off is taken as corrupt (2^32 receivers would need 32 GiB of
// coordinates alone in 3D).
static const int64_t kMaxReceivers = (int64_t)1 << 32;

// checksrcrecposition (src/EikSolv/eikchecks.jl:4-50) and ExtraParams checks (eikstructs.jl:270-277)
int load_sources(eik_ctx* ctx, eik_model* M, int64_t nsrc, const double* coordsrc, const int64_t* rec_off,
                 const double* coordrec, const eik_params* P, SourceSet& S) {
    const int nd = M->g.ndim;
    if (nsrc < 0 || (nsrc > 0 && (!coordsrc || !rec_off))) return ctx->fail(EIK_EINVAL, "bad source arguments");
    if (!P) return ctx->fail(EIK_EINVAL, "null params");
    if (P->radiussmoothgradsrc < 0) return ctx->fail(EIK_EASSERT, "radiussmoothgradsrc>= 0");
    if (P->smoothgradkern < 0) return ctx->fail(EIK_EASSERT, "smoothgradkern>=0");
    if (P->smoothgradkern > 0 && P->smoothgradkern % 2 == 0)
        return ctx->fail(EIK_EINVAL, "ExtraParams: smoothgradkern must be an odd number");
    if (P->precision != EIK_PREC_FP64 && P->precision != EIK_PREC_MIXED && P->precision != EIK_PREC_FP32)
        return ctx->fail(EIK_EINVAL, "bad precision");
    if (P->fast != 0 && P->fast != 1) return ctx->fail(EIK_EINVAL, "fast must be 0 or 1");
    if (P->refinearoundsrc && (P->downscalefactor < 1 || P->noderadius < 0))
        return ctx->fail(EIK_EINVAL, "bad grid refinement parameters");
    if (P->refinearoundsrc && M->g.kind == EIK_GRID3D_SPHERE)
        return ctx->fail(EIK_EINVAL, "createfinegrid(): spherical coordinates in 3D still work in progress...");
    S.nsrc = nsrc;
    S.nd = nd;
    // rec_off is read only when there are sources (it may be NULL otherwise); its prefix sums are
    // checked before anything is sized from them, and no allocation failure crosses the ABI
    if (nsrc > 0) {
        if (rec_off[0] != 0) return ctx->fail(EIK_EINVAL, "rec_off[0] must be 0");
        for (int64_t s = 0; s < nsrc; s++)
            if (rec_off[s + 1] < rec_off[s]) return ctx->fail(EIK_EINVAL, "rec_off must be non-decreasing");
        if (rec_off[nsrc] > kMaxReceivers) return ctx->fail(EIK_EINVAL, "too many receivers (%lld)", (long long)rec_off[nsrc]);
    }
    try {
        S.src.assign(coordsrc, coordsrc + nsrc * nd);
        if (nsrc > 0)
            S.rec_off.assign(rec_off, rec_off + nsrc + 1);
        else
            S.rec_off.assign(1, 0);
        const int64_t nrec = S.nrec();
        if (nrec > 0 && !coordrec) return ctx->fail(EIK_EINVAL, "null receivers");
        S.rec.assign(coordrec, coordrec + nrec * nd);
    } catch (const std::bad_alloc&) {
        return ctx->fail(EIK_ENOMEM, "out of host memory (sources and receivers)");
    }
    const int64_t nrec = S.nrec();
    for (int d = 0; d < nd; d++) {
        const double lo = gcoord(M->g, d, 0), hi = gcoord(M->g, d, M->g.n[d] - 1);
        for (int64_t s = 0; s < nsrc; s++) {
            const double x = S.src[s * nd + d];
            if (!(lo <= x && x <= hi)) return ctx->fail(EIK_EASSERT, "source %lld outside the grid", (long long)s);
        }
        for (int64_t r = 0; r < nrec; r++) {
            const double x = S.rec[r * nd + d];
            if (!(lo <= x && x <= hi)) return ctx->fail(EIK_EASSERT, "receiver %lld outside the grid", (long long)r);
        }
    }
    return EIK_OK;
}

// The FP32 precision keeps the solve near the source in FP64.  Without refinement that region is
// an FP64 box of +-kNearSourceRadius nodes at the coarse spacing, solved with the refinement
// machinery at downscalefactor 1: the nodes it hands to the coarse solve are those the heap
// accepts before the front reaches the box faces, whose values do not depend on the box, so the
// coarse traveltimes are the reference's unrefined ones (the gradient then follows the refined
// formulas near the source).  The mixed precision needs no box: its traveltimes are FP64, so the
// nodes within kNearSourceRadius of the source box simply use the FP64 stencil in the coarse solve
// (forward_solve), and the problem, gradient included, stays the unrefined one.
eik_params effective_params(const eik_params* p, int kind) {
    eik_params e = *p;
    e.precision = run_precision(e.precision, kind);
    if (e.precision == EIK_PREC_FP32 && !e.refinearoundsrc) {
        e.refinearoundsrc = 1;
        e.downscalefactor = 1;
        e.noderadius = kNearSourceRadius;
    }
    return e;
}

}  // namespace eik

using namespace eik;

extern "C" {

const char* eik_version(void) { return "libeik 0.1 (gfx950)"; }

const char* eik_last_error(const eik_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

// Streams a large traveltime or gradient call's sources are split over by default (EIK_STREAMS = 1
// keeps one; eik_ctx_set_streams per context).  Each half runs its own round loop on its own
// stream, so while the last tiles of one half's round finish, the other half's round kernel takes
// the CUs they leave idle (201x201x101 x 16 sources: +11 % FP64, +14 % FP32; 600x400 x 256: +28 %;
// 120x150x150 spherical x 32: +20 %; four streams are slower than two).  Calls below kSplitNodes
// source-nodes stay on one stream: there the halves' rounds are too small to fill the device and
// the split costs time (1-2 x 2^22 source-nodes: 600x400 x 18 and x 32 sources 20-23 % slower,
// 161x161x81 x 4 10 %, 101x101x51 x 9 14 %; 300x220 x 4: 6 %), from about 3.7 x 2^22 it gains
// (600x400 x 64 +4 %, 161x161x81 x 8 +4 %, the larger workloads above).  Results do not depend on
// the split: every source is solved on its own, and a gradient's sums and its source-to-source
// carry-over run over both halves in source order.
constexpr int64_t kSplitNodes = (int64_t)3 << 22;
static int split_streams() {
    static const int v = [] {
        const char* e = getenv("EIK_STREAMS");
        return e ? std::min(2, std::max(1, atoi(e))) : 2;
    }();
    return v;
}

// stream, counters and timers of a context (a peer shares its owner's tbase)
static bool ctx_init(eik_ctx* ctx, int device, bool with_tbase) {
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&ctx->d_small, kSmallBytes) != hipSuccess ||
        hipHostMalloc(&ctx->h_small, kSmallBytes, hipHostMallocDefault) != hipSuccess ||
        hipEventCreate(&ctx->ev[0]) != hipSuccess || hipEventCreate(&ctx->ev[1]) != hipSuccess)
        return false;
    for (auto& e : ctx->kev)
        if (hipEventCreate(&e) != hipSuccess) return false;
    return !with_tbase || (hipEventCreate(&ctx->tbase) == hipSuccess && hipEventCreate(&ctx->tend) == hipSuccess);
}

int eik_ctx_create(int32_t device, eik_ctx** out) {
    if (!out) return EIK_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return EIK_EDEVICE;
    if (device < 0 || device >= ndev) return EIK_EINVAL;
    eik_ctx* ctx = new eik_ctx();
    if (!ctx_init(ctx, device, true)) {
        eik_ctx_destroy(ctx);
        return EIK_EDEVICE;
    }
    ctx->split = split_streams() == 2;
    // the no-allocation entry point's scratch block: this device's address, resolved once
    scratch_setup(device);
    *out = ctx;
    return EIK_OK;
}

void eik_ctx_destroy(eik_ctx* ctx) {
    if (!ctx) return;
    if (ctx->peer) eik_ctx_destroy(ctx->peer);
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->bufs) kv.second.release();
    if (ctx->d_small) (void)hipFree(ctx->d_small);
    if (ctx->h_small) (void)hipHostFree(ctx->h_small);
    for (int i = 0; i < 2; i++)
        if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
    for (auto& e : ctx->kev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->tbase && !ctx->owner) (void)hipEventDestroy(ctx->tbase);
    if (ctx->tend) (void)hipEventDestroy(ctx->tend);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

static int check_grid(eik_ctx* ctx, const eik_grid* g) {
    if (!g) return ctx->fail(EIK_EINVAL, "null grid");
    if (g->ndim < 2 || g->ndim > 3) return ctx->fail(EIK_EASSERT, "2<=ndims(vel)<=3");
    const bool k2 = g->kind == EIK_GRID2D_CART || g->kind == EIK_GRID2D_SPHERE;
    const bool k3 = g->kind == EIK_GRID3D_CART || g->kind == EIK_GRID3D_SPHERE;
    if (!(k2 || k3)) return ctx->fail(EIK_EINVAL, "unknown grid kind %d", g->kind);
    if ((k2 && g->ndim != 2) || (k3 && g->ndim != 3)) return ctx->fail(EIK_EASSERT, "Ndim==length(grd.grsize)");
    for (int d = 0; d < g->ndim; d++) {
        if (g->n[d] < 1) return ctx->fail(EIK_EASSERT, "all(grsize.>0)");
        // receiver interpolation needs a cell along every axis (the reference indexes i+1)
        if (g->n[d] < 2) return ctx->fail(EIK_EASSERT, "every grid axis needs at least 2 nodes");
        if (g->n[d] > (int64_t)1 << 30) return ctx->fail(EIK_EINVAL, "grid too large");
        if (!(g->d[d] > 0.0) && !(d > 0 && (g->kind == EIK_GRID2D_CART || g->kind == EIK_GRID3D_CART)))
            return ctx->fail(EIK_EASSERT, "grid spacing must be > 0");
    }
    return EIK_OK;
}

static int update_vel_locked(eik_model* M, const double* vel) {
    eik_ctx* ctx = M->ctx;
    double sm = 0.0;
    for (int64_t i = 0; i < M->N; i++) {
        if (!(vel[i] > 0.0)) return ctx->fail(EIK_EASSERT, "all(vel.>0.0)");
        sm += 1.0 / vel[i];
    }
    M->smean = sm / (double)M->N;
    (void)hipSetDevice(ctx->device);
    CK(hipMemcpyAsync(M->d_vel, vel, M->N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (model_slowness(M)) return ctx->fail(EIK_EDEVICE, "slowness kernel failed");
    CK(hipStreamSynchronize(ctx->stream));
    return EIK_OK;
}

int eik_model_create(eik_ctx* ctx, const eik_grid* grid, const double* vel, eik_model** out) {
    if (!ctx || !out || !vel) return EIK_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    *out = nullptr;
    int rc = check_grid(ctx, grid);
    if (rc) return rc;
    eik_model* M = new eik_model();
    M->ctx = ctx;
    M->g = *grid;
    if (grid->kind == EIK_GRID2D_CART || grid->kind == EIK_GRID3D_CART) M->g.d[1] = M->g.d[2] = grid->d[0];
    M->N = 1;
    for (int d = 0; d < grid->ndim; d++) M->N *= grid->n[d];
    for (int d = grid->ndim; d < 3; d++) M->g.n[d] = 1;
    (void)hipSetDevice(ctx->device);
    if (hipMalloc(&M->d_vel, M->N * sizeof(double)) != hipSuccess ||
        hipMalloc(&M->d_slow64, M->N * sizeof(double)) != hipSuccess ||
        hipMalloc(&M->d_slow32, M->N * sizeof(float)) != hipSuccess) {
        eik_model_destroy(M);
        return ctx->fail(EIK_ENOMEM, "out of device memory (model)");
    }
    if (grid->kind == EIK_GRID3D_SPHERE) {
        // sin(theta_j) of the 3D sphere metric r sin(theta) dphi, from the host libm
        std::vector<double> st((size_t)M->g.n[1]);
        for (int64_t j = 0; j < M->g.n[1]; j++) st[j] = sin(deg2rad(gcoord(M->g, 1, j)));
        if (hipMalloc(&M->d_sinth, st.size() * sizeof(double)) != hipSuccess ||
            hipMemcpy(M->d_sinth, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            eik_model_destroy(M);
            return ctx->fail(EIK_ENOMEM, "out of device memory (model)");
        }
    }
    rc = update_vel_locked(M, vel);
    if (rc) {
        eik_model_destroy(M);
        return rc;
    }
    *out = M;
    return EIK_OK;
}

int eik_model_update_vel(eik_model* M, const double* vel) {
    if (!M || !vel) return EIK_EINVAL;
    std::lock_guard<std::mutex> lk(M->ctx->mu);
    return update_vel_locked(M, vel);
}

void eik_model_destroy(eik_model* M) {
    if (!M) return;
    if (M->ctx) (void)hipSetDevice(M->ctx->device);
    if (M->d_vel) (void)hipFree(M->d_vel);
    if (M->d_slow64) (void)hipFree(M->d_slow64);
    if (M->d_slow32) (void)hipFree(M->d_slow32);
    if (M->d_sinth) (void)hipFree(M->d_sinth);
    delete M;
}

// marks a context and its peer as running one split call until the scope ends
struct SplitScope {
    eik_ctx* c;
    explicit SplitScope(eik_ctx* ctx) : c(ctx) { c->in_split = c->peer->in_split = true; }
    ~SplitScope() { c->in_split = c->peer->in_split = false; }
};

// The second half of a split call runs on a thread of its own; if no thread can be started, it
// runs on the calling thread after the first half (still on the peer's stream).
static std::thread start_thread(const std::function<void()>& f) {
    try {
        return std::thread(f);
    } catch (const std::system_error&) {
        return std::thread();
    }
}
static void finish_thread(std::thread& th, const std::function<void()>& f) {
    if (th.joinable())
        th.join();
    else
        f();
}

// the context's peer, created on first use
static eik_ctx* get_peer(eik_ctx* ctx) {
    if (!ctx->peer) {
        eik_ctx* p = new eik_ctx();
        p->owner = ctx;
        if (!ctx_init(p, ctx->device, false)) {
            eik_ctx_destroy(p);
            return nullptr;
        }
        p->tbase = ctx->tbase;
        ctx->peer = p;
    }
    return ctx->peer;
}

// sources [a, b) of S, receivers re-based
static SourceSet sub_sources(const SourceSet& S, int64_t a, int64_t b) {
    SourceSet T;
    T.nsrc = b - a;
    T.nd = S.nd;
    T.src.assign(S.src.begin() + a * S.nd, S.src.begin() + b * S.nd);
    T.rec_off.resize((size_t)(b - a + 1));
    for (int64_t i = 0; i <= b - a; i++) T.rec_off[i] = S.rec_off[a + i] - S.rec_off[a];
    T.rec.assign(S.rec.begin() + S.rec_off[a] * S.nd, S.rec.begin() + S.rec_off[b] * S.nd);
    return T;
}

// a peer's statistics folded into its owner's: counts and launch durations add, phase times and
// the settling stage take the larger ([17] and [18], unions over both streams' launch spans, are
// the caller's)
static void merge_peer_stats(eik_ctx* ctx) {
    const double* q = ctx->peer->stats;
    double* s = ctx->stats;
    for (int i : {1, 2, 3, 11}) s[i] = std::max(s[i], q[i]);
    for (int i : {4, 5, 6, 7, 8, 9, 10, 12, 13, 14, 15, 16}) s[i] += q[i];
}

int eik_traveltime(eik_ctx* ctx, eik_model* model, int64_t nsrc, const double* coordsrc, const int64_t* rec_off,
                   const double* coordrec, const eik_params* params_in, double* ttpicks, double* ttgrid) {
    if (!ctx || !model) return EIK_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    SourceSet S;
    int rc = load_sources(ctx, model, nsrc, coordsrc, rec_off, coordrec, params_in, S);
    if (rc) return rc;
    const eik_params pe = effective_params(params_in, model->g.kind);
    const eik_params* params = &pe;
    if (S.nrec() > 0 && !ttpicks) return ctx->fail(EIK_EINVAL, "null picks buffer");
    if (nsrc == 0) return EIK_OK;
    stats_begin(ctx);
    const bool split = ctx->split && nsrc >= 2 && nsrc * model->N >= kSplitNodes;
    eik_ctx* peer = split ? get_peer(ctx) : nullptr;
    if (!peer) {
        rc = forward_solve(ctx, model, S, params->precision, params);
        if (rc) return rc;
        rc = forward_picks(ctx, model, S, params->precision, ttpicks, ttgrid);
        ctx->stats[0] = call_ms(ctx);
        ctx->stats[17] = busy_ms(ctx, false, false);
        ctx->stats[18] = busy_ms(ctx, false, true);
        return rc;
    }
    // the second half of the sources on the peer's stream, from a thread of its own
    const int64_t h = (nsrc + 1) / 2;
    const SourceSet S0 = sub_sources(S, 0, h), S1 = sub_sources(S, h, nsrc);
    double* picks1 = ttpicks ? ttpicks + S.rec_off[h] : nullptr;
    double* grid1 = ttgrid ? ttgrid + h * model->N : nullptr;
    stats_begin(peer);
    SplitScope scope(ctx);
    CK(hipEventSynchronize(ctx->tbase));  // the peer's launch spans are timed from it
    int rc1 = EIK_OK;
    const auto second = [&] {
        (void)hipSetDevice(peer->device);
        rc1 = forward_solve(peer, model, S1, params->precision, params);
        if (rc1 == EIK_OK) rc1 = forward_picks(peer, model, S1, params->precision, picks1, grid1);
    };
    std::thread th = start_thread(second);
    rc = forward_solve(ctx, model, S0, params->precision, params);
    if (rc == EIK_OK) rc = forward_picks(ctx, model, S0, params->precision, ttpicks, ttgrid);
    finish_thread(th, second);
    if (rc) return rc;
    if (rc1) return ctx->fail(rc1, "%s", peer->err.c_str());
    merge_peer_stats(ctx);
    ctx->stats[0] = call_ms(ctx);
    ctx->stats[17] = busy_ms(ctx, true, false);
    ctx->stats[18] = busy_ms(ctx, true, true);
    return EIK_OK;
}

int eik_traveltime_batch(eik_ctx* ctx, int32_t nreq, const eik_request* reqs) {
    if (!ctx || nreq < 0 || (nreq > 0 && !reqs)) return EIK_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (nreq == 0) return EIK_OK;
    std::vector<SourceSet> SS((size_t)nreq);
    std::vector<FwdGroup> G;
    int64_t nrec = 0;
    for (int32_t i = 0; i < nreq; i++) {
        const eik_request& q = reqs[i];
        if (q.type != EIK_REQ_TRAVELTIME || !q.model) return ctx->fail(EIK_EINVAL, "request %d: not a traveltime request", i);
        if (std::memcmp(&q.params, &reqs[0].params, sizeof(eik_params)) != 0)
            return ctx->fail(EIK_EINVAL, "request %d: params differ from request 0", i);
        if (q.model->g.kind != reqs[0].model->g.kind || q.model->g.ndim != reqs[0].model->g.ndim)
            return ctx->fail(EIK_EINVAL, "request %d: grid kind differs from request 0", i);
        const int rc = load_sources(ctx, q.model, q.nsrc, q.coordsrc, q.rec_off, q.coordrec, &q.params, SS[i]);
        if (rc) return rc;
        if (SS[i].nrec() > 0 && !q.ttpicks) return ctx->fail(EIK_EINVAL, "request %d: null picks buffer", i);
        if (SS[i].nsrc > 0) G.push_back(FwdGroup{q.model, &SS[i]});
        nrec += SS[i].nrec();
    }
    if (G.empty()) return EIK_OK;
    const eik_params pe = effective_params(&reqs[0].params, reqs[0].model->g.kind);
    stats_begin(ctx);
    int rc = forward_solve(ctx, G, pe.precision, &pe);
    if (rc) return rc;
    std::vector<double> picks((size_t)std::max<int64_t>(nrec, 1));
    rc = forward_picks(ctx, G, pe.precision, picks.data(), nullptr);
    ctx->stats[0] = call_ms(ctx);
    ctx->stats[17] = busy_ms(ctx, false, false);
    ctx->stats[18] = busy_ms(ctx, false, true);
    if (rc) return rc;
    int64_t off = 0;
    for (int32_t i = 0; i < nreq; i++) {
        const int64_t n = SS[i].nsrc > 0 ? SS[i].nrec() : 0;
        if (n > 0) std::memcpy(reqs[i].ttpicks, picks.data() + off, (size_t)n * sizeof(double));
        off += n;
    }
    return EIK_OK;
}

int eik_misfit(eik_ctx* ctx, eik_model* model, int64_t nsrc, const double* coordsrc, const int64_t* rec_off,
               const double* coordrec, const double* pickobs, const double* stdobs, const eik_params* params,
               double* misf) {
    if (!ctx || !model || !misf) return EIK_EINVAL;
    int64_t nrec = 0;
    {
        // validate everything (sources, rec_off, observations) before reading any of it
        std::lock_guard<std::mutex> lk(ctx->mu);
        SourceSet S;
        const int rc = load_sources(ctx, model, nsrc, coordsrc, rec_off, coordrec, params, S);
        if (rc) return rc;
        nrec = S.nrec();
        if (nrec > 0 && (!pickobs || !stdobs)) return ctx->fail(EIK_EINVAL, "null observations");
    }
    std::vector<double> picks((size_t)std::max<int64_t>(nrec, 1));
    int rc = eik_traveltime(ctx, model, nsrc, coordsrc, rec_off, coordrec, params, picks.data(), nullptr);
    if (rc) return rc;
    double m = 0.0;
    for (int64_t s = 0; s < nsrc; s++) {
        double ms = 0.0;
        for (int64_t r = rec_off[s]; r < rec_off[s + 1]; r++) {
            const double dd = picks[r] - pickobs[r];
            ms += (dd * dd) / (stdobs[r] * stdobs[r]);
        }
        m += ms;
    }
    *misf = 0.5 * m;
    return EIK_OK;
}

int eik_gradient(eik_ctx* ctx, eik_model* model, int64_t nsrc, const double* coordsrc, const int64_t* rec_off,
                 const double* coordrec, const double* pickobs, const double* stdobs, int32_t whichgrad,
                 const eik_params* params_in, double* gradvel, double* gradsrc, double* misf) {
    if (!ctx || !model || !misf) return EIK_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (whichgrad < EIK_GRADVEL || whichgrad > EIK_GRADVELANDSRCLOC)
        return ctx->fail(EIK_EASSERT, "Symbol 'whichgrad' not in (:gradvel, :gradsrcloc, :gradvelandsrcloc)");
    SourceSet S;
    int rc = load_sources(ctx, model, nsrc, coordsrc, rec_off, coordrec, params_in, S);
    if (rc) return rc;
    const eik_params pe = effective_params(params_in, model->g.kind);
    const eik_params* params = &pe;
    const int kind = model->g.kind;
    // Grid3DSphere: the reference raises here (createsparsederivativematrices!, eikmainfwd.jl:885-887);
    // this library differentiates its own 3D spherical forward (r sin(theta) dphi metric), :gradvel only
    if ((kind == EIK_GRID2D_SPHERE || kind == EIK_GRID3D_SPHERE) && whichgrad != EIK_GRADVEL)
        return ctx->fail(EIK_EINVAL, "gradsrcloc is not available on spherical grids");
    if (S.nrec() > 0 && (!pickobs || !stdobs)) return ctx->fail(EIK_EINVAL, "null observations");
    if (whichgrad != EIK_GRADSRCLOC && !gradvel) return ctx->fail(EIK_EINVAL, "null gradvel buffer");
    if (whichgrad != EIK_GRADVEL && !gradsrc) return ctx->fail(EIK_EINVAL, "null gradsrc buffer");
    if (nsrc == 0) {
        *misf = 0.0;
        if (gradvel && whichgrad != EIK_GRADSRCLOC) memset(gradvel, 0, model->N * sizeof(double));
        return EIK_OK;
    }
    stats_begin(ctx);
    const bool split = ctx->split && nsrc >= 2 && nsrc * model->N >= kSplitNodes;
    eik_ctx* peer = split ? get_peer(ctx) : nullptr;
    if (!peer) {
        rc = gradient_solve(ctx, model, S, params, whichgrad, pickobs, stdobs, gradvel, gradsrc, misf);
        ctx->stats[0] = call_ms(ctx);
        ctx->stats[17] = busy_ms(ctx, false, false);
        ctx->stats[18] = busy_ms(ctx, false, true);
        return rc;
    }
    // the second half of the sources on the peer's stream; the sums over sources (misfit, gradvel)
    // then run here over both halves in source order, as they do unsplit
    const int64_t h = (nsrc + 1) / 2, r1 = S.rec_off[h];
    const SourceSet S0 = sub_sources(S, 0, h), S1 = sub_sources(S, h, nsrc);
    const int nd = model->g.ndim;
    GradPart parts[2] = {};
    stats_begin(peer);
    SplitScope scope(ctx);
    CK(hipEventSynchronize(ctx->tbase));
    int rc1 = EIK_OK;
    const auto second = [&] {
        (void)hipSetDevice(peer->device);
        rc1 = gradient_part(peer, model, S1, params, whichgrad, pickobs ? pickobs + r1 : nullptr,
                            stdobs ? stdobs + r1 : nullptr, gradsrc ? gradsrc + h * nd : nullptr, true, &parts[1]);
    };
    std::thread th = start_thread(second);
    rc = gradient_part(ctx, model, S0, params, whichgrad, pickobs, stdobs, gradsrc, false, &parts[0]);
    finish_thread(th, second);
    if (rc) return rc;
    if (rc1) return ctx->fail(rc1, "%s", peer->err.c_str());
    merge_peer_stats(ctx);
    rc = gradient_total(ctx, model, params, whichgrad, parts, 2, S.rec_off.data(), nsrc, gradvel, misf);
    ctx->stats[0] = call_ms(ctx);
    ctx->stats[17] = busy_ms(ctx, true, false);
    ctx->stats[18] = busy_ms(ctx, true, true);
    return rc;
}

int eik_ctx_set_streams(eik_ctx* ctx, int32_t nstreams) {
    if (!ctx || nstreams < 1 || nstreams > 2) return EIK_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->split = nstreams == 2;
    return EIK_OK;
}

int eik_mem_info(eik_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes) {
    if (!ctx || !free_bytes || !total_bytes) return EIK_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    size_t f = 0, t = 0;
    CK(hipMemGetInfo(&f, &t));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return EIK_OK;
}

int eik_last_stats(const eik_ctx* ctx, double* out, int32_t nout) {
    if (!ctx || !out) return EIK_EINVAL;
    for (int i = 0; i < nout && i < kNumStats; i++) out[i] = ctx->stats[i];
    return EIK_OK;
}

}  // extern "C"
