// eik_internal.h -- types shared by the kernels and the C-ABI host code of libeik.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/eik.h"
#include "eik_spans.h"
#include "eik_stencil.h"

// ------------------------------------------------------------------ device-side descriptors
namespace eik {

// One grid solved by the tiled solver: a source's coarse grid or its refined near-source box.
struct InstDev {
    int32_t n[3];        // grid nodes per axis (n[2] = 1 in 2D)
    int32_t t[3];        // forward-solver tiles per axis
    int64_t tt_off;      // element offset of this grid in its traveltime pool
    int64_t s_off;       // element offset of its slowness in the slowness pool
    int64_t tile_base;   // first global forward tile id (tile flags / lists)
    int32_t at[3];       // adjoint tiles per axis
    int32_t pad3;
    int64_t atile_base;  // first global adjoint tile id
    double h[3];         // node spacing per axis (Cartesian hgrid; spherical dr, dtheta[deg], dphi[deg])
    double o[3];         // origin (cooinit)
    double dthrad;       // spherical: dtheta in radians
    double dphrad;       // spherical 3D: dphi in radians
    int32_t kind;        // EIK_GRID*
    int32_t pad;
    // adjoint: nodes that pop no later than the refinement stop form the system; the stop's pop
    // key (K, S, D) -- (+inf, -inf, 0) on coarse grids
    double tstopK;
    double tstopS;
    int32_t tstopD;
    int32_t pad2;
    const double* sinth; // 3D sphere: sin(theta_j) per theta index, computed on the host (else null)
    double band;         // band scheduler: tiles within this traveltime of the lowest pending one run
    double nband;        // the same for the nodes updated inside a tile
    double aband;        // adjoint rounds: tiles within this traveltime of the highest pending one run
    // mixed precision: nodes inside [xlo, xhi] (per axis) use the FP64 stencil (the near-source
    // region of an unrefined solve); empty when xlo > xhi
    int32_t xlo[3], xhi[3];
};

// whether node g lies in the instance's FP64 region
__host__ __device__ __forceinline__ bool in_fp64_region(const InstDev& I, const int* g) {
    return g[0] >= I.xlo[0] && g[0] <= I.xhi[0] && g[1] >= I.xlo[1] && g[1] <= I.xhi[1] && g[2] >= I.xlo[2] &&
           g[2] <= I.xhi[2];
}

// Refined box around one source (createfinegrid, src/EikSolv/eikmainfwd.jl:950-1104)
struct FineDev {
    int32_t i1c[3];      // coarse index of the box origin
    int32_t nc[3];       // coarse grid size (to index the coarse velocity)
    int32_t df;          // downscalefactor
    int32_t outmin[3];   // box clipped at the model's lower border along each axis
    int32_t outmax[3];   // ... upper border
    int32_t ncorn;       // source-box corners on the fine grid
    int32_t corn[8][3];
    int32_t cinst;       // coarse instance of the same source
    const double* vel;   // velocity of the source's model (coarse grid)
};

// A source (init) point of one instance: coarse source-box corner, transferred coarse node, or
// fine-grid corner.  node is the linear index inside the instance's grid.
struct SrcPt {
    int32_t inst;
    int32_t aux;    // transferred points: fine-grid node of the same position; -1 otherwise
    int64_t node;
};

// Taper box of one source (smoothgradaroundsrc!, src/EikSolv/utils.jl:121-338)
struct TaperDev {
    int32_t lo[3], hi[3];  // index box (may reach outside the grid)
    double src[3];
    double rmax;
    int32_t on;            // radius > 0
    int32_t pad;
};

// One source-box corner to initialise (sourceboxloctt!, eikmainfwd.jl:1195-1322)
struct CornerDev {
    int32_t inst;
    int32_t ijk[3];
    double src[3];
    double dist;  // distance corner-source, computed on the host (the same libm as the oracle)
    const double* vel;  // velocity of the instance's model (coarse grid)
};

// fine index -> index inside the coarse window, nearest-neighbour rule of createfinegrid
// (eikmainfwd.jl:1039-1041: round up when the remainder is >= df/2+1), 0-based
__host__ __device__ __forceinline__ int nn_coarse(int i0, int df) {
    const int i_f = i0 + 1;
    const int di = (i_f - 1) / df;
    const int ri = i_f - di * df;
    return ((double)ri >= (double)df / 2.0 + 1.0) ? di + 1 : di;
}

__host__ __device__ __forceinline__ double deg2rad(double x) { return x * (M_PI / 180.0); }

// distance2points (src/EikSolv/eikutils.jl:47-73); 3D sphere: chord length (this library's semantics)
__host__ __device__ __forceinline__ double distance2points(int kind, int nd, const double* a, const double* b) {
#pragma clang fp contract(off)
    if (kind == EIK_GRID2D_SPHERE) {
        const double r1 = a[0], t1 = a[1], r2 = b[0], t2 = b[1];
        return sqrt(r1 * r1 + r2 * r2 - 2.0 * r1 * r2 * cos(deg2rad(t1 - t2)));
    }
    if (kind == EIK_GRID3D_SPHERE) {
        const double xa = a[0] * sin(deg2rad(a[1])) * cos(deg2rad(a[2]));
        const double ya = a[0] * sin(deg2rad(a[1])) * sin(deg2rad(a[2]));
        const double za = a[0] * cos(deg2rad(a[1]));
        const double xb = b[0] * sin(deg2rad(b[1])) * cos(deg2rad(b[2]));
        const double yb = b[0] * sin(deg2rad(b[1])) * sin(deg2rad(b[2]));
        const double zb = b[0] * cos(deg2rad(b[1]));
        return sqrt((xa - xb) * (xa - xb) + (ya - yb) * (ya - yb) + (za - zb) * (za - zb));
    }
    double s = 0.0;
    for (int d = 0; d < nd; d++) s += (a[d] - b[d]) * (a[d] - b[d]);
    return sqrt(s);
}

// Tile edges: the forward solver relaxes one-wave tiles (8x8 in 2D, 4x4x4 in 3D); the adjoint
// solve uses 16x16 / 8x8x8 tiles.
__host__ __device__ constexpr int fwd_tile_edge(int nd) { return nd == 2 ? 8 : 4; }
__host__ __device__ constexpr int adj_tile_edge(int nd) { return nd == 2 ? 16 : 8; }

// mark the adjoint tile of node g and the face-neighbour tiles whose stencil reaches it (one
// byte per tile)
// Adjoint tile flags (one byte per tile) and pending values (float bits of the largest traveltime
// of a node whose right-hand side or lambda changed, rounded up: the adjoint runs from late to
// early traveltimes).  Seeding node g (traveltime t) flags its tile and the face neighbours within
// reach of its stencil.
__device__ __forceinline__ void seed_tiles(const InstDev& I, int nd, const int* g, uint8_t* flags, unsigned* apend,
                                           double t) {
    const int TX = adj_tile_edge(nd), TY = TX, TZ = nd == 2 ? 1 : TX;
    const int tx = g[0] / TX, ty = g[1] / TY, tz = nd == 3 ? g[2] / TZ : 0;
    const int64_t t0 = I.at[0], t01 = (int64_t)I.at[0] * I.at[1];
    const int64_t tl = tx + t0 * ty + t01 * tz;
    uint8_t* fl = flags + I.atile_base;
    unsigned* pd = apend + I.atile_base;
    const unsigned pv = __float_as_uint(__double2float_ru(fabs(t)));
    auto mark = [&](int64_t k) {
        fl[k] = 1;
        atomicMax(&pd[k], pv);
    };
    mark(tl);
    const int lx = g[0] - tx * TX, ly = g[1] - ty * TY, lz = nd == 3 ? g[2] - tz * TZ : 0;
    if (lx < 2 && tx > 0) mark(tl - 1);
    if (lx >= TX - 2 && tx < I.at[0] - 1) mark(tl + 1);
    if (ly < 2 && ty > 0) mark(tl - t0);
    if (ly >= TY - 2 && ty < I.at[1] - 1) mark(tl + t0);
    if (nd == 3) {
        if (lz < 2 && tz > 0) mark(tl - t01);
        if (lz >= TZ - 2 && tz < I.at[2] - 1) mark(tl + t01);
    }
}

// Forward tile flags (one word per tile): bits 0-5 = faces (x-, x+, y-, y+, z-, z+) whose 2-deep
// halo changed since the tile last ran, kFullBit = relax every node.  The pending value is the
// smallest changed key, as float bits rounded down.
constexpr unsigned kFullBit = 0x40u;
// kKeyBit: nodes were left pending at the band limit; relax every node whose key reaches the
// pending value and that has a reached neighbour.
constexpr unsigned kKeyBit = 0x80u;

// Seed node g with value val: its forward tile and the face neighbours whose stencil reaches it
// relax every node in their next round.
__device__ __forceinline__ void seed_tiles_fwd(const InstDev& I, int nd, const int* g, unsigned* flags,
                                               unsigned* pend, double val) {
    const int TX = fwd_tile_edge(nd), TY = TX, TZ = nd == 2 ? 1 : TX;
    const int tx = g[0] / TX, ty = g[1] / TY, tz = nd == 3 ? g[2] / TZ : 0;
    const int64_t t0 = I.t[0], t01 = (int64_t)I.t[0] * I.t[1];
    const int64_t tl = tx + t0 * ty + t01 * tz;
    const unsigned pv = __float_as_uint(__double2float_rd(fabs(val)));
    auto mark = [&](int64_t t) {
        atomicOr(&flags[I.tile_base + t], kFullBit);
        atomicMin(&pend[I.tile_base + t], pv);
    };
    mark(tl);
    const int lx = g[0] - tx * TX, ly = g[1] - ty * TY, lz = nd == 3 ? g[2] - tz * TZ : 0;
    if (lx < 2 && tx > 0) mark(tl - 1);
    if (lx >= TX - 2 && tx < I.t[0] - 1) mark(tl + 1);
    if (ly < 2 && ty > 0) mark(tl - t0);
    if (ly >= TY - 2 && ty < I.t[1] - 1) mark(tl + t0);
    if (nd == 3) {
        if (lz < 2 && tz > 0) mark(tl - t01);
        if (lz >= TZ - 2 && tz < I.t[2] - 1) mark(tl + t01);
    }
}

// per-node stencil spacing (eikforward2D.jl:115-119, eikforward3D.jl:168-172)
template <typename T>
__device__ __forceinline__ void node_spacing(const InstDev& I, const int* g, T* dh) {
#pragma clang fp contract(off)
    if (I.kind == EIK_GRID2D_CART || I.kind == EIK_GRID3D_CART) {
        dh[0] = dh[1] = dh[2] = (T)I.h[0];
        return;
    }
    const double r = fma((double)g[0], I.h[0], I.o[0]);
    dh[0] = (T)I.h[0];
    dh[1] = (T)(r * I.dthrad);
    if (I.kind == EIK_GRID3D_SPHERE) {
        dh[2] = (T)(r * I.sinth[g[1]] * I.dphrad);
    } else {
        dh[2] = (T)0;
    }
}

// Receiver interpolation on a traveltime grid (|value| strips the init-node sign).
// bilinear_interp (eikforward2D.jl:17-87); trilinear_interp (eikforward3D.jl:17-140) keeps the
// reference's weight xd*(1-yd)*(1-zd) for corner f101 so picks and gradients stay consistent
// with it.  coef/idx (optional) receive the weights and node indices in the reference's order.
template <typename T>
__host__ __device__ __forceinline__ double interp_receiver(const T* tt, const InstDev& I, int nd, int kind,
                                                          const double* pt, double* coef, int64_t* idx) {
#pragma clang fp contract(off)
    const bool sph = (kind == EIK_GRID2D_SPHERE || kind == EIK_GRID3D_SPHERE);
    if (nd == 2) {
        const double dx = I.h[0], dy = sph ? I.h[1] : I.h[0];
        const double xh = (pt[0] - I.o[0]) / dx, yh = (pt[1] - I.o[1]) / dy;
        int64_t i = (int64_t)floor(xh + 1.0), j = (int64_t)floor(yh + 1.0);
        const int64_t nx = I.n[0], ny = I.n[1];
        if (i == nx) i -= 1;
        else if (i > nx) i -= 2;
        if (j == ny) j -= 1;
        else if (j > ny) j -= 2;
        const double xd = xh - (double)(i - 1), yd = yh - (double)(j - 1);
        const int64_t l00 = (i - 1) + nx * (j - 1), l10 = l00 + 1, l01 = l00 + nx, l11 = l01 + 1;
        if (coef) {
            coef[0] = (1.0 - xd) * (1.0 - yd);
            coef[1] = (1.0 - yd) * xd;
            coef[2] = (1.0 - xd) * yd;
            coef[3] = xd * yd;
            idx[0] = l00;
            idx[1] = l10;
            idx[2] = l01;
            idx[3] = l11;
        }
        if (!tt) return 0.0;
        const double f00 = fabs((double)tt[l00]), f10 = fabs((double)tt[l10]), f01 = fabs((double)tt[l01]),
                     f11 = fabs((double)tt[l11]);
        return f00 * (1.0 - xd) * (1.0 - yd) + f10 * (1.0 - yd) * xd + f01 * (1.0 - xd) * yd + f11 * xd * yd;
    }
    const double dx = I.h[0], dy = sph ? I.h[1] : I.h[0], dz = sph ? I.h[2] : I.h[0];
    const double xh = (pt[0] - I.o[0]) / dx, yh = (pt[1] - I.o[1]) / dy, zh = (pt[2] - I.o[2]) / dz;
    const double x = pt[0] - I.o[0], y = pt[1] - I.o[1], z = pt[2] - I.o[2];
    int64_t i = (int64_t)floor(xh + 1.0), j = (int64_t)floor(yh + 1.0), k = (int64_t)floor(zh + 1.0);
    const int64_t nx = I.n[0], ny = I.n[1], nz = I.n[2];
    if (i == nx) i -= 1;
    else if (i > nx) i -= 2;
    if (j == ny) j -= 1;
    else if (j > ny) j -= 2;
    if (k == nz) k -= 1;
    else if (k > nz) k -= 2;
    const double x0 = (double)(i - 1) * dx, y0 = (double)(j - 1) * dy, z0 = (double)(k - 1) * dz;
    const double x1 = (double)i * dx, y1 = (double)j * dy, z1 = (double)k * dz;
    const double xd = (x - x0) / (x1 - x0), yd = (y - y0) / (y1 - y0), zd = (z - z0) / (z1 - z0);
    const int64_t nxy = nx * ny;
    const int64_t l000 = (i - 1) + nx * (j - 1) + nxy * (k - 1);
    const int64_t l100 = l000 + 1, l010 = l000 + nx, l001 = l000 + nxy, l110 = l010 + 1, l011 = l010 + nxy,
                  l101 = l001 + 1, l111 = l011 + 1;
    if (coef) {
        coef[0] = (1 - xd) * (1 - yd) * (1 - zd);
        coef[1] = xd * (1 - yd) * (1 - zd);
        coef[2] = (1 - xd) * yd * (1 - zd);
        coef[3] = (1 - xd) * (1 - yd) * zd;
        coef[4] = xd * yd * (1 - zd);
        coef[5] = (1 - xd) * yd * zd;
        coef[6] = xd * (1 - yd) * (1 - zd);
        coef[7] = xd * yd * zd;
        idx[0] = l000;
        idx[1] = l100;
        idx[2] = l010;
        idx[3] = l001;
        idx[4] = l110;
        idx[5] = l011;
        idx[6] = l101;
        idx[7] = l111;
    }
    if (!tt) return 0.0;
    const double f000 = fabs((double)tt[l000]), f100 = fabs((double)tt[l100]), f010 = fabs((double)tt[l010]),
                 f001 = fabs((double)tt[l001]), f110 = fabs((double)tt[l110]), f011 = fabs((double)tt[l011]),
                 f101 = fabs((double)tt[l101]), f111 = fabs((double)tt[l111]);
    return f000 * (1 - xd) * (1 - yd) * (1 - zd) + f100 * xd * (1 - yd) * (1 - zd) +
           f010 * (1 - xd) * yd * (1 - zd) + f001 * (1 - xd) * (1 - yd) * zd + f110 * xd * yd * (1 - zd) +
           f011 * (1 - xd) * yd * zd + f101 * xd * (1 - yd) * (1 - zd) + f111 * xd * yd * zd;
}

}  // namespace eik

// ------------------------------------------------------------------ host-side objects
// A fixed block of device memory handed out in 256-byte aligned pieces and never freed piecewise
// (the no-allocation entry point's scratch); reset() starts a call over from an empty block.
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    void* carve(size_t n) {
        const size_t at = (used + 255) & ~(size_t)255;
        if (!base || at > cap || n > cap - at) return nullptr;
        used = at + n;
        return base + at;
    }
};

// A named device buffer of a context: grows with hipMalloc, or, when the context runs in arena
// mode, is carved from the arena (no allocation).
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    Arena* arena = nullptr;
    int ensure(size_t n) {
        if (n <= cap && p) return 0;
        const size_t want = n ? n : 16;
        if (arena) {
            p = arena->carve(want);
            cap = p ? want : 0;
            return p ? 0 : -1;
        }
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (hipMalloc(&p, want) != hipSuccess) {
            p = nullptr;
            return -1;
        }
        cap = want;
        return 0;
    }
    void release() {
        if (p && !arena) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// solver rounds enqueued between two host checks for convergence
constexpr int kRoundBatch = 8;
// rounds after which a traveltime solve starts settling (see fim_run_t)
constexpr int kSettleRound = 2000;
// device counters (ctx->d_small, 32-bit words): flagged tiles, sweeps, non-empty rounds; from word
// kCntSlots on, three round slots of kSlotWords words (list length, list cursor; a round uses slot
// round mod 3); from word kCnt64 on, 64-bit node updates, changed updates, tile activations
constexpr size_t kSmallBytes = 256;
constexpr int kCntFlagged = 2, kCntSweeps = 3, kCntRounds = 4, kCntSlots = 8, kCnt64 = 16;
constexpr int kSlotList = 0, kSlotCursor = 1, kSlotWords = 2;

// eik_last_stats entries
constexpr int kNumStats = 19;
// one round launch: start and end (ms after the call's base event), near-source (fine) solve or not
struct KSpan {
    float a, b;
    bool fine;
};

struct eik_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<std::string, DevBuf> bufs;
    void* d_small = nullptr;  // device counters
    void* h_small = nullptr;  // pinned host mirror of the counters
    hipEvent_t ev[2] = {nullptr, nullptr};   // phase timer
    hipEvent_t kev[2 * kRoundBatch] = {};    // round-kernel timers (start, end) of one round batch
    double stats[kNumStats] = {};
    // Round-launch spans of the current call (ms after tbase, an event recorded when the call
    // starts): with the sources split over this context and its peer, the launches of the two
    // streams overlap, and the union of the spans is the time the round kernel held the GPU.
    hipEvent_t tbase = nullptr, tend = nullptr;
    std::vector<KSpan> kspans;
    bool span_fine = false;  // the launches being timed belong to the near-source solve
    // a second context on its own stream that solves part of a call's sources concurrently
    // (created on first use, owned by this context, only used while this context's lock is held)
    eik_ctx* peer = nullptr;
    eik_ctx* owner = nullptr;  // set on a peer: the context it belongs to (whose tbase it shares)
    bool split = true;         // large calls may split their sources with the peer
    bool in_split = false;     // the current call is split (set on the context and its peer)
    int xcap = 1;  // capacity per source of the transferred-point list
    std::mutex mu;
    // arena mode (the no-allocation entry point): buffers are carved from this block instead
    Arena* arena = nullptr;
    std::map<std::string, DevBuf> abufs;
    DevBuf& buf(const char* name) {
        if (!arena) return bufs[name];
        DevBuf& b = abufs[name];
        b.arena = arena;
        return b;
    }
    int fail(int code, const char* fmt, ...) {
        char b[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(b, sizeof(b), fmt, ap);
        va_end(ap);
        err = b;
        return code;
    }
};

namespace eik {
// device bytes a context's reusable pools hold (they grow to the largest call seen and are kept)
inline int64_t ctx_pool_bytes(const eik_ctx* ctx) {
    int64_t b = 0;
    for (const auto& kv : ctx->bufs) b += (int64_t)kv.second.cap;
    return b + (ctx->peer ? ctx_pool_bytes(ctx->peer) : 0);
}
// a call's statistics start from zero and its round-launch spans are timed from now (a peer's
// from its owner's base event, which the owner records)
inline void stats_begin(eik_ctx* ctx) {
    for (double& x : ctx->stats) x = 0.0;
    ctx->kspans.clear();
    if (ctx->tbase && !ctx->owner) (void)hipEventRecord(ctx->tbase, ctx->stream);
}
// device time of the call so far (from its tbase, on the context stream; waits for the stream)
inline double call_ms(eik_ctx* ctx) {
    float ms = 0.f;
    if (!ctx->tbase || !ctx->tend || hipEventRecord(ctx->tend, ctx->stream) != hipSuccess ||
        hipEventSynchronize(ctx->tend) != hipSuccess || hipEventElapsedTime(&ms, ctx->tbase, ctx->tend) != hipSuccess)
        return 0.0;
    return (double)ms;
}
// length of the union of the round-launch spans of a call on this context and its peer (ms):
// the coarse solve's launches, and the near-source solve's too if `fine`
inline double busy_ms(const eik_ctx* ctx, bool with_peer, bool fine) {
    std::vector<std::pair<float, float>> v;
    for (const eik_ctx* c : {ctx, with_peer ? (const eik_ctx*)ctx->peer : nullptr})
        if (c)
            for (const KSpan& k : c->kspans)
                if (fine || !k.fine) v.emplace_back(k.a, k.b);
    return span_union(std::move(v));
}
// release a context's reusable pools once its stream is idle (the next call allocates again)
inline void ctx_trim(eik_ctx* ctx) {
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->bufs) kv.second.release();
    if (ctx->peer) ctx_trim(ctx->peer);
}
}  // namespace eik

struct eik_model {
    eik_ctx* ctx = nullptr;
    eik_grid g{};
    int64_t N = 0;
    double* d_vel = nullptr;
    double* d_slow64 = nullptr;
    float* d_slow32 = nullptr;
    double smean = 1.0;  // mean slowness (band scheduler scale)
    double* d_sinth = nullptr;  // 3D sphere: sin(theta) per theta index
};

#define CK(x)                                                                                    \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) return ctx->fail(EIK_EDEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

namespace eik {

struct SourceSet {
    int64_t nsrc = 0;
    int nd = 2;
    std::vector<double> src;      // nsrc*nd
    std::vector<int64_t> rec_off; // nsrc+1
    std::vector<double> rec;      // nrec*nd
    int64_t nrec() const { return rec_off.empty() ? 0 : rec_off.back(); }
};

// elapsed device time between marks on the context stream (one Timer at a time per context: they
// share its two events; a call's total is call_ms)
struct Timer {
    eik_ctx* c;
    explicit Timer(eik_ctx* ctx) : c(ctx) {
        (void)hipEventRecord(c->ev[0], c->stream);
    }
    double mark() {
        (void)hipEventRecord(c->ev[1], c->stream);
        (void)hipEventSynchronize(c->ev[1]);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
        std::swap(c->ev[0], c->ev[1]);
        return (double)ms;
    }
};

// host geometry (eik_geom.cpp)
void make_inst(const eik_grid& g, InstDev& I);
void set_tiles(InstDev& I, int nd);  // forward and adjoint tile counts from I.n
int find_enclosing_box(const eik_grid& g, const double* src, int corn[8][3]);
void build_fine(const eik_grid& g, const double* src, int df, int noderadius, FineDev& F, InstDev& I);

// forward (eik_fim.hip)
// The sources of several models solved in shared launches (models of one grid kind): instances
// are numbered group by group, source by source.
struct FwdGroup {
    eik_model* M;
    const SourceSet* S;
};
int forward_solve(eik_ctx* ctx, const std::vector<FwdGroup>& G, int prec, const eik_params* P);
int forward_picks(eik_ctx* ctx, const std::vector<FwdGroup>& G, int prec, double* picks, double* ttgrid);
inline int forward_solve(eik_ctx* ctx, eik_model* M, const SourceSet& S, int prec, const eik_params* P) {
    return forward_solve(ctx, std::vector<FwdGroup>{{M, &S}}, prec, P);
}
inline int forward_picks(eik_ctx* ctx, eik_model* M, const SourceSet& S, int prec, double* picks, double* ttgrid) {
    return forward_picks(ctx, std::vector<FwdGroup>{{M, &S}}, prec, picks, ttgrid);
}
int model_slowness(eik_model* M);
// A gradient call in two parts: the per-source part (forward, adjoint, source terms and the
// per-source gradient images, misfit terms per receiver), then the sums over sources in source
// order (misfit, gradvel, smoothing), over the parts of one call (split over two contexts) in turn.
struct GradPart {
    double* img;           // nsrc gradient images (N doubles each), device
    const double* mterm;   // nrec misfit terms, device
    int64_t nsrc, nrec;
    // a part that continues another one leaves the source-node carry-over and the taper of its
    // images to gradient_total (they chain from the previous part's last source): its device
    // instances, source points, their offsets, tapers and source terms
    bool deferred;
    const void *inst, *pts, *poff, *tap, *t2, *t1;
};
int gradient_part(eik_ctx* ctx, eik_model* M, const SourceSet& S, const eik_params* P, int whichgrad,
                  const double* pickobs, const double* stdobs, double* gradsrc, bool defer, GradPart* out);
int gradient_total(eik_ctx* ctx, eik_model* M, const eik_params* P, int whichgrad, const GradPart* parts, int nparts,
                   const int64_t* rec_off, int64_t nsrc, double* gradvel, double* misf);
int gradient_solve(eik_ctx* ctx, eik_model* M, const SourceSet& S, const eik_params* P, int whichgrad,
                   const double* pickobs, const double* stdobs, double* gradvel, double* gradsrc, double* misf);
// fill kernel (eik_fim.hip), also used by the adjoint rounds
__global__ void k_fill_u32(unsigned* __restrict__ p, int64_t n, unsigned v);
struct FimStats {
    int64_t rounds = 0;        // solver rounds (one launch each)
    double activations = 0;    // tile activations (blocks launched)
    double sweeps = 0;         // local relaxation sweeps summed over activations
    double updates = 0;        // node updates (stencil replays) summed over activations
    double changed = 0;        // node updates that changed the node
    double keychanged = 0;     // of those, the ones that changed only the pop key
    int settled = 0;           // settling stage reached (0: exact; k: tolerance 10^(k-1) x base)
    int verifications = 0;     // verification rounds (every tile relaxed once more at convergence)
    double kernel_ms = 0;      // summed duration of the round launches
    int64_t launches = 0;      // round-kernel launches (rounds that found no tile included)
};
// traveltime pool plus its pop-key pools (K, S in the pool's precision, D as bytes)
struct TTPools {
    void* tt;
    void* k;
    void* s;
    uint8_t* d;
};
// near: some instance has an FP64 region (unrefined mixed solve; InstDev::xlo/xhi)
int fim_run(eik_ctx* ctx, int nd, int prec, int kind, const InstDev* d_inst, int ninst, int64_t ntiles,
            const TTPools& pools, const void* slowpool, bool exact, bool near, FimStats* fs);

// no-allocation scratch (eik_noalloc.hip): resolve this device's copy once (context creation)
void scratch_setup(int device);

// argument checks shared by the entry points (eik_capi.hip): checksrcrecposition and the
// ExtraParams checks; the parameters a call actually runs with (mixed precision's FP64 box)
int load_sources(eik_ctx* ctx, eik_model* M, int64_t nsrc, const double* coordsrc, const int64_t* rec_off,
                 const double* coordrec, const eik_params* P, SourceSet& S);
eik_params effective_params(const eik_params* p, int kind);
// nodes within this many of the source box stay FP64 in the reduced precisions
constexpr int32_t kNearSourceRadius = 4;
// The precision a grid kind runs with: Grid2DSphere keeps the FP64 stencil under the mixed
// precision (its discrete adjoint amplifies rounding-level traveltime changes ~1e4-fold, so FP32
// increments move its gradients beyond 1e-3).
inline int run_precision(int prec, int kind) {
    return (prec == EIK_PREC_MIXED && kind == EIK_GRID2D_SPHERE) ? EIK_PREC_FP64 : prec;
}

}  // namespace eik
