// eik_fim.hip -- forward traveltime engine of libeik (gfx950).
//
// Tiled fast-iterative solver whose local update replays the reference's heap order (see
// eik_stencil.h), batched over many sources and over the per-source refined grids in one launch
// per round.  Layout in HBM:
//   model : vel (FP64), slowness (FP64 and FP32), N = prod(grsize) each, index 1 fastest;
//   pools : one traveltime grid per source instance (FP64, or FP32 in mixed precision), and one
//           FP64 traveltime + FP64 slowness grid per refined near-source box.
// Init (source) nodes are stored with the sign bit set (value -t); they count as accepted from
// the start and never change.  Unreached nodes hold +inf.
#include <array>

#include "eik_internal.h"

namespace eik {

// ================================================================= tile geometry
// A tile is relaxed by one wavefront, one node per lane: 8x8 nodes in 2D, 4x4x4 in 3D, staged in
// LDS with a 2-deep halo on every face (the 2nd-order stencil reads two nodes along each axis;
// edges and corners of the halo are never read).  Small tiles keep the relaxation close to the
// front (fewer repeated updates) and let many independent tiles share a CU.
template <int ND>
struct TileShape;
template <>
struct TileShape<2> {
    static constexpr int TX = 8, TY = 8, TZ = 1, H = 2;
    static constexpr int LX = TX + 2 * H, LY = TY + 2 * H, LZ = 1;
    static constexpr int NODES = TX * TY;
    static constexpr int MAXIT = 64;
};
template <>
struct TileShape<3> {
    static constexpr int TX = 4, TY = 4, TZ = 4, H = 2;
    static constexpr int LX = TX + 2 * H, LY = TY + 2 * H, LZ = TZ + 2 * H;
    static constexpr int NODES = TX * TY * TZ;
    static constexpr int MAXIT = 64;
};
static_assert(TileShape<2>::TX == fwd_tile_edge(2) && TileShape<3>::TX == fwd_tile_edge(3), "tile edge");
constexpr int kTileThreads = 64;
// A run of consecutive list entries claimed by one block is at most 1/kRunDiv of what is left per
// block (and at most the EIK_CLAIM maximum).  Longer runs relax more tiles right after their
// neighbour and claim less often; shorter ones let the blocks finish together.  Measured against
// 4: 2 is 7 % faster on 201x201x101 x 16 sources (FP64; mixed 5 %, FP32 4 %) and 4.5 % on 600x400 x
// 256; with the x neighbours' activations merged into the run (k_fim_round), 1 is another 1-3 %
// faster on 201x201x101 and 1.5 % on the 120x150x150 spherical grid (the maximum of 8 entries
// binds; 12 is slower).
constexpr unsigned kRunDiv = 1;
// Waves per SIMD the register allocation targets.  The kernel is bound by VALU/SALU issue, not
// by latency: forcing 3 waves (168 registers) on the FP32 3D kernel measured 8 % slower than 2.
template <int ND, typename T>
struct RoundOcc {
    static constexpr int value = 2;
};

__device__ __forceinline__ int find_inst(const InstDev* __restrict__ inst, int ninst, int64_t gid) {
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (inst[mid].tile_base <= gid)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Minimum of v over the 64 lanes of the wave, wave-uniform; every lane must be active.  Quad
// swaps and row rotations (DPP) reduce each row of 16 lanes, and the four row minima are read out
// to scalar registers.  (An LDS atomicMin per lane instead compiles to a loop over the active
// lanes, several scalar instructions per lane.)
__device__ __forceinline__ unsigned wave_min_u32(unsigned v) {
    v = min(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true));   // quad_perm [1,0,3,2]
    v = min(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true));   // quad_perm [2,3,0,1]
    v = min(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x124, 0xF, 0xF, true));  // row_ror:4
    v = min(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, true));  // row_ror:8
    const unsigned a = (unsigned)__builtin_amdgcn_readlane((int)v, 0), b = (unsigned)__builtin_amdgcn_readlane((int)v, 16);
    const unsigned c = (unsigned)__builtin_amdgcn_readlane((int)v, 32), d = (unsigned)__builtin_amdgcn_readlane((int)v, 48);
    return min(min(a, b), min(c, d));
}

template <typename T>
__device__ __forceinline__ bool is_fixed(T raw) {
    return signbit(raw);
}

// Settling: a change whose value and key move by at most `settle` relative to the old ones is
// written but wakes no other node.  In FP64 with `fast` settle = 1e-12 from the first round
// (most of those changes are rounding ripples that would re-wake the whole cone downstream);
// otherwise 0 until the solve has run far more rounds than any converging one: on rare models the
// replayed update then circulates changes of ~1e-13 relative through a few nodes instead of
// reaching the heap's fixed point (see fim_run_t).
template <typename T>
__device__ __forceinline__ bool tiny_change(T v0, const Key<T>& k0, T v1, const Key<T>& k1, T settle) {
    if (!(settle > T(0)) || !(fabs(v0) < tinf<T>())) return false;
    return fabs(fabs(v1) - fabs(v0)) <= settle * fabs(v0) && fabs(k1.K - k0.K) <= settle * fabs(k0.K);
}

// A pop key K as a float no larger than it (changes and pending values are compared against node
// keys conservatively, so rounding must never move them up).  Keys here are >= 0 or +inf: the
// nearest float, stepped one unit towards zero when it lies above K, is K rounded down (+inf stays).
__device__ __forceinline__ float key_down(double k) {
    const float f = (float)k;
    return (double)f > k ? __uint_as_float(__float_as_uint(f) - 1u) : f;
}
__device__ __forceinline__ float key_down(float k) { return k; }
// ... and a node's own key as a float no smaller than it: a trigger c <= key_up(K) includes every
// c <= K (at most a spurious re-evaluation, which leaves a node whose inputs did not change as it
// is).  For K >= 0: the nearest float, stepped one unit up when it lies below K.
__device__ __forceinline__ float key_up(double k) {
    const float f = (float)k;
    return (double)f < k ? __uint_as_float(__float_as_uint(f) + 1u) : f;
}
__device__ __forceinline__ float key_up(float k) { return k; }

// LDS image of a tile.  Compact layout (the pop keys K and S, which only the rare cascade path
// reads around a node): the NODES core nodes (x fastest), then for each face k = 2*axis + l (l = 0:
// + side, l = 1: - side; the neighbour-slot numbering of eik_stencil.h) its two halo layers of FA
// nodes each, nearer layer first; the edges and corners of the halo are never read by the stencil
// and get no slot (256 slots instead of 8^3 for a 3D tile).  The staging loads go by compact slot.
template <int ND>
struct TileLds {
    static constexpr int E = TileShape<ND>::TX;  // tiles are squares / cubes
    static constexpr int NN = TileShape<ND>::NODES;
    static constexpr int FA = NN / E;
    static constexpr int NL = NN + 4 * ND * FA;
    static constexpr int stride(int ax) { return ax == 0 ? 1 : (ax == 1 ? E : E * E); }
    __device__ static int coord(int n, int ax) { return ax == 0 ? n % E : (ax == 1 ? (n / E) % E : n / (E * E)); }
    // the coordinates of core node n other than `ax`, as an index into a face layer
    __device__ static int face_index(int n, int ax) {
        if (ax == 0) return n / E;
        if (ax == 1) return n % E + E * (n / (E * E));
        return n % (E * E);
    }
    // slot of the node d steps (d = 1, 2) from core node n across face k
    __device__ static int nbr(int n, int k, int d) {
        const int ax = k >> 1, l = k & 1;
        const int c = coord(n, ax);
        const bool in = l ? c - d >= 0 : c + d < E;
        const int q = l ? d - 1 - c : c + d - E;
        return in ? n + (l ? -d : d) * stride(ax) : NN + (2 * ax + l) * 2 * FA + q * FA + face_index(n, ax);
    }
    // The values, depth bytes and change keys sit in a padded box instead: the core plus two
    // layers on every side (edges and corners unused), so every stencil neighbour of a core node
    // lies at a fixed offset from its slot and a lane reads all of them from one address register
    // with immediate offsets.
    static constexpr int P = E + 4;
    static constexpr int PV = ND == 3 ? P * P * P : P * P;
    static constexpr int pstride(int ax) { return ax == 0 ? 1 : (ax == 1 ? P : P * P); }
    __device__ static int pslot(int x, int y, int z) { return (x + 2) + P * (y + 2) + (ND == 3 ? P * P * (z + 2) : 0); }
    // padded-box offset of the node d steps across face k
    static constexpr int poff(int k, int d) { return ((k & 1) ? -d : d) * pstride(k >> 1); }
    // tile-local position of halo slot c (NN <= c < NL); coordinates may be -2, -1, E, E + 1
    __device__ static void halo_pos(int c, int* p) {
        const int r = c - NN;
        const int f = r / (2 * FA), q = (r / FA) & 1, u = r % FA;
        const int ax = f >> 1, l = f & 1;
        const int a1 = ax == 0 ? 1 : 0, a2 = ax == 2 ? 1 : 2;  // the other axes, in order
        p[0] = p[1] = p[2] = 0;
        p[ax] = l ? -1 - q : E + q;
        p[a1] = u % E;
        if (ND == 3) p[a2] = u / E;
    }
};

// One round: every listed tile is relaxed in LDS until nothing within the band is left.  One
// wavefront per tile, one node per lane (lane n holds the same tile-local node in every tile, so
// its LDS slots and stencil neighbours are fixed for the launch).  Each node carries a pending
// trigger key (+inf: nothing pending):
//   at entry: 0 for every node if the tile was seeded or did not converge last time; else the
//            tile's pending value, for nodes whose key reaches it and that lie within two of a
//            face whose halo changed (or, after a band cut, that have a reached neighbour);
//   later:   the smallest key (before or after the change) of a stencil node changed in the
//            previous sweep, if it does not exceed the node's own key.
// A change whose old and new keys both exceed a node's key leaves that node's heap replay as it
// is (the replay stops before reaching such a neighbour), which is why the key filter is exact.
// A sweep updates the nodes whose trigger lies within the instance's band (lowest pending tile
// value + band width); nodes beyond it stay pending and the tile is queued again with their
// smallest trigger.  Faces whose 2-deep layer changed activate the neighbour tile with that
// face's bit and the smallest changed key as its pending value.  Blocks take list entries
// (local tile, instance, pending value, face bits) from a cursor, claiming the next entry while
// the current tile runs.  SPH: spherical grids (per-node metric); Cartesian grids share one
// metric, held in scalar registers.  MIX: the mixed precision (FP64 values, keys and comparisons,
// the quadratic's increment in FP32; eik_stencil.h).  NEAR: some instance has an FP64 region (an
// unrefined mixed solve) whose nodes take the FP64 stencil; compiled only where it can occur, since
// that second replay costs registers.
template <int ND, typename T, bool SPH, bool MIX, bool NEAR>
__global__ void __attribute__((amdgpu_flat_work_group_size(kTileThreads, kTileThreads),
                               amdgpu_waves_per_eu(RoundOcc<ND, T>::value))) k_fim_round(
    const uint4* __restrict__ list, unsigned* __restrict__ count, unsigned* __restrict__ slot,
    unsigned* __restrict__ slot_reset, const InstDev* __restrict__ inst, int ninst, const unsigned* __restrict__ wmin,
    unsigned* __restrict__ wnext, unsigned* __restrict__ wreset, T* __restrict__ ttpool, T* __restrict__ kpool,
    T* __restrict__ spool, uint8_t* __restrict__ dpool, const T* __restrict__ slowpool, unsigned* __restrict__ flags,
    unsigned* __restrict__ pend, unsigned long long* __restrict__ counters, T settle, unsigned chunk) {
    using S = TileShape<ND>;
    using L = TileLds<ND>;
    using A = Arith<T, MIX>;  // arithmetic of the quadratic
    constexpr int NN = S::NODES, NL = L::NL, E = L::E;
    static_assert(NN == kTileThreads, "one node per lane");
    static_assert(NL % NN == 0, "staging loop");
    constexpr int PV = L::PV;
    // The tile image.  The key arrays come first: the padded arrays then start at least two
    // layers' worth of bytes into LDS, so each neighbour's offset from a lane's slot, plus the
    // array's own address, is a non-negative immediate.
    struct Image {
        T k[NL];              // pop key K (the value itself outside cascades), compact layout
        T s[NL];              // pop key S (cascade nodes), compact layout
        T t[PV];              // value, sign bit = init node
        float chg[2][PV];     // smaller of old/new key of a core node changed in a sweep; +inf elsewhere
        uint8_t d[PV];        // pop key D
    };
    __shared__ Image img;
    __shared__ uint4 s_run[32];  // list entries of the claimed run
    T* const s_k = img.k;
    T* const s_s = img.s;
    T* const s_t = img.t;
    uint8_t* const s_d = img.d;
    const int n = threadIdx.x;
    const int cx = L::coord(n, 0), cy = L::coord(n, 1), cz = ND == 3 ? L::coord(n, 2) : 0;
    const int ps = L::pslot(cx, cy, cz);  // this lane's node in the padded arrays
    // faces this node lies within two of: bit 2a = - side of axis a, bit 2a + 1 = + side
    unsigned fmask = (unsigned)(cx < 2) | (unsigned)(cx >= E - 2) << 1 | (unsigned)(cy < 2) << 2 |
                     (unsigned)(cy >= E - 2) << 3;
    if (ND == 3) fmask |= (unsigned)(cz < 2) << 4 | (unsigned)(cz >= E - 2) << 5;
    const unsigned nlist = slot[kSlotList];
    unsigned tot_upd = 0u, tot_vchg = 0u, tot_kchg = 0u, tot_act = 0u, tot_sweeps = 0u;
    if (blockIdx.x == 0) {
        if (n == 0 && nlist) atomicAdd(&count[kCntRounds], 1u);
        // the buffers of the round after next (last used by the previous round, which has ended)
        for (int i = n; i < ninst; i += kTileThreads) wreset[i] = 0x7f800000u;
        if (n == 0) {
            slot_reset[kSlotList] = 0u;
            slot_reset[kSlotCursor] = 0u;
        }
    }
    // change keys outside the core stay +inf (core slots are reset per tile)
    for (int i = n; i < 2 * PV; i += kTileThreads) (&img.chg[0][0])[i] = EIK_INFF;
    // Entries are claimed in runs of consecutive list entries, which are neighbouring tiles (the
    // list is built in tile order): a block relaxing a tile right after its neighbour sees that
    // neighbour's new values, which saves work.  Runs shrink towards the end of the list (at most
    // `chunk`, at least 1, what is left per block) so the blocks finish
    // together.  Inside a run the tiles go in increasing pending value (upstream first).  The next
    // run is claimed while the last tile of the current one runs.  The run's entries (at most 32)
    // are read once into LDS, and choosing the next tile is a wave minimum over them.
    const unsigned gsz = gridDim.x;
    auto run_len = [&](unsigned at) {
        const unsigned left = at < nlist ? nlist - at : 0u;
        const unsigned r = left / (kRunDiv * gsz);
        return r < 1u ? 1u : (r > chunk ? chunk : r);
    };
    unsigned claim = 0u, want = run_len(0u);
    if (n == 0) claim = atomicAdd(&slot[kSlotCursor], want);
    unsigned rbase = (unsigned)__builtin_amdgcn_readfirstlane((int)claim);
    unsigned rlen = rbase < nlist ? min(want, nlist - rbase) : 0u, rdone = 0u;
    if ((unsigned)n < rlen) s_run[n] = list[rbase + n];
    __syncthreads();
    while (rbase < nlist) {
        // the run's next tile: lowest pending value among those not yet relaxed (lowest lane on ties)
        const bool open = (unsigned)n < rlen && !((rdone >> (n & 31)) & 1u);
        const unsigned mine = open ? s_run[n].z : 0xffffffffu;
        const unsigned best = wave_min_u32(mine);  // lanes 0-31 hold the run
        const unsigned long long cand = __ballot(open && mine == best);
        const int bk = cand ? __ffsll((long long)cand) - 1 : 0;
        rdone |= 1u << bk;
        const uint4 e = s_run[bk];
        const bool last = (unsigned)__popc(rdone) == rlen;
        const InstDev& I = inst[e.y];
        const float pend0 = __uint_as_float(e.z);
        const unsigned fbits = e.w;
        const float lim = __uint_as_float(wmin[e.y]) + (float)I.nband;
        const unsigned tl = e.x;
        const unsigned tix = tl % (unsigned)I.t[0];
        const unsigned tiy = (tl / (unsigned)I.t[0]) % (unsigned)I.t[1];
        const unsigned tiz = ND == 3 ? tl / ((unsigned)I.t[0] * (unsigned)I.t[1]) : 0u;
        const int ox = (int)tix * E, oy = (int)tiy * E, oz = (int)tiz * E;
        T* __restrict__ tt = ttpool + I.tt_off;
        T* __restrict__ kk = kpool + I.tt_off;
        T* __restrict__ ss = spool + I.tt_off;
        uint8_t* __restrict__ dd = dpool + I.tt_off;
        const T* __restrict__ slow = slowpool + I.s_off;
        const int nx = I.n[0], ny = I.n[1], nz = ND == 3 ? I.n[2] : 1;
        const int64_t nxy = (int64_t)nx * ny;

        // Stage the core and the face halos (this lane's own node first).  The values and depth
        // bytes of the whole tile image are loaded before any is used, from addresses clamped into
        // the grid (slots outside it are replaced by an unreached node afterwards): one memory round
        // trip.  The key pools are read only for cascade nodes (depth byte set), which are rare: a
        // wave with none makes no second trip and reads no key bytes.  An ordinary node's key is
        // its value.
        constexpr int NC = NL / NN - 1;  // halo chunks of NN slots
        const int g[3] = {ox + cx, oy + cy, oz + cz};
        const bool inside = g[0] < nx && g[1] < ny && g[2] < nz;
        const int64_t gl = g[0] + (int64_t)nx * g[1] + nxy * g[2];
        const int64_t glc = min(g[0], nx - 1) + (int64_t)nx * min(g[1], ny - 1) + nxy * min(g[2], nz - 1);
        T v0 = tt[glc], sl = slow[glc];
        int d0 = dd[glc];
        T hv[NC];
        int hd[NC];
        int64_t hl[NC];
        bool hin[NC];
#pragma unroll
        for (int h = 0; h < NC; h++) {
            int p[3];
            L::halo_pos((h + 1) * NN + n, p);
            const int hx = ox + p[0], hy = oy + p[1], hz = oz + p[2];
            hin[h] = hx >= 0 && hx < nx && hy >= 0 && hy < ny && hz >= 0 && hz < nz;
            hl[h] = min(max(hx, 0), nx - 1) + (int64_t)nx * min(max(hy, 0), ny - 1) + nxy * min(max(hz, 0), nz - 1);
            hv[h] = tt[hl[h]];
            hd[h] = dd[hl[h]];
        }
        if (!inside) {
            v0 = tinf<T>();
            sl = T(0);
            d0 = 0;
        }
        Key<T> k0{fabs(v0), -tinf<T>(), d0};
        if (d0) {
            k0.K = kk[glc];
            k0.S = ss[glc];
        }
        T hk[NC], hs[NC];
#pragma unroll
        for (int h = 0; h < NC; h++) {
            if (!hin[h]) {
                hv[h] = tinf<T>();
                hd[h] = 0;
            }
            hk[h] = fabs(hv[h]);
            hs[h] = -tinf<T>();
            if (hd[h]) {
                hk[h] = kk[hl[h]];
                hs[h] = ss[hl[h]];
            }
        }
#pragma unroll
        for (int h = 0; h < NC; h++) {
            const int c = (h + 1) * NN + n;
            int p[3];
            L::halo_pos(c, p);
            const int q = L::pslot(p[0], p[1], p[2]);
            s_t[q] = hv[h];
            s_k[c] = hk[h];
            s_s[c] = hs[h];
            s_d[q] = (uint8_t)hd[h];
        }
        s_t[ps] = v0;
        s_k[n] = k0.K;
        s_s[n] = k0.S;
        s_d[ps] = (uint8_t)k0.D;
        img.chg[0][ps] = EIK_INFF;
        img.chg[1][ps] = EIK_INFF;
        // stencil neighbours inside the grid (ok1) and usable 2-away nodes (ok2), per slot k
        unsigned ok1 = 0u, ok2 = 0u;
#pragma unroll
        for (int k = 0; k < 2 * ND; k++) {
            const int ax = k >> 1;
            int nb[3] = {g[0], g[1], g[2]};
            nb[ax] += (k & 1) ? -1 : 1;
            const int lim_ax = ax == 0 ? nx : (ax == 1 ? ny : nz);
            const bool in1 = nb[ax] >= 0 && nb[ax] <= lim_ax - 1;
            bool in2 = in1 && nb[0] >= 1 && nb[0] <= nx - 2 && nb[1] >= 1 && nb[1] <= ny - 2;
            if (ND == 3) in2 = in2 && nb[2] >= 1 && nb[2] <= nz - 2;
            ok1 |= (unsigned)in1 << k;
            ok2 |= (unsigned)in2 << k;
        }
        __syncthreads();
        // the next run is claimed once the image is in LDS: the atomic's round trip then overlaps
        // the sweeps, which touch no global memory (claimed earlier, waiting for the image would
        // also wait for it: memory returns are counted in order)
        if (last) {
            want = run_len(rbase + rlen);
            if (n == 0) claim = atomicAdd(&slot[kSlotCursor], want);
        }
        bool act = inside && !is_fixed(v0);
        // plain image: no init node and no cascade key anywhere in the core or the halo.  The
        // sweeps then read values only (no sign bits to strip, no depth bytes); it ends with the
        // first cascade key a sweep writes.
        bool odd_node = is_fixed(v0) || d0 != 0;
#pragma unroll
        for (int h = 0; h < NC; h++) odd_node = odd_node || is_fixed(hv[h]) || hd[h] != 0;
        bool plain = !(MIX && NEAR) && __ballot(odd_node) == 0ull;
        float ptr = EIK_INFF;  // pending trigger key
        if (act) {
            if (fbits & kFullBit) {
                ptr = 0.f;
            } else {
                bool reached = false;
#pragma unroll
                for (int k = 0; k < 2 * ND; k++) reached = reached | (fabs(s_t[ps + L::poff(k, 1)]) < tinf<T>());
                const bool cand = (fmask & fbits) != 0u || ((fbits & kKeyBit) != 0u && reached);
                if (cand && (T)pend0 <= k0.K) ptr = pend0;
            }
        }

        Metric<A> mcart;
#pragma unroll
        for (int ax = 0; ax < 3; ax++) mcart.dh[ax] = (A)I.h[0];
        metric_coefs(mcart, ND);
        const A sa = (A)sl;
        int it = 0;
        for (; it < S::MAXIT; it++) {
            const int cb = it & 1, pb = cb ^ 1;
            // 1. collect the triggers of the previous sweep; a node within the band is updated
            if (act && it > 0) {
                // the smallest change key not above the node's own key (a tree of minima: the keys
                // are never NaN, and +inf stands for none)
                const float kx = key_up(s_k[n]);
                float m[4 * ND];
#pragma unroll
                for (int k = 0; k < 2 * ND; k++) {
#pragma unroll
                    for (int d = 1; d <= 2; d++) {
                        const float c = img.chg[pb][ps + L::poff(k, d)];
                        m[2 * k + d - 1] = c <= kx ? c : EIK_INFF;
                    }
                }
#pragma unroll
                for (int w = 1; w < 4 * ND; w *= 2)
#pragma unroll
                    for (int i = 0; i + w < 4 * ND; i += 2 * w) m[i] = fminf(m[i], m[i + w]);
                ptr = fminf(ptr, m[0]);
            }
            const bool need = act & (ptr < EIK_INFF) & (ptr <= lim);
            ptr = need ? EIK_INFF : ptr;
            __syncthreads();  // every lane has read the previous sweep's changes
            img.chg[pb][ps] = EIK_INFF;
            if (__ballot(need) == 0ull) break;
            // 2. update
            bool newkey = false;  // this lane wrote a cascade key
            if (need && plain) {
                T v1[2 * ND], v2[2 * ND];
#pragma unroll
                for (int k = 0; k < 2 * ND; k++) {
                    v1[k] = s_t[ps + L::poff(k, 1)];
                    v2[k] = s_t[ps + L::poff(k, 2)];
                }
                auto nokeys = [](int, bool) { return Key<T>{T(0), T(0), 0}; };  // never called: no cascade key
                int code[3];
                Key<T> nk;
                T nv;
                if constexpr (SPH) {
                    Metric<A> mt;
                    node_spacing<A>(I, g, mt.dh);
                    metric_coefs(mt, ND);
                    nv = replay<ND, T, MIX>(v1, v2, nokeys, 0u, 0u, 0u, 0u, 0u, ok1, ok2, sa, mt, code, &nk);
                } else {
                    nv = replay<ND, T, MIX>(v1, v2, nokeys, 0u, 0u, 0u, 0u, 0u, ok1, ok2, sa, mcart, code, &nk);
                }
                tot_upd++;
                const T cur = s_t[ps];
                const T kold = s_k[n];
                // an ordinary node's key is (value, -inf, 0)
                const bool vchg = nv == nv && nv != cur && !(nv >= tinf<T>() && cur >= tinf<T>());
                const bool kchg = nv < tinf<T>() && (nk.K != kold || nk.S != -tinf<T>() || nk.D != 0);
                if (vchg || kchg) {
                    tot_vchg += vchg ? 1u : 0u;
                    tot_kchg += vchg ? 0u : 1u;
                    const bool quiet = tiny_change<T>(cur, Key<T>{kold, -tinf<T>(), 0}, nv, nk, settle);
                    s_t[ps] = nv;
                    s_k[n] = nk.K;
                    s_s[n] = nk.S;
                    s_d[ps] = (uint8_t)nk.D;
                    newkey = nk.D != 0;
                    if (!quiet) img.chg[cb][ps] = key_down(kold < nk.K ? kold : nk.K);
                }
            } else if (need) {
                unsigned f1 = 0u, f2 = 0u, d1 = 0u, d2 = 0u, dcap = 0u;
                T v1[2 * ND], v2[2 * ND];
                // Slots outside the grid are staged as +inf with depth 0, and every use of a
                // neighbour below is masked by ok1 / ok2, so the reads need no bounds select.
#pragma unroll
                for (int k = 0; k < 2 * ND; k++) {
                    const int o1 = ps + L::poff(k, 1), o2 = ps + L::poff(k, 2);
                    const T a = s_t[o1];
                    const T b = s_t[o2];
                    const int da = (int)s_d[o1], db = (int)s_d[o2];
                    f1 |= (unsigned)is_fixed(a) << k;
                    f2 |= (unsigned)is_fixed(b) << k;
                    d1 |= (unsigned)(da != 0) << k;
                    d2 |= (unsigned)(db != 0) << k;
                    dcap |= (unsigned)(da >= kMaxCascade) << k;
                    v1[k] = fabs(a);
                    v2[k] = fabs(b);
                }
                // pop keys straight from LDS (only the cascade path reads them)
                auto keys = [&](int k, bool two) {
                    const int o = L::nbr(n, k, two ? 2 : 1);
                    return Key<T>{s_k[o], s_s[o], (int)s_d[ps + L::poff(k, two ? 2 : 1)]};
                };
                int code[3];
                Key<T> nk;
                T nv = tinf<T>();
                bool near64 = false;  // mixed precision near the source: the FP64 stencil
                if constexpr (MIX && NEAR) {
                    near64 = in_fp64_region(I, g);
                    if (near64) {
                        Metric<double> m64;
                        node_spacing<double>(I, g, m64.dh);
                        metric_coefs(m64, ND);
                        nv = replay<ND, T, false>(v1, v2, keys, d1, d2, dcap, f1, f2, ok1, ok2, (double)sl, m64, code,
                                                  &nk);
                    }
                }
                if (near64) {
                } else if constexpr (SPH) {
                    Metric<A> mt;
                    node_spacing<A>(I, g, mt.dh);
                    metric_coefs(mt, ND);
                    nv = replay<ND, T, MIX>(v1, v2, keys, d1, d2, dcap, f1, f2, ok1, ok2, sa, mt, code, &nk);
                } else {
                    nv = replay<ND, T, MIX>(v1, v2, keys, d1, d2, dcap, f1, f2, ok1, ok2, sa, mcart, code, &nk);
                }
                tot_upd++;
                const T cur = s_t[ps];
                const T kold = s_k[n];
                const T sold = s_s[n];
                const int dold = (int)s_d[ps];
                const bool vchg = nv == nv && nv != cur && !(nv >= tinf<T>() && cur >= tinf<T>());
                const bool kchg = nv < tinf<T>() && (nk.K != kold || nk.S != sold || nk.D != dold);
                if (vchg || kchg) {
                    tot_vchg += vchg ? 1u : 0u;
                    tot_kchg += vchg ? 0u : 1u;
                    const bool quiet = tiny_change<T>(cur, Key<T>{kold, sold, dold}, nv, nk, settle);
                    s_t[ps] = nv;
                    s_k[n] = nk.K;
                    s_s[n] = nk.S;
                    s_d[ps] = (uint8_t)nk.D;
                    if (!quiet) img.chg[cb][ps] = key_down(kold < nk.K ? kold : nk.K);
                }
            }
            plain = plain && __ballot(newkey) == 0ull;
            __syncthreads();
        }
        const bool converged = it < S::MAXIT;
        // the next run's entries go to LDS now, before this tile's stores and atomics are issued
        // (waiting for them later would then wait for those too); s_run is not read again until
        // the next tile is chosen, after the barrier below
        unsigned nbase = 0u, nlen = 0u;
        if (last) {
            nbase = (unsigned)__builtin_amdgcn_readfirstlane((int)claim);
            nlen = nbase < nlist ? min(want, nlist - nbase) : 0u;
            if ((unsigned)n < nlen) s_run[n] = list[nbase + n];
        }

        // write back a changed node; its smaller key (old or new) wakes the neighbour tiles
        unsigned pvl = 0x7f800000u;
        if (act) {
            const T v = s_t[ps];
            const Key<T> kn{s_k[n], s_s[n], (int)s_d[ps]};
            if (!(v == v0 && keq(kn, k0))) {
                const bool quiet = tiny_change<T>(v0, k0, v, kn, settle);
                // every field of the node is written (K = value, S = -inf for an ordinary node, and
                // the key pools start as an unreached node's key): a reader racing this write sees each
                // field from the node's old or new state, never a leftover key, so the wake-up rule
                // below (the smaller of the old and new keys) covers whatever it read
                tt[gl] = v;
                dd[gl] = (uint8_t)kn.D;
                kk[gl] = kn.K;
                ss[gl] = kn.S;
                if (!quiet) pvl = __float_as_uint(key_down(k0.K < kn.K ? k0.K : kn.K));
            }
        }
        // smallest changed key per face ([f]: nodes within two of face f) and over the tile ([6]),
        // and the smallest deferred trigger ([7]); +inf: none
        unsigned fmin[8];
#pragma unroll
        for (int f = 0; f < 6; f++)
            fmin[f] = f < 2 * ND ? wave_min_u32(((fmask >> f) & 1u) ? pvl : 0x7f800000u) : 0x7f800000u;
        fmin[6] = wave_min_u32(pvl);
        fmin[7] = wave_min_u32(ptr < EIK_INFF ? __float_as_uint(ptr) : 0x7f800000u);
        if (n == 0) {
            unsigned* fl = flags + I.tile_base;
            unsigned* pd = pend + I.tile_base;
            const unsigned t0 = (unsigned)I.t[0], t01 = (unsigned)I.t[0] * (unsigned)I.t[1];
            unsigned amin = 0x7f800000u;  // smallest pending value this tile hands out
            auto act_tile = [&](unsigned t, unsigned bit, unsigned v) {
                atomicOr(&fl[t], bit);
                atomicMin(&pd[t], v);
                amin = min(amin, v);
            };
            // An x neighbour that is the adjacent entry of this run and still to run takes the
            // activation there, in place of a flag for the next round: its entry merges the face
            // bit and the pending value.  The block wrote this tile, so that relaxation reads the
            // new values.  (Not for the run's last tile: s_run already holds the next run.
            // Relaxing an entry that has already run once more, in the same run, made the runs
            // longer than the rounds they saved: 12 % slower.)
            auto run_merge = [&](int k, unsigned t, unsigned bit, unsigned v) {
                if (last || k < 0 || k >= (int)rlen || ((rdone >> k) & 1u)) return false;
                const uint4 q = s_run[k];
                if (q.x != t || q.y != e.y) return false;
                s_run[k] = make_uint4(q.x, q.y, min(q.z, v), q.w | bit);
                return true;
            };
            // the x- neighbour sees this tile across its x+ face, and so on
            if (fmin[0] != 0x7f800000u && tix > 0 && !run_merge(bk - 1, tl - 1, 1u << 1, fmin[0]))
                act_tile(tl - 1, 1u << 1, fmin[0]);
            if (fmin[1] != 0x7f800000u && tix < (unsigned)I.t[0] - 1 && !run_merge(bk + 1, tl + 1, 1u << 0, fmin[1]))
                act_tile(tl + 1, 1u << 0, fmin[1]);
            if (fmin[2] != 0x7f800000u && tiy > 0) act_tile(tl - t0, 1u << 3, fmin[2]);
            if (fmin[3] != 0x7f800000u && tiy < (unsigned)I.t[1] - 1) act_tile(tl + t0, 1u << 2, fmin[3]);
            if (ND == 3) {
                if (fmin[4] != 0x7f800000u && tiz > 0) act_tile(tl - t01, 1u << 5, fmin[4]);
                if (fmin[5] != 0x7f800000u && tiz < (unsigned)I.t[2] - 1) act_tile(tl + t01, 1u << 4, fmin[5]);
            }
            if (!converged)
                act_tile(tl, kFullBit, fmin[6] == 0x7f800000u ? 0u : fmin[6]);
            else if (fmin[7] != 0x7f800000u)
                act_tile(tl, kKeyBit, fmin[7]);
            // the next round's lowest pending value of the instance (with the tiles the band
            // selection left flagged, which k_band_select adds)
            if (amin != 0x7f800000u) atomicMin(&wnext[e.y], amin);
        }
        tot_sweeps += (unsigned)it;
        tot_act++;
        __syncthreads();  // LDS is reused by the next tile
        if (last) {
            rbase = nbase;
            rlen = nlen;
            rdone = 0u;
        }
    }
    // per-launch statistics: one atomic per counter and block
    unsigned long long upd = tot_upd, chg = (unsigned long long)tot_vchg + tot_kchg, kchg = tot_kchg;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        upd += __shfl_xor(upd, o);
        chg += __shfl_xor(chg, o);
        kchg += __shfl_xor(kchg, o);
    }
    if (n == 0 && tot_act) {
        atomicAdd(&count[kCntSweeps], tot_sweeps);
        atomicAdd(&counters[0], upd);
        atomicAdd(&counters[1], chg);
        atomicAdd(&counters[2], (unsigned long long)tot_act);
        atomicAdd(&counters[3], kchg);
    }
}

// Band scheduling over the active tiles, one tile per thread.  A block of BS consecutive tiles
// almost always lies in one instance: its minimum and flagged count are then reduced in LDS and
// published with one atomic each (atomics on a few hot addresses from every wave serialise);
// blocks that straddle instances reduce per wave.
// Block size: while the other stream's round kernel holds the CUs (a split call), a scan block
// starts only where all its waves fit beside it, and the FP64 and mixed round kernels leave about
// one wave's registers per SIMD free: 256 threads (one wave per SIMD) then beat 1024 in 3D
// (201x201x101 x 16 sources: FP64 +2.4 %, mixed +4.4 %, FP64 gradient +2 %; 120x150x150 sphere x 32:
// FP64 +6 %); FP32 keeps 1024 (256 is 1.3 % slower there), and so does 2D (600x400 x 256: 256 is
// 5 % slower); 128 is slower everywhere.
template <int ND, typename T, bool MIX>
constexpr int scan_threads() {
    return (ND == 2 || (sizeof(T) == 4 && !MIX)) ? 1024 : 256;
}
// Instances whose first tile ids a scan block stages in LDS (one coalesced read instead of a
// binary search through global memory, one dependent load per step); beyond it the search reads
// the instance table itself.
constexpr int kScanLdsInst = 1024;

__device__ __forceinline__ int find_tile_base(const int64_t* __restrict__ tb, int ninst, int64_t gid) {
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tb[mid] <= gid)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// The instance a tile belongs to: from the LDS copy of the first tile ids when it holds them all.
struct InstFinder {
    const InstDev* inst;
    const int64_t* tb;  // LDS, or null
    int ninst;
    __device__ int operator()(int64_t gid) const {
        return tb ? find_tile_base(tb, ninst, gid) : find_inst(inst, ninst, gid);
    }
};

__device__ __forceinline__ InstFinder stage_tile_bases(const InstDev* __restrict__ inst, int ninst, int64_t* s_tb) {
    const bool lds = ninst <= kScanLdsInst;
    if (lds)
        for (int k = threadIdx.x; k < ninst; k += blockDim.x) s_tb[k] = inst[k].tile_base;
    __syncthreads();
    return InstFinder{inst, lds ? s_tb : nullptr, ninst};
}

// The lowest pending value per instance (into wmin, unless null) and the number of flagged tiles.
template <int BS>
__global__ void __launch_bounds__(BS) k_band_min(const unsigned* __restrict__ flags,
                                                         const unsigned* __restrict__ pend, int64_t ntiles,
                                                         const InstDev* __restrict__ inst, int ninst,
                                                         unsigned* __restrict__ wmin, unsigned* __restrict__ nflagged) {
    __shared__ unsigned s_min[BS / 64], s_cnt[BS / 64];
    __shared__ int64_t s_tb[kScanLdsInst];
    const int64_t b0 = (int64_t)blockIdx.x * BS;
    const int64_t t = b0 + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the tile's words are read before the instance table is searched (no dependence)
    const unsigned fl = t < ntiles ? flags[t] : 0u;
    const unsigned pd = t < ntiles ? pend[t] : 0x7f800000u;
    const InstFinder find = stage_tile_bases(inst, ninst, s_tb);
    const int ilo = find(b0);
    const int ihi = find((b0 + BS < ntiles ? b0 + BS : ntiles) - 1);
    const bool on = fl != 0u;
    unsigned long long m = __ballot(on);
    const unsigned v = on ? pd : 0x7f800000u;
    const int i = on ? (ilo == ihi ? ilo : find(t)) : -1;
    const unsigned wm = wmin && ilo == ihi ? wave_min_u32(v) : 0x7f800000u;
    if (lane == 0) {
        s_cnt[wave] = (unsigned)__popcll(m);
        s_min[wave] = wm;
    }
    if (wmin && ilo != ihi) {
        while (m) {
            const int leader = __ffsll((long long)m) - 1;
            const int li = __shfl(i, leader);
            unsigned r = (i == li) ? v : 0x7f800000u;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) r = min(r, (unsigned)__shfl_xor((int)r, o));
            if (lane == leader) atomicMin(&wmin[li], r);
            m &= ~__ballot(i == li);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned c = 0u, mn = 0x7f800000u;
        for (int w = 0; w < BS / 64; w++) {
            c += s_cnt[w];
            mn = min(mn, s_min[w]);
        }
        if (c) atomicAdd(nflagged, c);
        if (mn != 0x7f800000u) atomicMin(&wmin[ilo], mn);
    }
}

// List the active tiles whose pending value is within the instance's band of its lowest (wmin),
// in tile order, as (tile within its instance, instance, pending value, face bits), and clear
// them; the others wait for a later round, and their lowest pending value per instance goes to
// the next round's minimum (wnext; the round kernel adds the tiles it activates).  One list
// reservation per block.
template <int BS>
__global__ void __launch_bounds__(BS) k_band_select(unsigned* __restrict__ flags, unsigned* __restrict__ pend,
                                                            int64_t ntiles, const InstDev* __restrict__ inst,
                                                            int ninst, const unsigned* __restrict__ wmin,
                                                            unsigned* __restrict__ wnext, uint4* __restrict__ list,
                                                            unsigned* __restrict__ slot) {
    __shared__ unsigned s_wc[BS / 64], s_left[BS / 64];
    __shared__ unsigned s_base;
    __shared__ int64_t s_tb[kScanLdsInst];
    const int64_t b0 = (int64_t)blockIdx.x * BS;
    const int64_t t = b0 + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the tile's words are read before the instance table is searched (no dependence)
    const unsigned f = t < ntiles ? flags[t] : 0u;
    const unsigned pd = t < ntiles ? pend[t] : 0x7f800000u;
    const InstFinder find = stage_tile_bases(inst, ninst, s_tb);
    const int ilo = find(b0);
    const int ihi = find((b0 + BS < ntiles ? b0 + BS : ntiles) - 1);
    const float lim0 = __uint_as_float(wmin[ilo]) + (float)inst[ilo].band;
    bool sel = false;
    unsigned pv = 0x7f800000u;
    int i = ilo;
    if (f) {
        float lim = lim0;
        if (ilo != ihi) {
            i = find(t);
            lim = __uint_as_float(wmin[i]) + (float)inst[i].band;
        }
        pv = pd;
        sel = __uint_as_float(pv) <= lim;
    }
    const bool left = f && !sel;
    const unsigned long long m = __ballot(sel);
    const unsigned lm = ilo == ihi ? wave_min_u32(left ? pv : 0x7f800000u) : 0x7f800000u;
    if (lane == 0) {
        s_wc[wave] = (unsigned)__popcll(m);
        s_left[wave] = lm;
    }
    if (ilo != ihi) {
        unsigned long long q = __ballot(left);
        while (q) {
            const int leader = __ffsll((long long)q) - 1;
            const int li = __shfl(i, leader);
            unsigned r = (left && i == li) ? pv : 0x7f800000u;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) r = min(r, (unsigned)__shfl_xor((int)r, o));
            if (lane == leader) atomicMin(&wnext[li], r);
            q &= ~__ballot(left && i == li);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned tot = 0u, mn = 0x7f800000u;
        for (int w = 0; w < BS / 64; w++) {
            const unsigned c = s_wc[w];
            s_wc[w] = tot;
            tot += c;
            mn = min(mn, s_left[w]);
        }
        s_base = tot ? atomicAdd(&slot[kSlotList], tot) : 0u;
        if (mn != 0x7f800000u) atomicMin(&wnext[ilo], mn);
    }
    __syncthreads();
    if (sel) {
        const unsigned pos = s_base + s_wc[wave] + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
        list[pos] = make_uint4((unsigned)(t - inst[i].tile_base), (unsigned)i, pv, f);
        flags[t] = 0u;
        pend[t] = 0x7f800000u;
    }
}

__global__ void k_fill_u32(unsigned* __restrict__ p, int64_t n, unsigned v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_slowness(const double* __restrict__ vel, double* __restrict__ s64, float* __restrict__ s32,
                           int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double s = 1.0 / vel[i];
        s64[i] = s;
        s32[i] = (float)s;
    }
}

// Source-box corners: tt = dist(corner, src)/vel[corner] (sourceboxloctt!, eikmainfwd.jl:1195-1322).
// One thread per (instance, corner).  Works on the coarse grid (no refinement) and on the fine
// boxes (vel = the nearest-neighbour fine velocity, read through the box's map).
template <typename T>
__global__ void k_corners(const CornerDev* __restrict__ cs, int ncs, const InstDev* __restrict__ inst,
                          T* __restrict__ ttpool, const FineDev* __restrict__ fine, unsigned* __restrict__ flags,
                          unsigned* __restrict__ pend, int nd, int kind) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncs) return;
    const CornerDev C = cs[q];
    const InstDev I = inst[C.inst];
    const int g[3] = {C.ijk[0], C.ijk[1], C.ijk[2]};
    const int64_t l = g[0] + (int64_t)I.n[0] * (g[1] + (int64_t)I.n[1] * g[2]);
    double v;
    if (fine) {
        const FineDev F = fine[C.inst];
        int cg[3] = {0, 0, 0};
        for (int d = 0; d < nd; d++) cg[d] = F.i1c[d] + nn_coarse(g[d], F.df);
        v = C.vel[cg[0] + (int64_t)F.nc[0] * (cg[1] + (int64_t)F.nc[1] * cg[2])];
    } else {
        v = C.vel[l];
    }
    const double dist = C.dist;
    ttpool[I.tt_off + l] = -(T)(dist / v);
    seed_tiles_fwd(I, nd, g, flags, pend, dist / v);
}

// nearest-neighbour fine slowness of every refined box (createfinegrid, eikmainfwd.jl:1035-1056)
__global__ void k_fine_setup(const FineDev* __restrict__ fine, int nfine, const InstDev* __restrict__ inst,
                             double* __restrict__ slowf, double* __restrict__ ttf, int nd) {
    const int f = blockIdx.y;
    if (f >= nfine) return;
    const FineDev F = fine[f];
    const InstDev I = inst[f];
    const int64_t n = (int64_t)I.n[0] * I.n[1] * I.n[2];
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
        int g[3];
        g[0] = (int)(l % I.n[0]);
        g[1] = (int)((l / I.n[0]) % I.n[1]);
        g[2] = (int)(l / ((int64_t)I.n[0] * I.n[1]));
        int cg[3] = {0, 0, 0};
        for (int d = 0; d < nd; d++) cg[d] = F.i1c[d] + nn_coarse(g[d], F.df);
        const double v = F.vel[cg[0] + (int64_t)F.nc[0] * (cg[1] + (int64_t)F.nc[1] * cg[2])];
        slowf[I.s_off + l] = 1.0 / v;
        ttf[I.tt_off + l] = EIK_INF;
    }
}

// Transfer of the refined solution to the coarse grid (eikmainfwd.jl:689-754).  The heap FMM on
// the fine box stops at the first accepted node on a box face that is not a model border, once
// at least one coarse-coincident node has been transferred; the coarse init set is every
// coarse-coincident fine node accepted up to and including that node.  With the converged fine
// solution and its pop keys: k_c = min key over coarse-coincident nodes, k_stop = min key over
// face nodes with key >= k_c, transfer = coarse-coincident nodes with key <= k_stop.  One block
// per source; the transferred points are listed in coarse-window order (xlist, cap entries per
// source) and k_stop is stored in the fine instance for the adjoint.
template <typename T>
__global__ void __launch_bounds__(256) k_transfer(const FineDev* __restrict__ fine, InstDev* __restrict__ finst,
                                                  TTPools fp, const InstDev* __restrict__ cinst, T* __restrict__ ttc,
                                                  unsigned* __restrict__ cflags, unsigned* __restrict__ cpend,
                                                  int* __restrict__ ntransfer, SrcPt* __restrict__ xlist, int cap,
                                                  int nd) {
    const int f = blockIdx.x;
    const FineDev F = fine[f];
    const InstDev I = finst[f];
    const InstDev C = cinst[F.cinst];
    const int64_t n = (int64_t)I.n[0] * I.n[1] * I.n[2];
    const double* ttf = (const double*)fp.tt + I.tt_off;
    const double* kf = (const double*)fp.k + I.tt_off;
    const double* sf = (const double*)fp.s + I.tt_off;
    const uint8_t* df = fp.d + I.tt_off;
    __shared__ Key<double> red[256];
    auto coord = [&](int64_t l, int* g) {
        g[0] = (int)(l % I.n[0]);
        g[1] = (int)((l / I.n[0]) % I.n[1]);
        g[2] = (int)(l / ((int64_t)I.n[0] * I.n[1]));
    };
    auto coincident = [&](const int* g) {
        for (int d = 0; d < nd; d++)
            if (g[d] % F.df != 0) return false;
        return true;
    };
    // pop key of a fine node; init nodes pop before every other node
    auto keyof = [&](int64_t l) {
        const double raw = ttf[l];
        if (signbit(raw)) return Key<double>{-EIK_INF, -EIK_INF, 0};
        if (df[l]) return Key<double>{kf[l], sf[l], (int)df[l]};
        return Key<double>{raw, -EIK_INF, 0};
    };
    auto blockmin = [&](Key<double> v) {
        red[threadIdx.x] = v;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s && klt(red[threadIdx.x + s], red[threadIdx.x])) red[threadIdx.x] = red[threadIdx.x + s];
            __syncthreads();
        }
        const Key<double> r = red[0];
        __syncthreads();
        return r;
    };
    const Key<double> kinf{EIK_INF, EIK_INF, 0};
    Key<double> m = kinf;
    for (int64_t l = threadIdx.x; l < n; l += 256) {
        int g[3];
        coord(l, g);
        if (coincident(g)) {
            const Key<double> k = keyof(l);
            if (klt(k, m)) m = k;
        }
    }
    const Key<double> kc = blockmin(m);
    m = kinf;
    for (int64_t l = threadIdx.x; l < n; l += 256) {
        int g[3];
        coord(l, g);
        if (signbit(ttf[l]) || !(ttf[l] < EIK_INF)) continue;  // init nodes are never popped
        bool face = false;
        for (int d = 0; d < nd; d++) {
            if (g[d] == 0 && !F.outmin[d]) face = true;
            if (g[d] == I.n[d] - 1 && !F.outmax[d]) face = true;
        }
        const Key<double> k = keyof(l);
        if (face && !klt(k, kc) && klt(k, m)) m = k;
    }
    const Key<double> ks = blockmin(m);
    if (threadIdx.x != 0) return;
    finst[f].tstopK = ks.K;
    finst[f].tstopS = ks.S;
    finst[f].tstopD = ks.D;
    int nw[3] = {1, 1, 1};
    for (int d = 0; d < nd; d++) nw[d] = (I.n[d] - 1) / F.df + 1;
    int cnt = 0;
    for (int wz = 0; wz < nw[2]; wz++)
        for (int wy = 0; wy < nw[1]; wy++)
            for (int wx = 0; wx < nw[0]; wx++) {
                const int64_t lf = (int64_t)wx * F.df +
                                   (int64_t)I.n[0] * ((int64_t)wy * F.df + (int64_t)I.n[1] * ((int64_t)wz * F.df));
                if (!(fabs(ttf[lf]) < EIK_INF) || klt(ks, keyof(lf))) continue;
                const double t = fabs(ttf[lf]);
                const int cg[3] = {F.i1c[0] + wx, F.i1c[1] + wy, F.i1c[2] + wz};
                const int64_t lc = cg[0] + (int64_t)C.n[0] * (cg[1] + (int64_t)C.n[1] * cg[2]);
                ttc[C.tt_off + lc] = -(T)t;
                seed_tiles_fwd(C, nd, cg, cflags, cpend, t);
                if (cnt < cap) xlist[(int64_t)f * cap + cnt] = SrcPt{F.cinst, (int32_t)lf, lc};
                cnt++;
            }
    ntransfer[f] = cnt;
}

// receiver picks: bilinear_interp (eikforward2D.jl:17-87) / trilinear_interp (eikforward3D.jl:17-140)
template <typename T>
__global__ void k_receivers(const double* __restrict__ rec, const int* __restrict__ rsrc, int64_t nrec,
                            const InstDev* __restrict__ inst, const T* __restrict__ ttpool, int nd, int kind,
                            double* __restrict__ picks) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    const InstDev I = inst[rsrc[r]];
    const T* tt = ttpool + I.tt_off;
    picks[r] = interp_receiver<T>(tt, I, nd, kind, rec + r * nd, nullptr, nullptr);
}

// final |tt| of a pool (removes the init-node sign), converted to FP64
template <typename T>
__global__ void k_export(const T* __restrict__ src, double* __restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = fabs((double)src[i]);
}

// ================================================================= host-side solver driver

static unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }




// Longest run of consecutive list entries a block claims at once (EIK_CLAIM overrides).
static unsigned claim_chunk() {
    static const unsigned v = [] {
        const char* e = getenv("EIK_CLAIM");
        const int c = e ? atoi(e) : 8;
        return (unsigned)(c < 1 ? 1 : (c > 32 ? 32 : c));
    }();
    return v;
}

// Relative size below which an FP32 change wakes no other node (EIK_F32_SETTLE overrides).
static double f32_settle() {
    static const double v = [] {
        const char* e = getenv("EIK_F32_SETTLE");
        return e ? atof(e) : 0.0;
    }();
    return v;
}

// Relative size below which a mixed-precision change wakes no other node (EIK_MIX_SETTLE overrides;
// default 0: every change propagates and the mixed solve replays the mixed heap FMM exactly).
static double mix_settle() {
    static const double v = [] {
        const char* e = getenv("EIK_MIX_SETTLE");
        return e ? atof(e) : 0.0;
    }();
    return v;
}

// Rounds are enqueued in batches of kRoundBatch without host synchronisation, two launches each:
// k_band_select lists the tiles within the band of their instance's lowest pending value, and the
// round kernel runs them on a fixed grid of resident blocks that take tiles from the list.  The
// lowest pending value per instance for the next round is gathered on the way (the tiles the
// selection leaves flagged, and the tiles the round activates), in three buffers used round
// robin: round r reads buffer r % 3, fills (r + 1) % 3 and clears (r + 2) % 3, which round r - 1
// was the last to use; the list length and cursor rotate the same way.  After a batch the host
// reads how many tiles are still flagged; none means converged.
template <int ND, typename T, bool SPH, bool MIX, bool NEAR>
static int fim_run_t(eik_ctx* ctx, const InstDev* d_inst, int ninst, int64_t ntiles, const TTPools& P,
                     const T* slowpool, bool exact, FimStats* fs) {
    DevBuf& list = ctx->buf("fim_tile_list");
    DevBuf& wm = ctx->buf("fim_wmin");
    if (list.ensure((size_t)ntiles * sizeof(uint4)) || wm.ensure((size_t)3 * ninst * sizeof(unsigned)))
        return ctx->fail(EIK_ENOMEM, "out of device memory (tile list)");
    unsigned* flags = (unsigned*)ctx->buf("fim_tflags").p;
    unsigned* pend = (unsigned*)ctx->buf("fim_pend").p;
    unsigned* d_cnt = (unsigned*)ctx->d_small;
    unsigned* h_cnt = (unsigned*)ctx->h_small;
    unsigned long long* d_c64 = (unsigned long long*)(d_cnt + kCnt64);
    hipStream_t st = ctx->stream;
    // resident blocks of this instantiation on the device (EIK_GRID overrides); the statics are
    // initialised once, thread-safely, by the first caller.  More concurrent tiles are not
    // faster: neighbouring tiles relaxed at the same time see each other's old values, and the
    // runs each block claims get shorter.  Measured optimum per CU: 8 blocks in 3D (201x201x101
    // and 120x150x150 grids: 6 blocks 5 % slower on the first, 2 % faster on the second), 4-6 in
    // 2D (600x400 x 256 sources: 8 blocks 15 % slower in FP64; FP32 occupancy allows 12, 35 %
    // slower than 6).  A call split over two streams runs two launches at once: on 3D Cartesian
    // grids 7 blocks per CU per launch then win in FP64 and mixed precision and 6 in FP32
    // (201x201x101 x 16: +2.5 % and +4.6 % against 8); on the sphere 8 stay best (120x150x150 x 32:
    // 6 blocks 7 % slower).
    static const std::array<int, 3> dev = [ctx] {
        int ncu = 0, occ = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) ncu = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_fim_round<ND, T, SPH, MIX, NEAR>, kTileThreads, 0) != hipSuccess)
            occ = 1;
        const char* e = getenv("EIK_GRID");
        return std::array<int, 3>{std::max(1, ncu), std::max(1, occ), e ? std::max(1, atoi(e)) : 0};
    }();
    const int per_cu = ND == 2 ? 5 : (SPH || !ctx->in_split) ? 8 : (sizeof(T) == 4 && !MIX) ? 6 : 7;
    const int grid = dev[2] > 0 ? dev[2] : dev[0] * std::min(per_cu, dev[1]);
    if (getenv("EIK_VERBOSE"))
        fprintf(stderr, "libeik: k_fim_round<%d,%s,%d,%d> %d CUs x %d blocks (occupancy %d) -> grid %d\n", ND,
                sizeof(T) == 8 ? "double" : "float", (int)SPH, (int)MIX, dev[0], per_cu, dev[1], grid);
    const int g = (int)std::min<int64_t>(grid, ntiles);
    constexpr int kSB = scan_threads<ND, T, MIX>();
    FimStats stt{};
    CK(hipMemsetAsync(d_cnt, 0, kSmallBytes, st));
    // Settling starts after kSettleRound rounds (converging solves of the benchmark, test and
    // serving workloads need < 500): changes below kSettle relative no longer propagate, which ends the
    // rare rounding-level cycles at ~1e-13 (FP64) / ~1e-6 (FP32) of the heap's values.
    const int64_t maxbatches = 50000 / kRoundBatch;
    const char* es = getenv("EIK_SETTLE_ROUND");  // test hook: settle from an earlier round
    const int64_t settle_batch = (es ? std::max(0, atoi(es)) : kSettleRound) / kRoundBatch;
    const T settle_tol = sizeof(T) == 8 ? T(1e-12) : T(4e-6);
    // FP64 with `fast`: changes below settle_tol stop propagating from the first round.  FP32
    // (mixed precision): changes below f32_settle() relative (a few ulps) never propagate; they
    // are rounding ripples, and at heap near-ties within FP32 resolution they can otherwise
    // circulate for thousands of rounds before the settling stage ends them.
    const T base_settle = MIX ? T(mix_settle()) : (sizeof(T) == 8 ? (!exact ? settle_tol : T(0)) : T(f32_settle()));
    // Verification (exact solves only): when no tile is left flagged, one more round relaxes every
    // tile of every instance from scratch (every node re-evaluated).  A node can end a round below its fixed
    // point after reading a neighbour that was being written in the same round; its too-low key
    // then filters out the very changes that would correct it.  A verification round that changes
    // nothing proves the state a fixed point of the replay (nothing is written, so no read races a
    // write); otherwise the iteration goes on and verifies again.
    bool verifying = false;
    unsigned rounds_at_verify = 0u;
    unsigned* W = (unsigned*)wm.p;      // 3 x ninst lowest pending values, by round mod 3
    unsigned* SL = d_cnt + kCntSlots;   // 3 round slots (list length, cursor), by round mod 3
    // round 0's lowest pending values from the seeded tiles; the other two buffers start empty
    k_fill_u32<<<nblocks(3 * (int64_t)ninst, 256), 256, 0, st>>>(W, 3 * (int64_t)ninst, 0x7f800000u);
    k_band_min<kSB><<<nblocks(ntiles, kSB), kSB, 0, st>>>(flags, pend, ntiles, d_inst, ninst, W,
                                                                      d_cnt + kCntFlagged);
    CK(hipGetLastError());
    int64_t round = 0;
    for (int64_t batch = 0;; batch++) {
        if (batch >= maxbatches) return ctx->fail(EIK_ENOCONV, "traveltime solve did not converge");
        // the tolerance grows tenfold every 500 rounds while the cycle persists
        T settle = base_settle;
        if (batch >= settle_batch) {
            settle = settle_tol;
            for (int64_t b = settle_batch + 500 / kRoundBatch; b <= batch; b += 500 / kRoundBatch) settle *= T(10);
            stt.settled = std::max(stt.settled, (int)(1 + (batch - settle_batch) / (500 / kRoundBatch)));
        }
        for (int r = 0; r < kRoundBatch; r++, round++) {
            const int c = (int)(round % 3), nx = (int)((round + 1) % 3), rs = (int)((round + 2) % 3);
            k_band_select<kSB><<<nblocks(ntiles, kSB), kSB, 0, st>>>(
                flags, pend, ntiles, d_inst, ninst, W + (size_t)c * ninst, W + (size_t)nx * ninst, (uint4*)list.p,
                SL + c * kSlotWords);
            CK(hipEventRecord(ctx->kev[2 * r], st));
            k_fim_round<ND, T, SPH, MIX, NEAR><<<g, kTileThreads, 0, st>>>(
                (const uint4*)list.p, d_cnt, SL + c * kSlotWords, SL + rs * kSlotWords, d_inst, ninst,
                W + (size_t)c * ninst, W + (size_t)nx * ninst, W + (size_t)rs * ninst, (T*)P.tt, (T*)P.k, (T*)P.s, P.d,
                slowpool, flags, pend, d_c64, settle, claim_chunk());
            CK(hipGetLastError());
            CK(hipEventRecord(ctx->kev[2 * r + 1], st));
        }
        // the tiles still flagged (the next round's lowest pending values are already gathered)
        CK(hipMemsetAsync(d_cnt + kCntFlagged, 0, sizeof(unsigned), st));
        k_band_min<kSB><<<nblocks(ntiles, kSB), kSB, 0, st>>>(flags, pend, ntiles, d_inst, ninst, nullptr,
                                                                          d_cnt + kCntFlagged);
        CK(hipMemcpyAsync(h_cnt, d_cnt, kSmallBytes, hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        for (int r = 0; r < kRoundBatch; r++) {
            float ms = 0.f;
            CK(hipEventElapsedTime(&ms, ctx->kev[2 * r], ctx->kev[2 * r + 1]));
            stt.kernel_ms += ms;
            if (ctx->tbase) {
                float a = 0.f, b = 0.f;
                CK(hipEventElapsedTime(&a, ctx->tbase, ctx->kev[2 * r]));
                CK(hipEventElapsedTime(&b, ctx->tbase, ctx->kev[2 * r + 1]));
                ctx->kspans.push_back(KSpan{a, b, ctx->span_fine});
            }
        }
        stt.launches += kRoundBatch;
        if (h_cnt[kCntFlagged] == 0u) {
            // a solve that lets small changes settle (fast FP64, the settling stages) is not meant
            // to reach the exact fixed point: nothing to verify
            if (settle > T(0)) break;
            if (verifying && h_cnt[kCntRounds] <= rounds_at_verify + 1u) break;  // verified
            k_fill_u32<<<nblocks(ntiles, 256), 256, 0, st>>>(flags, ntiles, kFullBit);
            k_fill_u32<<<nblocks(ntiles, 256), 256, 0, st>>>(pend, ntiles, 0u);
            // every tile now pends at 0: the next round's lowest pending value
            k_fill_u32<<<nblocks(ninst, 256), 256, 0, st>>>(W + (size_t)(round % 3) * ninst, ninst, 0u);
            CK(hipGetLastError());
            verifying = true;
            rounds_at_verify = h_cnt[kCntRounds];
            stt.verifications++;
        }
    }
    const unsigned long long* h_c64 = (const unsigned long long*)(h_cnt + kCnt64);
    stt.rounds = (int64_t)h_cnt[kCntRounds];
    stt.sweeps = (double)h_cnt[kCntSweeps];
    stt.updates = (double)h_c64[0];
    stt.changed = (double)h_c64[1];
    stt.activations = (double)h_c64[2];
    stt.keychanged = (double)h_c64[3];
    if (fs) *fs = stt;
    return EIK_OK;
}

template <int ND, typename T, bool MIX>
static int fim_run_k(eik_ctx* ctx, bool sph, bool near, const InstDev* d_inst, int ninst, int64_t ntiles,
                     const TTPools& P, const void* slowpool, bool exact, FimStats* fs) {
    if constexpr (MIX) {
        if (near)
            return sph ? fim_run_t<ND, T, true, true, true>(ctx, d_inst, ninst, ntiles, P, (const T*)slowpool, exact, fs)
                       : fim_run_t<ND, T, false, true, true>(ctx, d_inst, ninst, ntiles, P, (const T*)slowpool, exact, fs);
    }
    return sph ? fim_run_t<ND, T, true, MIX, false>(ctx, d_inst, ninst, ntiles, P, (const T*)slowpool, exact, fs)
               : fim_run_t<ND, T, false, MIX, false>(ctx, d_inst, ninst, ntiles, P, (const T*)slowpool, exact, fs);
}

template <int ND>
static int fim_run_nd(eik_ctx* ctx, int prec, bool sph, bool near, const InstDev* d_inst, int ninst, int64_t ntiles,
                      const TTPools& P, const void* slowpool, bool exact, FimStats* fs) {
    if (prec == EIK_PREC_FP32)
        return fim_run_k<ND, float, false>(ctx, sph, false, d_inst, ninst, ntiles, P, slowpool, exact, fs);
    if (prec == EIK_PREC_MIXED)
        return fim_run_k<ND, double, true>(ctx, sph, near, d_inst, ninst, ntiles, P, slowpool, exact, fs);
    return fim_run_k<ND, double, false>(ctx, sph, false, d_inst, ninst, ntiles, P, slowpool, exact, fs);
}

int fim_run(eik_ctx* ctx, int nd, int prec, int kind, const InstDev* d_inst, int ninst, int64_t ntiles,
            const TTPools& pools, const void* slowpool, bool exact, bool near, FimStats* fs) {
    const bool sph = kind == EIK_GRID2D_SPHERE || kind == EIK_GRID3D_SPHERE;
    return nd == 2 ? fim_run_nd<2>(ctx, prec, sph, near, d_inst, ninst, ntiles, pools, slowpool, exact, fs)
                   : fim_run_nd<3>(ctx, prec, sph, near, d_inst, ninst, ntiles, pools, slowpool, exact, fs);
}

// ----------------------------------------------------------------- forward pass of one call
// Band width of the tile scheduler, in node crossing times (spacing x mean slowness).  A round
// runs only the active tiles whose pending traveltime lies within the band of the lowest pending
// one, so tiles far ahead of the front are not relaxed against values that will still change.
// A wider band means fewer rounds and more updates per node.  Measured optima, FP64: 6 in 3D
// (201x201x101 x 16 sources: 6-8 within 0.3 % of each other and 4 % faster than 4, 10 slower; the
// 120x150x150 spherical grid 1 % slower at 6 than at 4-5), 16 in 2D (600x400 x 256 sources: 16-20
// 2.5 % faster than 12, which is 17 % faster than 4; 8x8 tiles are twice as wide).  Mixed and
// FP32: below.  EIK_BAND overrides the multiplier (0 or less: no band, every active tile runs).
static double band_mult(int nd, int prec, bool sph) {
    static const double env = [] {
        const char* e = getenv("EIK_BAND");
        return e ? atof(e) : NAN;
    }();
    if (env == env) return env;
    if (nd == 2) return 16.0;  // mixed on 600x400 x 256 sources: 3 % faster than 12, like FP64
    if (prec == EIK_PREC_FP64) return 6.0;
    // mixed and FP32 on 201x201x101 x 16 sources: 8 crossings 2-3 % faster than 6; FP32 gains
    // another 2 % at 10, where mixed loses 1.5 %; on the 120x150x150 spherical grid 8 is 2 % slower
    // than 6
    if (sph) return 6.0;
    return prec == EIK_PREC_FP32 ? 10.0 : 8.0;
}

// EIK_NODEBAND: the same for the nodes inside a tile (default 0: a listed tile relaxes all of
// its nodes).
static double nodeband_mult() {
    static const double m = [] {
        const char* e = getenv("EIK_NODEBAND");
        return e ? atof(e) : 0.0;
    }();
    return m;
}

// EIK_ADJ_BAND: the same for the adjoint rounds, which run from late to early traveltimes
// (0 or less: every flagged tile runs).  4 and 8 measure the same on 201x201x101 and 600x400;
// 4 is 2 % faster on the 120x150x150 spherical grid.
static double adjband_mult() {
    static const double m = [] {
        const char* e = getenv("EIK_ADJ_BAND");
        return e ? atof(e) : 4.0;
    }();
    return m;
}

// A node crossing is the mean node spacing over the axes times the mean slowness.  On spherical
// grids the angular spacings are arcs at the middle radius (and middle colatitude for phi): on a
// 120x150x150 grid with dr = 2 km and 0.1 degree steps at r ~ 6250 km they are ~11 km, and a band
// counted in crossings of dr alone was 13 % (FP64) / 24 % (FP32) slower.
static double node_spacing_mean(const InstDev& I, int nd) {
    if (I.kind != EIK_GRID2D_SPHERE && I.kind != EIK_GRID3D_SPHERE) return I.h[0];
    const double rm = I.o[0] + 0.5 * (I.n[0] - 1) * I.h[0];
    double sum = I.h[0] + rm * I.dthrad;
    if (nd == 3) {
        const double thm = (I.o[1] + 0.5 * (I.n[1] - 1) * I.h[1]) * (M_PI / 180.0);
        sum += rm * sin(thm) * I.dphrad;
    }
    return sum / nd;
}

static void set_band(InstDev& I, int nd, int prec, double smean) {
    const bool sph = I.kind == EIK_GRID2D_SPHERE || I.kind == EIK_GRID3D_SPHERE;
    const double m = band_mult(nd, prec, sph), mn = nodeband_mult(), ma = adjband_mult();
    const double cross = node_spacing_mean(I, nd) * smean;
    I.band = m > 0.0 ? m * cross : INFINITY;
    I.nband = mn > 0.0 ? mn * cross : INFINITY;
    I.aband = ma > 0.0 ? ma * cross : INFINITY;
}

// distance from a source-box corner to the source, on the host (sourceboxloctt!,
// eikmainfwd.jl:1195-1322): the same libm as the oracle, so the init values agree bit for bit
static double corner_dist(const InstDev& I, int nd, int kind, const int* ijk, const double* src) {
    double xyz[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < nd; d++) xyz[d] = fma((double)ijk[d], I.h[d], I.o[d]);
    return distance2points(kind, nd, xyz, src);
}

int forward_solve(eik_ctx* ctx, const std::vector<FwdGroup>& G, int prec, const eik_params* P) {
    const eik_model* M0 = G[0].M;
    const int nd = M0->g.ndim;
    const int kind = M0->g.kind;
    for (const FwdGroup& g : G)
        if (g.M->g.ndim != nd || g.M->g.kind != kind) return ctx->fail(EIK_EINVAL, "batched models of different grid kinds");
    // per source: its model and its coordinates
    std::vector<const eik_model*> smod;
    std::vector<double> sxyz;
    for (const FwdGroup& g : G)
        for (int64_t s = 0; s < g.S->nsrc; s++) {
            smod.push_back(g.M);
            for (int d = 0; d < nd; d++) sxyz.push_back(g.S->src[s * nd + d]);
        }
    const int64_t nsrc = (int64_t)smod.size();
    hipStream_t st = ctx->stream;
    Timer tm(ctx);

    // slowness of the models: a single model's own arrays, or the models' arrays side by side
    const bool f32 = prec == EIK_PREC_FP32;  // FP32 traveltime and key pools
    const size_t esz = f32 ? sizeof(float) : sizeof(double);
    const void* slow = f32 ? (const void*)M0->d_slow32 : (const void*)M0->d_slow64;
    std::vector<int64_t> soff(G.size(), 0);
    if (G.size() > 1) {
        int64_t tot = 0;
        for (size_t k = 0; k < G.size(); k++) {
            soff[k] = tot;
            tot += G[k].M->N;
        }
        DevBuf& sb = ctx->buf("slow_batch");
        if (sb.ensure((size_t)tot * esz)) return ctx->fail(EIK_ENOMEM, "out of device memory (batched models)");
        for (size_t k = 0; k < G.size(); k++) {
            const void* src = f32 ? (const void*)G[k].M->d_slow32 : (const void*)G[k].M->d_slow64;
            CK(hipMemcpyAsync((char*)sb.p + soff[k] * esz, src, (size_t)G[k].M->N * esz, hipMemcpyDeviceToDevice, st));
        }
        slow = sb.p;
    }

    // coarse instances: one grid per source, sharing its model's slowness
    std::vector<InstDev> cinst(nsrc);
    int64_t ntiles_c = 0, antiles_c = 0, totN = 0;
    {
        int64_t s = 0;
        for (size_t k = 0; k < G.size(); k++)
            for (int64_t j = 0; j < G[k].S->nsrc; j++, s++) {
                const eik_model* M = G[k].M;
                InstDev& I = cinst[s];
                make_inst(M->g, I);
                I.sinth = M->d_sinth;
                if (kind == EIK_GRID3D_SPHERE && !I.sinth)
                    return ctx->fail(EIK_EINVAL, "3D sphere model without its metric table");
                set_tiles(I, nd);
                I.tt_off = totN;
                I.s_off = soff[k];
                I.tile_base = ntiles_c;
                I.atile_base = antiles_c;
                set_band(I, nd, prec, M->smean);
                if (prec == EIK_PREC_MIXED && !P->refinearoundsrc) {
                    // unrefined mixed solve: the stencil stays FP64 within kNearSourceRadius nodes of
                    // the source box
                    int corn[8][3];
                    const int nc = find_enclosing_box(M->g, sxyz.data() + s * nd, corn);
                    for (int d = 0; d < 3; d++) {
                        int lo = corn[0][d], hi = corn[0][d];
                        for (int c = 1; c < nc; c++) {
                            lo = std::min(lo, corn[c][d]);
                            hi = std::max(hi, corn[c][d]);
                        }
                        I.xlo[d] = d < nd ? lo - kNearSourceRadius : 0;
                        I.xhi[d] = d < nd ? hi + kNearSourceRadius : 0;
                    }
                }
                totN += M->N;
                ntiles_c += (int64_t)I.t[0] * I.t[1] * I.t[2];
                antiles_c += (int64_t)I.at[0] * I.at[1] * I.at[2];
            }
    }
    DevBuf& ttc = ctx->buf("tt_coarse");
    DevBuf& kc = ctx->buf("key_K");
    DevBuf& sc = ctx->buf("key_S");
    DevBuf& dc = ctx->buf("key_D");
    if (ttc.ensure((size_t)totN * esz) || kc.ensure((size_t)totN * esz) || sc.ensure((size_t)totN * esz) ||
        dc.ensure((size_t)totN))
        return ctx->fail(EIK_ENOMEM, "out of device memory (traveltime grids)");
    CK(hipMemsetAsync(dc.p, 0, (size_t)totN, st));
    // key pools start as an unreached node's key (+inf, -inf)
    if (f32) {
        k_fill<float><<<nblocks(totN, 256), 256, 0, st>>>((float*)kc.p, totN, EIK_INFF);
        k_fill<float><<<nblocks(totN, 256), 256, 0, st>>>((float*)sc.p, totN, -EIK_INFF);
    } else {
        k_fill<double><<<nblocks(totN, 256), 256, 0, st>>>((double*)kc.p, totN, EIK_INF);
        k_fill<double><<<nblocks(totN, 256), 256, 0, st>>>((double*)sc.p, totN, -EIK_INF);
    }
    const TTPools cpools{ttc.p, kc.p, sc.p, (uint8_t*)dc.p};
    DevBuf& dci = ctx->buf("inst_coarse");
    if (dci.ensure(cinst.size() * sizeof(InstDev))) return ctx->fail(EIK_ENOMEM, "out of device memory");
    CK(hipMemcpyAsync(dci.p, cinst.data(), cinst.size() * sizeof(InstDev), hipMemcpyHostToDevice, st));

    int64_t maxt = ntiles_c, maxat = antiles_c;
    std::vector<FineDev> fine;
    std::vector<InstDev> finst;
    int64_t ntiles_f = 0, antiles_f = 0, nfine_pts = 0;
    if (P->refinearoundsrc) {
        fine.resize(nsrc);
        finst.resize(nsrc);
        for (int64_t s = 0; s < nsrc; s++) {
            build_fine(smod[s]->g, sxyz.data() + s * nd, P->downscalefactor, P->noderadius, fine[s], finst[s]);
            fine[s].cinst = (int)s;
            fine[s].vel = smod[s]->d_vel;
            finst[s].tt_off = nfine_pts;
            finst[s].s_off = nfine_pts;
            finst[s].tile_base = ntiles_f;
            finst[s].atile_base = antiles_f;
            set_band(finst[s], nd, EIK_PREC_FP64, smod[s]->smean);  // refined boxes solve in FP64
            nfine_pts += (int64_t)finst[s].n[0] * finst[s].n[1] * finst[s].n[2];
            ntiles_f += (int64_t)finst[s].t[0] * finst[s].t[1] * finst[s].t[2];
            antiles_f += (int64_t)finst[s].at[0] * finst[s].at[1] * finst[s].at[2];
        }
        maxt = std::max(maxt, ntiles_f);
        maxat = std::max(maxat, antiles_f);
        if (kind == EIK_GRID3D_SPHERE) {
            // the boxes' own sin(theta) tables (a 3D sphere is refined only by the mixed-precision
            // FP64 box), from the host libm like the model's
            std::vector<double> tab;
            std::vector<size_t> off(nsrc);
            for (int64_t s = 0; s < nsrc; s++) {
                off[s] = tab.size();
                for (int j = 0; j < finst[s].n[1]; j++)
                    tab.push_back(sin(deg2rad(fma((double)j, finst[s].h[1], finst[s].o[1]))));
            }
            DevBuf& ft = ctx->buf("fine_sinth");
            if (ft.ensure(tab.size() * sizeof(double))) return ctx->fail(EIK_ENOMEM, "out of device memory");
            CK(hipMemcpy(ft.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
            for (int64_t s = 0; s < nsrc; s++) finst[s].sinth = (const double*)ft.p + off[s];
        }
    }
    DevBuf& aflags = ctx->buf("fim_flags");  // adjoint rounds, one byte per adjoint tile
    DevBuf& apend = ctx->buf("adj_pend");    // and its pending traveltime
    if (aflags.ensure((size_t)maxat) || apend.ensure((size_t)maxat * sizeof(unsigned)))
        return ctx->fail(EIK_ENOMEM, "out of device memory (tile flags)");
    CK(hipMemsetAsync(aflags.p, 0, (size_t)maxat, st));
    CK(hipMemsetAsync(apend.p, 0, (size_t)maxat * sizeof(unsigned), st));
    DevBuf& flags = ctx->buf("fim_tflags");
    if (flags.ensure((size_t)maxt * sizeof(unsigned))) return ctx->fail(EIK_ENOMEM, "out of device memory (tile flags)");
    CK(hipMemsetAsync(flags.p, 0, (size_t)maxt * sizeof(unsigned), st));
    DevBuf& pend = ctx->buf("fim_pend");
    if (pend.ensure((size_t)maxt * sizeof(unsigned))) return ctx->fail(EIK_ENOMEM, "out of device memory (tile flags)");
    k_fill_u32<<<nblocks(maxt, 256), 256, 0, st>>>((unsigned*)pend.p, maxt, 0x7f800000u);

    // coarse grids start at +inf
    if (prec != EIK_PREC_FP32)
        k_fill<double><<<nblocks(totN, 256), 256, 0, st>>>((double*)ttc.p, totN, EIK_INF);
    else
        k_fill<float><<<nblocks(totN, 256), 256, 0, st>>>((float*)ttc.p, totN, EIK_INFF);

    std::vector<CornerDev> corners;
    if (P->refinearoundsrc) {
        DevBuf& dfi = ctx->buf("inst_fine");
        DevBuf& dfd = ctx->buf("fine_desc");
        DevBuf& ttf = ctx->buf("tt_fine");
        DevBuf& slf = ctx->buf("slow_fine");
        DevBuf& kf = ctx->buf("keyf_K");
        DevBuf& sf = ctx->buf("keyf_S");
        DevBuf& df = ctx->buf("keyf_D");
        if (dfi.ensure(finst.size() * sizeof(InstDev)) || dfd.ensure(fine.size() * sizeof(FineDev)) ||
            ttf.ensure((size_t)nfine_pts * sizeof(double)) || slf.ensure((size_t)nfine_pts * sizeof(double)) ||
            kf.ensure((size_t)nfine_pts * sizeof(double)) || sf.ensure((size_t)nfine_pts * sizeof(double)) ||
            df.ensure((size_t)nfine_pts))
            return ctx->fail(EIK_ENOMEM, "out of device memory (refined grids)");
        CK(hipMemsetAsync(df.p, 0, (size_t)nfine_pts, st));
        k_fill<double><<<nblocks(nfine_pts, 256), 256, 0, st>>>((double*)kf.p, nfine_pts, EIK_INF);
        k_fill<double><<<nblocks(nfine_pts, 256), 256, 0, st>>>((double*)sf.p, nfine_pts, -EIK_INF);
        const TTPools fpools{ttf.p, kf.p, sf.p, (uint8_t*)df.p};
        CK(hipMemcpyAsync(dfi.p, finst.data(), finst.size() * sizeof(InstDev), hipMemcpyHostToDevice, st));
        CK(hipMemcpyAsync(dfd.p, fine.data(), fine.size() * sizeof(FineDev), hipMemcpyHostToDevice, st));
        dim3 gs(64, (unsigned)nsrc);
        k_fine_setup<<<gs, 256, 0, st>>>((const FineDev*)dfd.p, (int)nsrc, (const InstDev*)dfi.p, (double*)slf.p,
                                         (double*)ttf.p, nd);
        for (int64_t s = 0; s < nsrc; s++) {
            for (int c = 0; c < fine[s].ncorn; c++) {
                CornerDev C{};
                C.inst = (int)s;
                for (int d = 0; d < 3; d++) C.ijk[d] = fine[s].corn[c][d];
                for (int d = 0; d < nd; d++) C.src[d] = sxyz[s * nd + d];
                C.dist = corner_dist(finst[s], nd, kind, C.ijk, C.src);
                C.vel = smod[s]->d_vel;
                corners.push_back(C);
            }
        }
        DevBuf& dcs = ctx->buf("corners");
        if (dcs.ensure(corners.size() * sizeof(CornerDev))) return ctx->fail(EIK_ENOMEM, "out of device memory");
        CK(hipMemcpyAsync(dcs.p, corners.data(), corners.size() * sizeof(CornerDev), hipMemcpyHostToDevice, st));
        k_corners<double><<<nblocks((int64_t)corners.size(), 64), 64, 0, st>>>(
            (const CornerDev*)dcs.p, (int)corners.size(), (const InstDev*)dfi.p, (double*)ttf.p,
            (const FineDev*)dfd.p, (unsigned*)flags.p, (unsigned*)pend.p, nd, kind);
        CK(hipGetLastError());
        tm.mark();
        FimStats fst{};
        ctx->span_fine = true;
        int rc = fim_run(ctx, nd, EIK_PREC_FP64, kind, (const InstDev*)dfi.p, (int)nsrc, ntiles_f, fpools, slf.p,
                         P->fast == 0, false, &fst);
        ctx->span_fine = false;
        if (rc) return rc;
        ctx->stats[1] = tm.mark();
        ctx->stats[14] = fst.kernel_ms;
        ctx->stats[15] = (double)fst.launches;
        ctx->stats[16] = fst.activations;
        DevBuf& ntr = ctx->buf("ntransfer");
        DevBuf& xl = ctx->buf("xlist");
        int cap = 1;
        for (int64_t s = 0; s < nsrc; s++) {
            int w = 1;
            for (int d = 0; d < nd; d++) w *= (finst[s].n[d] - 1) / fine[s].df + 1;
            cap = std::max(cap, w);
        }
        ctx->xcap = cap;
        if (ntr.ensure((size_t)nsrc * sizeof(int)) || xl.ensure((size_t)nsrc * cap * sizeof(SrcPt)))
            return ctx->fail(EIK_ENOMEM, "out of device memory");
        if (prec != EIK_PREC_FP32)
            k_transfer<double><<<(unsigned)nsrc, 256, 0, st>>>((const FineDev*)dfd.p, (InstDev*)dfi.p, fpools,
                                                              (const InstDev*)dci.p, (double*)ttc.p, (unsigned*)flags.p,
                                                              (unsigned*)pend.p, (int*)ntr.p, (SrcPt*)xl.p, cap, nd);
        else
            k_transfer<float><<<(unsigned)nsrc, 256, 0, st>>>((const FineDev*)dfd.p, (InstDev*)dfi.p, fpools,
                                                             (const InstDev*)dci.p, (float*)ttc.p, (unsigned*)flags.p,
                                                             (unsigned*)pend.p, (int*)ntr.p, (SrcPt*)xl.p, cap, nd);
        CK(hipGetLastError());
    } else {
        for (int64_t s = 0; s < nsrc; s++) {
            int corn[8][3];
            const int nc = find_enclosing_box(smod[s]->g, sxyz.data() + s * nd, corn);
            for (int c = 0; c < nc; c++) {
                CornerDev C{};
                C.inst = (int)s;
                for (int d = 0; d < 3; d++) C.ijk[d] = corn[c][d];
                for (int d = 0; d < nd; d++) C.src[d] = sxyz[s * nd + d];
                C.dist = corner_dist(cinst[s], nd, kind, C.ijk, C.src);
                C.vel = smod[s]->d_vel;
                corners.push_back(C);
            }
        }
        DevBuf& dcs = ctx->buf("corners");
        if (dcs.ensure(corners.size() * sizeof(CornerDev))) return ctx->fail(EIK_ENOMEM, "out of device memory");
        CK(hipMemcpyAsync(dcs.p, corners.data(), corners.size() * sizeof(CornerDev), hipMemcpyHostToDevice, st));
        if (prec != EIK_PREC_FP32)
            k_corners<double><<<nblocks((int64_t)corners.size(), 64), 64, 0, st>>>(
                (const CornerDev*)dcs.p, (int)corners.size(), (const InstDev*)dci.p, (double*)ttc.p, nullptr,
                (unsigned*)flags.p, (unsigned*)pend.p, nd, kind);
        else
            k_corners<float><<<nblocks((int64_t)corners.size(), 64), 64, 0, st>>>(
                (const CornerDev*)dcs.p, (int)corners.size(), (const InstDev*)dci.p, (float*)ttc.p, nullptr,
                (unsigned*)flags.p, (unsigned*)pend.p, nd, kind);
        CK(hipGetLastError());
        tm.mark();
    }
    FimStats cst{};
    // unrefined mixed solves give their instances an FP64 region around the source (above)
    const bool near = prec == EIK_PREC_MIXED && !P->refinearoundsrc;
    int rc = fim_run(ctx, nd, prec, kind, (const InstDev*)dci.p, (int)nsrc, ntiles_c, cpools, slow, P->fast == 0,
                     near, &cst);
    if (rc) return rc;
    ctx->stats[2] = tm.mark();
    ctx->stats[4] = cst.activations;
    ctx->stats[5] = (double)cst.rounds;
    ctx->stats[6] = cst.kernel_ms;
    ctx->stats[7] = cst.sweeps;
    ctx->stats[8] = cst.updates;
    ctx->stats[9] = cst.changed;
    ctx->stats[10] = cst.keychanged;
    ctx->stats[11] = cst.settled;
    ctx->stats[12] = cst.verifications;
    ctx->stats[13] = (double)cst.launches;
    return EIK_OK;
}

int forward_picks(eik_ctx* ctx, const std::vector<FwdGroup>& G, int prec, double* picks, double* ttgrid) {
    hipStream_t st = ctx->stream;
    const int nd = G[0].M->g.ndim;
    const int kind = G[0].M->g.kind;
    // receivers of every source, in group and source order
    std::vector<double> rec;
    std::vector<int> rs;
    std::vector<int64_t> N;
    int s0 = 0;
    for (const FwdGroup& g : G) {
        rec.insert(rec.end(), g.S->rec.begin(), g.S->rec.end());
        for (int64_t s = 0; s < g.S->nsrc; s++) {
            for (int64_t r = g.S->rec_off[s]; r < g.S->rec_off[s + 1]; r++) rs.push_back(s0 + (int)s);
            N.push_back(g.M->N);
        }
        s0 += (int)g.S->nsrc;
    }
    const int64_t nrec = (int64_t)rs.size();
    if (nrec > 0) {
        DevBuf& drec = ctx->buf("rec");
        DevBuf& drs = ctx->buf("recsrc");
        DevBuf& dpk = ctx->buf("picks");
        if (drec.ensure((size_t)nrec * nd * sizeof(double)) || drs.ensure((size_t)nrec * sizeof(int)) ||
            dpk.ensure((size_t)nrec * sizeof(double)))
            return ctx->fail(EIK_ENOMEM, "out of device memory (receivers)");
        CK(hipMemcpyAsync(drec.p, rec.data(), (size_t)nrec * nd * sizeof(double), hipMemcpyHostToDevice, st));
        CK(hipMemcpyAsync(drs.p, rs.data(), (size_t)nrec * sizeof(int), hipMemcpyHostToDevice, st));
        const InstDev* dci = (const InstDev*)ctx->buf("inst_coarse").p;
        void* ttc = ctx->buf("tt_coarse").p;
        if (prec != EIK_PREC_FP32)
            k_receivers<double><<<nblocks(nrec, 128), 128, 0, st>>>((const double*)drec.p, (const int*)drs.p, nrec, dci,
                                                                    (const double*)ttc, nd, kind, (double*)dpk.p);
        else
            k_receivers<float><<<nblocks(nrec, 128), 128, 0, st>>>((const double*)drec.p, (const int*)drs.p, nrec, dci,
                                                                   (const float*)ttc, nd, kind, (double*)dpk.p);
        CK(hipGetLastError());
        CK(hipMemcpyAsync(picks, dpk.p, (size_t)nrec * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    if (ttgrid) {
        int64_t maxN = 0;
        for (int64_t n : N) maxN = std::max(maxN, n);
        DevBuf& ex = ctx->buf("export");
        if (ex.ensure((size_t)maxN * sizeof(double))) return ctx->fail(EIK_ENOMEM, "out of device memory");
        void* ttc = ctx->buf("tt_coarse").p;
        int64_t off = 0;
        for (size_t s = 0; s < N.size(); s++) {
            if (prec != EIK_PREC_FP32)
                k_export<double><<<nblocks(N[s], 256), 256, 0, st>>>((const double*)ttc + off, (double*)ex.p, N[s]);
            else
                k_export<float><<<nblocks(N[s], 256), 256, 0, st>>>((const float*)ttc + off, (double*)ex.p, N[s]);
            CK(hipMemcpyAsync(ttgrid + off, ex.p, (size_t)N[s] * sizeof(double), hipMemcpyDeviceToHost, st));
            off += N[s];
        }
    }
    CK(hipStreamSynchronize(st));
    return EIK_OK;
}

int model_slowness(eik_model* M) {
    k_slowness<<<nblocks(M->N, 256), 256, 0, M->ctx->stream>>>(M->d_vel, M->d_slow64, M->d_slow32, M->N);
    return hipGetLastError() == hipSuccess ? EIK_OK : EIK_EDEVICE;
}

}  // namespace eik
