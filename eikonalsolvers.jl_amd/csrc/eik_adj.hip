// eik_adj.hip -- discrete-adjoint gradient of libeik (gfx950).
//
// The reference assembles per source the derivative matrices D_d in FMM order
// (createsparsederivativematrices!, src/EikSolv/eikmainfwd.jl:827-946), forms
// A = sum_d 2 diag(D_d t) D_d (calc2diagDttD, eikmaingrad.jl:751-868), and back-substitutes
// A^T lambda = -P_B^T r (solveadjointeq, :488-535).  Here:
//   * k_adj_prep replays every node's last stencil evaluation at the converged traveltimes
//     (the codes the heap recorded) and stores G_d = 2*(D_d t)_i, so A_ij = G_d(i) * coef;
//   * the triangular solve runs as tiled rounds like the forward, upwind:
//       lambda_j = (b_j - sum_{i: j in stencil(i)} A_ij lambda_i) / A_jj,
//     each node gathering from the <= 4*ND rows whose stencil reaches it (no atomics).  The
//     dependency graph is acyclic, so the iteration terminates with the exact solution;
//   * with refinement the Nh right-hand sides of the fine adjoint (computestuff_finegrid!,
//     :540-632) only ever appear contracted with q = dpsi/du_s of the coarse solve, so one fine
//     solve with rhs -H_B^T q gives the same gradient (calc_duh_dva_finegrid!, :636-741,
//     calc_dus_dsr_finegrid, :1184-1231).
#include <cstring>

#include "eik_internal.h"

namespace eik {

constexpr int kAB = 256;

// ------------------------------------------------------------------ per-node helpers
__device__ __forceinline__ int unpack_code(uint16_t c, int d) { return (int)((c >> (3 * d)) & 7u) - 2; }

// spacing along axis d at node g (arc length for the angular axes of spherical grids)
__device__ __forceinline__ double axis_h(const InstDev& I, const int* g, int d) {
#pragma clang fp contract(off)
    if (I.kind == EIK_GRID2D_CART || I.kind == EIK_GRID3D_CART || d == 0) return I.h[0];
    const double r = fma((double)g[0], I.h[0], I.o[0]);
    if (d == 1) return r * I.dthrad;
    return r * I.sinth[g[1]] * I.dphrad;  // the forward metric's sin(theta_j) table
}

// coefficient of stencil position p (0 = the row's own node) for code c on axis d
// (setcoeffderiv!, eikmaingrad.jl:1037-1090): signcod * firstord/secondord[p]
__device__ __forceinline__ double stencil_coef(int c, int p, double h) {
#pragma clang fp contract(off)
    const double sg = (c > 0) ? -1.0 : 1.0;
    const int ac = c < 0 ? -c : c;
    double cf;
    if (ac == 1)
        cf = (p == 0) ? -1.0 / h : 1.0 / h;
    else
        cf = (p == 0) ? -3.0 / (2.0 * h) : (p == 1 ? 4.0 / (2.0 * h) : -1.0 / (2.0 * h));
    return sg * cf;
}

template <typename T>
__device__ __forceinline__ double tabs(T v) {
    return fabs((double)v);
}

// pop key of node l (init nodes pop first)
template <typename T>
__device__ __forceinline__ Key<T> node_key(const T* tt, const T* kk, const T* ss, const uint8_t* dd, int64_t l) {
    const T raw = tt[l];
    if (signbit(raw)) return Key<T>{-tinf<T>(), -tinf<T>(), 0};
    if (dd[l]) return Key<T>{kk[l], ss[l], (int)dd[l]};
    return Key<T>{raw, -tinf<T>(), 0};
}

// the node popped no later than the refinement stop (always true on coarse grids)
template <typename T>
__device__ __forceinline__ bool before_stop(const InstDev& I, const Key<T>& k) {
    const Key<double> kd{(double)k.K, (double)k.S, k.D};
    return !klt(Key<double>{I.tstopK, I.tstopS, I.tstopD}, kd);
}

// Codes and G_d of every node: replay of the final stencil evaluation (eikmainfwd.jl:794-809)
// and 2*(D_d t)_i accumulated over the stencil in acceptance order (far node first).  NEAR: the
// instances may have an FP64 region (an unrefined mixed solve), whose nodes replay the FP64
// stencil; compiled only where it can occur (the second replay costs registers).
template <int ND, typename T, bool MIX, bool NEAR>
__global__ void __launch_bounds__(kAB) k_adj_prep(const InstDev* __restrict__ inst, TTPools pools,
                                                  const T* __restrict__ slowpool, uint16_t* __restrict__ code_out,
                                                  double* __restrict__ G, int64_t gstride) {
#pragma clang fp contract(off)
    const InstDev I = inst[blockIdx.y];
    const int64_t n = (int64_t)I.n[0] * I.n[1] * I.n[2];
    const T* tt = (const T*)pools.tt + I.tt_off;
    const T* kk = (const T*)pools.k + I.tt_off;
    const T* ss = (const T*)pools.s + I.tt_off;
    const uint8_t* dd = pools.d + I.tt_off;
    const T* slow = slowpool + I.s_off;
    const int64_t nx = I.n[0], nxy = (int64_t)I.n[0] * I.n[1];
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
        const int g[3] = {(int)(l % I.n[0]), (int)((l / I.n[0]) % I.n[1]), (int)(l / nxy)};
        const T raw = tt[l];
        const double t0 = tabs(raw);
        uint16_t pc = 0;
        double Gd[3] = {0.0, 0.0, 0.0};
        if (!signbit(raw) && t0 < EIK_INF && before_stop(I, node_key(tt, kk, ss, dd, l))) {
            T v1[2 * ND], v2[2 * ND];
            Key<T> k1[2 * ND], k2[2 * ND];
            unsigned f1 = 0, f2 = 0, ok1 = 0, ok2 = 0;
#pragma unroll
            for (int ax = 0; ax < ND; ax++) {
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    const int k = 2 * ax + s;
                    int nb[3] = {g[0], g[1], g[2]};
                    nb[ax] += (s == 0) ? 1 : -1;
                    const bool in1 = nb[ax] >= 0 && nb[ax] <= I.n[ax] - 1;
                    bool in2 = in1;
#pragma unroll
                    for (int d = 0; d < ND; d++) in2 = in2 && nb[d] >= 1 && nb[d] <= I.n[d] - 2;
                    T a = tinf<T>(), b = tinf<T>();
                    Key<T> ka{tinf<T>(), -tinf<T>(), 0}, kb{tinf<T>(), -tinf<T>(), 0};
                    if (in1) {
                        const int64_t q1 = nb[0] + nx * nb[1] + nxy * nb[2];
                        a = tt[q1];
                        ka = node_key(tt, kk, ss, dd, q1);
                    }
                    if (in2) {
                        int nb2[3] = {g[0], g[1], g[2]};
                        nb2[ax] += (s == 0) ? 2 : -2;
                        const int64_t q2 = nb2[0] + nx * nb2[1] + nxy * nb2[2];
                        b = tt[q2];
                        kb = node_key(tt, kk, ss, dd, q2);
                    }
                    k1[k] = ka;
                    k2[k] = kb;
                    ok1 |= (unsigned)in1 << k;
                    ok2 |= (unsigned)in2 << k;
                    f1 |= (unsigned)signbit(a) << k;
                    f2 |= (unsigned)signbit(b) << k;
                    v1[k] = fabs(a);
                    v2[k] = fabs(b);
                }
            }
            using A = Arith<T, MIX>;
            int code[3] = {0, 0, 0};
            Key<T> kz;
            bool near64 = false;  // the forward's FP64 near-source stencil
            if constexpr (MIX && NEAR) {
                near64 = in_fp64_region(I, g);
                if (near64) {
                    Metric<double> m64;
                    node_spacing<double>(I, g, m64.dh);
                    metric_coefs(m64, ND);
                    (void)replay<ND, T, false>(v1, v2, k1, k2, f1, f2, ok1, ok2, (double)slow[l], m64, code, &kz);
                }
            }
            if (!near64) {
                Metric<A> mt;
                node_spacing<A>(I, g, mt.dh);
                metric_coefs(mt, ND);
                (void)replay<ND, T, MIX>(v1, v2, k1, k2, f1, f2, ok1, ok2, (A)slow[l], mt, code, &kz);
            }
#pragma unroll
            for (int d = 0; d < ND; d++) {
                const int c = code[d];
                pc |= (uint16_t)((c + 2) << (3 * d));
                if (c == 0) continue;
                const int ac = c < 0 ? -c : c;
                const int sg = (c > 0) ? -1 : 1;
                double tmp1 = 0.0;
                for (int p = ac; p >= 0; p--) {
                    int q[3] = {g[0], g[1], g[2]};
                    q[d] += sg * p;
                    const double cf = stencil_coef(c, p, axis_h(I, q, d));
                    tmp1 += cf * tabs(tt[q[0] + nx * q[1] + nxy * q[2]]);
                }
                Gd[d] = 2.0 * tmp1;
            }
        } else {
#pragma unroll
            for (int d = 0; d < ND; d++) pc |= (uint16_t)(2 << (3 * d));
        }
        code_out[I.tt_off + l] = pc;
#pragma unroll
        for (int d = 0; d < ND; d++) G[d * gstride + I.tt_off + l] = Gd[d];
    }
}

// Coefficient A_ij with which row i (at offset o = i - j along axis d) references node j,
// or 0 when j is not in i's stencil.
__device__ __forceinline__ double contrib(const InstDev& I, const int* gi, int d, int o, uint16_t ci, double Gi) {
    const int c = unpack_code(ci, d);
    if (c == 0) return 0.0;
    if ((c > 0) != (o > 0)) return 0.0;
    const int ac = c < 0 ? -c : c, ao = o < 0 ? -o : o;
    if (ao > ac) return 0.0;
    return Gi * stencil_coef(c, ao, axis_h(I, gi, d));
}

// diagonal A_jj = sum_d G_d(j) * coef(self)
template <int ND>
__device__ __forceinline__ double adj_diag(const InstDev& I, const int* g, uint16_t cj, const double* Gj) {
#pragma clang fp contract(off)
    double a = 0.0;
#pragma unroll
    for (int d = 0; d < ND; d++) {
        const int c = unpack_code(cj, d);
        if (c == 0) continue;
        a += Gj[d] * stencil_coef(c, 0, axis_h(I, g, d));
    }
    return a;
}

template <int ND>
struct AdjTile;
template <>
struct AdjTile<2> {
    static constexpr int TX = 16, TY = 16, TZ = 1, H = 2, LX = 20, LY = 20, LZ = 1, NP = 1, MAXIT = 128;
};
template <>
struct AdjTile<3> {
    static constexpr int TX = 8, TY = 8, TZ = 8, H = 2, LX = 12, LY = 12, LZ = 12, NP = 2, MAXIT = 96;
};
static_assert(AdjTile<2>::TX == adj_tile_edge(2) && AdjTile<3>::TX == adj_tile_edge(3), "adjoint tile edge");

// one adjoint round over the listed tiles; a neighbour woken by a changed row gets that row's
// traveltime (rounded up) as its pending value
// Adjoint round counters (ctx->d_small words, while the adjoint runs): two list-length slots used
// by round parity, then sweeps, tile activations and non-empty rounds.
constexpr int kAdjSlots = 0, kAdjSweeps = 2, kAdjActs = 3, kAdjRounds = 4, kAdjWords = 5;
constexpr int kAdjBatch = 8;  // adjoint rounds enqueued between two host checks

template <int ND, typename T>
__global__ void __launch_bounds__(kAB) k_adj_round(const int64_t* __restrict__ list, const unsigned* __restrict__ slot,
                                                   unsigned* __restrict__ slot_reset, unsigned* __restrict__ wreset,
                                                   const InstDev* __restrict__ inst, int ninst,
                                                   const uint16_t* __restrict__ codep,
                                                   const double* __restrict__ Gp, int64_t gstride,
                                                   const double* __restrict__ bp, double* __restrict__ lamp,
                                                   const T* __restrict__ ttpool, uint8_t* __restrict__ flags,
                                                   unsigned* __restrict__ apend, unsigned* __restrict__ stats) {
#pragma clang fp contract(off)
    using S = AdjTile<ND>;
    constexpr int NL = S::LX * S::LY * S::LZ;
    // G along axis d is read only by rows along d, so s_G[d] holds the core extended by the halo
    // on the two faces across d (L = T + 2H along d, T along the other axes)
    constexpr int NG = S::LX * S::TY * (ND == 3 ? S::TZ : 1);
    static_assert(S::TX == S::TY && (ND == 2 || S::TZ == S::TX), "cubic adjoint tiles");
    __shared__ double s_lam[NL];
    __shared__ double s_G[ND][NG];
    __shared__ uint16_t s_code[NL];
    __shared__ unsigned s_fmax[7];
    // slot of local node (halo coordinates lc) in s_G[d]: the other axes drop their halo
    auto gslot = [](int d, const int lc[3]) {
        int s = 0, m = 1;
    #pragma unroll
        for (int a = 0; a < ND; a++) {
            s += (a == d ? lc[a] : lc[a] - S::H) * m;
            m *= a == d ? S::LX : S::TX;
        }
        return s;
    };  // per face (and whole tile): largest traveltime of a changed row
    const int tid = threadIdx.x;
    const unsigned nlist = *slot;
    if (blockIdx.x == 0) {
        // the next round's list length and per-instance maxima (last used by the previous round)
        for (int i = tid; i < ninst; i += kAB) wreset[i] = 0u;
        if (tid == 0) {
            *slot_reset = 0u;
            if (nlist) {
                atomicAdd(&stats[kAdjRounds], 1u);
                atomicAdd(&stats[kAdjActs], nlist);
            }
        }
    }
    // a fixed grid strides over the listed tiles, one tile per block at a time
    for (unsigned e = blockIdx.x; e < nlist; e += gridDim.x) {
        const int64_t gid = list[e];
        int lo = 0, hi = ninst - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (inst[mid].atile_base <= gid)
                lo = mid;
            else
                hi = mid - 1;
        }
        const InstDev I = inst[lo];
        const int64_t tl = gid - I.atile_base;
        const int tix = (int)(tl % I.at[0]), tiy = (int)((tl / I.at[0]) % I.at[1]);
        const int tiz = ND == 3 ? (int)(tl / ((int64_t)I.at[0] * I.at[1])) : 0;
        const int ox = tix * S::TX, oy = tiy * S::TY, oz = tiz * S::TZ;
        const int64_t nx = I.n[0], nxy = (int64_t)I.n[0] * I.n[1];
        const uint16_t* code = codep + I.tt_off;
        const double* b = bp + I.tt_off;
        double* lam = lamp + I.tt_off;
        if (tid < 7) s_fmax[tid] = 0u;
        for (int c = tid; c < NL; c += kAB) {
            const int lx = c % S::LX, ly = (c / S::LX) % S::LY, lz = ND == 3 ? c / (S::LX * S::LY) : 0;
            int nout = (lx < S::H || lx >= S::H + S::TX) + (ly < S::H || ly >= S::H + S::TY);
            if (ND == 3) nout += (lz < S::H || lz >= S::H + S::TZ);
            double lv = 0.0;
            uint16_t cv = 0x1B6;  // all codes zero ((0+2) in every 3-bit field)
            double gv[3] = {0.0, 0.0, 0.0};
            if (nout <= 1) {
                const int gx = ox + lx - S::H, gy = oy + ly - S::H, gz = ND == 3 ? oz + lz - S::H : 0;
                if (gx >= 0 && gx < I.n[0] && gy >= 0 && gy < I.n[1] && (ND == 2 || (gz >= 0 && gz < I.n[2]))) {
                    const int64_t l = gx + nx * gy + nxy * gz;
                    lv = lam[l];
                    cv = code[l];
    #pragma unroll
                    for (int d = 0; d < ND; d++) gv[d] = Gp[d * gstride + I.tt_off + l];
                }
            }
            s_lam[c] = lv;
            s_code[c] = cv;
            const int lc[3] = {lx, ly, lz};
    #pragma unroll
            for (int d = 0; d < ND; d++) {
                bool in_d = true;
    #pragma unroll
                for (int a = 0; a < ND; a++)
                    if (a != d) in_d = in_d && lc[a] >= S::H && lc[a] < S::H + S::TX;
                if (in_d) s_G[d][gslot(d, lc)] = gv[d];
            }
        }
        __syncthreads();

        constexpr int NS = 4 * ND;  // candidate rows: axis d, offset +-1, +-2
        double w[S::NP][NS];
        int off[NS];
        double diag[S::NP], bj[S::NP], l0[S::NP];
        bool act[S::NP];
        int lidx[S::NP], gxyz[S::NP][3];
        const int strides[3] = {1, S::LX, S::LX * S::LY};
        const int gstrides[3] = {1, S::TX, S::TX * S::TY};  // along axis d within s_G[d]
    #pragma unroll
        for (int k = 0; k < NS; k++) {
            const int d = k / 4, o = (k % 4 < 2) ? ((k % 2) ? -1 : 1) : ((k % 2) ? -2 : 2);
            off[k] = o * strides[d];
        }
    #pragma unroll
        for (int p = 0; p < S::NP; p++) {
            const int lx = tid % S::TX, ly = (tid / S::TX) % S::TY;
            const int lz = ND == 3 ? (tid / (S::TX * S::TY)) + p * (S::TZ / S::NP) : 0;
            const int g[3] = {ox + lx, oy + ly, oz + lz};
            gxyz[p][0] = g[0];
            gxyz[p][1] = g[1];
            gxyz[p][2] = g[2];
            lidx[p] = (lx + S::H) + S::LX * ((ly + S::H) + S::LY * (ND == 3 ? lz + S::H : 0));
            const int lcore[3] = {lx + S::H, ly + S::H, ND == 3 ? lz + S::H : 0};
            const bool inside = g[0] < I.n[0] && g[1] < I.n[1] && (ND == 2 || g[2] < I.n[2]);
            const uint16_t cj = s_code[lidx[p]];
            bool hasrow = false;
    #pragma unroll
            for (int d = 0; d < ND; d++) hasrow = hasrow || unpack_code(cj, d) != 0;
            act[p] = inside && hasrow;
            diag[p] = 1.0;
            bj[p] = 0.0;
    #pragma unroll
            for (int k = 0; k < NS; k++) w[p][k] = 0.0;
            if (act[p]) {
                double Gj[3] = {0, 0, 0};
    #pragma unroll
                for (int d = 0; d < ND; d++) Gj[d] = s_G[d][gslot(d, lcore)];
                diag[p] = adj_diag<ND>(I, g, cj, Gj);
                bj[p] = b[g[0] + nx * g[1] + nxy * g[2]];
    #pragma unroll
                for (int k = 0; k < NS; k++) {
                    const int d = k / 4, o = (k % 4 < 2) ? ((k % 2) ? -1 : 1) : ((k % 2) ? -2 : 2);
                    int gi[3] = {g[0], g[1], g[2]};
                    gi[d] += o;
                    if (gi[d] < 0 || gi[d] > I.n[d] - 1) continue;
                    const int li = lidx[p] + off[k];
                    w[p][k] = contrib(I, gi, d, o, s_code[li], s_G[d][gslot(d, lcore) + o * gstrides[d]]);
                }
            }
            l0[p] = s_lam[lidx[p]];
        }
        __syncthreads();

        int it = 0;
        for (; it < S::MAXIT; it++) {
            int any = 0;
    #pragma unroll
            for (int p = 0; p < S::NP; p++) {
                if (!act[p]) continue;
                double x = bj[p];
    #pragma unroll
                for (int k = 0; k < NS; k++)
                    if (w[p][k] != 0.0) x -= s_lam[lidx[p] + off[k]] * w[p][k];
                x = x / diag[p];
                if (x != s_lam[lidx[p]]) {
                    s_lam[lidx[p]] = x;
                    any = 1;
                }
            }
            if (!__syncthreads_or(any)) break;
        }
    #pragma unroll
        for (int p = 0; p < S::NP; p++) {
            if (!act[p]) continue;
            const double v = s_lam[lidx[p]];
            if (v != l0[p]) {
                const int gx = gxyz[p][0], gy = gxyz[p][1], gz = gxyz[p][2];
                const int64_t l = gx + nx * gy + nxy * gz;
                lam[l] = v;
                const unsigned tv = __float_as_uint(__double2float_ru(fabs((double)ttpool[I.tt_off + l])));
                atomicMax(&s_fmax[6], tv);
                // a changed row only feeds the nodes its own stencil reads (code c along an axis: the
                // |c| nodes on the side opposite to the sign of c), so a neighbour tile is woken only
                // when that stencil crosses into it
                const int lc[3] = {gx - ox, gy - oy, gz - oz};
                const int te[3] = {S::TX, S::TY, S::TZ};
                const uint16_t cr = s_code[lidx[p]];
    #pragma unroll
                for (int d = 0; d < ND; d++) {
                    const int c = unpack_code(cr, d);
                    if (c > 0 && lc[d] < c) atomicMax(&s_fmax[2 * d], tv);
                    if (c < 0 && lc[d] - c >= te[d]) atomicMax(&s_fmax[2 * d + 1], tv);
                }
            }
        }
        __syncthreads();
        if (tid == 0) {
            uint8_t* fl = flags + I.atile_base;
            unsigned* pd = apend + I.atile_base;
            const int64_t t0 = I.at[0], t01 = (int64_t)I.at[0] * I.at[1];
            auto wake = [&](int64_t t, unsigned v) {
                fl[t] = 1;
                atomicMax(&pd[t], v);
            };
            if (s_fmax[0] && tix > 0) wake(tl - 1, s_fmax[0]);
            if (s_fmax[1] && tix < I.at[0] - 1) wake(tl + 1, s_fmax[1]);
            if (s_fmax[2] && tiy > 0) wake(tl - t0, s_fmax[2]);
            if (s_fmax[3] && tiy < I.at[1] - 1) wake(tl + t0, s_fmax[3]);
            if (ND == 3) {
                if (s_fmax[4] && tiz > 0) wake(tl - t01, s_fmax[4]);
                if (s_fmax[5] && tiz < I.at[2] - 1) wake(tl + t01, s_fmax[5]);
            }
            if (it >= S::MAXIT) wake(tl, s_fmax[6]);
            atomicAdd(&stats[kAdjSweeps], (unsigned)it);
        }
        __syncthreads();  // LDS is reused by the next tile
    }
}

// Band scheduling of the adjoint rounds (late traveltimes first).  Pass 1: highest pending value
// of the flagged tiles per instance.
__device__ __forceinline__ int find_ainst(const InstDev* __restrict__ inst, int ninst, int64_t gid) {
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (inst[mid].atile_base <= gid)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void k_adj_band_max(const uint8_t* __restrict__ flags, const unsigned* __restrict__ apend, int64_t ntiles,
                               const InstDev* __restrict__ inst, int ninst, unsigned* __restrict__ wmax) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = t < ntiles && flags[t];
    const int i = on ? find_ainst(inst, ninst, t) : -1;
    unsigned v = on ? apend[t] : 0u;
    // one atomic per wave and instance (the tiles of a wave almost always share one)
    unsigned long long m = __ballot(on);
    const int lane = threadIdx.x & 63;
    while (m) {
        const int leader = __ffsll((long long)m) - 1;
        const int li = __shfl(i, leader);
        unsigned r = (i == li) ? v : 0u;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) r = max(r, (unsigned)__shfl_xor((int)r, o));
        if (lane == leader) atomicMax(&wmax[li], r);
        m &= ~__ballot(i == li);
    }
}

// Pass 2: list the flagged tiles within the band below their instance's highest pending value and
// clear them; one list reservation per wave.
__global__ void k_adj_band_select(uint8_t* __restrict__ flags, unsigned* __restrict__ apend, int64_t ntiles,
                                  const InstDev* __restrict__ inst, int ninst, const unsigned* __restrict__ wmax,
                                  int64_t* __restrict__ list, unsigned* __restrict__ count) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool sel = false;
    if (t < ntiles && flags[t]) {
        const int i = find_ainst(inst, ninst, t);
        const float lim = __uint_as_float(wmax[i]) - (float)inst[i].aband;
        sel = __uint_as_float(apend[t]) >= lim;
    }
    const unsigned long long m = __ballot(sel);
    const int lane = threadIdx.x & 63;
    unsigned base = 0u;
    if (lane == 0 && m) base = atomicAdd(count, (unsigned)__popcll(m));
    base = __shfl(base, 0);
    if (sel) {
        list[base + (unsigned)__popcll(m & ((1ull << lane) - 1ull))] = t;
        flags[t] = 0;
        apend[t] = 0u;
    }
}

// sum over the rows that reference node j of A_ij * lambda_i (dpsi/du_s at source points:
// transpose(sum(twoDttDS)) * lambda, eikmaingrad.jl:341-344)
template <int ND>
__device__ double gather_rows(const InstDev& I, const int* g, const uint16_t* code, const double* Gp, int64_t gstride,
                              const double* lam) {
#pragma clang fp contract(off)
    const int64_t nx = I.n[0], nxy = (int64_t)I.n[0] * I.n[1];
    double acc = 0.0;
    for (int k = 0; k < 4 * ND; k++) {
        const int d = k / 4, o = (k % 4 < 2) ? ((k % 2) ? -1 : 1) : ((k % 2) ? -2 : 2);
        int gi[3] = {g[0], g[1], g[2]};
        gi[d] += o;
        if (gi[d] < 0 || gi[d] > I.n[d] - 1) continue;
        const int64_t li = gi[0] + nx * gi[1] + nxy * gi[2];
        const double a = contrib(I, gi, d, o, code[I.tt_off + li], Gp[d * gstride + I.tt_off + li]);
        if (a != 0.0) acc += a * lam[I.tt_off + li];
    }
    return acc;
}

// receivers: picks, residual factors, rhs b = -P^T (pd / std^2), per-source misfit terms
template <typename T>
__global__ void k_adj_rhs(const double* __restrict__ rec, const int* __restrict__ rsrc, int64_t nrec,
                          const InstDev* __restrict__ inst, const T* __restrict__ ttpool, int nd, int kind,
                          const double* __restrict__ pobs, const double* __restrict__ sobs, double* __restrict__ bpool,
                          double* __restrict__ mterm, uint8_t* __restrict__ flags, unsigned* __restrict__ apend) {
#pragma clang fp contract(off)
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    const InstDev I = inst[rsrc[r]];
    double cf[8];
    int64_t idx[8];
    const double pick = interp_receiver<T>(ttpool + I.tt_off, I, nd, kind, rec + r * nd, cf, idx);
    const double pd = pick - pobs[r];
    const double s2 = sobs[r] * sobs[r];
    mterm[r] = (pd * pd) / s2;
    const double fact = pd / s2;
    const int nc = nd == 2 ? 4 : 8;
    for (int c = 0; c < nc; c++) {
        atomicAdd(&bpool[I.tt_off + idx[c]], -cf[c] * fact);
        const int64_t l = idx[c];
        const int g[3] = {(int)(l % I.n[0]), (int)((l / I.n[0]) % I.n[1]), (int)(l / ((int64_t)I.n[0] * I.n[1]))};
        seed_tiles(I, nd, g, flags, apend, (double)ttpool[I.tt_off + l]);
    }
}

// per-source misfit sum in receiver order, then the total in source order (calcgradsomesrc,
// eikmaingrad.jl:237: totmisf += 0.5*misfit1src)
// One block: each source's receiver terms are summed in receiver order by one thread, and the
// sources' halves are added in source order by thread 0, chunk by chunk: the summation order of
// a serial loop over sources and receivers, so the same result bit for bit.
constexpr int kMisfitThreads = 256;
__global__ void __launch_bounds__(kMisfitThreads) k_misfit(const double* __restrict__ mterm,
                                                            const int64_t* __restrict__ rec_off, int64_t nsrc,
                                                            double* __restrict__ out) {
    __shared__ double s_m[kMisfitThreads];
    double tot = 0.0;
    for (int64_t base = 0; base < nsrc; base += kMisfitThreads) {
        const int64_t s = base + threadIdx.x;
        double m = 0.0;
        if (s < nsrc)
            for (int64_t r = rec_off[s]; r < rec_off[s + 1]; r++) m += mterm[r];
        s_m[threadIdx.x] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int64_t cnt = nsrc - base < kMisfitThreads ? nsrc - base : kMisfitThreads;
            for (int64_t k = 0; k < cnt; k++) tot += 0.5 * s_m[k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = tot;
}

// source points of every instance: dpsiup = sum_i A^S_iq lambda_i, dpsidus = (P_A^T fact)_q
template <int ND>
__global__ void k_srcterms(const SrcPt* __restrict__ pts, int npts, const InstDev* __restrict__ inst,
                           const uint16_t* __restrict__ code, const double* __restrict__ Gp, int64_t gstride,
                           const double* __restrict__ lam, const double* __restrict__ b, double* __restrict__ dpsiup,
                           double* __restrict__ dpsidus) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npts) return;
    const SrcPt P = pts[q];
    const InstDev I = inst[P.inst];
    const int64_t l = P.node;
    const int g[3] = {(int)(l % I.n[0]), (int)((l / I.n[0]) % I.n[1]), (int)(l / ((int64_t)I.n[0] * I.n[1]))};
    dpsiup[q] = gather_rows<ND>(I, g, code, Gp, gstride, lam);
    dpsidus[q] = -b[I.tt_off + l];
}


}  // namespace eik

namespace eik {

// scatter of per-point results into a pool-shaped array (value at the point's node)
__global__ void k_scatter_pts(const SrcPt* __restrict__ pts, int npts, const InstDev* __restrict__ inst,
                              const double* __restrict__ val, double* __restrict__ pool) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < npts) pool[inst[pts[q].inst].tt_off + pts[q].node] = val[q];
}

// T2Vb (eikmaingrad.jl:417-423): 2*lambda/v^3 at every non-source node of every source image
template <typename T>
__global__ void k_grad_img(const InstDev* __restrict__ inst, const T* __restrict__ ttpool,
                           const double* __restrict__ lam, const double* __restrict__ vel, double* __restrict__ img,
                           int64_t N) {
#pragma clang fp contract(off)
    const InstDev I = inst[blockIdx.y];
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < N; l += (int64_t)gridDim.x * blockDim.x) {
        const T raw = ttpool[I.tt_off + l];
        double g = 0.0;
        if (!signbit(raw)) {
            const double v = vel[l];
            g = 2.0 * lam[I.tt_off + l] / (v * v * v);
        }
        img[I.tt_off + l] = g;
    }
}

// dus/dva = -diag(dist/vel^2) at the coarse source-box corners (no refinement, :428-479)
__global__ void k_src_own(const SrcPt* __restrict__ pts, int npts, const InstDev* __restrict__ inst,
                          const double* __restrict__ vel, const double* __restrict__ srcxyz, int nd, int kind,
                          const double* __restrict__ dpsiup, const double* __restrict__ dpsidus,
                          double* __restrict__ t2, double* __restrict__ t1) {
#pragma clang fp contract(off)
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npts) return;
    const SrcPt P = pts[q];
    const InstDev I = inst[P.inst];
    const int64_t l = P.node;
    const int g[3] = {(int)(l % I.n[0]), (int)((l / I.n[0]) % I.n[1]), (int)(l / ((int64_t)I.n[0] * I.n[1]))};
    double xyz[3];
    for (int d = 0; d < nd; d++) xyz[d] = fma((double)g[d], I.h[d], I.o[d]);
    const double dist = distance2points(kind, nd, xyz, srcxyz + P.inst * nd);
    const double v = vel[l];
    const double dd = -dist * (1.0 / (v * v));
    t2[q] = dd * dpsiup[q];
    t1[q] = dd * dpsidus[q];
}

// fine rhs -H_B^T q and H_A^T q for the transferred points (q = dpsiup + dpsidus, coarse)
__global__ void k_fine_rhs(const SrcPt* __restrict__ xp, int nxp, const InstDev* __restrict__ finst,
                           const double* __restrict__ ttf, const double* __restrict__ dpsiup,
                           const double* __restrict__ dpsidus, double* __restrict__ bf, double* __restrict__ haq,
                           uint8_t* __restrict__ flags, unsigned* __restrict__ apend, int nd) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nxp) return;
    const SrcPt P = xp[q];
    const InstDev I = finst[P.inst];
    const int64_t lf = P.aux;
    const double qv = dpsiup[q] + dpsidus[q];
    if (signbit(ttf[I.tt_off + lf])) {
        haq[I.tt_off + lf] = qv;
    } else {
        bf[I.tt_off + lf] = -qv;
        const int g[3] = {(int)(lf % I.n[0]), (int)((lf / I.n[0]) % I.n[1]), (int)(lf / ((int64_t)I.n[0] * I.n[1]))};
        seed_tiles(I, nd, g, flags, apend, ttf[I.tt_off + lf]);
    }
}

// z_p of the fine grid: 2*mu/w^3 (T2Wb) at accepted non-source nodes, dtau/dw * (dgi + H_A^T q)
// (T2Wa + T1Wa) at the fine source-box corners
__global__ void k_fine_z(const InstDev* __restrict__ finst, const FineDev* __restrict__ fine, TTPools fp,
                         const double* __restrict__ ttf, const double* __restrict__ mu, const double* __restrict__ dgi,
                         const double* __restrict__ haq, const double* __restrict__ velc,
                         const double* __restrict__ srcxyz, int nd, int kind, double* __restrict__ z) {
#pragma clang fp contract(off)
    const int f = blockIdx.y;
    const InstDev I = finst[f];
    const FineDev F = fine[f];
    const int64_t n = (int64_t)I.n[0] * I.n[1] * I.n[2];
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (int64_t)gridDim.x * blockDim.x) {
        const double raw = ttf[I.tt_off + l];
        const int g[3] = {(int)(l % I.n[0]), (int)((l / I.n[0]) % I.n[1]), (int)(l / ((int64_t)I.n[0] * I.n[1]))};
        int cg[3] = {0, 0, 0};
        for (int d = 0; d < nd; d++) cg[d] = F.i1c[d] + nn_coarse(g[d], F.df);
        const double w = velc[cg[0] + (int64_t)F.nc[0] * (cg[1] + (int64_t)F.nc[1] * cg[2])];
        double zv = 0.0;
        if (signbit(raw)) {
            double xyz[3];
            for (int d = 0; d < nd; d++) xyz[d] = fma((double)g[d], I.h[d], I.o[d]);
            const double dist = distance2points(kind, nd, xyz, srcxyz + f * nd);
            const double dd = -dist * (1.0 / (w * w));
            zv = dd * (dgi[I.tt_off + l] + haq[I.tt_off + l]);
        } else if (raw < EIK_INF && before_stop(I, node_key((const double*)fp.tt + I.tt_off, (const double*)fp.k + I.tt_off,
                                                             (const double*)fp.s + I.tt_off, fp.d + I.tt_off, l))) {
            zv = 2.0 * mu[I.tt_off + l] / (w * w * w);
        }
        z[I.tt_off + l] = zv;
    }
}

// coarse source point a gathers z over the accepted fine nodes whose nearest coarse node is a
// (dwq_dva keeps only coarse nodes that are transferred points, eikmaingrad.jl:717-733)
__global__ void k_fine_gather(const SrcPt* __restrict__ xp, int nxp, const InstDev* __restrict__ finst,
                              const FineDev* __restrict__ fine, TTPools fp, const double* __restrict__ ttf,
                              const double* __restrict__ z, double* __restrict__ t2, double* __restrict__ t1, int nd) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nxp) return;
    const SrcPt P = xp[q];
    const InstDev I = finst[P.inst];
    const FineDev F = fine[P.inst];
    const int64_t lf = P.aux;
    const int g[3] = {(int)(lf % I.n[0]), (int)((lf / I.n[0]) % I.n[1]), (int)(lf / ((int64_t)I.n[0] * I.n[1]))};
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, w[3] = {0, 0, 0};
    for (int d = 0; d < nd; d++) {
        w[d] = g[d] / F.df;
        lo[d] = max(0, (w[d] - 1) * F.df);
        hi[d] = min(I.n[d] - 1, (w[d] + 1) * F.df);
    }
    double acc = 0.0;
    for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = lo[0]; i <= hi[0]; i++) {
                if (nn_coarse(i, F.df) != w[0] || nn_coarse(j, F.df) != w[1]) continue;
                if (nd == 3 && nn_coarse(k, F.df) != w[2]) continue;
                const int64_t l = i + (int64_t)I.n[0] * (j + (int64_t)I.n[1] * k);
                if (!(fabs(ttf[I.tt_off + l]) < EIK_INF)) continue;
                if (!before_stop(I, node_key((const double*)fp.tt + I.tt_off, (const double*)fp.k + I.tt_off,
                                             (const double*)fp.s + I.tt_off, fp.d + I.tt_off, l)))
                    continue;
                acc += z[I.tt_off + l];
            }
    t2[q] = acc;
    t1[q] = 0.0;
}

// taper factor of a source at node g (smoothgradaroundsrc!, utils.jl:121-338)
__device__ __forceinline__ double taper_factor(const TaperDev& T, const InstDev& I, int nd, int kind, const int* g) {
#pragma clang fp contract(off)
    if (!T.on) return 1.0;
    for (int d = 0; d < nd; d++)
        if (g[d] < T.lo[d] || g[d] > T.hi[d]) return 1.0;
    double r;
    if (kind == EIK_GRID3D_SPHERE) {  // chord distance, as the 3D spherical forward measures it
        const double a[3] = {fma((double)g[0], I.h[0], I.o[0]), fma((double)g[1], I.h[1], I.o[1]),
                             fma((double)g[2], I.h[2], I.o[2])};
        r = distance2points(kind, 3, a, T.src);
    } else if (kind == EIK_GRID2D_SPHERE) {
        const double xs = T.src[0] * cos(deg2rad(T.src[1])), ys = T.src[0] * sin(deg2rad(T.src[1]));
        const double rr = fma((double)g[0], I.h[0], I.o[0]), th = fma((double)g[1], I.h[1], I.o[1]);
        const double xc = rr * cos(deg2rad(th)), yc = rr * sin(deg2rad(th));
        r = sqrt((xc - xs) * (xc - xs) + (ys - yc) * (ys - yc));
    } else if (nd == 2) {
        const double xc = fma((double)g[0], I.h[0], I.o[0]), yc = fma((double)g[1], I.h[1], I.o[1]);
        r = sqrt((xc - T.src[0]) * (xc - T.src[0]) + (T.src[1] - yc) * (T.src[1] - yc));
    } else {
        const double xc = fma((double)g[0], I.h[0], I.o[0]), yc = fma((double)g[1], I.h[1], I.o[1]),
                     zc = fma((double)g[2], I.h[2], I.o[2]);
        r = sqrt((xc - T.src[0]) * (xc - T.src[0]) + (T.src[1] - yc) * (T.src[1] - yc) +
                 (T.src[2] - zc) * (T.src[2] - zc));
    }
    return (r <= T.rmax) ? r / T.rmax : 1.0;
}

// The reference reuses one per-chunk gradient buffer across sources and only overwrites the
// non-source nodes (T2Vb), so a source node starts from the previous source's tapered value
// before T2Va and T1Va are added (calcgradsomesrc, eikmaingrad.jl:182-247).  Sequential over
// sources, parallel over each source's points.  carry (or null): the tapered image of the source
// before the first one here (k_taper multiplies by the same factor, so it holds the same bits).
__global__ void k_grad_quirk(const SrcPt* __restrict__ pts, const int* __restrict__ poff, int nsrc,
                             const InstDev* __restrict__ inst, const TaperDev* __restrict__ tap, int nd, int kind,
                             const double* __restrict__ t2, const double* __restrict__ t1, double* __restrict__ img,
                             const double* __restrict__ carry) {
#pragma clang fp contract(off)
    for (int s = 0; s < nsrc; s++) {
        for (int q = poff[s] + threadIdx.x; q < poff[s + 1]; q += blockDim.x) {
            const InstDev I = inst[s];
            const int64_t l = pts[q].node;
            double prev = s == 0 && carry ? carry[l] : 0.0;
            if (s > 0) {
                const int g[3] = {(int)(l % I.n[0]), (int)((l / I.n[0]) % I.n[1]),
                                  (int)(l / ((int64_t)I.n[0] * I.n[1]))};
                prev = taper_factor(tap[s - 1], I, nd, kind, g) * img[inst[s - 1].tt_off + l];
            }
            img[I.tt_off + l] = (prev + t2[q]) + t1[q];
        }
        __syncthreads();
    }
}

__global__ void k_taper(const InstDev* __restrict__ inst, const TaperDev* __restrict__ tap, int nd, int kind,
                        double* __restrict__ img) {
    const int s = blockIdx.x;
    const TaperDev T = tap[s];
    if (!T.on) return;
    const InstDev I = inst[s];
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < nd; d++) {
        lo[d] = max(T.lo[d], 0);
        hi[d] = min(T.hi[d], I.n[d] - 1);
        if (hi[d] < lo[d]) return;
    }
    const int ex = hi[0] - lo[0] + 1, ey = hi[1] - lo[1] + 1, ez = hi[2] - lo[2] + 1;
    for (int c = threadIdx.x; c < ex * ey * ez; c += blockDim.x) {
        const int g[3] = {lo[0] + c % ex, lo[1] + (c / ex) % ey, lo[2] + c / (ex * ey)};
        const double a = taper_factor(T, I, nd, kind, g);
        if (a != 1.0) img[I.tt_off + g[0] + (int64_t)I.n[0] * (g[1] + (int64_t)I.n[1] * g[2])] *= a;
    }
}

// gradvelall .+= gradvel1, source by source (eikmaingrad.jl:246), over the images of the first
// part's sources, then the second's
__global__ void k_sum_sources(const double* __restrict__ img0, int n0, const double* __restrict__ img1, int n1,
                              int64_t N, double* __restrict__ out) {
#pragma clang fp contract(off)
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= N) return;
    double a = 0.0;
    for (int s = 0; s < n0; s++) a += img0[(int64_t)s * N + l];
    for (int s = 0; s < n1; s++) a += img1[(int64_t)s * N + l];
    out[l] = a;
}

// zero-padded Gaussian smoothing of the summed gradient (smoothgradient, utils.jl:73-96)
__global__ void k_smooth(const double* __restrict__ in, double* __restrict__ out, const double* __restrict__ ker,
                         int l, int nd, int n0, int n1, int n2) {
#pragma clang fp contract(off)
    const int64_t N = (int64_t)n0 * n1 * n2;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    const int g[3] = {(int)(idx % n0), (int)((idx / n0) % n1), (int)(idx / ((int64_t)n0 * n1))};
    const int m = l / 2;
    const int n[3] = {n0, n1, n2};
    int keep[3], nk = 0;
    for (int d = 0; d < nd; d++)
        if (n[d] != 1) keep[nk++] = d;
    int64_t kn = 1;
    for (int d = 0; d < nk; d++) kn *= l;
    double acc = 0.0;
    for (int64_t q = 0; q < kn; q++) {
        int64_t r = q;
        int c[3] = {g[0], g[1], g[2]};
        bool ok = true;
        for (int d = 0; d < nk; d++) {
            const int o = (int)(r % l) - m;
            r /= l;
            c[keep[d]] += o;
            if (c[keep[d]] < 0 || c[keep[d]] > n[keep[d]] - 1) ok = false;
        }
        if (ok) acc += in[c[0] + (int64_t)n0 * (c[1] + (int64_t)n1 * c[2])] * ker[q];
    }
    out[idx] = acc;
}

// partderivttsrcpos (eikmaingrad.jl:1235-1273)
__device__ __forceinline__ void partder(int nd, const double* xyzpt, const double* src, double vel, double hgrid,
                                        double* der) {
#pragma clang fp contract(off)
    double xs[3];
    for (int d = 0; d < nd; d++) {
        double al = fmod(src[d], hgrid);
        if (al < 0) al += hgrid;
        xs[d] = (al <= 2.220446049250313e-16) ? src[d] + 1e-6 * hgrid : src[d];
    }
    double s2 = 0.0;
    for (int d = 0; d < nd; d++) s2 += (xyzpt[d] - xs[d]) * (xyzpt[d] - xs[d]);
    const double denom = vel * sqrt(s2);
    for (int d = 0; d < nd; d++) der[d] = (denom == 0.0) ? 0.0 : -(xyzpt[d] - xs[d]) / denom;
}

// gradient w.r.t. the source position (calcgrads_singlesrc!, :362-399); one thread per source.
// Without refinement over the coarse corners; with refinement over the fine corners, where
// dgi + H_A^T q equals (du_h/dtau_s)^T q + H_A^T q of calc_dus_dsr_finegrid.
__global__ void k_gradsrc(int nsrc, int nd, const SrcPt* __restrict__ pts, const int* __restrict__ poff,
                          const InstDev* __restrict__ inst, const double* __restrict__ vel,
                          const double* __restrict__ srcxyz, const double* __restrict__ a1,
                          const double* __restrict__ a2, const FineDev* __restrict__ fine,
                          const double* __restrict__ velc, const double* __restrict__ haq, double* __restrict__ out) {
#pragma clang fp contract(off)
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsrc) return;
    double A[3] = {0, 0, 0}, B[3] = {0, 0, 0};
    for (int q = poff[s]; q < poff[s + 1]; q++) {
        const SrcPt P = pts[q];
        const InstDev I = inst[P.inst];
        const int64_t l = P.node;
        const int g[3] = {(int)(l % I.n[0]), (int)((l / I.n[0]) % I.n[1]), (int)(l / ((int64_t)I.n[0] * I.n[1]))};
        double xyz[3], v;
        for (int d = 0; d < nd; d++) xyz[d] = fma((double)g[d], I.h[d], I.o[d]);
        if (fine) {
            const FineDev F = fine[P.inst];
            int cg[3] = {0, 0, 0};
            for (int d = 0; d < nd; d++) cg[d] = F.i1c[d] + nn_coarse(g[d], F.df);
            v = velc[cg[0] + (int64_t)F.nc[0] * (cg[1] + (int64_t)F.nc[1] * cg[2])];
        } else {
            v = vel[l];
        }
        double der[3];
        partder(nd, xyz, srcxyz + s * nd, v, I.h[0], der);
        if (fine) {
            const double w = a1[q] + haq[I.tt_off + l];
            for (int d = 0; d < nd; d++) A[d] += der[d] * w;
        } else {
            for (int d = 0; d < nd; d++) {
                A[d] += der[d] * a1[q];
                B[d] += der[d] * a2[q];
            }
        }
    }
    for (int d = 0; d < nd; d++) out[s * nd + d] = A[d] + B[d];
}

// ================================================================= host driver

static unsigned nb(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <int ND, typename T>
static int adj_run_t(eik_ctx* ctx, const InstDev* d_inst, int ninst, int64_t ntiles, const uint16_t* code,
                     const double* G, int64_t gstride, const double* b, double* lam, const T* ttpool, FimStats* fs) {
    DevBuf& list = ctx->buf("fim_list");
    DevBuf& wm = ctx->buf("adj_wmax");
    if (list.ensure((size_t)ntiles * sizeof(int64_t)) || wm.ensure((size_t)2 * ninst * sizeof(unsigned)))
        return ctx->fail(EIK_ENOMEM, "out of device memory (tile list)");
    uint8_t* flags = (uint8_t*)ctx->buf("fim_flags").p;
    unsigned* apend = (unsigned*)ctx->buf("adj_pend").p;
    unsigned* d_cnt = (unsigned*)ctx->d_small;
    unsigned* h_cnt = (unsigned*)ctx->h_small;
    unsigned* W = (unsigned*)wm.p;  // 2 x ninst highest pending values, by round parity
    hipStream_t st = ctx->stream;
    // resident blocks that stride over each round's list (EIK_GRID does not apply here)
    static const int grid = [ctx] {
        int ncu = 0, occ = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) ncu = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_adj_round<ND, T>, kAB, 0) != hipSuccess) occ = 1;
        return std::max(1, ncu * std::max(1, occ));
    }();
    // Rounds are enqueued kAdjBatch at a time with no host synchronisation: the round kernel reads
    // its list length on the device and clears the next round's length and maxima (round parity;
    // both were last used by the previous round, which has ended).  After a batch the host reads
    // the last round's list length; none listed means converged (a round with nothing listed wakes
    // nothing, so every later round is empty too).
    CK(hipMemsetAsync(d_cnt, 0, kAdjWords * sizeof(unsigned), st));
    k_fill_u32<<<nb(2 * (int64_t)ninst, 256), 256, 0, st>>>(W, 2 * (int64_t)ninst, 0u);
    FimStats stt{};
    int64_t round = 0;
    for (;;) {
        for (int r = 0; r < kAdjBatch; r++, round++) {
            const int p = (int)(round & 1);
            k_adj_band_max<<<nb(ntiles, 256), 256, 0, st>>>(flags, apend, ntiles, d_inst, ninst, W + (size_t)p * ninst);
            k_adj_band_select<<<nb(ntiles, 256), 256, 0, st>>>(flags, apend, ntiles, d_inst, ninst,
                                                               (const unsigned*)W + (size_t)p * ninst,
                                                               (int64_t*)list.p, d_cnt + kAdjSlots + p);
            k_adj_round<ND, T><<<grid, kAB, 0, st>>>((const int64_t*)list.p, d_cnt + kAdjSlots + p,
                                                      d_cnt + kAdjSlots + (p ^ 1), W + (size_t)(p ^ 1) * ninst, d_inst,
                                                      ninst, code, G, gstride, b, lam, ttpool, flags, apend, d_cnt);
            CK(hipGetLastError());
        }
        CK(hipMemcpyAsync(h_cnt, d_cnt, kAdjWords * sizeof(unsigned), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        if (h_cnt[kAdjSlots + (int)((round - 1) & 1)] == 0u) break;
        if (round > 50000) return ctx->fail(EIK_ENOCONV, "adjoint solve did not converge");
    }
    stt.rounds = (int64_t)h_cnt[kAdjRounds];
    stt.activations = (double)h_cnt[kAdjActs];
    stt.sweeps = (double)h_cnt[kAdjSweeps];
    if (fs) *fs = stt;
    return EIK_OK;
}

// ttpool: the traveltime pool of the instances (FP64, or FP32 for the coarse grids in mixed
// precision); it orders the rounds
static int adj_run(eik_ctx* ctx, int nd, int prec, const InstDev* d_inst, int ninst, int64_t ntiles,
                   const uint16_t* code, const double* G, int64_t gstride, const double* b, double* lam,
                   const void* ttpool, FimStats* fs) {
    if (prec != EIK_PREC_FP32)
        return nd == 2 ? adj_run_t<2, double>(ctx, d_inst, ninst, ntiles, code, G, gstride, b, lam,
                                              (const double*)ttpool, fs)
                       : adj_run_t<3, double>(ctx, d_inst, ninst, ntiles, code, G, gstride, b, lam,
                                              (const double*)ttpool, fs);
    return nd == 2 ? adj_run_t<2, float>(ctx, d_inst, ninst, ntiles, code, G, gstride, b, lam, (const float*)ttpool, fs)
                   : adj_run_t<3, float>(ctx, d_inst, ninst, ntiles, code, G, gstride, b, lam, (const float*)ttpool, fs);
}

template <typename T, bool MIX, bool NEAR = false>
static void launch_prep(int nd, hipStream_t st, const InstDev* inst, int ninst, const TTPools& pools,
                        const void* slow, uint16_t* code, double* G, int64_t gstride) {
    dim3 grid(256, (unsigned)ninst);
    if (nd == 2)
        k_adj_prep<2, T, MIX, NEAR><<<grid, kAB, 0, st>>>(inst, pools, (const T*)slow, code, G, gstride);
    else
        k_adj_prep<3, T, MIX, NEAR><<<grid, kAB, 0, st>>>(inst, pools, (const T*)slow, code, G, gstride);
}

static void taper_box(const eik_grid& g, const double* src, int R, TaperDev& T) {
    std::memset(&T, 0, sizeof(T));
    T.on = R > 0;
    for (int d = 0; d < g.ndim; d++) T.src[d] = src[d];
    T.rmax = R * g.d[0];
    if (!T.on) return;
    int c[8][3];
    const int nc = find_enclosing_box(g, src, c);
    if (nc == 1) {
        for (int d = 0; d < g.ndim; d++) {
            T.lo[d] = c[0][d] - R;
            T.hi[d] = c[0][d] + R;
        }
    } else {
        T.lo[0] = c[0][0] - R;
        T.hi[0] = c[1][0] + R;
        T.lo[1] = c[0][1] - R;
        T.hi[1] = c[2][1] + R;
        if (g.ndim == 3 && g.kind == EIK_GRID3D_SPHERE) {  // no reference: the box's own k range
            T.lo[2] = c[0][2] - R;
            T.hi[2] = c[nc - 1][2] + R;
        } else if (g.ndim == 3) {  // the reference builds the k range from the i index (utils.jl:257-258)
            T.lo[2] = c[0][0] - R;
            T.hi[2] = c[3][0] + R;
        }
    }
}

int gradient_part(eik_ctx* ctx, eik_model* M, const SourceSet& S, const eik_params* P, int whichgrad,
                  const double* pickobs, const double* stdobs, double* gradsrc, bool defer, GradPart* out) {
    const int nd = M->g.ndim, kind = M->g.kind, prec = P->precision;
    const int64_t N = M->N, nsrc = S.nsrc, nrec = S.nrec();
    const bool refine = P->refinearoundsrc != 0;
    hipStream_t st = ctx->stream;
    int rc = forward_solve(ctx, M, S, prec, P);
    if (rc) return rc;
    Timer tm(ctx);
    const InstDev* dci = (const InstDev*)ctx->buf("inst_coarse").p;
    void* ttc = ctx->buf("tt_coarse").p;
    const void* slow = prec == EIK_PREC_FP32 ? (const void*)M->d_slow32 : (const void*)M->d_slow64;
    const int TX = adj_tile_edge(nd), TZ = nd == 2 ? 1 : TX;
    const int64_t tiles_per = (int64_t)((M->g.n[0] + TX - 1) / TX) * ((M->g.n[1] + TX - 1) / TX) *
                              (nd == 3 ? (M->g.n[2] + TZ - 1) / TZ : 1);
    const int64_t P_c = nsrc * N;

    DevBuf &code_c = ctx->buf("adj_code"), &G_c = ctx->buf("adj_G"), &lam_c = ctx->buf("adj_lam"),
           &b_c = ctx->buf("adj_b");
    if (code_c.ensure((size_t)P_c * sizeof(uint16_t)) || G_c.ensure((size_t)P_c * nd * sizeof(double)) ||
        lam_c.ensure((size_t)P_c * sizeof(double)) || b_c.ensure((size_t)P_c * sizeof(double)))
        return ctx->fail(EIK_ENOMEM, "out of device memory (adjoint pools)");
    CK(hipMemsetAsync(lam_c.p, 0, (size_t)P_c * sizeof(double), st));
    CK(hipMemsetAsync(b_c.p, 0, (size_t)P_c * sizeof(double), st));
    const TTPools cpools{ttc, ctx->buf("key_K").p, ctx->buf("key_S").p, (uint8_t*)ctx->buf("key_D").p};
    if (prec == EIK_PREC_FP64)
        launch_prep<double, false>(nd, st, dci, (int)nsrc, cpools, slow, (uint16_t*)code_c.p, (double*)G_c.p, P_c);
    else if (prec == EIK_PREC_MIXED && !refine)  // unrefined mixed: FP64 regions around the sources
        launch_prep<double, true, true>(nd, st, dci, (int)nsrc, cpools, slow, (uint16_t*)code_c.p, (double*)G_c.p, P_c);
    else if (prec == EIK_PREC_MIXED)
        launch_prep<double, true>(nd, st, dci, (int)nsrc, cpools, slow, (uint16_t*)code_c.p, (double*)G_c.p, P_c);
    else
        launch_prep<float, false>(nd, st, dci, (int)nsrc, cpools, slow, (uint16_t*)code_c.p, (double*)G_c.p, P_c);
    CK(hipGetLastError());

    // receivers: rhs, misfit
    DevBuf &drec = ctx->buf("rec"), &drs = ctx->buf("recsrc"), &dpo = ctx->buf("pobs"), &dso = ctx->buf("sobs"),
           &dmt = ctx->buf("mterm"), &droff = ctx->buf("rec_off");
    if (drec.ensure((size_t)std::max<int64_t>(nrec, 1) * nd * sizeof(double)) ||
        drs.ensure((size_t)std::max<int64_t>(nrec, 1) * sizeof(int)) ||
        dpo.ensure((size_t)std::max<int64_t>(nrec, 1) * sizeof(double)) ||
        dso.ensure((size_t)std::max<int64_t>(nrec, 1) * sizeof(double)) ||
        dmt.ensure((size_t)std::max<int64_t>(nrec, 1) * sizeof(double)) ||
        droff.ensure((size_t)(nsrc + 1) * sizeof(int64_t)))
        return ctx->fail(EIK_ENOMEM, "out of device memory (receivers)");
    std::vector<int> rs((size_t)std::max<int64_t>(nrec, 1));
    for (int64_t s = 0; s < nsrc; s++)
        for (int64_t r = S.rec_off[s]; r < S.rec_off[s + 1]; r++) rs[r] = (int)s;
    if (nrec > 0) {
        CK(hipMemcpyAsync(drec.p, S.rec.data(), (size_t)nrec * nd * sizeof(double), hipMemcpyHostToDevice, st));
        CK(hipMemcpyAsync(drs.p, rs.data(), (size_t)nrec * sizeof(int), hipMemcpyHostToDevice, st));
        CK(hipMemcpyAsync(dpo.p, pickobs, (size_t)nrec * sizeof(double), hipMemcpyHostToDevice, st));
        CK(hipMemcpyAsync(dso.p, stdobs, (size_t)nrec * sizeof(double), hipMemcpyHostToDevice, st));
    }
    CK(hipMemcpyAsync(droff.p, S.rec_off.data(), (size_t)(nsrc + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    unsigned* apend = (unsigned*)ctx->buf("adj_pend").p;
    uint8_t* flags = (uint8_t*)ctx->buf("fim_flags").p;
    if (nrec > 0) {
        if (prec != EIK_PREC_FP32)
            k_adj_rhs<double><<<nb(nrec, 128), 128, 0, st>>>((const double*)drec.p, (const int*)drs.p, nrec, dci,
                                                             (const double*)ttc, nd, kind, (const double*)dpo.p,
                                                             (const double*)dso.p, (double*)b_c.p, (double*)dmt.p, flags, apend);
        else
            k_adj_rhs<float><<<nb(nrec, 128), 128, 0, st>>>((const double*)drec.p, (const int*)drs.p, nrec, dci,
                                                            (const float*)ttc, nd, kind, (const double*)dpo.p,
                                                            (const double*)dso.p, (double*)b_c.p, (double*)dmt.p, flags, apend);
        CK(hipGetLastError());
    }
    *out = GradPart{};
    out->mterm = (const double*)dmt.p;
    out->nrec = nrec;
    out->nsrc = nsrc;

    FimStats ast{};
    rc = adj_run(ctx, nd, prec, dci, (int)nsrc, nsrc * tiles_per, (const uint16_t*)code_c.p, (const double*)G_c.p,
                 P_c, (const double*)b_c.p, (double*)lam_c.p, ttc, &ast);
    if (rc) return rc;
    if (getenv("EIK_VERBOSE"))
        fprintf(stderr, "libeik: adjoint %lld rounds, %.0f tile activations of %lld tiles, %.2f sweeps per activation\n",
                (long long)ast.rounds, ast.activations, (long long)(nsrc * tiles_per),
                ast.sweeps / std::max(1.0, ast.activations));

    // coarse source points, grouped by source
    std::vector<SrcPt> pts;
    std::vector<int> poff(nsrc + 1, 0);
    if (refine) {
        const int cap = ctx->xcap;
        std::vector<int> ntr(nsrc);
        std::vector<SrcPt> xl((size_t)nsrc * cap);
        CK(hipMemcpyAsync(ntr.data(), ctx->buf("ntransfer").p, nsrc * sizeof(int), hipMemcpyDeviceToHost, st));
        CK(hipMemcpyAsync(xl.data(), ctx->buf("xlist").p, xl.size() * sizeof(SrcPt), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        for (int64_t s = 0; s < nsrc; s++) {
            if (ntr[s] > cap) return ctx->fail(EIK_EINVAL, "transferred-point list overflow");
            for (int k = 0; k < ntr[s]; k++) pts.push_back(xl[(size_t)s * cap + k]);
            poff[s + 1] = (int)pts.size();
        }
    } else {
        for (int64_t s = 0; s < nsrc; s++) {
            int c[8][3];
            const int ncn = find_enclosing_box(M->g, S.src.data() + s * nd, c);
            for (int k = 0; k < ncn; k++)
                pts.push_back(SrcPt{(int32_t)s, -1, c[k][0] + M->g.n[0] * (c[k][1] + M->g.n[1] * c[k][2])});
            poff[s + 1] = (int)pts.size();
        }
    }
    const int npts = (int)pts.size();
    DevBuf &dpts = ctx->buf("pts"), &dpoff = ctx->buf("poff"), &dup = ctx->buf("dpsiup"), &dus = ctx->buf("dpsidus"),
           &dt2 = ctx->buf("t2"), &dt1 = ctx->buf("t1"), &dsrc = ctx->buf("srcxyz");
    if (dpts.ensure(std::max(npts, 1) * sizeof(SrcPt)) || dpoff.ensure((nsrc + 1) * sizeof(int)) ||
        dup.ensure(std::max(npts, 1) * sizeof(double)) || dus.ensure(std::max(npts, 1) * sizeof(double)) ||
        dt2.ensure(std::max(npts, 1) * sizeof(double)) || dt1.ensure(std::max(npts, 1) * sizeof(double)) ||
        dsrc.ensure((size_t)nsrc * nd * sizeof(double)))
        return ctx->fail(EIK_ENOMEM, "out of device memory (source points)");
    CK(hipMemcpyAsync(dpts.p, pts.data(), npts * sizeof(SrcPt), hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(dpoff.p, poff.data(), (nsrc + 1) * sizeof(int), hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(dsrc.p, S.src.data(), (size_t)nsrc * nd * sizeof(double), hipMemcpyHostToDevice, st));
    if (npts > 0) {
        if (nd == 2)
            k_srcterms<2><<<nb(npts, 64), 64, 0, st>>>((const SrcPt*)dpts.p, npts, dci, (const uint16_t*)code_c.p,
                                                        (const double*)G_c.p, P_c, (const double*)lam_c.p,
                                                        (const double*)b_c.p, (double*)dup.p, (double*)dus.p);
        else
            k_srcterms<3><<<nb(npts, 64), 64, 0, st>>>((const SrcPt*)dpts.p, npts, dci, (const uint16_t*)code_c.p,
                                                        (const double*)G_c.p, P_c, (const double*)lam_c.p,
                                                        (const double*)b_c.p, (double*)dup.p, (double*)dus.p);
        CK(hipGetLastError());
    }

    // fine-grid adjoint (refinement)
    std::vector<SrcPt> fpts;
    std::vector<int> fpoff(nsrc + 1, 0);
    FimStats fst{};
    if (refine) {
        const InstDev* dfi = (const InstDev*)ctx->buf("inst_fine").p;
        const FineDev* dfd = (const FineDev*)ctx->buf("fine_desc").p;
        const double* ttf = (const double*)ctx->buf("tt_fine").p;
        const double* slf = (const double*)ctx->buf("slow_fine").p;
        std::vector<InstDev> finst(nsrc);
        std::vector<FineDev> fdesc(nsrc);
        CK(hipMemcpyAsync(finst.data(), dfi, nsrc * sizeof(InstDev), hipMemcpyDeviceToHost, st));
        CK(hipMemcpyAsync(fdesc.data(), dfd, nsrc * sizeof(FineDev), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        int64_t P_f = 0, ntiles_f = 0;
        for (int64_t s = 0; s < nsrc; s++) {
            P_f = std::max(P_f, finst[s].tt_off + (int64_t)finst[s].n[0] * finst[s].n[1] * finst[s].n[2]);
            ntiles_f = std::max(ntiles_f, finst[s].atile_base + (int64_t)finst[s].at[0] * finst[s].at[1] * finst[s].at[2]);
            for (int k = 0; k < fdesc[s].ncorn; k++) {
                const int* c = fdesc[s].corn[k];
                fpts.push_back(SrcPt{(int32_t)s, -1, c[0] + (int64_t)finst[s].n[0] * (c[1] + (int64_t)finst[s].n[1] * c[2])});
            }
            fpoff[s + 1] = (int)fpts.size();
        }
        const int nfp = (int)fpts.size();
        DevBuf &code_f = ctx->buf("adjf_code"), &G_f = ctx->buf("adjf_G"), &lam_f = ctx->buf("adjf_lam"),
               &b_f = ctx->buf("adjf_b"), &haq = ctx->buf("adjf_haq"), &dgi = ctx->buf("adjf_dgi"),
               &zf = ctx->buf("adjf_z"), &dfp = ctx->buf("fpts"), &dfpv = ctx->buf("fpts_val"),
               &dfpu = ctx->buf("fpts_u");
        if (code_f.ensure((size_t)P_f * sizeof(uint16_t)) || G_f.ensure((size_t)P_f * nd * sizeof(double)) ||
            lam_f.ensure((size_t)P_f * sizeof(double)) || b_f.ensure((size_t)P_f * sizeof(double)) ||
            haq.ensure((size_t)P_f * sizeof(double)) || dgi.ensure((size_t)P_f * sizeof(double)) ||
            zf.ensure((size_t)P_f * sizeof(double)) || dfp.ensure(std::max(nfp, 1) * sizeof(SrcPt)) ||
            dfpv.ensure(std::max(nfp, 1) * sizeof(double)) || dfpu.ensure(std::max(nfp, 1) * sizeof(double)))
            return ctx->fail(EIK_ENOMEM, "out of device memory (fine adjoint)");
        CK(hipMemsetAsync(lam_f.p, 0, (size_t)P_f * sizeof(double), st));
        CK(hipMemsetAsync(b_f.p, 0, (size_t)P_f * sizeof(double), st));
        CK(hipMemsetAsync(haq.p, 0, (size_t)P_f * sizeof(double), st));
        CK(hipMemsetAsync(dgi.p, 0, (size_t)P_f * sizeof(double), st));
        CK(hipMemcpyAsync(dfp.p, fpts.data(), nfp * sizeof(SrcPt), hipMemcpyHostToDevice, st));
        const TTPools fpools{(void*)ttf, ctx->buf("keyf_K").p, ctx->buf("keyf_S").p, (uint8_t*)ctx->buf("keyf_D").p};
        launch_prep<double, false>(nd, st, dfi, (int)nsrc, fpools, slf, (uint16_t*)code_f.p, (double*)G_f.p, P_f);
        if (npts > 0)
            k_fine_rhs<<<nb(npts, 64), 64, 0, st>>>((const SrcPt*)dpts.p, npts, dfi, ttf, (const double*)dup.p,
                                                    (const double*)dus.p, (double*)b_f.p, (double*)haq.p, flags, apend,
                                                    nd);
        CK(hipGetLastError());
        rc = adj_run(ctx, nd, EIK_PREC_FP64, dfi, (int)nsrc, ntiles_f, (const uint16_t*)code_f.p, (const double*)G_f.p,
                     P_f, (const double*)b_f.p, (double*)lam_f.p, ttf, &fst);
        if (rc) return rc;
        if (nfp > 0) {
            if (nd == 2)
                k_srcterms<2><<<nb(nfp, 64), 64, 0, st>>>((const SrcPt*)dfp.p, nfp, dfi, (const uint16_t*)code_f.p,
                                                           (const double*)G_f.p, P_f, (const double*)lam_f.p,
                                                           (const double*)b_f.p, (double*)dfpv.p, (double*)dfpu.p);
            else
                k_srcterms<3><<<nb(nfp, 64), 64, 0, st>>>((const SrcPt*)dfp.p, nfp, dfi, (const uint16_t*)code_f.p,
                                                           (const double*)G_f.p, P_f, (const double*)lam_f.p,
                                                           (const double*)b_f.p, (double*)dfpv.p, (double*)dfpu.p);
            k_scatter_pts<<<nb(nfp, 64), 64, 0, st>>>((const SrcPt*)dfp.p, nfp, dfi, (const double*)dfpv.p,
                                                      (double*)dgi.p);
        }
        dim3 gz(256, (unsigned)nsrc);
        k_fine_z<<<gz, 256, 0, st>>>(dfi, dfd, fpools, ttf, (const double*)lam_f.p, (const double*)dgi.p,
                                     (const double*)haq.p, M->d_vel, (const double*)dsrc.p, nd, kind, (double*)zf.p);
        if (npts > 0)
            k_fine_gather<<<nb(npts, 64), 64, 0, st>>>((const SrcPt*)dpts.p, npts, dfi, dfd, fpools, ttf, (const double*)zf.p,
                                                       (double*)dt2.p, (double*)dt1.p, nd);
        CK(hipGetLastError());
        if (gradsrc && whichgrad != EIK_GRADVEL) {
            DevBuf& dgs = ctx->buf("gradsrc");
            DevBuf& dfpo = ctx->buf("fpoff");
            if (dgs.ensure((size_t)nsrc * nd * sizeof(double)) || dfpo.ensure((nsrc + 1) * sizeof(int)))
                return ctx->fail(EIK_ENOMEM, "out of device memory");
            CK(hipMemcpyAsync(dfpo.p, fpoff.data(), (nsrc + 1) * sizeof(int), hipMemcpyHostToDevice, st));
            k_gradsrc<<<nb(nsrc, 64), 64, 0, st>>>((int)nsrc, nd, (const SrcPt*)dfp.p, (const int*)dfpo.p, dfi,
                                                   M->d_vel, (const double*)dsrc.p, (const double*)dfpv.p, nullptr,
                                                   dfd, M->d_vel, (const double*)haq.p, (double*)dgs.p);
            CK(hipGetLastError());
            CK(hipMemcpyAsync(gradsrc, dgs.p, (size_t)nsrc * nd * sizeof(double), hipMemcpyDeviceToHost, st));
        }
    } else {
        if (npts > 0)
            k_src_own<<<nb(npts, 64), 64, 0, st>>>((const SrcPt*)dpts.p, npts, dci, M->d_vel, (const double*)dsrc.p, nd,
                                                   kind, (const double*)dup.p, (const double*)dus.p, (double*)dt2.p,
                                                   (double*)dt1.p);
        if (gradsrc && whichgrad != EIK_GRADVEL) {
            DevBuf& dgs = ctx->buf("gradsrc");
            if (dgs.ensure((size_t)nsrc * nd * sizeof(double))) return ctx->fail(EIK_ENOMEM, "out of device memory");
            k_gradsrc<<<nb(nsrc, 64), 64, 0, st>>>((int)nsrc, nd, (const SrcPt*)dpts.p, (const int*)dpoff.p, dci,
                                                   M->d_vel, (const double*)dsrc.p, (const double*)dup.p,
                                                   (const double*)dus.p, nullptr, nullptr, nullptr, (double*)dgs.p);
            CK(hipGetLastError());
            CK(hipMemcpyAsync(gradsrc, dgs.p, (size_t)nsrc * nd * sizeof(double), hipMemcpyDeviceToHost, st));
        }
    }

    if (whichgrad != EIK_GRADSRCLOC) {
        DevBuf &img = ctx->buf("grad_img"), &dtap = ctx->buf("taper");
        if (img.ensure((size_t)P_c * sizeof(double)) || dtap.ensure((size_t)nsrc * sizeof(TaperDev)))
            return ctx->fail(EIK_ENOMEM, "out of device memory (gradient images)");
        std::vector<TaperDev> tap(nsrc);
        for (int64_t s = 0; s < nsrc; s++) taper_box(M->g, S.src.data() + s * nd, P->radiussmoothgradsrc, tap[s]);
        CK(hipMemcpyAsync(dtap.p, tap.data(), nsrc * sizeof(TaperDev), hipMemcpyHostToDevice, st));
        dim3 gi(512, (unsigned)nsrc);
        if (prec != EIK_PREC_FP32)
            k_grad_img<double><<<gi, 256, 0, st>>>(dci, (const double*)ttc, (const double*)lam_c.p, M->d_vel,
                                                   (double*)img.p, N);
        else
            k_grad_img<float><<<gi, 256, 0, st>>>(dci, (const float*)ttc, (const double*)lam_c.p, M->d_vel,
                                                  (double*)img.p, N);
        if (defer) {
            out->deferred = true;
            out->inst = dci;
            out->pts = dpts.p;
            out->poff = dpoff.p;
            out->tap = dtap.p;
            out->t2 = dt2.p;
            out->t1 = dt1.p;
        } else {
            k_grad_quirk<<<1, 256, 0, st>>>((const SrcPt*)dpts.p, (const int*)dpoff.p, (int)nsrc, dci,
                                            (const TaperDev*)dtap.p, nd, kind, (const double*)dt2.p, (const double*)dt1.p,
                                            (double*)img.p, nullptr);
            k_taper<<<(unsigned)nsrc, 256, 0, st>>>(dci, (const TaperDev*)dtap.p, nd, kind, (double*)img.p);
        }
        CK(hipGetLastError());
        out->img = (double*)img.p;
    }
    CK(hipStreamSynchronize(st));
    ctx->stats[3] = tm.mark();
    return EIK_OK;
}

int gradient_total(eik_ctx* ctx, eik_model* M, const eik_params* P, int whichgrad, const GradPart* parts, int nparts,
                   const int64_t* rec_off, int64_t nsrc, double* gradvel, double* misf) {
    const int nd = M->g.ndim;
    const int64_t N = M->N, nrec = rec_off[nsrc];
    hipStream_t st = ctx->stream;
    Timer tm(ctx);
    // misfit over the receivers' terms of every part, source by source
    DevBuf &dmt = ctx->buf("mterm_all"), &droff = ctx->buf("rec_off_all"), &dmis = ctx->buf("misf");
    if (dmt.ensure((size_t)std::max<int64_t>(nrec, 1) * sizeof(double)) ||
        droff.ensure((size_t)(nsrc + 1) * sizeof(int64_t)) || dmis.ensure(sizeof(double)))
        return ctx->fail(EIK_ENOMEM, "out of device memory (misfit)");
    int64_t off = 0;
    for (int i = 0; i < nparts; i++) {
        if (parts[i].nrec > 0)
            CK(hipMemcpyAsync((double*)dmt.p + off, parts[i].mterm, (size_t)parts[i].nrec * sizeof(double),
                              hipMemcpyDeviceToDevice, st));
        off += parts[i].nrec;
    }
    CK(hipMemcpyAsync(droff.p, rec_off, (size_t)(nsrc + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    k_misfit<<<1, kMisfitThreads, 0, st>>>((const double*)dmt.p, (const int64_t*)droff.p, nsrc, (double*)dmis.p);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(misf, dmis.p, sizeof(double), hipMemcpyDeviceToHost, st));
    if (gradvel && whichgrad != EIK_GRADSRCLOC) {
        DevBuf& gout = ctx->buf("grad_out");
        if (gout.ensure((size_t)N * 2 * sizeof(double))) return ctx->fail(EIK_ENOMEM, "out of device memory (gradient)");
        // a continuing part's source-node carry-over (from the previous part's last tapered image), then its taper
        for (int i = 1; i < nparts; i++) {
            const GradPart& q = parts[i];
            if (!q.deferred || q.nsrc == 0) continue;
            const InstDev* qi = (const InstDev*)q.inst;
            k_grad_quirk<<<1, 256, 0, st>>>((const SrcPt*)q.pts, (const int*)q.poff, (int)q.nsrc, qi,
                                            (const TaperDev*)q.tap, nd, M->g.kind, (const double*)q.t2,
                                            (const double*)q.t1, q.img, parts[i - 1].img + (parts[i - 1].nsrc - 1) * N);
            k_taper<<<(unsigned)q.nsrc, 256, 0, st>>>(qi, (const TaperDev*)q.tap, nd, M->g.kind, q.img);
            CK(hipGetLastError());
        }
        k_sum_sources<<<nb(N, 256), 256, 0, st>>>(parts[0].img, (int)parts[0].nsrc, nparts > 1 ? parts[1].img : nullptr,
                                                  nparts > 1 ? (int)parts[1].nsrc : 0, N, (double*)gout.p);
        CK(hipGetLastError());
        double* res = (double*)gout.p;
        if (P->smoothgradkern > 0) {
            const int l = P->smoothgradkern;
            int nk = 0;
            for (int d = 0; d < nd; d++) nk += M->g.n[d] != 1;
            int64_t kn = 1;
            for (int d = 0; d < nk; d++) kn *= l;
            std::vector<double> ker(kn);
            const int m = l / 2;
            const double s2 = ((double)l / 2.0) - sqrt((double)l / 2.0);
            double ks = 0.0;
            for (int64_t q = 0; q < kn; q++) {
                int64_t r = q;
                double rr2 = 0.0;
                for (int d = 0; d < nk; d++) {
                    const int64_t pos = r % l - m;
                    r /= l;
                    rr2 += (double)(pos * pos);
                }
                const double rr = sqrt(rr2);
                ker[q] = exp(-(rr * rr) / (2.0 * s2));
                ks += ker[q];
            }
            for (auto& x : ker) x /= ks;
            DevBuf& dk = ctx->buf("smooth_ker");
            if (dk.ensure(kn * sizeof(double))) return ctx->fail(EIK_ENOMEM, "out of device memory");
            CK(hipMemcpyAsync(dk.p, ker.data(), kn * sizeof(double), hipMemcpyHostToDevice, st));
            k_smooth<<<nb(N, 256), 256, 0, st>>>((const double*)gout.p, (double*)gout.p + N, (const double*)dk.p, l, nd,
                                                 (int)M->g.n[0], (int)M->g.n[1], (int)M->g.n[2]);
            CK(hipGetLastError());
            res = (double*)gout.p + N;
        }
        CK(hipMemcpyAsync(gradvel, res, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    CK(hipStreamSynchronize(st));
    ctx->stats[3] += tm.mark();
    return EIK_OK;
}

int gradient_solve(eik_ctx* ctx, eik_model* M, const SourceSet& S, const eik_params* P, int whichgrad,
                   const double* pickobs, const double* stdobs, double* gradvel, double* gradsrc, double* misf) {
    GradPart part{};
    const int rc = gradient_part(ctx, M, S, P, whichgrad, pickobs, stdobs, gradsrc, false, &part);
    if (rc) return rc;
    return gradient_total(ctx, M, P, whichgrad, &part, 1, S.rec_off.data(), S.nsrc, gradvel, misf);
}

}  // namespace eik
