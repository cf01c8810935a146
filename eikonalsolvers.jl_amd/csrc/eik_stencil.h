// eik_stencil.h -- the local traveltime update used by the tiled fast-iterative solver.
//
// The reference accepts nodes from a binary heap (ttFMM_core!, src/EikSolv/eikmainfwd.jl:630-814)
// and recomputes a narrow-band node with calcttpt_2ndord! (eikforward2D.jl:98-433,
// eikforward3D.jl:151-532) every time one of its neighbours is accepted, keeping the LAST value
// (update_node_minheap! replaces the key, src/BinHeap/BinHeaps.jl:394-430).  A node's final
// traveltime is therefore a function of its neighbours' values and of the ORDER in which they
// were accepted.  replay() re-enacts that sequence locally from the current neighbour values:
//   1. if any adjacent node is a source/init node (accepted before the march starts), evaluate
//      the stencil with only those accepted (initial narrow band, eikmainfwd.jl:590-621);
//   2. take the other adjacent neighbours in the heap's pop order (their pop keys, below; equal
//      keys form one batch, eikmainfwd.jl:647-666); while the next one pops before this node it
//      would be accepted first, so re-evaluate with every node that pops no later than it
//      counted as accepted (this also decides whether the 2-away nodes of the 2nd-order stencil
//      were accepted at that moment);
//   3. stop at the first neighbour that does not pop first: the heap would pop this node first.
// At the fixed point of this update over the whole grid every node holds the value the heap
// FMM gives it; in FP64 the arithmetic below follows the reference's operation order so the
// results agree bit for bit.
//
// Neighbour slot k = 2*axis + l, l = 0 for the + side and 1 for the - side (the reference scans
// the + side first).  ok1 bit k: the adjacent node is inside the grid (isonb1st false); ok2 bit k:
// that adjacent node is off every boundary plane (isonb2nd false, eikutils.jl:129-162), so the
// 2-away node may be used.  f1/f2 bits: the adjacent / 2-away node is a fixed (init) node.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define EIK_INF (__builtin_huge_val())
#define EIK_INFF (__builtin_huge_valf())

namespace eik {

// ---------------------------------------------------------------- FP64 (reference arithmetic)
template <int ND>
__device__ __forceinline__ double eval64(const double* v1, const double* v2, unsigned acc1, unsigned acc2,
                                         double s, const double* dh, int* code) {
#pragma clang fp contract(off)
    double alpha = 0.0, beta = 0.0, gamma = -(s * s);
#pragma unroll
    for (int d = 0; d < ND; d++) code[d] = 0;
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        bool use1 = false, use2 = false;
        double cv1 = EIK_INF, cv2 = EIK_INF;
#pragma unroll
        for (int l = 0; l < 2; l++) {
            const int k = 2 * ax + l;
            const int sh = (l == 0) ? 1 : -1;
            const bool a1 = (acc1 >> k) & 1u;
            if (a1) {
                const double tv1 = v1[k];
                if (tv1 < cv1) {
                    cv1 = tv1;
                    use1 = true;
                    code[ax] = -sh;
                    const bool a2 = (acc2 >> k) & 1u;
                    if (a2) {
                        const double tv2 = v2[k];
                        if (tv2 <= cv1) {
                            cv2 = tv2;
                            use2 = true;
                            code[ax] = -2 * sh;
                        } else {
                            cv2 = EIK_INF;
                            use2 = false;
                            code[ax] = -sh;
                        }
                    }
                }
            }
        }
        const double h = dh[ax];
        if (use2 && use1) {
            const double tmpa2 = 1.0 / 3.0 * (4.0 * cv1 - cv2);
            const double ca = 9.0 / (4.0 * (h * h));
            alpha += ca;
            beta += (-2.0 * ca * tmpa2);
            gamma += ca * (tmpa2 * tmpa2);
        } else if (use1) {
            const double ca = 1.0 / (h * h);
            alpha += ca;
            beta += (-2.0 * ca * cv1);
            gamma += ca * (cv1 * cv1);
        }
    }
    double sqarg = beta * beta - 4.0 * alpha * gamma;
    if (sqarg < 0.0) {  // all axes 1st order (eikforward2D.jl:265-337)
#pragma unroll
        for (int d = 0; d < ND; d++) code[d] = 0;
        alpha = 0.0;
        beta = 0.0;
        gamma = -(s * s);
#pragma unroll
        for (int ax = 0; ax < ND; ax++) {
            bool use1 = false;
            double cv1 = EIK_INF;
#pragma unroll
            for (int l = 0; l < 2; l++) {
                const int k = 2 * ax + l;
                const bool a1 = (acc1 >> k) & 1u;
                if (a1 && v1[k] < cv1) {
                    cv1 = v1[k];
                    use1 = true;
                    code[ax] = (l == 0) ? -1 : 1;
                }
            }
            if (use1) {
                const double h = dh[ax];
                const double ca = 1.0 / (h * h);
                alpha += ca;
                beta += (-2.0 * ca * cv1);
                gamma += ca * (cv1 * cv1);
            }
        }
        sqarg = beta * beta - 4.0 * alpha * gamma;
    }
    if (sqarg < 0.0) {  // one dimension (eikforward2D.jl:340-421)
        double cvv[3] = {EIK_INF, EIK_INF, EIK_INF};
        int cc[3] = {0, 0, 0};
#pragma unroll
        for (int ax = 0; ax < ND; ax++) {
#pragma unroll
            for (int l = 0; l < 2; l++) {
                const int k = 2 * ax + l;
                const bool a1 = (acc1 >> k) & 1u;
                if (a1 && v1[k] < cvv[ax]) {
                    cvv[ax] = v1[k];
                    cc[ax] = (l == 0) ? -1 : 1;
                }
            }
        }
        int idc = 0;
#pragma unroll
        for (int ax = 1; ax < ND; ax++)
            if (cvv[ax] < cvv[idc]) idc = ax;
#pragma unroll
        for (int d = 0; d < ND; d++) code[d] = (d == idc) ? cc[d] : 0;
        return (dh[idc] * s) + cvv[idc];
    }
    const double tmpsq = sqrt(sqarg);
    const double t1 = (-beta + tmpsq) / (2.0 * alpha);
    const double t2 = (-beta - tmpsq) / (2.0 * alpha);
    return fmax(t1, t2);
}

// ---------------------------------------------------------------- FP32 (shifted quadratic)
// Same neighbour selection as eval64.  The quadratic sum_a alpha_a (T - c_a)^2 = s^2 is solved
// for delta = T - r, r = the smallest chosen neighbour value, with c_a - r formed from
// differences of nearby traveltimes, so FP32 keeps ~7 digits of the increment instead of losing
// them to the cancellation in beta^2 - 4 alpha gamma (both terms ~ (alpha T)^2).
template <int ND>
__device__ __forceinline__ float eval32(const float* v1, const float* v2, unsigned acc1, unsigned acc2, float s,
                                        const float* dh, int* code) {
    float t1a[3], t2a[3], al[3];
    int mode[3];  // 0 unused, 1 first order, 2 second order
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        bool use1 = false, use2 = false;
        float cv1 = EIK_INFF, cv2 = EIK_INFF;
        code[ax] = 0;
#pragma unroll
        for (int l = 0; l < 2; l++) {
            const int k = 2 * ax + l;
            const int sh = (l == 0) ? 1 : -1;
            const bool a1 = (acc1 >> k) & 1u;
            if (a1) {
                const float tv1 = v1[k];
                if (tv1 < cv1) {
                    cv1 = tv1;
                    use1 = true;
                    code[ax] = -sh;
                    const bool a2 = (acc2 >> k) & 1u;
                    if (a2) {
                        const float tv2 = v2[k];
                        if (tv2 <= cv1) {
                            cv2 = tv2;
                            use2 = true;
                            code[ax] = -2 * sh;
                        } else {
                            cv2 = EIK_INFF;
                            use2 = false;
                            code[ax] = -sh;
                        }
                    }
                }
            }
        }
        const float ih2 = 1.0f / (dh[ax] * dh[ax]);
        mode[ax] = (use1 && use2) ? 2 : (use1 ? 1 : 0);
        t1a[ax] = cv1;
        t2a[ax] = cv2;
        al[ax] = (mode[ax] == 2) ? 2.25f * ih2 : ih2;
    }
    float r = EIK_INFF;
#pragma unroll
    for (int ax = 0; ax < ND; ax++)
        if (mode[ax]) r = fminf(r, t1a[ax]);
    float A = 0.f, B = 0.f, C = -(s * s);
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        if (!mode[ax]) continue;
        float d = t1a[ax] - r;
        if (mode[ax] == 2) d += (t1a[ax] - t2a[ax]) * (1.0f / 3.0f);
        A += al[ax];
        B += al[ax] * d;
        C += al[ax] * d * d;
    }
    float disc = B * B - A * C;
    if (disc < 0.f) {  // all axes 1st order
        A = 0.f;
        B = 0.f;
        C = -(s * s);
        r = EIK_INFF;
        float c1[3];
        bool u[3];
#pragma unroll
        for (int ax = 0; ax < ND; ax++) {
            c1[ax] = EIK_INFF;
            u[ax] = false;
            code[ax] = 0;
#pragma unroll
            for (int l = 0; l < 2; l++) {
                const int k = 2 * ax + l;
                const bool a1 = (acc1 >> k) & 1u;
                if (a1 && v1[k] < c1[ax]) {
                    c1[ax] = v1[k];
                    u[ax] = true;
                    code[ax] = (l == 0) ? -1 : 1;
                }
            }
            if (u[ax]) r = fminf(r, c1[ax]);
        }
#pragma unroll
        for (int ax = 0; ax < ND; ax++) {
            if (!u[ax]) continue;
            const float ih2 = 1.0f / (dh[ax] * dh[ax]);
            const float d = c1[ax] - r;
            A += ih2;
            B += ih2 * d;
            C += ih2 * d * d;
        }
        disc = B * B - A * C;
        if (disc < 0.f) {  // one dimension
            int idc = 0;
#pragma unroll
            for (int ax = 1; ax < ND; ax++)
                if (c1[ax] < c1[idc]) idc = ax;
            const int keep = code[idc];
#pragma unroll
            for (int d = 0; d < ND; d++) code[d] = (d == idc) ? keep : 0;
            return c1[idc] + dh[idc] * s;
        }
    }
    return r + (B + sqrtf(disc)) / A;
}

// ---------------------------------------------------------------- heap pop keys
// The heap pops nodes in value order except after a non-causal 2nd-order update: a node whose
// new value falls below the value of the neighbour whose acceptance triggered it is popped right
// after that neighbour, and such nodes can trigger further ones ("cascade").  The pop order is
// reproduced by the key (K, S, D): K = level (for an ordinary node its own value), S = sub-level
// inside a cascade (-inf for ordinary nodes), D = cascade depth.  A node whose replay ends below
// the key of its trigger c gets (K_c, max(t, S_c), D_c + 1); otherwise (t, -inf, 0).  Keys compare
// lexicographically; equal keys pop in one batch.
template <typename T>
struct Key {
    T K, S;
    int D;
};
template <typename T>
__device__ __forceinline__ bool klt(const Key<T>& a, const Key<T>& b) {
    if (a.K != b.K) return a.K < b.K;
    if (a.S != b.S) return a.S < b.S;
    return a.D < b.D;
}
template <typename T>
__device__ __forceinline__ bool keq(const Key<T>& a, const Key<T>& b) {
    return a.K == b.K && a.S == b.S && a.D == b.D;
}
// A cascade can never be deeper than this; a longer chain can only come from two nodes feeding
// each other stale keys during the iteration, and ignoring such a candidate breaks the cycle.
constexpr int kMaxCascade = 16;

template <int ND, typename T>
__device__ __forceinline__ T eval_any(const T* v1, const T* v2, unsigned acc1, unsigned acc2, T s, const T* dh,
                                      int* code);
template <>
__device__ __forceinline__ double eval_any<2, double>(const double* v1, const double* v2, unsigned acc1, unsigned acc2,
                                                      double s, const double* dh, int* code) {
    return eval64<2>(v1, v2, acc1, acc2, s, dh, code);
}
template <>
__device__ __forceinline__ double eval_any<3, double>(const double* v1, const double* v2, unsigned acc1, unsigned acc2,
                                                      double s, const double* dh, int* code) {
    return eval64<3>(v1, v2, acc1, acc2, s, dh, code);
}
template <>
__device__ __forceinline__ float eval_any<2, float>(const float* v1, const float* v2, unsigned acc1, unsigned acc2,
                                                    float s, const float* dh, int* code) {
    return eval32<2>(v1, v2, acc1, acc2, s, dh, code);
}
template <>
__device__ __forceinline__ float eval_any<3, float>(const float* v1, const float* v2, unsigned acc1, unsigned acc2,
                                                    float s, const float* dh, int* code) {
    return eval32<3>(v1, v2, acc1, acc2, s, dh, code);
}

template <typename T>
__device__ __forceinline__ T tinf();
template <>
__device__ __forceinline__ double tinf<double>() {
    return EIK_INF;
}
template <>
__device__ __forceinline__ float tinf<float>() {
    return EIK_INFF;
}

// Local replay of the heap (see the file comment).  v1/v2: adjacent / 2-away values; k1/k2 their
// keys; f1/f2 fixed (init) bits; ok1/ok2 grid-boundary bits.  Returns the value and sets *kout to
// the node's pop key and code to the derivative codes of the last stencil evaluation.
template <int ND, typename T>
__device__ __forceinline__ T replay(const T* v1, const T* v2, const Key<T>* k1, const Key<T>* k2, unsigned f1,
                                    unsigned f2, unsigned ok1, unsigned ok2, T s, const T* dh, int* code,
                                    Key<T>* kout) {
    int ord[2 * ND];
    int nc = 0;
    bool haveinit = false;
#pragma unroll
    for (int k = 0; k < 2 * ND; k++) {
        if (!((ok1 >> k) & 1u)) continue;
        if ((f1 >> k) & 1u) {
            haveinit = true;
        } else if (v1[k] < tinf<T>() && k1[k].D < kMaxCascade) {
            int j = nc++;
            while (j > 0 && klt(k1[k], k1[ord[j - 1]])) {
                ord[j] = ord[j - 1];
                j--;
            }
            ord[j] = k;
        }
    }
    T cur = tinf<T>();
    Key<T> kz{tinf<T>(), -tinf<T>(), 0};
#pragma unroll
    for (int d = 0; d < ND; d++) code[d] = 0;
    if (haveinit) {
        cur = eval_any<ND, T>(v1, v2, f1 & ok1, f2 & ok2, s, dh, code);
        kz = Key<T>{cur, -tinf<T>(), 0};
    }
    for (int i = 0; i < nc; i++) {
        const Key<T> c = k1[ord[i]];
        if (i + 1 < nc && keq(k1[ord[i + 1]], c)) continue;  // one batch of equal keys
        if (!klt(c, kz)) break;
        unsigned acc1 = 0u, acc2 = 0u;
#pragma unroll
        for (int k = 0; k < 2 * ND; k++) {
            const bool a1 = ((ok1 >> k) & 1u) && (((f1 >> k) & 1u) || (v1[k] < tinf<T>() && !klt(c, k1[k])));
            const bool a2 = ((ok2 >> k) & 1u) && (((f2 >> k) & 1u) || (v2[k] < tinf<T>() && !klt(c, k2[k])));
            acc1 |= (unsigned)a1 << k;
            acc2 |= (unsigned)a2 << k;
        }
        cur = eval_any<ND, T>(v1, v2, acc1, acc2, s, dh, code);
        if (cur >= c.K)
            kz = Key<T>{cur, -tinf<T>(), 0};
        else
            kz = Key<T>{c.K, cur > c.S ? cur : c.S, c.D + 1};
    }
    *kout = kz;
    return cur;
}

}  // namespace eik
