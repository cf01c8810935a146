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
// Arithmetic.  FP64 (eval64) follows the reference bit for bit.  The mixed precision (MIX, eval_mix)
// keeps values, pop keys and every comparison of the replay in FP64 and forms only the quadratic's
// increment over the smallest chosen neighbour in FP32; FP32 storage (eval32) does everything in FP32.
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


// Arithmetic type of the quadratic: the storage type, except FP32 for the mixed precision.
template <typename T, bool MIX>
struct ArithOf {
    using type = T;
};
template <>
struct ArithOf<double, true> {
    using type = float;
};
template <typename T, bool MIX>
using Arith = typename ArithOf<T, MIX>::type;

// Per-node stencil metric: spacing per axis and the quadratic coefficients of the 1st-order
// (1/h^2) and 2nd-order (9/(4 h^2)) one-sided differences, computed once per node update rather
// than once per stencil evaluation.
template <typename T>
struct Metric {
    T dh[3], c1[3], c2[3];
};
__device__ __forceinline__ void metric_coefs(Metric<double>& m, int nd) {
#pragma clang fp contract(off)
    for (int ax = 0; ax < nd; ax++) {
        const double h = m.dh[ax];
        m.c1[ax] = 1.0 / (h * h);
        m.c2[ax] = 9.0 / (4.0 * (h * h));
    }
}
__device__ __forceinline__ void metric_coefs(Metric<float>& m, int nd) {
    for (int ax = 0; ax < nd; ax++) {
        const float ih2 = 1.0f / (m.dh[ax] * m.dh[ax]);
        m.c1[ax] = ih2;
        m.c2[ax] = 2.25f * ih2;
    }
}

// Neighbour choice along one axis, as calcttpt_2ndord! scans it (+ side first, strict <): cv1 =
// the smaller accepted adjacent value, 2nd order if that side's 2-away node is accepted and not
// larger.  When the - side wins without a usable 2-away node the + side's 2nd-order choice is
// kept (the reference does not reset it).  Written with selects only, so a wave never diverges.
template <typename T>
struct AxisPick {
    T cv1, cv2;
    bool use1, use2;
    int codeA, codeB;  // derivative code of the mixed-order and of the all-1st-order stencil
};
template <typename T>
__device__ __forceinline__ AxisPick<T> pick_axis(const T* v1, const T* v2, unsigned acc1, unsigned acc2, int ax) {
    const int kp = 2 * ax, km = 2 * ax + 1;
    AxisPick<T> p;
    // + side (cv1 starts at +inf)
    const bool p1 = (bool)((acc1 >> kp) & 1u) & (v1[kp] < tinf<T>());
    const T c1p = p1 ? v1[kp] : tinf<T>();
    const bool p2 = p1 & (bool)((acc2 >> kp) & 1u);
    const bool u2p = p2 & (v2[kp] <= c1p);
    // - side
    const bool q1 = (bool)((acc1 >> km) & 1u) & (v1[km] < c1p);
    const bool q2 = q1 & (bool)((acc2 >> km) & 1u);
    p.cv1 = q1 ? v1[km] : c1p;
    const bool u2m = q2 & (v2[km] <= p.cv1);
    p.use1 = p1 | q1;
    p.use2 = q2 ? u2m : u2p;
    p.cv2 = q2 ? (u2m ? v2[km] : tinf<T>()) : (u2p ? v2[kp] : tinf<T>());
    p.codeA = q1 ? (u2m ? 2 : 1) : (p1 ? (u2p ? -2 : -1) : 0);
    p.codeB = q1 ? 1 : (p1 ? -1 : 0);
    return p;
}

// ---------------------------------------------------------------- FP64 (reference arithmetic)
// calcttpt_2ndord! with the reference's operation order: the mixed-order quadratic, its
// all-1st-order fallback when the discriminant is negative, then the 1-D fallback t1 + h s.
// The fallbacks are rare, so they are computed only when a lane of the wave needs them; one
// square root and one division are taken on the selected quadratic; max(t1, t2) is always t1
// here (alpha > 0, rounding is monotone).
template <int ND>
__device__ __forceinline__ double eval64(const double* v1, const double* v2, unsigned acc1, unsigned acc2,
                                         double s, const Metric<double>& m, int* code) {
#pragma clang fp contract(off)
    double alpha = 0.0, beta = 0.0, gamma = -(s * s);
    AxisPick<double> p[3];
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        p[ax] = pick_axis<double>(v1, v2, acc1, acc2, ax);
        const bool sec = p[ax].use1 && p[ax].use2;
        const double tmpa2 = 1.0 / 3.0 * (4.0 * p[ax].cv1 - p[ax].cv2);
        const double ca = sec ? m.c2[ax] : m.c1[ax];
        const double tv = sec ? tmpa2 : p[ax].cv1;
        alpha = p[ax].use1 ? alpha + ca : alpha;
        beta = p[ax].use1 ? beta + (-2.0 * ca * tv) : beta;
        gamma = p[ax].use1 ? gamma + ca * (tv * tv) : gamma;
        code[ax] = p[ax].codeA;
    }
    const double sqA = beta * beta - 4.0 * alpha * gamma;
    double a = alpha, b = beta, q = sqA, oned = 0.0;
    bool useoned = false;
    if (__ballot(sqA < 0.0) != 0ull) {
        double alpha1 = 0.0, beta1 = 0.0, gamma1 = -(s * s);
        double cmin = EIK_INF, hmin = m.dh[0];
        int idc = 0;
#pragma unroll
        for (int ax = 0; ax < ND; ax++) {
            const double cb = m.c1[ax];
            alpha1 = p[ax].use1 ? alpha1 + cb : alpha1;
            beta1 = p[ax].use1 ? beta1 + (-2.0 * cb * p[ax].cv1) : beta1;
            gamma1 = p[ax].use1 ? gamma1 + cb * (p[ax].cv1 * p[ax].cv1) : gamma1;
            // 1-D fallback: first axis with the smallest adjacent value
            const bool better = (ax == 0) || p[ax].cv1 < cmin;
            idc = better ? ax : idc;
            cmin = better ? p[ax].cv1 : cmin;
            hmin = better ? m.dh[ax] : hmin;
        }
        const double sqB = beta1 * beta1 - 4.0 * alpha1 * gamma1;
        const bool useA = !(sqA < 0.0);
        const bool useB = !useA && !(sqB < 0.0);
        a = useA ? alpha : alpha1;
        b = useA ? beta : beta1;
        q = useA ? sqA : sqB;
        useoned = !useA && !useB;
        oned = (hmin * s) + cmin;
#pragma unroll
        for (int d = 0; d < ND; d++)
            code[d] = useA ? p[d].codeA : (useB ? p[d].codeB : (d == idc ? p[d].codeB : 0));
    }
    const double t = (-b + sqrt(q)) / (2.0 * a);
    return useoned ? oned : t;
}

// ---------------------------------------------------------------- FP32 (shifted quadratic)
// Same neighbour selection as eval64.  The quadratic sum_a alpha_a (T - c_a)^2 = s^2 is solved
// for delta = T - r, r = the smallest chosen neighbour value, with c_a - r formed from
// differences of nearby traveltimes, so FP32 keeps ~7 digits of the increment instead of losing
// them to the cancellation in beta^2 - 4 alpha gamma (both terms ~ (alpha T)^2).
template <int ND>
__device__ __forceinline__ float eval32(const float* v1, const float* v2, unsigned acc1, unsigned acc2, float s,
                                        const Metric<float>& m, int* code) {
    AxisPick<float> p[3];
    float r = EIK_INFF;
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        p[ax] = pick_axis<float>(v1, v2, acc1, acc2, ax);
        r = p[ax].use1 ? fminf(r, p[ax].cv1) : r;
    }
    float A = 0.f, B = 0.f, C = -(s * s);
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        const bool sec = p[ax].use1 && p[ax].use2;
        const float al = sec ? m.c2[ax] : m.c1[ax];
        const float d1 = p[ax].cv1 - r;
        const float d = sec ? d1 + (p[ax].cv1 - p[ax].cv2) * (1.0f / 3.0f) : d1;
        A = p[ax].use1 ? A + al : A;
        B = p[ax].use1 ? B + al * d : B;
        C = p[ax].use1 ? C + al * d * d : C;
        code[ax] = p[ax].codeA;
    }
    const float discA = B * B - A * C;
    float a = A, b = B, q = discA, oned = 0.f;
    bool useoned = false;
    if (__ballot(discA < 0.f) != 0ull) {
        float A1 = 0.f, B1 = 0.f, C1 = -(s * s);
        float cmin = EIK_INFF, hmin = m.dh[0];
        int idc = 0;
#pragma unroll
        for (int ax = 0; ax < ND; ax++) {
            const float ih2 = m.c1[ax];
            const float d1 = p[ax].cv1 - r;
            A1 = p[ax].use1 ? A1 + ih2 : A1;
            B1 = p[ax].use1 ? B1 + ih2 * d1 : B1;
            C1 = p[ax].use1 ? C1 + ih2 * d1 * d1 : C1;
            const bool better = (ax == 0) || p[ax].cv1 < cmin;
            idc = better ? ax : idc;
            cmin = better ? p[ax].cv1 : cmin;
            hmin = better ? m.dh[ax] : hmin;
        }
        const float discB = B1 * B1 - A1 * C1;
        const bool useA = !(discA < 0.f);
        const bool useB = !useA && !(discB < 0.f);
        a = useA ? A : A1;
        b = useA ? B : B1;
        q = useA ? discA : discB;
        useoned = !useA && !useB;
        oned = cmin + hmin * s;
#pragma unroll
        for (int d = 0; d < ND; d++)
            code[d] = useA ? p[d].codeA : (useB ? p[d].codeB : (d == idc ? p[d].codeB : 0));
    }
    const float t = r + (b + sqrtf(q)) / a;
    return useoned ? oned : t;
}

// ---------------------------------------------------------------- mixed (FP64 values, FP32 increment)
// Same neighbour selection as eval64 (in FP64).  The quadratic is eval32's shifted form, solved in
// FP32 for delta = T - r with the chosen values' offsets c_a - r formed in FP64 and rounded once
// to FP32; the result r + delta is formed in FP64, so a value carries the FP64 precision of its
// smallest neighbour plus an FP32-accurate increment (~1e-7 of a node crossing), and the replay's
// comparisons, which decide the heap order, stay exact.  The 1-D fallback is cmin + (h s in FP32).
template <int ND>
__device__ __forceinline__ double eval_mix(const double* v1, const double* v2, unsigned acc1, unsigned acc2, float s,
                                           const Metric<float>& m, int* code) {
#pragma clang fp contract(off)
    AxisPick<double> p[3];
    double r = EIK_INF;
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        p[ax] = pick_axis<double>(v1, v2, acc1, acc2, ax);
        r = (p[ax].use1 && p[ax].cv1 < r) ? p[ax].cv1 : r;
    }
    float A = 0.f, B = 0.f, C = -(s * s);
    float d1v[3];
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        const bool sec = p[ax].use1 && p[ax].use2;
        const float al = sec ? m.c2[ax] : m.c1[ax];
        const float d1 = (float)(p[ax].cv1 - r);
        const float d = sec ? d1 + (float)(p[ax].cv1 - p[ax].cv2) * (1.0f / 3.0f) : d1;
        d1v[ax] = d1;
        A = p[ax].use1 ? A + al : A;
        B = p[ax].use1 ? B + al * d : B;
        C = p[ax].use1 ? C + al * d * d : C;
        code[ax] = p[ax].codeA;
    }
    const float discA = B * B - A * C;
    float a = A, b = B, q = discA, hs = 0.f;
    double cmin = EIK_INF;
    bool useoned = false;
    if (__ballot(discA < 0.f) != 0ull) {
        float A1 = 0.f, B1 = 0.f, C1 = -(s * s);
        float hmin = m.dh[0];
        int idc = 0;
#pragma unroll
        for (int ax = 0; ax < ND; ax++) {
            const float ih2 = m.c1[ax];
            const float d1 = d1v[ax];
            A1 = p[ax].use1 ? A1 + ih2 : A1;
            B1 = p[ax].use1 ? B1 + ih2 * d1 : B1;
            C1 = p[ax].use1 ? C1 + ih2 * d1 * d1 : C1;
            const bool better = (ax == 0) || p[ax].cv1 < cmin;
            idc = better ? ax : idc;
            cmin = better ? p[ax].cv1 : cmin;
            hmin = better ? m.dh[ax] : hmin;
        }
        const float discB = B1 * B1 - A1 * C1;
        const bool useA = !(discA < 0.f);
        const bool useB = !useA && !(discB < 0.f);
        a = useA ? A : A1;
        b = useA ? B : B1;
        q = useA ? discA : discB;
        useoned = !useA && !useB;
        hs = hmin * s;
#pragma unroll
        for (int d = 0; d < ND; d++)
            code[d] = useA ? p[d].codeA : (useB ? p[d].codeB : (d == idc ? p[d].codeB : 0));
    }
    const double t = r + (double)((b + sqrtf(q)) / a);
    return useoned ? cmin + (double)hs : t;
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
// (bitwise & and | keep the comparisons branch-free)
template <typename T>
__device__ __forceinline__ bool klt(const Key<T>& a, const Key<T>& b) {
    return (a.K < b.K) | ((a.K == b.K) & ((a.S < b.S) | ((a.S == b.S) & (a.D < b.D))));
}
template <typename T>
__device__ __forceinline__ bool keq(const Key<T>& a, const Key<T>& b) {
    return (a.K == b.K) & (a.S == b.S) & (a.D == b.D);
}
// A cascade can never be deeper than this; a longer chain can only come from two nodes feeding
// each other stale keys during the iteration, and ignoring such a candidate breaks the cycle.
constexpr int kMaxCascade = 16;

// Pop-order comparisons in FP32: a traveltime within kTieUlps ulps below another is treated as
// equal to it (one heap batch: neither node pops before the other, and an update that lands just
// below its trigger is not a cascade).  Near such ties FP32 cannot decide the heap's order anyway,
// and exact comparisons there let two nodes flip each other's choice round after round.  FP64
// compares exactly (the heap's own order).
constexpr float kTieUlps = 4.0f;
__device__ __forceinline__ double tie_below(double x) { return x; }
__device__ __forceinline__ float tie_below(float x) {
    return fabsf(x) < EIK_INFF ? x - kTieUlps * 1.1920929e-7f * fabsf(x) : x;
}
// Whether an update `cur` triggered by a neighbour popping at `c` pops after it as an ordinary node
// (key = its own value).  FP64: only when strictly above; an update landing exactly on its trigger
// pops in the heap's next batch, i.e. right after the trigger's, which the cascade key (c, cur, 1)
// expresses (an ordinary key equal to c would put it in the trigger's own batch).  FP32: within
// the tie margin counts as above.
__device__ __forceinline__ bool pops_after(double cur, double c) { return cur > c; }
__device__ __forceinline__ bool pops_after(float cur, float c) { return cur >= tie_below(c); }

template <int ND, bool MIX>
__device__ __forceinline__ double eval_any(const double* v1, const double* v2, unsigned acc1, unsigned acc2,
                                           Arith<double, MIX> s, const Metric<Arith<double, MIX>>& m, int* code) {
    if constexpr (MIX)
        return eval_mix<ND>(v1, v2, acc1, acc2, s, m, code);
    else
        return eval64<ND>(v1, v2, acc1, acc2, s, m, code);
}
template <int ND, bool MIX>
__device__ __forceinline__ float eval_any(const float* v1, const float* v2, unsigned acc1, unsigned acc2, float s,
                                          const Metric<float>& m, int* code) {
    return eval32<ND>(v1, v2, acc1, acc2, s, m, code);
}

// ---------------------------------------------------------------- replay by axes
// Without init neighbours and cascade keys the candidates are accepted in increasing value order,
// so once an axis enters the stencil its side (the smaller candidate, + side on ties) and its
// 2nd-order choice stay fixed: a later candidate on that axis is never chosen, and the chosen
// side's 2-away node is usable iff it is off the boundary planes and not above the chosen value
// (it was then accepted no later than the chosen node).  The replay's evaluations therefore differ
// only in which axes are in use.  Each axis' contribution to the quadratic is formed once, and the
// replay takes one step per axis instead of one per candidate: a candidate that adds no axis
// re-evaluates the same sums, so the value and the key stay as they are.
//
// One case breaks this (`odd`): the - side wins without a usable 2-away node of its own while the
// + side is a candidate with one; pick_axis then takes the + side's 2nd-order choice once the + side
// is accepted (calcttpt_2ndord! does not reset it).  A wave with such a lane, or with a lane next
// to an init node, runs the per-candidate replay.
template <typename T, typename A = T>
struct AxisTerm {
    T cv1;          // chosen neighbour's value: the axis enters at this candidate (+inf: never)
    A ca, bt, gt;   // contributions to alpha, beta and gamma (eval64's, eval32's or eval_mix's form)
    A d1;           // eval32 / eval_mix: cv1 - r (the all-1st-order fallback's offset)
    int codeA, codeB;
};

template <int ND, typename T, bool MIX = false>
__device__ __forceinline__ bool axis_terms(const T* v1, const T* v2, unsigned cand, unsigned ok2,
                                           const Metric<Arith<T, MIX>>& m, AxisTerm<T, Arith<T, MIX>>* at) {
#pragma clang fp contract(off)
    bool odd = false;
    bool sc[ND];
    T w2[ND];
    T r = tinf<T>();
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        const int kp = 2 * ax, km = 2 * ax + 1;
        const bool cp = (cand >> kp) & 1u, cm = (cand >> km) & 1u;
        const bool okp = (ok2 >> kp) & 1u, okm = (ok2 >> km) & 1u;
        const T vp = cp ? v1[kp] : tinf<T>();
        const bool useM = cm & (v1[km] < vp);
        const T cv1 = useM ? v1[km] : vp;
        w2[ax] = useM ? v2[km] : v2[kp];
        sc[ax] = (useM ? okm : okp) & (w2[ax] <= cv1) & (cv1 < tinf<T>());
        odd |= useM & !sc[ax] & cp & okp & (v2[kp] <= v1[kp]);
        at[ax].cv1 = cv1;
        at[ax].codeA = useM ? (sc[ax] ? 2 : 1) : (cp ? (sc[ax] ? -2 : -1) : 0);
        at[ax].codeB = useM ? 1 : (cp ? -1 : 0);
        r = (cv1 < r) ? cv1 : r;
    }
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        const T cv1 = at[ax].cv1;
        if constexpr (sizeof(T) == 8 && !MIX) {
            const double tmpa2 = 1.0 / 3.0 * (4.0 * cv1 - w2[ax]);
            const double ca = sc[ax] ? m.c2[ax] : m.c1[ax];
            const double tv = sc[ax] ? tmpa2 : cv1;
            at[ax].ca = ca;
            at[ax].bt = -2.0 * ca * tv;
            at[ax].gt = ca * (tv * tv);
            at[ax].d1 = 0.0;
        } else if constexpr (MIX) {
            // eval_mix's terms: offsets from r formed in FP64, rounded once to FP32
            const float al = sc[ax] ? m.c2[ax] : m.c1[ax];
            const float d1 = (float)(cv1 - r);
            const float d = sc[ax] ? d1 + (float)(cv1 - w2[ax]) * (1.0f / 3.0f) : d1;
            at[ax].ca = al;
            at[ax].bt = al * d;
            at[ax].gt = al * d * d;
            at[ax].d1 = d1;
        } else {
            const float al = sc[ax] ? m.c2[ax] : m.c1[ax];
            const float d1 = cv1 - r;
            const float d = sc[ax] ? d1 + (cv1 - w2[ax]) * (1.0f / 3.0f) : d1;
            at[ax].ca = al;
            at[ax].bt = al * d;
            at[ax].gt = al * d * d;
            at[ax].d1 = d1;
        }
    }
    return odd;
}

// eval64 / eval32 over the axes in `on` from their precomputed contributions (same operations in
// the same order, so the same result).  FB: form the fallback for a negative discriminant (the
// caller knows whether any lane needs it); disc_on is the discriminant alone.
template <int ND>
__device__ __forceinline__ double disc_on(const AxisTerm<double, double>* at, unsigned on, double s) {
#pragma clang fp contract(off)
    double alpha = 0.0, beta = 0.0, gamma = -(s * s);
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        const bool u = (on >> ax) & 1u;
        alpha = u ? alpha + at[ax].ca : alpha;
        beta = u ? beta + at[ax].bt : beta;
        gamma = u ? gamma + at[ax].gt : gamma;
    }
    return beta * beta - 4.0 * alpha * gamma;
}
template <int ND, bool FB>
__device__ __forceinline__ double eval_on(const AxisTerm<double, double>* at, unsigned on, double s,
                                          const Metric<double>& m, int* code) {
#pragma clang fp contract(off)
    double alpha = 0.0, beta = 0.0, gamma = -(s * s);
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        const bool u = (on >> ax) & 1u;
        alpha = u ? alpha + at[ax].ca : alpha;
        beta = u ? beta + at[ax].bt : beta;
        gamma = u ? gamma + at[ax].gt : gamma;
        code[ax] = u ? at[ax].codeA : 0;
    }
    const double sqA = beta * beta - 4.0 * alpha * gamma;
    double a = alpha, b = beta, q = sqA, oned = 0.0;
    bool useoned = false;
    if (FB) {
        double alpha1 = 0.0, beta1 = 0.0, gamma1 = -(s * s);
        double cmin = EIK_INF, hmin = m.dh[0];
        int idc = 0;
#pragma unroll
        for (int ax = 0; ax < ND; ax++) {
            const bool u = (on >> ax) & 1u;
            const double cv1 = u ? at[ax].cv1 : EIK_INF;
            const double cb = m.c1[ax];
            alpha1 = u ? alpha1 + cb : alpha1;
            beta1 = u ? beta1 + (-2.0 * cb * cv1) : beta1;
            gamma1 = u ? gamma1 + cb * (cv1 * cv1) : gamma1;
            const bool better = (ax == 0) || cv1 < cmin;
            idc = better ? ax : idc;
            cmin = better ? cv1 : cmin;
            hmin = better ? m.dh[ax] : hmin;
        }
        const double sqB = beta1 * beta1 - 4.0 * alpha1 * gamma1;
        const bool useA = !(sqA < 0.0);
        const bool useB = !useA && !(sqB < 0.0);
        a = useA ? alpha : alpha1;
        b = useA ? beta : beta1;
        q = useA ? sqA : sqB;
        useoned = !useA && !useB;
        oned = (hmin * s) + cmin;
#pragma unroll
        for (int d = 0; d < ND; d++) {
            const int cB = ((on >> d) & 1u) ? at[d].codeB : 0;
            code[d] = useA ? code[d] : (useB ? cB : (d == idc ? cB : 0));
        }
    }
    const double t = (-b + sqrt(q)) / (2.0 * a);
    return useoned ? oned : t;
}
template <int ND>
__device__ __forceinline__ float disc_on(const AxisTerm<float, float>* at, unsigned on, float s) {
    float A = 0.f, B = 0.f, C = -(s * s);
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        const bool u = (on >> ax) & 1u;
        A = u ? A + at[ax].ca : A;
        B = u ? B + at[ax].bt : B;
        C = u ? C + at[ax].gt : C;
    }
    return B * B - A * C;
}
template <int ND, bool FB>
__device__ __forceinline__ float eval_on(const AxisTerm<float, float>* at, unsigned on, float s, const Metric<float>& m,
                                         int* code) {
    float r = EIK_INFF;
#pragma unroll
    for (int ax = 0; ax < ND; ax++) r = ((on >> ax) & 1u) ? fminf(r, at[ax].cv1) : r;
    float A = 0.f, B = 0.f, C = -(s * s);
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        const bool u = (on >> ax) & 1u;
        A = u ? A + at[ax].ca : A;
        B = u ? B + at[ax].bt : B;
        C = u ? C + at[ax].gt : C;
        code[ax] = u ? at[ax].codeA : 0;
    }
    const float discA = B * B - A * C;
    float a = A, b = B, q = discA, oned = 0.f;
    bool useoned = false;
    if (FB) {
        float A1 = 0.f, B1 = 0.f, C1 = -(s * s);
        float cmin = EIK_INFF, hmin = m.dh[0];
        int idc = 0;
#pragma unroll
        for (int ax = 0; ax < ND; ax++) {
            const bool u = (on >> ax) & 1u;
            const float cv1 = u ? at[ax].cv1 : EIK_INFF;
            const float ih2 = m.c1[ax];
            const float d1 = cv1 - r;
            A1 = u ? A1 + ih2 : A1;
            B1 = u ? B1 + ih2 * d1 : B1;
            C1 = u ? C1 + ih2 * d1 * d1 : C1;
            const bool better = (ax == 0) || cv1 < cmin;
            idc = better ? ax : idc;
            cmin = better ? cv1 : cmin;
            hmin = better ? m.dh[ax] : hmin;
        }
        const float discB = B1 * B1 - A1 * C1;
        const bool useA = !(discA < 0.f);
        const bool useB = !useA && !(discB < 0.f);
        a = useA ? A : A1;
        b = useA ? B : B1;
        q = useA ? discA : discB;
        useoned = !useA && !useB;
        oned = cmin + hmin * s;
#pragma unroll
        for (int d = 0; d < ND; d++) {
            const int cB = ((on >> d) & 1u) ? at[d].codeB : 0;
            code[d] = useA ? code[d] : (useB ? cB : (d == idc ? cB : 0));
        }
    }
    const float t = r + (b + sqrtf(q)) / a;
    return useoned ? oned : t;
}

// eval_mix over the axes in `on` from their precomputed contributions (the first axis in use is
// the smallest candidate, so r is the same shift axis_terms used).
template <int ND>
__device__ __forceinline__ float disc_on(const AxisTerm<double, float>* at, unsigned on, float s) {
    float A = 0.f, B = 0.f, C = -(s * s);
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        const bool u = (on >> ax) & 1u;
        A = u ? A + at[ax].ca : A;
        B = u ? B + at[ax].bt : B;
        C = u ? C + at[ax].gt : C;
    }
    return B * B - A * C;
}
template <int ND, bool FB>
__device__ __forceinline__ double eval_on(const AxisTerm<double, float>* at, unsigned on, float s,
                                          const Metric<float>& m, int* code) {
#pragma clang fp contract(off)
    double r = EIK_INF;
#pragma unroll
    for (int ax = 0; ax < ND; ax++) r = (((on >> ax) & 1u) && at[ax].cv1 < r) ? at[ax].cv1 : r;
    float A = 0.f, B = 0.f, C = -(s * s);
#pragma unroll
    for (int ax = 0; ax < ND; ax++) {
        const bool u = (on >> ax) & 1u;
        A = u ? A + at[ax].ca : A;
        B = u ? B + at[ax].bt : B;
        C = u ? C + at[ax].gt : C;
        code[ax] = u ? at[ax].codeA : 0;
    }
    const float discA = B * B - A * C;
    float a = A, b = B, q = discA, hs = 0.f;
    double cmin = EIK_INF;
    bool useoned = false;
    if (FB) {
        float A1 = 0.f, B1 = 0.f, C1 = -(s * s);
        float hmin = m.dh[0];
        int idc = 0;
#pragma unroll
        for (int ax = 0; ax < ND; ax++) {
            const bool u = (on >> ax) & 1u;
            const double cv1 = u ? at[ax].cv1 : EIK_INF;
            const float ih2 = m.c1[ax];
            const float d1 = at[ax].d1;
            A1 = u ? A1 + ih2 : A1;
            B1 = u ? B1 + ih2 * d1 : B1;
            C1 = u ? C1 + ih2 * d1 * d1 : C1;
            const bool better = (ax == 0) || cv1 < cmin;
            idc = better ? ax : idc;
            cmin = better ? cv1 : cmin;
            hmin = better ? m.dh[ax] : hmin;
        }
        const float discB = B1 * B1 - A1 * C1;
        const bool useA = !(discA < 0.f);
        const bool useB = !useA && !(discB < 0.f);
        a = useA ? A : A1;
        b = useA ? B : B1;
        q = useA ? discA : discB;
        useoned = !useA && !useB;
        hs = hmin * s;
#pragma unroll
        for (int d = 0; d < ND; d++) {
            const int cB = ((on >> d) & 1u) ? at[d].codeB : 0;
            code[d] = useA ? code[d] : (useB ? cB : (d == idc ? cB : 0));
        }
    }
    const double t = r + (double)((b + sqrtf(q)) / a);
    return useoned ? cmin + (double)hs : t;
}

// replay_plain's sequence, one step per axis (see above).  The steps' axis sets follow from the
// sorted candidate values, so every step is evaluated at once, the evaluations independent of each
// other (one wave's long square-root and divide chains interleave), and the heap's sequence then
// keeps the last step it reaches; a step beyond it is evaluated and dropped.  The fallback of a
// negative discriminant is formed for every step when some lane needs it for any.
template <int ND, typename T, bool MIX = false>
__device__ __forceinline__ T replay_axes(const AxisTerm<T, Arith<T, MIX>>* at, Arith<T, MIX> s,
                                         const Metric<Arith<T, MIX>>& m, int* code, Key<T>* kout) {
    T c[ND];          // distinct candidate values, increasing (+inf: none)
    unsigned on[ND];  // the axes in use once c[k] has popped
    T prev = -tinf<T>();
#pragma unroll
    for (int k = 0; k < ND; k++) {
        T ck = tinf<T>();
#pragma unroll
        for (int ax = 0; ax < ND; ax++) ck = (at[ax].cv1 > prev && at[ax].cv1 < ck) ? at[ax].cv1 : ck;
        unsigned o = 0u;
#pragma unroll
        for (int ax = 0; ax < ND; ax++) o |= (unsigned)(at[ax].cv1 <= ck) << ax;
        c[k] = ck;
        on[k] = o;
        prev = ck;
    }
    bool neg = false;
#pragma unroll
    for (int k = 0; k < ND; k++) neg = neg | ((c[k] < tinf<T>()) & (disc_on<ND>(at, on[k], s) < 0));
    T e[ND];
    int cd[ND][3];
    if (__ballot(neg) != 0ull) {
#pragma unroll
        for (int k = 0; k < ND; k++) e[k] = eval_on<ND, true>(at, on[k], s, m, cd[k]);
    } else {
#pragma unroll
        for (int k = 0; k < ND; k++) e[k] = eval_on<ND, false>(at, on[k], s, m, cd[k]);
    }
    T cur = tinf<T>();
    T kzK = tinf<T>(), kzS = -tinf<T>();
    int kzD = 0;
#pragma unroll
    for (int d = 0; d < ND; d++) code[d] = 0;
    bool go = true;
#pragma unroll
    for (int k = 0; k < ND; k++) {
        go = go && (c[k] < tinf<T>()) && kzD == 0 && (c[k] < tie_below(kzK));
        if (go) {
            cur = e[k];
            const bool ord = pops_after(cur, c[k]);
            kzK = ord ? cur : c[k];
            kzS = ord ? -tinf<T>() : cur;
            kzD = ord ? 0 : 1;
#pragma unroll
            for (int d = 0; d < ND; d++) code[d] = cd[k][d];
        }
    }
    *kout = Key<T>{kzK, kzS, kzD};
    return cur;
}

// Local replay of the heap (see the file comment).  v1/v2: adjacent / 2-away values; k1/k2 their
// keys; f1/f2 fixed (init) bits; ok1/ok2 grid-boundary bits.  Returns the value and sets *kout to
// the node's pop key and code to the derivative codes of the last stencil evaluation.
//
// Outside cascades every key is (value, -inf, 0) and the order is the value order, so when no
// neighbour of any lane in the wave carries a cascade key the replay compares values only
// (replay_plain); otherwise the full key comparison runs (replay_keys).  Both give the same
// result on such neighbourhoods.
template <int ND, typename T, bool MIX = false>
__device__ __forceinline__ T replay_plain(const T* v1, const T* v2, unsigned cand, bool haveinit, unsigned f1,
                                          unsigned f2, unsigned ok1, unsigned ok2, Arith<T, MIX> s,
                                          const Metric<Arith<T, MIX>>& m, int* code, Key<T>* kout) {
    T cur = tinf<T>();
    T kzK = tinf<T>(), kzS = -tinf<T>();
    int kzD = 0;
#pragma unroll
    for (int d = 0; d < ND; d++) code[d] = 0;
    if (haveinit) {
        cur = eval_any<ND, MIX>(v1, v2, f1 & ok1, f2 & ok2, s, m, code);
        kzK = cur;
    }
    // distinct candidate values in increasing order; after a step that ends below its trigger
    // (cascade key), the next candidate never pops first
    T prev = -tinf<T>();
    unsigned acc1 = f1 & ok1;
#pragma unroll 1
    for (int step = 0; step < 2 * ND && cand; step++) {
        T c = tinf<T>();
#pragma unroll
        for (int k = 0; k < 2 * ND; k++) c = (((cand >> k) & 1u) && v1[k] > prev && v1[k] < c) ? v1[k] : c;
        if (!(c < tinf<T>()) || kzD != 0 || !(c < tie_below(kzK))) break;
        prev = c;
        unsigned acc2 = 0u;
#pragma unroll
        for (int k = 0; k < 2 * ND; k++) {
            acc1 |= (unsigned)(((cand >> k) & 1u) && v1[k] == c) << k;
            const bool a2 = (bool)((ok2 >> k) & 1u) & ((bool)((f2 >> k) & 1u) | (v2[k] <= c));
            acc2 |= (unsigned)a2 << k;
        }
        cur = eval_any<ND, MIX>(v1, v2, acc1, acc2, s, m, code);
        const bool ord = pops_after(cur, c);
        kzK = ord ? cur : c;
        kzS = ord ? -tinf<T>() : cur;
        kzD = ord ? 0 : 1;
    }
    *kout = Key<T>{kzK, kzS, kzD};
    return cur;
}

// keys(k, two) gives the pop key of slot k's adjacent (two = false) or 2-away (two = true)
// node; it is called only on the rare cascade path, so callers can read keys from memory there
// instead of holding them in registers.  d1/d2: slots whose adjacent / 2-away node has a cascade
// key (D != 0); dcap: adjacent nodes at the cascade depth cap.
template <int ND, typename T, bool MIX, class KeyAt>
__device__ __forceinline__ T replay(const T* v1, const T* v2, KeyAt keys, unsigned d1, unsigned d2, unsigned dcap,
                                    unsigned f1, unsigned f2, unsigned ok1, unsigned ok2, Arith<T, MIX> s,
                                    const Metric<Arith<T, MIX>>& m, int* code, Key<T>* kout) {
    // candidates: adjacent nodes inside the grid, not init, reached, cascade depth below the cap
    unsigned cand = 0u;
    bool haveinit = false;
    bool keyed = false;  // a neighbour in use carries a cascade key
#pragma unroll
    for (int k = 0; k < 2 * ND; k++) {
        const bool in = (ok1 >> k) & 1u, fx = (f1 >> k) & 1u, cp = (dcap >> k) & 1u;
        haveinit |= in & fx;
        const bool reached1 = in & !fx & (v1[k] < tinf<T>());
        cand |= (unsigned)(reached1 & !cp) << k;
        const bool reached2 = (bool)((ok2 >> k) & 1u) & !((f2 >> k) & 1u) & (v2[k] < tinf<T>());
        keyed |= (reached1 & (bool)((d1 >> k) & 1u)) | (reached2 & (bool)((d2 >> k) & 1u));
    }
    if (__ballot(keyed) == 0ull) {
        if (__ballot(haveinit) == 0ull) {
            AxisTerm<T, Arith<T, MIX>> at[ND];
            const bool odd = axis_terms<ND, T, MIX>(v1, v2, cand, ok2, m, at);
            if (__ballot(odd) == 0ull) return replay_axes<ND, T, MIX>(at, s, m, code, kout);
        }
        return replay_plain<ND, T, MIX>(v1, v2, cand, haveinit, f1, f2, ok1, ok2, s, m, code, kout);
    }
    T cur = tinf<T>();
    Key<T> kz{tinf<T>(), -tinf<T>(), 0};
#pragma unroll
    for (int d = 0; d < ND; d++) code[d] = 0;
    if (haveinit) {
        cur = eval_any<ND, MIX>(v1, v2, f1 & ok1, f2 & ok2, s, m, code);
        kz = Key<T>{cur, -tinf<T>(), 0};
    }
    // distinct candidate keys in increasing order (equal keys pop as one batch); each step scans
    // the 2*ND slots with static indices
    Key<T> prev{-tinf<T>(), -tinf<T>(), -1};
#pragma unroll 1
    for (int step = 0; step < 2 * ND && cand; step++) {
        Key<T> c{tinf<T>(), tinf<T>(), 1 << 30};
        bool found = false;
#pragma unroll
        for (int k = 0; k < 2 * ND; k++) {
            const Key<T> kk = keys(k, false);
            const bool take = (bool)((cand >> k) & 1u) & klt(prev, kk) & (!found | klt(kk, c));
            c.K = take ? kk.K : c.K;
            c.S = take ? kk.S : c.S;
            c.D = take ? kk.D : c.D;
            found |= take;
        }
        if (!found || !((c.K < tie_below(kz.K)) | ((c.K == kz.K) & klt(c, kz)))) break;
        prev = c;
        unsigned acc1 = 0u, acc2 = 0u;
#pragma unroll
        for (int k = 0; k < 2 * ND; k++) {
            const bool a1 = (bool)((ok1 >> k) & 1u) &
                            ((bool)((f1 >> k) & 1u) | ((v1[k] < tinf<T>()) & !klt(c, keys(k, false))));
            const bool a2 = (bool)((ok2 >> k) & 1u) &
                            ((bool)((f2 >> k) & 1u) | ((v2[k] < tinf<T>()) & !klt(c, keys(k, true))));
            acc1 |= (unsigned)a1 << k;
            acc2 |= (unsigned)a2 << k;
        }
        cur = eval_any<ND, MIX>(v1, v2, acc1, acc2, s, m, code);
        const bool ord = pops_after(cur, c.K);
        kz.K = ord ? cur : c.K;
        kz.S = ord ? -tinf<T>() : (cur > c.S ? cur : c.S);
        kz.D = ord ? 0 : c.D + 1;
    }
    *kout = kz;
    return cur;
}

// Same with the keys held in arrays (k1: adjacent, k2: 2-away).
template <int ND, typename T, bool MIX = false>
__device__ __forceinline__ T replay(const T* v1, const T* v2, const Key<T>* k1, const Key<T>* k2, unsigned f1,
                                    unsigned f2, unsigned ok1, unsigned ok2, Arith<T, MIX> s,
                                    const Metric<Arith<T, MIX>>& m, int* code, Key<T>* kout) {
    unsigned d1 = 0u, d2 = 0u, dcap = 0u;
#pragma unroll
    for (int k = 0; k < 2 * ND; k++) {
        d1 |= (unsigned)(k1[k].D != 0) << k;
        d2 |= (unsigned)(k2[k].D != 0) << k;
        dcap |= (unsigned)(k1[k].D >= kMaxCascade) << k;
    }
    auto keys = [&](int k, bool two) { return two ? k2[k] : k1[k]; };
    return replay<ND, T, MIX>(v1, v2, keys, d1, d2, dcap, f1, f2, ok1, ok2, s, m, code, kout);
}

}  // namespace eik
