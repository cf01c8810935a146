// Host check of the stencil header: the axis replay (axis_terms + replay_axes) returns the same
// value, pop key and derivative codes, bit for bit, as the per-candidate replay (replay_plain) on
// random neighbourhoods: discrete values with many ties, smooth fronts, unreached and off-grid
// neighbours, fixed 2-away nodes, Cartesian and per-axis spacings, FP64, FP32 and the mixed
// arithmetic (FP64 values, FP32 increment).  Neighbourhoods
// with an "odd" axis (which the device code routes to replay_plain) are counted, not compared,
// unless the first argument is "all" (then they must differ somewhere: the check is sensitive).
#include "eik_stencil.h"
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <random>
using namespace eik;
template <typename T> static bool same(T a, T b) { return memcmp(&a, &b, sizeof(T)) == 0; }
static bool g_all = false;
template <int ND, typename T, bool MIX = false>
static long run(std::mt19937_64& rng, long n, long& nodd, long& nused, bool sph) {
    using A = Arith<T, MIX>;
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long bad = 0;
    for (long it = 0; it < n; it++) {
        T v1[6], v2[6];
        unsigned ok1 = 0, ok2 = 0, f2 = 0, cand = 0;
        const double t0 = 1.0 + 20.0 * U(rng);
        const double step = (it & 1) ? 0.37 : 0.05 + U(rng);
        const int mode = (int)(U(rng) * 3);
        for (int k = 0; k < 2 * ND; k++) {
            const bool in1 = U(rng) < 0.95, in2 = in1 && U(rng) < 0.9;
            ok1 |= (unsigned)in1 << k;
            ok2 |= (unsigned)in2 << k;
            double a, b;
            if (mode == 0) {  // discrete values: ties
                a = t0 + step * (int)(U(rng) * 7 - 3);
                b = t0 + step * (int)(U(rng) * 9 - 4);
            } else {          // smooth front along a random direction
                const double dir = (k & 1) ? -1.0 : 1.0;
                const double g = (U(rng) - 0.5) * 1.2;
                a = t0 + dir * g + (mode == 2 ? (U(rng) - 0.5) * 0.02 : 0.0);
                b = t0 + 2.0 * dir * g + (mode == 2 ? (U(rng) - 0.5) * 0.05 : 0.0);
            }
            if (U(rng) < 0.1) a = INFINITY;
            if (U(rng) < 0.1) b = INFINITY;
            if (U(rng) < 0.03) f2 |= 1u << k;
            v1[k] = (T)a;
            v2[k] = (T)b;
            if (!in1) v1[k] = (T)INFINITY;
            if (!in2) v2[k] = (T)INFINITY;
            if (in1 && v1[k] < (T)INFINITY) cand |= 1u << k;
        }
        Metric<A> m;
        for (int ax = 0; ax < 3; ax++) m.dh[ax] = sph ? (A)(0.5 + U(rng)) : (A)1.0;
        if (!sph) { A h = (A)(0.25 + U(rng)); for (int ax = 0; ax < 3; ax++) m.dh[ax] = h; }
        metric_coefs(m, ND);
        const A s = (A)(0.15 + 0.6 * U(rng));
        int c0[3] = {9, 9, 9}, c1[3] = {9, 9, 9};
        Key<T> k0, k1;
        const T r0 = replay_plain<ND, T, MIX>(v1, v2, cand, false, 0u, f2, ok1, ok2, s, m, c0, &k0);
        AxisTerm<T, A> at[ND];
        const bool odd = axis_terms<ND, T, MIX>(v1, v2, cand, ok2, m, at);
        if (odd) { nodd++; if (!g_all) continue; }
        nused++;
        const T r1 = replay_axes<ND, T, MIX>(at, s, m, c1, &k1);
        bool ok = same(r0, r1) && same(k0.K, k1.K) && same(k0.S, k1.S) && k0.D == k1.D;
        for (int d = 0; d < ND; d++) ok = ok && c0[d] == c1[d];
        if (!ok) {
            if (bad < 5 && !g_all) {
                printf("MISMATCH nd=%d sz=%d: %.17g vs %.17g  K %.17g/%.17g D %d/%d codes", ND, (int)sizeof(T),
                       (double)r0, (double)r1, (double)k0.K, (double)k1.K, k0.D, k1.D);
                for (int d = 0; d < ND; d++) printf(" %d/%d", c0[d], c1[d]);
                printf("\n");
            }
            bad++;
        }
    }
    return bad;
}
int main(int argc, char** argv) {
    g_all = argc > 1 && strcmp(argv[1], "all") == 0;
    long n = argc > 2 ? atol(argv[2]) : 200000;
    std::mt19937_64 rng(12345);
    long tot = 0;
    for (int sph = 0; sph < 2; sph++) {
        long o, u;
        o = u = 0; tot += run<2, double>(rng, n, o, u, sph); printf("2D f64 sph=%d used %ld odd %ld\n", sph, u, o);
        o = u = 0; tot += run<3, double>(rng, n, o, u, sph); printf("3D f64 sph=%d used %ld odd %ld\n", sph, u, o);
        o = u = 0; tot += run<2, float>(rng, n, o, u, sph); printf("2D f32 sph=%d used %ld odd %ld\n", sph, u, o);
        o = u = 0; tot += run<3, float>(rng, n, o, u, sph); printf("3D f32 sph=%d used %ld odd %ld\n", sph, u, o);
        o = u = 0; tot += run<2, double, true>(rng, n, o, u, sph); printf("2D mix sph=%d used %ld odd %ld\n", sph, u, o);
        o = u = 0; tot += run<3, double, true>(rng, n, o, u, sph); printf("3D mix sph=%d used %ld odd %ld\n", sph, u, o);
    }
    printf("mismatches: %ld\n", tot);
    return g_all ? (tot == 0) : (tot != 0);
}
