// Host check that the device stencil's mixed arithmetic (eik_stencil.h: eval_mix, and the axis
// replay's eval_on on the same neighbourhood) equals the oracle's mixed stencil (og_stencil_mixed)
// bit for bit: value and derivative codes, on random neighbourhoods of small Cartesian, 2D-sphere
// and 3D-sphere grids (ties, smooth fronts, unreached and boundary neighbours, nodes next to the
// grid faces, where the reference's isonb2nd rule withholds the 2nd-order stencil).  With the
// argument "fp64" it compares against the FP64 stencil instead, which must differ somewhere.
#include "eik_stencil.h"
extern "C" {
#include "eik_oracle.h"
}
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
using namespace eik;

static bool same(double a, double b) { return memcmp(&a, &b, sizeof(double)) == 0; }
static bool g_vs_fp64 = false;  // compare with the FP64 stencil instead (must differ: sensitivity)

template <int ND>
static long run(std::mt19937_64& rng, long n, int kind, long& used) {
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long bad = 0;
    const int E = 7;  // 7 nodes per axis
    og_grid g{};
    g.kind = kind;
    g.ndim = ND;
    for (int d = 0; d < 3; d++) g.n[d] = d < ND ? E : 1;
    const int64_t N = ND == 2 ? E * E : E * E * E;
    std::vector<double> tt(N), vel(N);
    std::vector<uint8_t> st(N);
    for (long it = 0; it < n; it++) {
        if (kind == OG_CART2D || kind == OG_CART3D) {
            const double h = 0.25 + U(rng);
            g.d[0] = g.d[1] = g.d[2] = h;
            g.o[0] = g.o[1] = g.o[2] = 0.0;
        } else {
            g.d[0] = 0.5 + 2.0 * U(rng);        // dr
            g.d[1] = 0.05 + 0.2 * U(rng);       // dtheta, degrees
            g.d[2] = 0.05 + 0.2 * U(rng);       // dphi, degrees
            g.o[0] = 6000.0 + 300.0 * U(rng);
            g.o[1] = 20.0 + 100.0 * U(rng);
            g.o[2] = 10.0 * U(rng);
        }
        // the node: interior, or next to a face (then the reference withholds 2nd order there)
        int64_t ijk[3] = {0, 0, 0};
        for (int d = 0; d < ND; d++) ijk[d] = U(rng) < 0.2 ? (U(rng) < 0.5 ? 1 : E - 2) : 2 + (int)(U(rng) * 3);
        const int mode = (int)(U(rng) * 3);
        const double t0 = 1.0 + 20.0 * U(rng), step = 0.05 + U(rng);
        double gr[3];
        for (int d = 0; d < 3; d++) gr[d] = (U(rng) - 0.5) * 1.2;
        for (int64_t l = 0; l < N; l++) {
            const int64_t q[3] = {l % E, (l / E) % E, ND == 3 ? l / (E * E) : 0};
            double a;
            if (mode == 0) {
                a = t0 + step * (int)(U(rng) * 7 - 3);  // discrete values: ties
            } else {
                a = t0;
                for (int d = 0; d < ND; d++) a += gr[d] * (double)(q[d] - ijk[d]);
                if (mode == 2) a += (U(rng) - 0.5) * 0.05;
            }
            tt[l] = a > 0.0 ? a : 0.0;
            st[l] = U(rng) < 0.8 ? 2 : (U(rng) < 0.5 ? 1 : 0);
            vel[l] = 1.5 + 4.5 * U(rng);
        }
        const int64_t c = ND == 2 ? ijk[0] + E * ijk[1] : ijk[0] + E * (ijk[1] + E * ijk[2]);
        st[c] = 1;
        int32_t oc[3] = {9, 9, 9};
        const double ro = g_vs_fp64 ? og_stencil(&g, vel.data(), tt.data(), st.data(), ijk, oc)
                                    : og_stencil_mixed(&g, vel.data(), tt.data(), st.data(), ijk, oc);
        // the device's inputs: adjacent / 2-away values and the accepted / usable masks
        double v1[6], v2[6];
        unsigned acc1 = 0, acc2 = 0;
        for (int k = 0; k < 2 * ND; k++) {
            const int ax = k >> 1, sh = (k & 1) ? -1 : 1;
            int64_t a[3] = {ijk[0], ijk[1], ijk[2]}, b[3] = {ijk[0], ijk[1], ijk[2]};
            a[ax] += sh;
            b[ax] += 2 * sh;
            const bool in1 = a[ax] >= 0 && a[ax] <= E - 1;
            bool in2 = in1;
            for (int d = 0; d < ND; d++) in2 = in2 && a[d] >= 1 && a[d] <= E - 2;
            const int64_t la = ND == 2 ? a[0] + E * a[1] : a[0] + E * (a[1] + E * a[2]);
            const int64_t lb = ND == 2 ? b[0] + E * b[1] : b[0] + E * (b[1] + E * b[2]);
            v1[k] = in1 ? tt[la] : INFINITY;
            v2[k] = in2 ? tt[lb] : INFINITY;
            acc1 |= (unsigned)(in1 && st[la] == 2) << k;
            acc2 |= (unsigned)(in2 && st[lb] == 2) << k;
        }
        Metric<float> m;
        if (kind == OG_CART2D || kind == OG_CART3D) {
            for (int d = 0; d < 3; d++) m.dh[d] = (float)g.d[0];
        } else {
            const double r = fma((double)ijk[0], g.d[0], g.o[0]);
            m.dh[0] = (float)g.d[0];
            m.dh[1] = (float)(r * (g.d[1] * (M_PI / 180.0)));
            const double th = fma((double)ijk[1], g.d[1], g.o[1]);
            m.dh[2] = ND == 3 ? (float)(r * sin(th * (M_PI / 180.0)) * (g.d[2] * (M_PI / 180.0))) : 0.f;
        }
        metric_coefs(m, ND);
        const float s = (float)(1.0 / vel[c]);
        int dc[3] = {9, 9, 9};
        const double rd = eval_mix<ND>(v1, v2, acc1, acc2, s, m, dc);
        if (!(acc1 & ((1u << (2 * ND)) - 1u))) continue;  // no accepted neighbour: not a stencil call
        used++;
        bool ok = same(ro, rd);
        for (int d = 0; d < ND; d++) ok = ok && oc[d] == dc[d];
        if (!ok) {
            if (bad < 5 && !g_vs_fp64)
                printf("MISMATCH nd=%d kind=%d: oracle %.17g device %.17g codes %d %d %d / %d %d %d\n", ND, kind, ro,
                       rd, oc[0], oc[1], oc[2], dc[0], dc[1], dc[2]);
            bad++;
        }
    }
    return bad;
}

int main(int argc, char** argv) {
    const long n = argc > 1 ? atol(argv[1]) : 100000;
    g_vs_fp64 = argc > 2 && strcmp(argv[2], "fp64") == 0;
    std::mt19937_64 rng(2024);
    long tot = 0, u;
    u = 0; tot += run<2>(rng, n, OG_CART2D, u); printf("cart2d used %ld\n", u);
    u = 0; tot += run<3>(rng, n, OG_CART3D, u); printf("cart3d used %ld\n", u);
    u = 0; tot += run<2>(rng, n, OG_SPH2D, u); printf("sph2d used %ld\n", u);
    u = 0; tot += run<3>(rng, n, OG_SPH3D, u); printf("sph3d used %ld\n", u);
    printf("mismatches: %ld\n", tot);
    return g_vs_fp64 ? (tot == 0) : (tot != 0);
}
