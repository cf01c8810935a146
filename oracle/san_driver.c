/*
 * san_driver.c -- runs every entry point of the CPU oracle under AddressSanitizer and
 * UndefinedBehaviorSanitizer (TEST INFRASTRUCTURE ONLY; built by `make -C oracle sanitize`).
 *
 * Small Cartesian 2D/3D and spherical 2D/3D cases, with and without refinement, every whichgrad,
 * the Gaussian smoothing, the single-source recorder, the brute-force FMM and the bare stencil.
 * Each call must succeed and return finite values; the misfit of eikgradient must equal
 * eikttimemisfit's (eikmaingrad.jl:237 and :1325 compute the same sum).  Any sanitizer report
 * aborts the process with a non-zero status.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "eik_oracle.h"

static int failures = 0;

#define CHECK(cond, ...)                          \
    do {                                          \
        if (!(cond)) {                            \
            fprintf(stderr, "FAIL: " __VA_ARGS__); \
            fprintf(stderr, "\n");                \
            failures++;                           \
        }                                         \
    } while (0)

static int all_finite(const double* a, int64_t n) {
    for (int64_t i = 0; i < n; i++)
        if (!isfinite(a[i])) return 0;
    return 1;
}

/* velocity: smooth positive field, index 1 fastest */
static double* make_vel(const og_grid* g) {
    const int64_t N = g->n[0] * g->n[1] * g->n[2];
    double* v = malloc((size_t)N * sizeof(double));
    for (int64_t l = 0; l < N; l++) {
        const int64_t i = l % g->n[0], j = (l / g->n[0]) % g->n[1], k = l / (g->n[0] * g->n[1]);
        v[l] = 2.0 + 0.03 * (double)j + 0.2 * sin(0.3 * (double)i) + 0.1 * cos(0.2 * (double)k);
    }
    return v;
}

static void run_case(const char* name, const og_grid* g, const double* src, int nsrc, const double* rec, int nrecps,
                     int refine, int smooth, int grads) {
    const int nd = g->ndim;
    const int64_t N = g->n[0] * g->n[1] * g->n[2];
    double* vel = make_vel(g);
    int64_t* off = malloc((size_t)(nsrc + 1) * sizeof(int64_t));
    double* coordrec = malloc((size_t)nsrc * nrecps * nd * sizeof(double));
    for (int s = 0; s <= nsrc; s++) off[s] = (int64_t)s * nrecps;
    for (int s = 0; s < nsrc; s++) memcpy(coordrec + (size_t)s * nrecps * nd, rec, (size_t)nrecps * nd * sizeof(double));
    const int64_t nrec = off[nsrc];
    og_params p = {refine, 5, 3, 3, smooth, 2};
    double* picks = malloc((size_t)nrec * sizeof(double));
    double* tt = malloc((size_t)nsrc * N * sizeof(double));
    int rc = og_traveltime(g, vel, nsrc, src, off, coordrec, &p, picks, tt);
    CHECK(rc == 0, "%s: og_traveltime rc %d (%s)", name, rc, og_last_error());
    CHECK(all_finite(picks, nrec) && all_finite(tt, nsrc * N), "%s: non-finite traveltimes", name);
    double* obs = malloc((size_t)nrec * sizeof(double));
    double* sd = malloc((size_t)nrec * sizeof(double));
    for (int64_t r = 0; r < nrec; r++) {
        obs[r] = picks[r] * 1.02 + 0.01;
        sd[r] = 0.05;
    }
    double mref = 0.0;
    rc = og_misfit(g, vel, nsrc, src, off, coordrec, obs, sd, &p, &mref);
    CHECK(rc == 0 && isfinite(mref), "%s: og_misfit rc %d", name, rc);
    for (int wg = OG_GRADVEL; wg <= OG_GRADBOTH; wg++) {
        if (!((grads >> (wg - 1)) & 1)) continue;
        double* gv = malloc((size_t)N * sizeof(double));
        double* gs = malloc((size_t)nsrc * nd * sizeof(double));
        double m = 0.0;
        rc = og_gradient(g, vel, nsrc, src, off, coordrec, obs, sd, wg, &p, wg != OG_GRADSRCLOC ? gv : NULL,
                         wg != OG_GRADVEL ? gs : NULL, &m);
        CHECK(rc == 0, "%s: og_gradient(%d) rc %d (%s)", name, wg, rc, og_last_error());
        if (wg != OG_GRADSRCLOC) CHECK(all_finite(gv, N), "%s: non-finite gradvel", name);
        if (wg != OG_GRADVEL) CHECK(all_finite(gs, (int64_t)nsrc * nd), "%s: non-finite gradsrc", name);
        CHECK(fabs(m - mref) <= 1e-12 * fabs(mref), "%s: misfit %.17g vs %.17g", name, m, mref);
        free(gv);
        free(gs);
    }
    /* single-source recorder and, without refinement on Cartesian grids, the brute-force FMM */
    double* t1 = malloc((size_t)N * sizeof(double));
    int64_t* order = malloc((size_t)N * sizeof(int64_t));
    int8_t* codes = malloc((size_t)N * nd);
    int64_t nsp = 0;
    const int64_t nacc = og_fmm_record(g, vel, src, &p, t1, order, codes, &nsp);
    CHECK(nacc > 0 && nacc <= N, "%s: og_fmm_record accepted %lld", name, (long long)nacc);
    if (!refine && g->kind <= OG_CART3D && N <= 4096) {
        double* t2 = malloc((size_t)N * sizeof(double));
        int64_t* o2 = malloc((size_t)N * sizeof(int64_t));
        rc = og_fmm_bruteforce(g, vel, src, t2, o2);
        CHECK(rc == 0, "%s: og_fmm_bruteforce rc %d", name, rc);
        for (int64_t l = 0; l < N; l++) CHECK(t1[l] == t2[l], "%s: brute force differs at %lld", name, (long long)l);
        free(t2);
        free(o2);
    }
    /* the bare stencil on the converged state, at an interior node */
    uint8_t* st = malloc((size_t)N);
    memset(st, 2, (size_t)N);
    int64_t ijk[3] = {g->n[0] / 2, g->n[1] / 2, nd == 3 ? g->n[2] / 2 : 0};
    int32_t cd[3] = {0, 0, 0};
    const double ts = og_stencil(g, vel, t1, st, ijk, cd);
    CHECK(isfinite(ts), "%s: og_stencil non-finite", name);
    /* smoothing with a few odd kernels */
    double* sm = malloc((size_t)N * sizeof(double));
    for (int l = 3; l <= 7; l += 2) {
        rc = og_smoothgradient(g, l, t1, sm);
        CHECK(rc == 0 && all_finite(sm, N), "%s: og_smoothgradient(%d)", name, l);
    }
    printf("%s: ok\n", name);
    free(sm);
    free(st);
    free(t1);
    free(order);
    free(codes);
    free(obs);
    free(sd);
    free(picks);
    free(tt);
    free(coordrec);
    free(off);
    free(vel);
}

int main(void) {
    {
        og_grid g = {OG_CART2D, 2, {40, 30, 1}, {0.5, 0.5, 0.5}, {0.0, 0.0, 0.0}};
        const double src[] = {7.3, 9.1, 12.0, 3.0};
        const double rec[] = {1.0, 1.0, 18.0, 0.5, 19.5, 14.5};
        run_case("cart2d", &g, src, 2, rec, 3, 0, 0, 7);
        run_case("cart2d refined", &g, src, 2, rec, 3, 1, 3, 7);
    }
    {
        og_grid g = {OG_CART3D, 3, {14, 12, 10}, {1.0, 1.0, 1.0}, {0.0, 0.0, 0.0}};
        const double src[] = {6.4, 5.2, 7.7, 3.0, 3.0, 2.0};
        const double rec[] = {1.0, 1.0, 0.0, 12.5, 10.0, 0.0, 7.0, 3.5, 0.0};
        run_case("cart3d", &g, src, 2, rec, 3, 0, 0, 7);
        run_case("cart3d refined", &g, src, 2, rec, 3, 1, 5, 7);
    }
    {
        og_grid g = {OG_SPH2D, 2, {30, 40, 1}, {2.0, 0.2, 1.0}, {6300.0, 40.0, 0.0}};
        const double src[] = {6330.0, 43.1};
        const double rec[] = {6358.0, 40.5, 6358.0, 46.9};
        run_case("sphere2d", &g, src, 1, rec, 2, 0, 0, 1);
        run_case("sphere2d refined", &g, src, 1, rec, 2, 1, 0, 1);
    }
    {
        og_grid g = {OG_SPH3D, 3, {12, 14, 14}, {2.0, 0.2, 0.2}, {6340.0, 30.0, 10.0}};
        const double src[] = {6352.0, 31.3, 11.1};
        const double rec[] = {6362.0, 30.2, 10.4, 6362.0, 32.4, 12.3};
        run_case("sphere3d", &g, src, 1, rec, 2, 0, 0, 1);
    }
    if (failures) {
        fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    printf("all oracle entry points ran clean\n");
    return 0;
}
