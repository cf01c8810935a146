/*
 * eik_oracle.c -- CPU oracle for the eikonal hot path (TEST INFRASTRUCTURE ONLY).
 *
 * Plain-C FP64 restatement of EikonalSolvers.jl v0.15.3 (src/EikSolv/eik*.jl, utils.jl and src/BinHeap/BinHeaps.jl).
 * Each function cites the reference file:line it follows.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this library; the product library never does.
 *
 * Index convention: 0-based everywhere; the reference is 1-based, so every boundary test is
 * shifted by one (e.g. isonbord's "ib<2 || ib>n1-1" becomes "ib<1 || ib>n1-2").
 */
#include "eik_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define HUGEV INFINITY

static __thread char g_err[512];
static int64_t g_warn = 0;

const char* og_last_error(void) { return g_err; }
int64_t og_warning_count(void) { return g_warn; }
void og_reset_warnings(void) { g_warn = 0; }

static int fail(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return -1;
}
static void warn_hit(void) {
#pragma omp atomic
    g_warn++;
}

/* ------------------------------------------------------------------ grid helpers */

static inline int64_t npts(const og_grid* g) {
    return g->ndim == 2 ? g->n[0] * g->n[1] : g->n[0] * g->n[1] * g->n[2];
}
/* cart2lin (eikutils.jl:6-19) */
static inline int64_t lin(const og_grid* g, const int64_t* ijk) {
    if (g->ndim == 2) return ijk[0] + g->n[0] * ijk[1];
    return ijk[0] + g->n[0] * (ijk[1] + g->n[1] * ijk[2]);
}
/* lin2cart! (eikutils.jl:24-43) */
static inline void unlin(const og_grid* g, int64_t l, int64_t* ijk) {
    ijk[0] = l % g->n[0];
    int64_t r = l / g->n[0];
    if (g->ndim == 2) {
        ijk[1] = r;
        ijk[2] = 0;
    } else {
        ijk[1] = r % g->n[1];
        ijk[2] = r / g->n[1];
    }
}
/* grd.x[i] / grd.r[i] ... ranges (eikstructs.jl:48-52): start + (i-1)*step, correctly rounded */
static inline double coord(const og_grid* g, int d, int64_t i0) { return fma((double)i0, g->d[d], g->o[d]); }
static inline double deg2rad(double x) { return x * (M_PI / 180.0); }
static inline int is_sph(const og_grid* g) { return g->kind == OG_SPH2D || g->kind == OG_SPH3D; }

/* Test hook (og_plant_mistake): a known mistake planted in the spherical metric so that the
 * tests can show that their pins detect it.  0 = none (the only value any real use sees). */
static int planted_mistake = 0;
int og_plant_mistake(int which) {
    const int old = planted_mistake;
    planted_mistake = which;
    return old;
}

/* per-node spacing of the stencil (eikforward2D.jl:115-119, eikforward3D.jl:168-172).
 * OG_SPH3D uses the engine's own metric r*sin(theta)*dphi (the reference's 3D sphere is
 * non-functional). */
static inline void node_spacing(const og_grid* g, const int64_t* ijk, double* dh) {
    if (!is_sph(g)) {
        dh[0] = dh[1] = dh[2] = g->d[0];
        return;
    }
    double r = coord(g, 0, ijk[0]);
    dh[0] = g->d[0];
    dh[1] = r * deg2rad(g->d[1]);
    if (planted_mistake == OG_MISTAKE_DEGREES) dh[1] = r * g->d[1];
    if (planted_mistake == OG_MISTAKE_NO_RADIUS) dh[1] = deg2rad(g->d[1]);
    if (g->ndim == 3) {
        double th = coord(g, 1, ijk[1]);
        dh[2] = r * sin(deg2rad(th)) * deg2rad(g->d[2]);
        if (planted_mistake == OG_MISTAKE_DEGREES) dh[2] = r * sin(deg2rad(th)) * g->d[2];
        if (planted_mistake == OG_MISTAKE_NO_RADIUS) dh[2] = sin(deg2rad(th)) * deg2rad(g->d[2]);
        if (planted_mistake == OG_MISTAKE_NO_SINTHETA) dh[2] = r * deg2rad(g->d[2]);
    }
}

/* isonbord (eikutils.jl:129-162): 1st-order test = outside the grid; 2nd-order test = the
 * first neighbour lies on ANY boundary plane. */
static inline void isonbord(const og_grid* g, const int64_t* nb, int* onb1, int* onb2) {
    int a = 0, b = 0;
    for (int d = 0; d < g->ndim; d++) {
        if (nb[d] < 0 || nb[d] > g->n[d] - 1) a = 1;
        if (nb[d] < 1 || nb[d] > g->n[d] - 2) b = 1;
    }
    *onb1 = a;
    *onb2 = b;
}

/* ------------------------------------------------------------------ local stencil */

/* calcttpt_2ndord! (eikforward2D.jl:98-433, eikforward3D.jl:151-532).  Same operation order
 * as the reference so FP64 results agree bit for bit. */
static double stencil(const og_grid* g, const double* vel, const double* tt, const uint8_t* st,
                      const int64_t* ijk, int32_t* code) {
    const int nd = g->ndim;
    double dh[3];
    node_spacing(g, ijk, dh);
    const double s = 1.0 / vel[lin(g, ijk)];
    double alpha = 0.0, beta = 0.0, gamma = -(s * s);
    for (int d = 0; d < nd; d++) code[d] = 0;

    for (int ax = 0; ax < nd; ax++) {
        int use1 = 0, use2 = 0;
        double cv1 = HUGEV, cv2 = HUGEV;
        for (int l = 0; l < 2; l++) {
            const int sh = (l == 0) ? 1 : -1;
            int64_t nb[3] = {ijk[0], ijk[1], ijk[2]};
            nb[ax] += sh;
            int onb1, onb2;
            isonbord(g, nb, &onb1, &onb2);
            if (!onb1 && st[lin(g, nb)] == 2) {
                double tv1 = tt[lin(g, nb)];
                if (tv1 < cv1) {
                    cv1 = tv1;
                    use1 = 1;
                    code[ax] = -sh;
                    int64_t nb2[3] = {ijk[0], ijk[1], ijk[2]};
                    nb2[ax] += 2 * sh;
                    if (!onb2 && st[lin(g, nb2)] == 2) {
                        double tv2 = tt[lin(g, nb2)];
                        if (tv2 <= cv1) {
                            cv2 = tv2;
                            use2 = 1;
                            code[ax] = -2 * sh;
                        } else {
                            cv2 = HUGEV;
                            use2 = 0;
                            code[ax] = -sh;
                        }
                    }
                }
            }
        }
        const double h = dh[ax];
        if (use2 && use1) {
            double tmpa2 = 1.0 / 3.0 * (4.0 * cv1 - cv2);
            double ca = 9.0 / (4.0 * (h * h));
            alpha += ca;
            beta += (-2.0 * ca * tmpa2);
            gamma += ca * (tmpa2 * tmpa2);
        } else if (use1) {
            double ca = 1.0 / (h * h);
            alpha += ca;
            beta += (-2.0 * ca * cv1);
            gamma += ca * (cv1 * cv1);
        }
    }
    double sqarg = beta * beta - 4.0 * alpha * gamma;

    if (sqarg < 0.0) { /* revert to 1st order on all axes (eikforward2D.jl:265-337) */
        for (int d = 0; d < nd; d++) code[d] = 0;
        alpha = 0.0;
        beta = 0.0;
        gamma = -(s * s);
        for (int ax = 0; ax < nd; ax++) {
            int use1 = 0;
            double cv1 = HUGEV;
            for (int l = 0; l < 2; l++) {
                const int sh = (l == 0) ? 1 : -1;
                int64_t nb[3] = {ijk[0], ijk[1], ijk[2]};
                nb[ax] += sh;
                int onb1, onb2;
                isonbord(g, nb, &onb1, &onb2);
                if (!onb1 && st[lin(g, nb)] == 2) {
                    double tv1 = tt[lin(g, nb)];
                    if (tv1 < cv1) {
                        cv1 = tv1;
                        use1 = 1;
                        code[ax] = -sh;
                    }
                }
            }
            if (use1) {
                const double h = dh[ax];
                double ca = 1.0 / (h * h);
                alpha += ca;
                beta += (-2.0 * ca * cv1);
                gamma += ca * (cv1 * cv1);
            }
        }
        sqarg = beta * beta - 4.0 * alpha * gamma;
    }
    if (sqarg < 0.0) { /* drop to one dimension (eikforward2D.jl:340-421) */
        warn_hit();
        double cvv[3] = {HUGEV, HUGEV, HUGEV};
        for (int d = 0; d < nd; d++) code[d] = 0;
        for (int ax = 0; ax < nd; ax++) {
            for (int l = 0; l < 2; l++) {
                const int sh = (l == 0) ? 1 : -1;
                int64_t nb[3] = {ijk[0], ijk[1], ijk[2]};
                nb[ax] += sh;
                int onb1, onb2;
                isonbord(g, nb, &onb1, &onb2);
                if (!onb1 && st[lin(g, nb)] == 2) {
                    double tv1 = tt[lin(g, nb)];
                    if (tv1 < cvv[ax]) {
                        cvv[ax] = tv1;
                        code[ax] = -sh;
                    }
                }
            }
        }
        int idc = 0; /* argmin: first minimum */
        for (int ax = 1; ax < nd; ax++)
            if (cvv[ax] < cvv[idc]) idc = ax;
        double t = (dh[idc] * s) + cvv[idc];
        int32_t keep = code[idc];
        for (int d = 0; d < nd; d++) code[d] = 0;
        code[idc] = keep;
        return t;
    }
    double tmpsq = sqrt(sqarg);
    double t1 = (-beta + tmpsq) / (2.0 * alpha);
    double t2 = (-beta - tmpsq) / (2.0 * alpha);
    return fmax(t1, t2);
}

/* The mixed precision's stencil (the engine's EIK_PREC_MIXED; no reference counterpart): the same
 * neighbour choice as stencil() in FP64, the quadratic solved in FP32 for the increment over the
 * smallest chosen value r, with the offsets c_a - r formed in FP64 and rounded once to FP32, the
 * result r + increment formed in FP64.  1st-order fallback in FP32; 1-D fallback
 * cmin + (float)h * s.  Slowness and metric are the FP64 values rounded to FP32; the metric
 * coefficients are 1/h^2 and 2.25/h^2 in FP32. */
static double stencil_mix(const og_grid* g, const double* vel, const double* tt, const uint8_t* st,
                          const int64_t* ijk, int32_t* code) {
    const int nd = g->ndim;
    double dh[3];
    node_spacing(g, ijk, dh);
    const float s = (float)(1.0 / vel[lin(g, ijk)]);
    int use1v[3] = {0, 0, 0}, use2v[3] = {0, 0, 0};
    double cv1v[3] = {HUGEV, HUGEV, HUGEV}, cv2v[3] = {HUGEV, HUGEV, HUGEV};
    int32_t codeA[3] = {0, 0, 0}, codeB[3] = {0, 0, 0};
    for (int ax = 0; ax < nd; ax++) {
        for (int l = 0; l < 2; l++) {
            const int sh = (l == 0) ? 1 : -1;
            int64_t nb[3] = {ijk[0], ijk[1], ijk[2]};
            nb[ax] += sh;
            int onb1, onb2;
            isonbord(g, nb, &onb1, &onb2);
            if (!onb1 && st[lin(g, nb)] == 2) {
                double tv1 = tt[lin(g, nb)];
                if (tv1 < cv1v[ax]) {
                    cv1v[ax] = tv1;
                    use1v[ax] = 1;
                    codeA[ax] = -sh;
                    codeB[ax] = -sh;
                    int64_t nb2[3] = {ijk[0], ijk[1], ijk[2]};
                    nb2[ax] += 2 * sh;
                    if (!onb2 && st[lin(g, nb2)] == 2) {
                        double tv2 = tt[lin(g, nb2)];
                        if (tv2 <= cv1v[ax]) {
                            cv2v[ax] = tv2;
                            use2v[ax] = 1;
                            codeA[ax] = -2 * sh;
                        } else {
                            cv2v[ax] = HUGEV;
                            use2v[ax] = 0;
                            codeA[ax] = -sh;
                        }
                    }
                }
            }
        }
    }
    float c1[3], c2[3], hf[3];
    for (int ax = 0; ax < nd; ax++) {
        hf[ax] = (float)dh[ax];
        c1[ax] = 1.0f / (hf[ax] * hf[ax]);
        c2[ax] = 2.25f * c1[ax];
    }
    double r = HUGEV;
    for (int ax = 0; ax < nd; ax++)
        if (use1v[ax] && cv1v[ax] < r) r = cv1v[ax];
    float A = 0.f, B = 0.f, C = -(s * s), d1v[3] = {0.f, 0.f, 0.f};
    for (int ax = 0; ax < nd; ax++) {
        if (!use1v[ax]) continue;
        const int sec = use2v[ax];
        const float al = sec ? c2[ax] : c1[ax];
        const float d1 = (float)(cv1v[ax] - r);
        const float d = sec ? d1 + (float)(cv1v[ax] - cv2v[ax]) * (1.0f / 3.0f) : d1;
        d1v[ax] = d1;
        A = A + al;
        B = B + al * d;
        C = C + al * d * d;
    }
    for (int d = 0; d < nd; d++) code[d] = codeA[d];
    float disc = B * B - A * C;
    if (!(disc < 0.f)) return r + (double)((B + sqrtf(disc)) / A);
    float A1 = 0.f, B1 = 0.f, C1 = -(s * s);
    for (int ax = 0; ax < nd; ax++) {
        if (!use1v[ax]) continue;
        A1 = A1 + c1[ax];
        B1 = B1 + c1[ax] * d1v[ax];
        C1 = C1 + c1[ax] * d1v[ax] * d1v[ax];
    }
    for (int d = 0; d < nd; d++) code[d] = codeB[d];
    disc = B1 * B1 - A1 * C1;
    if (!(disc < 0.f)) return r + (double)((B1 + sqrtf(disc)) / A1);
    warn_hit();
    int idc = 0; /* first axis with the smallest adjacent value */
    for (int ax = 1; ax < nd; ax++)
        if (cv1v[ax] < cv1v[idc]) idc = ax;
    for (int d = 0; d < nd; d++) code[d] = d == idc ? codeB[d] : 0;
    return cv1v[idc] + (double)(hf[idc] * s);
}

/* The mixed precision's arithmetic for node ijk: FP64 inside the near-source region xb (the source
 * box +- 4 nodes; NULL: none), the mixed stencil elsewhere. */
typedef struct {
    int64_t lo[3], hi[3];
} xbox_t;
static inline int in_xbox(const xbox_t* xb, const int64_t* ijk, int nd) {
    if (!xb) return 0;
    for (int d = 0; d < nd; d++)
        if (ijk[d] < xb->lo[d] || ijk[d] > xb->hi[d]) return 0;
    return 1;
}
static inline double stencil_any(const og_grid* g, const double* vel, const double* tt, const uint8_t* st,
                                 const int64_t* ijk, int32_t* code, int mix, const xbox_t* xb) {
    return (mix && !in_xbox(xb, ijk, g->ndim)) ? stencil_mix(g, vel, tt, st, ijk, code)
                                                : stencil(g, vel, tt, st, ijk, code);
}

double og_stencil(const og_grid* g, const double* vel, const double* tt, const uint8_t* status,
                  const int64_t* ijk, int32_t* codes) {
    int64_t p[3] = {ijk[0], ijk[1], g->ndim == 3 ? ijk[2] : 0};
    return stencil(g, vel, tt, status, p, codes);
}

double og_stencil_mixed(const og_grid* g, const double* vel, const double* tt, const uint8_t* status,
                        const int64_t* ijk, int32_t* codes) {
    int64_t p[3] = {ijk[0], ijk[1], g->ndim == 3 ? ijk[2] : 0};
    return stencil_mix(g, vel, tt, status, p, codes);
}

/* ------------------------------------------------------------------ binary min-heap */
/* BinHeapMin with handles (BinHeaps.jl:56-62, 313-485); 1-based positions, 0 = absent. */
typedef struct {
    int64_t nmax, nh;
    double* v;    /* [nmax+1] */
    int64_t* hd;  /* [nmax+1] */
    int64_t* ix;  /* [nmax] indexed by handle */
} heap_t;

static int hp_init(heap_t* h, int64_t nmax) {
    h->nmax = nmax;
    h->nh = 0;
    h->v = (double*)calloc((size_t)nmax + 1, sizeof(double));
    h->hd = (int64_t*)calloc((size_t)nmax + 1, sizeof(int64_t));
    h->ix = (int64_t*)calloc((size_t)nmax, sizeof(int64_t));
    return (h->v && h->hd && h->ix) ? 0 : -1;
}
static void hp_free(heap_t* h) {
    free(h->v);
    free(h->hd);
    free(h->ix);
}
static inline void hp_swap(heap_t* h, int64_t p, int64_t q) { /* swap_nodes_heap! (BinHeaps.jl:93-113) */
    double tv = h->v[p];
    int64_t th = h->hd[p];
    h->v[p] = h->v[q];
    h->v[q] = tv;
    h->hd[p] = h->hd[q];
    h->hd[q] = th;
    h->ix[h->hd[p]] = p;
    h->ix[h->hd[q]] = q;
}
static void hp_heapify(heap_t* h, int64_t i) { /* min_heapify! (BinHeaps.jl:364-385) */
    for (;;) {
        int64_t l = 2 * i, r = 2 * i + 1, sm;
        sm = (l <= h->nh && h->v[l] < h->v[i]) ? l : i;
        if (r <= h->nh && h->v[r] < h->v[sm]) sm = r;
        if (sm == i) return;
        hp_swap(h, i, sm);
        i = sm;
    }
}
static void hp_insert(heap_t* h, double val, int64_t handle) { /* insert_minheap! (:437-460) */
    h->nh++;
    h->v[h->nh] = val;
    h->hd[h->nh] = handle;
    h->ix[handle] = h->nh;
    int64_t i = h->nh;
    while (i > 1 && h->v[i / 2] > h->v[i]) {
        hp_swap(h, i, i / 2);
        i /= 2;
    }
}
static void hp_update(heap_t* h, double val, int64_t handle) { /* update_node_minheap! (:394-430) */
    int64_t i = h->ix[handle];
    h->v[i] = val;
    while (i > 1 && h->v[i / 2] > h->v[i]) {
        hp_swap(h, i, i / 2);
        i /= 2;
    }
    if (i < h->nh) hp_heapify(h, i);
}
static int64_t hp_pop(heap_t* h, double* val) { /* pop_minheap! (:467-485) */
    int64_t top = h->hd[1];
    *val = h->v[1];
    h->v[1] = h->v[h->nh];
    h->hd[1] = h->hd[h->nh];
    h->ix[h->hd[h->nh]] = 1;
    h->v[h->nh] = 0.0;
    h->hd[h->nh] = 0;
    h->ix[top] = 0;
    h->nh--;
    hp_heapify(h, 1);
    return top;
}

/* ------------------------------------------------------------------ per-grid state */

typedef struct {
    int64_t n;
    double* tt;
    uint8_t* st;
    heap_t hp;
} fmm_t;

static int fmm_alloc(fmm_t* f, int64_t n) {
    f->n = n;
    f->tt = (double*)malloc((size_t)n * sizeof(double));
    f->st = (uint8_t*)malloc((size_t)n);
    if (!f->tt || !f->st || hp_init(&f->hp, n)) return -1;
    return 0;
}
static void fmm_free(fmm_t* f) {
    free(f->tt);
    free(f->st);
    hp_free(&f->hp);
}
static void fmm_reset(fmm_t* f) { /* ttFMM_hiord! (eikmainfwd.jl:249-253) */
    for (int64_t i = 0; i < f->n; i++) {
        f->tt[i] = HUGEV;
        f->st[i] = 0;
    }
    f->hp.nh = 0;
}

/* discrete-adjoint recording (AdjointVars*, eikstructs.jl:286-406) */
typedef struct {
    int64_t n;
    int nd;
    int64_t* lfmm2grid; /* FMM index -> grid index, -1 unset */
    int64_t* lgrid2fmm; /* grid index -> FMM index */
    double* ttfmm;      /* tt in FMM order, -1 unset */
    uint8_t* onsrc;     /* by FMM index */
    uint8_t* onh;       /* by FMM index */
    int8_t* code;       /* by grid index, n*nd */
    int64_t lastcomputed;
    /* derivative matrices D_d: rows = non-source FMM rows, cols = FMM index (CSR, sorted) */
    int64_t nrows, ncols, nsrc;
    int64_t* iptr[3];
    int64_t* jc[3];
    double* vv[3];
} adj_t;

static int adj_alloc(adj_t* a, int64_t n, int nd) {
    memset(a, 0, sizeof(*a));
    a->n = n;
    a->nd = nd;
    a->lfmm2grid = (int64_t*)malloc((size_t)n * sizeof(int64_t));
    a->lgrid2fmm = (int64_t*)malloc((size_t)n * sizeof(int64_t));
    a->ttfmm = (double*)malloc((size_t)n * sizeof(double));
    a->onsrc = (uint8_t*)malloc((size_t)n);
    a->onh = (uint8_t*)malloc((size_t)n);
    a->code = (int8_t*)malloc((size_t)n * nd);
    for (int d = 0; d < nd; d++) {
        a->iptr[d] = (int64_t*)malloc(((size_t)n + 1) * sizeof(int64_t));
        a->jc[d] = (int64_t*)malloc((size_t)n * 3 * sizeof(int64_t));
        a->vv[d] = (double*)malloc((size_t)n * 3 * sizeof(double));
        if (!a->iptr[d] || !a->jc[d] || !a->vv[d]) return -1;
    }
    if (!a->lfmm2grid || !a->lgrid2fmm || !a->ttfmm || !a->onsrc || !a->onh || !a->code) return -1;
    for (int64_t i = 0; i < n; i++) a->lfmm2grid[i] = -1;
    return 0;
}
static void adj_free(adj_t* a) {
    free(a->lfmm2grid);
    free(a->lgrid2fmm);
    free(a->ttfmm);
    free(a->onsrc);
    free(a->onh);
    free(a->code);
    for (int d = 0; d < a->nd; d++) {
        free(a->iptr[d]);
        free(a->jc[d]);
        free(a->vv[d]);
    }
}
static void adj_reset(adj_t* a) { /* ttFMM_hiord! (eikmainfwd.jl:257-279) */
    for (int64_t i = 0; i < a->n; i++) {
        a->onsrc[i] = 0;
        a->onh[i] = 0;
        a->ttfmm[i] = -1.0;
    }
    memset(a->code, 0, (size_t)a->n * a->nd);
    a->lastcomputed = 0;
}

/* source box (SourceBoxParams, eikstructs.jl:411-420) */
typedef struct {
    int ncorn;
    int64_t ijk[8][3];
    double velcorn[8], distcorn[8];
} srcbox_t;

/* refinement variables (SrcRefinVars, eikstructs.jl:492-500) */
typedef struct {
    og_grid gf;
    int df;
    int64_t ijk1c[3];
    int outmin[3], outmax[3];
    int64_t nwin[3];
    int64_t* nn_idxcoarse; /* per fine node: linear index inside the coarse window */
    double* velfine;
} refvars_t;

static void refvars_free(refvars_t* r) {
    free(r->nn_idxcoarse);
    free(r->velfine);
}

/* ------------------------------------------------------------------ geometry */

/* findenclosingbox (eikmainfwd.jl:1114-1186).  Returns the corner count (1 or 2^ndim). */
static int findenclosingbox(const og_grid* g, const double* src, int64_t corn[8][3]) {
    const int nd = g->ndim;
    int64_t ijk[3] = {0, 0, 0};
    int onnode = 1;
    for (int d = 0; d < nd; d++) {
        double x = src[d] - g->o[d];
        double r = fmod(x, g->d[d]);
        double q = nearbyint((x - r) / g->d[d]); /* Julia div(x,y) for floats */
        ijk[d] = (int64_t)floor(q);
        if (!(fabs(r) <= 1e-8)) onnode = 0;
    }
    if (onnode) {
        for (int d = 0; d < 3; d++) corn[0][d] = (d < nd) ? ijk[d] : 0;
        return 1;
    }
    for (int d = 0; d < nd; d++)
        if (ijk[d] == g->n[d] - 1) ijk[d] -= 1;
    static const int off2[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
    static const int off3[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                   {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
    int nc = 1 << nd;
    for (int c = 0; c < nc; c++)
        for (int d = 0; d < 3; d++)
            corn[c][d] = (d < nd) ? ijk[d] + (nd == 2 ? off2[c][d] : off3[c][d]) : 0;
    return nc;
}

/* distance2points (eikutils.jl:47-73); OG_SPH3D: the engine's chord metric */
static double distance2points(const og_grid* g, const double* a, const double* b) {
    if (g->kind == OG_SPH2D) {
        double r1 = a[0], t1 = a[1], r2 = b[0], t2 = b[1];
        if (planted_mistake == OG_MISTAKE_INIT_DEGREES) return sqrt(r1 * r1 + r2 * r2 - 2.0 * r1 * r2 * cos(t1 - t2));
        return sqrt(r1 * r1 + r2 * r2 - 2.0 * r1 * r2 * cos(deg2rad(t1 - t2)));
    }
    if (g->kind == OG_SPH3D) {
        double xa = a[0] * sin(deg2rad(a[1])) * cos(deg2rad(a[2]));
        double ya = a[0] * sin(deg2rad(a[1])) * sin(deg2rad(a[2]));
        double za = a[0] * cos(deg2rad(a[1]));
        double xb = b[0] * sin(deg2rad(b[1])) * cos(deg2rad(b[2]));
        double yb = b[0] * sin(deg2rad(b[1])) * sin(deg2rad(b[2]));
        double zb = b[0] * cos(deg2rad(b[1]));
        return sqrt((xa - xb) * (xa - xb) + (ya - yb) * (ya - yb) + (za - zb) * (za - zb));
    }
    double s = 0.0;
    for (int d = 0; d < g->ndim; d++) s += (a[d] - b[d]) * (a[d] - b[d]);
    return sqrt(s);
}

/* sourceboxloctt! (eikmainfwd.jl:1195-1322); the "halfway" shift (:1258-1319) only rewrites a
 * field nothing reads, so it reduces to the warning. */
static int sourceboxloctt(const og_grid* g, const double* vel, const double* src, fmm_t* f, srcbox_t* sb) {
    sb->ncorn = findenclosingbox(g, src, sb->ijk);
    for (int c = 0; c < sb->ncorn; c++) {
        int64_t l = lin(g, sb->ijk[c]);
        f->st[l] = 2;
        double xyz[3];
        for (int d = 0; d < g->ndim; d++) xyz[d] = coord(g, d, sb->ijk[c][d]);
        double dist = distance2points(g, xyz, src);
        sb->velcorn[c] = vel[l];
        sb->distcorn[c] = dist;
        f->tt[l] = dist / vel[l];
    }
    if (sb->ncorn > 1) {
        for (int d = 0; d < g->ndim; d++) {
            double r = fmod(src[d] - g->o[d], g->d[d]);
            double hh = g->d[d] / 2.0;
            if (fabs(r - hh) <= 1.4901161193847656e-08 * fmax(fabs(r), fabs(hh))) {
                warn_hit();
                break;
            }
        }
    }
    return 0;
}

/* createfinegrid (eikmainfwd.jl:950-1104) */
static int createfinegrid(const og_grid* g, const double* src, const double* vel, int df, int noderadius,
                          refvars_t* R) {
    const int nd = g->ndim;
    int64_t corn[8][3];
    int nc = findenclosingbox(g, src, corn);
    int64_t i1[3] = {0, 0, 0}, i2[3] = {0, 0, 0};
    for (int d = 0; d < nd; d++) {
        int64_t mn = corn[0][d], mx = corn[0][d];
        for (int c = 1; c < nc; c++) {
            if (corn[c][d] < mn) mn = corn[c][d];
            if (corn[c][d] > mx) mx = corn[c][d];
        }
        int64_t v1 = mn - noderadius, v2 = mx + noderadius;
        R->outmin[d] = v1 < 0;
        R->outmax[d] = v2 > g->n[d] - 1;
        i1[d] = R->outmin[d] ? 0 : v1;
        i2[d] = R->outmax[d] ? g->n[d] - 1 : v2;
    }
    memset(&R->gf, 0, sizeof(R->gf));
    R->gf.kind = g->kind;
    R->gf.ndim = nd;
    R->df = df;
    for (int d = 0; d < 3; d++) {
        R->ijk1c[d] = (d < nd) ? i1[d] : 0;
        R->nwin[d] = (d < nd) ? i2[d] - i1[d] + 1 : 1;
        R->gf.n[d] = (d < nd) ? (i2[d] - i1[d]) * df + 1 : 1;
    }
    for (int d = 0; d < nd; d++) {
        R->gf.d[d] = g->d[d] / df;
        R->gf.o[d] = coord(g, d, i1[d]);
    }
    if (!is_sph(g))
        for (int d = 0; d < 3; d++) R->gf.d[d] = g->d[0] / df;
    int64_t nf = npts(&R->gf);
    R->nn_idxcoarse = (int64_t*)malloc((size_t)nf * sizeof(int64_t));
    R->velfine = (double*)malloc((size_t)nf * sizeof(double));
    if (!R->nn_idxcoarse || !R->velfine) return fail("createfinegrid: out of memory");
    for (int64_t l = 0; l < nf; l++) {
        int64_t fi[3], ci[3] = {0, 0, 0};
        unlin(&R->gf, l, fi);
        for (int d = 0; d < nd; d++) {
            int64_t i_f = fi[d] + 1; /* 1-based as in the reference */
            int64_t di = (i_f - 1) / df;
            int64_t ri = i_f - di * df;
            ci[d] = ((double)ri >= (double)df / 2.0 + 1.0) ? di + 1 : di; /* 0-based window index */
        }
        int64_t jw = (nd == 2) ? ci[0] + R->nwin[0] * ci[1] : ci[0] + R->nwin[0] * (ci[1] + R->nwin[1] * ci[2]);
        R->nn_idxcoarse[l] = jw;
        int64_t cg[3] = {ci[0] + i1[0], ci[1] + i1[1], (nd == 3) ? ci[2] + i1[2] : 0};
        R->velfine[l] = vel[lin(g, cg)];
    }
    return 0;
}

/* ------------------------------------------------------------------ derivative matrices */

/* createsparsederivativematrices! + setcoeffderiv! + addrowCSRmat!
 * (eikmainfwd.jl:827-946, eikmaingrad.jl:946-1131) */
static void build_deriv(const og_grid* g, adj_t* A, const uint8_t* st) {
    const int nd = g->ndim;
    int64_t nsrc = 0;
    for (int64_t i = 0; i < A->n; i++) nsrc += A->onsrc[i];
    int64_t nsteps = A->n;
    for (int64_t i = 0; i < A->n; i++) {
        int64_t ifm = A->lfmm2grid[i];
        if (ifm < 0) {
            nsteps = i;
            break;
        }
        A->lgrid2fmm[ifm] = i;
    }
    A->nsrc = nsrc;
    A->nrows = nsteps - nsrc;
    A->ncols = nsteps;
    for (int d = 0; d < nd; d++) A->iptr[d][0] = 0;
    for (int64_t irow = 0; irow < A->nrows; irow++) {
        int64_t iorig = A->lfmm2grid[irow + nsrc];
        int64_t ijk[3] = {0, 0, 0};
        unlin(g, iorig, ijk);
        int32_t codes[3] = {0, 0, 0};
        if (st[iorig] == 2)
            for (int d = 0; d < nd; d++) codes[d] = A->code[iorig * nd + d];
        for (int d = 0; d < nd; d++) {
            int64_t base = A->iptr[d][irow];
            int32_t c = codes[d];
            if (c == 0) {
                A->iptr[d][irow + 1] = base;
                continue;
            }
            int ac = c < 0 ? -c : c;
            int nnz = ac + 1;
            int sg = (c > 0) ? -1 : 1; /* signcod = -sign(code) */
            /* coefficient set (CoeffDerivCartesian / CoeffDerivSpherical2D): hgrid or dr, and for
             * the angular axes the arc length at the stencil node's radius */
            int64_t cols[3];
            double vals[3];
            for (int p = 0; p < nnz; p++) {
                int64_t q[3] = {ijk[0], ijk[1], ijk[2]};
                q[d] += (int64_t)sg * p;
                double hc = g->d[0];
                if (is_sph(g) && d >= 1) {
                    double r = coord(g, 0, q[0]);
                    hc = r * deg2rad(g->d[1]);
                    if (d == 2) hc = r * sin(deg2rad(coord(g, 1, q[1]))) * deg2rad(g->d[2]);
                }
                double cf;
                if (ac == 1)
                    cf = (p == 0) ? -1.0 / hc : 1.0 / hc;
                else
                    cf = (p == 0) ? -3.0 / (2.0 * hc) : (p == 1 ? 4.0 / (2.0 * hc) : -1.0 / (2.0 * hc));
                vals[p] = (double)sg * cf;
                cols[p] = A->lgrid2fmm[lin(g, q)];
            }
            /* sortperm! on the column indices */
            for (int a = 1; a < nnz; a++)
                for (int b = a; b > 0 && cols[b - 1] > cols[b]; b--) {
                    int64_t tc = cols[b];
                    cols[b] = cols[b - 1];
                    cols[b - 1] = tc;
                    double tv = vals[b];
                    vals[b] = vals[b - 1];
                    vals[b - 1] = tv;
                }
            for (int p = 0; p < nnz; p++) {
                A->jc[d][base + p] = cols[p];
                A->vv[d][base + p] = vals[p];
            }
            A->iptr[d][irow + 1] = base + nnz;
        }
    }
}

/* ------------------------------------------------------------------ FMM core */

typedef struct {
    const refvars_t* R; /* non-NULL while running on the fine grid */
    const og_grid* gc;  /* coarse grid */
    fmm_t* Fc;
    adj_t* Ac;
    int64_t* cinit; /* out: coarse init list (grid indices) */
    int64_t ncinit;
    int stopped;
} fine_ctx_t;

static const int neigh2[4][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
static const int neigh3[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};

static inline int is_coarse_node(const refvars_t* R, const int64_t* fijk, int nd, int64_t* cijk) {
    /* isacoarsegridnode (eikutils.jl:78-119) */
    for (int d = 0; d < nd; d++)
        if (fijk[d] % R->df != 0) return 0;
    for (int d = 0; d < 3; d++) cijk[d] = (d < nd) ? fijk[d] / R->df + R->ijk1c[d] : 0;
    return 1;
}

/* ttFMM_core! (eikmainfwd.jl:427-823) */
static int fmm_core(const og_grid* g, const double* vel, fmm_t* F, const int64_t* init, int64_t ninit,
                    adj_t* A, fine_ctx_t* fc, int mix, const xbox_t* xb) {
    const int nd = g->ndim;
    const int nnb = 2 * nd;
    const int64_t N = npts(g);
    int32_t idD[3];

    if (A) { /* eikmainfwd.jl:500-547 (the sortperm is a no-op: index l gets data for l) */
        A->lastcomputed = ninit;
        for (int64_t l = 0; l < ninit; l++) {
            A->lfmm2grid[l] = init[l];
            A->ttfmm[l] = F->tt[init[l]];
            A->onsrc[l] = 1;
        }
    }
    int64_t ntransfer = 0, cnt_adjc = 0;
    if (fc) { /* eikmainfwd.jl:554-581 */
        for (int64_t l = 0; l < ninit; l++) {
            int64_t fijk[3], cijk[3];
            unlin(g, init[l], fijk);
            if (is_coarse_node(fc->R, fijk, nd, cijk)) {
                int64_t lc = lin(fc->gc, cijk);
                fc->Fc->tt[lc] = F->tt[init[l]];
                fc->Fc->st[lc] = F->st[init[l]];
                fc->cinit[ntransfer++] = lc;
                if (A) A->onh[l] = 1;
            }
        }
    }
    /* initial narrow band (eikmainfwd.jl:590-621) */
    for (int64_t l = 0; l < ninit; l++) {
        int64_t b[3];
        unlin(g, init[l], b);
        for (int ne = 0; ne < nnb; ne++) {
            int64_t c[3] = {b[0], b[1], b[2]};
            for (int d = 0; d < nd; d++) c[d] += (nd == 2) ? neigh2[ne][d] : neigh3[ne][d];
            int oob = 0;
            for (int d = 0; d < nd; d++)
                if (c[d] < 0 || c[d] > g->n[d] - 1) oob = 1;
            if (oob) continue;
            int64_t lc = lin(g, c);
            if (F->st[lc] == 0) {
                double t = stencil_any(g, vel, F->tt, F->st, c, idD, mix, xb);
                hp_insert(&F->hp, t, lc);
                F->st[lc] = 1;
                if (A)
                    for (int d = 0; d < nd; d++) A->code[lc * nd + d] = (int8_t)idD[d];
            }
        }
    }
    /* main loop (eikmainfwd.jl:630-814) */
    int firstwarning = 1;
    int64_t node = ninit;
    int64_t cap = 128, *popped = (int64_t*)malloc((size_t)cap * sizeof(int64_t));
    if (!popped) return fail("fmm_core: out of memory");
    for (int64_t iter = 0; iter < N - ninit; iter++) {
        if (F->hp.nh < 1) break;
        int same = 1;
        int64_t np = 0;
        while (same) {
            node++;
            if (F->hp.nh < 1) break;
            double t;
            int64_t acc = hp_pop(&F->hp, &t);
            same = (F->hp.v[1] == t);
            int64_t aijk[3] = {0, 0, 0};
            unlin(g, acc, aijk);
            F->st[acc] = 2;
            F->tt[acc] = t;
            if (A) {
                A->lfmm2grid[node - 1] = acc;
                A->ttfmm[node - 1] = t;
                A->lastcomputed = node;
            }
            if (fc) {
                int64_t cijk[3];
                if (is_coarse_node(fc->R, aijk, nd, cijk)) {
                    int64_t lc = lin(fc->gc, cijk);
                    fc->Fc->tt[lc] = F->tt[acc];
                    fc->Fc->st[lc] = F->st[acc];
                    fc->cinit[ntransfer++] = lc;
                    if (A) {
                        A->onh[node - 1] = 1;
                        if (fc->Ac) {
                            fc->Ac->lfmm2grid[cnt_adjc] = lc;
                            fc->Ac->ttfmm[cnt_adjc] = F->tt[acc];
                        }
                        cnt_adjc++;
                    }
                }
                int onface = 0; /* eikmainfwd.jl:724 */
                for (int d = 0; d < nd; d++) {
                    if (aijk[d] == 0 && !fc->R->outmin[d]) onface = 1;
                    if (aijk[d] == g->n[d] - 1 && !fc->R->outmax[d]) onface = 1;
                }
                if (onface) {
                    if (ntransfer < 1) { /* eikmainfwd.jl:729-735: accepted but not expanded */
                        if (firstwarning) {
                            warn_hit();
                            firstwarning = 0;
                        }
                        continue;
                    }
                    if (A) build_deriv(g, A, F->st);
                    for (int64_t i = 0; i < fc->Fc->n; i++)
                        if (fc->Fc->st[i] == 1) fc->Fc->st[i] = 0;
                    fc->ncinit = ntransfer;
                    fc->stopped = 1;
                    free(popped);
                    return 0;
                }
            }
            if (np == cap) {
                cap *= 2;
                int64_t* pp = (int64_t*)realloc(popped, (size_t)cap * sizeof(int64_t));
                if (!pp) {
                    free(popped);
                    return fail("fmm_core: out of memory");
                }
                popped = pp;
            }
            popped[np++] = acc;
        }
        for (int64_t ip = 0; ip < np; ip++) {
            int64_t b[3];
            unlin(g, popped[ip], b);
            for (int ne = 0; ne < nnb; ne++) {
                int64_t c[3] = {b[0], b[1], b[2]};
                for (int d = 0; d < nd; d++) c[d] += (nd == 2) ? neigh2[ne][d] : neigh3[ne][d];
                int oob = 0;
                for (int d = 0; d < nd; d++)
                    if (c[d] < 0 || c[d] > g->n[d] - 1) oob = 1;
                if (oob) continue;
                int64_t lc = lin(g, c);
                if (F->st[lc] == 0) {
                    double t = stencil_any(g, vel, F->tt, F->st, c, idD, mix, xb);
                    hp_insert(&F->hp, t, lc);
                    F->st[lc] = 1;
                    if (A)
                        for (int d = 0; d < nd; d++) A->code[lc * nd + d] = (int8_t)idD[d];
                } else if (F->st[lc] == 1) {
                    double t = stencil_any(g, vel, F->tt, F->st, c, idD, mix, xb);
                    hp_update(&F->hp, t, lc);
                    if (A)
                        for (int d = 0; d < nd; d++) A->code[lc * nd + d] = (int8_t)idD[d];
                }
            }
        }
    }
    free(popped);
    if (fc) fc->ncinit = ntransfer; /* whole fine grid done without a stop */
    if (A) build_deriv(g, A, F->st);
    return 0;
}

/* ------------------------------------------------------------------ one source */

/* per-source fine-grid products kept for the gradient (runrefinementaroundsrc!, :330-405) */
typedef struct {
    refvars_t R;
    fmm_t Ff;
    adj_t Af;
    srcbox_t sbf;
    int have;
} finegrid_t;

static void finegrid_free(finegrid_t* fg) {
    if (!fg->have) return;
    refvars_free(&fg->R);
    fmm_free(&fg->Ff);
    if (fg->Af.lfmm2grid) adj_free(&fg->Af);
    fg->have = 0;
}

/* ttFMM_hiord! (eikmainfwd.jl:230-326).  On return F holds tt; cinit/ncinit hold the coarse
 * source points (ijksrc); sb the coarse source box when not refining. */
static int ttfmm_hiord(const og_grid* g, const double* vel, const double* src, const og_params* p,
                       fmm_t* F, adj_t* A, int64_t* cinit, int64_t* ncinit, srcbox_t* sb, finegrid_t* fg) {
    fmm_reset(F);
    if (A) adj_reset(A);
    if (p->refine) {
        memset(fg, 0, sizeof(*fg));
        if (createfinegrid(g, src, vel, p->df, p->noderadius, &fg->R)) return -1;
        fg->have = 1;
        int64_t nf = npts(&fg->R.gf);
        if (fmm_alloc(&fg->Ff, nf)) return fail("out of memory (fine grid)");
        fmm_reset(&fg->Ff);
        adj_t* Af = NULL;
        if (A) {
            if (adj_alloc(&fg->Af, nf, g->ndim)) return fail("out of memory (fine adjoint)");
            adj_reset(&fg->Af);
            Af = &fg->Af;
        }
        sourceboxloctt(&fg->R.gf, fg->R.velfine, src, &fg->Ff, &fg->sbf);
        int64_t finit[8];
        for (int c = 0; c < fg->sbf.ncorn; c++) finit[c] = lin(&fg->R.gf, fg->sbf.ijk[c]);
        fine_ctx_t fc;
        memset(&fc, 0, sizeof(fc));
        fc.R = &fg->R;
        fc.gc = g;
        fc.Fc = F;
        fc.Ac = A;
        fc.cinit = cinit;
        /* the refined box is solved in FP64 in both precisions */
        if (fmm_core(&fg->R.gf, fg->R.velfine, &fg->Ff, finit, fg->sbf.ncorn, Af, &fc, 0, NULL)) return -1;
        *ncinit = fc.ncinit;
    } else {
        sourceboxloctt(g, vel, src, F, sb);
        for (int c = 0; c < sb->ncorn; c++) cinit[c] = lin(g, sb->ijk[c]);
        *ncinit = sb->ncorn;
    }
    /* mixed precision without refinement: FP64 stencil within +-4 nodes of the source box */
    xbox_t xb;
    const int mix = p->precision == OG_PREC_MIXED;
    const int usebox = mix && !p->refine;
    if (usebox) {
        for (int d = 0; d < 3; d++) {
            xb.lo[d] = INT64_MAX;
            xb.hi[d] = INT64_MIN;
        }
        for (int c = 0; c < sb->ncorn; c++)
            for (int d = 0; d < g->ndim; d++) {
                if (sb->ijk[c][d] - OG_MIXED_FP64_RADIUS < xb.lo[d]) xb.lo[d] = sb->ijk[c][d] - OG_MIXED_FP64_RADIUS;
                if (sb->ijk[c][d] + OG_MIXED_FP64_RADIUS > xb.hi[d]) xb.hi[d] = sb->ijk[c][d] + OG_MIXED_FP64_RADIUS;
            }
    }
    return fmm_core(g, vel, F, cinit, *ncinit, A, NULL, mix, usebox ? &xb : NULL);
}

/* ------------------------------------------------------------------ receivers */

typedef struct {
    int ncoe;
    int64_t idx[8];
    double cf[8];
} interp_t;

/* bilinear_interp (eikforward2D.jl:17-87) and trilinear_interp (eikforward3D.jl:17-140,
 * including the f101 weight of the reference).  value = forward-path evaluation order. */
static double interp(const og_grid* g, const double* f, const double* pt, interp_t* ic) {
    if (g->ndim == 2) {
        double dx = g->d[0], dy = g->d[1];
        if (!is_sph(g)) dy = g->d[0];
        double xh = (pt[0] - g->o[0]) / dx, yh = (pt[1] - g->o[1]) / dy;
        int64_t i = (int64_t)floor(xh + 1.0), j = (int64_t)floor(yh + 1.0);
        int64_t nx = g->n[0], ny = g->n[1];
        if (i == nx) i -= 1; else if (i > nx) i -= 2;
        if (j == ny) j -= 1; else if (j > ny) j -= 2;
        double xd = xh - (double)(i - 1), yd = yh - (double)(j - 1);
        int64_t i0 = i - 1, j0 = j - 1;
        int64_t l00 = i0 + nx * j0, l10 = l00 + 1, l01 = l00 + nx, l11 = l01 + 1;
        if (ic) {
            ic->ncoe = 4;
            ic->cf[0] = (1.0 - xd) * (1.0 - yd);
            ic->cf[1] = (1.0 - yd) * xd;
            ic->cf[2] = (1.0 - xd) * yd;
            ic->cf[3] = xd * yd;
            ic->idx[0] = l00;
            ic->idx[1] = l10;
            ic->idx[2] = l01;
            ic->idx[3] = l11;
        }
        if (!f) return 0.0;
        return f[l00] * (1.0 - xd) * (1.0 - yd) + f[l10] * (1.0 - yd) * xd + f[l01] * (1.0 - xd) * yd +
               f[l11] * xd * yd;
    }
    double dx = g->d[0], dy = g->d[1], dz = g->d[2];
    if (!is_sph(g)) dy = dz = g->d[0];
    double xh = (pt[0] - g->o[0]) / dx, yh = (pt[1] - g->o[1]) / dy, zh = (pt[2] - g->o[2]) / dz;
    double x = pt[0] - g->o[0], y = pt[1] - g->o[1], z = pt[2] - g->o[2];
    int64_t i = (int64_t)floor(xh + 1.0), j = (int64_t)floor(yh + 1.0), k = (int64_t)floor(zh + 1.0);
    int64_t nx = g->n[0], ny = g->n[1], nz = g->n[2];
    if (i == nx) i -= 1; else if (i > nx) i -= 2;
    if (j == ny) j -= 1; else if (j > ny) j -= 2;
    if (k == nz) k -= 1; else if (k > nz) k -= 2;
    double x0 = (double)(i - 1) * dx, y0 = (double)(j - 1) * dy, z0 = (double)(k - 1) * dz;
    double x1 = (double)i * dx, y1 = (double)j * dy, z1 = (double)k * dz;
    double xd = (x - x0) / (x1 - x0), yd = (y - y0) / (y1 - y0), zd = (z - z0) / (z1 - z0);
    int64_t i0 = i - 1, j0 = j - 1, k0 = k - 1, nxy = nx * ny;
    int64_t l000 = i0 + nx * j0 + nxy * k0;
    int64_t l100 = l000 + 1, l010 = l000 + nx, l001 = l000 + nxy, l110 = l010 + 1, l011 = l010 + nxy,
            l101 = l001 + 1, l111 = l011 + 1;
    if (ic) {
        ic->ncoe = 8;
        ic->cf[0] = (1 - xd) * (1 - yd) * (1 - zd);
        ic->cf[1] = xd * (1 - yd) * (1 - zd);
        ic->cf[2] = (1 - xd) * yd * (1 - zd);
        ic->cf[3] = (1 - xd) * (1 - yd) * zd;
        ic->cf[4] = xd * yd * (1 - zd);
        ic->cf[5] = (1 - xd) * yd * zd;
        ic->cf[6] = xd * (1 - yd) * (1 - zd);
        ic->cf[7] = xd * yd * zd;
        ic->idx[0] = l000;
        ic->idx[1] = l100;
        ic->idx[2] = l010;
        ic->idx[3] = l001;
        ic->idx[4] = l110;
        ic->idx[5] = l011;
        ic->idx[6] = l101;
        ic->idx[7] = l111;
    }
    if (!f) return 0.0;
    return f[l000] * (1 - xd) * (1 - yd) * (1 - zd) + f[l100] * xd * (1 - yd) * (1 - zd) +
           f[l010] * (1 - xd) * yd * (1 - zd) + f[l001] * (1 - xd) * (1 - yd) * zd +
           f[l110] * xd * yd * (1 - zd) + f[l011] * (1 - xd) * yd * zd + f[l101] * xd * (1 - yd) * (1 - zd) +
           f[l111] * xd * yd * zd;
}

/* ------------------------------------------------------------------ validation */

static int check_inputs(const og_grid* g, const double* vel, int64_t nsrc, const double* coordsrc,
                        const int64_t* rec_off, const double* coordrec, const og_params* p) {
    if (g->ndim < 2 || g->ndim > 3) return fail("AssertionError: 2<=ndims(vel)<=3");
    if (g->kind == OG_SPH3D && g->ndim != 3) return fail("bad grid kind");
    for (int d = 0; d < g->ndim; d++)
        if (g->n[d] < 1) return fail("AssertionError: all(grsize.>0)");
    int64_t N = npts(g);
    for (int64_t i = 0; i < N; i++)
        if (!(vel[i] > 0.0)) return fail("AssertionError: all(vel.>0.0)");
    if (p->radiussmooth < 0) return fail("AssertionError: radiussmoothgradsrc>=0");
    if (p->smoothkern < 0) return fail("AssertionError: smoothgradkern>=0");
    if (p->smoothkern > 0 && p->smoothkern % 2 == 0) return fail("ExtraParams: smoothgradkern must be an odd number");
    if (p->refine && (p->df < 1 || p->noderadius < 0)) return fail("bad grid refinement parameters");
    /* checksrcrecposition (eikchecks.jl:4-50) */
    for (int d = 0; d < g->ndim; d++) {
        double lo = coord(g, d, 0), hi = coord(g, d, g->n[d] - 1);
        if (!is_sph(g)) {
            lo = fma(0.0, g->d[0], g->o[d]);
            hi = fma((double)(g->n[d] - 1), g->d[0], g->o[d]);
        }
        for (int64_t s = 0; s < nsrc; s++) {
            double x = coordsrc[s * g->ndim + d];
            if (!(lo <= x && x <= hi)) return fail("AssertionError: source %lld outside the grid", (long long)s);
        }
        for (int64_t r = 0; r < rec_off[nsrc]; r++) {
            double x = coordrec[r * g->ndim + d];
            if (!(lo <= x && x <= hi)) return fail("AssertionError: receiver %lld outside the grid", (long long)r);
        }
    }
    return 0;
}

/* distribsrcs (eikutils.jl:171-190): contiguous chunks, remainder to the first chunks */
static int64_t distribsrcs(int64_t nsrc, int64_t nw, int64_t* lo, int64_t* hi) {
    if (nw < 1) nw = 1;
    int64_t nch = nsrc >= nw ? nw : nsrc;
    int64_t base = nsrc >= nw ? nsrc / nw : 1, rest = nsrc >= nw ? nsrc % nw : 0;
    int64_t s = 0;
    for (int64_t c = 0; c < nch; c++) {
        int64_t sz = base + (c < rest ? 1 : 0);
        lo[c] = s;
        hi[c] = s + sz;
        s += sz;
    }
    return nch;
}

/* The parameters a call runs with: as given, except that Grid2DSphere keeps the FP64 stencil under
 * the mixed precision (the engine's rule: its adjoint amplifies rounding-level changes ~1e4-fold). */
static og_params effective_params(const og_grid* g, const og_params* p) {
    og_params e = *p;
    if (e.precision == OG_PREC_MIXED && g->kind == OG_SPH2D) e.precision = OG_PREC_FP64;
    return e;
}
static int sph3d_refine_unsupported(const og_grid* g, const og_params* p) {
    return g->kind == OG_SPH3D && p->refine;
}

/* ------------------------------------------------------------------ forward API */

/* ttforwsomesrc (eikmainfwd.jl:152-199) */
static int fwd_chunk(const og_grid* g, const double* vel, int64_t s0, int64_t s1, const double* coordsrc,
                     const int64_t* rec_off, const double* coordrec, const og_params* p, double* ttpicks,
                     double* ttgrid) {
    const int64_t N = npts(g);
    fmm_t F;
    if (fmm_alloc(&F, N)) return fail("out of memory");
    int64_t* cinit = (int64_t*)malloc((size_t)N * sizeof(int64_t));
    int rc = cinit ? 0 : fail("out of memory");
    for (int64_t s = s0; s < s1 && rc == 0; s++) {
        srcbox_t sb;
        finegrid_t fg;
        memset(&fg, 0, sizeof(fg));
        int64_t nci = 0;
        rc = ttfmm_hiord(g, vel, coordsrc + s * g->ndim, p, &F, NULL, cinit, &nci, &sb, &fg);
        finegrid_free(&fg);
        if (rc) break;
        for (int64_t r = rec_off[s]; r < rec_off[s + 1]; r++)
            ttpicks[r] = interp(g, F.tt, coordrec + r * g->ndim, NULL);
        if (ttgrid) memcpy(ttgrid + s * N, F.tt, (size_t)N * sizeof(double));
    }
    free(cinit);
    fmm_free(&F);
    return rc;
}

int og_traveltime(const og_grid* g, const double* vel, int64_t nsrc, const double* coordsrc,
                  const int64_t* rec_off, const double* coordrec, const og_params* p_in, double* ttpicks,
                  double* ttgrid) {
    const og_params pe = effective_params(g, p_in);
    const og_params* p = &pe;
    if (check_inputs(g, vel, nsrc, coordsrc, rec_off, coordrec, p)) return -1;
    if (sph3d_refine_unsupported(g, p))
        return fail("createfinegrid(): spherical coordinates in 3D still work in progress...");
    int64_t nt = p->nthreads > 1 ? p->nthreads : 1;
    int64_t* lo = (int64_t*)malloc((size_t)(nt + 1) * sizeof(int64_t));
    int64_t* hi = (int64_t*)malloc((size_t)(nt + 1) * sizeof(int64_t));
    int64_t nch = distribsrcs(nsrc, nt, lo, hi);
    int rc = 0;
#pragma omp parallel for schedule(static, 1) num_threads((int)nt) if (nt > 1) reduction(| : rc)
    for (int64_t c = 0; c < nch; c++)
        rc |= fwd_chunk(g, vel, lo[c], hi[c], coordsrc, rec_off, coordrec, p, ttpicks, ttgrid) ? 1 : 0;
    free(lo);
    free(hi);
    return rc ? -1 : 0;
}

int og_misfit(const og_grid* g, const double* vel, int64_t nsrc, const double* coordsrc,
              const int64_t* rec_off, const double* coordrec, const double* pickobs, const double* stdobs,
              const og_params* p, double* misf) {
    int64_t nr = rec_off[nsrc];
    double* pk = (double*)malloc((size_t)(nr > 0 ? nr : 1) * sizeof(double));
    if (!pk) return fail("out of memory");
    int rc = og_traveltime(g, vel, nsrc, coordsrc, rec_off, coordrec, p, pk, NULL);
    if (rc == 0) {
        double m = 0.0;
        for (int64_t s = 0; s < nsrc; s++) {
            double ms = 0.0;
            for (int64_t r = rec_off[s]; r < rec_off[s + 1]; r++) {
                double dd = pk[r] - pickobs[r];
                ms += (dd * dd) / (stdobs[r] * stdobs[r]);
            }
            m += ms;
        }
        *misf = 0.5 * m;
    }
    free(pk);
    return rc;
}

/* ------------------------------------------------------------------ adjoint pieces */

/* calc2diagDttD (eikmaingrad.jl:751-868) for all dims, and the transposed-sum lhs of
 * solveadjointeq (:488-535) stored row-wise: row i of A (non-source rows, reduced columns). */
typedef struct {
    int64_t nrows, nsrc;
    int64_t* rp;    /* [nrows+1] */
    int64_t* col;   /* reduced column indices (<= row) */
    double* val;
    int64_t* srp;   /* source part: [nrows+1] */
    int64_t* scol;  /* source point index */
    double* sval;
} lhs_t;

static void lhs_free(lhs_t* L) {
    free(L->rp);
    free(L->col);
    free(L->val);
    free(L->srp);
    free(L->scol);
    free(L->sval);
}

static int build_lhs(const adj_t* A, const double* tt, lhs_t* L) {
    const int nd = A->nd;
    const int64_t nr = A->nrows, ns = A->nsrc;
    L->nrows = nr;
    L->nsrc = ns;
    L->rp = (int64_t*)malloc(((size_t)nr + 1) * sizeof(int64_t));
    L->srp = (int64_t*)malloc(((size_t)nr + 1) * sizeof(int64_t));
    L->col = (int64_t*)malloc(((size_t)nr * 7 + 1) * sizeof(int64_t));
    L->val = (double*)malloc(((size_t)nr * 7 + 1) * sizeof(double));
    L->scol = (int64_t*)malloc(((size_t)nr * 7 + 1) * sizeof(int64_t));
    L->sval = (double*)malloc(((size_t)nr * 7 + 1) * sizeof(double));
    if (!L->rp || !L->srp || !L->col || !L->val || !L->scol || !L->sval) return fail("out of memory");
    int64_t k = 0, ks = 0;
    L->rp[0] = 0;
    L->srp[0] = 0;
    for (int64_t i = 0; i < nr; i++) {
        int64_t c0 = k, s0 = ks;
        for (int d = 0; d < nd; d++) {
            double tmp1 = 0.0;
            for (int64_t l = A->iptr[d][i]; l < A->iptr[d][i + 1]; l++) tmp1 += A->vv[d][l] * tt[A->jc[d][l]];
            for (int64_t l = A->iptr[d][i]; l < A->iptr[d][i + 1]; l++) {
                int64_t j = A->jc[d][l];
                double v = 2.0 * tmp1 * A->vv[d][l];
                if (j >= ns) { /* tmplhs .+= ... : add into an existing (row,col) entry */
                    int64_t jr = j - ns, q;
                    for (q = c0; q < k; q++)
                        if (L->col[q] == jr) break;
                    if (q < k) {
                        L->val[q] = L->val[q] + v;
                    } else {
                        L->col[k] = jr;
                        L->val[k] = 0.0 + v;
                        k++;
                    }
                } else {
                    int64_t q;
                    for (q = s0; q < ks; q++)
                        if (L->scol[q] == j) break;
                    if (q < ks) {
                        L->sval[q] = L->sval[q] + v;
                    } else {
                        L->scol[ks] = j;
                        L->sval[ks] = v;
                        ks++;
                    }
                }
            }
        }
        L->rp[i + 1] = k;
        L->srp[i + 1] = ks;
    }
    return 0;
}

/* UpperTriangular(A^T) \ rhs by column-oriented back substitution, nrhs columns (leading dim nr) */
static int backsolve(const lhs_t* L, double* x, int64_t nrhs) {
    const int64_t nr = L->nrows;
    for (int64_t k = 0; k < nrhs; k++) {
        double* b = x + k * nr;
        for (int64_t j = nr - 1; j >= 0; j--) {
            double dg = 0.0;
            int found = 0;
            for (int64_t q = L->rp[j]; q < L->rp[j + 1]; q++)
                if (L->col[q] == j) {
                    dg = L->val[q];
                    found = 1;
                }
            if (!found || dg == 0.0) return fail("SingularException(%lld)", (long long)(j + 1));
            double bj = (b[j] /= dg);
            /* remaining entries of column j of A^T in decreasing row order */
            int64_t cnt = L->rp[j + 1] - L->rp[j];
            int64_t ord[8];
            int64_t m = 0;
            for (int64_t q = L->rp[j]; q < L->rp[j + 1]; q++)
                if (L->col[q] != j) ord[m++] = q;
            (void)cnt;
            for (int64_t a = 1; a < m; a++)
                for (int64_t c = a; c > 0 && L->col[ord[c - 1]] < L->col[ord[c]]; c--) {
                    int64_t t = ord[c];
                    ord[c] = ord[c - 1];
                    ord[c - 1] = t;
                }
            for (int64_t q = 0; q < m; q++) b[L->col[ord[q]]] -= bj * L->val[ord[q]];
        }
    }
    return 0;
}

/* partderivttsrcpos (eikmaingrad.jl:1235-1273) */
static void partderivttsrcpos(int nd, const double* xyzpt, const double* src, double vel, double hgrid,
                              double* der) {
    double xs[3];
    for (int d = 0; d < nd; d++) {
        double al = fmod(src[d], hgrid);
        if (al < 0) al += hgrid; /* Julia mod: sign of the divisor */
        if (al <= 2.220446049250313e-16) {
            xs[d] = src[d] + 1e-6 * hgrid;
            warn_hit();
        } else {
            xs[d] = src[d];
        }
    }
    double s2 = 0.0;
    for (int d = 0; d < nd; d++) s2 += (xyzpt[d] - xs[d]) * (xyzpt[d] - xs[d]);
    double denom = vel * sqrt(s2);
    if (denom == 0.0) {
        warn_hit();
        for (int d = 0; d < nd; d++) der[d] = 0.0;
        return;
    }
    for (int d = 0; d < nd; d++) der[d] = -(xyzpt[d] - xs[d]) / denom;
}

/* smoothgradaroundsrc! (utils.jl:121-338), including the 3D k-range taken from the i index */
static void taper(const og_grid* g, double* grad, const double* src, int R) {
    if (R == 0) return;
    int64_t corn[8][3];
    int nc = findenclosingbox(g, src, corn);
    const int nd = g->ndim;
    int64_t imin, imax, jmin, jmax, kmin = 0, kmax = 0;
    double rmax;
    if (nc == 1) {
        imin = corn[0][0] - R;
        imax = corn[0][0] + R;
        jmin = corn[0][1] - R;
        jmax = corn[0][1] + R;
        if (nd == 3) {
            kmin = corn[0][2] - R;
            kmax = corn[0][2] + R;
        }
    } else {
        imin = corn[0][0] - R;
        imax = corn[1][0] + R;
        jmin = corn[0][1] - R;
        jmax = corn[2][1] + R;
        if (nd == 3 && g->kind == OG_SPH3D) { /* no reference: the box's own k range */
            kmin = corn[0][2] - R;
            kmax = corn[nc - 1][2] + R;
        } else if (nd == 3) {
            kmin = corn[0][0] - R;
            kmax = corn[3][0] + R;
        }
    }
    if (g->kind == OG_SPH3D) { /* chord distance, as the 3D spherical forward measures it */
        rmax = R * g->d[0];
        for (int64_t k = kmin; k <= kmax; k++)
            for (int64_t j = jmin; j <= jmax; j++)
                for (int64_t i = imin; i <= imax; i++) {
                    if (i < 0 || i > g->n[0] - 1 || j < 0 || j > g->n[1] - 1 || k < 0 || k > g->n[2] - 1) continue;
                    const double a[3] = {coord(g, 0, i), coord(g, 1, j), coord(g, 2, k)};
                    const double r = distance2points(g, a, src);
                    if (r <= rmax) grad[i + g->n[0] * (j + g->n[1] * k)] *= r / rmax;
                }
        return;
    }
    if (g->kind == OG_SPH2D) {
        rmax = R * g->d[0];
        double xs = src[0] * cos(deg2rad(src[1])), ys = src[0] * sin(deg2rad(src[1]));
        for (int64_t j = jmin; j <= jmax; j++)
            for (int64_t i = imin; i <= imax; i++) {
                if (i < 0 || i > g->n[0] - 1 || j < 0 || j > g->n[1] - 1) continue;
                double r = coord(g, 0, i), th = coord(g, 1, j);
                double xc = r * cos(deg2rad(th)), yc = r * sin(deg2rad(th));
                double ra = sqrt((xc - xs) * (xc - xs) + (ys - yc) * (ys - yc));
                if (ra <= rmax) grad[i + g->n[0] * j] *= ra / rmax;
            }
        return;
    }
    rmax = R * g->d[0];
    if (nd == 2) {
        for (int64_t j = jmin; j <= jmax; j++)
            for (int64_t i = imin; i <= imax; i++) {
                if (i < 0 || i > g->n[0] - 1 || j < 0 || j > g->n[1] - 1) continue;
                double xc = coord(g, 0, i), yc = coord(g, 1, j);
                double r = sqrt((xc - src[0]) * (xc - src[0]) + (src[1] - yc) * (src[1] - yc));
                if (r <= rmax) grad[i + g->n[0] * j] *= r / rmax;
            }
        return;
    }
    for (int64_t k = kmin; k <= kmax; k++)
        for (int64_t j = jmin; j <= jmax; j++)
            for (int64_t i = imin; i <= imax; i++) {
                if (i < 0 || i > g->n[0] - 1 || j < 0 || j > g->n[1] - 1 || k < 0 || k > g->n[2] - 1) continue;
                double xc = coord(g, 0, i), yc = coord(g, 1, j), zc = coord(g, 2, k);
                double r = sqrt((xc - src[0]) * (xc - src[0]) + (src[1] - yc) * (src[1] - yc) +
                                (src[2] - zc) * (src[2] - zc));
                if (r <= rmax) grad[i + g->n[0] * (j + g->n[1] * k)] *= r / rmax;
            }
}

/* calcgrads_singlesrc! (eikmaingrad.jl:274-484) with computestuff_finegrid! (:540-632) and
 * calc_duh_dva_finegrid! (:636-741).  gradvel1 is the caller's persistent per-chunk buffer. */
static int grads_singlesrc(const og_grid* g, const double* vel, const double* src, const double* recs,
                           int64_t nrec, const double* pobs, const double* sobs, int whichgrad,
                           const og_params* p, fmm_t* F, adj_t* A, const int64_t* cinit, int64_t ncinit,
                           const srcbox_t* sb, finegrid_t* fg, double* gradvel1, double* gradsrc1,
                           double* misf1) {
    const int nd = g->ndim;
    const int64_t ns = A->nsrc, nr = A->nrows;
    int rc = 0;
    /* P (calcprojttfmmord, :878-937), misfit, rhs */
    double* fact = (double*)malloc((size_t)(nrec > 0 ? nrec : 1) * sizeof(double));
    double* rhs = (double*)calloc((size_t)(nr > 0 ? nr : 1), sizeof(double));
    double* dpsi_dus = (double*)calloc((size_t)(ns > 0 ? ns : 1), sizeof(double));
    double* dpsiup = (double*)calloc((size_t)(ns > 0 ? ns : 1), sizeof(double));
    lhs_t L;
    memset(&L, 0, sizeof(L));
    double* t2va = NULL;   /* T2Va per coarse source point (refinement) */
    double* t1va = NULL;   /* T1Va per coarse source point (refinement) */
    double* dusdsr = NULL; /* ns x nd */
    if (!fact || !rhs || !dpsi_dus || !dpsiup) {
        rc = fail("out of memory");
        goto out;
    }
    double ms = 0.0;
    for (int64_t r = 0; r < nrec; r++) {
        interp_t ic;
        interp(g, NULL, recs + r * nd, &ic);
        /* P*tt_fmmord: CSC matvec, accumulates by increasing FMM column */
        int64_t ord[8];
        for (int c = 0; c < ic.ncoe; c++) ord[c] = c;
        for (int a = 1; a < ic.ncoe; a++)
            for (int b = a; b > 0 && A->lgrid2fmm[ic.idx[ord[b - 1]]] > A->lgrid2fmm[ic.idx[ord[b]]]; b--) {
                int64_t t = ord[b];
                ord[b] = ord[b - 1];
                ord[b - 1] = t;
            }
        double pt = 0.0;
        for (int c = 0; c < ic.ncoe; c++) pt += ic.cf[ord[c]] * A->ttfmm[A->lgrid2fmm[ic.idx[ord[c]]]];
        double pd = pt - pobs[r];
        ms += (pd * pd) / (sobs[r] * sobs[r]);
        fact[r] = pd / (sobs[r] * sobs[r]);
    }
    *misf1 = ms;
    for (int64_t r = 0; r < nrec; r++) {
        interp_t ic;
        interp(g, NULL, recs + r * nd, &ic);
        for (int c = 0; c < ic.ncoe; c++) {
            int64_t j = A->lgrid2fmm[ic.idx[c]];
            if (j >= ns)
                rhs[j - ns] += -ic.cf[c] * fact[r];
            else
                dpsi_dus[j] += ic.cf[c] * fact[r];
        }
    }
    if ((rc = build_lhs(A, A->ttfmm, &L))) goto out;
    if ((rc = backsolve(&L, rhs, 1))) goto out; /* rhs now holds lambda */
    for (int64_t i = 0; i < nr; i++)
        for (int64_t q = L.srp[i]; q < L.srp[i + 1]; q++) dpsiup[L.scol[q]] += L.sval[q] * rhs[i];

    /* fine-grid products (refinement) */
    if (p->refine) {
        adj_t* Af = &fg->Af;
        const int64_t nacc = Af->lastcomputed, nsf = Af->nsrc, nrf = Af->nrows;
        int64_t* hidx = (int64_t*)malloc((size_t)(nacc > 0 ? nacc : 1) * sizeof(int64_t));
        int64_t nh = 0;
        for (int64_t q = 0; q < nacc; q++)
            if (Af->onh[q]) hidx[nh++] = q;
        if (nh != ns) {
            free(hidx);
            rc = fail("inconsistent fine/coarse source points (%lld vs %lld)", (long long)nh, (long long)ns);
            goto out;
        }
        lhs_t Lf;
        memset(&Lf, 0, sizeof(Lf));
        double* Lam = (double*)calloc((size_t)(nrf > 0 ? nrf : 1) * (size_t)(nh > 0 ? nh : 1), sizeof(double));
        double* duhdtau = (double*)calloc((size_t)(nh > 0 ? nh : 1) * (size_t)(nsf > 0 ? nsf : 1), sizeof(double));
        double* duhdva = (double*)calloc((size_t)(nh > 0 ? nh : 1) * (size_t)(nh > 0 ? nh : 1), sizeof(double));
        double* duhdwq_src = (double*)calloc((size_t)(nh > 0 ? nh : 1) * (size_t)(nsf > 0 ? nsf : 1), sizeof(double));
        if (!Lam || !duhdtau || !duhdva || !duhdwq_src) {
            rc = fail("out of memory");
            goto fine_out;
        }
        for (int64_t l = 0; l < nh; l++)
            if (hidx[l] >= nsf) Lam[l * nrf + (hidx[l] - nsf)] = -1.0; /* rhs = -H_B^T */
        if ((rc = build_lhs(Af, Af->ttfmm, &Lf))) goto fine_out;
        if ((rc = backsolve(&Lf, Lam, nh))) goto fine_out;
        /* du_h/dtau_s = Lambda^T * dgi/dtau_s (Nh x nsf) */
        for (int64_t h = 0; h < nh; h++)
            for (int64_t i = 0; i < nrf; i++) {
                double lv = Lam[h * nrf + i];
                if (lv == 0.0) continue;
                for (int64_t q = Lf.srp[i]; q < Lf.srp[i + 1]; q++) duhdtau[h * nsf + Lf.scol[q]] += lv * Lf.sval[q];
            }
        /* duh_dwq source columns: T2Wa + T1Wa (dtaus_dwa = -dist/vel^2 diagonal) */
        for (int64_t h = 0; h < nh; h++)
            for (int64_t q = 0; q < nsf; q++) {
                double dd = -fg->sbf.distcorn[q] * (1.0 / (fg->sbf.velcorn[q] * fg->sbf.velcorn[q]));
                double v = duhdtau[h * nsf + q] * dd;
                if (hidx[h] == q) v += dd;
                duhdwq_src[h * nsf + q] = 0.0 + v;
            }
        /* duh_dva = duh_dwq * dwq_dva; dwq_dva[p,j] = 1 iff NN(p) is H-point j's coarse node */
        for (int64_t j = 0; j < nh; j++) {
            int64_t qorig = fg->R.nn_idxcoarse[Af->lfmm2grid[hidx[j]]];
            for (int64_t pp = 0; pp < nacc; pp++) {
                int64_t porig = Af->lfmm2grid[pp];
                if (fg->R.nn_idxcoarse[porig] != qorig) continue;
                for (int64_t h = 0; h < nh; h++) {
                    double w;
                    if (pp < nsf) {
                        w = duhdwq_src[h * nsf + pp];
                    } else {
                        double wv = fg->R.velfine[porig];
                        w = 2.0 * Lam[h * nrf + (pp - nsf)] / (wv * wv * wv);
                    }
                    duhdva[h * nh + j] += w;
                }
            }
        }
        /* T2Va then T1Va at coarse source point a (:458-479): duh_dva^T dpsiup, duh_dva^T dpsi_dus.
         * They are added after T2Vb below, so keep them in t2va/t1va until then. */
        t2va = (double*)calloc((size_t)nh, sizeof(double));
        t1va = (double*)calloc((size_t)nh, sizeof(double));
        if (!t2va || !t1va) {
            rc = fail("out of memory");
            goto fine_out;
        }
        for (int64_t a = 0; a < nh; a++) {
            double t2 = 0.0, t1 = 0.0;
            for (int64_t h = 0; h < nh; h++) {
                t2 += duhdva[h * nh + a] * dpsiup[h];
                t1 += duhdva[h * nh + a] * dpsi_dus[h];
            }
            t2va[a] = t2;
            t1va[a] = t1;
        }
        if (whichgrad != OG_GRADVEL) { /* calc_dus_dsr_finegrid (:1184-1231) */
            dusdsr = (double*)calloc((size_t)nh * nd, sizeof(double));
            double* dtaus = (double*)calloc((size_t)(nsf > 0 ? nsf : 1) * nd, sizeof(double));
            if (!dusdsr || !dtaus) {
                free(dtaus);
                rc = fail("out of memory");
                goto fine_out;
            }
            for (int64_t q = 0; q < nsf; q++) {
                double xyz[3];
                for (int d = 0; d < nd; d++) xyz[d] = coord(&fg->R.gf, d, fg->sbf.ijk[q][d]);
                partderivttsrcpos(nd, xyz, src, fg->R.velfine[lin(&fg->R.gf, fg->sbf.ijk[q])], fg->R.gf.d[0],
                                  dtaus + q * nd);
            }
            for (int64_t h = 0; h < nh; h++)
                for (int d = 0; d < nd; d++) {
                    double t1 = 0.0;
                    for (int64_t q = 0; q < nsf; q++) t1 += duhdtau[h * nsf + q] * dtaus[q * nd + d];
                    double t2 = (hidx[h] < nsf) ? dtaus[hidx[h] * nd + d] : 0.0;
                    dusdsr[h * nd + d] = t1 + t2;
                }
            free(dtaus);
        }
    fine_out:
        free(hidx);
        free(Lam);
        free(duhdtau);
        free(duhdva);
        free(duhdwq_src);
        lhs_free(&Lf);
        if (rc) goto out;
    }

    if (whichgrad != OG_GRADVEL) {
        if (is_sph(g)) {
            rc = fail("gradsrcloc is not available on spherical grids (grd.hgrid undefined)");
            goto out;
        }
        if (!p->refine) { /* calc_dus_dsr (:1140-1180) */
            dusdsr = (double*)calloc((size_t)(ns > 0 ? ns : 1) * nd, sizeof(double));
            for (int64_t q = 0; q < ns; q++) {
                double xyz[3];
                for (int d = 0; d < nd; d++) xyz[d] = coord(g, d, sb->ijk[q][d]);
                partderivttsrcpos(nd, xyz, src, vel[lin(g, sb->ijk[q])], g->d[0], dusdsr + q * nd);
            }
        }
        for (int d = 0; d < nd; d++) {
            double a = 0.0, b = 0.0;
            for (int64_t q = 0; q < ns; q++) {
                a += dusdsr[q * nd + d] * dpsiup[q];
                b += dusdsr[q * nd + d] * dpsi_dus[q];
            }
            gradsrc1[d] = a + b;
        }
    }
    if (whichgrad != OG_GRADSRCLOC) {
        /* T2Vb (:417-423) overwrites every non-source node */
        for (int64_t i = 0; i < nr; i++) {
            int64_t io = A->lfmm2grid[i + ns];
            double v = vel[io];
            gradvel1[io] = 2.0 * rhs[i] / (v * v * v);
        }
        if (!p->refine) { /* T2Va, T1Va with dus_dva = -diag(dist/vel^2) (:428-479) */
            for (int64_t q = 0; q < ns; q++) {
                double dd = -sb->distcorn[q] * (1.0 / (sb->velcorn[q] * sb->velcorn[q]));
                gradvel1[cinit[q]] += dd * dpsiup[q];
            }
            for (int64_t q = 0; q < ns; q++) {
                double dd = -sb->distcorn[q] * (1.0 / (sb->velcorn[q] * sb->velcorn[q]));
                gradvel1[cinit[q]] += dd * dpsi_dus[q];
            }
        } else {
            for (int64_t q = 0; q < ns; q++) gradvel1[cinit[q]] += t2va[q];
            for (int64_t q = 0; q < ns; q++) gradvel1[cinit[q]] += t1va[q];
        }
    }
out:
    free(t2va);
    free(t1va);
    free(dusdsr);
    free(fact);
    free(rhs);
    free(dpsi_dus);
    free(dpsiup);
    lhs_free(&L);
    return rc;
}

/* calcgradsomesrc (eikmaingrad.jl:174-258) */
static int grad_chunk(const og_grid* g, const double* vel, int64_t s0, int64_t s1, const double* coordsrc,
                      const int64_t* rec_off, const double* coordrec, const double* pickobs,
                      const double* stdobs, int whichgrad, const og_params* p, double* gradall, double* gradsrc,
                      double* misf) {
    const int64_t N = npts(g);
    const int nd = g->ndim;
    fmm_t F;
    adj_t A;
    if (fmm_alloc(&F, N) || adj_alloc(&A, N, nd)) return fail("out of memory");
    double* gradvel1 = (double*)calloc((size_t)N, sizeof(double));
    int64_t* cinit = (int64_t*)malloc((size_t)N * sizeof(int64_t));
    int rc = (gradvel1 && cinit) ? 0 : fail("out of memory");
    double tot = 0.0;
    for (int64_t i = 0; i < N; i++) gradall[i] = 0.0;
    for (int64_t s = s0; s < s1 && rc == 0; s++) {
        srcbox_t sb;
        finegrid_t fg;
        memset(&fg, 0, sizeof(fg));
        int64_t nci = 0;
        const double* src = coordsrc + s * nd;
        rc = ttfmm_hiord(g, vel, src, p, &F, &A, cinit, &nci, &sb, &fg);
        if (rc == 0) {
            double m1 = 0.0, gs[3] = {0, 0, 0};
            rc = grads_singlesrc(g, vel, src, coordrec + rec_off[s] * nd, rec_off[s + 1] - rec_off[s],
                                 pickobs + rec_off[s], stdobs + rec_off[s], whichgrad, p, &F, &A, cinit, nci, &sb,
                                 &fg, gradvel1, gs, &m1);
            if (rc == 0) {
                tot += 0.5 * m1;
                if (whichgrad != OG_GRADVEL && gradsrc)
                    for (int d = 0; d < nd; d++) gradsrc[s * nd + d] = gs[d];
                if (whichgrad != OG_GRADSRCLOC) {
                    taper(g, gradvel1, src, p->radiussmooth);
                    for (int64_t i = 0; i < N; i++) gradall[i] += gradvel1[i];
                }
            }
        }
        finegrid_free(&fg);
    }
    *misf = tot;
    free(gradvel1);
    free(cinit);
    fmm_free(&F);
    adj_free(&A);
    return rc;
}

int og_smoothgradient(const og_grid* g, int32_t l, const double* in, double* out) {
    /* smoothgradient + gaussian_kernel + smoothimg (utils.jl:22-96): zero-padded convolution */
    if (l <= 0 || l % 2 == 0) return fail("gaussian_kernel(): l must be odd");
    int nd0 = g->ndim;
    int64_t sz[3] = {g->n[0], g->n[1], nd0 == 3 ? g->n[2] : 1};
    int64_t keep[3];
    int nd = 0;
    for (int d = 0; d < nd0; d++)
        if (sz[d] != 1) keep[nd++] = d;
    int m = l / 2;
    double s = ((double)l / 2.0) - sqrt((double)l / 2.0);
    int64_t nk = 1;
    for (int d = 0; d < nd; d++) nk *= l;
    double* ker = (double*)malloc((size_t)nk * sizeof(double));
    double ksum = 0.0;
    for (int64_t q = 0; q < nk; q++) {
        int64_t r = q;
        double r2 = 0.0;
        for (int d = 0; d < nd; d++) {
            int64_t pos = r % l - m;
            r /= l;
            r2 += (double)(pos * pos);
        }
        double rr = sqrt(r2);
        ker[q] = exp(-(rr * rr) / (2.0 * s));
        ksum += ker[q];
    }
    for (int64_t q = 0; q < nk; q++) ker[q] /= ksum;
    int64_t N = npts(g);
    for (int64_t i = 0; i < N; i++) {
        int64_t ijk[3] = {0, 0, 0};
        unlin(g, i, ijk);
        double acc = 0.0;
        for (int64_t q = 0; q < nk; q++) {
            int64_t r = q, c[3] = {ijk[0], ijk[1], ijk[2]};
            int ok = 1;
            for (int d = 0; d < nd; d++) {
                int64_t off = r % l - m;
                r /= l;
                c[keep[d]] += off;
                if (c[keep[d]] < 0 || c[keep[d]] > sz[keep[d]] - 1) ok = 0;
            }
            if (ok) acc += in[lin(g, c)] * ker[q];
        }
        out[i] = acc;
    }
    free(ker);
    return 0;
}

int og_gradient(const og_grid* g, const double* vel, int64_t nsrc, const double* coordsrc,
                const int64_t* rec_off, const double* coordrec, const double* pickobs, const double* stdobs,
                int whichgrad, const og_params* p_in, double* gradvel, double* gradsrc, double* misf) {
    const og_params pe = effective_params(g, p_in);
    const og_params* p = &pe;
    if (sph3d_refine_unsupported(g, p))
        return fail("createfinegrid(): spherical coordinates in 3D still work in progress...");
    if (whichgrad < 1 || whichgrad > 3)
        return fail("AssertionError: Symbol 'whichgrad' not in (:gradvel, :gradsrcloc, :gradvelandsrcloc)");
    if (check_inputs(g, vel, nsrc, coordsrc, rec_off, coordrec, p)) return -1;
    for (int64_t r = 0; r < rec_off[nsrc]; r++)
        if (!(stdobs[r] != 0.0)) return fail("stdobs must be non-zero");
    /* OG_SPH3D: the reference raises (createsparsederivativematrices!, eikmainfwd.jl:885-887); this
     * build differentiates its own 3D spherical forward, :gradvel only (as on 2D spheres) */
    if (is_sph(g) && whichgrad != OG_GRADVEL) return fail("gradsrcloc is not available on spherical grids");
    const int64_t N = npts(g);
    int64_t nt = p->nthreads > 1 ? p->nthreads : 1;
    int64_t lo[256], hi[256];
    if (nt > 256) nt = 256;
    int64_t nch = distribsrcs(nsrc, nt, lo, hi);
    double* parts = (double*)calloc((size_t)nch * (size_t)N, sizeof(double));
    double* pm = (double*)calloc((size_t)nch, sizeof(double));
    double* gs = (double*)calloc((size_t)(nsrc > 0 ? nsrc : 1) * g->ndim, sizeof(double));
    if (!parts || !pm || !gs) return fail("out of memory");
    int rc = 0;
#pragma omp parallel for schedule(static, 1) num_threads((int)nt) if (nt > 1) reduction(| : rc)
    for (int64_t c = 0; c < nch; c++)
        rc |= grad_chunk(g, vel, lo[c], hi[c], coordsrc, rec_off, coordrec, pickobs, stdobs, whichgrad, p,
                         parts + c * N, gs, pm + c)
                  ? 1
                  : 0;
    if (rc == 0) {
        double m = 0.0;
        for (int64_t c = 0; c < nch; c++) m += pm[c];
        *misf = m;
        if (gradvel) {
            for (int64_t i = 0; i < N; i++) {
                double a = 0.0;
                for (int64_t c = 0; c < nch; c++) a += parts[c * N + i];
                gradvel[i] = a;
            }
            if (p->smoothkern > 0) {
                double* tmp = (double*)malloc((size_t)N * sizeof(double));
                og_smoothgradient(g, p->smoothkern, gradvel, tmp);
                memcpy(gradvel, tmp, (size_t)N * sizeof(double));
                free(tmp);
            }
        }
        if (gradsrc) memcpy(gradsrc, gs, (size_t)nsrc * g->ndim * sizeof(double));
    }
    free(parts);
    free(pm);
    free(gs);
    return rc ? -1 : 0;
}

/* ------------------------------------------------------------------ test hooks */

int64_t og_fmm_record(const og_grid* g, const double* vel, const double* src, const og_params* p_in, double* tt,
                      int64_t* order, int8_t* codes, int64_t* nsrcpts) {
    const og_params pe = effective_params(g, p_in);
    const og_params* p = &pe;
    const int64_t N = npts(g);
    fmm_t F;
    adj_t A;
    if (fmm_alloc(&F, N) || adj_alloc(&A, N, g->ndim)) return fail("out of memory");
    int64_t* cinit = (int64_t*)malloc((size_t)N * sizeof(int64_t));
    srcbox_t sb;
    finegrid_t fg;
    memset(&fg, 0, sizeof(fg));
    int64_t nci = 0;
    int rc = ttfmm_hiord(g, vel, src, p, &F, &A, cinit, &nci, &sb, &fg);
    int64_t nacc = -1;
    if (rc == 0) {
        memcpy(tt, F.tt, (size_t)N * sizeof(double));
        if (order) memcpy(order, A.lfmm2grid, (size_t)N * sizeof(int64_t));
        if (codes) memcpy(codes, A.code, (size_t)N * g->ndim);
        if (nsrcpts) *nsrcpts = nci;
        nacc = A.lastcomputed;
    }
    finegrid_free(&fg);
    free(cinit);
    fmm_free(&F);
    adj_free(&A);
    return nacc;
}

int og_fmm_bruteforce(const og_grid* g, const double* vel, const double* src, double* tt, int64_t* order) {
    /* O(N^2) FMM: scan for the minimum narrow-band value, pop every node holding it (increasing
     * grid index), then update neighbours of the popped set (eikmainfwd.jl:630-814 semantics). */
    const int64_t N = npts(g);
    const int nd = g->ndim, nnb = 2 * nd;
    uint8_t* st = (uint8_t*)calloc((size_t)N, 1);
    double* key = (double*)malloc((size_t)N * sizeof(double));
    int64_t* bat = (int64_t*)malloc((size_t)N * sizeof(int64_t));
    fmm_t F;
    F.n = N;
    F.tt = tt;
    F.st = st;
    for (int64_t i = 0; i < N; i++) tt[i] = HUGEV;
    srcbox_t sb;
    sourceboxloctt(g, vel, src, &F, &sb);
    int64_t cnt = 0;
    for (int c = 0; c < sb.ncorn; c++) order[cnt++] = lin(g, sb.ijk[c]);
    for (int64_t i = cnt; i < N; i++) order[i] = -1;
    int32_t idD[3];
    /* initial band */
    for (int c = 0; c < sb.ncorn; c++) {
        for (int ne = 0; ne < nnb; ne++) {
            int64_t q[3] = {sb.ijk[c][0], sb.ijk[c][1], sb.ijk[c][2]};
            for (int d = 0; d < nd; d++) q[d] += (nd == 2) ? neigh2[ne][d] : neigh3[ne][d];
            int oob = 0;
            for (int d = 0; d < nd; d++)
                if (q[d] < 0 || q[d] > g->n[d] - 1) oob = 1;
            if (oob) continue;
            int64_t l = lin(g, q);
            if (st[l] == 0) {
                key[l] = stencil(g, vel, tt, st, q, idD);
                st[l] = 1;
            }
        }
    }
    for (;;) {
        double mn = HUGEV;
        int any = 0;
        for (int64_t i = 0; i < N; i++)
            if (st[i] == 1 && (!any || key[i] < mn)) {
                mn = key[i];
                any = 1;
            }
        if (!any) break;
        int64_t nb = 0;
        for (int64_t i = 0; i < N; i++)
            if (st[i] == 1 && key[i] == mn) bat[nb++] = i;
        for (int64_t b = 0; b < nb; b++) {
            st[bat[b]] = 2;
            tt[bat[b]] = mn;
            order[cnt++] = bat[b];
        }
        for (int64_t b = 0; b < nb; b++) {
            int64_t c0[3] = {0, 0, 0};
            unlin(g, bat[b], c0);
            for (int ne = 0; ne < nnb; ne++) {
                int64_t q[3] = {c0[0], c0[1], c0[2]};
                for (int d = 0; d < nd; d++) q[d] += (nd == 2) ? neigh2[ne][d] : neigh3[ne][d];
                int oob = 0;
                for (int d = 0; d < nd; d++)
                    if (q[d] < 0 || q[d] > g->n[d] - 1) oob = 1;
                if (oob) continue;
                int64_t l = lin(g, q);
                if (st[l] != 2) {
                    key[l] = stencil(g, vel, tt, st, q, idD);
                    st[l] = 1;
                }
            }
        }
    }
    free(st);
    free(key);
    free(bat);
    return 0;
}
