/*
 * eik_oracle.h -- CPU oracle for the eikonal hot path (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C, FP64 restatement of EikonalSolvers.jl v0.15.3's forward fast
 * marching (heap FMM, 2nd-order stencil, refined-grid source init) and of its
 * discrete-adjoint gradient.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or as
 * the timed CPU baseline; the product (libeik) never links or calls it.
 *
 * Conventions
 *   - arrays are column-major with index 1 fastest (reference cart2lin,
 *     src/EikSolv/eikutils.jl:6-19); all indices here are 0-based.
 *   - coordsrc is packed rows: coordsrc[s*ndim + d].
 *   - receivers of source s are rows rec_off[s] .. rec_off[s+1]-1 of coordrec
 *     (packed rows), picks are written in the same packed order.
 *   - return value 0 = ok, <0 = error (see og_last_error()).
 */
#ifndef EIK_ORACLE_H
#define EIK_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* grid kinds: reference eikstructs.jl:33-219 */
enum { OG_CART2D = 0, OG_CART3D = 1, OG_SPH2D = 2, OG_SPH3D = 3 };

typedef struct {
    int32_t kind;      /* OG_* */
    int32_t ndim;      /* 2 or 3 */
    int64_t n[3];      /* grsize */
    double  d[3];      /* Cartesian: hgrid (all three); spherical: dr, dtheta[deg], dphi[deg] */
    double  o[3];      /* cooinit */
} og_grid;

typedef struct {
    int32_t refine;        /* refinearoundsrc (eikstructs.jl:264) */
    int32_t df;            /* grdrefpars.downscalefactor (default 5) */
    int32_t noderadius;    /* grdrefpars.noderadius (default 3) */
    int32_t radiussmooth;  /* radiussmoothgradsrc (default 3) */
    int32_t smoothkern;    /* smoothgradkern (default 0, odd) */
    int32_t nthreads;      /* 1 = serial (:serial); >1 = source chunks as distribsrcs (:sharedmem) */
    int32_t precision;     /* OG_PREC_FP64 (the reference) or OG_PREC_MIXED (the engine's mixed mode) */
} og_params;

/* OG_PREC_MIXED restates the engine's mixed precision (no reference counterpart): the coarse stencil
 * keeps FP64 values and comparisons and solves the quadratic's increment in FP32; near the source it
 * stays FP64 (the refined box, or without refinement the nodes within OG_MIXED_FP64_RADIUS of the
 * source box).  The problem (init set, adjoint) is the one the parameters define, as in FP64. */
enum { OG_PREC_FP64 = 0, OG_PREC_MIXED = 1 };
#define OG_MIXED_FP64_RADIUS 4

/* whichgrad (eikmaingrad.jl:40) */
enum { OG_GRADVEL = 1, OG_GRADSRCLOC = 2, OG_GRADBOTH = 3 };

const char* og_last_error(void);
int64_t og_warning_count(void);   /* reference @warn sites hit since last reset */
void og_reset_warnings(void);

/* eiktraveltime (eikmainfwd.jl:33-143).  ttgrid (nsrc*N) may be NULL. */
int og_traveltime(const og_grid* g, const double* vel, int64_t nsrc, const double* coordsrc,
                  const int64_t* rec_off, const double* coordrec, const og_params* p,
                  double* ttpicks, double* ttgrid);

/* eikgradient (eikmaingrad.jl:31-164).  gradvel (N) / gradsrc (nsrc*ndim packed rows) may be
 * NULL when not requested; misf receives 0.5*sum(r^2/sigma^2). */
int og_gradient(const og_grid* g, const double* vel, int64_t nsrc, const double* coordsrc,
                const int64_t* rec_off, const double* coordrec, const double* pickobs,
                const double* stdobs, int whichgrad, const og_params* p,
                double* gradvel, double* gradsrc, double* misf);

/* eikttimemisfit (eikmaingrad.jl:1302-1328) */
int og_misfit(const og_grid* g, const double* vel, int64_t nsrc, const double* coordsrc,
              const int64_t* rec_off, const double* coordrec, const double* pickobs,
              const double* stdobs, const og_params* p, double* misf);

/* Single-source FMM with the order/codes recorded (for tests): tt (N), order (N, grid index in
 * acceptance order, -1 past the end), codes (N*ndim, node-major).  Returns number accepted. */
int64_t og_fmm_record(const og_grid* g, const double* vel, const double* src, const og_params* p,
                      double* tt, int64_t* order, int8_t* codes, int64_t* nsrcpts);

/* Brute-force FMM (O(N^2) scan for the minimum, same equal-value batch rule) without refinement,
 * used to pin the heap on tiny grids.  Ties inside a batch are taken in increasing grid index. */
int og_fmm_bruteforce(const og_grid* g, const double* vel, const double* src, double* tt,
                      int64_t* order);

/* The local 2nd-order stencil on a caller-provided state (eikforward2D.jl:98-433,
 * eikforward3D.jl:151-532).  ijk 0-based; status: 2 = accepted.  Writes codes[ndim]. */
double og_stencil(const og_grid* g, const double* vel, const double* tt, const uint8_t* status,
                  const int64_t* ijk, int32_t* codes);

/* The same with the mixed precision's arithmetic (OG_PREC_MIXED). */
double og_stencil_mixed(const og_grid* g, const double* vel, const double* tt, const uint8_t* status,
                        const int64_t* ijk, int32_t* codes);

/* Gaussian smoothing of a gradient (utils.jl:22-96), out-of-place. */
int og_smoothgradient(const og_grid* g, int32_t l, const double* in, double* out);

/* TEST HOOK: plant a known mistake in the spherical forward so that a test can show that its pins
 * detect it (returns the previous setting).  Never set outside tests; 0 = the oracle as written. */
enum {
    OG_MISTAKE_NONE = 0,
    OG_MISTAKE_DEGREES = 1,      /* angular spacings taken in degrees, not radians */
    OG_MISTAKE_NO_RADIUS = 2,    /* arc length without the node's radius r_i */
    OG_MISTAKE_NO_SINTHETA = 3,  /* 3D: phi arc without sin(theta) */
    OG_MISTAKE_INIT_DEGREES = 4  /* 2D: source-box distance (law of cosines) in degrees */
};
int og_plant_mistake(int which);

#ifdef __cplusplus
}
#endif
#endif
