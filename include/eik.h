/*
 * eik.h -- C ABI of libeik, the MI355X (gfx950) eikonal engine.
 *
 * This is the drop-in boundary under the host-language API of EikonalSolvers.jl.  Each entry
 * point replaces one public function of the reference (cited per declaration); the host mirror
 * (Python ctypes in eikonalsolvers.jl_amd/eikonalsolvers, or a Julia ccall wrapper) converts
 * the reference's arguments into the plain arrays below.
 *
 * Conventions
 *   - Arrays are column-major with index 1 fastest, exactly like the reference's Julia arrays
 *     (cart2lin, src/EikSolv/eikutils.jl:6-19).  Indices crossing this ABI are 0-based.
 *   - coordsrc: nsrc rows of ndim doubles (packed rows).  Receivers of source s are rows
 *     rec_off[s] .. rec_off[s+1]-1 of coordrec (packed rows); picks use the same packing.
 *   - All output buffers are caller-allocated host memory.  Device memory is owned by the
 *     context and model handles.
 *   - Return value: EIK_OK (0) or a negative EIK_E* code; eik_last_error() gives the message.
 *     No exception ever crosses the ABI.  Validation failures that the reference raises with
 *     @assert return EIK_EASSERT so the host wrapper can raise AssertionError.
 *   - A context serialises its calls; use one context per host thread for concurrency.
 */
#ifndef EIK_H
#define EIK_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EIK_OK 0
#define EIK_EASSERT (-1)      /* reference @assert (AssertionError) */
#define EIK_EINVAL (-2)       /* unsupported case (reference error(...) / ErrorException) */
#define EIK_ENOMEM (-3)       /* device or host allocation failed */
#define EIK_EDEVICE (-4)      /* HIP runtime error */
#define EIK_ENOCONV (-5)      /* iterative solve did not converge within its round limit */
#define EIK_ESIZE (-6)        /* caller's output buffer too small; the size needed has been written */

/* grid kinds (src/EikSolv/eikstructs.jl:33-219) */
#define EIK_GRID2D_CART 0
#define EIK_GRID3D_CART 1
#define EIK_GRID2D_SPHERE 2
#define EIK_GRID3D_SPHERE 3

/* arithmetic of the traveltime stencil */
#define EIK_PREC_FP64 0   /* FP64 everywhere: the reference's arithmetic */
#define EIK_PREC_MIXED 1  /* FP32 stencil increments on FP64 traveltimes (the heap order is decided in
                             FP64); FP64 near-source solve, adjoint and reductions */
#define EIK_PREC_FP32 2   /* FP32 traveltimes and stencil (throughput mode: heap near-ties within FP32
                             resolution can change a cone downstream); FP64 near-source solve, adjoint
                             and reductions */

/* whichgrad (src/EikSolv/eikmaingrad.jl:40) */
#define EIK_GRADVEL 1
#define EIK_GRADSRCLOC 2
#define EIK_GRADVELANDSRCLOC 3

typedef struct {
    int32_t kind;   /* EIK_GRID* */
    int32_t ndim;   /* 2 or 3 */
    int64_t n[3];   /* grsize */
    double d[3];    /* Cartesian: hgrid (all three); spherical: dr, dtheta[deg], dphi[deg] */
    double o[3];    /* cooinit */
} eik_grid;

/* ExtraParams (src/EikSolv/eikstructs.jl:246-282) plus precision */
typedef struct {
    int32_t refinearoundsrc;     /* default 1 */
    int32_t downscalefactor;     /* grdrefpars.downscalefactor, default 5 */
    int32_t noderadius;          /* grdrefpars.noderadius, default 3 */
    int32_t radiussmoothgradsrc; /* default 3 */
    int32_t smoothgradkern;      /* default 0 (odd when > 0) */
    int32_t precision;           /* EIK_PREC_* */
    int32_t fast;                /* FP64: 0 (default) = bit-identical to the heap FMM (every change
                                    propagates); 1 = changes below 1e-12 relative wake no other node:
                                    ~1/3 fewer updates, but a suppressed change can flip a decision at
                                    an exact heap tie (measured up to 5e-5 relative on a grid, inside
                                    the stated tolerances).  Ignored in mixed precision. */
    int32_t reserved;            /* must be 0 */
} eik_params;

typedef struct eik_ctx eik_ctx;
typedef struct eik_model eik_model;

/* Library/context lifetime.  No reference counterpart (the reference is stateless Julia code):
 * a context owns one HIP device, one stream and the device scratch pools. */
int eik_ctx_create(int32_t device, eik_ctx** out);
void eik_ctx_destroy(eik_ctx* ctx);
const char* eik_last_error(const eik_ctx* ctx);
const char* eik_version(void);

/* Resident velocity model.  Replaces the `vel` argument that the reference passes to every call
 * (eikmainfwd.jl:33, eikmaingrad.jl:31): the model is validated once (@assert all(vel.>0.0),
 * eikmainfwd.jl:45) and kept in HBM as FP64 velocity plus FP64/FP32 slowness. */
int eik_model_create(eik_ctx* ctx, const eik_grid* grid, const double* vel, eik_model** out);
int eik_model_update_vel(eik_model* model, const double* vel);
void eik_model_destroy(eik_model* model);

/* eiktraveltime (src/EikSolv/eikmainfwd.jl:33-143): picks for every source's receivers and,
 * when ttgrid != NULL, the nsrc traveltime grids (nsrc*N doubles, source-major). */
int eik_traveltime(eik_ctx* ctx, eik_model* model, int64_t nsrc, const double* coordsrc,
                   const int64_t* rec_off, const double* coordrec, const eik_params* params,
                   double* ttpicks, double* ttgrid);

/* eikgradient (src/EikSolv/eikmaingrad.jl:31-164).  gradvel (N doubles) and gradsrc (nsrc*ndim,
 * packed rows) may be NULL when not requested by whichgrad; misf = 0.5*sum((t-tobs)^2/std^2).
 * Spherical grids: EIK_GRADVEL only (EIK_EINVAL otherwise).  Grid3DSphere, where the reference
 * raises (createsparsederivativematrices!, eikmainfwd.jl:885-887), differentiates this library's
 * own 3D spherical forward. */
int eik_gradient(eik_ctx* ctx, eik_model* model, int64_t nsrc, const double* coordsrc,
                 const int64_t* rec_off, const double* coordrec, const double* pickobs,
                 const double* stdobs, int32_t whichgrad, const eik_params* params,
                 double* gradvel, double* gradsrc, double* misf);

/* eikttimemisfit (src/EikSolv/eikmaingrad.jl:1302-1328) */
int eik_misfit(eik_ctx* ctx, eik_model* model, int64_t nsrc, const double* coordsrc,
               const int64_t* rec_off, const double* coordrec, const double* pickobs,
               const double* stdobs, const eik_params* params, double* misf);

/* eik_traveltime without any device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, results and errors as eik_traveltime (FP64: bit for bit);
 * every device buffer of the call is carved from a fixed 128 MiB scratch block per GPU (a __device__
 * array held from library load until the process ends), and a call whose buffers would not fit is
 * rejected with EIK_ENOMEM before any device work (about one 201x201x101 source in FP64).  The block
 * keeps the last call's contents; calls on one GPU from any context of the process run one at a
 * time (the library serialises them; overlapping calls would otherwise race on the block). */
int eik_traveltime_noalloc(eik_ctx* ctx, eik_model* model, int64_t nsrc, const double* coordsrc,
                           const int64_t* rec_off, const double* coordrec, const eik_params* params,
                           double* ttpicks, double* ttgrid);

/* Free and total memory of the context's GPU (hipMemGetInfo through the HIP runtime this library uses),
 * for callers that size their calls; no reference counterpart. */
int eik_mem_info(eik_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);

/* Streams the context's large eik_traveltime and eik_gradient calls may split their sources over:
 * 2 (the default; EIK_STREAMS=1 in the environment makes it 1) or 1.  A split call solves the
 * second half of its sources on a second stream, concurrently, with the same results; callers
 * that keep the GPU busy from threads of their own may prefer 1.  No reference counterpart. */
int eik_ctx_set_streams(eik_ctx* ctx, int32_t nstreams);

/* Timing of the last call's device work (ms), split by phase; no reference counterpart.
 * out[0] = total, [1] = near-source (fine grid) solve, [2] = coarse solve, [3] = adjoint,
 * [4] = tile activations of the coarse solve (node-visits = activations x tile nodes),
 * [5] = coarse solve rounds (launches), [6] = summed duration of those launches (ms, HIP events
 * on the library's stream), [7] = local relaxation sweeps summed over the activations,
 * [8] = node updates (stencil replays) of the coarse solve, [9] = those that changed the node,
 * [10] = of those, key-only changes, [11] = settling stage of the coarse solve: 0 = exact; k > 0
 * = after 2000 rounds, changes below 10^(k-1) x 1e-12 relative (FP64; 4e-6 in FP32) stopped
 * propagating (the tolerance grows tenfold every 500 rounds).  Converging solves stay at 0.
 * [12] = verification rounds of the coarse solve (at convergence every tile is relaxed once more;
 * one round that changes nothing ends the solve), [13] = round-kernel launches of the coarse solve
 * (rounds that found no tile included), [14] = summed duration of the near-source (fine grid)
 * solve's round launches (ms), [15] = those launches, [16] = that solve's tile activations,
 * [17] = time in which at least one round launch of the coarse solve ran (ms, the union of the
 * launches' spans over both streams of a split call; [6] sums overlapping launches), [18] = the
 * same over the launches of both solves (near-source and coarse).
 * eik_traveltime splits two or more sources (at least 3 x 2^22 nodes over all of them) into two halves
 * on two streams (the context's own and a second one it creates on first use; not on a server's
 * slots): the counts above are then summed over both halves, [1], [2], [3] and [11] are the larger
 * of the two.  Values a call did not produce are 0. */
int eik_last_stats(const eik_ctx* ctx, double* out, int32_t nout);

/* tracerays (src/EikSolv/eikutils.jl:201-328): for every source's receivers, the ray descending
 * the gradient of the cubic B-spline interpolant of that source's traveltime grid (Interpolations
 * BSpline(Cubic(Line(OnGrid())))) in steps of steplen*hgrid (Cartesian) or steplen (Grid2DSphere),
 * until within hgrid*sqrt(2) (dr*sqrt(2)) of the source, whose coordinates end the ray.
 * ttime: nsrc grids (nsrc*N doubles, source-major, index 1 fastest), e.g. eik_traveltime's ttgrid.
 * ray_off (nrec+1) receives the prefix sums of the points per ray, rays in receiver order.  When
 * rays != NULL the points are written as packed rows (ray_off[nrec] x ndim doubles) if they fit in
 * cap_points, else EIK_ESIZE is returned with ray_off filled.  Grid3DSphere: EIK_EINVAL, as in the
 * reference.  A ray that has not reached its source after 20 domain diagonals of steps fails with
 * EIK_ENOCONV (the reference has no such limit). */
int eik_tracerays(eik_ctx* ctx, const eik_grid* grid, int64_t nsrc, const double* ttime, const double* coordsrc,
                  const int64_t* rec_off, const double* coordrec, double steplen, int64_t* ray_off, double* rays,
                  int64_t cap_points);

/* The step cap eik_tracerays applies to one ray on this grid (host-only query, no device work; no
 * reference counterpart, the reference's loop has no cap): 20 domain lengths of steps plus 1000,
 * where the domain length is the box diagonal (Cartesian) or the outer diameter 2 r_max
 * (Grid2DSphere) and the step is steplen*hgrid (Cartesian) or steplen (Grid2DSphere). */
int eik_tracerays_max_steps(const eik_grid* grid, double steplen, int64_t* out);

/* ---------------------------------------------------------------- request loop (serving)
 * No reference counterpart: the reference evaluates one call at a time.  A server runs
 * traveltime and gradient requests with continuous batching: nslots slots (each a context with
 * its own stream and reusable device pools) take queued requests as soon as a slot is free and
 * the request's device working set fits in what the memory budget has left; admitted requests
 * run concurrently and each batches all of its sources into shared launches.  What the server
 * holds on the device is, per slot, the larger of the pools the slot keeps from its largest call
 * and the working set of the request it runs; a request is admitted when that sum stays within
 * the budget (idle slots' pools are released when that makes the room).  mem_budget > 0 fixes the
 * budget in bytes; mem_budget = 0 takes the device's free memory (read at every admission) plus
 * what the slots hold, less a reserve of max(2 GiB, 1/32 of the device).  A request whose working
 * set exceeds the whole budget fails on its own with EIK_ENOMEM (eik_wait returns it).  The arrays a
 * request points to (inputs and outputs) must stay valid until eik_wait returns for it, and its
 * model must not be updated or destroyed meanwhile. */
#define EIK_REQ_TRAVELTIME 1 /* eiktraveltime: fills ttpicks */
#define EIK_REQ_GRADIENT 2   /* eikgradient: fills gradvel / gradsrc per whichgrad, and misf */

typedef struct {
    int32_t type;      /* EIK_REQ_* */
    int32_t whichgrad; /* EIK_GRAD* (gradient requests) */
    eik_model* model;
    int64_t nsrc;
    const double* coordsrc;
    const int64_t* rec_off;
    const double* coordrec;
    const double* pickobs; /* gradient requests */
    const double* stdobs;
    eik_params params;
    double* ttpicks;
    double* gradvel;
    double* gradsrc;
    double* misf;
} eik_request;

typedef struct eik_server eik_server;
int eik_server_create(int32_t device, int32_t nslots, int64_t mem_budget, eik_server** out);
/* waits for queued and running requests, then releases the slots */
void eik_server_destroy(eik_server* server);
int eik_submit(eik_server* server, const eik_request* request, int64_t* id);
/* The device working set the server counts for a request (host-only, no device work): reads only
 * type, model, nsrc and params. */
int eik_request_bytes(const eik_request* request, int64_t* out);
/* Blocks until request id has finished; returns its status (EIK_OK or the call's error code).
 * times_ms (optional, 3 doubles): latency from submission, queueing delay, run time; errbuf
 * (optional) receives the error message. */
int eik_wait(eik_server* server, int64_t id, double* times_ms, char* errbuf, int64_t errlen);
/* eik_wait, and stats (optional, up to 14 doubles): the eik_last_stats values of the call that ran
 * the request ([0..12]), then the number of requests that call held ([13]; > 1 for a shared
 * traveltime launch). */
int eik_wait_stats(eik_server* server, int64_t id, double* times_ms, double* stats, int32_t nstats, char* errbuf,
                   int64_t errlen);
/* Non-blocking: *done = 1 when request id has finished (eik_wait then returns at once), else 0.
 * EIK_EINVAL for an unknown id (never submitted, or already waited for). */
int eik_poll(eik_server* server, int64_t id, int32_t* done);
/* out[0] queued requests, [1] running, [2] admitted working-set bytes, [3] budget (the last
 * admission's), [4] shared launches run for two or more traveltime requests, [5] requests run in
 * them, [6] bytes the slots' pools hold, [7] requests failed as larger than the budget, [8] idle
 * slots whose pools were released to admit a request, [9] calls retried after running out of
 * device memory */
int eik_server_stats(eik_server* server, int64_t* out, int32_t nout);

/* Several traveltime requests solved in shared launches (the server's continuous batching; no
 * reference counterpart).  Every request must be EIK_REQ_TRAVELTIME with the same params; the
 * models must share grid kind and dimension (sizes, spacings and velocities may differ).  Each
 * request's picks go to its own ttpicks; results equal those of separate eik_traveltime calls
 * (FP64 bit for bit).  An invalid request fails the whole call. */
int eik_traveltime_batch(eik_ctx* ctx, int32_t nreq, const eik_request* reqs);

#ifdef __cplusplus
}
#endif
#endif
