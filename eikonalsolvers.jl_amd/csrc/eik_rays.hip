// eik_rays.hip -- ray tracing on traveltime grids (tracerays / tracerays_singlesrc,
// src/EikSolv/eikutils.jl:201-328).
//
// The reference descends the gradient of a cubic B-spline interpolant of each source's
// traveltime grid (Interpolations.jl, BSpline(Cubic()) with its default Line(OnGrid()) boundary,
// scaled to the node coordinates) from every receiver, in fixed steps, until the ray comes within
// h*sqrt(2) (dr*sqrt(2) on Grid2DSphere) of the source, and replaces that last point by the source.
// Here every source's grid is prefiltered on the device (one tridiagonal solve per grid line and
// axis) and every ray is one thread; a first pass counts each ray's points, a second writes them.
//
// B-spline coefficients: a line of n values y gets n + 2 coefficients c (one padding coefficient
// per end) with (c[i-1] + 4 c[i] + c[i+1]) / 6 = y[i-1] at the nodes and a zero second
// difference c[0] - 2 c[1] + c[2] = 0 at both end nodes (the Line boundary: the interpolant
// continues linearly).  That gives c[1] = y[0], c[n] = y[n-1], and a tridiagonal system for the
// interior coefficients.
#include "eik_internal.h"

namespace eik {

namespace {

constexpr int kRayThreads = 64;

struct RayGrid {
    int nd, kind;
    int n[3];
    double h[3], o[3], end[3];
    double rstep, thr;
    int64_t N, NP;       // nodes and padded coefficients per source
    int64_t pn[3];       // padded extents
    int64_t max_steps;   // a ray still above the threshold after this many steps fails
};

// one tridiagonal solve per line along `axis` of the padded coefficient array (in place): the
// line's data sit at padded positions 1..n, its n + 2 coefficients are written at 0..n+1.
// `work` holds n doubles per line (the Thomas algorithm's modified super-diagonal).
__global__ void k_prefilter(double* __restrict__ coef, double* __restrict__ work, RayGrid G, int64_t nsrc, int axis,
                            int64_t nlines_per_src) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nlines_per_src * nsrc) return;
    const int64_t s = q / nlines_per_src, l = q % nlines_per_src;
    const int n = G.n[axis];
    // the line's fixed padded indices: axes before `axis` run over their padded extent (already
    // filtered), axes after it over the data positions 1..n (still raw)
    int64_t idx[3] = {0, 0, 0};
    int64_t rem = l;
    for (int d = 0; d < G.nd; d++) {
        if (d == axis) continue;
        const int64_t ext = d < axis ? G.pn[d] : G.n[d];
        idx[d] = rem % ext + (d < axis ? 0 : 1);
        rem /= ext;
    }
    int64_t stride = 1;
    for (int d = 0; d < axis; d++) stride *= G.pn[d];
    int64_t base = s * G.NP;
    {
        int64_t st = 1;
        for (int d = 0; d < G.nd; d++) {
            if (d != axis) base += idx[d] * st;
            st *= G.pn[d];
        }
    }
    double* c = coef + base;
    auto at = [&](int i) -> double& { return c[(int64_t)i * stride]; };
    double* cp = work + q * (int64_t)n;
    const double y0 = at(1), yn = at(n);
    if (n >= 3) {
        // interior unknowns c[2..n-1]: (1/6) c[i-1] + (2/3) c[i] + (1/6) c[i+1] = y[i-1]
        const int m = n - 2;
        const double a = 1.0 / 6.0, b = 2.0 / 3.0;
        // forward sweep (Thomas); rhs overwrite in place at positions 2..n-1
        for (int k = 0; k < m; k++) {
            const int i = k + 2;
            double r = at(i);
            if (k == 0) r -= a * y0;
            if (k == m - 1) r -= a * yn;
            const double denom = k == 0 ? b : b - a * cp[k - 1];
            cp[k] = a / denom;
            at(i) = k == 0 ? r / denom : (r - a * at(i - 1)) / denom;
        }
        for (int k = m - 2; k >= 0; k--) at(k + 2) -= cp[k] * at(k + 3);
    }
    at(1) = y0;
    at(n) = yn;
    at(0) = 2.0 * y0 - at(2);
    at(n + 1) = 2.0 * yn - at(n - 1);
}

// value-free gradient of the interpolant at physical point p (grid coordinates: x, y, z or r,
// theta[deg]); weights of the cubic B-spline and their derivatives
__device__ void bspline_grad(const double* __restrict__ c, const RayGrid& G, const double* p, double* g) {
    int j[3] = {0, 0, 0};
    double w[3][4], dw[3][4];
    for (int d = 0; d < G.nd; d++) {
        const double u = (p[d] - G.o[d]) / G.h[d];
        int i = (int)floor(u);
        i = i < 0 ? 0 : (i > G.n[d] - 2 ? G.n[d] - 2 : i);
        const double t = u - (double)i, s = 1.0 - t;
        w[d][0] = s * s * s / 6.0;
        w[d][1] = 2.0 / 3.0 - t * t + 0.5 * t * t * t;
        w[d][2] = 2.0 / 3.0 - s * s + 0.5 * s * s * s;
        w[d][3] = t * t * t / 6.0;
        dw[d][0] = -0.5 * s * s;
        dw[d][1] = -2.0 * t + 1.5 * t * t;
        dw[d][2] = 2.0 * s - 1.5 * s * s;
        dw[d][3] = 0.5 * t * t;
        j[d] = i;  // padded index of the first of the four coefficients
    }
    if (G.nd == 2) {
        double gx = 0.0, gy = 0.0;
        for (int b = 0; b < 4; b++) {
            const double* row = c + (j[1] + b) * G.pn[0] + j[0];
            double vx = 0.0, vy = 0.0;
            for (int a = 0; a < 4; a++) {
                vx += dw[0][a] * row[a];
                vy += w[0][a] * row[a];
            }
            gx += w[1][b] * vx;
            gy += dw[1][b] * vy;
        }
        g[0] = gx / G.h[0];
        g[1] = gy / G.h[1];
        return;
    }
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int e = 0; e < 4; e++) {
        double px = 0.0, py = 0.0, pz = 0.0;
        for (int b = 0; b < 4; b++) {
            const double* row = c + ((int64_t)(j[2] + e) * G.pn[1] + (j[1] + b)) * G.pn[0] + j[0];
            double vx = 0.0, vv = 0.0;
            for (int a = 0; a < 4; a++) {
                vx += dw[0][a] * row[a];
                vv += w[0][a] * row[a];
            }
            px += w[1][b] * vx;
            py += dw[1][b] * vv;
            pz += w[1][b] * vv;
        }
        gx += w[2][e] * px;
        gy += w[2][e] * py;
        gz += dw[2][e] * pz;
    }
    g[0] = gx / G.h[0];
    g[1] = gy / G.h[1];
    g[2] = gz / G.h[2];
}

__device__ double ray_dist(const RayGrid& G, const double* a, const double* b) {
    if (G.kind == EIK_GRID2D_SPHERE) {
        const double dth = (a[1] - b[1]) * (M_PI / 180.0);
        return sqrt(a[0] * a[0] + b[0] * b[0] - 2.0 * a[0] * b[0] * cos(dth));
    }
    double s = 0.0;
    for (int d = 0; d < G.nd; d++) s += (a[d] - b[d]) * (a[d] - b[d]);
    return sqrt(s);
}

// One thread per ray.  out == nullptr: count the points (npts); else write them at out + off*nd,
// never more than the counted npts (the second pass repeats the first's arithmetic exactly).
__global__ void __launch_bounds__(kRayThreads) k_trace(const double* __restrict__ coef, RayGrid G,
                                                       const double* __restrict__ src, const double* __restrict__ rec,
                                                       const int* __restrict__ rsrc, int64_t nrec,
                                                       int64_t* __restrict__ npts, const int64_t* __restrict__ off,
                                                       double* __restrict__ out, unsigned* __restrict__ failed) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    const int s = rsrc[r];
    const double* c = coef + (int64_t)s * G.NP;
    const double* sp = src + (int64_t)s * G.nd;
    double p[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < G.nd; d++) p[d] = rec[r * G.nd + d];
    double* o = out ? out + off[r] * G.nd : nullptr;
    int64_t k = 0;  // index of the current point
    if (o)
        for (int d = 0; d < G.nd; d++) o[d] = p[d];
    const int64_t limit = out ? npts[r] - 1 : G.max_steps;
    double dist = ray_dist(G, p, sp);
    while (dist >= G.thr) {
        if (k >= limit) {
            if (!out) atomicAdd(failed, 1u);
            break;
        }
        double g[3];
        bspline_grad(c, G, p, g);
        double np[3];
        if (G.kind == EIK_GRID2D_SPHERE) {
            const double rc = p[0], th = p[1] * (M_PI / 180.0);
            const double gth = g[1] * (180.0 / M_PI);
            const double dx = cos(th) * g[0] - 1.0 / rc * sin(th) * gth;
            const double dy = sin(th) * g[0] + 1.0 / rc * cos(th) * gth;
            const double nrm = sqrt(dx * dx + dy * dy);
            const double x = rc * cos(th) - G.rstep * (dx / nrm);
            const double y = rc * sin(th) - G.rstep * (dy / nrm);
            np[0] = sqrt(x * x + y * y);
            np[1] = atan2(y, x) * (180.0 / M_PI);
        } else {
            double nrm = 0.0;
            for (int d = 0; d < G.nd; d++) nrm += g[d] * g[d];
            nrm = sqrt(nrm);
            for (int d = 0; d < G.nd; d++) np[d] = p[d] - G.rstep * (g[d] / nrm);
        }
        for (int d = 0; d < G.nd; d++) {
            if (np[d] < G.o[d])
                np[d] = G.o[d];
            else if (np[d] > G.end[d])
                np[d] = G.end[d];
            p[d] = np[d];
        }
        k++;
        if (o)
            for (int d = 0; d < G.nd; d++) o[k * G.nd + d] = p[d];
        dist = ray_dist(G, p, sp);
    }
    // the last point (within the threshold) becomes the source
    if (o)
        for (int d = 0; d < G.nd; d++) o[k * G.nd + d] = sp[d];
    if (!out) npts[r] = k + 1;
}

}  // namespace

// Step cap of one ray: a descending ray is never much longer than the domain, so 20 domain
// lengths of steps (plus 1000) are allowed.  The domain length is the diagonal of the Cartesian
// box (`diag`), or on a Grid2DSphere the outer circle's diameter 2 r_max, which bounds every
// chord inside the annulus sector.  Both are lengths; rstep is the step length in the same unit.
static int64_t ray_max_steps(int kind, double diag, double rmax, double rstep) {
    const double len = kind == EIK_GRID2D_SPHERE ? 2.0 * rmax : diag;
    return (int64_t)std::min(4.0e9, 20.0 * len / rstep + 1000.0);
}

}  // namespace eik

using namespace eik;

extern "C" int eik_tracerays_max_steps(const eik_grid* grid, double steplen, int64_t* out) {
    if (!grid || !out || grid->ndim < 2 || grid->ndim > 3) return EIK_EINVAL;
    if (!(steplen > 0.0)) return EIK_EASSERT;
    double diag = 0.0;
    for (int d = 0; d < grid->ndim; d++) {
        const double ext = (double)(grid->n[d] - 1) * grid->d[d];
        diag += ext * ext;
    }
    const bool sph = grid->kind == EIK_GRID2D_SPHERE;
    const double rmax = grid->o[0] + (double)(grid->n[0] - 1) * grid->d[0];
    *out = ray_max_steps(grid->kind, sqrt(diag), rmax, sph ? steplen : steplen * grid->d[0]);
    return EIK_OK;
}

extern "C" int eik_tracerays(eik_ctx* ctx, const eik_grid* grid, int64_t nsrc, const double* ttime,
                             const double* coordsrc, const int64_t* rec_off, const double* coordrec, double steplen,
                             int64_t* ray_off, double* rays, int64_t cap_points) {
    if (!ctx || !grid) return EIK_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    const int nd = grid->ndim;
    if (nd < 2 || nd > 3) return ctx->fail(EIK_EASSERT, "2<=ndims(ttime)<=3");
    if (grid->kind == EIK_GRID3D_SPHERE)
        return ctx->fail(EIK_EINVAL, "tracerays_singlesrc(): ray tracing for spherical grids not yet implemented.");
    if (grid->kind != EIK_GRID2D_CART && grid->kind != EIK_GRID3D_CART && grid->kind != EIK_GRID2D_SPHERE)
        return ctx->fail(EIK_EINVAL, "unknown grid kind %d", grid->kind);
    if (nsrc < 0 || (nsrc > 0 && (!ttime || !coordsrc || !rec_off || !ray_off)))
        return ctx->fail(EIK_EINVAL, "null argument");
    if (!(steplen > 0.0)) return ctx->fail(EIK_EASSERT, "steplen>0");
    RayGrid G{};
    G.nd = nd;
    G.kind = grid->kind;
    G.N = 1;
    G.NP = 1;
    double diag = 0.0;
    for (int d = 0; d < 3; d++) {
        G.n[d] = d < nd ? (int)grid->n[d] : 1;
        G.pn[d] = d < nd ? G.n[d] + 2 : 1;
        G.h[d] = grid->d[d];
        G.o[d] = grid->o[d];
        G.end[d] = grid->o[d] + (G.n[d] - 1) * grid->d[d];
        if (d < nd) {
            if (G.n[d] < 4) return ctx->fail(EIK_EASSERT, "ray tracing needs at least 4 nodes per axis");
            G.N *= G.n[d];
            G.NP *= G.pn[d];
            diag += (G.end[d] - G.o[d]) * (G.end[d] - G.o[d]);
        }
    }
    const bool sph = grid->kind == EIK_GRID2D_SPHERE;
    G.rstep = sph ? steplen : steplen * grid->d[0];
    G.thr = (sph ? grid->d[0] : grid->d[0]) * sqrt(2.0);
    G.max_steps = ray_max_steps(grid->kind, sqrt(diag), G.end[0], G.rstep);
    const int64_t nrec = nsrc > 0 ? rec_off[nsrc] : 0;
    ray_off[0] = 0;
    if (nsrc == 0 || nrec == 0) {
        for (int64_t r = 0; r <= nrec; r++) ray_off[r] = 0;
        return EIK_OK;
    }
    // every receiver starts inside the grid (the steps are clamped to it)
    for (int64_t r = 0; r < nrec; r++)
        for (int d = 0; d < nd; d++) {
            const double x = coordrec[r * nd + d];
            if (!(x >= G.o[d] && x <= G.end[d])) return ctx->fail(EIK_EASSERT, "receiver %lld outside the grid", (long long)r);
        }
    hipStream_t st = ctx->stream;
    DevBuf &dcoef = ctx->buf("ray_coef"), &dwork = ctx->buf("ray_work"), &dsrc = ctx->buf("ray_src"),
           &drec = ctx->buf("ray_rec"), &drs = ctx->buf("ray_rsrc"), &dnp = ctx->buf("ray_npts"),
           &doff = ctx->buf("ray_off"), &dout = ctx->buf("ray_out");
    int64_t maxlines = 0;
    for (int a = 0; a < nd; a++) {
        int64_t nl = 1;
        for (int d = 0; d < nd; d++)
            if (d != a) nl *= d < a ? G.pn[d] : G.n[d];
        maxlines = std::max(maxlines, nl * G.n[a]);
    }
    if (dcoef.ensure((size_t)(nsrc * G.NP) * sizeof(double)) || dwork.ensure((size_t)(nsrc * maxlines) * sizeof(double)) ||
        dsrc.ensure((size_t)nsrc * nd * sizeof(double)) || drec.ensure((size_t)nrec * nd * sizeof(double)) ||
        drs.ensure((size_t)nrec * sizeof(int)) || dnp.ensure((size_t)nrec * sizeof(int64_t)) ||
        doff.ensure((size_t)(nrec + 1) * sizeof(int64_t)))
        return ctx->fail(EIK_ENOMEM, "out of device memory (ray tracing)");
    // grids into the padded coefficient array (data at padded positions 1..n per axis)
    std::vector<int> rs((size_t)nrec);
    for (int64_t s = 0; s < nsrc; s++)
        for (int64_t r = rec_off[s]; r < rec_off[s + 1]; r++) rs[r] = (int)s;
    CK(hipMemsetAsync(dcoef.p, 0, (size_t)(nsrc * G.NP) * sizeof(double), st));
    for (int64_t s = 0; s < nsrc; s++) {
        const double* src = ttime + s * G.N;
        double* dst = (double*)dcoef.p + s * G.NP;
        if (nd == 2) {
            CK(hipMemcpy2DAsync(dst + G.pn[0] + 1, G.pn[0] * sizeof(double), src, G.n[0] * sizeof(double),
                                G.n[0] * sizeof(double), G.n[1], hipMemcpyHostToDevice, st));
        } else {
            for (int z = 0; z < G.n[2]; z++)
                CK(hipMemcpy2DAsync(dst + ((int64_t)(z + 1) * G.pn[1] + 1) * G.pn[0] + 1, G.pn[0] * sizeof(double),
                                    src + (int64_t)z * G.n[0] * G.n[1], G.n[0] * sizeof(double),
                                    G.n[0] * sizeof(double), G.n[1], hipMemcpyHostToDevice, st));
        }
    }
    CK(hipMemcpyAsync(dsrc.p, coordsrc, (size_t)nsrc * nd * sizeof(double), hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(drec.p, coordrec, (size_t)nrec * nd * sizeof(double), hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(drs.p, rs.data(), (size_t)nrec * sizeof(int), hipMemcpyHostToDevice, st));
    for (int a = 0; a < nd; a++) {
        int64_t nl = 1;
        for (int d = 0; d < nd; d++)
            if (d != a) nl *= d < a ? G.pn[d] : G.n[d];
        const int64_t tot = nl * nsrc;
        k_prefilter<<<(unsigned)((tot + 127) / 128), 128, 0, st>>>((double*)dcoef.p, (double*)dwork.p, G, nsrc, a, nl);
        CK(hipGetLastError());
    }
    unsigned* d_fail = (unsigned*)ctx->d_small;
    CK(hipMemsetAsync(d_fail, 0, sizeof(unsigned), st));
    const unsigned nb = (unsigned)((nrec + kRayThreads - 1) / kRayThreads);
    k_trace<<<nb, kRayThreads, 0, st>>>((const double*)dcoef.p, G, (const double*)dsrc.p, (const double*)drec.p,
                                        (const int*)drs.p, nrec, (int64_t*)dnp.p, nullptr, nullptr, d_fail);
    CK(hipGetLastError());
    std::vector<int64_t> np((size_t)nrec);
    CK(hipMemcpyAsync(np.data(), dnp.p, (size_t)nrec * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    CK(hipMemcpyAsync(ctx->h_small, d_fail, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (*(unsigned*)ctx->h_small)
        return ctx->fail(EIK_ENOCONV, "%u ray(s) did not reach their source within %lld steps",
                         *(unsigned*)ctx->h_small, (long long)G.max_steps);
    for (int64_t r = 0; r < nrec; r++) ray_off[r + 1] = ray_off[r] + np[r];
    const int64_t total = ray_off[nrec];
    if (!rays) return EIK_OK;
    if (total > cap_points) return ctx->fail(EIK_ESIZE, "ray buffer holds %lld points, %lld needed",
                                             (long long)cap_points, (long long)total);
    if (dout.ensure((size_t)total * nd * sizeof(double))) return ctx->fail(EIK_ENOMEM, "out of device memory (rays)");
    CK(hipMemcpyAsync(doff.p, ray_off, (size_t)(nrec + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    k_trace<<<nb, kRayThreads, 0, st>>>((const double*)dcoef.p, G, (const double*)dsrc.p, (const double*)drec.p,
                                        (const int*)drs.p, nrec, (int64_t*)dnp.p, (const int64_t*)doff.p,
                                        (double*)dout.p, d_fail);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(rays, dout.p, (size_t)total * nd * sizeof(double), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    return EIK_OK;
}
