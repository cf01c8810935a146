// eik_noalloc.hip -- eik_traveltime_noalloc: the traveltime call without any device allocation.
//
// For callers that cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  Every device buffer the forward solve needs (traveltime
// and pop-key pools, tile flags and lists, refined boxes, receivers, picks) is carved from one
// file-scope scratch block instead of the context's growing pools.  The block is a __device__
// array: each GPU the library runs on holds its own copy, at its own address, from module load
// until the process ends.  Its address on a device is resolved once, at context creation, with
// hipGetSymbolAddress; host code only ever uses that resolved pointer (the array's name in host
// code is a host-side stand-in, not a device address).
#include <cstring>

#include "eik_internal.h"

namespace eik {

// Size of the scratch block: 128 MiB per GPU, enough for a 201x201x101 model with one source in
// FP64 (~105 MiB with refinement) or a 300x220 model with 4 sources (~7 MiB).
constexpr size_t kScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
struct ScratchSlot {
    std::once_flag once;
    char* ptr = nullptr;       // this device's copy of `scratch`
    hipError_t rc = hipSuccess;
    std::mutex use;            // one no-allocation call at a time on this device
};
ScratchSlot g_scratch[kMaxDevices];

size_t al(size_t n) { return ((n ? n : 16) + 255) & ~(size_t)255; }

// Upper bound of the device bytes forward_solve + forward_picks carve for this call (every buffer
// rounded up to the carving alignment; the tile list and the per-instance minimum are counted for
// the refined and the coarse solve separately, as they are ensured once per solve).
size_t forward_bytes(const eik_model* M, const SourceSet& S, const eik_params& P, bool ttgrid) {
    const int nd = M->g.ndim;
    const int64_t nsrc = S.nsrc;
    const size_t esz = P.precision == EIK_PREC_FP32 ? sizeof(float) : sizeof(double);
    InstDev I;
    make_inst(M->g, I);
    set_tiles(I, nd);
    const int64_t N = M->N;
    const int64_t tiles = (int64_t)I.t[0] * I.t[1] * I.t[2], atiles = (int64_t)I.at[0] * I.at[1] * I.at[2];
    const int64_t totN = nsrc * N;
    size_t b = 3 * al((size_t)totN * esz) + al((size_t)totN) + al((size_t)nsrc * sizeof(InstDev));
    int64_t maxt = nsrc * tiles, maxat = nsrc * atiles, ntf = 0, natf = 0;
    if (P.refinearoundsrc) {
        int64_t nfine = 0, sinth = 0, corn = 0;
        int cap = 1;
        for (int64_t s = 0; s < nsrc; s++) {
            FineDev F;
            InstDev FI;
            build_fine(M->g, S.src.data() + s * nd, P.downscalefactor, P.noderadius, F, FI);
            nfine += (int64_t)FI.n[0] * FI.n[1] * FI.n[2];
            ntf += (int64_t)FI.t[0] * FI.t[1] * FI.t[2];
            natf += (int64_t)FI.at[0] * FI.at[1] * FI.at[2];
            sinth += FI.n[1];
            corn += F.ncorn;
            int w = 1;
            for (int d = 0; d < nd; d++) w *= (FI.n[d] - 1) / F.df + 1;
            cap = std::max(cap, w);
        }
        b += 2 * al((size_t)nsrc * sizeof(InstDev)) + al((size_t)nsrc * sizeof(FineDev));
        b += 4 * al((size_t)nfine * sizeof(double)) + al((size_t)nfine);
        if (M->g.kind == EIK_GRID3D_SPHERE) b += al((size_t)sinth * sizeof(double));
        b += al((size_t)corn * sizeof(CornerDev)) + al((size_t)nsrc * sizeof(int)) + al((size_t)nsrc * cap * sizeof(SrcPt));
        b += al((size_t)ntf * sizeof(uint4)) + al((size_t)3 * nsrc * sizeof(unsigned));
        maxt = std::max(maxt, ntf);
        maxat = std::max(maxat, natf);
    } else {
        b += al((size_t)nsrc * (1 << nd) * sizeof(CornerDev));
    }
    // adjoint-tile flags (cleared by every forward call), forward tile flags and pending values,
    // the coarse solve's tile list and per-instance minima (three, used round robin)
    b += al((size_t)maxat) + al((size_t)maxat * sizeof(unsigned)) + 2 * al((size_t)maxt * sizeof(unsigned));
    b += al((size_t)nsrc * tiles * sizeof(uint4)) + al((size_t)3 * nsrc * sizeof(unsigned));
    const int64_t nrec = S.nrec();
    if (nrec > 0) b += al((size_t)nrec * nd * sizeof(double)) + al((size_t)nrec * sizeof(int)) + al((size_t)nrec * sizeof(double));
    if (ttgrid) b += al((size_t)N * sizeof(double));
    return b;
}
}  // namespace

void scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    ScratchSlot& S = g_scratch[device];
    std::call_once(S.once, [&S] {
        void* p = nullptr;
        // the caller (eik_ctx_create) has made `device` current
        S.rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        S.ptr = S.rc == hipSuccess ? (char*)p : nullptr;
    });
}

}  // namespace eik

using namespace eik;

extern "C" int eik_traveltime_noalloc(eik_ctx* ctx, eik_model* model, int64_t nsrc, const double* coordsrc,
                                      const int64_t* rec_off, const double* coordrec, const eik_params* params_in,
                                      double* ttpicks, double* ttgrid) {
    if (!ctx || !model) return EIK_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return ctx->fail(EIK_EDEVICE, "device %d out of range", ctx->device);
    ScratchSlot& slot = g_scratch[ctx->device];
    if (!slot.ptr)
        return ctx->fail(EIK_EDEVICE, "no-allocation scratch unavailable on device %d: hipGetSymbolAddress: %s",
                         ctx->device, hipGetErrorString(slot.rc));
    SourceSet S;
    int rc = load_sources(ctx, model, nsrc, coordsrc, rec_off, coordrec, params_in, S);
    if (rc) return rc;
    const eik_params pe = effective_params(params_in, model->g.kind);
    if (S.nrec() > 0 && !ttpicks) return ctx->fail(EIK_EINVAL, "null picks buffer");
    if (nsrc == 0) return EIK_OK;
    const size_t need = forward_bytes(model, S, pe, ttgrid != nullptr);
    if (need > kScratchBytes)
        return ctx->fail(EIK_ENOMEM,
                         "eik_traveltime_noalloc: this call needs %.1f MiB of device scratch, more than its %zu MiB "
                         "limit (use eik_traveltime, or fewer sources per call)",
                         (double)need / 1048576.0, kScratchBytes >> 20);
    // one call at a time on this GPU in this process: the block is shared by every context here
    std::lock_guard<std::mutex> dl(slot.use);
    Arena A;
    A.base = slot.ptr;
    A.cap = kScratchBytes;
    ctx->abufs.clear();
    ctx->arena = &A;
    stats_begin(ctx);
    rc = forward_solve(ctx, model, S, pe.precision, &pe);
    if (rc == EIK_OK) rc = forward_picks(ctx, model, S, pe.precision, ttpicks, ttgrid);
    if (rc == EIK_OK) {
        ctx->stats[0] = call_ms(ctx);
        ctx->stats[17] = busy_ms(ctx, false, false);
        ctx->stats[18] = busy_ms(ctx, false, true);
    }
    // nothing of this call may still run on the block when the next call takes it
    (void)hipStreamSynchronize(ctx->stream);
    ctx->arena = nullptr;
    ctx->abufs.clear();
    if (rc == EIK_ENOMEM)
        return ctx->fail(EIK_ENOMEM, "eik_traveltime_noalloc: the %zu MiB device scratch is too small for this call (%s)",
                         kScratchBytes >> 20, ctx->err.c_str());
    return rc;
}
