/*
 * xo_noalloc.hip — xo_batch_run_noalloc: the batch run of xo_batch_run for
 * callers that cannot allocate device memory. The caller passes the LUT and
 * the accumulator as device arrays; everything else the batch kernel needs
 * (the safe-prime multipliers, the instance tables, the work items and the
 * counters) lives in one static device array per GPU.
 *
 * Layout of the array (bytes):
 *   [0, kTableBytes)            seed-table multipliers, rows 1.. (written
 *                               once per device at its first call)
 *   [kTableBytes, +kCountBytes) item counter and run statistics
 *   then, each 256-byte aligned: layer records, headers, work items.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/xomc.h"
#include "xo_batch.h"
#include "xo_ctx.h"
#include "xo_internal.h"
#include "xo_kernels.h"
#include "xo_seedhash.h"

namespace {

constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
constexpr size_t kTableBytes = align_up(sizeof(uint32_t) * XO_SEED_TABLE_ROWS);
constexpr size_t kCountBytes = align_up(sizeof(unsigned long long) * (1 + XO_STAT_COUNT));
constexpr size_t kScratchBytes = XO_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes > kTableBytes + kCountBytes, "scratch too small");

/* Each GPU holds its own copy of this array for the life of the process. */
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;

/* Device address of `scratch` on each device (hipGetSymbolAddress, taken
 * once per device with that device current) and whether its seed table has
 * been written. In host code the array's name is not a device address. */
struct DeviceScratch {
    unsigned char *ptr = nullptr;
    bool table_ready = false;
};
DeviceScratch g_scratch[kMaxDevices];
std::mutex g_scratch_mu;

/* Set-up of the array on ctx's device (its first call there): the symbol's
 * address, then the seed table copied in on ctx's stream. The caller holds
 * ctx->mu and has made the device current. */
int scratch_for(xo_ctx *ctx, unsigned char **out) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) {
        xo::set_error("xo_batch_run_noalloc: device %d beyond the %d supported", ctx->device,
                      kMaxDevices);
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    DeviceScratch &d = g_scratch[ctx->device];
    if (!d.ptr) {
        void *p = nullptr;
        XO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        d.ptr = (unsigned char *)p;
    }
    if (!d.table_ready) {
        std::vector<uint32_t> fora(XO_NUM_PRIMES);
        if (xo::table_get(XO_NUM_PRIMES, fora.data()) != 0) {
            xo::set_error("xo_batch_run_noalloc: safe-prime table unavailable");
            return -1;
        }
        XO_HIP(hipMemcpyAsync(d.ptr, fora.data() + 1, sizeof(uint32_t) * XO_SEED_TABLE_ROWS,
                              hipMemcpyHostToDevice, ctx->stream));
        XO_HIP(hipStreamSynchronize(ctx->stream));
        d.table_ready = true;
    }
    *out = d.ptr;
    return 0;
}

} // namespace

extern "C" {

int xo_batch_run_noalloc(xo_ctx *ctx, const xo_batch_instance *inst, uint32_t ninst,
                         const float *lut_dev, uint32_t nlut, const xo_batch_opts *opts,
                         uint64_t *accum_dev, uint64_t naccum, xo_run_report *report) {
    auto t0 = std::chrono::steady_clock::now();
    xo::clear_error();
    if (!ctx || !inst || !opts || !accum_dev || ninst == 0 || naccum == 0) {
        xo::set_error("xo_batch_run_noalloc: null or empty argument");
        return -1;
    }
    if (nlut && !lut_dev) {
        xo::set_error("xo_batch_run_noalloc: %u LUT floats but no LUT array", nlut);
        return -1;
    }
    if (naccum > 0xFFFFFFFFull / 2) {
        xo::set_error("xo_batch_run_noalloc: %llu accumulator bins (at most %llu)",
                      (unsigned long long)naccum, 0xFFFFFFFFull / 2);
        return -1;
    }
    std::vector<xo::LayerL> layers;
    std::vector<xo::BatchHdr> hdr;
    std::vector<xo::BatchItem> items;
    uint64_t total = 0;
    uint32_t block = 0;
    if (xo::pack_batch("xo_batch_run_noalloc", inst, ninst, lut_dev != nullptr, nlut, opts,
                       naccum, layers, hdr, items, &total, &block))
        return -1;
    const size_t lbytes = layers.size() * sizeof(xo::LayerL);
    const size_t hbytes = hdr.size() * sizeof(xo::BatchHdr);
    const size_t ibytes = items.size() * sizeof(xo::BatchItem);
    const size_t off_layers = kTableBytes + kCountBytes;
    const size_t off_hdr = off_layers + align_up(lbytes);
    const size_t off_items = off_hdr + align_up(hbytes);
    const size_t need = off_items + align_up(ibytes ? ibytes : 1);
    if (need > kScratchBytes) {
        xo::set_error("xo_batch_run_noalloc: the batch needs %llu bytes of scratch (%u "
                      "instances, %llu work items), the static array holds %llu; raise the "
                      "chunk size or split the batch",
                      (unsigned long long)need, ninst, (unsigned long long)items.size(),
                      (unsigned long long)kScratchBytes);
        return -1;
    }

    std::lock_guard<std::mutex> lock(ctx->mu);
    XO_HIP(hipSetDevice(ctx->device));
    unsigned char *base = nullptr;
    if (scratch_for(ctx, &base)) return -1;
    unsigned long long *counters = (unsigned long long *)(base + kTableBytes);
    XO_HIP(hipMemcpyAsync(base + off_layers, layers.data(), lbytes, hipMemcpyHostToDevice,
                          ctx->stream));
    XO_HIP(hipMemcpyAsync(base + off_hdr, hdr.data(), hbytes, hipMemcpyHostToDevice,
                          ctx->stream));
    if (ibytes)
        XO_HIP(hipMemcpyAsync(base + off_items, items.data(), ibytes, hipMemcpyHostToDevice,
                              ctx->stream));

    xo::BatchParams P;
    xo::batch_params(*opts, P);
    P.layers = (const xo::LayerL *)(base + off_layers);
    P.hdr = (const xo::BatchHdr *)(base + off_hdr);
    P.items = (const xo::BatchItem *)(base + off_items);
    P.lut = nlut ? lut_dev : nullptr;
    P.table = (const uint32_t *)base;
    P.item_counter = (unsigned int *)counters;
    P.accum = (unsigned long long *)accum_dev;
    P.stats = counters + 1;
    P.nitems = (uint32_t)items.size();
    P.any_fluence = xo::any_fluence(hdr.data(), hdr.size()) ? 1 : 0;
    uint32_t blocks = 0;
    if (xo::batch_grid(ctx, block, *opts, &blocks)) return -1;

    /* the previous call's counters and the caller's accumulator start at
     * zero (fill kernels, no memset) */
    XO_HIP(xo::launch_zero_u32((uint32_t *)accum_dev, (uint32_t)(naccum * 2), ctx->stream));
    XO_HIP(xo::launch_zero_u32((uint32_t *)counters, (uint32_t)(2 * (1 + XO_STAT_COUNT)),
                               ctx->stream));
    XO_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (total > 0) XO_HIP(xo::launch_batch(P, blocks, block, ctx->stream));
    XO_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    unsigned long long stats[1 + XO_STAT_COUNT];
    XO_HIP(hipMemcpyAsync(stats, counters, sizeof(stats), hipMemcpyDeviceToHost, ctx->stream));
    XO_HIP(hipStreamSynchronize(ctx->stream));
    float kms = 0.f;
    XO_HIP(hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
    xo::fill_report(report, stats + 1, blocks * block, kms, xo::ms_since(t0));
    return 0;
}

} // extern "C"
