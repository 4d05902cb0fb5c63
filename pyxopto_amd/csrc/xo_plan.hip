/*
 * xo_plan.hip — a batch uploaded once and launched many times. Creation
 * validates and packs the instances exactly as xo_batch_run does, uploads
 * the layer records, headers, work items and LUT into buffers the plan
 * owns, and captures the launch — accumulator and counter clears (fill
 * kernels, no memset nodes) and the batch kernel — as a hipGraph. Each run
 * replays the graph and downloads the accumulators: the same integers as
 * xo_batch_run on the same arguments.
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

struct xo_batch_plan {
    xo_ctx *ctx = nullptr;
    void *layers = nullptr, *hdr = nullptr, *items = nullptr;
    float *lut = nullptr;
    unsigned long long *accum = nullptr;
    unsigned long long *counters = nullptr;
    uint64_t naccum = 0, total = 0;
    uint32_t blocks = 0, block = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void plan_free(xo_batch_plan *p) {
    if (!p) return;
    if (p->ctx) (void)hipSetDevice(p->ctx->device);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    void *bufs[] = {p->layers, p->hdr, p->items, p->lut, p->accum, p->counters};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete p;
}

/* Upload, parameters and capture; returns -1 (message set) on error. */
int plan_build(xo_batch_plan *p, const std::vector<xo::LayerL> &layers,
               const std::vector<xo::BatchHdr> &hdr, const std::vector<xo::BatchItem> &items,
               const float *lut, uint32_t nlut, const xo_batch_opts &opts) {
    xo_ctx *ctx = p->ctx;
    const size_t lbytes = layers.size() * sizeof(xo::LayerL);
    const size_t hbytes = hdr.size() * sizeof(xo::BatchHdr);
    const size_t ibytes = (items.empty() ? 1 : items.size()) * sizeof(xo::BatchItem);
    XO_HIP(hipMalloc(&p->layers, lbytes));
    XO_HIP(hipMalloc(&p->hdr, hbytes));
    XO_HIP(hipMalloc(&p->items, ibytes));
    XO_HIP(hipMalloc((void **)&p->accum, p->naccum * 8));
    XO_HIP(hipMalloc((void **)&p->counters, sizeof(unsigned long long) * (1 + XO_STAT_COUNT)));
    if (nlut) XO_HIP(hipMalloc((void **)&p->lut, (size_t)nlut * 4));
    XO_HIP(hipMemcpyAsync(p->layers, layers.data(), lbytes, hipMemcpyHostToDevice, ctx->stream));
    XO_HIP(hipMemcpyAsync(p->hdr, hdr.data(), hbytes, hipMemcpyHostToDevice, ctx->stream));
    if (!items.empty())
        XO_HIP(hipMemcpyAsync(p->items, items.data(), items.size() * sizeof(xo::BatchItem),
                              hipMemcpyHostToDevice, ctx->stream));
    if (nlut)
        XO_HIP(hipMemcpyAsync(p->lut, lut, (size_t)nlut * 4, hipMemcpyHostToDevice, ctx->stream));
    XO_HIP(hipStreamSynchronize(ctx->stream));

    xo::BatchParams P;
    xo::batch_params(opts, P);
    P.layers = (const xo::LayerL *)p->layers;
    P.hdr = (const xo::BatchHdr *)p->hdr;
    P.items = (const xo::BatchItem *)p->items;
    P.lut = nlut ? p->lut : nullptr;
    P.table = ctx->table;
    P.item_counter = (unsigned int *)p->counters;
    P.accum = p->accum;
    P.stats = p->counters + 1;
    P.nitems = (uint32_t)items.size();
    P.any_fluence = xo::any_fluence(hdr.data(), hdr.size()) ? 1 : 0;
    if (xo::batch_grid(ctx, p->block, opts, &p->blocks)) return -1;

    XO_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e1 = xo::launch_zero_u32((uint32_t *)p->accum, (uint32_t)(p->naccum * 2),
                                        ctx->stream);
    hipError_t e2 = xo::launch_zero_u32((uint32_t *)p->counters, 2u * (1u + XO_STAT_COUNT),
                                        ctx->stream);
    hipError_t e3 = p->total > 0 ? xo::launch_batch(P, p->blocks, p->block, ctx->stream)
                                 : hipSuccess;
    hipError_t e4 = hipStreamEndCapture(ctx->stream, &p->graph);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
        xo::set_error("xo_batch_plan_create: capture failed: %s",
                      hipGetErrorString(e4 != hipSuccess ? e4
                                        : e1 != hipSuccess ? e1
                                        : e2 != hipSuccess ? e2 : e3));
        return -1;
    }
    XO_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    return 0;
}

} // namespace

extern "C" {

xo_batch_plan *xo_batch_plan_create(xo_ctx *ctx, const xo_batch_instance *inst,
                                    uint32_t ninst, const float *lut, uint32_t nlut,
                                    const xo_batch_opts *opts, uint64_t naccum) {
    xo::clear_error();
    if (!ctx || !inst || !opts || ninst == 0 || naccum == 0) {
        xo::set_error("xo_batch_plan_create: null or empty argument");
        return nullptr;
    }
    if (naccum > 0xFFFFFFFFull / 2) {
        xo::set_error("xo_batch_plan_create: %llu accumulator bins (at most %llu)",
                      (unsigned long long)naccum, 0xFFFFFFFFull / 2);
        return nullptr;
    }
    std::vector<xo::LayerL> layers;
    std::vector<xo::BatchHdr> hdr;
    std::vector<xo::BatchItem> items;
    xo_batch_plan *p = new xo_batch_plan();
    p->ctx = ctx;
    p->naccum = naccum;
    if (xo::pack_batch("xo_batch_plan_create", inst, ninst, lut != nullptr, nlut, opts, naccum,
                       layers, hdr, items, &p->total, &p->block)) {
        delete p;
        return nullptr;
    }
    int rc;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        rc = hipSetDevice(ctx->device) != hipSuccess ? -1 : 0;
        if (rc) xo::set_error("xo_batch_plan_create: hipSetDevice failed");
        if (!rc) rc = xo::ensure_table(ctx);
        if (!rc) rc = plan_build(p, layers, hdr, items, lut, nlut, *opts);
    }
    if (rc) {
        plan_free(p);
        return nullptr;
    }
    return p;
}

int xo_batch_plan_run(xo_batch_plan *p, uint64_t *accum, xo_run_report *report) {
    auto t0 = std::chrono::steady_clock::now();
    xo::clear_error();
    if (!p || !accum) {
        xo::set_error("xo_batch_plan_run: null argument");
        return -1;
    }
    xo_ctx *ctx = p->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    XO_HIP(hipSetDevice(ctx->device));
    XO_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    XO_HIP(hipGraphLaunch(p->exec, ctx->stream));
    XO_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    unsigned long long stats[1 + XO_STAT_COUNT];
    XO_HIP(hipMemcpyAsync(accum, p->accum, p->naccum * 8, hipMemcpyDeviceToHost, ctx->stream));
    XO_HIP(hipMemcpyAsync(stats, p->counters, sizeof(stats), hipMemcpyDeviceToHost,
                          ctx->stream));
    XO_HIP(hipStreamSynchronize(ctx->stream));
    float kms = 0.f;
    XO_HIP(hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
    xo::fill_report(report, stats + 1, p->blocks * p->block, kms, xo::ms_since(t0));
    return 0;
}

void xo_batch_plan_destroy(xo_batch_plan *p) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lock(p->ctx->mu);
        (void)hipSetDevice(p->ctx->device);
        (void)hipStreamSynchronize(p->ctx->stream);
    }
    plan_free(p);
}

} // extern "C"
