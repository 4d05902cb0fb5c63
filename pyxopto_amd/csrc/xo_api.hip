/*
 * xo_api.hip — the C ABI of libxomc.so: contexts (device-resident RNG
 * streams and buffers, one HIP stream each) and the run entry points that
 * replace Mc.run's upload / McKernel enqueue / download sequence
 * (xopto/mcml/mc.py:730-1018, xopto/mcvox/mc.py:772-1068).
 */
#include <hip/hip_runtime.h>

#include <algorithm>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/xomc.h"
#include "xo_batch.h"
#include "xo_internal.h"
#include "xo_kernels.h"
#include "xo_seedhash.h"

static_assert(sizeof(xo_pf) == 32, "xo_pf layout");
static_assert(sizeof(xo_layer) == 80, "xo_layer layout");
static_assert(sizeof(xo_source_line) == 64, "xo_source_line layout");
static_assert(sizeof(xo_source_gaussian) == 72, "xo_source_gaussian layout");
static_assert(sizeof(xo_detector) == 48, "xo_detector layout");
static_assert(sizeof(xo_fluence) == 52, "xo_fluence layout");
static_assert(sizeof(xo_material) == 64, "xo_material layout");
static_assert(sizeof(xo_voxel_cfg) == 48, "xo_voxel_cfg layout");
static_assert(sizeof(xo_source_fiber) == 32, "xo_source_fiber layout");
static_assert(sizeof(xo_batch_instance) == 1024, "xo_batch_instance layout");
static_assert(sizeof(xo_trace) == 32, "xo_trace layout");
static_assert(sizeof(xo_fiber_axis) == 48, "xo_fiber_axis layout");
static_assert(sizeof(xo_sampling_volume) == 64, "xo_sampling_volume layout");
static_assert(sizeof(xo::McvoxParams) <= 4096, "kernel argument too large");
static_assert(sizeof(xo::McmlParams) <= 4096, "kernel argument too large");

namespace xo {

static thread_local std::string g_error;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
}

/* Entry points start here: the message, and the HIP error a failed call of
 * an earlier, unrelated entry point (a destructor's, whose errors are
 * ignored) may have left for hipGetLastError, which the launch checks
 * read. */
void clear_error() {
    g_error.clear();
    (void)hipGetLastError();
}

} // namespace xo


#include "xo_ctx.h"

namespace xo {

int ensure_buffer(void **ptr, size_t *cap, size_t bytes) {
    if (bytes <= *cap && *ptr) return 0;
    if (*ptr) XO_HIP(hipFree(*ptr));
    *ptr = nullptr;
    *cap = 0;
    size_t alloc = bytes < 256 ? 256 : bytes;
    XO_HIP(hipMalloc(ptr, alloc));
    *cap = alloc;
    return 0;
}

int check_detector(const char *name, const xo_detector &d, uint64_t naccum,
                   uint64_t *end) {
    if (d.kind == XO_DET_NONE) return 0;
    if (d.kind != XO_DET_TOTAL && d.kind != XO_DET_RADIAL) {
        xo::set_error("%s detector: unknown kind %d", name, d.kind);
        return -1;
    }
    uint64_t n = d.kind == XO_DET_TOTAL ? 1u : d.n;
    if (n == 0 || (uint64_t)d.offset + n > naccum) {
        xo::set_error("%s detector: bins [%u, %llu) exceed the accumulator (%llu)", name,
                      d.offset, (unsigned long long)((uint64_t)d.offset + n),
                      (unsigned long long)naccum);
        return -1;
    }
    if ((uint64_t)d.offset + n > *end) *end = (uint64_t)d.offset + n;
    return 0;
}

/* The fluence bins must follow the detector bins (which may be staged in
 * LDS) inside the accumulator. */
int check_fluence(const char *fn, const xo_fluence &f, uint64_t det_end, uint64_t naccum) {
    if (f.enabled == XO_FLUENCE_NONE) return 0;
    if (f.enabled != XO_FLUENCE_XYZ && f.enabled != XO_FLUENCE_RZ) {
        xo::set_error("%s: unknown fluence kind %d", fn, f.enabled);
        return -1;
    }
    const uint64_t nf = (uint64_t)f.shape[0] * f.shape[1] * f.shape[2];
    if (nf == 0 || f.offset < det_end || (uint64_t)f.offset + nf > naccum) {
        xo::set_error("%s: fluence bins [%u, +%llu) must follow the detector bins (%llu) "
                      "inside the accumulator (%llu)",
                      fn, f.offset, (unsigned long long)nf, (unsigned long long)det_end,
                      (unsigned long long)naccum);
        return -1;
    }
    return 0;
}

/* The LDS fluence box of a voxel run: kFlBox bins of a Cartesian grid
 * around the point where the source enters it (centred in x and y,
 * starting at that depth), clamped into the grid; off for other grids, or
 * when the grid or the LDS budget is too small. */
void place_fluence_box(const xo_mcvox_cfg &cfg, McvoxParams &P) {
    P.fl_box_on = 0;
    const xo_fluence &F = cfg.fluence;
    if (F.enabled != XO_FLUENCE_XYZ || F.shape[0] < (uint32_t)kFlBoxX ||
        F.shape[1] < (uint32_t)kFlBoxY || F.shape[2] < (uint32_t)kFlBoxZ ||
        mcvox_lds_bytes(P) > 32768u)
        return;
    const float *sp = cfg.source_kind == XO_SRC_LINE ? cfg.line.position : cfg.gaussian.position;
    const int32_t ext[3] = {kFlBoxX, kFlBoxY, kFlBoxZ};
    const int32_t back[3] = {kFlBoxX / 2, kFlBoxY / 2, 0};
    for (int k = 0; k < 3; ++k) {
        const double f = ((double)sp[k] - F.top_left[k]) * F.inv_step[k];
        int64_t c = std::isfinite(f) ? (int64_t)std::floor(f) - back[k] : 0;
        const int64_t hi = (int64_t)F.shape[k] - ext[k];
        c = c < 0 ? 0 : (c > hi ? hi : c);
        P.fl_box[k] = (int32_t)c;
    }
    P.fl_box_on = 1;
}

int check_pf(int pf_kind, const xo_pf &pf, uint32_t nlut, const char *what, uint32_t idx) {
    if (pf_kind == XO_PF_LUT) {
        if (pf.size < 2 || (uint64_t)pf.offset + pf.size > nlut) {
            xo::set_error("%s %u: lookup table [%u, %u+%u) outside the LUT array (%u)", what,
                          idx, pf.offset, pf.offset, pf.size, nlut);
            return -1;
        }
    } else if (pf_kind != XO_PF_HG && pf_kind != XO_PF_MHG) {
        xo::set_error("unknown phase function kind %d", pf_kind);
        return -1;
    } else if (!(pf.g >= -1.0f && pf.g <= 1.0f)) {
        /* the Python classes clip g to [-1, 1]; the sampling relies on it */
        xo::set_error("%s %u: anisotropy g = %g outside [-1, 1]", what, idx, (double)pf.g);
        return -1;
    }
    return 0;
}

/* The seed-table multipliers (rows 1..) on the device, for keyed packets. */
int ensure_table(xo_ctx *ctx) {
    if (ctx->table) return 0;
    std::vector<uint32_t> fora(XO_NUM_PRIMES);
    if (xo::table_get(XO_NUM_PRIMES, fora.data()) != 0) {
        xo::set_error("safe-prime table unavailable");
        return -1;
    }
    XO_HIP(hipMalloc(&ctx->table, sizeof(uint32_t) * XO_SEED_TABLE_ROWS));
    XO_HIP(hipMemcpy(ctx->table, fora.data() + 1, sizeof(uint32_t) * XO_SEED_TABLE_ROWS,
                     hipMemcpyHostToDevice));
    return 0;
}

uint32_t choose_lanes(xo_ctx *ctx, uint32_t requested, int blocks_per_cu, uint64_t npackets) {
    uint32_t lanes = requested;
    if (lanes == 0) {
        uint64_t persistent =
            (uint64_t)(blocks_per_cu > 0 ? blocks_per_cu : 1) * ctx->cus * XO_BLOCK;
        /* a run of fewer than kPacketsPerLane packets per lane of the
         * persistent grid: one lane per kPacketsPerLane packets, at least
         * 65,536 (one wave per SIMD); the extra resident waves would only
         * slow the issue of the ones that work (1 mm slab, 1e6 packets:
         * 0.353 ms on 196,608 lanes against 0.392 ms on the persistent
         * 458,752; 1e4 packets: 0.168 ms on 65,536 against 0.228 ms) */
        constexpr uint64_t kPacketsPerLane = 5;
        const uint64_t want =
            ((npackets / kPacketsPerLane + XO_BLOCK - 1) / XO_BLOCK) * XO_BLOCK;
        const uint64_t floor = 65536u;
        if (want < persistent) persistent = want > floor ? want : floor;
        lanes = (uint32_t)(persistent < ctx->max_lanes ? persistent : ctx->max_lanes);
    }
    if (lanes > ctx->max_lanes) lanes = ctx->max_lanes;
    return lanes;
}

uint64_t brick_bytes(const int32_t shape[3]) {
    const uint64_t nbx = ((uint64_t)shape[0] + 3) / 4, nby = ((uint64_t)shape[1] + 3) / 4,
                   nbz = ((uint64_t)shape[2] + 3) / 4;
    return nbx * nby * nbz * 64u;
}

int pack_voxels(const int32_t *voxels, const int32_t shape[3], uint32_t num_materials,
                std::vector<uint8_t> &out, const char *fn) {
    const uint32_t nx = (uint32_t)shape[0], ny = (uint32_t)shape[1], nz = (uint32_t)shape[2];
    const uint32_t nbx = (nx + 3) / 4, nby = (ny + 3) / 4;
    out.assign(brick_bytes(shape), 0u);
    for (uint32_t z = 0; z < nz; ++z)
        for (uint32_t y = 0; y < ny; ++y) {
            const int32_t *row = voxels + ((uint64_t)z * ny + y) * nx;
            const uint64_t base = ((uint64_t)(z >> 2) * nby + (y >> 2)) * nbx;
            const uint32_t inner = ((z & 3u) << 4) | ((y & 3u) << 2);
            for (uint32_t x = 0; x < nx; ++x) {
                const int32_t m = row[x];
                if (m < 0 || (uint32_t)m >= num_materials) {
                    xo::set_error("%s: voxel (x %u, y %u, z %u) has material index %d "
                                  "(valid 0..%u)",
                                  fn, x, y, z, m, num_materials - 1);
                    return -1;
                }
                out[(base + (x >> 2)) * 64u + (inner | (x & 3u))] = (uint8_t)m;
            }
        }
    return 0;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
        .count();
}

void fill_report(xo_run_report *r, const unsigned long long *stats, uint32_t lanes,
                 float kernel_ms, double total_ms) {
    if (!r) return;
    r->packets = stats[XO_STAT_PACKETS];
    r->steps = stats[XO_STAT_STEPS];
    r->boundary_steps = stats[XO_STAT_BOUNDARY];
    r->voxel_crossings = stats[XO_STAT_CROSSINGS];
    r->deposits = stats[XO_STAT_DEPOSITS];
    r->truncated = stats[XO_STAT_TRUNCATED];
    r->lanes = lanes;
    r->threads_used = (uint32_t)stats[XO_STAT_THREADS];
    r->kernel_ms = kernel_ms;
    r->total_ms = total_ms;
}

int pack_batch(const char *fn, const xo_batch_instance *inst, uint32_t ninst, bool have_lut,
               uint32_t nlut, const xo_batch_opts *opts, uint64_t naccum,
               std::vector<LayerL> &layers, std::vector<BatchHdr> &hdr,
               std::vector<BatchItem> &items, uint64_t *total, uint32_t *block_out) {
    const uint32_t block = opts->block ? opts->block : 256u;
    if (block != 64 && block != 128 && block != 256) {
        set_error("%s: block must be 64, 128 or 256 (got %u)", fn, block);
        return -1;
    }
    const uint64_t chunk = opts->chunk ? opts->chunk : 1024u;
    if (opts->method != XO_METHOD_AW && opts->method != XO_METHOD_AR &&
        opts->method != XO_METHOD_MBL) {
        set_error("%s: unknown stepping method %d", fn, opts->method);
        return -1;
    }
    layers.assign((size_t)ninst * XO_BATCH_MAX_LAYERS, LayerL{});
    hdr.assign(ninst, BatchHdr{});
    items.clear();
    *total = 0;
    for (uint32_t i = 0; i < ninst; ++i) {
        const xo_batch_instance &I = inst[i];
        if (I.num_layers < 3 || I.num_layers > XO_BATCH_MAX_LAYERS) {
            set_error("instance %u: %u layers (3..%d supported)", i, I.num_layers,
                      XO_BATCH_MAX_LAYERS);
            return -1;
        }
        for (uint32_t l = 0; l < I.num_layers; ++l)
            if (check_pf(I.pf_kind, I.layers[l].pf, nlut, "layer", l)) return -1;
        if (I.pf_kind == XO_PF_LUT && !have_lut) {
            set_error("instance %u: a lookup-table phase function needs the LUT", i);
            return -1;
        }
        if ((uint64_t)I.accum_base + I.accum_count > naccum) {
            set_error("instance %u: bins [%u, +%u) outside the accumulator (%llu)", i,
                      I.accum_base, I.accum_count, (unsigned long long)naccum);
            return -1;
        }
        uint64_t end = 0;
        const xo_detector *dets[3] = {&I.top, &I.bottom, &I.specular};
        for (int k = 0; k < 3; ++k) {
            if (check_detector("instance", *dets[k], naccum, &end)) return -1;
            if (dets[k]->kind != XO_DET_NONE &&
                (dets[k]->offset < I.accum_base ||
                 (uint64_t)dets[k]->offset +
                         (dets[k]->kind == XO_DET_TOTAL ? 1u : dets[k]->n) >
                     (uint64_t)I.accum_base + I.accum_count)) {
                set_error("instance %u: detector bins outside its range", i);
                return -1;
            }
        }
        for (uint32_t l = 0; l < XO_BATCH_MAX_LAYERS; ++l)
            layers[(size_t)i * XO_BATCH_MAX_LAYERS + l] =
                layer_to_device(I.layers[l < I.num_layers ? l : I.num_layers - 1]);
        if (I.source_kind != XO_SRC_LINE && I.source_kind != XO_SRC_FIBER_NI &&
            I.source_kind != XO_SRC_FIBER) {
            set_error("instance %u: unknown source kind %d", i, I.source_kind);
            return -1;
        }
        if (I.source_kind == XO_SRC_FIBER && !(I.fiber_axis.direction[2] > 0.0f)) {
            set_error("instance %u: the fiber direction must point into the sample (z > 0)",
                      i);
            return -1;
        }
        if (I.fluence.enabled != XO_FLUENCE_NONE) {
            if (check_fluence("instance", I.fluence, end, naccum)) return -1;
            const xo_fluence &F = I.fluence;
            const uint64_t nf = (uint64_t)F.shape[0] * F.shape[1] * F.shape[2];
            if ((uint64_t)F.offset + nf > (uint64_t)I.accum_base + I.accum_count) {
                set_error("instance %u: fluence bins outside its range", i);
                return -1;
            }
        }
        const xo_detector det[3] = {I.top, I.bottom, I.specular};
        hdr[i] = batch_header(I, det, 0u);
        for (uint64_t first = 0; first < I.npackets; first += chunk) {
            uint64_t cnt = I.npackets - first < chunk ? I.npackets - first : chunk;
            items.push_back(BatchItem{i, (uint32_t)cnt, first});
        }
        *total += I.npackets;
        if (items.size() >= 0xFFFFFFF0ull) {
            set_error("%s: too many work items; raise the chunk size", fn);
            return -1;
        }
    }

    /* Waves claim consecutive items: grouping the items of instances that
     * run the same code (phase-function kind, source, layer count) keeps the
     * lanes of a wave on one path when it moves from item to item. The
     * order does not change any result. */
    std::stable_sort(items.begin(), items.end(), [inst](const BatchItem &p, const BatchItem &q) {
        return item_class(inst[p.inst]) < item_class(inst[q.inst]);
    });
    *block_out = block;
    return 0;
}

void batch_params(const xo_batch_opts &opts, BatchParams &P) {
    std::memset(&P, 0, sizeof(P));
    P.w_min = opts.w_min;
    P.chance = opts.lottery_chance;
    P.use_lottery = opts.use_lottery;
    P.refill_min = opts.refill_min ? opts.refill_min : 4u;
    P.defer_min = opts.defer_min ? opts.defer_min : kDeferMinDefault;
    P.method = opts.method;
    P.fast_math = (opts.flags & XO_FLAG_FAST_MATH) ? 1 : 0;
}

int batch_grid(xo_ctx *ctx, uint32_t block, const xo_batch_opts &opts, uint32_t *blocks) {
    int bpc = 1;
    XO_HIP(occupancy_batch(block, &bpc));
    if (opts.blocks_per_cu && (int)opts.blocks_per_cu < bpc) bpc = (int)opts.blocks_per_cu;
    if (bpc < 1) bpc = 1;
    *blocks = (uint32_t)bpc * ctx->cus;
    return 0;
}

} // namespace xo

using xo::check_detector;
using xo::check_fluence;
using xo::check_pf;
using xo::choose_lanes;
using xo::ensure_buffer;
using xo::ensure_table;
using xo::ms_since;
using xo::fill_report;

extern "C" {

int xo_abi_version(void) { return XO_ABI_VERSION; }

const char *xo_last_error(void) { return xo::g_error.c_str(); }

int xo_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int xo_struct_sizes(uint32_t *out, int n) {
    const uint32_t sizes[] = {
        (uint32_t)sizeof(xo_pf),          (uint32_t)sizeof(xo_layer),
        (uint32_t)sizeof(xo_source_line), (uint32_t)sizeof(xo_source_gaussian),
        (uint32_t)sizeof(xo_detector),    (uint32_t)sizeof(xo_fluence),
        (uint32_t)sizeof(xo_mcml_cfg),    (uint32_t)sizeof(xo_material),
        (uint32_t)sizeof(xo_voxel_cfg),   (uint32_t)sizeof(xo_mcvox_cfg),
        (uint32_t)sizeof(xo_run_report),  (uint32_t)sizeof(xo_batch_instance),
        (uint32_t)sizeof(xo_batch_opts),  (uint32_t)sizeof(xo_serve_opts),
        (uint32_t)sizeof(xo_serve_stats), (uint32_t)sizeof(xo_source_fiber),
        (uint32_t)sizeof(xo_trace),       (uint32_t)sizeof(xo_sampling_volume),
        (uint32_t)sizeof(xo_fiber_axis)};
    int count = (int)(sizeof(sizes) / sizeof(sizes[0]));
    if (n < count) count = n;
    for (int i = 0; i < count; ++i) out[i] = sizes[i];
    return count;
}

xo_ctx *xo_ctx_create(int device, uint32_t max_lanes, uint64_t xinit) {
    xo::clear_error();
    if (max_lanes == 0 || max_lanes > XO_MAX_SEEDS) max_lanes = XO_MAX_SEEDS;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        xo::set_error("no HIP device is available");
        return nullptr;
    }
    if (device < 0 || device >= ndev) {
        xo::set_error("device %d out of range (%d devices)", device, ndev);
        return nullptr;
    }
    std::vector<uint64_t> hx(max_lanes);
    std::vector<uint32_t> ha(max_lanes);
    int rc = xo_init_rng(max_lanes, xinit, hx.data(), ha.data());
    if (rc == 1) {
        xo::set_error("The provided initializer %llu is not valid. Try a different 64-bit "
                      "unsigned integer!",
                      (unsigned long long)xinit);
        return nullptr;
    }
    if (rc != 0) return nullptr;

    xo_ctx *ctx = new xo_ctx();
    ctx->device = device;
    ctx->max_lanes = max_lanes;
    auto fail = [&](const char *what, hipError_t e) -> xo_ctx * {
        xo::set_error("%s failed: %s", what, hipGetErrorString(e));
        xo_ctx_destroy(ctx);
        return nullptr;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return fail("hipSetDevice", e);
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess)
        return fail("hipGetDeviceProperties", e);
    ctx->cus = (uint32_t)prop.multiProcessorCount;
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail("hipStreamCreate", e);
    if ((e = hipEventCreate(&ctx->ev0)) != hipSuccess) return fail("hipEventCreate", e);
    if ((e = hipEventCreate(&ctx->ev1)) != hipSuccess) return fail("hipEventCreate", e);
    if ((e = hipMalloc(&ctx->rng_x, sizeof(uint64_t) * max_lanes)) != hipSuccess)
        return fail("hipMalloc(rng_x)", e);
    if ((e = hipMalloc(&ctx->rng_a, sizeof(uint32_t) * max_lanes)) != hipSuccess)
        return fail("hipMalloc(rng_a)", e);
    if ((e = hipMalloc(&ctx->counters, sizeof(unsigned long long) * (1 + XO_STAT_COUNT))) !=
        hipSuccess)
        return fail("hipMalloc(counters)", e);
    if ((e = hipMemcpy(ctx->rng_x, hx.data(), sizeof(uint64_t) * max_lanes,
                       hipMemcpyHostToDevice)) != hipSuccess)
        return fail("hipMemcpy(rng_x)", e);
    if ((e = hipMemcpy(ctx->rng_a, ha.data(), sizeof(uint32_t) * max_lanes,
                       hipMemcpyHostToDevice)) != hipSuccess)
        return fail("hipMemcpy(rng_a)", e);
    return ctx;
}

void xo_ctx_destroy(xo_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->rng_x) (void)hipFree(ctx->rng_x);
    if (ctx->rng_a) (void)hipFree(ctx->rng_a);
    if (ctx->counters) (void)hipFree(ctx->counters);
    if (ctx->accum) (void)hipFree(ctx->accum);
    if (ctx->lut) (void)hipFree(ctx->lut);
    if (ctx->voxels) (void)hipFree(ctx->voxels);
    if (ctx->table) (void)hipFree(ctx->table);
    if (ctx->batch_layers) (void)hipFree(ctx->batch_layers);
    if (ctx->batch_hdr) (void)hipFree(ctx->batch_hdr);
    if (ctx->batch_items) (void)hipFree(ctx->batch_items);
    if (ctx->trace_data) (void)hipFree(ctx->trace_data);
    if (ctx->trace_n) (void)hipFree(ctx->trace_n);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int xo_ctx_info(xo_ctx *ctx, uint32_t *persistent_lanes, uint32_t *max_lanes,
                uint32_t *compute_units) {
    if (!ctx) {
        xo::set_error("null context");
        return -1;
    }
    int bpc = 1;
    XO_HIP(hipSetDevice(ctx->device));
    XO_HIP(xo::occupancy_mcml(XO_PF_HG, XO_MAX_LAYERS * 64 + 8 * 102, false, &bpc));
    if (persistent_lanes) *persistent_lanes = choose_lanes(ctx, 0, bpc, ~0ull);
    if (max_lanes) *max_lanes = ctx->max_lanes;
    if (compute_units) *compute_units = ctx->cus;
    return 0;
}

int xo_ctx_get_rng(xo_ctx *ctx, uint32_t n, uint64_t *x_out, uint32_t *a_out) {
    if (!ctx || n > ctx->max_lanes) {
        xo::set_error("xo_ctx_get_rng: bad context or n");
        return -1;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    XO_HIP(hipSetDevice(ctx->device));
    if (x_out)
        XO_HIP(hipMemcpyAsync(x_out, ctx->rng_x, sizeof(uint64_t) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
    if (a_out)
        XO_HIP(hipMemcpyAsync(a_out, ctx->rng_a, sizeof(uint32_t) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
    XO_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int xo_ctx_set_rng(xo_ctx *ctx, uint32_t n, const uint64_t *x, const uint32_t *a) {
    if (!ctx || n > ctx->max_lanes) {
        xo::set_error("xo_ctx_set_rng: bad context or n");
        return -1;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    XO_HIP(hipSetDevice(ctx->device));
    if (x)
        XO_HIP(hipMemcpyAsync(ctx->rng_x, x, sizeof(uint64_t) * n, hipMemcpyHostToDevice,
                              ctx->stream));
    if (a)
        XO_HIP(hipMemcpyAsync(ctx->rng_a, a, sizeof(uint32_t) * n, hipMemcpyHostToDevice,
                              ctx->stream));
    XO_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}


/* Trace arguments of xo_mcml_run_trace / xo_mcvox_run_trace; *floats: the
 * size of the trace data. */
static int check_trace(const char *fn, const xo_trace *trace, const float *data,
                       const int32_t *n, uint64_t npackets, uint64_t *floats) {
    if (!data || !n) {
        xo::set_error("%s: null trace buffer", fn);
        return -1;
    }
    if (trace->max_events < 1 ||
        (trace->options != XO_TRACE_START && trace->options != XO_TRACE_END &&
         trace->options != (XO_TRACE_START | XO_TRACE_END) &&
         trace->options != XO_TRACE_ALL)) {
        xo::set_error("%s: max_events %d / options %u invalid", fn, trace->max_events,
                      trace->options);
        return -1;
    }
    const uint64_t per = (uint64_t)trace->max_events * XO_TRACE_ENTRY_LEN;
    if (npackets > (UINT64_MAX / 4u) / per) {
        xo::set_error("%s: trace of %llu packets too large", fn, (unsigned long long)npackets);
        return -1;
    }
    *floats = npackets * per;
    return 0;
}

/* Device trace buffers of one run, zeroed on the stream (entries past a
 * packet's last event read as zeros). */
static int trace_buffers(xo_ctx *ctx, uint64_t floats, uint64_t npackets) {
    if (!npackets) return 0;
    if (ensure_buffer((void **)&ctx->trace_data, &ctx->trace_data_cap, floats * 4) ||
        ensure_buffer((void **)&ctx->trace_n, &ctx->trace_n_cap, npackets * 4))
        return -1;
    XO_HIP(hipMemsetAsync(ctx->trace_data, 0, floats * 4, ctx->stream));
    return 0;
}

static int trace_download(xo_ctx *ctx, float *data, int32_t *n, uint64_t floats,
                          uint64_t npackets) {
    if (!npackets) return 0;
    XO_HIP(hipMemcpyAsync(data, ctx->trace_data, floats * 4, hipMemcpyDeviceToHost,
                          ctx->stream));
    XO_HIP(hipMemcpyAsync(n, ctx->trace_n, npackets * 4, hipMemcpyDeviceToHost, ctx->stream));
    return 0;
}

/* xo_mcml_run and xo_mcml_run_trace; trace NULL: no trace. */
static int mcml_run(xo_ctx *ctx, const xo_mcml_cfg *cfg, const float *lut, uint32_t nlut,
                    uint64_t npackets, uint32_t lanes, uint32_t flags, const xo_trace *trace,
                    float *trace_data, int32_t *trace_n, uint64_t *accum, uint64_t naccum,
                    xo_run_report *report) {
    auto t0 = std::chrono::steady_clock::now();
    xo::clear_error();
    if (!ctx || !cfg || !accum) {
        xo::set_error("xo_mcml_run: null argument");
        return -1;
    }
    uint64_t trace_floats = 0;
    if (trace && check_trace("xo_mcml_run_trace", trace, trace_data, trace_n, npackets,
                             &trace_floats))
        return -1;
    if (cfg->num_layers < 3 || cfg->num_layers > XO_MAX_LAYERS) {
        xo::set_error("xo_mcml_run: %u layers (3..%d supported)", cfg->num_layers,
                      XO_MAX_LAYERS);
        return -1;
    }
    if (cfg->method != XO_METHOD_AW && cfg->method != XO_METHOD_AR &&
        cfg->method != XO_METHOD_MBL) {
        xo::set_error("xo_mcml_run: unknown stepping method %d", cfg->method);
        return -1;
    }
    if (cfg->source_kind != XO_SRC_LINE && cfg->source_kind != XO_SRC_FIBER_NI &&
        cfg->source_kind != XO_SRC_FIBER) {
        xo::set_error("xo_mcml_run: unknown source kind %d", cfg->source_kind);
        return -1;
    }
    if (cfg->source_kind == XO_SRC_FIBER && !(cfg->fiber_axis.direction[2] > 0.0f)) {
        xo::set_error("xo_mcml_run: the fiber direction must point into the sample (z > 0)");
        return -1;
    }
    if (naccum == 0) {
        xo::set_error("xo_mcml_run: empty accumulator");
        return -1;
    }
    for (uint32_t i = 0; i < cfg->num_layers; ++i)
        if (check_pf(cfg->pf_kind, cfg->layers[i].pf, nlut, "layer", i)) return -1;
    if (cfg->pf_kind == XO_PF_LUT && !lut) {
        xo::set_error("xo_mcml_run: a lookup-table phase function needs the LUT array");
        return -1;
    }
    uint64_t det_end = 0;
    if (check_detector("top", cfg->top, naccum, &det_end) ||
        check_detector("bottom", cfg->bottom, naccum, &det_end) ||
        check_detector("specular", cfg->specular, naccum, &det_end))
        return -1;
    if (check_fluence("xo_mcml_run", cfg->fluence, det_end, naccum)) return -1;

    std::lock_guard<std::mutex> lock(ctx->mu);
    XO_HIP(hipSetDevice(ctx->device));
    if (ensure_buffer((void **)&ctx->accum, &ctx->accum_cap, naccum * 8)) return -1;
    if (nlut && ensure_buffer((void **)&ctx->lut, &ctx->lut_cap, (size_t)nlut * 4)) return -1;
    if (trace && trace_buffers(ctx, trace_floats, npackets)) return -1;

    if ((flags & XO_FLAG_PACKET_KEYED) && ensure_table(ctx)) return -1;
    xo::McmlParams P;
    std::memset(&P, 0, sizeof(P));
    P.cfg = *cfg;
    if (trace) {
        P.trace = *trace;
        P.trace_data = ctx->trace_data;
        P.trace_n = ctx->trace_n;
    }
    P.table = ctx->table;
    P.lut = nlut ? ctx->lut : nullptr;
    P.rng_x = ctx->rng_x;
    P.rng_a = ctx->rng_a;
    P.counter = ctx->counters;
    P.stats = ctx->counters + 1;
    P.accum = ctx->accum;
    P.npackets = npackets;
    P.flags = flags;
    P.lds_bins = (!(flags & XO_FLAG_NO_LDS_BINS) && naccum <= XO_LDS_BINS_MAX)
                     ? (uint32_t)naccum
                     : 0u;
    P.lds_lut = (nlut && nlut <= XO_LDS_LUT_MAX) ? nlut : 0u;
    int bpc = 1;
    XO_HIP(xo::occupancy_mcml(cfg->pf_kind, xo::mcml_lds_bytes(P),
                                (flags & XO_FLAG_FAST_MATH) != 0, &bpc));
    P.lanes = choose_lanes(ctx, lanes, bpc, npackets);
    uint32_t blocks = (P.lanes + XO_BLOCK - 1) / XO_BLOCK;

    if (nlut)
        XO_HIP(hipMemcpyAsync(ctx->lut, lut, (size_t)nlut * 4, hipMemcpyHostToDevice,
                              ctx->stream));
    XO_HIP(hipMemsetAsync(ctx->accum, 0, naccum * 8, ctx->stream));
    XO_HIP(hipMemsetAsync(ctx->counters, 0, sizeof(unsigned long long) * (1 + XO_STAT_COUNT),
                          ctx->stream));
    XO_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (npackets > 0) XO_HIP(xo::launch_mcml(P, blocks, ctx->stream));
    XO_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    unsigned long long stats[1 + XO_STAT_COUNT];
    XO_HIP(hipMemcpyAsync(accum, ctx->accum, naccum * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (trace && trace_download(ctx, trace_data, trace_n, trace_floats, npackets)) return -1;
    XO_HIP(hipMemcpyAsync(stats, ctx->counters, sizeof(stats), hipMemcpyDeviceToHost,
                          ctx->stream));
    XO_HIP(hipStreamSynchronize(ctx->stream));
    float kms = 0.f;
    XO_HIP(hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
    fill_report(report, stats + 1, P.lanes, kms, ms_since(t0));
    return 0;
}

int xo_mcml_run(xo_ctx *ctx, const xo_mcml_cfg *cfg, const float *lut, uint32_t nlut,
                uint64_t npackets, uint32_t lanes, uint32_t flags, uint64_t *accum,
                uint64_t naccum, xo_run_report *report) {
    return mcml_run(ctx, cfg, lut, nlut, npackets, lanes, flags, nullptr, nullptr, nullptr,
                    accum, naccum, report);
}

int xo_mcml_run_trace(xo_ctx *ctx, const xo_mcml_cfg *cfg, const float *lut, uint32_t nlut,
                      uint64_t npackets, uint32_t lanes, uint32_t flags, const xo_trace *trace,
                      float *trace_data, int32_t *trace_n, uint64_t *accum, uint64_t naccum,
                      xo_run_report *report) {
    if (!trace) {
        xo::clear_error();
        xo::set_error("xo_mcml_run_trace: null trace");
        return -1;
    }
    return mcml_run(ctx, cfg, lut, nlut, npackets, lanes, flags, trace, trace_data, trace_n,
                    accum, naccum, report);
}

/* xo_mcvox_run and xo_mcvox_run_trace; trace NULL: no trace. */
static int mcvox_run(xo_ctx *ctx, const xo_mcvox_cfg *cfg, const int32_t *voxels,
                     uint64_t nvoxels, const float *lut, uint32_t nlut, uint64_t npackets,
                     uint32_t lanes, uint32_t flags, const xo_trace *trace, float *trace_data,
                     int32_t *trace_n, uint64_t *accum, uint64_t naccum,
                     xo_run_report *report) {
    auto t0 = std::chrono::steady_clock::now();
    xo::clear_error();
    if (!ctx || !cfg || !accum) {
        xo::set_error("xo_mcvox_run: null argument");
        return -1;
    }
    uint64_t trace_floats = 0;
    if (trace && check_trace("xo_mcvox_run_trace", trace, trace_data, trace_n, npackets,
                             &trace_floats))
        return -1;
    if (cfg->num_materials < 1 || cfg->num_materials > XO_MAX_MATERIALS) {
        xo::set_error("xo_mcvox_run: %u materials (1..%d supported)", cfg->num_materials,
                      XO_MAX_MATERIALS);
        return -1;
    }
    if (cfg->method != XO_METHOD_AW && cfg->method != XO_METHOD_AR &&
        cfg->method != XO_METHOD_MBL) {
        xo::set_error("xo_mcvox_run: unknown stepping method %d", cfg->method);
        return -1;
    }
    if (cfg->source_kind != XO_SRC_LINE && cfg->source_kind != XO_SRC_GAUSSIAN) {
        xo::set_error("xo_mcvox_run: unknown source kind %d", cfg->source_kind);
        return -1;
    }
    const int32_t *shape = cfg->voxels.shape;
    if (shape[0] <= 0 || shape[1] <= 0 || shape[2] <= 0) {
        xo::set_error("xo_mcvox_run: invalid voxel grid shape");
        return -1;
    }
    const uint64_t nvox = (uint64_t)shape[0] * shape[1] * shape[2];
    if (naccum == 0) {
        xo::set_error("xo_mcvox_run: empty accumulator");
        return -1;
    }
    for (uint32_t i = 0; i < cfg->num_materials; ++i)
        if (check_pf(cfg->pf_kind, cfg->materials[i].pf, nlut, "material", i)) return -1;
    if (cfg->pf_kind == XO_PF_LUT && !lut) {
        xo::set_error("xo_mcvox_run: a lookup-table phase function needs the LUT array");
        return -1;
    }
    uint64_t det_end = 0;
    if (check_detector("top", cfg->top, naccum, &det_end) ||
        check_detector("bottom", cfg->bottom, naccum, &det_end) ||
        check_detector("specular", cfg->specular, naccum, &det_end))
        return -1;
    if (cfg->fluence.enabled) {
        const uint32_t *fs = cfg->fluence.shape;
        uint64_t nf = (uint64_t)fs[0] * fs[1] * fs[2];
        if (nf == 0 || cfg->fluence.offset < det_end ||
            (uint64_t)cfg->fluence.offset + nf > naccum) {
            xo::set_error("xo_mcvox_run: fluence bins [%u, +%llu) must follow the detector "
                          "bins (%llu) inside the accumulator (%llu)",
                          cfg->fluence.offset, (unsigned long long)nf,
                          (unsigned long long)det_end, (unsigned long long)naccum);
            return -1;
        }
    }

    std::lock_guard<std::mutex> lock(ctx->mu);
    XO_HIP(hipSetDevice(ctx->device));
    if (voxels) {
        if (nvoxels != nvox) {
            xo::set_error("xo_mcvox_run: %llu voxels given, grid has %llu",
                          (unsigned long long)nvoxels, (unsigned long long)nvox);
            return -1;
        }
        std::vector<uint8_t> packed;
        if (xo::pack_voxels(voxels, shape, cfg->num_materials, packed, "xo_mcvox_run"))
            return -1;
        ctx->vox_valid = false;
        if (ensure_buffer((void **)&ctx->voxels, &ctx->voxels_cap, packed.size())) return -1;
        XO_HIP(hipMemcpyAsync(ctx->voxels, packed.data(), packed.size(), hipMemcpyHostToDevice,
                              ctx->stream));
        XO_HIP(hipStreamSynchronize(ctx->stream));
        std::memcpy(ctx->vox_shape, shape, sizeof(ctx->vox_shape));
        ctx->vox_valid = true;
    } else if (!ctx->vox_valid || std::memcmp(ctx->vox_shape, shape, sizeof(ctx->vox_shape))) {
        xo::set_error("xo_mcvox_run: no voxel grid of this shape has been uploaded");
        return -1;
    }
    if (ensure_buffer((void **)&ctx->accum, &ctx->accum_cap, naccum * 8)) return -1;
    if (nlut && ensure_buffer((void **)&ctx->lut, &ctx->lut_cap, (size_t)nlut * 4)) return -1;

    if (trace && trace_buffers(ctx, trace_floats, npackets)) return -1;

    if ((flags & XO_FLAG_PACKET_KEYED) && ensure_table(ctx)) return -1;
    xo::McvoxParams P;
    std::memset(&P, 0, sizeof(P));
    P.cfg = *cfg;
    if (trace) {
        P.trace = *trace;
        P.trace_data = ctx->trace_data;
        P.trace_n = ctx->trace_n;
    }
    P.table = ctx->table;
    P.voxels = ctx->voxels;
    P.lut = nlut ? ctx->lut : nullptr;
    P.rng_x = ctx->rng_x;
    P.rng_a = ctx->rng_a;
    P.counter = ctx->counters;
    P.stats = ctx->counters + 1;
    P.accum = ctx->accum;
    P.npackets = npackets;
    P.flags = flags;
    P.lds_bins = (!(flags & XO_FLAG_NO_LDS_BINS) && det_end <= XO_LDS_BINS_MAX)
                     ? (uint32_t)det_end
                     : 0u;
    P.lds_lut = (nlut && nlut <= XO_LDS_LUT_MAX) ? nlut : 0u;
    xo::place_fluence_box(*cfg, P);
    int bpc = 1;
    XO_HIP(xo::occupancy_mcvox(cfg->pf_kind, xo::mcvox_lds_bytes(P),
                                 (flags & XO_FLAG_FAST_MATH) != 0, &bpc));
    P.lanes = choose_lanes(ctx, lanes, bpc, npackets);
    uint32_t blocks = (P.lanes + XO_BLOCK - 1) / XO_BLOCK;

    if (nlut)
        XO_HIP(hipMemcpyAsync(ctx->lut, lut, (size_t)nlut * 4, hipMemcpyHostToDevice,
                              ctx->stream));
    XO_HIP(hipMemsetAsync(ctx->accum, 0, naccum * 8, ctx->stream));
    XO_HIP(hipMemsetAsync(ctx->counters, 0, sizeof(unsigned long long) * (1 + XO_STAT_COUNT),
                          ctx->stream));
    XO_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (npackets > 0) XO_HIP(xo::launch_mcvox(P, blocks, ctx->stream));
    XO_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    unsigned long long stats[1 + XO_STAT_COUNT];
    XO_HIP(hipMemcpyAsync(accum, ctx->accum, naccum * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (trace && trace_download(ctx, trace_data, trace_n, trace_floats, npackets)) return -1;
    XO_HIP(hipMemcpyAsync(stats, ctx->counters, sizeof(stats), hipMemcpyDeviceToHost,
                          ctx->stream));
    XO_HIP(hipStreamSynchronize(ctx->stream));
    float kms = 0.f;
    XO_HIP(hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
    fill_report(report, stats + 1, P.lanes, kms, ms_since(t0));
    return 0;
}

int xo_mcvox_run(xo_ctx *ctx, const xo_mcvox_cfg *cfg, const int32_t *voxels,
                 uint64_t nvoxels, const float *lut, uint32_t nlut, uint64_t npackets,
                 uint32_t lanes, uint32_t flags, uint64_t *accum, uint64_t naccum,
                 xo_run_report *report) {
    return mcvox_run(ctx, cfg, voxels, nvoxels, lut, nlut, npackets, lanes, flags, nullptr,
                     nullptr, nullptr, accum, naccum, report);
}

int xo_mcvox_run_trace(xo_ctx *ctx, const xo_mcvox_cfg *cfg, const int32_t *voxels,
                       uint64_t nvoxels, const float *lut, uint32_t nlut, uint64_t npackets,
                       uint32_t lanes, uint32_t flags, const xo_trace *trace, float *trace_data,
                       int32_t *trace_n, uint64_t *accum, uint64_t naccum,
                       xo_run_report *report) {
    if (!trace) {
        xo::clear_error();
        xo::set_error("xo_mcvox_run_trace: null trace");
        return -1;
    }
    return mcvox_run(ctx, cfg, voxels, nvoxels, lut, nlut, npackets, lanes, flags, trace,
                     trace_data, trace_n, accum, naccum, report);
}

int xo_packet_seed(uint64_t seed, uint64_t p, uint64_t *x, uint32_t *a) {
    static std::vector<uint32_t> table;
    static std::mutex m;
    std::lock_guard<std::mutex> lock(m);
    if (table.empty()) {
        std::vector<uint32_t> fora(XO_NUM_PRIMES);
        if (xo::table_get(XO_NUM_PRIMES, fora.data()) != 0) {
            xo::set_error("xo_packet_seed: safe-prime table unavailable");
            return -1;
        }
        table.assign(fora.begin() + 1, fora.end());
    }
    xo_packet_state(seed, p, table.data(), x, a);
    return 0;
}

int xo_batch_run(xo_ctx *ctx, const xo_batch_instance *inst, uint32_t ninst,
                 const float *lut, uint32_t nlut, const xo_batch_opts *opts,
                 uint64_t *accum, uint64_t naccum, xo_run_report *report) {
    auto t0 = std::chrono::steady_clock::now();
    xo::clear_error();
    if (!ctx || !inst || !opts || !accum || ninst == 0 || naccum == 0) {
        xo::set_error("xo_batch_run: null or empty argument");
        return -1;
    }
    std::vector<xo::LayerL> layers;
    std::vector<xo::BatchHdr> hdr;
    std::vector<xo::BatchItem> items;
    uint64_t total = 0;
    uint32_t block = 0;
    if (xo::pack_batch("xo_batch_run", inst, ninst, lut != nullptr, nlut, opts, naccum,
                       layers, hdr, items, &total, &block))
        return -1;

    std::lock_guard<std::mutex> lock(ctx->mu);
    XO_HIP(hipSetDevice(ctx->device));
    if (ensure_table(ctx)) return -1;
    const size_t lbytes = layers.size() * sizeof(xo::LayerL);
    const size_t hbytes = hdr.size() * sizeof(xo::BatchHdr);
    const size_t ibytes = (items.empty() ? 1 : items.size()) * sizeof(xo::BatchItem);
    if (ensure_buffer(&ctx->batch_layers, &ctx->batch_layers_cap, lbytes) ||
        ensure_buffer(&ctx->batch_hdr, &ctx->batch_hdr_cap, hbytes) ||
        ensure_buffer(&ctx->batch_items, &ctx->batch_items_cap, ibytes) ||
        ensure_buffer((void **)&ctx->accum, &ctx->accum_cap, naccum * 8))
        return -1;
    if (nlut && ensure_buffer((void **)&ctx->lut, &ctx->lut_cap, (size_t)nlut * 4)) return -1;
    XO_HIP(hipMemcpyAsync(ctx->batch_layers, layers.data(), lbytes, hipMemcpyHostToDevice,
                          ctx->stream));
    XO_HIP(hipMemcpyAsync(ctx->batch_hdr, hdr.data(), hbytes, hipMemcpyHostToDevice,
                          ctx->stream));
    if (!items.empty())
        XO_HIP(hipMemcpyAsync(ctx->batch_items, items.data(),
                              items.size() * sizeof(xo::BatchItem), hipMemcpyHostToDevice,
                              ctx->stream));
    if (nlut)
        XO_HIP(hipMemcpyAsync(ctx->lut, lut, (size_t)nlut * 4, hipMemcpyHostToDevice,
                              ctx->stream));

    xo::BatchParams P;
    xo::batch_params(*opts, P);
    P.layers = (const xo::LayerL *)ctx->batch_layers;
    P.hdr = (const xo::BatchHdr *)ctx->batch_hdr;
    P.items = (const xo::BatchItem *)ctx->batch_items;
    P.lut = nlut ? ctx->lut : nullptr;
    P.table = ctx->table;
    P.item_counter = (unsigned int *)ctx->counters;
    P.accum = ctx->accum;
    P.stats = ctx->counters + 1;
    P.nitems = (uint32_t)items.size();
    P.any_fluence = xo::any_fluence(hdr.data(), hdr.size()) ? 1 : 0;
    uint32_t blocks = 0;
    if (xo::batch_grid(ctx, block, *opts, &blocks)) return -1;

    XO_HIP(hipMemsetAsync(ctx->accum, 0, naccum * 8, ctx->stream));
    XO_HIP(hipMemsetAsync(ctx->counters, 0, sizeof(unsigned long long) * (1 + XO_STAT_COUNT),
                          ctx->stream));
    XO_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (total > 0) XO_HIP(xo::launch_batch(P, blocks, block, ctx->stream));
    XO_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    unsigned long long stats[1 + XO_STAT_COUNT];
    XO_HIP(hipMemcpyAsync(accum, ctx->accum, naccum * 8, hipMemcpyDeviceToHost, ctx->stream));
    XO_HIP(hipMemcpyAsync(stats, ctx->counters, sizeof(stats), hipMemcpyDeviceToHost,
                          ctx->stream));
    XO_HIP(hipStreamSynchronize(ctx->stream));
    float kms = 0.f;
    XO_HIP(hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
    fill_report(report, stats + 1, blocks * block, kms, ms_since(t0));
    return 0;
}

} // extern "C"
