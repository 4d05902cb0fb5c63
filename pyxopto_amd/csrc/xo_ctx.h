/*
 * xo_ctx.h — the context object behind the opaque xo_ctx handle and the
 * host helpers shared by the translation units of libxomc.so.
 */
#ifndef XO_CTX_H
#define XO_CTX_H

#include <hip/hip_runtime.h>

#include <chrono>
#include <mutex>
#include <vector>

#include "../../include/xomc.h"
#include "xo_batch.h"
#include "xo_internal.h"

/* Return -1 from the enclosing function, with the message set, when a HIP
 * call fails. */
#define XO_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            xo::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),      \
                          __FILE__, __LINE__);                                        \
            return -1;                                                                \
        }                                                                             \
    } while (0)

struct xo_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t max_lanes = 0;
    uint32_t cus = 0;
    uint64_t *rng_x = nullptr;
    uint32_t *rng_a = nullptr;
    unsigned long long *counters = nullptr; /* [0] packet counter, [1..] stats */
    unsigned long long *accum = nullptr;
    size_t accum_cap = 0;
    float *lut = nullptr;
    size_t lut_cap = 0;
    uint8_t *voxels = nullptr;
    size_t voxels_cap = 0;
    int32_t vox_shape[3] = {0, 0, 0};

    bool vox_valid = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    /* batch buffers */
    uint32_t *table = nullptr;          /* seed multipliers, rows 1..      */
    void *batch_layers = nullptr;
    size_t batch_layers_cap = 0;
    void *batch_hdr = nullptr;
    size_t batch_hdr_cap = 0;
    void *batch_items = nullptr;
    size_t batch_items_cap = 0;
    /* trace and sampling-volume buffers */
    float *trace_data = nullptr;
    size_t trace_data_cap = 0;
    int32_t *trace_n = nullptr;
    size_t trace_n_cap = 0;
    std::mutex mu;
};

namespace xo {

int ensure_buffer(void **ptr, size_t *cap, size_t bytes);
int ensure_table(xo_ctx *ctx);
int check_detector(const char *name, const xo_detector &d, uint64_t naccum, uint64_t *end);
int check_pf(int pf_kind, const xo_pf &pf, uint32_t nlut, const char *what, uint32_t idx);
int check_fluence(const char *fn, const xo_fluence &f, uint64_t det_end, uint64_t naccum);
/* Lanes of a single-simulation launch: `requested` if not 0 (capped by the
 * context's streams), else the persistent grid, fewer for a run of fewer
 * packets than that. */
uint32_t choose_lanes(xo_ctx *ctx, uint32_t requested, int blocks_per_cu, uint64_t npackets);
double ms_since(std::chrono::steady_clock::time_point t0);
void fill_report(xo_run_report *r, const unsigned long long *stats, uint32_t lanes,
                 float kernel_ms, double total_ms);

/* Validate a batch (fn names the entry point in messages) and build its
 * device tables: layer records [instance][XO_BATCH_MAX_LAYERS], per-instance
 * headers and the work items (chunks of opts->chunk packets, grouped by
 * instance class). *total: packets of all instances; *block: threads per
 * block. Returns -1 (message set) on error. */
int pack_batch(const char *fn, const xo_batch_instance *inst, uint32_t ninst, bool have_lut,
               uint32_t nlut, const xo_batch_opts *opts, uint64_t naccum,
               std::vector<LayerL> &layers, std::vector<BatchHdr> &hdr,
               std::vector<BatchItem> &items, uint64_t *total, uint32_t *block);
/* The launch options of a batch as kernel parameters (buffers left null). */
void batch_params(const xo_batch_opts &opts, BatchParams &P);
/* Blocks of the persistent batch grid for this block size and options. */
int batch_grid(xo_ctx *ctx, uint32_t block, const xo_batch_opts &opts, uint32_t *blocks);

/* Device layout of a voxel grid: uint8 material indices in 4x4x4 bricks of
 * 64 bytes, bricks in [z][y][x] order, voxel (x, y, z) at byte
 * 64 brick(x>>2, y>>2, z>>2) + 16 (z&3) + 4 (y&3) + (x&3). shape is
 * {nx, ny, nz}; the host array is int32[nz][ny][nx]. */
uint64_t brick_bytes(const int32_t shape[3]);
/* Returns -1 (message set) if a voxel's material is outside
 * [0, num_materials). */
int pack_voxels(const int32_t *voxels, const int32_t shape[3], uint32_t num_materials,
                std::vector<uint8_t> &out, const char *fn);

} // namespace xo

#endif /* XO_CTX_H */
