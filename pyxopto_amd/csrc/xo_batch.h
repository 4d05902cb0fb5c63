/*
 * xo_batch.h — device-side form of a batch (built on the host from the
 * xo_batch_instance array) and the launch parameters of the batch kernel.
 */
#ifndef XO_BATCH_H
#define XO_BATCH_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>

#include "../../include/xomc.h"
#include "xo_device.h"
#include "xo_kernels.h"

namespace xo {

/* Per-instance constants the lanes read at packet launch and escape. */
struct BatchHdr {
    int32_t num_layers;
    int32_t pf_kind;
    float rmax2;
    float reflectance;
    float src_pos[3];           /* Line: surface point; fiber: core centre */
    float dir_sample[3];
    float dir_reflected[3];
    int32_t src_kind;           /* XO_SRC_LINE, _FIBER_NI or _FIBER */
    uint64_t seed;
    xo_detector top, bottom, specular;
    xo_source_fiber fiber;
    xo_fiber_axis fiber_axis;
    xo_fluence fluence;         /* enabled 0: none */
};

/* Host: instances of one class run the same branches of the transport
 * step (phase-function kind, source kind, number of layers). */
inline int item_class(const xo_batch_instance &I) {
    return (I.pf_kind * 4 + I.source_kind) * 16 + (int)I.num_layers;
}

/* Host: the per-instance header of a batch instance; accum_shift is added
 * to every detector offset. */
inline BatchHdr batch_header(const xo_batch_instance &I, const xo_detector det[3],
                             uint32_t accum_shift) {
    BatchHdr H;
    std::memset(&H, 0, sizeof(H));
    H.num_layers = (int32_t)I.num_layers;
    H.pf_kind = I.pf_kind;
    H.rmax2 = I.rmax * I.rmax;
    H.src_kind = I.source_kind;
    H.fiber = I.fiber;
    H.fiber_axis = I.fiber_axis;
    const bool fiber = I.source_kind == XO_SRC_FIBER_NI || I.source_kind == XO_SRC_FIBER;
    H.reflectance = fiber ? 0.0f : I.source.reflectance;
    for (int k = 0; k < 3; ++k) {
        H.src_pos[k] = fiber ? I.fiber.position[k] : I.source.position[k];
        H.dir_sample[k] = I.source.direction_sample[k];
        H.dir_reflected[k] = I.source.direction_reflected[k];
    }
    H.seed = I.seed;
    H.top = det[0];
    H.bottom = det[1];
    H.specular = det[2];
    H.top.offset += accum_shift;
    H.bottom.offset += accum_shift;
    H.specular.offset += accum_shift;
    H.fluence = I.fluence;
    if (H.fluence.enabled != XO_FLUENCE_NONE) H.fluence.offset += accum_shift;
    return H;
}

/* Host: whether any instance of a batch has a fluence grid. */
inline bool any_fluence(const BatchHdr *hdr, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (hdr[i].fluence.enabled != XO_FLUENCE_NONE) return true;
    return false;
}

/* One work-queue item: packets [first, first + count) of one instance. */
struct BatchItem {
    uint32_t inst;
    uint32_t count;
    uint64_t first;
};

/* Default of xo_batch_opts.defer_min. */
constexpr uint32_t kDeferMinDefault = 16;

struct BatchParams {
    const LayerL *layers;        /* [instance][XO_BATCH_MAX_LAYERS] */
    const BatchHdr *hdr;
    const BatchItem *items;
    const float *lut;
    const uint32_t *table;       /* seed-table multipliers, rows 1.. */
    unsigned int *item_counter;
    unsigned long long *accum;
    unsigned long long *stats;
    const uint32_t *nitems_ptr;  /* if set, the item count is read here    */
    uint32_t nitems;
    float w_min;
    float chance;
    int32_t use_lottery;
    uint32_t refill_min;         /* idle lanes that trigger a refill (>= 1) */
    uint32_t defer_min;          /* lanes holding a boundary event that make
                                    the wave handle them (>= 1)             */
    int32_t method;              /* enum xo_method of every instance        */
    int32_t fast_math;           /* 1: MathFast arithmetic, 0: MathExact    */
    int32_t any_fluence;         /* 1 if some instance has a fluence grid   */
};

hipError_t launch_batch(const BatchParams &P, uint32_t blocks, uint32_t block,
                        hipStream_t stream);
hipError_t occupancy_batch(uint32_t block, int *blocks_per_cu);
/* Zero n 32-bit words (a kernel, so it can be recorded in a graph). */
hipError_t launch_zero_u32(uint32_t *ptr, uint32_t n, hipStream_t stream);

/* Host conversion of one layer into the device layout. */
inline LayerL layer_to_device(const xo_layer &L) {
    LayerL o;
    o.top = L.top;
    o.bottom = L.bottom;
    o.inv_mut = L.inv_mut;
    o.mua_inv_mut = L.mua_inv_mut;
    o.n = L.n;
    o.cc_top = L.cos_critical_top;
    o.cc_bottom = L.cos_critical_bottom;
    o.g = L.pf.g;
    o.beta = L.pf.beta;
    o.la = L.pf.a;
    o.lb = L.pf.b;
    o.lc = L.pf.c;
    o.loff = L.pf.offset;
    o.lsize_m1 = L.pf.size > 0u ? L.pf.size - 1u : 0u;
    o.mua = L.mua;
    o.mus = L.mus;
    return o;
}

} // namespace xo

#endif /* XO_BATCH_H */
