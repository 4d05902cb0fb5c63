/*
 * xo_batch.hip — one persistent launch for many layered-media instances.
 *
 * Work queue: items (instance, first packet, packet count), claimed by whole
 * wavefronts with one atomic each. Lanes are independent: a lane whose packet
 * ends takes the next packet of its wave's current item; when the item has
 * no unclaimed packet left the wave claims the next item at once, so lanes
 * of one wave may be on different instances and no lane waits for a slow
 * packet of another. Each lane carries its instance index; the instance's
 * layers (64 B each, a few instances per wave at a time) are read through
 * the L1/L2 caches (the per-step fields of the current layer stay in
 * registers between boundary events), detector deposits are u64 atomics
 * to HBM.
 *
 * Each packet's MWC generator is keyed on (instance seed, packet index)
 * (xo_seedhash.h), so the integer accumulators of an instance are a pure
 * function of the instance: identical alone, in any mix, at any launch shape.
 * The transport step is the layered kernel's (mcml.template.c:514-788).
 */
#include <hip/hip_runtime.h>

#include "xo_batch.h"
#include "xo_device.h"
#include "xo_seedhash.h"

namespace xo {

struct BatchBlockStats {
    unsigned long long v[XO_STAT_COUNT];
};

/* Phase-function sampling by the instance's kind. */
template <class M>
XO_DEV float pf_sample_rt(int kind, const PfL &pf, const float *lut, uint64_t &x, uint32_t a,
                          float &xi_fi) {
    if (kind == XO_PF_HG) return pf_sample<XO_PF_HG, M>(pf, lut, x, a, xi_fi);
    if (kind == XO_PF_MHG) return pf_sample<XO_PF_MHG, M>(pf, lut, x, a, xi_fi);
    return pf_sample<XO_PF_LUT, M>(pf, lut, x, a, xi_fi);
}

/* The per-step fields of a packet's current layer, kept in registers: top,
 * bottom, 1/mut and mua/mut (the first 16 bytes of the record) and the
 * phase-function parameters. */
XO_DEV void layer_state(const LayerL &Q, float4 &lt, PfL &pf) {
    lt = *reinterpret_cast<const float4 *>(&Q);
    pf = PfL{Q.g, Q.beta, Q.la, Q.lb, Q.lc, Q.loff, Q.lsize_m1};
}

/* Pending events of the split transport step. */
constexpr uint32_t kEvNone = 0, kEvInteraction = 1, kEvBoundary = 2;
/* Boundary events are handled once P.defer_min lanes of the wave hold one,
 * or after kDeferWait iterations (bounded wait, every event is handled). */
constexpr uint32_t kDeferWait = 8;
/* Transport rounds (B then C) per loop iteration: 1.66e9 -> 1.92e9
 * packets/s on the 4096-instance batch going from 1 to 6 (2: 1.83e9, 3:
 * 1.89e9, 8: 1.93e9; ending the rounds early once a quarter of the lanes
 * wait: 1.86e9). */
constexpr int kStepsPerIteration = 6;

/* End of an event (mcml.template.c:756-782): the simulation radius and the
 * step limit; a finished packet frees its lane and adds its steps to the
 * lane's count (every step ends in exactly one event). */
__device__ __forceinline__ void end_event(bool done, bool &alive, float px, float py, float pz,
                                          float spx, float spy, float spz, float rmax2,
                                          uint32_t &pkt_steps, uint32_t *truncated,
                                          uint32_t *steps) {
    /* no distance can exceed an infinite radius (the default): skip the
     * test, which changes nothing */
    if (rmax2 < XO_INF_F) {
        const float ex = px - spx, ey = py - spy, ez = pz - spz;
        if (ex * ex + ey * ey + ez * ez > rmax2) done = true;
    }
    ++pkt_steps;
    if (!done && pkt_steps >= XO_MAX_PACKET_STEPS) {
        done = true;
        atomicAdd(truncated, 1u);
    }
    if (done) {
        alive = false;
        atomicAdd(steps, pkt_steps);
    }
}

/* At most 64 VGPRs (8 waves per SIMD): the compiler fits the loop without
 * spills, and the extra wave hides more of the layer and LUT reads (+2.5 %
 * packets/s over 72 VGPRs and 7 waves). METHOD is the stepping method of
 * every instance (xo_method), M the arithmetic (MathExact, MathFast), FLU
 * whether any instance has a fluence grid. */
template <int METHOD, class M, bool FLU>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
batch_kernel(const BatchParams P) {
    __shared__ BatchBlockStats sstats;
    /* per-thread event counts in LDS (packets, steps, boundary events,
     * truncated packets and, with fluence grids, deposits), added with LDS
     * atomics: kept in registers they would take four or five of the
     * loop's 64 */
    constexpr int kCounts = FLU ? 5 : 4;
    __shared__ uint32_t scount[kCounts][256];
    for (int k = 0; k < kCounts; ++k) scount[k][threadIdx.x] = 0u;
    if (threadIdx.x < XO_STAT_COUNT) sstats.v[threadIdx.x] = 0ull;
    __syncthreads();
    const uint32_t nitems = P.nitems_ptr ? *P.nitems_ptr : P.nitems;

    const Bins bins{nullptr, P.accum};
    const uint32_t lane = lane_id();

    /* wave-uniform queue state */
    uint32_t cur_inst = 0;
    uint64_t cur_next = 0, cur_end = 0;
    bool queue_empty = false;

    /* lane state */
    bool alive = false;
    uint32_t inst = 0;
    const LayerL *layers = P.layers;
    int32_t nl = 3, pf_kind = 0;
    float rmax2 = 0.f, spx = 0.f, spy = 0.f, spz = 0.f;
    float px = 0.f, py = 0.f, pz = 0.f, dx = 0.f, dy = 0.f, dz = 0.f, w = 0.f;
    int32_t layer = 1;
    uint64_t x = 0;
    uint32_t a = 0;
    uint32_t pkt_steps = 0;
    /* the per-step fields of the packet's current layer (layer_state):
     * read again only when the layer may change (a new packet, a boundary
     * event), not in every step (+14 % packets/s; the compiler spills a
     * few rarely used registers to stay at 64) */
    float4 lt = make_float4(0.f, 0.f, 0.f, 0.f);
    PfL lpf{0.f, 0.f, 0.f, 0.f, 0.f, 0u, 0u};
    uint32_t *const n_trunc = &scount[3][threadIdx.x];
    uint32_t *const n_steps = &scount[1][threadIdx.x];
    uint32_t pend = kEvNone;  /* event the lane's last step ended on     */
    float pstep = 0.f;        /* drawn step of a pending boundary event  */
    uint32_t bwait = 0;       /* iterations boundary events have waited  */

    for (;;) {
        /* ---- hand out packets to the lanes that need one, once at least
         * refill_min of them (or every lane) wait: one refill serves many
         * lanes, the rest of the time no lane runs the launch code ---- */
        const bool want = !alive && !queue_empty;
        const uint64_t m = __ballot(want);
        const uint32_t nwant = (uint32_t)__builtin_popcountll(m);
        const uint32_t nalive = (uint32_t)__builtin_popcountll(__ballot(alive));
        const uint32_t rank = prefix_count(m);
        uint32_t assigned = 0;
        bool launch = false;
        uint64_t my_p = 0;
        const bool refill = nwant >= P.refill_min || nalive == 0;
        while (refill && assigned < nwant) {
            if (cur_next >= cur_end) {
                uint32_t it = 0;
                if (lane == 0) it = atomicAdd(P.item_counter, 1u);
                it = (uint32_t)__builtin_amdgcn_readlane((int)it, 0);
                if (it >= nitems) {
                    queue_empty = true;
                    break;
                }
                const BatchItem item = P.items[it];
                cur_inst = item.inst;
                cur_next = item.first;
                cur_end = item.first + item.count;
                /* every packet of an instance deposits the same specular
                 * integer at the same bin: one atomic for the whole item */
                if (lane == 0) {
                    const BatchHdr &H = P.hdr[cur_inst];
                    if (H.src_kind == XO_SRC_LINE)
                        detector_deposit_n<M>(H.specular, bins, H.src_pos[0], H.src_pos[1],
                                           H.dir_reflected[0], H.dir_reflected[1],
                                           H.dir_reflected[2], H.reflectance, item.count);
                }
            }
            const uint64_t avail = cur_end - cur_next;
            const uint32_t left = nwant - assigned;
            const uint32_t take = avail < (uint64_t)left ? (uint32_t)avail : left;
            if (want && rank >= assigned && rank < assigned + take) {
                launch = true;
                inst = cur_inst;
                my_p = cur_next + (rank - assigned);
            }
            cur_next += take;
            assigned += take;
        }
        if (launch) {
            const BatchHdr &H = P.hdr[inst];
            layers = P.layers + (size_t)inst * XO_BATCH_MAX_LAYERS;
            nl = H.num_layers;
            pf_kind = H.pf_kind;
            rmax2 = H.rmax2;
            spx = H.src_pos[0];
            spy = H.src_pos[1];
            spz = H.src_pos[2];
            xo_packet_state(H.seed, my_p, P.table, &x, &a);
            if (H.src_kind == XO_SRC_FIBER_NI) {
                const float n1 = layers[1].n;
                const float R =
                    fiber_ni_launch<M>(H.fiber, n1, x, a, px, py, pz, dx, dy, dz, w);
                if (H.specular.kind != XO_DET_NONE) {
                    float rx, ry, rz;
                    fiber_specular_direction<M>(dx, dy, dz, n1, H.fiber.n, rx, ry, rz);
                    detector_deposit<M>(H.specular, bins, px, py, rx, ry, rz, R);
                }
            } else if (H.src_kind == XO_SRC_FIBER) {
                /* UniformFiber, xopto/mcml/mcsource/fiber.py:265-330 */
                const float R = fiber_launch<M>(H.fiber, H.fiber_axis, layers[1].n, x, a, px,
                                                py, pz, dx, dy, dz, w);
                detector_deposit<M>(H.specular, bins, px, py, dx, dy, dz, R);
            } else {
                /* Line source, xopto/mcml/mcsource/line.py:95-110 */
                w = 1.0f - H.reflectance;
                px = spx;
                py = spy;
                pz = spz;
                dx = H.dir_sample[0];
                dy = H.dir_sample[1];
                dz = H.dir_sample[2];
            }
            layer = 1;
            layer_state(layers[1], lt, lpf);
            pkt_steps = 0;
            alive = true;
            atomicAdd(&scount[0][threadIdx.x], 1u);
        }
        if (!__any(alive)) {
            if (queue_empty) break;
            continue;
        }
        if (!alive) continue;

        /* an instance's fluence deposit (fluence.py:102-147); the grid is
         * read from the instance header only when the batch has one */
        auto fluence_at = [&](float qx, float qy, float qz, float deposit, float mua) {
            const xo_fluence &F = P.hdr[inst].fluence;
            if (F.enabled != XO_FLUENCE_NONE &&
                fluence_deposit<M>(F, P.accum, qx, qy, qz, deposit, mua))
                atomicAdd(&scount[kCounts - 1][threadIdx.x], 1u);
        };
        /* microscopic Beer-Lambert: absorption along the whole step
         * (mcml.template.c:584-606); with a fluence grid the deposit goes
         * to a point drawn along the step, as in the layered kernel */
        auto mbl_absorb = [&](const LayerL &L, float step) {
            const float mua = L.mua;
            const float fraction = 1.0f - M::exp(-(mua * step));
            const float deposit = fraction * w;
            w -= deposit;
            if (FLU && P.hdr[inst].fluence.enabled != XO_FLUENCE_NONE) {
                const float back =
                    (mua != 0.0f)
                        ? step - M::div(-M::log_unit(1.0f - rng_next(x, a) * fraction), mua)
                        : 0.0f;
                fluence_at(px - back * dx, py - back * dy, pz - back * dz, deposit, mua);
            }
        };
        /* microscopic Beer-Lambert: roulette after the boundary or the
         * scattering, whether or not the packet has left
         * (mcml.template.c:652-667) */
        auto mbl_roulette = [&](bool done) {
            if (METHOD == XO_METHOD_MBL && roulette(w, x, a, P.w_min, P.chance, P.use_lottery))
                done = true;
            return done;
        };

        /* ---- transport steps, mcml.template.c:514-788, in three phases
         * per iteration. (B) lanes without a pending event draw a step; a
         * step that stays inside the layer moves the packet and leaves an
         * interaction event, a step that reaches a layer boundary is kept
         * pending. (C) every interaction event is handled. (A, at the start
         * of the next iterations) boundary events, a few percent of the
         * steps in turbid layers, are handled together once defer_min lanes
         * of the wave hold one, every alive lane holds one, or they waited
         * kDeferWait iterations: the clipping of the step to the boundary,
         * the move and the Fresnel event then run for those lanes only,
         * instead of in nearly every iteration for two or three lanes.
         * Each packet's arithmetic and draws are those of the reference's
         * loop; only when the wave issues them changes. ---- */
        {
            const uint32_t nb = (uint32_t)__builtin_popcountll(__ballot(pend == kEvBoundary));
            const uint32_t nal = (uint32_t)__builtin_popcountll(__ballot(true));
            /* the wait count of the first active lane (lanes that were idle
             * hold stale counts), so the decision is wave-uniform */
            const uint32_t bw = (uint32_t)__builtin_amdgcn_readfirstlane((int)bwait);
            const bool run_b = nb != 0 && (nb >= P.defer_min || nb == nal || bw >= kDeferWait);
            bwait = (nb != 0 && !run_b) ? bw + 1u : 0u;
            if (run_b && pend == kEvBoundary) {
                /* (A) the step as drawn, clipped to the boundary it reaches */
                const LayerL L = layers[layer];
                float step = pstep;
                int32_t next = layer;
                if (__builtin_fmaf(step, dz, pz) < L.top) {
                    --next;
                    if (fabsf(dz) != 0.0f) step = M::div(L.top - pz, dz);
                }
                if (__builtin_fmaf(step, dz, pz) >= L.bottom) {
                    ++next;
                    if (fabsf(dz) != 0.0f) step = M::div(L.bottom - pz, dz);
                }
                px = __builtin_fmaf(dx, step, px);
                py = __builtin_fmaf(dy, step, py);
                pz = __builtin_fmaf(dz, step, pz);
                if (layer < next) pz = L.bottom;
                if (layer > next) pz = L.top;
                if (METHOD == XO_METHOD_MBL) mbl_absorb(L, step);
                if (next != layer) {
                    atomicAdd(&scount[2][threadIdx.x], 1u);
                    mcml_boundary<M>(layers, layer, next, dx, dy, dz, x, a);
                    bool done = false;
                    if (!(layer > 0 && layer < nl - 1)) {
                        const BatchHdr &H = P.hdr[inst];
                        if (layer <= 0)
                            detector_deposit<M>(H.top, bins, px, py, dx, dy, dz, w);
                        else
                            detector_deposit<M>(H.bottom, bins, px, py, dx, dy, dz, w);
                        done = true;
                    }
                    done = mbl_roulette(done);
                    pend = kEvNone;
                    layer_state(layers[layer], lt, lpf);
                    end_event(done, alive, px, py, pz, spx, spy, spz, rmax2, pkt_steps, n_trunc,
                              n_steps);
                } else {
                    /* crossed both planes of a degenerate layer: an
                     * interaction at the clipped point, handled below */
                    pend = kEvInteraction;
                }
            }
        }
        /* (B) and (C) kStepsPerIteration times per iteration: the loop-top
         * work (refill and boundary ballots, the wave-uniform decisions)
         * is shared by that many steps; a lane whose packet ends or whose
         * step reaches a boundary sits out the remaining rounds */
#pragma unroll
        for (int rep = 0; rep < kStepsPerIteration; ++rep) {
            if (alive && pend == kEvNone) {
                /* (B) */
                float step;
                if (METHOD == XO_METHOD_MBL)
                    step = M::div(-M::log_unit(rng_next(x, a)),
                                  layers[layer].mus); /* mcml.template.c:529-531 */
                else
                    step = -M::log_unit(rng_next(x, a)) * lt.z;
                step = fminf(step, XO_FLT_MAX);
                const float zend = __builtin_fmaf(step, dz, pz);
                if (zend < lt.x || zend >= lt.y) {
                    pend = kEvBoundary;
                    pstep = step;
                } else {
                    px = __builtin_fmaf(dx, step, px);
                    py = __builtin_fmaf(dy, step, py);
                    pz = zend;
                    if (METHOD == XO_METHOD_MBL) mbl_absorb(layers[layer], step);
                    pend = kEvInteraction;
                }
            }
            if (alive && pend == kEvInteraction) {
                /* (C) absorption and scattering, mcml.template.c:696-753 */
                const LayerL &L = layers[layer];
                bool done = false;
                if (METHOD == XO_METHOD_AR && rng_next(x, a) < lt.w) {
                    /* albedo rejection: the whole weight is absorbed here */
                    if (FLU) fluence_at(px, py, pz, w, L.mua);
                    w = 0.0f;
                    done = true;
                } else {
                    if (METHOD == XO_METHOD_AW) {
                        float deposit = w * lt.w;
                        w -= deposit;
                        if (FLU) fluence_at(px, py, pz, deposit, L.mua);
                    }
                    float xi_fi;
                    float ct = pf_sample_rt<M>(pf_kind, lpf, P.lut, x, a, xi_fi);
                    scatter_direction<M>(dx, dy, dz, ct, xi_fi);
                    if (METHOD == XO_METHOD_AW)
                        done = roulette(w, x, a, P.w_min, P.chance, P.use_lottery);
                }
                done = mbl_roulette(done);
                pend = kEvNone;
                end_event(done, alive, px, py, pz, spx, spy, spz, rmax2, pkt_steps, n_trunc,
                          n_steps);
            }
        }
    }

    const int kinds[5] = {XO_STAT_PACKETS, XO_STAT_STEPS, XO_STAT_BOUNDARY, XO_STAT_TRUNCATED,
                          XO_STAT_DEPOSITS};
    for (int k = 0; k < kCounts; ++k)
        if (scount[k][threadIdx.x])
            atomicAdd(&sstats.v[kinds[k]], (unsigned long long)scount[k][threadIdx.x]);
    __syncthreads();
    if (threadIdx.x < XO_STAT_COUNT && sstats.v[threadIdx.x])
        atomicAdd(&P.stats[threadIdx.x], sstats.v[threadIdx.x]);
}

template <class M, bool FLU>
static hipError_t launch_batch_m(const BatchParams &P, uint32_t blocks, uint32_t block,
                                 hipStream_t stream) {
    switch (P.method) {
    case XO_METHOD_AW:
        hipLaunchKernelGGL((batch_kernel<XO_METHOD_AW, M, FLU>), dim3(blocks), dim3(block), 0,
                           stream, P);
        break;
    case XO_METHOD_AR:
        hipLaunchKernelGGL((batch_kernel<XO_METHOD_AR, M, FLU>), dim3(blocks), dim3(block), 0,
                           stream, P);
        break;
    case XO_METHOD_MBL:
        hipLaunchKernelGGL((batch_kernel<XO_METHOD_MBL, M, FLU>), dim3(blocks), dim3(block), 0,
                           stream, P);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

/* The kernel is specialized on whether any instance has a fluence grid, so
 * batches without one run the transport loop with no trace of it. */
hipError_t launch_batch(const BatchParams &P, uint32_t blocks, uint32_t block,
                        hipStream_t stream) {
    if (P.any_fluence)
        return P.fast_math ? launch_batch_m<MathFast, true>(P, blocks, block, stream)
                           : launch_batch_m<MathExact, true>(P, blocks, block, stream);
    return P.fast_math ? launch_batch_m<MathFast, false>(P, blocks, block, stream)
                       : launch_batch_m<MathExact, false>(P, blocks, block, stream);
}

__global__ void zero_u32_kernel(uint32_t *ptr, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ptr[i] = 0u;
}

hipError_t launch_zero_u32(uint32_t *ptr, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, ptr, n);
    return hipGetLastError();
}

hipError_t occupancy_batch(uint32_t block, int *blocks_per_cu) {
    /* the albedo-weight build stands for all three methods */
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        blocks_per_cu, batch_kernel<XO_METHOD_AW, MathExact, false>, (int)block, 0);
}

} // namespace xo
