/*
 * xo_kernels.hip — persistent photon-packet transport kernels for gfx950.
 *
 * One lane = one packet in flight, its state (position, direction, weight,
 * layer/voxel, MWC state) in VGPRs. Lanes take a new packet when theirs ends:
 * with the dynamic schedule, finished lanes of a wave claim consecutive
 * packet indices with one atomic per wave (ballot + mbcnt prefix) instead of
 * one atomic per packet (the reference's pkt_cnt_atomic_inc,
 * mcml.template.c:460, 790-812); with the static schedule lane g runs packets
 * g, g+L, ... . The layer/material table is staged once per block in LDS;
 * detector bins are privatized in LDS when they fit and flushed with one u64
 * atomic per bin and block; fluence deposits are u64 atomics to HBM.
 * Everything is binary32 VALU work; nothing here is a contraction, so no MFMA.
 */
#include <hip/hip_runtime.h>

#include "xo_device.h"
#include "xo_kernels.h"
#include "xo_seedhash.h"

namespace xo {

/* ------------------------------------------------------------------ */
/* mcml                                                               */
/* ------------------------------------------------------------------ */

/* Block-level prologue/epilogue shared by the kernels. */
struct BlockStats {
    unsigned long long v[XO_STAT_COUNT];
};

XO_DEV void stats_flush(BlockStats *ss, const uint32_t *mine,
                        unsigned long long *gstats) {
    for (int i = 0; i < XO_STAT_COUNT; ++i)
        if (mine[i]) atomicAdd(&ss->v[i], (unsigned long long)mine[i]);
    __syncthreads();
    if (threadIdx.x < XO_STAT_COUNT && ss->v[threadIdx.x])
        atomicAdd(&gstats[threadIdx.x], ss->v[threadIdx.x]);
}

/* Pending events of the split transport step of the layered kernel;
 * boundary events wait until kDeferMin lanes of a wave hold one (at most
 * kDeferWait iterations). */
constexpr uint32_t kEvNone = 0, kEvInteraction = 1, kEvBoundary = 2;
constexpr uint32_t kDeferMin = 12, kDeferWait = 8;
/* Transport rounds (B then C) per loop iteration of the layered kernel. */
constexpr int kRoundsPerIteration = 6;

template <int PF, int METHOD, bool TRACE, class M>
__global__ void __launch_bounds__(XO_BLOCK) mcml_kernel(const McmlParams P) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ BlockStats sstats;
    __shared__ LayerEventCfgL sev;
    LayerL *sl = reinterpret_cast<LayerL *>(smem);
    unsigned long long *sbins =
        reinterpret_cast<unsigned long long *>(smem + XO_MAX_LAYERS * sizeof(LayerL));
    float *slut = reinterpret_cast<float *>(sbins + P.lds_bins);

    const uint32_t tid = threadIdx.x;
    const uint32_t nl = P.cfg.num_layers;
    if (tid < nl) {
        const xo_layer &L = P.cfg.layers[tid];
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
        sl[tid] = o;
    }
    for (uint32_t i = tid; i < P.lds_bins; i += XO_BLOCK) sbins[i] = 0ull;
    for (uint32_t i = tid; i < P.lds_lut; i += XO_BLOCK) slut[i] = P.lut[i];
    if (tid < XO_STAT_COUNT) sstats.v[tid] = 0ull;
    if (tid == 0) {
        sev.top = P.cfg.top;
        sev.bottom = P.cfg.bottom;
        sev.specular = P.cfg.specular;
        sev.fluence = P.cfg.fluence;
        sev.line = P.cfg.source;
        sev.fiber = P.cfg.fiber;
        sev.fiber_axis = P.cfg.fiber_axis;
    }
    __syncthreads();

    const Bins bins{P.lds_bins ? sbins : nullptr, P.accum};
    const float *lut = P.lds_lut ? slut : P.lut;

    const uint32_t gid = blockIdx.x * XO_BLOCK + tid;
    const bool lane_ok = gid < P.lanes;
    const bool static_sched = (P.flags & XO_FLAG_STATIC_SCHEDULE) != 0u;
    const bool keyed = (P.flags & XO_FLAG_PACKET_KEYED) != 0u;
    uint64_t x = 0;
    uint32_t a = 0;
    uint64_t my_n = 0, my_k = 0;
    if (lane_ok) {
        if (!keyed) {
            x = P.rng_x[gid];
            a = P.rng_a[gid];
        }
        if (static_sched)
            my_n = P.npackets / P.lanes + ((uint64_t)gid < P.npackets % P.lanes ? 1u : 0u);
    }

    const float rmax2 = P.cfg.rmax * P.cfg.rmax;
    const int32_t skind = P.cfg.source_kind;
    const bool line_src = skind == XO_SRC_LINE;
    const float *sp = line_src ? P.cfg.source.position : P.cfg.fiber.position;
    const float spx = sp[0];
    const float spy = sp[1];
    const float spz = sp[2];
    const float w_min = P.cfg.w_min, chance = P.cfg.lottery_chance;
    const int use_lottery = P.cfg.use_lottery;
    const bool use_fluence = P.cfg.fluence.enabled != XO_FLUENCE_NONE;
    const bool use_specular = P.cfg.specular.kind != XO_DET_NONE;

    bool active = lane_ok;
    WaveDispenser wq;
    bool alive = false;
    float px = 0.f, py = 0.f, pz = 0.f, dx = 0.f, dy = 0.f, dz = 0.f, w = 0.f;
    int32_t layer = 1;
    uint32_t st[XO_STAT_COUNT] = {0, 0, 0, 0, 0, 0, 0};
    uint32_t pkt_steps = 0;
    TraceState ts{0, 0, 0, 0.0f};
    const xo_trace &T = P.trace;
    uint32_t pend = kEvNone; /* event the lane's last step ended on      */
    float pstep = 0.f;       /* drawn step of a pending boundary event   */
    uint32_t bwait = 0;      /* iterations boundary events have waited   */
    /* the per-step fields of the packet's current layer (top, bottom,
     * 1/mut, mua/mut) and the phase-function parameters its kind uses, in
     * registers: read from LDS again only when the layer may change (a new
     * packet, a boundary event), not in every step (+5 % on the 1 mm slab) */
    float4 lt = make_float4(0.f, 0.f, 0.f, 0.f);
    PfL lpf{0.f, 0.f, 0.f, 0.f, 0.f, 0u, 0u};
    auto layer_state = [&](const LayerL &Q) {
        lt = make_float4(Q.top, Q.bottom, Q.inv_mut, Q.mua_inv_mut);
        if (PF == XO_PF_HG) lpf.g = Q.g;
        if (PF == XO_PF_MHG) {
            lpf.g = Q.g;
            lpf.beta = Q.beta;
        }
        if (PF == XO_PF_LUT) lpf = PfL{0.f, 0.f, Q.la, Q.lb, Q.lc, Q.loff, Q.lsize_m1};
    };

    for (;;) {
        /* ---- refill lanes whose packet ended ---- */
        bool want = active && !alive;
        bool launch = false;
        uint64_t pkt_index = 0;
        if (static_sched) {
            if (want) {
                if (my_k < my_n) {
                    pkt_index = (uint64_t)gid + my_k * (uint64_t)P.lanes;
                    ++my_k;
                    launch = true;
                } else {
                    active = false;
                }
            }
        } else {
            bool retire = false;
            uint64_t claimed = 0;
            launch = dispense(wq, want, alive, P.counter, P.npackets, claim_size(P), kRefillMin,
                              pkt_index, retire, claimed);
            if (retire) active = false;
            /* every Line packet deposits the same specular integer at the
             * same bin: one atomic per claimed block */
            if (claimed && line_src &&
                lane_id() == (uint32_t)__builtin_ctzll(__ballot(true))) {
                const xo_source_line &src = lds_fresh(sev.line);
                detector_deposit_n<M>(lds_fresh(sev.specular), bins, src.position[0],
                                      src.position[1],
                                   src.direction_reflected[0], src.direction_reflected[1],
                                   src.direction_reflected[2], src.reflectance,
                                   (uint32_t)claimed);
            }
        }
        if (launch) {
            if (keyed)
                xo_packet_state(P.cfg.packet_seed, P.cfg.first_packet + pkt_index, P.table, &x,
                                &a);
            if (skind == XO_SRC_FIBER_NI) {
                const float n1 = sl[1].n;
                const xo_source_fiber &fib = lds_fresh(sev.fiber);
                const float R = fiber_ni_launch<M>(fib, n1, x, a, px, py, pz, dx, dy,
                                                   dz, w);
                if (use_specular) {
                    float rx, ry, rz;
                    fiber_specular_direction<M>(dx, dy, dz, n1, fib.n, rx, ry, rz);
                    detector_deposit<M>(lds_fresh(sev.specular), bins, px, py, rx, ry, rz, R);
                }
            } else if (skind == XO_SRC_FIBER) {
                const float R = fiber_launch<M>(lds_fresh(sev.fiber), lds_fresh(sev.fiber_axis),
                                                sl[1].n, x, a, px, py, pz, dx, dy, dz, w);
                detector_deposit<M>(lds_fresh(sev.specular), bins, px, py, dx, dy, dz, R);
            } else {
                /* Line source, xopto/mcml/mcsource/line.py:95-110 */
                const xo_source_line &src = lds_fresh(sev.line);
                w = 1.0f - src.reflectance;
                px = src.position[0];
                py = src.position[1];
                pz = src.position[2];
                dx = src.direction_sample[0];
                dy = src.direction_sample[1];
                dz = src.direction_sample[2];
                if (static_sched)
                    detector_deposit<M>(lds_fresh(sev.specular), bins, px, py, src.direction_reflected[0],
                                     src.direction_reflected[1], src.direction_reflected[2],
                                     src.reflectance);
            }
            layer = 1;
            layer_state(sl[1]);
            alive = true;
            pkt_steps = 0;
            st[XO_STAT_PACKETS]++;
            if (TRACE) {
                /* mcml.template.c:513-520, 791-810 */
                ts = TraceState{pkt_index, 0u, XO_EVENT_LAUNCH, 0.0f};
                if (T.options & XO_TRACE_START)
                    trace_event(T, P.trace_data, ts, px, py, pz, dx, dy, dz, w);
            }
        }
        if (!__any(alive)) break;
        if (!alive) continue;
        /* ---- transport steps, the while(!done) loop of
         * mcml.template.c:514-788, in three phases per iteration: (B)
         * lanes without a pending event draw a step; a step inside the layer
         * moves the packet (and, for microscopic Beer-Lambert, absorbs along
         * it) and leaves an interaction event, a step reaching a boundary is
         * kept pending; (C) interaction events are handled; (A, at the top
         * of later iterations) boundary events are handled together once
         * kDeferMin lanes of the wave hold one, every live lane holds one,
         * or they waited kDeferWait iterations. Each packet's arithmetic and
         * draws follow the reference's loop; only when the wave issues them
         * changes. ---- */
        auto mbl_absorb = [&](const LayerL &L, float step) {
            /* absorption along the whole step, mcml.template.c:584-606 */
            const float mua = L.mua;
            const float fraction = 1.0f - M::exp(-(mua * step));
            const float deposit = fraction * w;
            w -= deposit;
            if (TRACE) ts.ev |= XO_EVENT_ABSORPTION;
            if (use_fluence) {
                const float back =
                    (mua != 0.0f)
                        ? step - M::div(-M::log_unit(1.0f - rng_next(x, a) * fraction), mua)
                        : 0.0f;
                if (fluence_deposit<M>(lds_fresh(sev.fluence), P.accum, px - back * dx,
                                       py - back * dy, pz - back * dz, deposit, mua))
                    st[XO_STAT_DEPOSITS]++;
            }
        };
        auto end_step = [&](bool done) {
            /* microscopic Beer-Lambert: roulette after the boundary or the
             * scattering, whether or not the packet has left (as the
             * reference does, mcml.template.c:652-667) */
            if (METHOD == XO_METHOD_MBL && roulette(w, x, a, w_min, chance, use_lottery))
                done = true;
            float ex = px - spx, ey = py - spy, ez = pz - spz;
            if (rmax2 < XO_INF_F && ex * ex + ey * ey + ez * ez > rmax2) {
                done = true;
                if (TRACE) ts.ev |= XO_EVENT_ESCAPED;
            }
            if (!done && ++pkt_steps >= XO_MAX_PACKET_STEPS) {
                done = true;
                st[XO_STAT_TRUNCATED]++;
            }
            if (TRACE) {
                /* mcml.template.c:765-781 */
                if (done) ts.ev |= XO_EVENT_TERMINATED;
                if (T.options == XO_TRACE_ALL || ((T.options & XO_TRACE_END) && done))
                    trace_event(T, P.trace_data, ts, px, py, pz, dx, dy, dz, w);
                ts.ev = 0u;
                if (done) P.trace_n[ts.pkt] = (int32_t)ts.count;
            }
            pend = kEvNone;
            if (done) alive = false;
        };
        {
            const uint32_t nb = (uint32_t)__builtin_popcountll(__ballot(pend == kEvBoundary));
            const uint32_t nal = (uint32_t)__builtin_popcountll(__ballot(true));
            const uint32_t bw = (uint32_t)__builtin_amdgcn_readfirstlane((int)bwait);
            const bool run_b = nb != 0 && (nb >= kDeferMin || nb == nal || bw >= kDeferWait);
            bwait = (nb != 0 && !run_b) ? bw + 1u : 0u;
            if (run_b && pend == kEvBoundary) {
                /* (A) the step as drawn, clipped to the boundary it reaches */
                const LayerL &L = sl[layer];
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
                /* optical path length, mcml.template.c:570-574 */
                if (TRACE && T.plon) ts.pl += L.n * step;
                if (layer < next) pz = L.bottom;
                if (layer > next) pz = L.top;
                if (METHOD == XO_METHOD_MBL) mbl_absorb(L, step);
                if (next != layer) {
                    st[XO_STAT_BOUNDARY]++;
                    const int32_t before = layer;
                    mcml_boundary<M>(sl, layer, next, dx, dy, dz, x, a);
                    layer_state(sl[layer]);
                    if (TRACE)
                        ts.ev |= XO_EVENT_BOUNDARY_HIT |
                                 (layer != before ? XO_EVENT_REFRACTION : XO_EVENT_REFLECTION);
                    bool done = false;
                    if (!(layer > 0 && layer < (int32_t)nl - 1)) {
                        if (layer <= 0)
                            detector_deposit<M>(lds_fresh(sev.top), bins, px, py, dx, dy, dz, w);
                        else
                            detector_deposit<M>(lds_fresh(sev.bottom), bins, px, py, dx, dy, dz,
                                                w);
                        done = true;
                    }
                    end_step(done);
                } else {
                    /* crossed both planes of a degenerate layer: an
                     * interaction at the clipped point, handled below */
                    pend = kEvInteraction;
                }
            }
        }
        /* (B) and (C) kRoundsPerIteration times per iteration, as in the
         * batch kernel: the loop-top work is shared by that many steps
         * (1 mm slab 0.464 -> 0.403 ms per 1e6 packets) */
#pragma unroll
        for (int rep = 0; rep < kRoundsPerIteration; ++rep) {
            if (alive && pend == kEvNone) {
                /* (B) */
                const LayerL &L = sl[layer];
                st[XO_STAT_STEPS]++;
                float step;
                if (METHOD == XO_METHOD_MBL)
                    step = M::div(-M::log_unit(rng_next(x, a)), L.mus); /* mcml.template.c:529-531 */
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
                    if (TRACE && T.plon) ts.pl += L.n * step;
                    if (METHOD == XO_METHOD_MBL) mbl_absorb(L, step);
                    pend = kEvInteraction;
                }
            }
            if (alive && pend == kEvInteraction) {
                /* (C) */
                const LayerL &L = sl[layer];
                bool done = false;
                if (METHOD == XO_METHOD_AR && rng_next(x, a) < lt.w) {
                    /* albedo rejection: the whole weight is absorbed here,
                     * mcml.template.c:696-708 */
                    const float deposit = w;
                    w -= deposit;
                    if (TRACE) ts.ev |= XO_EVENT_ABSORPTION;
                    if (use_fluence &&
                        fluence_deposit<M>(lds_fresh(sev.fluence), P.accum, px, py, pz, deposit,
                                           L.mua))
                        st[XO_STAT_DEPOSITS]++;
                    done = true;
                } else {
                    if (METHOD == XO_METHOD_AW) {
                        const float deposit = w * lt.w;
                        w -= deposit;
                        if (TRACE) ts.ev |= XO_EVENT_ABSORPTION;
                        if (use_fluence &&
                            fluence_deposit<M>(lds_fresh(sev.fluence), P.accum, px, py, pz, deposit,
                                               L.mua))
                            st[XO_STAT_DEPOSITS]++;
                    }
                    float xi_fi;
                    float ct = pf_sample<PF, M>(lpf, lut, x, a, xi_fi);
                    scatter_direction<M>(dx, dy, dz, ct, xi_fi);
                    if (TRACE) ts.ev |= XO_EVENT_SCATTERING;
                    if (METHOD == XO_METHOD_AW && roulette(w, x, a, w_min, chance, use_lottery))
                        done = true;
                }
                end_step(done);
            }
        }
    }

    if (lane_ok) {
        if (!keyed) P.rng_x[gid] = x;
        if (st[XO_STAT_PACKETS]) st[XO_STAT_THREADS] = 1;
    }
    stats_flush(&sstats, st, P.stats);
    if (P.lds_bins) {
        for (uint32_t i = tid; i < P.lds_bins; i += XO_BLOCK)
            if (sbins[i]) atomicAdd(&P.accum[i], sbins[i]);
    }
}

/* ------------------------------------------------------------------ */
/* mcvox                                                              */
/* ------------------------------------------------------------------ */

struct VoxGeom {
    float tl[3], br[3], size[3];
    int32_t shape[3];
};

XO_DEV bool vox_valid(const VoxGeom &g, int32_t ix, int32_t iy, int32_t iz) {
    return ix >= 0 && ix < g.shape[0] && iy >= 0 && iy < g.shape[1] && iz >= 0 &&
           iz < g.shape[2];
}

/* The grid is stored in 4x4x4 bricks of 64 bytes (xo_ctx.h,
 * xo::pack_voxels): a packet crossing into a neighbour voxel along any axis
 * stays in the same brick three times in four, so most gathers hit a cache
 * line the lane (or its wave) has just read. */
XO_DEV uint32_t vox_material(const uint8_t *vox, const VoxGeom &g, int32_t ix,
                             int32_t iy, int32_t iz) {
    const uint32_t nbx = ((uint32_t)g.shape[0] + 3u) >> 2;
    const uint32_t nby = ((uint32_t)g.shape[1] + 3u) >> 2;
    const size_t brick = ((size_t)((uint32_t)iz >> 2) * nby + ((uint32_t)iy >> 2)) * nbx +
                         ((uint32_t)ix >> 2);
    const uint32_t off = (((uint32_t)iz & 3u) << 4) | (((uint32_t)iy & 3u) << 2) |
                         ((uint32_t)ix & 3u);
    return vox[brick * 64u + off];
}

/* mcsim_sample_box_intersect, mcvox.template.c:94-162 */
XO_DEV bool box_intersect(const VoxGeom &g, const float pos[3], const float dir[3],
                          float inter[3], float nrm[3]) {
    float inv[3];
    for (int k = 0; k < 3; ++k) inv[k] = (dir[k] != 0.0f) ? 1.0f / dir[k] : XO_INF_F;
    float t1 = (g.tl[0] - pos[0]) * inv[0];
    float t2 = (g.br[0] - pos[0]) * inv[0];
    float t3 = (g.tl[1] - pos[1]) * inv[1];
    float t4 = (g.br[1] - pos[1]) * inv[1];
    float t5 = (g.tl[2] - pos[2]) * inv[2];
    float t6 = (g.br[2] - pos[2]) * inv[2];
    float txmin = fminf(t1, t2), txmax = fmaxf(t1, t2);
    float tymin = fminf(t3, t4), tymax = fmaxf(t3, t4);
    float tzmin = fminf(t5, t6), tzmax = fmaxf(t5, t6);
    float tmin = fmaxf(fmaxf(txmin, tymin), tzmin);
    float tmax = fminf(fminf(txmax, tymax), tzmax);
    if (tmin > tmax) return false;
    nrm[0] = (float)(txmin >= tymin && txmin >= tzmin);
    nrm[1] = (float)((nrm[0] == 0.0f) && tymin >= tzmin);
    nrm[2] = (float)(nrm[0] == 0.0f && nrm[1] == 0.0f);
    nrm[0] *= fsign(dir[0]);
    nrm[1] *= fsign(dir[1]);
    nrm[2] *= fsign(dir[2]);
    float t = (tmin > 0.0f) ? tmin : tmax;
    for (int k = 0; k < 3; ++k) inter[k] = pos[k] + t * dir[k];
    for (int k = 0; k < 3; ++k)
        inter[k] = fminf(fmaxf(inter[k], g.tl[k]), g.br[k] - 1.1920929e-07f);
    return tmax >= tmin && tmax >= 0.0f;
}

/* Parked interactions of the voxel kernel are handled once this many lanes
 * of a wave hold one, or after kVoxDeferWait iterations. */
constexpr uint32_t kVoxDeferMin = 24, kVoxDeferWait = 8;
/* Steps per loop iteration of the voxel kernel (at most). */
constexpr int kVoxRoundsPerIteration = 4;

template <int PF, int METHOD, bool TRACE, class M>
__global__ void __launch_bounds__(XO_BLOCK) mcvox_kernel(const McvoxParams P) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ BlockStats sstats;
    __shared__ VoxEventCfgL sev;
    MaterialL *sm = reinterpret_cast<MaterialL *>(smem);
    unsigned long long *sfl = reinterpret_cast<unsigned long long *>(
        smem + XO_MAX_MATERIALS * sizeof(MaterialL));
    unsigned long long *sbins = sfl + (P.fl_box_on ? kFlBoxBins : 0u);
    float *slut = reinterpret_cast<float *>(sbins + P.lds_bins);

    const uint32_t tid = threadIdx.x;
    const uint32_t nm = P.cfg.num_materials;
    if (tid < nm) {
        const xo_material &Mi = P.cfg.materials[tid];
        MaterialL o;
        o.inv_mut = Mi.inv_mut;
        o.mua_inv_mut = Mi.mua_inv_mut;
        o.n = Mi.n;
        o.g = Mi.pf.g;
        o.beta = Mi.pf.beta;
        o.la = Mi.pf.a;
        o.lb = Mi.pf.b;
        o.lc = Mi.pf.c;
        o.loff = Mi.pf.offset;
        o.lsize_m1 = Mi.pf.size > 0u ? Mi.pf.size - 1u : 0u;
        o.mua = Mi.mua;
        o.mus = Mi.mus;
        sm[tid] = o;
    }
    for (uint32_t i = tid; i < P.lds_bins; i += XO_BLOCK) sbins[i] = 0ull;
    if (P.fl_box_on)
        for (uint32_t i = tid; i < kFlBoxBins; i += XO_BLOCK) sfl[i] = 0ull;
    for (uint32_t i = tid; i < P.lds_lut; i += XO_BLOCK) slut[i] = P.lut[i];
    if (tid < XO_STAT_COUNT) sstats.v[tid] = 0ull;
    if (tid == 0) {
        sev.top = P.cfg.top;
        sev.bottom = P.cfg.bottom;
        sev.specular = P.cfg.specular;
        sev.fluence = P.cfg.fluence;
        sev.line = P.cfg.line;
        sev.gaussian = P.cfg.gaussian;
    }
    __syncthreads();

    const Bins bins{P.lds_bins ? sbins : nullptr, P.accum};
    const float *lut = P.lds_lut ? slut : P.lut;
    const uint8_t *vox = P.voxels;

    VoxGeom g;
    for (int k = 0; k < 3; ++k) {
        g.tl[k] = P.cfg.voxels.top_left[k];
        g.br[k] = P.cfg.voxels.bottom_right[k];
        g.size[k] = P.cfg.voxels.size[k];
        g.shape[k] = P.cfg.voxels.shape[k];
    }
    const bool use_fluence = P.cfg.fluence.enabled != 0;
    /* a fluence deposit: through the LDS box when there is one */
    FluenceCache fcache;
    auto deposit_fluence = [&](float qx, float qy, float qz, float wgt, float mua) {
        if (P.fl_box_on)
            return fluence_deposit_boxed<M>(lds_fresh(sev.fluence), P.accum, sfl, P.fl_box,
                                            kFlBoxX, kFlBoxY, kFlBoxZ, qx, qy, qz, wgt, mua,
                                            fcache);
        return fluence_deposit<M>(lds_fresh(sev.fluence), P.accum, qx, qy, qz, wgt, mua,
                                  &fcache);
    };

    const uint32_t gid = blockIdx.x * XO_BLOCK + tid;
    const bool lane_ok = gid < P.lanes;
    const bool static_sched = (P.flags & XO_FLAG_STATIC_SCHEDULE) != 0u;
    const bool keyed = (P.flags & XO_FLAG_PACKET_KEYED) != 0u;
    uint64_t x = 0;
    uint32_t a = 0;
    uint64_t my_n = 0, my_k = 0;
    if (lane_ok) {
        if (!keyed) {
            x = P.rng_x[gid];
            a = P.rng_a[gid];
        }
        if (static_sched)
            my_n = P.npackets / P.lanes + ((uint64_t)gid < P.npackets % P.lanes ? 1u : 0u);
    }

    const bool line_src = P.cfg.source_kind == XO_SRC_LINE;
    const float *sp = line_src ? P.cfg.line.position : P.cfg.gaussian.position;
    const float spx = sp[0], spy = sp[1], spz = sp[2];
    const float rmax2 = P.cfg.rmax * P.cfg.rmax;
    const float w_min = P.cfg.w_min, chance = P.cfg.lottery_chance;
    const int use_lottery = P.cfg.use_lottery;

    bool active = lane_ok;
    WaveDispenser wq;
    bool alive = false;
    float p[3] = {0.f, 0.f, 0.f}, d[3] = {0.f, 0.f, 0.f};
    float w = 0.f;
    int32_t vi[3] = {0, 0, 0};
    uint32_t mi = 0;
    uint32_t st[XO_STAT_COUNT] = {0, 0, 0, 0, 0, 0, 0};
    uint32_t pkt_steps = 0;
    TraceState ts{0, 0, 0, 0.0f};
    const xo_trace &T = P.trace;
    bool pend_int = false; /* the lane's step ended on an interaction      */
    uint32_t iwait = 0;    /* iterations parked interactions have waited   */

    for (;;) {
        bool want = active && !alive;
        bool launch = false;
        uint64_t pkt_index = 0;
        if (static_sched) {
            if (want) {
                if (my_k < my_n) {
                    pkt_index = (uint64_t)gid + my_k * (uint64_t)P.lanes;
                    ++my_k;
                    launch = true;
                } else {
                    active = false;
                }
            }
        } else {
            bool retire = false;
            uint64_t claimed = 0;
            launch = dispense(wq, want, alive, P.counter, P.npackets, claim_size(P), kRefillMin,
                              pkt_index, retire, claimed);
            if (retire) active = false;
            if (claimed && line_src && lane_id() == (uint32_t)__builtin_ctzll(__ballot(true))) {
                const xo_source_line &src = lds_fresh(sev.line);
                detector_deposit_n<M>(lds_fresh(sev.specular), bins, src.position[0],
                                      src.position[1],
                                   src.direction_reflected[0], src.direction_reflected[1],
                                   src.direction_reflected[2], src.reflectance,
                                   (uint32_t)claimed);
            }
        }
        if (launch) {
            if (keyed)
                xo_packet_state(P.cfg.packet_seed, P.cfg.first_packet + pkt_index, P.table, &x,
                                &a);
            if (line_src) {
                /* xopto/mcvox/mcsource/line.py:84-104 */
                const xo_source_line &src = lds_fresh(sev.line);
                w = 1.0f - src.reflectance;
                for (int k = 0; k < 3; ++k) {
                    p[k] = src.position[k];
                    d[k] = src.direction_sample[k];
                }
                if (static_sched)
                    detector_deposit<M>(lds_fresh(sev.specular), bins, p[0], p[1],
                                     src.direction_reflected[0], src.direction_reflected[1],
                                     src.direction_reflected[2], src.reflectance);
            } else {
                /* xopto/mcvox/mcsource/gaussianbeam.py:112-189 */
                const xo_source_gaussian &src = lds_fresh(sev.gaussian);
                float specular_r = 0.0f;
                float r = M::sqrt(-2.0f * M::log_unit(1.0f - rng_next(x, a)));
                r = fminf(r, src.clip);
                float sin_fi, cos_fi;
                M::sincos_turns(rng_next(x, a), &sin_fi, &cos_fi);
                float ps0 = r * cos_fi * src.sigma[0];
                float ps1 = r * sin_fi * src.sigma[1];
                float ps2 = 0.0f;
                const float *T = src.transformation;
                float pm[3];
                for (int k = 0; k < 3; ++k)
                    pm[k] = T[3 * k] * ps0 + T[3 * k + 1] * ps1 + T[3 * k + 2] * ps2;
                for (int k = 0; k < 3; ++k) pm[k] += src.position[k];
                float dir[3] = {src.direction[0], src.direction[1], src.direction[2]};
                bool inside = pm[0] >= g.tl[0] && pm[0] < g.br[0] && pm[1] >= g.tl[1] &&
                              pm[1] < g.br[1] && pm[2] >= g.tl[2] && pm[2] < g.br[2];
                if (inside) {
                    dir[0] = -dir[0];
                    dir[1] = -dir[1];
                    dir[2] = -dir[2];
                }
                float inter[3], nrm[3];
                if (box_intersect(g, pm, dir, inter, nrm)) {
                    for (int k = 0; k < 3; ++k) {
                        p[k] = inter[k];
                        dir[k] = -dir[k];
                        nrm[k] = -nrm[k];
                    }
                    float n_out = sm[0].n;
                    int32_t ci[3];
                    for (int k = 0; k < 3; ++k) {
                        int32_t t = f2i_sat((inter[k] - g.tl[k]) / g.size[k]);
                        ci[k] = t < 0 ? 0 : (t > g.shape[k] - 1 ? g.shape[k] - 1 : t);
                    }
                    uint32_t cm = vox_material(vox, g, ci[0], ci[1], ci[2]);
                    float n_in = sm[cm].n;
                    float refr[3] = {dir[0], dir[1], dir[2]};
                    {
                        /* refract_safe, mcbase.template.c:1487-1508 */
                        float cos1 = dir[0] * nrm[0] + dir[1] * nrm[1] + dir[2] * nrm[2];
                        float n1_d_n2 = M::div(n_out, n_in);
                        float s2 = n1_d_n2 * n1_d_n2 * (1.0f - cos1 * cos1);
                        if (!(s2 > 1.0f)) {
                            float k = fsign(cos1) *
                                      (n1_d_n2 * fabsf(cos1) - M::sqrt(1.0f - s2));
                            for (int q = 0; q < 3; ++q) refr[q] = n1_d_n2 * dir[q] - k * nrm[q];
                        }
                    }
                    for (int k = 0; k < 3; ++k) d[k] = refr[k];
                    specular_r = reflectance<M>(n_out, n_in,
                                                nrm[0] * dir[0] + nrm[1] * dir[1] + nrm[2] * dir[2],
                                                cos_critical<M>(n_out, n_in));
                    detector_deposit<M>(lds_fresh(sev.specular), bins, p[0], p[1], dir[0], dir[1], dir[2],
                                     specular_r);
                } else {
                    for (int k = 0; k < 3; ++k) p[k] = g.tl[k];
                    d[0] = 0.0f;
                    d[1] = 0.0f;
                    d[2] = 1.0f;
                    specular_r = 1.0f;
                }
                w = 1.0f - specular_r;
            }
            /* mcsim_set_voxel_index_from_position, mcvox.template.c:245-251 */
            for (int k = 0; k < 3; ++k) vi[k] = f2i_sat((p[k] - g.tl[k]) / g.size[k]);
            mi = vox_valid(g, vi[0], vi[1], vi[2]) ? vox_material(vox, g, vi[0], vi[1], vi[2])
                                                    : 0u;
            alive = true;
            pkt_steps = 0;
            st[XO_STAT_PACKETS]++;
            if (TRACE) {
                /* mcvox.template.c:750-756, 1015-1034 */
                ts = TraceState{pkt_index, 0u, XO_EVENT_LAUNCH, 0.0f};
                if (T.options & XO_TRACE_START)
                    trace_event(T, P.trace_data, ts, p[0], p[1], p[2], d[0], d[1], d[2], w);
            }
        }
        if (!__any(alive)) break;
        /* ---- interactions are deferred: a lane whose step ends inside its
         * voxel parks the interaction (absorption, scattering, roulette),
         * and the parked interactions of the wave are handled together once
         * kVoxDeferMin lanes hold one, every live lane holds one, or they
         * waited kVoxDeferWait iterations. Most steps end on a voxel face
         * (a few instructions when the material does not change), so
         * otherwise the long interaction code would run in nearly every
         * iteration for a handful of lanes. Each packet's arithmetic and
         * draws are those of the reference's loop; only when the wave
         * issues them changes. The decision is taken by every lane of the
         * wave, so the wait count stays wave-uniform. ---- */
        const uint32_t n_int = (uint32_t)__builtin_popcountll(__ballot(alive && pend_int));
        const uint32_t n_alive = (uint32_t)__builtin_popcountll(__ballot(alive));
        const bool run_int = n_int != 0 && (n_int >= kVoxDeferMin || n_int == n_alive ||
                                            iwait >= kVoxDeferWait);
        iwait = (n_int != 0 && !run_int) ? iwait + 1u : 0u;
        if (!alive) continue;

        /* end of one iteration of while(!done), mcvox.template.c:865-1005 */
        auto end_iteration = [&](bool done) {
            /* microscopic Beer-Lambert: roulette after the boundary or the
             * scattering, whether or not the packet has left
             * (mcvox.template.c:865-881) */
            if (METHOD == XO_METHOD_MBL && roulette(w, x, a, w_min, chance, use_lottery))
                done = true;
            float ex = p[0] - spx, ey = p[1] - spy, ez = p[2] - spz;
            if ((rmax2 < XO_INF_F && ex * ex + ey * ey + ez * ez > rmax2) || w <= 0.0f) {
                done = true;
                if (TRACE) ts.ev |= XO_EVENT_ESCAPED;
            }
            if (!done && ++pkt_steps >= XO_MAX_PACKET_STEPS) {
                done = true;
                st[XO_STAT_TRUNCATED]++;
            }
            if (TRACE) {
                /* mcvox.template.c:989-1005 */
                if (done) ts.ev |= XO_EVENT_TERMINATED;
                if (T.options == XO_TRACE_ALL || ((T.options & XO_TRACE_END) && done))
                    trace_event(T, P.trace_data, ts, p[0], p[1], p[2], d[0], d[1], d[2], w);
                ts.ev = 0u;
                if (done) P.trace_n[ts.pkt] = (int32_t)ts.count;
            }
            if (done) alive = false;
        };

        if (run_int && pend_int) {
            /* the parked interaction, mcvox.template.c:925-968 */
            const MaterialL &Mt = sm[mi];
            bool done = false;
            if (METHOD == XO_METHOD_AR && rng_next(x, a) < Mt.mua_inv_mut) {
                /* albedo rejection: the whole weight is absorbed here,
                 * mcvox.template.c:925-944 */
                const float deposit = w;
                w -= deposit;
                if (TRACE) ts.ev |= XO_EVENT_ABSORPTION;
                if (use_fluence && deposit_fluence(p[0], p[1], p[2], deposit, Mt.mua))
                    st[XO_STAT_DEPOSITS]++;
                done = true;
            } else {
                if (METHOD == XO_METHOD_AW) {
                    float deposit = w * Mt.mua_inv_mut;
                    w -= deposit;
                    if (TRACE) ts.ev |= XO_EVENT_ABSORPTION;
                    /* mcsim_fluence_deposit_at, fluence.py:102-147 */
                    if (use_fluence && deposit_fluence(p[0], p[1], p[2], deposit, Mt.mua))
                        st[XO_STAT_DEPOSITS]++;
                }
                PfL pf{Mt.g, Mt.beta, Mt.la, Mt.lb, Mt.lc, Mt.loff, Mt.lsize_m1};
                float xi_fi;
                float ct = pf_sample<PF, M>(pf, lut, x, a, xi_fi);
                scatter_direction<M>(d[0], d[1], d[2], ct, xi_fi);
                if (TRACE) ts.ev |= XO_EVENT_SCATTERING;
                if (METHOD == XO_METHOD_AW && roulette(w, x, a, w_min, chance, use_lottery))
                    done = true;
            }
            pend_int = false;
            end_iteration(done);
        }
        if (!alive || pend_int) continue;

        /* up to kVoxRoundsPerIteration steps per iteration: a lane goes on
         * stepping across voxel faces until its step ends inside a voxel
         * (the parked interaction) or its packet ends; the loop-top work
         * is shared by those steps (+4 % on the 200^3 run) */
#pragma unroll
        for (int rep = 0; rep < kVoxRoundsPerIteration; ++rep) {
            if (!alive || pend_int) break;
            /* ---- one step of while(!done), mcvox.template.c:748-924 ---- */
            const MaterialL &Mt = sm[mi];
            st[XO_STAT_STEPS]++;
            float step;
            if (METHOD == XO_METHOD_MBL)
                step = M::div(-M::log_unit(rng_next(x, a)), Mt.mus); /* mcvox.template.c:763-766 */
            else
                step = -M::log_unit(rng_next(x, a)) * Mt.inv_mut;
            step = fminf(step, XO_FLT_MAX);
            /* mcsim_intersect, mcvox.template.c:173-196 */
            float dist[3];
            for (int k = 0; k < 3; ++k) {
                float side = (float)((int32_t)(d[k] >= 0.0f) + vi[k]);
                dist[k] = g.tl[k] + side * g.size[k] - p[k];
                dist[k] = (d[k] != 0.0f) ? M::div(dist[k], d[k]) : XO_INF_F;
            }
            float dd = fminf(dist[0], fminf(dist[1], dist[2]));
            float d_ok = fminf(dd, step);
            p[0] = __builtin_fmaf(d_ok, d[0], p[0]);
            p[1] = __builtin_fmaf(d_ok, d[1], p[1]);
            p[2] = __builtin_fmaf(d_ok, d[2], p[2]);
            /* optical path length, mcvox.template.c:792-798 */
            if (TRACE && T.plon) ts.pl += Mt.n * d_ok;

            bool done = false;
            if (METHOD == XO_METHOD_MBL) {
                /* absorption along the step inside this voxel,
                 * mcvox.template.c:800-818 */
                const float mua = Mt.mua;
                const float fraction = 1.0f - M::exp(-(mua * d_ok));
                const float deposit = fraction * w;
                w -= deposit;
                if (TRACE) ts.ev |= XO_EVENT_ABSORPTION;
                if (use_fluence) {
                    const float back =
                        (mua != 0.0f)
                            ? d_ok - M::div(-M::log_unit(1.0f - rng_next(x, a) * fraction), mua)
                            : 0.0f;
                    if (deposit_fluence(p[0] - back * d[0], p[1] - back * d[1],
                                        p[2] - back * d[2], deposit, mua))
                        st[XO_STAT_DEPOSITS]++;
                }
            }
            if (dd < step) {
                st[XO_STAT_BOUNDARY]++;
                /* mcsim_boundary, mcvox.template.c:275-388 */
                int32_t nrm[3];
                nrm[0] = (dist[0] <= dist[1]) && (dist[0] <= dist[2]);
                nrm[1] = nrm[0] ? 0 : ((dist[1] <= dist[0]) && (dist[1] <= dist[2]));
                nrm[2] = !(nrm[0] + nrm[1]);
                for (int k = 0; k < 3; ++k)
                    if (d[k] < 0.0f) nrm[k] = -nrm[k];
                int32_t nv0 = vi[0] + nrm[0], nv1 = vi[1] + nrm[1], nv2 = vi[2] + nrm[2];
                bool escaping = !vox_valid(g, nv0, nv1, nv2);
                uint32_t next_mi = !escaping ? vox_material(vox, g, nv0, nv1, nv2) : 0u;
                bool crossed = false;
                if (mi == next_mi && !escaping) {
                    crossed = true;
                } else {
                    float n1 = sm[mi].n, n2 = sm[next_mi].n;
                    if (n1 == n2) {
                        crossed = true;
                    } else {
                        float cc = cos_critical<M>(n1, n2);
                        float cos1 = (float)nrm[0] * d[0] + (float)nrm[1] * d[1] +
                                     (float)nrm[2] * d[2];
                        float fn0 = (float)nrm[0], fn1 = (float)nrm[1], fn2 = (float)nrm[2];
                        bool refracted = false;
                        if (cos1 > cc) {
                            float R = reflectance<M>(n1, n2, cos1, cc);
                            if (R < rng_next(x, a)) {
                                /* refract, mcbase.template.c:1393-1410 */
                                float c1 = d[0] * fn0 + d[1] * fn1 + d[2] * fn2;
                                float n1_d_n2 = M::div(n1, n2);
                                float s2 = n1_d_n2 * n1_d_n2 * (1.0f - c1 * c1);
                                float k = fsign(c1) * (n1_d_n2 * fabsf(c1) - M::sqrt(1.0f - s2));
                                d[0] = n1_d_n2 * d[0] - k * fn0;
                                d[1] = n1_d_n2 * d[1] - k * fn1;
                                d[2] = n1_d_n2 * d[2] - k * fn2;
                                refracted = true;
                                crossed = true;
                            }
                        }
                        if (!refracted) {
                            /* reflect, mcbase.template.c:1342-1350 */
                            float pn2 = 2.0f * (d[0] * fn0 + d[1] * fn1 + d[2] * fn2);
                            d[0] = d[0] - fn0 * pn2;
                            d[1] = d[1] - fn1 * pn2;
                            d[2] = d[2] - fn2 * pn2;
                        }
                    }
                }
                if (TRACE)
                    ts.ev |= XO_EVENT_BOUNDARY_HIT |
                             (crossed ? XO_EVENT_REFRACTION : XO_EVENT_REFLECTION);
                if (crossed) {
                    vi[0] = nv0;
                    vi[1] = nv1;
                    vi[2] = nv2;
                    mi = next_mi;
                    st[XO_STAT_CROSSINGS]++;
                    if (escaping) {
                        if (vi[2] < 0)
                            detector_deposit<M>(lds_fresh(sev.top), bins, p[0], p[1], d[0], d[1], d[2],
                                                w);
                        else if (vi[2] >= g.shape[2])
                            detector_deposit<M>(lds_fresh(sev.bottom), bins, p[0], p[1], d[0], d[1],
                                                d[2], w);
                        done = true;
                    }
                }
                end_iteration(done);
            } else {
                pend_int = true;
            }
        }
    }

    cache_flush(fcache, P.accum);
    if (lane_ok) {
        if (!keyed) P.rng_x[gid] = x;
        if (st[XO_STAT_PACKETS]) st[XO_STAT_THREADS] = 1;
    }
    stats_flush(&sstats, st, P.stats);
    if (P.lds_bins) {
        for (uint32_t i = tid; i < P.lds_bins; i += XO_BLOCK)
            if (sbins[i]) atomicAdd(&P.accum[i], sbins[i]);
    }
    if (P.fl_box_on) {
        const xo_fluence &F = P.cfg.fluence;
        for (uint32_t i = tid; i < kFlBoxBins; i += XO_BLOCK) {
            if (!sfl[i]) continue;
            const uint64_t ix = (uint64_t)P.fl_box[0] + i % kFlBoxX;
            const uint64_t iy = (uint64_t)P.fl_box[1] + (i / kFlBoxX) % kFlBoxY;
            const uint64_t iz = (uint64_t)P.fl_box[2] + i / (kFlBoxX * kFlBoxY);
            atomicAdd(&P.accum[F.offset + (iz * F.shape[1] + iy) * F.shape[0] + ix], sfl[i]);
        }
    }
}

/* ------------------------------------------------------------------ */
/* launchers                                                          */
/* ------------------------------------------------------------------ */
size_t mcml_lds_bytes(const McmlParams &P) {
    return XO_MAX_LAYERS * sizeof(LayerL) + 8u * P.lds_bins + 4u * P.lds_lut;
}

size_t mcvox_lds_bytes(const McvoxParams &P) {
    return XO_MAX_MATERIALS * sizeof(MaterialL) + (P.fl_box_on ? 8u * kFlBoxBins : 0u) +
           8u * P.lds_bins + 4u * P.lds_lut;
}

template <int PF, bool TRACE, class M>
static hipError_t launch_mcml_pf(const McmlParams &P, uint32_t blocks, size_t lds,
                                 hipStream_t stream) {
    switch (P.cfg.method) {
    case XO_METHOD_AW:
        hipLaunchKernelGGL((mcml_kernel<PF, XO_METHOD_AW, TRACE, M>), dim3(blocks),
                           dim3(XO_BLOCK), lds, stream, P);
        break;
    case XO_METHOD_AR:
        hipLaunchKernelGGL((mcml_kernel<PF, XO_METHOD_AR, TRACE, M>), dim3(blocks),
                           dim3(XO_BLOCK), lds, stream, P);
        break;
    case XO_METHOD_MBL:
        hipLaunchKernelGGL((mcml_kernel<PF, XO_METHOD_MBL, TRACE, M>), dim3(blocks),
                           dim3(XO_BLOCK), lds, stream, P);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

/* Traced runs use the exact arithmetic whatever the flags say. */
template <int PF>
static hipError_t launch_mcml_pf(const McmlParams &P, uint32_t blocks, size_t lds,
                                 hipStream_t stream) {
    if (P.trace.max_events > 0)
        return launch_mcml_pf<PF, true, MathExact>(P, blocks, lds, stream);
    if (P.flags & XO_FLAG_FAST_MATH)
        return launch_mcml_pf<PF, false, MathFast>(P, blocks, lds, stream);
    return launch_mcml_pf<PF, false, MathExact>(P, blocks, lds, stream);
}

hipError_t launch_mcml(const McmlParams &P, uint32_t blocks, hipStream_t stream) {
    size_t lds = mcml_lds_bytes(P);
    switch (P.cfg.pf_kind) {
    case XO_PF_HG:
        return launch_mcml_pf<XO_PF_HG>(P, blocks, lds, stream);
    case XO_PF_MHG:
        return launch_mcml_pf<XO_PF_MHG>(P, blocks, lds, stream);
    case XO_PF_LUT:
        return launch_mcml_pf<XO_PF_LUT>(P, blocks, lds, stream);
    default:
        return hipErrorInvalidValue;
    }
}

template <int PF, bool TRACE, class M>
static hipError_t launch_mcvox_pf(const McvoxParams &P, uint32_t blocks, size_t lds,
                                 hipStream_t stream) {
    switch (P.cfg.method) {
    case XO_METHOD_AW:
        hipLaunchKernelGGL((mcvox_kernel<PF, XO_METHOD_AW, TRACE, M>), dim3(blocks),
                           dim3(XO_BLOCK), lds, stream, P);
        break;
    case XO_METHOD_AR:
        hipLaunchKernelGGL((mcvox_kernel<PF, XO_METHOD_AR, TRACE, M>), dim3(blocks),
                           dim3(XO_BLOCK), lds, stream, P);
        break;
    case XO_METHOD_MBL:
        hipLaunchKernelGGL((mcvox_kernel<PF, XO_METHOD_MBL, TRACE, M>), dim3(blocks),
                           dim3(XO_BLOCK), lds, stream, P);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

/* Traced runs use the exact arithmetic whatever the flags say. */
template <int PF>
static hipError_t launch_mcvox_pf(const McvoxParams &P, uint32_t blocks, size_t lds,
                                 hipStream_t stream) {
    if (P.trace.max_events > 0)
        return launch_mcvox_pf<PF, true, MathExact>(P, blocks, lds, stream);
    if (P.flags & XO_FLAG_FAST_MATH)
        return launch_mcvox_pf<PF, false, MathFast>(P, blocks, lds, stream);
    return launch_mcvox_pf<PF, false, MathExact>(P, blocks, lds, stream);
}

hipError_t launch_mcvox(const McvoxParams &P, uint32_t blocks, hipStream_t stream) {
    size_t lds = mcvox_lds_bytes(P);
    switch (P.cfg.pf_kind) {
    case XO_PF_HG:
        return launch_mcvox_pf<XO_PF_HG>(P, blocks, lds, stream);
    case XO_PF_MHG:
        return launch_mcvox_pf<XO_PF_MHG>(P, blocks, lds, stream);
    case XO_PF_LUT:
        return launch_mcvox_pf<XO_PF_LUT>(P, blocks, lds, stream);
    default:
        return hipErrorInvalidValue;
    }
}

/* Blocks per CU of the albedo-weight build of the arithmetic that will
 * run (it stands for all three methods: same registers within a few). */
template <class M>
static hipError_t occupancy_mcml_m(int pf_kind, size_t lds, int *blocks_per_cu) {
    switch (pf_kind) {
    case XO_PF_MHG:
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            blocks_per_cu, mcml_kernel<XO_PF_MHG, XO_METHOD_AW, false, M>, XO_BLOCK, lds);
    case XO_PF_LUT:
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            blocks_per_cu, mcml_kernel<XO_PF_LUT, XO_METHOD_AW, false, M>, XO_BLOCK, lds);
    default:
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            blocks_per_cu, mcml_kernel<XO_PF_HG, XO_METHOD_AW, false, M>, XO_BLOCK, lds);
    }
}

template <class M>
static hipError_t occupancy_mcvox_m(int pf_kind, size_t lds, int *blocks_per_cu) {
    switch (pf_kind) {
    case XO_PF_MHG:
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            blocks_per_cu, mcvox_kernel<XO_PF_MHG, XO_METHOD_AW, false, M>, XO_BLOCK, lds);
    case XO_PF_LUT:
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            blocks_per_cu, mcvox_kernel<XO_PF_LUT, XO_METHOD_AW, false, M>, XO_BLOCK, lds);
    default:
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            blocks_per_cu, mcvox_kernel<XO_PF_HG, XO_METHOD_AW, false, M>, XO_BLOCK, lds);
    }
}

hipError_t occupancy_mcml(int pf_kind, size_t lds, bool fast, int *blocks_per_cu) {
    return fast ? occupancy_mcml_m<MathFast>(pf_kind, lds, blocks_per_cu)
                : occupancy_mcml_m<MathExact>(pf_kind, lds, blocks_per_cu);
}

hipError_t occupancy_mcvox(int pf_kind, size_t lds, bool fast, int *blocks_per_cu) {
    return fast ? occupancy_mcvox_m<MathFast>(pf_kind, lds, blocks_per_cu)
                : occupancy_mcvox_m<MathExact>(pf_kind, lds, blocks_per_cu);
}

} // namespace xo
