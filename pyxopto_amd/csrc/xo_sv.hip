/*
 * xo_sv.hip — sampling volume of a photon packet trace: the SamplingVolume
 * kernel of xopto/mcbase/kernel/mcsv.template.c:236-420 (helpers :27-205).
 *
 * One lane walks one packet's trace at a time, segment by segment between
 * consecutive events, through the voxel lattice of the sampling volume
 * (3D DDA). Each voxel the path leaves gets the integer
 * round(terminal weight * path length in the voxel * multiplier * k) with a
 * u64 atomic in HBM. Lanes claim packets from a global counter, one atomic
 * per packet (traces differ in length by orders of magnitude). The walk is
 * bounded: a packet still walking after XO_MAX_PACKET_STEPS voxel or event
 * steps is stopped and counted as truncated.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>

#include "../../include/xomc.h"
#include "xo_ctx.h"
#include "xo_device.h"
#include "xo_kernels.h"

namespace {

struct SvParams {
    xo_sampling_volume sv;
    const float *data;        /* [packet][max_events][XO_TRACE_ENTRY_LEN] */
    const int32_t *n;         /* [packet]                                  */
    unsigned long long *accum;
    unsigned long long *total;   /* total weight (one u64)                  */
    unsigned long long *counter; /* packet counter                          */
    unsigned long long *stats;   /* XO_STAT_*                               */
    uint64_t npackets;
    int32_t max_events;
};

struct Ev {
    float p[3];
    float d[3];
};

__device__ __forceinline__ void load_event(const float *e, Ev &ev) {
    /* sva_get_event, mcsv.template.c:27-45 */
    const float4 a = *reinterpret_cast<const float4 *>(e);
    const float2 b = *reinterpret_cast<const float2 *>(e + 4);
    ev.p[0] = a.x;
    ev.p[1] = a.y;
    ev.p[2] = a.z;
    ev.d[0] = a.w;
    ev.d[1] = b.x;
    ev.d[2] = b.y;
}

/* mc_distance_point3f (mcbase.template.c:914-916, 1061-1064) */
__device__ __forceinline__ float distance(const float a[3], const float b[3]) {
    const float x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return __builtin_sqrtf(x * x + y * y + z * z);
}

/* (uint32_t)(value + 0.5) of the reference, saturating (a slightly negative
 * path length from rounding gives 0). */
__device__ __forceinline__ uint32_t to_weight(float v) { return xo::f2u_sat(v + 0.5f); }

__global__ void __launch_bounds__(256) sv_kernel(const SvParams P) {
    const xo_sampling_volume &S = P.sv;
    const float kf = (float)S.k;
    uint32_t walked = 0, truncated = 0, packets = 0;
    for (;;) {
        const uint64_t pkt = atomicAdd(P.counter, 1ull);
        if (pkt >= P.npackets) break;
        ++packets;
        const int32_t count = P.n[pkt];
        const int32_t nv = count < P.max_events ? count : P.max_events;
        if (nv <= 0) continue;
        const float *base = P.data + pkt * (uint64_t)P.max_events * XO_TRACE_ENTRY_LEN;
        const float end_w = base[(uint64_t)(nv - 1) * XO_TRACE_ENTRY_LEN + S.end_field];
        atomicAdd(P.total, (unsigned long long)to_weight(end_w * kf));

        Ev e1, e2;
        int32_t idx = 0;
        load_event(base, e1);
        load_event(base + (uint64_t)(nv > 1 ? 1 : 0) * XO_TRACE_ENTRY_LEN, e2);
        float d12 = distance(e1.p, e2.p);
        int32_t v[3];
        for (int k = 0; k < 3; ++k) v[k] = xo::f2i_sat((e1.p[k] - S.top_left[k]) / S.voxel_size[k]);
        float l_voxel = 0.0f;
        uint32_t steps = 0;
        for (;;) {
            /* sva_intersect, mcsv.template.c:95-115 */
            float dist[3];
            for (int k = 0; k < 3; ++k) {
                float side = (float)((int64_t)(e1.d[k] >= 0.0f) + v[k]);
                dist[k] = S.top_left[k] + side * S.voxel_size[k] - e1.p[k];
                dist[k] = (e1.d[k] != 0.0f) ? dist[k] / e1.d[k] : XO_INF_F;
            }
            const float d_voxel = fminf(dist[2], fminf(dist[0], dist[1]));
            const float d_move = fminf(d_voxel, d12);
            l_voxel += d_move;
            const bool valid = v[0] >= 0 && v[1] >= 0 && v[2] >= 0 &&
                               (uint32_t)v[0] < S.shape[0] && (uint32_t)v[1] < S.shape[1] &&
                               (uint32_t)v[2] < S.shape[2];
            bool last = false;
            if (d_voxel < d12) {
                /* leaving this voxel before the next event, :330-350 */
                if (valid) {
                    const uint32_t w = to_weight(end_w * l_voxel * S.multiplier * kf);
                    if (w) {
                        const uint64_t flat =
                            ((uint64_t)v[2] * S.shape[1] + (uint64_t)v[1]) * S.shape[0] + v[0];
                        atomicAdd(&P.accum[S.offset + flat], (unsigned long long)w);
                    }
                }
                e1.p[0] += e1.d[0] * d_move;
                e1.p[1] += e1.d[1] * d_move;
                e1.p[2] += e1.d[2] * d_move;
                d12 -= d_move;
                /* sva_next_voxel, :130-152 */
                int32_t nx = (dist[0] <= dist[1]) && (dist[0] <= dist[2]);
                int32_t ny = nx ? 0 : ((dist[1] <= dist[0]) && (dist[1] <= dist[2]));
                int32_t nz = !(nx + ny);
                v[0] += e1.d[0] < 0.0f ? -nx : nx;
                v[1] += e1.d[1] < 0.0f ? -ny : ny;
                v[2] += e1.d[2] < 0.0f ? -nz : nz;
                l_voxel = 0.0f;
            } else {
                /* the next event comes first, :351-372 */
                e1 = e2;
                ++idx;
                if (idx < nv - 1) {
                    load_event(base + (uint64_t)(idx + 1) * XO_TRACE_ENTRY_LEN, e2);
                    d12 = distance(e1.p, e2.p);
                } else {
                    if (valid) {
                        const uint32_t w = to_weight(end_w * l_voxel * S.multiplier * kf);
                        if (w) {
                            const uint64_t flat =
                                ((uint64_t)v[2] * S.shape[1] + (uint64_t)v[1]) * S.shape[0] +
                                v[0];
                            atomicAdd(&P.accum[S.offset + flat], (unsigned long long)w);
                        }
                    }
                    last = true;
                }
            }
            ++walked;
            if (last) break;
            if (++steps >= XO_MAX_PACKET_STEPS) {
                ++truncated;
                break;
            }
        }
    }
    if (walked) atomicAdd(&P.stats[XO_STAT_STEPS], (unsigned long long)walked);
    if (packets) atomicAdd(&P.stats[XO_STAT_PACKETS], (unsigned long long)packets);
    if (truncated) atomicAdd(&P.stats[XO_STAT_TRUNCATED], (unsigned long long)truncated);
}

} // namespace

extern "C" int xo_sampling_volume_run(xo_ctx *ctx, const xo_sampling_volume *sv,
                                      int32_t max_events, const float *trace_data,
                                      const int32_t *trace_n, uint64_t npackets,
                                      uint64_t *accum, uint64_t naccum,
                                      uint64_t *total_weight, xo_run_report *report) {
    auto t0 = std::chrono::steady_clock::now();
    xo::clear_error();
    if (!ctx || !sv || !accum || !total_weight || (npackets && (!trace_data || !trace_n))) {
        xo::set_error("xo_sampling_volume_run: null argument");
        return -1;
    }
    if (max_events < 1) {
        xo::set_error("xo_sampling_volume_run: max_events must be >= 1");
        return -1;
    }
    if (sv->end_field != 6 && sv->end_field != 7) {
        xo::set_error("xo_sampling_volume_run: end_field must be 6 (w) or 7 (pl)");
        return -1;
    }
    const uint64_t bins = (uint64_t)sv->shape[0] * sv->shape[1] * sv->shape[2];
    if (bins == 0 || sv->offset > naccum || bins > naccum - sv->offset) {
        xo::set_error("xo_sampling_volume_run: %llu bins at offset %u do not fit %llu",
                      (unsigned long long)bins, sv->offset, (unsigned long long)naccum);
        return -1;
    }
    const uint64_t per = (uint64_t)max_events * XO_TRACE_ENTRY_LEN;
    if (npackets > (UINT64_MAX / 4u) / per) {
        xo::set_error("xo_sampling_volume_run: trace too large");
        return -1;
    }
    const uint64_t floats = npackets * per;

    std::lock_guard<std::mutex> lock(ctx->mu);
    XO_HIP(hipSetDevice(ctx->device));
    /* the bins, then the total weight */
    if (xo::ensure_buffer((void **)&ctx->accum, &ctx->accum_cap, (naccum + 1) * 8)) return -1;
    if (npackets &&
        (xo::ensure_buffer((void **)&ctx->trace_data, &ctx->trace_data_cap, floats * 4) ||
         xo::ensure_buffer((void **)&ctx->trace_n, &ctx->trace_n_cap, npackets * 4)))
        return -1;
    SvParams P;
    std::memset(&P, 0, sizeof(P));
    P.sv = *sv;
    P.data = ctx->trace_data;
    P.n = ctx->trace_n;
    P.accum = ctx->accum;
    P.total = ctx->accum + naccum;
    P.counter = ctx->counters;
    P.stats = ctx->counters + 1;
    P.npackets = npackets;
    P.max_events = max_events;
    const uint32_t blocks = ctx->cus * 8u;
    if (npackets) {
        XO_HIP(hipMemcpyAsync(ctx->trace_data, trace_data, floats * 4, hipMemcpyHostToDevice,
                              ctx->stream));
        XO_HIP(hipMemcpyAsync(ctx->trace_n, trace_n, npackets * 4, hipMemcpyHostToDevice,
                              ctx->stream));
    }
    XO_HIP(hipMemsetAsync(ctx->accum, 0, (naccum + 1) * 8, ctx->stream));
    XO_HIP(hipMemsetAsync(ctx->counters, 0, sizeof(unsigned long long) * (1 + XO_STAT_COUNT),
                          ctx->stream));
    XO_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (npackets) {
        hipLaunchKernelGGL(sv_kernel, dim3(blocks), dim3(256), 0, ctx->stream, P);
        XO_HIP(hipGetLastError());
    }
    XO_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    unsigned long long counters[1 + XO_STAT_COUNT];
    unsigned long long total = 0;
    XO_HIP(hipMemcpyAsync(accum, ctx->accum, naccum * 8, hipMemcpyDeviceToHost, ctx->stream));
    XO_HIP(hipMemcpyAsync(&total, ctx->accum + naccum, 8, hipMemcpyDeviceToHost, ctx->stream));
    XO_HIP(hipMemcpyAsync(counters, ctx->counters, sizeof(counters), hipMemcpyDeviceToHost,
                          ctx->stream));
    XO_HIP(hipStreamSynchronize(ctx->stream));
    float kms = 0.f;
    XO_HIP(hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
    *total_weight = total;
    if (report) {
        std::memset(report, 0, sizeof(*report));
        report->packets = counters[1 + XO_STAT_PACKETS];
        report->steps = counters[1 + XO_STAT_STEPS];
        report->truncated = counters[1 + XO_STAT_TRUNCATED];
        report->lanes = blocks * 256u;
        report->kernel_ms = kms;
        report->total_ms = xo::ms_since(t0);
    }
    return 0;
}
