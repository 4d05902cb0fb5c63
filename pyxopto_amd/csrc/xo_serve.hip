/*
 * xo_serve.hip — continuous-batching server for simulation requests.
 *
 * Requests (a layered instance, or a voxelized simulation) are admitted
 * when their accumulators, lookup tables and voxel grid fit in fixed device
 * arenas; their accumulators stay resident until the result is taken.
 * Layered requests run in ticks: each tick gives every active layered
 * request a fair share of a packet quantum as items of one batch_kernel
 * launch (a hipGraph captured once per tick slot: counter reset + launch,
 * with the item count read from device memory). Two tick slots alternate,
 * so the next tick is queued while the current one runs and newly admitted
 * requests join the next tick: a short request never waits for a long one
 * to finish. Voxel requests run beside the ticks as keyed mcvox launches of
 * at most vox_tick_packets packets, on four streams, at most two queued
 * per request and sixteen in flight.
 *
 * Every packet draws from the generator keyed on (request seed, packet
 * index), so a request's integers do not depend on what else is served.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/xomc.h"
#include "xo_batch.h"
#include "xo_ctx.h"
#include "xo_internal.h"
#include "xo_kernels.h"

namespace {

/* First-fit allocator over [0, capacity) with coalescing. */
class Arena {
  public:
    explicit Arena(uint64_t capacity = 0) { reset(capacity); }
    void reset(uint64_t capacity) {
        free_.clear();
        cap_ = capacity;
        used_ = 0;
        if (capacity) free_[0] = capacity;
    }
    bool alloc(uint64_t n, uint64_t *off) {
        if (n == 0) {
            *off = 0;
            return true;
        }
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                uint64_t rest = it->second - n;
                uint64_t start = it->first + n;
                free_.erase(it);
                if (rest) free_[start] = rest;
                used_ += n;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        if (n == 0) return;
        used_ -= n;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + n == next->first) {
            n += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += n;
                return;
            }
        }
        free_[off] = n;
    }
    uint64_t used() const { return used_; }
    uint64_t capacity() const { return cap_; }

  private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t cap_ = 0, used_ = 0;
};

enum { KIND_MCML = 0, KIND_MCVOX = 1 };

struct Request {
    uint64_t id = 0;
    int kind = KIND_MCML;
    uint64_t npackets = 0, next = 0;
    uint32_t inflight = 0; /* ticks holding work of this request */
    bool cancelled = false, completed = false;
    uint64_t accum_off = 0, accum_n = 0;
    uint64_t lut_off = 0, lut_n = 0;
    uint64_t vox_off = 0, vox_n = 0;
    uint32_t slot = 0;
    int cls = 0; /* xo::item_class of a layered request */
    uint32_t lds_bins = 0;
    xo_mcvox_cfg vcfg;
    uint64_t steps = 0, boundary = 0, crossings = 0, deposits = 0;
};

constexpr int kSlots = 2;
constexpr int kVoxSlots = 16;  /* voxel launches in flight at once */
constexpr int kVoxStreams = 4; /* voxel launches of different requests overlap */
constexpr int kVoxPerRequest = 2; /* queued launches per voxel request */
constexpr int kCounterWords = 1 + XO_STAT_COUNT; /* u64 words per counter block */

/* One voxel launch in flight: voxel requests run beside the layered ticks,
 * each on its own stream, so a long voxel launch never holds a tick back. */
struct VoxSlot {
    bool busy = false;
    uint64_t id = 0;
    uint64_t packets = 0;
    hipEvent_t ev = nullptr;
    unsigned long long *counters = nullptr;
};

struct TickSlot {
    xo::BatchItem *items_host = nullptr;  /* pinned */
    uint32_t *nitems_host = nullptr;      /* pinned */
    xo::BatchItem *items_dev = nullptr;
    uint32_t *nitems_dev = nullptr;
    unsigned long long *ml_counters = nullptr; /* [item counter | stats] */
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev_ml = nullptr, ev_start = nullptr;
    bool busy = false;
    uint32_t nitems = 0;
    std::vector<uint64_t> requests; /* ids with work in this tick */
    uint64_t packets = 0;
};

} // namespace

struct xo_server {
    xo_ctx *ctx = nullptr;
    xo_serve_opts opts;
    hipStream_t s_ml = nullptr, s_io = nullptr;
    hipStream_t s_vox[kVoxStreams] = {};
    Arena accum_arena, lut_arena, vox_arena;
    unsigned long long *accum = nullptr;
    float *lut = nullptr;
    uint8_t *vox = nullptr;
    xo::LayerL *layers = nullptr;
    xo::BatchHdr *hdr = nullptr;
    std::vector<uint32_t> free_slots;
    TickSlot ticks[kSlots];
    int next_tick = 0;
    VoxSlot vox_slots[kVoxSlots];
    unsigned long long *vox_counters = nullptr; /* kVoxSlots blocks */
    uint32_t ml_blocks = 0;
    std::unordered_map<uint64_t, Request> reqs;
    std::deque<uint64_t> order; /* active requests, admission order */
    std::vector<uint64_t> done;
    uint64_t next_id = 1;
    xo_serve_stats stats;
    std::mutex mu;
};

namespace {

int fail_hip(const char *what, hipError_t e) {
    xo::set_error("%s failed: %s", what, hipGetErrorString(e));
    return -1;
}

#define XO_CHECK(call)                                                                \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) return fail_hip(#call, e_);                             \
    } while (0)

int build_graph(xo_server *S, int k) {
    TickSlot &T = S->ticks[k];
    xo::BatchParams P;
    std::memset(&P, 0, sizeof(P));
    P.layers = S->layers;
    P.hdr = S->hdr;
    P.items = T.items_dev;
    P.lut = S->lut;
    P.table = S->ctx->table;
    P.item_counter = (unsigned int *)T.ml_counters;
    P.accum = S->accum;
    P.stats = T.ml_counters + 1;
    P.nitems_ptr = T.nitems_dev;
    P.nitems = 0;
    P.w_min = S->opts.w_min;
    P.chance = S->opts.lottery_chance;
    P.use_lottery = S->opts.use_lottery;
    P.refill_min = 4u;
    P.defer_min = xo::kDeferMinDefault;
    P.fast_math = S->opts.fast_math ? 1 : 0;
    XO_CHECK(hipStreamBeginCapture(S->s_ml, hipStreamCaptureModeThreadLocal));
    hipError_t e1 = xo::launch_zero_u32((uint32_t *)T.ml_counters, 2 * kCounterWords, S->s_ml);
    hipError_t e2 = xo::launch_batch(P, S->ml_blocks, S->opts.block, S->s_ml);
    hipError_t e3 = hipStreamEndCapture(S->s_ml, &T.graph);
    if (e1 != hipSuccess) return fail_hip("launch_zero_u32 (capture)", e1);
    if (e2 != hipSuccess) return fail_hip("launch_batch (capture)", e2);
    if (e3 != hipSuccess) return fail_hip("hipStreamEndCapture", e3);
    XO_CHECK(hipGraphInstantiate(&T.exec, T.graph, nullptr, nullptr, 0));
    return 0;
}

/* A request that can never fit is an error, not a wait. */
int check_capacity(xo_server *S, const char *fn, uint64_t naccum, uint64_t nlut,
                   uint64_t nvox) {
    if (naccum > S->accum_arena.capacity() || nlut > S->lut_arena.capacity() ||
        nvox > S->vox_arena.capacity()) {
        xo::set_error("%s: request needs %llu bins, %llu LUT floats, %llu voxels; the server "
                      "holds %llu, %llu, %llu",
                      fn, (unsigned long long)naccum, (unsigned long long)nlut,
                      (unsigned long long)nvox,
                      (unsigned long long)S->accum_arena.capacity(),
                      (unsigned long long)S->lut_arena.capacity(),
                      (unsigned long long)S->vox_arena.capacity());
        return -1;
    }
    return 0;
}

void free_request_memory(xo_server *S, Request &R) {
    S->accum_arena.release(R.accum_off, R.accum_n);
    S->lut_arena.release(R.lut_off, R.lut_n);
    S->vox_arena.release(R.vox_off, R.vox_n);
    if (R.kind == KIND_MCML) S->free_slots.push_back(R.slot);
    R.accum_n = R.lut_n = R.vox_n = 0;
}

/* One unit of a request's work (a tick's items or a voxel launch) has
 * finished: complete the request when nothing of it is left. */
void work_done(xo_server *S, uint64_t id) {
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return;
    Request &R = it->second;
    if (R.inflight) R.inflight--;
    if (R.inflight == 0 && (R.next >= R.npackets || R.cancelled) && !R.completed) {
        R.completed = true;
        if (R.cancelled) {
            free_request_memory(S, R);
            S->reqs.erase(it);
        } else {
            S->done.push_back(id);
        }
    }
}

/* Retire finished ticks and voxel launches: fold their counters into the
 * statistics and the requests. Non-blocking unless wait is set. */
int retire(xo_server *S, bool wait) {
    for (int k = 0; k < kSlots; ++k) {
        TickSlot &T = S->ticks[k];
        if (!T.busy) continue;
        if (wait) {
            XO_CHECK(hipEventSynchronize(T.ev_ml));
        } else {
            hipError_t a = hipEventQuery(T.ev_ml);
            if (a == hipErrorNotReady) continue;
            if (a != hipSuccess) return fail_hip("hipEventQuery", a);
        }
        unsigned long long c[kCounterWords];
        XO_CHECK(hipMemcpy(c, T.ml_counters, sizeof(c), hipMemcpyDeviceToHost));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, T.ev_start, T.ev_ml) == hipSuccess)
            S->stats.busy_ms += ms;
        S->stats.ticks++;
        S->stats.packets += T.packets;
        S->stats.steps += c[1 + XO_STAT_STEPS];
        for (uint64_t id : T.requests) work_done(S, id);
        T.busy = false;
        T.requests.clear();
    }
    for (auto &V : S->vox_slots) {
        if (!V.busy) continue;
        if (wait) {
            XO_CHECK(hipEventSynchronize(V.ev));
        } else {
            hipError_t a = hipEventQuery(V.ev);
            if (a == hipErrorNotReady) continue;
            if (a != hipSuccess) return fail_hip("hipEventQuery", a);
        }
        unsigned long long c[kCounterWords];
        XO_CHECK(hipMemcpy(c, V.counters, sizeof(c), hipMemcpyDeviceToHost));
        S->stats.packets += V.packets;
        S->stats.steps += c[1 + XO_STAT_STEPS];
        V.busy = false;
        work_done(S, V.id);
    }
    return 0;
}

int launch_vox(xo_server *S, Request &R, uint64_t first, uint64_t count,
               unsigned long long *counters, hipStream_t stream) {
    xo::McvoxParams P;
    std::memset(&P, 0, sizeof(P));
    P.cfg = R.vcfg;
    P.cfg.packet_seed = R.vcfg.packet_seed;
    P.cfg.first_packet = first;
    P.table = S->ctx->table;
    P.voxels = S->vox + R.vox_off;
    P.lut = R.lut_n ? S->lut + R.lut_off : nullptr;
    P.rng_x = S->ctx->rng_x; /* unused with keyed packets */
    P.rng_a = S->ctx->rng_a;
    P.counter = counters;
    P.stats = counters + 1;
    P.accum = S->accum + R.accum_off;
    P.npackets = count;
    P.flags = XO_FLAG_PACKET_KEYED | (S->opts.fast_math ? XO_FLAG_FAST_MATH : 0u);
    P.lds_bins = R.lds_bins;
    P.lds_lut = (R.lut_n && R.lut_n <= XO_LDS_LUT_MAX) ? (uint32_t)R.lut_n : 0u;
    int bpc = 1;
    XO_CHECK(xo::occupancy_mcvox(R.vcfg.pf_kind, xo::mcvox_lds_bytes(P),
                                 S->opts.fast_math != 0, &bpc));
    uint64_t resident = (uint64_t)std::max(bpc, 1) * S->ctx->cus * XO_BLOCK;
    uint64_t lanes = std::min<uint64_t>(std::min<uint64_t>(count, resident), S->opts.vox_lanes);
    lanes = std::max<uint64_t>(lanes, 64u);
    P.lanes = (uint32_t)lanes;
    XO_CHECK(xo::launch_zero_u32((uint32_t *)counters, 2 * kCounterWords, stream));
    XO_CHECK(xo::launch_mcvox(P, (uint32_t)((lanes + XO_BLOCK - 1) / XO_BLOCK), stream));
    return 0;
}

/* Form and launch one tick in slot k. Returns 1 if there was no work. */
int dispatch_tick(xo_server *S, int k) {
    TickSlot &T = S->ticks[k];
    std::vector<uint64_t> ml;
    for (uint64_t id : S->order) {
        auto it = S->reqs.find(id);
        if (it == S->reqs.end()) continue;
        const Request &R = it->second;
        if (R.cancelled || R.next >= R.npackets || R.kind != KIND_MCML) continue;
        ml.push_back(id);
    }
    if (ml.empty()) return 1;
    /* requests of one class next to each other (same branches, see
     * xo::item_class): waves moving between their items stay coherent */
    std::stable_sort(ml.begin(), ml.end(), [S](uint64_t p, uint64_t q) {
        return S->reqs[p].cls < S->reqs[q].cls;
    });
    T.nitems = 0;
    T.packets = 0;
    /* at most opts.chunk packets per item, but enough items for every wave
     * of the persistent grid to get one */
    const uint64_t waves = (uint64_t)S->ml_blocks * (S->opts.block / 64u);
    const uint64_t chunk = std::max<uint64_t>(
        256u, std::min<uint64_t>(S->opts.chunk, S->opts.tick_packets / std::max<uint64_t>(waves, 1)));
    {
        uint64_t share = S->opts.tick_packets / ml.size();
        share = std::max<uint64_t>(share, chunk);
        for (uint64_t id : ml) {
            Request &R = S->reqs[id];
            uint64_t q = std::min<uint64_t>(share, R.npackets - R.next);
            uint64_t given = 0;
            while (given < q && T.nitems < S->opts.max_items) {
                uint64_t c = std::min<uint64_t>(chunk, q - given);
                T.items_host[T.nitems++] = xo::BatchItem{R.slot, (uint32_t)c, R.next + given};
                given += c;
            }
            if (given == 0) break;
            R.next += given;
            R.inflight++;
            T.requests.push_back(id);
            T.packets += given;
            if (T.nitems >= S->opts.max_items) break;
        }
    }
    *T.nitems_host = T.nitems;
    XO_CHECK(hipEventRecord(T.ev_start, S->s_ml));
    XO_CHECK(hipMemcpyAsync(T.items_dev, T.items_host,
                            sizeof(xo::BatchItem) * std::max<uint32_t>(T.nitems, 1u),
                            hipMemcpyHostToDevice, S->s_ml));
    XO_CHECK(hipMemcpyAsync(T.nitems_dev, T.nitems_host, sizeof(uint32_t),
                            hipMemcpyHostToDevice, S->s_ml));
    if (T.nitems > 0) {
        if (S->opts.use_graph) {
            XO_CHECK(hipGraphLaunch(T.exec, S->s_ml));
        } else {
            XO_CHECK(xo::launch_zero_u32((uint32_t *)T.ml_counters, 2 * kCounterWords, S->s_ml));
            xo::BatchParams P;
            std::memset(&P, 0, sizeof(P));
            P.layers = S->layers;
            P.hdr = S->hdr;
            P.items = T.items_dev;
            P.lut = S->lut;
            P.table = S->ctx->table;
            P.item_counter = (unsigned int *)T.ml_counters;
            P.accum = S->accum;
            P.stats = T.ml_counters + 1;
            P.nitems_ptr = T.nitems_dev;
            P.w_min = S->opts.w_min;
            P.chance = S->opts.lottery_chance;
            P.use_lottery = S->opts.use_lottery;
            P.refill_min = 4u;
            P.defer_min = xo::kDeferMinDefault;
            P.fast_math = S->opts.fast_math ? 1 : 0;
            XO_CHECK(xo::launch_batch(P, S->ml_blocks, S->opts.block, S->s_ml));
        }
    } else {
        XO_CHECK(xo::launch_zero_u32((uint32_t *)T.ml_counters, 2 * kCounterWords, S->s_ml));
    }
    XO_CHECK(hipEventRecord(T.ev_ml, S->s_ml));
    T.busy = true;
    return 0;
}

/* Queue voxel launches: up to kVoxPerRequest per request, each of at most
 * vox_tick_packets packets, on the request's stream (so its launches run
 * in packet order), while launch slots are free. */
int dispatch_vox(xo_server *S) {
    for (uint64_t id : S->order) {
        auto it = S->reqs.find(id);
        if (it == S->reqs.end()) continue;
        Request &R = it->second;
        if (R.kind != KIND_MCVOX) continue;
        while (!R.cancelled && R.next < R.npackets && R.inflight < kVoxPerRequest) {
            VoxSlot *V = nullptr;
            for (auto &slot : S->vox_slots)
                if (!slot.busy) {
                    V = &slot;
                    break;
                }
            if (!V) return 0;
            const uint64_t q = std::min<uint64_t>(S->opts.vox_tick_packets, R.npackets - R.next);
            hipStream_t st = S->s_vox[id % kVoxStreams];
            if (launch_vox(S, R, R.next, q, V->counters, st)) return -1;
            XO_CHECK(hipEventRecord(V->ev, st));
            V->busy = true;
            V->id = id;
            V->packets = q;
            R.next += q;
            R.inflight++;
        }
    }
    return 0;
}

void prune_order(xo_server *S) {
    std::deque<uint64_t> keep;
    for (uint64_t id : S->order) {
        auto it = S->reqs.find(id);
        if (it == S->reqs.end()) continue;
        const Request &R = it->second;
        if (R.completed || (R.next >= R.npackets && R.inflight == 0)) continue;
        keep.push_back(id);
    }
    S->order.swap(keep);
}

} // namespace

extern "C" {

void xo_serve_default_opts(xo_serve_opts *o) {
    std::memset(o, 0, sizeof(*o));
    o->accum_capacity = 64ull << 20;   /* 512 MiB of u64 bins */
    o->lut_capacity = 16ull << 20;     /* 64 MiB of floats */
    o->voxel_capacity = 256ull << 20;  /* 256 MiB of uint8 voxels */
    o->instance_slots = 16384;
    o->max_items = 65536;
    o->tick_packets = 16000000;
    o->vox_tick_packets = 500000;
    o->vox_lanes = 131072;
    o->chunk = 2048;
    o->block = 256;
    o->blocks_per_cu = 0;
    o->w_min = 1e-4f;
    o->lottery_chance = 0.1f;
    o->use_lottery = 1;
    o->use_graph = 1;
}

void xo_serve_destroy(xo_server *S) {
    if (!S) return;
    (void)hipSetDevice(S->ctx->device);
    if (S->s_ml) (void)hipStreamSynchronize(S->s_ml);
    for (auto s : S->s_vox)
        if (s) (void)hipStreamSynchronize(s);
    for (auto &T : S->ticks) {
        if (T.exec) (void)hipGraphExecDestroy(T.exec);
        if (T.graph) (void)hipGraphDestroy(T.graph);
        if (T.items_host) (void)hipHostFree(T.items_host);
        if (T.nitems_host) (void)hipHostFree(T.nitems_host);
        if (T.items_dev) (void)hipFree(T.items_dev);
        if (T.nitems_dev) (void)hipFree(T.nitems_dev);
        if (T.ml_counters) (void)hipFree(T.ml_counters);
        if (T.ev_ml) (void)hipEventDestroy(T.ev_ml);
        if (T.ev_start) (void)hipEventDestroy(T.ev_start);
    }
    for (auto &V : S->vox_slots)
        if (V.ev) (void)hipEventDestroy(V.ev);
    if (S->vox_counters) (void)hipFree(S->vox_counters);
    if (S->accum) (void)hipFree(S->accum);
    if (S->lut) (void)hipFree(S->lut);
    if (S->vox) (void)hipFree(S->vox);
    if (S->layers) (void)hipFree(S->layers);
    if (S->hdr) (void)hipFree(S->hdr);
    if (S->s_ml) (void)hipStreamDestroy(S->s_ml);
    for (auto s : S->s_vox)
        if (s) (void)hipStreamDestroy(s);
    if (S->s_io) (void)hipStreamDestroy(S->s_io);
    delete S;
}

xo_server *xo_serve_create(xo_ctx *ctx, const xo_serve_opts *opts) {
    xo::clear_error();
    if (!ctx) {
        xo::set_error("xo_serve_create: null context");
        return nullptr;
    }
    xo_server *S = new xo_server();
    S->ctx = ctx;
    if (opts)
        S->opts = *opts;
    else
        xo_serve_default_opts(&S->opts);
    xo_serve_opts &o = S->opts;
    if (o.block != 64 && o.block != 128 && o.block != 256) o.block = 256;
    if (!o.chunk) o.chunk = 2048;
    if (!o.max_items) o.max_items = 65536;
    if (!o.instance_slots) o.instance_slots = 16384;
    if (!o.vox_lanes) o.vox_lanes = 131072;
    std::memset(&S->stats, 0, sizeof(S->stats));
    auto bad = [&](const char *what, hipError_t e) -> xo_server * {
        xo::set_error("xo_serve_create: %s failed: %s", what, hipGetErrorString(e));
        xo_serve_destroy(S);
        return nullptr;
    };
    hipError_t e;
    if ((e = hipSetDevice(ctx->device)) != hipSuccess) return bad("hipSetDevice", e);
    if (xo::ensure_table(ctx)) {
        xo_serve_destroy(S);
        return nullptr;
    }
    if ((e = hipStreamCreateWithFlags(&S->s_ml, hipStreamNonBlocking)) != hipSuccess)
        return bad("stream", e);
    for (auto &sv : S->s_vox)
        if ((e = hipStreamCreateWithFlags(&sv, hipStreamNonBlocking)) != hipSuccess)
            return bad("stream", e);
    if ((e = hipStreamCreateWithFlags(&S->s_io, hipStreamNonBlocking)) != hipSuccess)
        return bad("stream", e);
    if ((e = hipMalloc(&S->accum, 8 * o.accum_capacity)) != hipSuccess) return bad("accum", e);
    if ((e = hipMalloc(&S->lut, 4 * std::max<uint64_t>(o.lut_capacity, 1))) != hipSuccess)
        return bad("lut", e);
    if ((e = hipMalloc(&S->vox, std::max<uint64_t>(o.voxel_capacity, 1))) != hipSuccess)
        return bad("voxels", e);
    if ((e = hipMalloc(&S->layers, sizeof(xo::LayerL) * XO_BATCH_MAX_LAYERS *
                                       (size_t)o.instance_slots)) != hipSuccess)
        return bad("layers", e);
    if ((e = hipMalloc(&S->hdr, sizeof(xo::BatchHdr) * (size_t)o.instance_slots)) !=
        hipSuccess)
        return bad("headers", e);
    S->accum_arena.reset(o.accum_capacity);
    S->lut_arena.reset(o.lut_capacity);
    S->vox_arena.reset(o.voxel_capacity);
    for (uint32_t i = o.instance_slots; i > 0; --i) S->free_slots.push_back(i - 1);
    int bpc = 1;
    if ((e = xo::occupancy_batch(o.block, &bpc)) != hipSuccess) return bad("occupancy", e);
    if (o.blocks_per_cu && (int)o.blocks_per_cu < bpc) bpc = (int)o.blocks_per_cu;
    S->ml_blocks = (uint32_t)std::max(bpc, 1) * ctx->cus;
    for (int k = 0; k < kSlots; ++k) {
        TickSlot &T = S->ticks[k];
        if ((e = hipHostMalloc((void **)&T.items_host, sizeof(xo::BatchItem) * o.max_items)) !=
            hipSuccess)
            return bad("pinned items", e);
        if ((e = hipHostMalloc((void **)&T.nitems_host, sizeof(uint32_t))) != hipSuccess)
            return bad("pinned count", e);
        if ((e = hipMalloc(&T.items_dev, sizeof(xo::BatchItem) * o.max_items)) != hipSuccess)
            return bad("items", e);
        if ((e = hipMalloc(&T.nitems_dev, sizeof(uint32_t))) != hipSuccess)
            return bad("count", e);
        if ((e = hipMalloc(&T.ml_counters, 8 * kCounterWords)) != hipSuccess)
            return bad("counters", e);
        if ((e = hipEventCreate(&T.ev_ml)) != hipSuccess) return bad("event", e);
        if ((e = hipEventCreate(&T.ev_start)) != hipSuccess) return bad("event", e);
        if (o.use_graph && build_graph(S, k)) {
            xo_serve_destroy(S);
            return nullptr;
        }
    }
    if ((e = hipMalloc(&S->vox_counters, 8 * kCounterWords * kVoxSlots)) != hipSuccess)
        return bad("counters", e);
    for (int v = 0; v < kVoxSlots; ++v) {
        S->vox_slots[v].counters = S->vox_counters + (size_t)kCounterWords * v;
        if ((e = hipEventCreateWithFlags(&S->vox_slots[v].ev, hipEventDisableTiming)) !=
            hipSuccess)
            return bad("event", e);
    }
    return S;
}

int xo_serve_admit_mcml(xo_server *S, const xo_batch_instance *inst, const float *lut,
                        uint32_t nlut, uint64_t *id_out) {
    xo::clear_error();
    if (!S || !inst || !id_out) {
        xo::set_error("xo_serve_admit_mcml: null argument");
        return -1;
    }
    const xo_batch_instance &I = *inst;
    if (I.num_layers < 3 || I.num_layers > XO_BATCH_MAX_LAYERS) {
        xo::set_error("xo_serve_admit_mcml: %u layers (3..%d supported)", I.num_layers,
                      XO_BATCH_MAX_LAYERS);
        return -1;
    }
    if (I.source_kind != XO_SRC_LINE && I.source_kind != XO_SRC_FIBER_NI &&
        I.source_kind != XO_SRC_FIBER) {
        xo::set_error("xo_serve_admit_mcml: unknown source kind %d", I.source_kind);
        return -1;
    }
    if (I.source_kind == XO_SRC_FIBER && !(I.fiber_axis.direction[2] > 0.0f)) {
        xo::set_error("xo_serve_admit_mcml: the fiber direction must point into the sample");
        return -1;
    }
    for (uint32_t l = 0; l < I.num_layers; ++l)
        if (xo::check_pf(I.pf_kind, I.layers[l].pf, nlut, "layer", l)) return -1;
    if (I.fluence.enabled != XO_FLUENCE_NONE) {
        xo::set_error("xo_serve_admit_mcml: layered requests with a fluence grid are not "
                      "served (run them with xo_batch_run)");
        return -1;
    }
    if (I.pf_kind == XO_PF_LUT && !lut) {
        xo::set_error("xo_serve_admit_mcml: a lookup-table phase function needs the LUT");
        return -1;
    }
    /* detector offsets are relative to accum_base, as in a batch */
    const uint64_t naccum = I.accum_count;
    xo_detector det[3] = {I.top, I.bottom, I.specular};
    const char *names[3] = {"top", "bottom", "specular"};
    uint64_t end = 0;
    if (naccum == 0) {
        xo::set_error("xo_serve_admit_mcml: no accumulator bins");
        return -1;
    }
    for (int k = 0; k < 3; ++k) {
        if (det[k].kind != XO_DET_NONE) {
            if (det[k].offset < I.accum_base) {
                xo::set_error("xo_serve_admit_mcml: %s detector below accum_base", names[k]);
                return -1;
            }
            det[k].offset -= I.accum_base;
        }
        if (xo::check_detector(names[k], det[k], naccum, &end)) return -1;
    }
    std::lock_guard<std::mutex> lock(S->mu);
    if (check_capacity(S, "xo_serve_admit_mcml", naccum, nlut, 0)) return -1;
    XO_CHECK(hipSetDevice(S->ctx->device));
    if (S->free_slots.empty()) return 1;
    Request R;
    R.kind = KIND_MCML;
    if (!S->accum_arena.alloc(naccum, &R.accum_off)) return 1;
    R.accum_n = naccum;
    if (!S->lut_arena.alloc(nlut, &R.lut_off)) {
        S->accum_arena.release(R.accum_off, naccum);
        return 1;
    }
    R.lut_n = nlut;
    R.slot = S->free_slots.back();
    S->free_slots.pop_back();
    R.cls = xo::item_class(I);
    R.npackets = I.npackets;
    R.id = S->next_id++;
    /* rebase into the arenas */
    xo::LayerL L[XO_BATCH_MAX_LAYERS];
    for (uint32_t l = 0; l < XO_BATCH_MAX_LAYERS; ++l) {
        xo_layer lay = I.layers[l < I.num_layers ? l : I.num_layers - 1];
        lay.pf.offset += (uint32_t)R.lut_off;
        L[l] = xo::layer_to_device(lay);
    }
    const xo::BatchHdr H = xo::batch_header(I, det, (uint32_t)R.accum_off);
    /* Uploads go on the I/O stream, which holds no tick, so admission does
     * not wait for the running tick; the slot and arena ranges are unused by
     * any tick in flight, and the host synchronizes before the request can
     * join a tick. */
    XO_CHECK(hipMemcpyAsync(S->layers + (size_t)R.slot * XO_BATCH_MAX_LAYERS, L, sizeof(L),
                            hipMemcpyHostToDevice, S->s_io));
    XO_CHECK(hipMemcpyAsync(S->hdr + R.slot, &H, sizeof(H), hipMemcpyHostToDevice, S->s_io));
    if (nlut)
        XO_CHECK(hipMemcpyAsync(S->lut + R.lut_off, lut, 4ull * nlut, hipMemcpyHostToDevice,
                                S->s_io));
    XO_CHECK(xo::launch_zero_u32((uint32_t *)(S->accum + R.accum_off), (uint32_t)(2 * naccum),
                                 S->s_io));
    XO_CHECK(hipStreamSynchronize(S->s_io));
    *id_out = R.id;
    if (R.npackets == 0) {
        R.completed = true; /* nothing to run: complete at once */
        S->done.push_back(R.id);
    } else {
        S->order.push_back(R.id);
    }
    S->reqs[R.id] = R;
    S->stats.admitted++;
    return 0;
}

int xo_serve_admit_mcvox(xo_server *S, const xo_mcvox_cfg *cfg, const int32_t *voxels,
                         uint64_t nvoxels, const float *lut, uint32_t nlut, uint64_t naccum,
                         uint64_t npackets, uint64_t seed, uint64_t *id_out) {
    xo::clear_error();
    if (!S || !cfg || !voxels || !id_out || naccum == 0) {
        xo::set_error("xo_serve_admit_mcvox: null or empty argument");
        return -1;
    }
    const int32_t *shape = cfg->voxels.shape;
    if (shape[0] <= 0 || shape[1] <= 0 || shape[2] <= 0 ||
        (uint64_t)shape[0] * shape[1] * shape[2] != nvoxels) {
        xo::set_error("xo_serve_admit_mcvox: voxel count does not match the grid");
        return -1;
    }
    if (cfg->num_materials < 1 || cfg->num_materials > XO_MAX_MATERIALS ||
        (cfg->method != XO_METHOD_AW && cfg->method != XO_METHOD_AR &&
         cfg->method != XO_METHOD_MBL) ||
        (cfg->source_kind != XO_SRC_LINE && cfg->source_kind != XO_SRC_GAUSSIAN)) {
        xo::set_error("xo_serve_admit_mcvox: invalid configuration");
        return -1;
    }
    for (uint32_t i = 0; i < cfg->num_materials; ++i)
        if (xo::check_pf(cfg->pf_kind, cfg->materials[i].pf, nlut, "material", i)) return -1;
    if (cfg->pf_kind == XO_PF_LUT && !lut) {
        xo::set_error("xo_serve_admit_mcvox: a lookup-table phase function needs the LUT");
        return -1;
    }
    uint64_t end = 0;
    if (xo::check_detector("top", cfg->top, naccum, &end) ||
        xo::check_detector("bottom", cfg->bottom, naccum, &end) ||
        xo::check_detector("specular", cfg->specular, naccum, &end))
        return -1;
    if (cfg->fluence.enabled) {
        const uint32_t *fs = cfg->fluence.shape;
        uint64_t nf = (uint64_t)fs[0] * fs[1] * fs[2];
        if (nf == 0 || cfg->fluence.offset < end || (uint64_t)cfg->fluence.offset + nf > naccum) {
            xo::set_error("xo_serve_admit_mcvox: fluence bins must follow the detector bins "
                          "inside the accumulator");
            return -1;
        }
    }
    std::vector<uint8_t> packed;
    if (xo::pack_voxels(voxels, shape, cfg->num_materials, packed, "xo_serve_admit_mcvox"))
        return -1;
    const uint64_t nbytes = packed.size();
    std::lock_guard<std::mutex> lock(S->mu);
    if (check_capacity(S, "xo_serve_admit_mcvox", naccum, nlut, nbytes)) return -1;
    XO_CHECK(hipSetDevice(S->ctx->device));
    Request R;
    R.kind = KIND_MCVOX;
    if (!S->accum_arena.alloc(naccum, &R.accum_off)) return 1;
    R.accum_n = naccum;
    if (!S->lut_arena.alloc(nlut, &R.lut_off)) {
        S->accum_arena.release(R.accum_off, naccum);
        return 1;
    }
    R.lut_n = nlut;
    if (!S->vox_arena.alloc(nbytes, &R.vox_off)) {
        S->accum_arena.release(R.accum_off, naccum);
        S->lut_arena.release(R.lut_off, nlut);
        return 1;
    }
    R.vox_n = nbytes;
    R.vcfg = *cfg;
    R.vcfg.packet_seed = seed;
    R.lds_bins = end <= XO_LDS_BINS_MAX ? (uint32_t)end : 0u;
    R.npackets = npackets;
    R.id = S->next_id++;
    XO_CHECK(hipMemcpyAsync(S->vox + R.vox_off, packed.data(), nbytes, hipMemcpyHostToDevice,
                            S->s_io));
    if (nlut)
        XO_CHECK(hipMemcpyAsync(S->lut + R.lut_off, lut, 4ull * nlut, hipMemcpyHostToDevice,
                                S->s_io));
    XO_CHECK(xo::launch_zero_u32((uint32_t *)(S->accum + R.accum_off), (uint32_t)(2 * naccum),
                                 S->s_io));
    XO_CHECK(hipStreamSynchronize(S->s_io));
    *id_out = R.id;
    if (R.npackets == 0) {
        R.completed = true; /* nothing to run: complete at once */
        S->done.push_back(R.id);
    } else {
        S->order.push_back(R.id);
    }
    S->reqs[R.id] = R;
    S->stats.admitted++;
    return 0;
}

int xo_serve_step(xo_server *S, uint32_t *active) {
    xo::clear_error();
    if (!S) {
        xo::set_error("xo_serve_step: null server");
        return -1;
    }
    std::lock_guard<std::mutex> lock(S->mu);
    XO_CHECK(hipSetDevice(S->ctx->device));
    if (retire(S, false)) return -1;
    for (int tries = 0; tries < kSlots; ++tries) {
        TickSlot &T = S->ticks[S->next_tick];
        if (T.busy) break;
        int rc = dispatch_tick(S, S->next_tick);
        if (rc < 0) return -1;
        if (rc == 1) break;
        S->next_tick = (S->next_tick + 1) % kSlots;
    }
    if (dispatch_vox(S)) return -1;
    prune_order(S);
    if (active) {
        uint32_t n = 0;
        for (auto &kv : S->reqs)
            if (!kv.second.completed) ++n;
        *active = n;
    }
    return 0;
}

int xo_serve_wait(xo_server *S) {
    if (!S) return -1;
    std::lock_guard<std::mutex> lock(S->mu);
    XO_CHECK(hipSetDevice(S->ctx->device));
    return retire(S, true);
}

int xo_serve_poll(xo_server *S, uint64_t *ids, uint32_t max, uint32_t *n) {
    if (!S || !n) {
        xo::set_error("xo_serve_poll: null argument");
        return -1;
    }
    std::lock_guard<std::mutex> lock(S->mu);
    uint32_t k = 0;
    while (k < max && !S->done.empty()) {
        ids[k++] = S->done.front();
        S->done.erase(S->done.begin());
    }
    *n = k;
    return 0;
}

int xo_serve_result(xo_server *S, uint64_t id, uint64_t *accum, uint64_t naccum) {
    xo::clear_error();
    if (!S || !accum) {
        xo::set_error("xo_serve_result: null argument");
        return -1;
    }
    std::lock_guard<std::mutex> lock(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end() || !it->second.completed) {
        xo::set_error("xo_serve_result: request %llu is not complete",
                      (unsigned long long)id);
        return -1;
    }
    Request &R = it->second;
    if (naccum != R.accum_n) {
        xo::set_error("xo_serve_result: request has %llu bins, %llu requested",
                      (unsigned long long)R.accum_n, (unsigned long long)naccum);
        return -1;
    }
    XO_CHECK(hipSetDevice(S->ctx->device));
    XO_CHECK(hipMemcpyAsync(accum, S->accum + R.accum_off, 8 * naccum, hipMemcpyDeviceToHost,
                            S->s_io));
    XO_CHECK(hipStreamSynchronize(S->s_io));
    free_request_memory(S, R);
    S->reqs.erase(it);
    S->stats.completed++;
    return 0;
}

int xo_serve_cancel(xo_server *S, uint64_t id) {
    if (!S) return -1;
    std::lock_guard<std::mutex> lock(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) {
        xo::set_error("xo_serve_cancel: unknown request %llu", (unsigned long long)id);
        return -1;
    }
    Request &R = it->second;
    R.cancelled = true;
    if (R.inflight == 0) {
        if (R.completed) {
            auto d = std::find(S->done.begin(), S->done.end(), id);
            if (d != S->done.end()) S->done.erase(d);
        }
        free_request_memory(S, R);
        S->reqs.erase(it);
    }
    S->stats.cancelled++;
    return 0;
}

int xo_serve_get_stats(xo_server *S, xo_serve_stats *out) {
    if (!S || !out) return -1;
    std::lock_guard<std::mutex> lock(S->mu);
    *out = S->stats;
    out->accum_used = S->accum_arena.used();
    out->voxel_used = S->vox_arena.used();
    out->lut_used = S->lut_arena.used();
    return 0;
}

} // extern "C"
