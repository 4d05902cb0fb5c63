/*
 * xomc.h — C ABI of libxomc.so, the MI355X (gfx950) photon-packet transport
 * engine for layered (mcml) and voxelized (mcvox) turbid media.
 *
 * Every entry point takes plain pointers and sizes. No torch, no HIP types.
 * The structs below are the packed forms of the PyXOpto plugin objects; each
 * cites the reference structure it replaces (paths relative to the PyXOpto
 * source tree). All floats are IEEE binary32, all accumulators are uint64
 * fixed point with K = 0x7FFFFF (xopto/mcbase/kernel/mcbase.template.h:201-205).
 *
 * Status codes: 0 on success, negative on error; xo_last_error() returns the
 * message of the last failing call on the calling thread.
 */
#ifndef XOMC_H
#define XOMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XO_ABI_VERSION 1

/* Fixed-point accumulator scale: weight_to_int(w) = (uint32)(w*K + 0.5).
 * xopto/mcml/kernel/mcml.template.h:606; mcbase.template.h:201-205. */
#define XO_ACCU_K 0x7FFFFF

/* Number of rows of the safe-prime MWC seed table
 * (xopto/data/primes/safeprimes_base32_500k.npz, 500000 rows; the first row
 * seeds the others, so 499999 generator streams are available:
 * xopto/cl/clrng.py:438-439). */
#define XO_NUM_PRIMES 500000u
#define XO_MAX_SEEDS (XO_NUM_PRIMES - 1u)
/* First multiplier of the table, searched downwards from here
 * (xopto/src/rng/find_primes.c:50). */
#define XO_FIRST_PRIME_A 4294967118u

/* A packet still in flight after this many transport steps is terminated
 * (counted in xo_run_report.truncated). The reference has no such limit; it
 * only matters for packets trapped by total internal reflection in a
 * non-absorbing, non-scattering layer, which would otherwise never end. */
#define XO_MAX_PACKET_STEPS (1u << 26)

#define XO_MAX_LAYERS 16
#define XO_MAX_MATERIALS 32
#define XO_BATCH_MAX_LAYERS 8
/* Bytes of the static device array behind xo_batch_run_noalloc (per GPU). */
#define XO_NOALLOC_SCRATCH_BYTES (16u << 20)

/* Scattering phase function kinds (one kind per simulation, as
 * xopto/mcml/mclayer/layer.py:862-872 requires). */
enum xo_pf_kind {
    XO_PF_HG = 0,   /* xopto/mcbase/mcpf/hg.py:70-97   */
    XO_PF_MHG = 1,  /* xopto/mcbase/mcpf/mhg.py:76-110 */
    XO_PF_LUT = 2   /* xopto/mcbase/mcpf/lut.py:111-150 */
};

/* Surface detector kinds. */
enum xo_det_kind {
    XO_DET_NONE = 0,
    XO_DET_TOTAL = 1,  /* xopto/mcml/mcdetector/total.py:86-108   */
    XO_DET_RADIAL = 2  /* xopto/mcml/mcdetector/radial.py:117-150 */
};

/* Photon packet source kinds. */
enum xo_src_kind {
    XO_SRC_LINE = 0,     /* xopto/mcml/mcsource/line.py:95-110, mcvox/mcsource/line.py:84-104 */
    XO_SRC_GAUSSIAN = 1, /* xopto/mcvox/mcsource/gaussianbeam.py:112-189 */
    XO_SRC_FIBER_NI = 2, /* xopto/mcml/mcsource/fiberni.py:237-290 (layered media) */
    XO_SRC_FIBER = 3     /* xopto/mcml/mcsource/fiber.py:252-330 (layered media)   */
};

/* Stepping method (xopto/mcbase/mcoptions.py:269-318; the branches of
 * mcml.template.c:529-755, mcvox.template.c:763-977). Layered, voxelized
 * and batched runs take all three. */
enum xo_method {
    XO_METHOD_AW = 0,  /* albedo weight                 */
    XO_METHOD_AR = 1,  /* albedo rejection              */
    XO_METHOD_MBL = 2  /* microscopic Beer-Lambert      */
};

/* Fluence accumulator geometry (xo_fluence.enabled). */
enum xo_fluence_kind {
    XO_FLUENCE_NONE = 0,
    XO_FLUENCE_XYZ = 1, /* Fluence, xopto/mcbase/mcfluence/fluence.py       */
    XO_FLUENCE_RZ = 2   /* FluenceRz, xopto/mcbase/mcfluence/fluencerz.py   */
};

/* Photon packet trace options (MC_USE_TRACE; xopto/mcbase/mctrace.py:444-455):
 * which states of a packet are logged. */
enum xo_trace_options {
    XO_TRACE_NONE = 0,
    XO_TRACE_START = 1, /* the state right after the launch                */
    XO_TRACE_END = 2,   /* the state when the packet terminates            */
    XO_TRACE_ALL = 7    /* launch, every step and termination              */
};

/* Event flags of one transport step (MC_EVENT_*,
 * xopto/mcbase/kernel/mcbase.template.h:664-681). */
enum xo_event_flags {
    XO_EVENT_REFLECTION = 1u,
    XO_EVENT_REFRACTION = 2u,
    XO_EVENT_BOUNDARY_HIT = 4u,
    XO_EVENT_LAUNCH = 8u,
    XO_EVENT_ABSORPTION = 16u,
    XO_EVENT_SCATTERING = 32u,
    XO_EVENT_TERMINATED = 64u,
    XO_EVENT_ESCAPED = 128u
};

/* Floats per trace entry: x, y, z, px, py, pz, w, pl (TRACE_ENTRY_LEN). */
#define XO_TRACE_ENTRY_LEN 8

/* Run flags. */
enum xo_run_flags {
    /* Lane g runs packets g, g+L, g+2L, ... (L lanes) instead of claiming them
     * from the shared counter: results become a deterministic function of the
     * seeds, the configuration and L. */
    XO_FLAG_STATIC_SCHEDULE = 1u << 0,
    /* Keep detector bins in global memory even if they fit in LDS. */
    XO_FLAG_NO_LDS_BINS = 1u << 1,
    /* Packet p (counted from cfg->first_packet) draws from the generator
     * keyed on (cfg->packet_seed, p) (xo_packet_seed) instead of the lane's
     * stream; the context's lane streams are left untouched. Makes a run a
     * function of the configuration alone, like a batch instance. */
    XO_FLAG_PACKET_KEYED = 1u << 2,
    /* Hardware approximations (v_log_f32, v_exp_f32, v_sin_f32/v_cos_f32,
     * v_sqrt_f32, v_rsq_f32, v_rcp_f32) for the logarithm, exponential,
     * sine/cosine, square root and division of the transport step instead
     * of the fixed-sequence, correctly rounded arithmetic: the build the
     * reference makes with -cl-fast-relaxed-math (xopto/cl/cloptions.py:
     * 137-152; test/performance.py:104). Results then agree with the exact
     * build statistically, not bit for bit. Ignored by traced runs. */
    XO_FLAG_FAST_MATH = 1u << 3
};

/* Phase-function parameters; the union of McPf for Hg, MHg and Lut. */
typedef struct xo_pf {
    float g;          /* Hg, MHg: anisotropy                       */
    float beta;       /* MHg: fraction of the HG term              */
    float a, b, c;    /* Lut: nonlinear index transform parameters */
    uint32_t offset;  /* Lut: first entry in the fp lookup array   */
    uint32_t size;    /* Lut: number of entries                    */
    uint32_t reserved;
} xo_pf; /* 32 B */

/* One layer of the stack; McLayer (xopto/mcml/mclayer/layer.py:40-72). Layer 0
 * and layer num_layers-1 are the surrounding media. */
typedef struct xo_layer {
    float thickness;
    float top;                  /* z of the top surface (-inf for layer 0)    */
    float bottom;               /* z of the bottom surface (+inf for last)    */
    float n;
    float cos_critical_top;
    float cos_critical_bottom;
    float mus;
    float mua;
    float inv_mut;              /* 1/(mua+mus), inf if both are 0             */
    float mua_inv_mut;          /* mua/(mua+mus), 1 if mus is 0               */
    float reserved[2];
    xo_pf pf;
} xo_layer; /* 80 B */

/* Line source; McSource of xopto/mcml/mcsource/line.py:33-64 (the mcvox Line
 * source uses the same packed fields, xopto/mcvox/mcsource/line.py). */
typedef struct xo_source_line {
    float position[3];
    float direction_medium[3];
    float direction_sample[3];
    float direction_reflected[3];
    float reflectance;
    float reserved[3];
} xo_source_line; /* 64 B */

/* Optical fiber in contact with the sample, uniform emission over the core
 * and over the acceptance cone; McSource of UniformFiberNI
 * (xopto/mcml/mcsource/fiberni.py:184-231, packed by cl_pack :341-375). */
typedef struct xo_source_fiber {
    float position[3];          /* core centre (z = 0)                        */
    float radius;               /* core radius                                */
    float cos_min;              /* cosine of the acceptance angle in air:
                                   sqrt(1 - NA^2)                             */
    float n;                    /* refractive index of the core               */
    float reserved[2];
} xo_source_fiber; /* 32 B */

/* Gaussian beam; McSource of xopto/mcvox/mcsource/gaussianbeam.py:33-85. */
typedef struct xo_source_gaussian {
    float transformation[9];    /* row-major 3x3, beam frame -> sample frame */
    float position[3];
    float direction[3];
    float sigma[2];
    float clip;
} xo_source_gaussian; /* 72 B */

/* A surface detector: Radial (radial.py:33-78) or Total (total.py:33-70). */
typedef struct xo_detector {
    int32_t kind;               /* enum xo_det_kind                           */
    float direction[3];
    float position[2];          /* Radial: centre                             */
    float r_min;                /* Radial: (log-)scaled start of the axis     */
    float inv_dr;               /* Radial: 1/step of the (log-)scaled axis    */
    float cos_min;
    uint32_t n;                 /* number of bins (1 for Total)               */
    uint32_t offset;            /* first bin in the accumulator buffer        */
    int32_t log_scale;
} xo_detector; /* 48 B */

/* Fluence accumulator; McFluence of xopto/mcbase/mcfluence/fluence.py:
 * 35-65 (XO_FLUENCE_XYZ: top_left, inv_step and shape per x, y, z) or of
 * fluencerz.py (XO_FLUENCE_RZ: top_left = centre, inv_step[0] = 1/dr,
 * inv_step[2] = 1/dz, shape = {n_r, 1, n_z}). rate: 0 deposits the absorbed
 * weight (mode 'deposition'), 1 the absorbed weight / mua of the medium the
 * packet is in, 0 where mua is 0 (mode 'fluence', MC_FLUENCE_MODE_RATE of
 * fluence.py:102-147 and fluencerz.py:104-160). */
typedef struct xo_fluence {
    int32_t enabled;            /* enum xo_fluence_kind                       */
    float inv_step[3];
    float top_left[3];
    uint32_t shape[3];
    uint32_t offset;
    int32_t k;
    int32_t rate;
} xo_fluence; /* 52 B */

/* Axis of a fiber source that meets the surface at an angle (UniformFiber,
 * xopto/mcml/mcsource/fiber.py:204-227): the rotation of the fiber frame
 * (z along the fiber) onto the sample frame, row major, and the fiber
 * direction (z > 0). Used with the core parameters of xo_source_fiber. */
typedef struct xo_fiber_axis {
    float transformation[9];
    float direction[3];
} xo_fiber_axis; /* 48 B */

/* Everything one layered-media run needs besides the LUT array. Compile-time
 * options of the reference (McMinimumPacketWeight, McPacketLotteryChance,
 * McUseLottery: xopto/mcbase/mcoptions.py:600-670) are runtime fields here. */
typedef struct xo_mcml_cfg {
    uint32_t num_layers;        /* including the two surrounding media (>= 3) */
    int32_t pf_kind;            /* enum xo_pf_kind                            */
    int32_t method;             /* enum xo_method                             */
    float rmax;                 /* simulation radius from the source (m)      */
    float w_min;                /* roulette threshold (1e-4)                  */
    float lottery_chance;       /* roulette survival probability (0.1)        */
    int32_t use_lottery;
    int32_t source_kind;        /* XO_SRC_LINE (source), XO_SRC_FIBER_NI
                                   (fiber) or XO_SRC_FIBER (fiber and
                                   fiber_axis)                                */
    xo_source_line source;
    xo_detector top, bottom, specular;
    xo_layer layers[XO_MAX_LAYERS];
    uint64_t packet_seed;       /* XO_FLAG_PACKET_KEYED only                  */
    uint64_t first_packet;      /* XO_FLAG_PACKET_KEYED only                  */
    xo_source_fiber fiber;
    xo_fluence fluence;         /* bins after every detector bin              */
    xo_fiber_axis fiber_axis;   /* XO_SRC_FIBER only                          */
} xo_mcml_cfg;

/* Photon packet trace of one run; McTrace of xopto/mcbase/mctrace.py:480-510
 * with the compile-time options of Trace.cl_options (:590-595) as fields.
 * Packet p's events go to entries [p][0 .. max_events-1] of the trace data;
 * an event past the last entry overwrites the last one; the event count
 * (which may exceed max_events) goes to the count array. */
typedef struct xo_trace {
    int32_t max_events;         /* entries per packet (>= 1)                  */
    uint32_t options;           /* enum xo_trace_options                      */
    uint32_t event_mask;        /* logged events (use_events only)            */
    int32_t use_events;         /* 0: every event passes the mask             */
    int32_t plon;               /* track the optical path length ('pl')       */
    int32_t reserved[3];
} xo_trace; /* 32 B */

/* Sampling volume accumulator; McSamplingVolume of
 * xopto/mcbase/mcsv.py:34-86. Bins [z][y][x] at accum[offset...]. */
typedef struct xo_sampling_volume {
    float top_left[3];
    float voxel_size[3];
    uint32_t shape[3];          /* x, y, z                                    */
    float multiplier;           /* 1 / smallest voxel size                    */
    uint32_t offset;
    int32_t k;                  /* float to integer weight factor (0x7FFFFF)  */
    int32_t end_field;          /* field of a packet's last trace entry read
                                   as its terminal weight: 7 (pl, what the
                                   reference reads, mcsv.template.c:278) or
                                   6 (w)                                      */
    int32_t reserved[3];
} xo_sampling_volume; /* 64 B */

/* A voxel material; McMaterial of xopto/mcbase/mcmaterial.py:37-63. */
typedef struct xo_material {
    float n, mus, mua, inv_mut, mua_inv_mut;
    float reserved[3];
    xo_pf pf;
} xo_material; /* 64 B */

/* Voxel grid; McVoxelConfig of xopto/mcvox/mcgeometry/voxel.py:96-120. The
 * material index array is int32[z][y][x] (C order). */
typedef struct xo_voxel_cfg {
    float top_left[3];
    float bottom_right[3];
    float size[3];
    int32_t shape[3];
} xo_voxel_cfg; /* 48 B */

typedef struct xo_mcvox_cfg {
    uint32_t num_materials;     /* material 0 is the surrounding medium       */
    int32_t pf_kind;
    int32_t method;
    float rmax;
    float w_min;
    float lottery_chance;
    int32_t use_lottery;
    int32_t source_kind;        /* enum xo_src_kind                           */
    xo_voxel_cfg voxels;
    xo_source_line line;
    xo_source_gaussian gaussian;
    xo_fluence fluence;
    xo_detector top, bottom, specular;
    xo_material materials[XO_MAX_MATERIALS];
    uint64_t packet_seed;       /* XO_FLAG_PACKET_KEYED only                  */
    uint64_t first_packet;      /* XO_FLAG_PACKET_KEYED only                  */
} xo_mcvox_cfg;

/* Run report; mirrors Mc.run_report (xopto/mcml/mc.py:329-332, 1003-1016)
 * plus the step counters the roofline needs. */
typedef struct xo_run_report {
    uint64_t packets;           /* packets launched                           */
    uint64_t steps;             /* iterations of the transport loop           */
    uint64_t boundary_steps;    /* of which ended on a boundary               */
    uint64_t voxel_crossings;   /* mcvox: voxel faces crossed (incl. steps)   */
    uint64_t deposits;          /* fluence deposits issued                    */
    uint64_t truncated;         /* packets stopped at XO_MAX_PACKET_STEPS     */
    uint32_t lanes;             /* lanes launched                             */
    uint32_t threads_used;      /* lanes that ran >= 1 packet (num_kernels)   */
    double kernel_ms;           /* device time of the transport kernel        */
    double total_ms;            /* host wall time of the whole call           */
} xo_run_report;

/* One layered-media simulation of a batch. There is no counterpart struct in
 * the reference, whose batch workflow calls Mc.run once per instance
 * (xopto/mcml/mcrun/helper.py:640-686); each field has the meaning of the
 * xo_mcml_cfg field of the same name. The generator of packet p is derived
 * from (seed, p) (xo_packet_seed), so an instance's result does not depend
 * on the batch it runs in or on the schedule. */
typedef struct xo_batch_instance {
    uint32_t num_layers;
    int32_t pf_kind;
    float rmax;
    uint32_t accum_base;        /* this instance's bins: [base, base+count)  */
    uint32_t accum_count;
    int32_t source_kind;        /* XO_SRC_LINE, _FIBER_NI or _FIBER           */
    uint64_t npackets;
    uint64_t seed;
    xo_source_line source;
    xo_detector top, bottom, specular; /* offsets are absolute              */
    xo_layer layers[XO_BATCH_MAX_LAYERS];
    xo_source_fiber fiber;
    xo_fiber_axis fiber_axis;   /* XO_SRC_FIBER: the fiber's axis            */
    xo_fluence fluence;         /* optional grid; its bins lie inside the
                                   instance's range, after its detector bins */
} xo_batch_instance; /* 1024 B */

/* Launch shape of a batch run (chosen by the autotuner). */
typedef struct xo_batch_opts {
    uint32_t chunk;             /* packets per work-queue item (>= 1)        */
    uint32_t block;             /* threads per block: 64, 128 or 256          */
    uint32_t blocks_per_cu;     /* 0: as many as fit                          */
    uint32_t flags;             /* XO_FLAG_FAST_MATH; other flags are ignored */
    float w_min;                /* roulette threshold for every instance     */
    float lottery_chance;
    int32_t use_lottery;
    uint32_t refill_min;        /* a wave hands out new packets once this
                                   many of its lanes are idle (0: 4)          */
    uint32_t defer_min;         /* a wave handles its boundary events once
                                   this many lanes hold one (0: 16)           */
    int32_t method;             /* enum xo_method of every instance           */
} xo_batch_opts;

/* ---------------- library ---------------- */

/* ABI version of the loaded library. */
int xo_abi_version(void);

/* Message of the last error on this thread ("" if none). */
const char *xo_last_error(void);

/* Number of visible HIP devices (0 if none). Replaces the OpenCL device
 * discovery of xopto/cl/clinfo.py:359-425. */
int xo_device_count(void);

/* sizeof() of each packed struct as the library was built; the analogue of
 * the sizeof_datatypes kernel (xopto/mcml/kernel/mcml.template.c:827-844).
 * Writes up to n values in the order: xo_pf, xo_layer, xo_source_line,
 * xo_source_gaussian, xo_detector, xo_fluence, xo_mcml_cfg, xo_material,
 * xo_voxel_cfg, xo_mcvox_cfg, xo_run_report, xo_batch_instance,
 * xo_batch_opts, xo_serve_opts, xo_serve_stats, xo_source_fiber, xo_trace,
 * xo_sampling_volume, xo_fiber_axis. Returns the count written. */
int xo_struct_sizes(uint32_t *out, int n);

/* ---------------- seeds (host) ---------------- */

/* The first n multipliers a of the safe-prime table: a counts down from
 * XO_FIRST_PRIME_A and is kept when a*2^32-1 and a*2^31-1 are both prime.
 * Regenerates xopto/data/primes/safeprimes_base32_500k.npz[:,0]
 * (xopto/src/rng/find_primes.c:43-93). n <= XO_NUM_PRIMES. */
int xo_safeprime_multipliers(uint32_t n, uint32_t *a_out);

/* Install the first n multipliers of the table computed earlier (for
 * example read back from a cache file) so that later calls do not search
 * again. a[0] must be XO_FIRST_PRIME_A. */
int xo_safeprime_preload(const uint32_t *a, uint32_t n);

/* MWC seeds for n streams: replaces init_RNG (xopto/src/rng/rng.cpp:64-103)
 * called by Random.seeds (xopto/cl/clrng.py:464-530) with the built-in
 * table. Returns 1 (and leaves x, a untouched) if xinit is not a valid
 * initializer, exactly as init_RNG does. n <= XO_MAX_SEEDS. */
int xo_init_rng(uint32_t n, uint64_t xinit, uint64_t *x_out, uint32_t *a_out);

/* Uniform floats from one MWC stream, as the RngKernel test harness
 * (xopto/mcbase/kernel/mcbase.template.c:1640-1646) computes them, on the
 * host. Advances *x. */
int xo_rng_host(uint64_t *x, uint32_t a, uint32_t n, float *out);

/* ---------------- simulation contexts (device) ---------------- */

typedef struct xo_ctx xo_ctx;

/* A context owns the device-resident state of one simulator object (Mc):
 * max_lanes MWC streams seeded by xinit on device `device`, the accumulator,
 * LUT and voxel buffers, and one HIP stream. Replaces ClWorker +
 * ClWorkerRngMixin (xopto/mcbase/mcworker.py:46-898, 1191-1248).
 * max_lanes = 0 selects XO_MAX_SEEDS. Returns NULL on error. */
xo_ctx *xo_ctx_create(int device, uint32_t max_lanes, uint64_t xinit);
void xo_ctx_destroy(xo_ctx *ctx);

/* Lanes the device runs concurrently for the transport kernels (the
 * persistent grid size) and the context's seed capacity. */
int xo_ctx_info(xo_ctx *ctx, uint32_t *persistent_lanes, uint32_t *max_lanes,
                uint32_t *compute_units);

/* Copy the mutable RNG state x of the first n lanes to the host
 * (Mc.run(copyseeds=True), xopto/mcml/mc.py:962-966) or overwrite it. */
int xo_ctx_get_rng(xo_ctx *ctx, uint32_t n, uint64_t *x_out, uint32_t *a_out);
int xo_ctx_set_rng(xo_ctx *ctx, uint32_t n, const uint64_t *x, const uint32_t *a);

/* Run `npackets` packets through a layered medium: the McKernel enqueue of
 * Mc.run (xopto/mcml/mc.py:922-953; kernel mcml.template.c:346-824).
 * lut/nlut: the fp lookup-table array (Lut phase functions), may be NULL.
 * lanes: 0 = the persistent grid, or one lane per 5 packets (at least
 * 65,536) for a run of fewer than 5 packets per lane of that grid, capped
 * by max_lanes; else the lane count
 * (maxthreads of Mc.run). accum/naccum: host output, zeroed and filled with
 * the integer accumulators (detector bins at their offsets). The RNG state
 * of the used lanes persists in the context across runs. */
int xo_mcml_run(xo_ctx *ctx, const xo_mcml_cfg *cfg,
                const float *lut, uint32_t nlut,
                uint64_t npackets, uint32_t lanes, uint32_t flags,
                uint64_t *accum, uint64_t naccum, xo_run_report *report);

/* xo_mcml_run with the photon packet trace (the MC_USE_TRACE build of
 * McKernel, xopto/mcml/kernel/mcml.template.c:513-520, 764-810, and
 * mcsim_trace_event / mcsim_trace_complete of xopto/mcbase/mctrace.py:
 * 519-566). trace_data: host output of npackets * max_events *
 * XO_TRACE_ENTRY_LEN floats; trace_n: host output of npackets event counts.
 * Row p holds packet p (the packet index of the run). */
int xo_mcml_run_trace(xo_ctx *ctx, const xo_mcml_cfg *cfg,
                      const float *lut, uint32_t nlut,
                      uint64_t npackets, uint32_t lanes, uint32_t flags,
                      const xo_trace *trace, float *trace_data, int32_t *trace_n,
                      uint64_t *accum, uint64_t naccum, xo_run_report *report);

/* Sampling volume of a trace: the SamplingVolume kernel of
 * xopto/mcbase/kernel/mcsv.template.c:236-420 called by Mc.sampling_volume
 * (xopto/mcml/mc.py:1040-1200). Every packet's path between consecutive
 * trace events is walked through the voxel grid; each voxel it crosses gets
 * round(terminal weight * path length in the voxel * multiplier * k). A
 * packet with no event is skipped. accum/naccum: host output (bins at
 * sv->offset); total_weight: sum of round(terminal weight * k). */
int xo_sampling_volume_run(xo_ctx *ctx, const xo_sampling_volume *sv,
                           int32_t max_events, const float *trace_data,
                           const int32_t *trace_n, uint64_t npackets,
                           uint64_t *accum, uint64_t naccum,
                           uint64_t *total_weight, xo_run_report *report);

/* Voxelized medium: McKernel of xopto/mcvox/kernel/mcvox.template.c:548-1048
 * called by xopto/mcvox/mc.py:772-1068. voxels is int32[shape z][y][x]; pass
 * NULL to reuse the grid uploaded by the previous call (the
 * Voxels.update_required() path, xopto/mcvox/mc.py:895-899). */
int xo_mcvox_run(xo_ctx *ctx, const xo_mcvox_cfg *cfg,
                 const int32_t *voxels, uint64_t nvoxels,
                 const float *lut, uint32_t nlut,
                 uint64_t npackets, uint32_t lanes, uint32_t flags,
                 uint64_t *accum, uint64_t naccum, xo_run_report *report);

/* xo_mcvox_run with the photon packet trace (the MC_USE_TRACE build of the
 * voxel McKernel, xopto/mcvox/kernel/mcvox.template.c:750-756, 986-1034);
 * trace buffers as for xo_mcml_run_trace. */
int xo_mcvox_run_trace(xo_ctx *ctx, const xo_mcvox_cfg *cfg,
                       const int32_t *voxels, uint64_t nvoxels,
                       const float *lut, uint32_t nlut,
                       uint64_t npackets, uint32_t lanes, uint32_t flags,
                       const xo_trace *trace, float *trace_data, int32_t *trace_n,
                       uint64_t *accum, uint64_t naccum, xo_run_report *report);

/* Run a batch of layered-media instances in one persistent launch: a device
 * work queue of (instance, chunk of packets) items; lanes take packets of
 * whichever item their wave holds and the wave takes the next item as soon
 * as the current one has no unclaimed packet left. Replaces the sequential
 * one-launch-per-instance loop of McRunHelper.run_batch
 * (xopto/mcml/mcrun/helper.py:640-686). accum/naccum: host output holding
 * every instance's bins; report->packets counts all instances. */
int xo_batch_run(xo_ctx *ctx, const xo_batch_instance *instances,
                 uint32_t ninstances, const float *lut, uint32_t nlut,
                 const xo_batch_opts *opts, uint64_t *accum, uint64_t naccum,
                 xo_run_report *report);

/* A batch uploaded once and launched many times, for a fixed-shape batch
 * run repeatedly (benchmarks, parameter sweeps that only change what the
 * host does with the results). xo_batch_plan_create validates and packs
 * the instances as xo_batch_run does, uploads the tables and the LUT into
 * device buffers the plan owns and captures the launch (accumulator and
 * counter clears by fill kernels, then the batch kernel) as a hipGraph;
 * xo_batch_plan_run replays it and copies the naccum accumulators to the
 * host: the integers of xo_batch_run on the same arguments. The reference
 * re-uploads every instance on every Mc.run of its batch loop
 * (xopto/mcml/mcrun/helper.py:640-686); no counterpart entry point.
 * ctx must outlive the plan; report->kernel_ms covers the replayed graph.
 * Returns NULL on error. */
typedef struct xo_batch_plan xo_batch_plan;
xo_batch_plan *xo_batch_plan_create(xo_ctx *ctx, const xo_batch_instance *instances,
                                    uint32_t ninstances, const float *lut, uint32_t nlut,
                                    const xo_batch_opts *opts, uint64_t naccum);
int xo_batch_plan_run(xo_batch_plan *plan, uint64_t *accum, xo_run_report *report);
void xo_batch_plan_destroy(xo_batch_plan *plan);

/* The MWC state (x, a) of packet p of a batch instance with key seed:
 * a = table multiplier (rows 1..XO_MAX_SEEDS) chosen by a hash of
 * (seed, p), x a valid state (0 <= c < a-1, x < 2^32-1) from a second hash.
 * Host copy of the device function, for checks. */
int xo_packet_seed(uint64_t seed, uint64_t p, uint64_t *x, uint32_t *a);

/* xo_batch_run that allocates no device memory. It has no counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * lut_dev: device array of nlut floats (NULL if nlut is 0); accum_dev:
 * device array of naccum u64, zeroed and filled by the call (the same
 * integers xo_batch_run returns). The seed table, the instance tables, the
 * work items and the counters live in a static device array of
 * XO_NOALLOC_SCRATCH_BYTES per GPU (about 2 MiB of it is the seed table,
 * written at the first call on a device); a batch whose tables do not fit
 * is rejected with an error (about 4096 instances at 8192-packet chunks
 * fit). Synchronous on ctx's stream. Calls that overlap on one GPU, from any
 * context, race on that array. */
int xo_batch_run_noalloc(xo_ctx *ctx, const xo_batch_instance *instances,
                         uint32_t ninstances, const float *lut_dev, uint32_t nlut,
                         const xo_batch_opts *opts, uint64_t *accum_dev, uint64_t naccum,
                         xo_run_report *report);

/* ---------------- serving (device) ---------------- */

/* A server keeps simulation requests resident on one device and runs them
 * by continuous batching. No counterpart in the reference, which runs one
 * Mc.run at a time per OpenCL queue (xopto/mcml/mc.py:730-1018); every
 * request's integers equal those of the corresponding xo_batch_run /
 * keyed xo_mcvox_run, whatever else is served with it. */
typedef struct xo_server xo_server;

typedef struct xo_serve_opts {
    uint64_t accum_capacity;    /* u64 accumulator bins of all requests      */
    uint64_t lut_capacity;      /* LUT floats of all requests                */
    uint64_t voxel_capacity;    /* bytes of voxel grids (one byte per voxel,
                                   grids padded to multiples of 4 per axis)  */
    uint64_t tick_packets;      /* layered packets per tick, shared fairly   */
    uint64_t vox_tick_packets;  /* packets per voxel request per tick        */
    uint32_t instance_slots;    /* layered requests resident at once         */
    uint32_t max_items;         /* work-queue items per tick                 */
    uint32_t vox_lanes;         /* lane cap of a voxel launch                */
    uint32_t chunk;             /* at most this many packets per work-queue
                                   item (fewer when a tick would otherwise
                                   leave waves of the grid without an item)  */
    uint32_t block;             /* threads per block of the batch kernel     */
    uint32_t blocks_per_cu;     /* 0: as many as fit                          */
    float w_min;                /* roulette threshold (layered requests)     */
    float lottery_chance;
    int32_t use_lottery;
    int32_t use_graph;          /* replay a captured graph per tick          */
    int32_t fast_math;          /* 1: the XO_FLAG_FAST_MATH arithmetic       */
    int32_t reserved;
} xo_serve_opts; /* 88 B */

typedef struct xo_serve_stats {
    uint64_t admitted, completed, cancelled;
    uint64_t ticks;             /* ticks retired                             */
    uint64_t packets;           /* packets of retired ticks                  */
    uint64_t steps;             /* transport iterations of retired ticks     */
    uint64_t accum_used, lut_used, voxel_used; /* arena occupancy now        */
    double busy_ms;             /* device time of the layered launches       */
} xo_serve_stats; /* 80 B */

void xo_serve_default_opts(xo_serve_opts *opts);

/* Allocate the arenas and tick buffers on ctx's device (opts NULL: the
 * defaults). The server uses ctx's seed table; ctx must outlive it.
 * Returns NULL on error. */
xo_server *xo_serve_create(xo_ctx *ctx, const xo_serve_opts *opts);
void xo_serve_destroy(xo_server *server);

/* Admit a layered request (detector offsets relative to inst->accum_base,
 * accum_count bins; lut/nlut its own LUT array). Returns 0 and the request
 * id, 1 when the arenas or slots are full right now (retry after results
 * are taken), negative on invalid input or on a request larger than the
 * server's arenas. */
int xo_serve_admit_mcml(xo_server *server, const xo_batch_instance *inst,
                        const float *lut, uint32_t nlut, uint64_t *id);

/* Admit a voxelized request: cfg as for xo_mcvox_run, voxels int32[z][y][x],
 * naccum accumulator bins, npackets packets keyed by seed. Returns as
 * xo_serve_admit_mcml. */
int xo_serve_admit_mcvox(xo_server *server, const xo_mcvox_cfg *cfg,
                         const int32_t *voxels, uint64_t nvoxels, const float *lut,
                         uint32_t nlut, uint64_t naccum, uint64_t npackets,
                         uint64_t seed, uint64_t *id);

/* Retire finished ticks and queue new ones (at most two in flight);
 * never blocks on the device. *active: requests not yet complete. */
int xo_serve_step(xo_server *server, uint32_t *active);

/* Block until every queued tick has finished, then retire them. */
int xo_serve_wait(xo_server *server);

/* Ids of requests completed since the last poll (up to max). */
int xo_serve_poll(xo_server *server, uint64_t *ids, uint32_t max, uint32_t *n);

/* Copy a completed request's naccum bins to the host and release it. */
int xo_serve_result(xo_server *server, uint64_t id, uint64_t *accum, uint64_t naccum);

/* Drop a request; its memory is released once no tick holds its work. */
int xo_serve_cancel(xo_server *server, uint64_t id);

int xo_serve_get_stats(xo_server *server, xo_serve_stats *stats);

/* ---------------- measured ceilings (device) ---------------- */

/* Device-to-device copy of `bytes` (16-byte vectors, one grid-stride
 * kernel), reps times after one warm-up: average ms per copy. The HBM
 * bandwidth the memory-bound figures are compared against. No counterpart
 * in the reference. */
int xo_probe_copy(xo_ctx *ctx, uint64_t bytes, uint32_t reps, double *ms_per_copy);

/* natomics u64 atomic adds of 1 at hashed random addresses of a zeroed
 * array of array_bytes: kernel ms, and the array's sum (== natomics when
 * every atomic landed). The HBM/L2 atomic rate the deposit accumulators
 * are compared against. No counterpart in the reference. */
int xo_probe_atomic_u64(xo_ctx *ctx, uint64_t array_bytes, uint64_t natomics,
                        uint64_t seed, double *ms, uint64_t *checksum);

/* xo_probe_atomic_u64 with the targets among the first hot_bytes of the
 * array and, when per_xcd, eight arrays (one per XCD): each wave adds into
 * the array of the XCD it runs on with workgroup-scope atomics. checksum is
 * the sum over every array. Whether the atomic rate depends on the
 * footprint or on cross-XCD coherence, which decides how the fluence
 * accumulators are laid out. No counterpart in the reference. */
int xo_probe_atomic_u64_xcd(xo_ctx *ctx, uint64_t array_bytes, uint64_t hot_bytes,
                            uint64_t natomics, uint64_t seed, int per_xcd, double *ms,
                            uint64_t *checksum);

/* Eight independent fp32 fused multiply-add chains per lane, iters rounds,
 * 32 blocks of 256 threads per CU: kernel ms and the lane-operations done.
 * The VALU issue ceiling the transport kernels are compared against. No
 * counterpart in the reference. */
int xo_probe_valu(xo_ctx *ctx, uint32_t iters, double *ms, double *lane_ops);

/* Check of the transport kernels' range-restricted correctly rounded square
 * root (every float +0 or in [2^-96, 2^96)) and division (ndiv random
 * operand pairs with exponents in [-60, 59]) against the general correctly
 * rounded operations, on the device: the count of square roots checked and
 * the mismatches of each. No counterpart in the reference. */
int xo_probe_cr_math(xo_ctx *ctx, uint64_t ndiv, uint64_t seed, uint64_t *sqrt_checked,
                     uint64_t *sqrt_mismatches, uint64_t *div_mismatches);

/* Functions of the transport kernels' arithmetic, for xo_probe_math. */
enum xo_math_fn {
    XO_MATH_LOG = 0,        /* natural logarithm                          */
    XO_MATH_SIN_TURNS = 1,  /* sin(2 pi x)                                */
    XO_MATH_COS_TURNS = 2,  /* cos(2 pi x)                                */
    XO_MATH_SQRT = 3,
    XO_MATH_RCP = 4,        /* 1 / x (the division of the kernels)        */
    XO_MATH_RSQRT = 5,      /* 1 / sqrt(x) (direction renormalization)    */
    XO_MATH_CBRT = 6,
    XO_MATH_EXP = 7,
    XO_MATH_COUNT = 8
};

/* out[i] = f(in[i]) on the device for the function `which` (enum
 * xo_math_fn) as the transport kernels compute it: fast = 0 the exact
 * arithmetic, fast = 1 the XO_FLAG_FAST_MATH arithmetic. The accuracy
 * check of the fast build. No counterpart in the reference. */
int xo_probe_math(xo_ctx *ctx, int fast, int which, uint32_t n, const float *in,
                  float *out);

/* ---------------- reflectance surrogate (device) ---------------- */

/* A fully connected ReLU network (widths dims[0..nlayers]) trained and
 * evaluated with bf16 MFMA products and fp32 accumulation; Adam on fp32
 * master weights. No counterpart in the reference, whose regression
 * utilities call TensorFlow (xopto/util/regression/ann.py). Parameters
 * are packed W_1, b_1, ..., W_n, b_n with W_l row-major [in][out]. */
typedef struct xo_mlp xo_mlp;

xo_mlp *xo_mlp_create(xo_ctx *ctx, const uint32_t *dims, uint32_t nlayers,
                      uint32_t max_batch);
void xo_mlp_destroy(xo_mlp *mlp);
uint64_t xo_mlp_param_count(xo_mlp *mlp);
/* Set (and reset the optimizer state) / read the fp32 parameters. */
int xo_mlp_set_params(xo_mlp *mlp, const float *params, uint64_t n);
int xo_mlp_get_params(xo_mlp *mlp, float *params, uint64_t n);
/* y[n][dims[nlayers]] = network(x[n][dims[0]]); kernel_ms may be NULL. */
int xo_mlp_forward(xo_mlp *mlp, const float *x, uint64_t n, float *y, double *kernel_ms);
/* Mean squared error of one batch (n <= max_batch rows) and its gradient
 * with respect to every parameter (packed like the parameters). */
int xo_mlp_gradient(xo_mlp *mlp, const float *x, const float *t, uint32_t n,
                    float *grad, float *loss);
/* Keep a training set on the device. */
int xo_mlp_set_data(xo_mlp *mlp, const float *x, const float *t, uint32_t n);
/* `steps` Adam steps (beta 0.9, 0.999, eps 1e-8) on minibatches of `batch`
 * rows drawn with replacement by a hash of (seed, step, row); when
 * use_graph, captured graphs of 64 steps and one of the remaining
 * steps % 64. losses[s]: the minibatch loss of step s. */
int xo_mlp_train(xo_mlp *mlp, uint32_t steps, uint32_t batch, float lr, uint64_t seed,
                 int use_graph, float *losses, double *kernel_ms);
/* C = op(A) op(B) with A and B rounded to bf16, fp32 accumulation: the
 * product the network runs on, exposed for tests. */
int xo_gemm_bf16(xo_ctx *ctx, int trans_a, int trans_b, uint32_t M, uint32_t N,
                 uint32_t K, const float *A, uint32_t lda, const float *B, uint32_t ldb,
                 float *C, uint32_t ldc);

#ifdef __cplusplus
}
#endif

#endif /* XOMC_H */
