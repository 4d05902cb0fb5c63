/*
 * xo_kernels.h — launch parameters of the transport kernels (host and
 * device view; plain data, passed by value as the single kernel argument).
 */
#ifndef XO_KERNELS_H
#define XO_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/xomc.h"

/* Threads per block: 4 wave64 wavefronts. */
#define XO_BLOCK 256

/* Per-run counters accumulated by the kernels. */
enum {
    XO_STAT_STEPS = 0,     /* transport-loop iterations                  */
    XO_STAT_BOUNDARY = 1,  /* iterations that ended on a boundary        */
    XO_STAT_PACKETS = 2,   /* packets launched                           */
    XO_STAT_THREADS = 3,   /* lanes that launched at least one packet    */
    XO_STAT_CROSSINGS = 4, /* mcvox: voxel faces crossed                 */
    XO_STAT_DEPOSITS = 5,  /* fluence atomics issued                     */
    XO_STAT_TRUNCATED = 6, /* packets stopped at XO_MAX_PACKET_STEPS     */
    XO_STAT_COUNT = 7
};

/* Largest detector-bin count privatized in LDS, and largest LUT (floats)
 * staged in LDS; beyond these the kernels use global memory. */
#define XO_LDS_BINS_MAX 1024u
#define XO_LDS_LUT_MAX 4096u

namespace xo {

struct McmlParams {
    xo_mcml_cfg cfg;
    xo_trace trace;        /* max_events 0: no trace                      */
    float *trace_data;     /* [packet][max_events][XO_TRACE_ENTRY_LEN]    */
    int32_t *trace_n;      /* [packet] event counts                       */
    const uint32_t *table; /* seed-table multipliers (XO_FLAG_PACKET_KEYED) */
    const float *lut;
    uint64_t *rng_x;
    const uint32_t *rng_a;
    unsigned long long *counter;
    unsigned long long *accum;
    unsigned long long *stats;
    uint64_t npackets;
    uint32_t lanes;
    uint32_t flags;
    uint32_t lds_bins;
    uint32_t lds_lut;
};

struct McvoxParams {
    xo_mcvox_cfg cfg;
    xo_trace trace;        /* max_events 0: no trace                      */
    float *trace_data;     /* [packet][max_events][XO_TRACE_ENTRY_LEN]    */
    int32_t *trace_n;      /* [packet] event counts                       */
    const uint32_t *table; /* seed-table multipliers (XO_FLAG_PACKET_KEYED) */
    const uint8_t *voxels; /* material index per voxel, [z][y][x] */
    const float *lut;
    uint64_t *rng_x;
    const uint32_t *rng_a;
    unsigned long long *counter;
    unsigned long long *accum;
    unsigned long long *stats;
    uint64_t npackets;
    uint32_t lanes;
    uint32_t flags;
    uint32_t lds_bins;
    uint32_t lds_lut;
    /* Cartesian fluence bins of one kFlBox-sized box of the grid (origin
     * fl_box, in bins) accumulate in LDS and are added to accum once per
     * block: the box around the beam's entry takes a large share of the
     * deposits. fl_box_on 0: every deposit goes to accum. */
    int32_t fl_box[3];
    uint32_t fl_box_on;
};

/* Extent of the LDS fluence box (x, y, z bins). */
constexpr int32_t kFlBoxX = 16, kFlBoxY = 16, kFlBoxZ = 8;
constexpr uint32_t kFlBoxBins = kFlBoxX * kFlBoxY * kFlBoxZ;

size_t mcml_lds_bytes(const McmlParams &P);
size_t mcvox_lds_bytes(const McvoxParams &P);
hipError_t launch_mcml(const McmlParams &P, uint32_t blocks, hipStream_t stream);
hipError_t launch_mcvox(const McvoxParams &P, uint32_t blocks, hipStream_t stream);
hipError_t occupancy_mcml(int pf_kind, size_t lds, bool fast, int *blocks_per_cu);
hipError_t occupancy_mcvox(int pf_kind, size_t lds, bool fast, int *blocks_per_cu);

} // namespace xo

#endif /* XO_KERNELS_H */
