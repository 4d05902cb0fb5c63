/*
 * xo_oracle.c — CPU oracle for the photon-packet transport path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * as the timed CPU baseline. The product (pyxopto_amd) never links, loads or
 * falls back to it.
 *
 * A plain C (C11 + OpenMP) restatement of PyXOpto's McKernel for layered
 * media (xopto/mcml/kernel/mcml.template.c:346-824) and voxelized media
 * (xopto/mcvox/kernel/mcvox.template.c:548-1048), with the plugin snippets
 * the configurations use, each function citing the reference lines it
 * follows. Packet state is binary32, accumulators are exact uint64 sums.
 *
 * keyed_table (may be NULL): with the seed-table multipliers of rows
 * 1..499999, packet p of a run draws from xo_packet_state(cfg->packet_seed,
 * cfg->first_packet + p) instead of its lane's stream (the engine's
 * XO_FLAG_PACKET_KEYED runs).
 *
 * Deliberate, documented differences from the reference text:
 *  - packets are assigned statically (lane g runs packets g, g+L, ...) instead
 *    of through the global atomic counter; every lane still owns one MWC
 *    stream, as in the reference;
 *  - float -> int conversions saturate (the reference's convert_int leaves
 *    out-of-range values implementation defined) and the Lut index is clamped
 *    to the table (the reference reads out of bounds when rounding puts the
 *    index at -1);
 *  - the planar step test and the move use fused multiply-add (the reference
 *    is built with -cl-mad-enable in its validation and timing runs).
 *
 * math_mode 0 (XO_OMATH_LIBM) evaluates log, sin/cos and cbrt with the C
 * library: an implementation independent of the product, used for the
 * statistical parity tests. math_mode 1 (XO_OMATH_KERNEL) uses the product's
 * fixed-sequence functions (pyxopto_amd/csrc/xo_math.h), which the GPU kernels
 * also use; with the static schedule a GPU run must then reproduce this
 * oracle's accumulators bit for bit.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "../include/xomc.h"
#include "../pyxopto_amd/csrc/xo_math.h"
#include "../pyxopto_amd/csrc/xo_seedhash.h"

#define XO_OMATH_LIBM 0
#define XO_OMATH_KERNEL 1

#define O_FP_COS_0 1.0f   /* FP_COS_0 = 1 - cos(90 deg), mcbase.template.h:512 */
#define O_FP_COS_90 0.0f
#define O_FP_RMIN 1e-12f  /* mcbase.template.h:476 */
#define O_2PI 6.283185307179586f

/* ------------------------------------------------------------------ */
/* math dispatch                                                      */
/* ------------------------------------------------------------------ */
static inline float o_log(float x, int mm) {
    return (mm == XO_OMATH_KERNEL) ? xo_logf(x) : logf(x);
}

/* sin/cos of the azimuth 2*pi*xi (the reference's FP_2PI*random then
 * sincos, hg.py:87, gaussianbeam.py:135) */
static inline void o_sincos_azimuth(float xi, float *s, float *c, int mm) {
    if (mm == XO_OMATH_KERNEL) {
        xo_sincos_turns(xi, s, c);
    } else {
        float fi = O_2PI * xi;
        *s = sinf(fi);
        *c = cosf(fi);
    }
}

static inline float o_cbrt(float x, int mm) {
    return (mm == XO_OMATH_KERNEL) ? xo_cbrtf(x) : cbrtf(x);
}

/* convert_int with saturation (mc_int, mcbase.template.h:620) */
static inline int32_t o_int(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483520.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}

/* convert_uint with saturation (mc_uint) */
static inline uint32_t o_uint(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967040.0f) return UINT32_MAX;
    return (uint32_t)x;
}

static inline int32_t o_clip(int32_t v, int32_t lo, int32_t hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

/* ------------------------------------------------------------------ */
/* RNG: fp_random_single, mcbase.template.c:1576-1594                  */
/* ------------------------------------------------------------------ */
static inline float o_rand(uint64_t *x, uint32_t a) {
    *x = (*x & 0xFFFFFFFFull) * (uint64_t)a + (*x >> 32);
    /* (float)(uint32)x / (float)0xFFFFFFFF; the divisor rounds to 2^32 */
    return (float)(uint32_t)(*x) / 4294967296.0f;
}

/* RngKernel (mcbase.template.c:1640-1646): n floats from one stream. */
int oracle_rng(uint64_t *x, uint32_t a, uint32_t n, float *out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = o_rand(x, a);
    return 0;
}

/* ------------------------------------------------------------------ */
/* accumulators                                                       */
/* ------------------------------------------------------------------ */
/* accu_64_deposit_32 (mcbase.template.c:58-61): an exact modular u64 add */
static inline void o_accu(uint64_t *acc, uint64_t index, uint32_t w) {
    __atomic_fetch_add(&acc[index], (uint64_t)w, __ATOMIC_RELAXED);
}

/* Radial (radial.py:117-150) and Total (total.py:86-108) deposits. */
static void o_detector_deposit(const xo_detector *d, uint64_t *acc,
                               float px, float py, float dx, float dy,
                               float dz, float w, int mm) {
    if (d->kind == XO_DET_NONE) return;
    uint32_t index = 0;
    if (d->kind == XO_DET_RADIAL) {
        float ddx = px - d->position[0];
        float ddy = py - d->position[1];
        float r = sqrtf(ddx * ddx + ddy * ddy);
        if (d->log_scale) r = o_log(fmaxf(r, O_FP_RMIN), mm);
        int32_t r_index = o_int((r - d->r_min) * d->inv_dr);
        index = (uint32_t)o_clip(r_index, 0, (int32_t)d->n - 1);
    }
    float cosv = fabsf(dx * d->direction[0] + dy * d->direction[1] +
                       dz * d->direction[2]);
    /* uint32_t ui32w = weight_to_int(weight)*(cos_min <= |dot|) */
    float wi = (w * (float)XO_ACCU_K + 0.5f) * (float)(d->cos_min <= cosv);
    uint32_t ui = o_uint(wi);
    if (ui > 0u) o_accu(acc, (uint64_t)d->offset + index, ui);
}

/* ------------------------------------------------------------------ */
/* scattering                                                         */
/* ------------------------------------------------------------------ */
/* mcsim_pf_sample_angles: returns cos(theta), azimuth as a fraction of a
 * turn in *xi_fi. HG hg.py:80-96, MHg mhg.py:87-109, Lut lut.py:125-150. */
static float o_pf_sample(int kind, const xo_pf *pf, const float *lut,
                         uint64_t *x, uint32_t a, float *xi_fi, int mm) {
    float cos_theta;
    *xi_fi = o_rand(x, a);
    if (kind == XO_PF_HG) {
        float g = pf->g;
        float k = (1.0f - g * g) / (1.0f + g * (2.0f * o_rand(x, a) - 1.0f));
        cos_theta = (1.0f + g * g - k * k) / (2.0f * g);
        if (g == 0.0f) cos_theta = 1.0f - 2.0f * o_rand(x, a);
        return fmaxf(fminf(cos_theta, 1.0f), -1.0f);
    } else if (kind == XO_PF_MHG) {
        float g = pf->g;
        if (o_rand(x, a) <= pf->beta) {
            float k = (1.0f - g * g) /
                      (1.0f + g * (2.0f * o_rand(x, a) - 1.0f));
            cos_theta = (1.0f + g * g - k * k) / (2.0f * g);
            if (g == 0.0f) cos_theta = 1.0f - 2.0f * o_rand(x, a);
        } else {
            cos_theta = o_cbrt(2.0f * o_rand(x, a) - 1.0f, mm);
        }
        /* mc_fclip = clamp(x, -1, 1) */
        return fminf(fmaxf(cos_theta, -1.0f), 1.0f);
    } else { /* XO_PF_LUT */
        uint32_t lut_size = pf->size - 1u;
        float fp_index = (pf->a / (o_rand(x, a) - pf->c) - pf->b + 1.0f) *
                         (float)lut_size * 0.5f;
        fp_index = fminf(fmaxf(fp_index, 0.0f), (float)lut_size);
        float fl = floorf(fp_index);
        float d = fp_index - fl;
        uint32_t index = (uint32_t)o_int(fl);
        uint32_t index2 = index + 1u < lut_size ? index + 1u : lut_size;
        return lut[pf->offset + index] * (1.0f - d) +
               lut[pf->offset + index2] * d;
    }
}

/* scatter_direction, mcbase.template.c:1522-1556 (+ mc_normalize_fpv3,
 * :975-981, applied in single precision) */
static void o_scatter_direction(float *dx, float *dy, float *dz,
                                float cos_theta, float xi_fi, int mm) {
    float sin_fi, cos_fi;
    float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    o_sincos_azimuth(xi_fi, &sin_fi, &cos_fi, mm);
    float stcf = sin_theta * cos_fi;
    float stsf = sin_theta * sin_fi;
    float px = *dx, py = *dy, pz = *dz;
    if (fabsf(pz) >= O_FP_COS_0) {
        *dx = stcf;
        *dy = stsf;
        *dz = copysignf(cos_theta, pz * cos_theta);
    } else {
        float k = sqrtf(1.0f - pz * pz);
        *dx = (stcf * px * pz - stsf * py) / k + px * cos_theta;
        *dy = (stcf * py * pz + stsf * px) / k + py * cos_theta;
        *dz = (-stcf) * k + pz * cos_theta;
    }
    float inv = 1.0f / sqrtf(*dx * *dx + *dy * *dy + *dz * *dz);
    *dx = *dx * inv;
    *dy = *dy * inv;
    *dz = *dz * inv;
}

/* ------------------------------------------------------------------ */
/* boundary physics, mcbase.template.c:1229-1520                       */
/* ------------------------------------------------------------------ */
static inline float o_cos_critical(float n1, float n2) {
    return (n1 > n2) ? sqrtf(1.0f - (n2 * n2) / (n1 * n1)) : 0.0f;
}

static float o_reflectance(float n1, float n2, float cos1, float cc) {
    float R = 1.0f;
    cos1 = fabsf(cos1);
    if (n1 == n2) return 0.0f;
    if (cos1 > cc) {
        float n1_d_n2 = n1 / n2;
        float sin1 = sqrtf(1.0f - cos1 * cos1);
        if (cos1 >= O_FP_COS_0) sin1 = 0.0f;
        float sin2 = fminf(1.0f, n1_d_n2 * sin1);
        float cos2 = sqrtf(1.0f - sin2 * sin2);
        float n_cos1 = n1_d_n2 * cos1;
        float n_cos2 = n1_d_n2 * cos2;
        float Rs = (n_cos1 - cos2) / (n_cos1 + cos2);
        Rs *= Rs;
        float Rp = (n_cos2 - cos1) / (n_cos2 + cos1);
        Rp *= Rp;
        R = 0.5f * (Rp + Rs);
        if (cos1 <= O_FP_COS_90 || sin2 == 1.0f) return 1.0f;
    }
    return R;
}

static inline float o_fsign(float x) { return (x >= 0.0f) ? 1.0f : -1.0f; }

/* refract (in place allowed), mcbase.template.c:1393-1410 */
static void o_refract(float p[3], const float n[3], float n1, float n2) {
    float cos1 = p[0] * n[0] + p[1] * n[1] + p[2] * n[2];
    float n1_d_n2 = n1 / n2;
    float sin2_squared = n1_d_n2 * n1_d_n2 * (1.0f - cos1 * cos1);
    float k = o_fsign(cos1) *
              (n1_d_n2 * fabsf(cos1) - sqrtf(1.0f - sin2_squared));
    p[0] = n1_d_n2 * p[0] - k * n[0];
    p[1] = n1_d_n2 * p[1] - k * n[1];
    p[2] = n1_d_n2 * p[2] - k * n[2];
}

/* refract_safe, mcbase.template.c:1487-1508: 1 on total reflection */
static int o_refract_safe(const float p[3], const float n[3], float n1,
                          float n2, float r[3]) {
    float cos1 = p[0] * n[0] + p[1] * n[1] + p[2] * n[2];
    float n1_d_n2 = n1 / n2;
    float sin2_squared = n1_d_n2 * n1_d_n2 * (1.0f - cos1 * cos1);
    if (sin2_squared > 1.0f) return 1;
    float k = o_fsign(cos1) *
              (n1_d_n2 * fabsf(cos1) - sqrtf(1.0f - sin2_squared));
    r[0] = n1_d_n2 * p[0] - k * n[0];
    r[1] = n1_d_n2 * p[1] - k * n[1];
    r[2] = n1_d_n2 * p[2] - k * n[2];
    return 0;
}

/* reflect (in place allowed), mcbase.template.c:1342-1350 */
static void o_reflect(float p[3], const float n[3]) {
    float p_n_2 = 2.0f * (p[0] * n[0] + p[1] * n[1] + p[2] * n[2]);
    p[0] = p[0] - n[0] * p_n_2;
    p[1] = p[1] - n[1] * p_n_2;
    p[2] = p[2] - n[2] * p_n_2;
}

/* ------------------------------------------------------------------ */
/* packet state                                                        */
/* ------------------------------------------------------------------ */
typedef struct o_pkt {
    float p[3];
    float d[3];
    float w;
    int32_t layer;          /* mcml layer index                         */
    int32_t vi[3];          /* mcvox voxel index                        */
    int32_t mi;             /* mcvox material index                     */
    uint64_t x;
    uint32_t a;
} o_pkt;

typedef struct o_counts {
    uint64_t steps, boundary, crossings, deposits, truncated;
} o_counts;

/* Roulette, mcml.template.c:738-753: returns 1 if the packet ends. */
static inline int o_roulette(o_pkt *s, float w_min, float chance,
                             int use_lottery) {
    if (s->w < w_min) {
        if (use_lottery) {
            if (o_rand(&s->x, s->a) > chance) return 1;
            s->w = s->w / chance;
        } else {
            return 1;
        }
    }
    return 0;
}

/* Photon packet trace (xopto/mcbase/mctrace.py:519-566); NULL when off. */
typedef struct o_trace {
    xo_trace cfg;
    float *data;    /* [packet][max_events][XO_TRACE_ENTRY_LEN] */
    int32_t *n;     /* [packet]                                  */
} o_trace;

typedef struct o_tstate {
    uint64_t pkt;
    uint32_t count;
    uint32_t ev;    /* event flags of the current step           */
    float pl;       /* optical path length                       */
} o_tstate;

/* mcsim_trace_event: log unless the event mask rejects this step; an event
 * past the last entry overwrites it (mctrace.py:527-560). */
static void o_trace_event(const o_trace *t, o_tstate *ts, const o_pkt *s) {
    if (t->cfg.use_events && !(t->cfg.event_mask & ts->ev)) return;
    uint32_t last = (uint32_t)t->cfg.max_events - 1u;
    uint32_t slot = ts->count < last ? ts->count : last;
    float *e = t->data +
               (ts->pkt * (uint64_t)t->cfg.max_events + slot) * XO_TRACE_ENTRY_LEN;
    e[0] = s->p[0];
    e[1] = s->p[1];
    e[2] = s->p[2];
    e[3] = s->d[0];
    e[4] = s->d[1];
    e[5] = s->d[2];
    e[6] = s->w;
    e[7] = t->cfg.plon ? ts->pl : 0.0f;
    ts->count++;
}

/* End of one transport step (mcml.template.c:765-781): termination flag,
 * the logged state, the event count of a finished packet. */
static void o_trace_step(const o_trace *t, o_tstate *ts, const o_pkt *s,
                         int done) {
    if (done) ts->ev |= XO_EVENT_TERMINATED;
    if (t->cfg.options == XO_TRACE_ALL ||
        ((t->cfg.options & XO_TRACE_END) && done))
        o_trace_event(t, ts, s);
    ts->ev = 0;
    if (done) t->n[ts->pkt] = (int32_t)ts->count;
}

/* Fluence deposit (deposition mode): the Cartesian voxel of
 * mcsim_fluence_deposit_at (xopto/mcbase/mcfluence/fluence.py:102-147) or
 * the (r, z) cell of fluencerz.py:112-160. */
static void o_fluence_deposit(const xo_fluence *f, uint64_t *acc,
                              const float p[3], float w, float mua, uint64_t *deposits) {
    uint64_t index;
    if (f->enabled == XO_FLUENCE_XYZ) {
        float fx = (p[0] - f->top_left[0]) * f->inv_step[0];
        float fy = (p[1] - f->top_left[1]) * f->inv_step[1];
        float fz = (p[2] - f->top_left[2]) * f->inv_step[2];
        if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx < (float)f->shape[0] &&
              fy < (float)f->shape[1] && fz < (float)f->shape[2]))
            return;
        uint64_t ix = o_uint(fx), iy = o_uint(fy), iz = o_uint(fz);
        index = (iz * f->shape[1] + iy) * f->shape[0] + ix;
    } else if (f->enabled == XO_FLUENCE_RZ) {
        float dx = p[0] - f->top_left[0];
        float dy = p[1] - f->top_left[1];
        float r = sqrtf(dx * dx + dy * dy);
        float dz = p[2] - f->top_left[2];
        float fr = r * f->inv_step[0];
        float fz = dz * f->inv_step[2];
        if (!(fr >= 0.0f && fz >= 0.0f && fr < (float)f->shape[0] && fz < (float)f->shape[2]))
            return;
        index = (uint64_t)o_uint(fz) * f->shape[0] + o_uint(fr);
    } else {
        return;
    }
    /* fluence-rate mode: weight / mua of the current medium, 0 if mua is 0
     * (xopto/mcbase/mcfluence/fluence.py:139-141) */
    if (f->rate) w *= (mua != 0.0f) ? 1.0f / mua : 0.0f;
    uint32_t ui = o_uint(w * (float)f->k + 0.5f);
    if (ui > 0u) {
        o_accu(acc, f->offset + index, ui);
        (*deposits)++;
    }
}

static inline float o_exp(float x, int mm) {
    return (mm == XO_OMATH_KERNEL) ? xo_expf(x) : expf(x);
}

/* ------------------------------------------------------------------ */
/* mcml                                                               */
/* ------------------------------------------------------------------ */
/* Line source launch, xopto/mcml/mcsource/line.py:95-110 */
static void o_mcml_launch(const xo_mcml_cfg *cfg, o_pkt *s, uint64_t *acc,
                          int mm) {
    const xo_source_line *src = &cfg->source;
    s->w = 1.0f - src->reflectance;
    s->p[0] = src->position[0];
    s->p[1] = src->position[1];
    s->p[2] = src->position[2];
    s->d[0] = src->direction_sample[0];
    s->d[1] = src->direction_sample[1];
    s->d[2] = src->direction_sample[2];
    o_detector_deposit(&cfg->specular, acc, s->p[0], s->p[1],
                       src->direction_reflected[0], src->direction_reflected[1],
                       src->direction_reflected[2], src->reflectance, mm);
    s->layer = 1;
}

/* reflectance_cos2, mcbase.template.c:1303-1346: Fresnel reflectance from
 * the cosine of the transmitted angle. */
static float o_reflectance_cos2(float n1, float n2, float cos2) {
    float R = 1.0f;
    cos2 = fabsf(cos2);
    if (n1 == n2) return 0.0f;
    float sin2 = sqrtf(1.0f - cos2 * cos2);
    if (cos2 >= O_FP_COS_0) sin2 = 0.0f;
    float sin1 = (n2 / n1) * sin2;
    if (sin1 < 1.0f) {
        float cos1 = sqrtf(1.0f - sin1 * sin1);
        float n1_d_n2 = n1 / n2;
        float n_cos1 = n1_d_n2 * cos1;
        float n_cos2 = n1_d_n2 * cos2;
        float Rs = (n_cos1 - cos2) / (n_cos1 + cos2);
        Rs *= Rs;
        float Rp = (n_cos2 - cos1) / (n_cos2 + cos1);
        Rp *= Rp;
        R = 0.5f * (Rp + Rs);
        if (cos1 <= O_FP_COS_90 || sin2 == 1.0f) return 1.0f;
    }
    return R;
}

/* UniformFiberNI launch, xopto/mcml/mcsource/fiberni.py:237-290: a point
 * uniform over the core, cos(theta) uniform over the acceptance cone,
 * refraction into layer 1, weight 1 - R; the specular part leaves along
 * refract((dx, dy, -dz), (0, 0, -1), n_sample, n_core)
 * (mcbase.template.c:1443-1462). */
static void o_fiber_launch(const xo_mcml_cfg *cfg, o_pkt *s, uint64_t *acc,
                           int mm) {
    const xo_source_fiber *src = &cfg->fiber;
    const float n1 = cfg->layers[1].n;
    float sin_fi, cos_fi;
    float r = sqrtf(o_rand(&s->x, s->a)) * src->radius;
    o_sincos_azimuth(o_rand(&s->x, s->a), &sin_fi, &cos_fi, mm);
    s->p[0] = src->position[0] + r * cos_fi;
    s->p[1] = src->position[1] + r * sin_fi;
    s->p[2] = 0.0f;
    o_sincos_azimuth(o_rand(&s->x, s->a), &sin_fi, &cos_fi, mm);
    float cos_theta = 1.0f - o_rand(&s->x, s->a) * (1.0f - src->cos_min);
    float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    sin_theta = sin_theta / n1;
    cos_theta = sqrtf(1.0f - sin_theta * sin_theta);
    s->d[0] = cos_fi * sin_theta;
    s->d[1] = sin_fi * sin_theta;
    s->d[2] = cos_theta;
    float R = o_reflectance_cos2(src->n, n1, s->d[2]);
    s->w = 1.0f - R;
    if (cfg->specular.kind != XO_DET_NONE) {
        float p[3] = {s->d[0], s->d[1], -s->d[2]};
        float nrm[3] = {0.0f, 0.0f, -1.0f};
        float cos1 = p[0] * nrm[0] + p[1] * nrm[1] + p[2] * nrm[2];
        float n1_d_n2 = n1 / src->n;
        float sin2_squared = n1_d_n2 * n1_d_n2 * (1.0f - cos1 * cos1);
        float k = ((cos1 >= 0.0f) ? 1.0f : -1.0f) *
                  (n1_d_n2 * fabsf(cos1) - sqrtf(1.0f - sin2_squared));
        float rd[3];
        for (int i = 0; i < 3; ++i) rd[i] = n1_d_n2 * p[i] - k * nrm[i];
        o_detector_deposit(&cfg->specular, acc, s->p[0], s->p[1], rd[0], rd[1],
                           rd[2], R, mm);
    }
    s->layer = 1;
}

/* UniformFiber launch, xopto/mcml/mcsource/fiber.py:265-330: the core
 * point rotated by the fiber axis and projected along it onto z = 0, the
 * direction rotated likewise. The reference tests the projected point's z
 * (0) against the critical cosine before refracting (:305-309), so the
 * direction is never refracted; the specular part leaves along it. */
static void o_fiber_axis_launch(const xo_mcml_cfg *cfg, o_pkt *s,
                                uint64_t *acc, int mm) {
    const xo_source_fiber *src = &cfg->fiber;
    const float *T = cfg->fiber_axis.transformation;
    const float *ax = cfg->fiber_axis.direction;
    const float n1 = cfg->layers[1].n;
    float sin_fi, cos_fi;
    float r = sqrtf(o_rand(&s->x, s->a)) * src->radius;
    o_sincos_azimuth(o_rand(&s->x, s->a), &sin_fi, &cos_fi, mm);
    float v[3] = {r * cos_fi, r * sin_fi, 0.0f}, m[3];
    for (int i = 0; i < 3; ++i)
        m[i] = T[3 * i] * v[0] + T[3 * i + 1] * v[1] + T[3 * i + 2] * v[2];
    float k = (0.0f - m[2]) / ax[2];
    m[0] += k * ax[0];
    m[1] += k * ax[1];
    s->p[0] = src->position[0] + m[0];
    s->p[1] = src->position[1] + m[1];
    s->p[2] = 0.0f;
    o_sincos_azimuth(o_rand(&s->x, s->a), &sin_fi, &cos_fi, mm);
    float cos_theta = 1.0f - o_rand(&s->x, s->a) * (1.0f - src->cos_min);
    float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    sin_theta = sin_theta / src->n;
    cos_theta = sqrtf(1.0f - sin_theta * sin_theta);
    float q[3] = {cos_fi * sin_theta, sin_fi * sin_theta, cos_theta};
    for (int i = 0; i < 3; ++i)
        s->d[i] = T[3 * i] * q[0] + T[3 * i + 1] * q[1] + T[3 * i + 2] * q[2];
    float R = o_reflectance(src->n, n1, s->d[2], o_cos_critical(src->n, n1));
    s->w = 1.0f - R;
    o_detector_deposit(&cfg->specular, acc, s->p[0], s->p[1], s->d[0], s->d[1],
                       s->d[2], R, mm);
    s->layer = 1;
}

/* mcsim_boundary, mcml.template.c:80-203 */
static void o_mcml_boundary(const xo_mcml_cfg *cfg, o_pkt *s, int32_t next) {
    const xo_layer *cur = &cfg->layers[s->layer];
    const xo_layer *nxt = &cfg->layers[s->layer + (s->d[2] >= 0.0f ? 1 : -1)];
    float cos_critical = (s->d[2] < 0.0f) ? cur->cos_critical_top
                                          : cur->cos_critical_bottom;
    float n2 = nxt->n;
    if (cur->n == n2) {
        s->layer = next;
        return;
    }
    s->d[2] = -s->d[2];
    float cos1 = fabsf(s->d[2]);
    if (cos1 > cos_critical) {
        if (cos1 > O_FP_COS_0) {
            float R = (cur->n - n2) / (cur->n + n2);
            if (R * R < o_rand(&s->x, s->a)) {
                s->layer = next;
                s->d[2] = -s->d[2];
            }
            return;
        }
        float n1_d_n2 = cur->n / n2;
        float sin1 = sqrtf(1.0f - cos1 * cos1);
        if (cos1 >= O_FP_COS_0) sin1 = 0.0f;
        float sin2 = fminf(1.0f, n1_d_n2 * sin1);
        float cos2 = sqrtf(1.0f - sin2 * sin2);
        float n_cos1 = n1_d_n2 * cos1;
        float n_cos2 = n1_d_n2 * cos2;
        float Rs = (n_cos1 - cos2) / (n_cos1 + cos2);
        Rs *= Rs;
        float Rp = (n_cos2 - cos1) / (n_cos2 + cos1);
        Rp *= Rp;
        float R = 0.5f * (Rs + Rp);
        if (cos1 <= O_FP_COS_90 || sin2 == 1.0f) R = 1.0f;
        if (R < o_rand(&s->x, s->a)) {
            s->layer = next;
            s->d[0] *= n1_d_n2;
            s->d[1] *= n1_d_n2;
            s->d[2] = -copysignf(cos2, s->d[2]);
        }
    }
}

/* One packet from launch to termination: the while(!done) loop of
 * mcml.template.c:514-788 with the three stepping methods; tr (may be NULL)
 * logs the packet as row pkt of the trace. */
static void o_mcml_packet(const xo_mcml_cfg *cfg, const float *lut, o_pkt *s,
                          uint64_t *acc, o_counts *cnt, int mm,
                          const o_trace *tr, uint64_t pkt) {
    const int32_t nl = (int32_t)cfg->num_layers;
    const float rmax2 = cfg->rmax * cfg->rmax;
    const float *sp = (cfg->source_kind == XO_SRC_LINE)
                          ? cfg->source.position
                          : cfg->fiber.position;
    if (cfg->source_kind == XO_SRC_FIBER_NI)
        o_fiber_launch(cfg, s, acc, mm);
    else if (cfg->source_kind == XO_SRC_FIBER)
        o_fiber_axis_launch(cfg, s, acc, mm);
    else
        o_mcml_launch(cfg, s, acc, mm);
    o_tstate ts = {pkt, 0u, XO_EVENT_LAUNCH, 0.0f};
    if (tr && (tr->cfg.options & XO_TRACE_START)) o_trace_event(tr, &ts, s);
    uint32_t pkt_steps = 0;
    for (;;) {
        const xo_layer *L = &cfg->layers[s->layer];
        int done = 0;
        cnt->steps++;
        float step;
        if (cfg->method == XO_METHOD_MBL) /* mcml.template.c:529-531 */
            step = (-o_log(o_rand(&s->x, s->a), mm)) / L->mus;
        else
            step = -o_log(o_rand(&s->x, s->a), mm) * L->inv_mut;
        step = fminf(step, XO_FLT_MAX);
        int32_t next = s->layer;
        if (fmaf(step, s->d[2], s->p[2]) < L->top) {
            --next;
            if (fabsf(s->d[2]) != 0.0f) step = (L->top - s->p[2]) / s->d[2];
        }
        if (fmaf(step, s->d[2], s->p[2]) >= L->bottom) {
            ++next;
            if (fabsf(s->d[2]) != 0.0f)
                step = (L->bottom - s->p[2]) / s->d[2];
        }
        s->p[0] = fmaf(s->d[0], step, s->p[0]);
        s->p[1] = fmaf(s->d[1], step, s->p[1]);
        s->p[2] = fmaf(s->d[2], step, s->p[2]);
        /* optical path length, mcml.template.c:570-574 */
        if (tr && tr->cfg.plon) ts.pl += L->n * step;
        if (s->layer < next) s->p[2] = L->bottom;
        if (s->layer > next) s->p[2] = L->top;

        const int method = cfg->method;
        if (method == XO_METHOD_MBL) {
            /* absorption along the whole step, mcml.template.c:584-606 */
            float mua = L->mua;
            float fraction = 1.0f - o_exp(-(mua * step), mm);
            float deposit = fraction * s->w;
            s->w -= deposit;
            ts.ev |= XO_EVENT_ABSORPTION;
            if (cfg->fluence.enabled) {
                float back = (mua != 0.0f)
                                 ? step - (-o_log(1.0f - o_rand(&s->x, s->a) * fraction,
                                                  mm)) / mua
                                 : 0.0f;
                float q[3] = {s->p[0] - back * s->d[0], s->p[1] - back * s->d[1],
                              s->p[2] - back * s->d[2]};
                o_fluence_deposit(&cfg->fluence, acc, q, deposit, mua, &cnt->deposits);
            }
        }
        if (next != s->layer) {
            cnt->boundary++;
            const int32_t before = s->layer;
            o_mcml_boundary(cfg, s, next);
            ts.ev |= XO_EVENT_BOUNDARY_HIT | (s->layer != before ? XO_EVENT_REFRACTION
                                                                 : XO_EVENT_REFLECTION);
            if (!(s->layer > 0 && s->layer < nl - 1)) {
                if (s->layer <= 0)
                    o_detector_deposit(&cfg->top, acc, s->p[0], s->p[1],
                                       s->d[0], s->d[1], s->d[2], s->w, mm);
                else if (s->layer >= nl - 1)
                    o_detector_deposit(&cfg->bottom, acc, s->p[0], s->p[1],
                                       s->d[0], s->d[1], s->d[2], s->w, mm);
                done = 1;
            }
        } else if (method == XO_METHOD_AR &&
                   o_rand(&s->x, s->a) < L->mua_inv_mut) {
            /* albedo rejection, mcml.template.c:696-708 */
            float deposit = s->w;
            s->w -= deposit;
            ts.ev |= XO_EVENT_ABSORPTION;
            if (cfg->fluence.enabled)
                o_fluence_deposit(&cfg->fluence, acc, s->p, deposit, L->mua, &cnt->deposits);
            done = 1;
        } else {
            if (method == XO_METHOD_AW) {
                float deposit = s->w * L->mua_inv_mut;
                s->w -= deposit;
                ts.ev |= XO_EVENT_ABSORPTION;
                if (cfg->fluence.enabled)
                    o_fluence_deposit(&cfg->fluence, acc, s->p, deposit, L->mua,
                                      &cnt->deposits);
            }
            float xi_fi;
            float ct = o_pf_sample(cfg->pf_kind, &L->pf, lut, &s->x, s->a,
                                   &xi_fi, mm);
            o_scatter_direction(&s->d[0], &s->d[1], &s->d[2], ct, xi_fi, mm);
            ts.ev |= XO_EVENT_SCATTERING;
            if (method == XO_METHOD_AW &&
                o_roulette(s, cfg->w_min, cfg->lottery_chance, cfg->use_lottery))
                done = 1;
        }
        /* microscopic Beer-Lambert: roulette after the boundary or the
         * scattering, whether or not the packet left (mcml.template.c:652-667) */
        if (method == XO_METHOD_MBL &&
            o_roulette(s, cfg->w_min, cfg->lottery_chance, cfg->use_lottery))
            done = 1;
        float ex = s->p[0] - sp[0], ey = s->p[1] - sp[1], ez = s->p[2] - sp[2];
        if (ex * ex + ey * ey + ez * ez > rmax2) {
            done = 1;
            ts.ev |= XO_EVENT_ESCAPED;
        }
        /* guard against packets that can never end (XO_MAX_PACKET_STEPS) */
        if (!done && ++pkt_steps >= XO_MAX_PACKET_STEPS) {
            cnt->truncated++;
            done = 1;
        }
        if (tr) o_trace_step(tr, &ts, s, done);
        if (done) return;
    }
}

static inline uint64_t o_lane_packets(uint64_t n, uint32_t lanes, uint32_t g) {
    return n / lanes + ((uint64_t)g < n % lanes ? 1u : 0u);
}

/* Run npackets packets on `lanes` lanes with the static schedule. rng_x is
 * updated in place (the reference writes the state back,
 * mcml.template.c:821). counts_out: steps, boundary steps, crossings,
 * deposits. */
static int mcml_run(const xo_mcml_cfg *cfg, const float *lut, uint32_t lanes,
                    uint64_t *rng_x, const uint32_t *rng_a, uint64_t npackets,
                    int math_mode, int nthreads, uint64_t *accum, uint64_t naccum,
                    uint64_t *counts_out, const uint32_t *keyed_table,
                    const o_trace *tr);

int oracle_mcml_run(const xo_mcml_cfg *cfg, const float *lut,
                    uint32_t lanes, uint64_t *rng_x, const uint32_t *rng_a,
                    uint64_t npackets, int math_mode, int nthreads,
                    uint64_t *accum, uint64_t naccum, uint64_t *counts_out,
                    const uint32_t *keyed_table) {
    return mcml_run(cfg, lut, lanes, rng_x, rng_a, npackets, math_mode, nthreads,
                    accum, naccum, counts_out, keyed_table, NULL);
}

/* oracle_mcml_run with the photon packet trace: trace_data holds npackets *
 * max_events * XO_TRACE_ENTRY_LEN floats (zeroed here), trace_n npackets
 * event counts; row p is packet p of the static schedule. */
int oracle_mcml_run_trace(const xo_mcml_cfg *cfg, const float *lut,
                          uint32_t lanes, uint64_t *rng_x, const uint32_t *rng_a,
                          uint64_t npackets, int math_mode, int nthreads,
                          uint64_t *accum, uint64_t naccum, uint64_t *counts_out,
                          const uint32_t *keyed_table, const xo_trace *trace,
                          float *trace_data, int32_t *trace_n) {
    if (!trace || trace->max_events < 1) return -1;
    o_trace t = {*trace, trace_data, trace_n};
    memset(trace_data, 0, npackets * (uint64_t)trace->max_events *
                              XO_TRACE_ENTRY_LEN * sizeof(float));
    return mcml_run(cfg, lut, lanes, rng_x, rng_a, npackets, math_mode, nthreads,
                    accum, naccum, counts_out, keyed_table, &t);
}

static int mcml_run(const xo_mcml_cfg *cfg, const float *lut, uint32_t lanes,
                    uint64_t *rng_x, const uint32_t *rng_a, uint64_t npackets,
                    int math_mode, int nthreads, uint64_t *accum, uint64_t naccum,
                    uint64_t *counts_out, const uint32_t *keyed_table,
                    const o_trace *tr) {
    if (cfg->num_layers < 3 || cfg->num_layers > XO_MAX_LAYERS || lanes == 0)
        return -1;
    memset(accum, 0, naccum * sizeof(uint64_t));
    uint64_t st = 0, sb = 0, sd = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : st, sb, sd)
    for (int64_t g = 0; g < (int64_t)lanes; ++g) {
        o_pkt s;
        o_counts c = {0, 0, 0, 0, 0};
        s.x = rng_x[g];
        s.a = rng_a[g];
        uint64_t n = o_lane_packets(npackets, lanes, (uint32_t)g);
        for (uint64_t k = 0; k < n; ++k) {
            if (keyed_table) /* packet g + k*lanes of the keyed run */
                xo_packet_state(cfg->packet_seed,
                                cfg->first_packet + (uint64_t)g + k * lanes,
                                keyed_table, &s.x, &s.a);
            o_mcml_packet(cfg, lut, &s, accum, &c, math_mode, tr,
                          (uint64_t)g + k * lanes);
        }
        if (!keyed_table) rng_x[g] = s.x;
        st += c.steps;
        sb += c.boundary;
        sd += c.deposits;
    }
    if (counts_out) {
        counts_out[0] = st;
        counts_out[1] = sb;
        counts_out[2] = 0;
        counts_out[3] = sd;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* mcvox                                                              */
/* ------------------------------------------------------------------ */
typedef struct o_vox {
    const xo_mcvox_cfg *cfg;
    const int32_t *voxels;
    const float *lut;
} o_vox;

static inline uint64_t o_flat(const xo_voxel_cfg *v, const int32_t i[3]) {
    /* mcvox.template.h:786-787 */
    return ((uint64_t)i[2] * (uint64_t)v->shape[1] + (uint64_t)i[1]) *
               (uint64_t)v->shape[0] + (uint64_t)i[0];
}

static inline int o_valid(const xo_voxel_cfg *v, const int32_t i[3]) {
    /* mcvox.template.h:965-970 */
    return i[0] >= 0 && i[0] < v->shape[0] && i[1] >= 0 &&
           i[1] < v->shape[1] && i[2] >= 0 && i[2] < v->shape[2];
}

/* mcsim_position_to_voxel_index, mcvox.template.c:206-222 */
static inline void o_pos_to_index(const xo_voxel_cfg *v, const float p[3],
                                  int32_t i[3]) {
    for (int k = 0; k < 3; ++k)
        i[k] = o_int((p[k] - v->top_left[k]) / v->size[k]);
}

/* mcsim_sample_box_contains, mcvox.template.h:1227-1234 */
static inline int o_box_contains(const xo_voxel_cfg *v, const float p[3]) {
    return p[0] >= v->top_left[0] && p[0] < v->bottom_right[0] &&
           p[1] >= v->top_left[1] && p[1] < v->bottom_right[1] &&
           p[2] >= v->top_left[2] && p[2] < v->bottom_right[2];
}

/* mcsim_sample_box_intersect, mcvox.template.c:94-162 */
static int o_box_intersect(const xo_voxel_cfg *v, const float pos[3],
                           const float dir[3], float inter[3], float nrm[3]) {
    float inv[3];
    for (int k = 0; k < 3; ++k)
        inv[k] = (dir[k] != 0.0f) ? 1.0f / dir[k] : XO_INF_F;
    float t1 = (v->top_left[0] - pos[0]) * inv[0];
    float t2 = (v->bottom_right[0] - pos[0]) * inv[0];
    float t3 = (v->top_left[1] - pos[1]) * inv[1];
    float t4 = (v->bottom_right[1] - pos[1]) * inv[1];
    float t5 = (v->top_left[2] - pos[2]) * inv[2];
    float t6 = (v->bottom_right[2] - pos[2]) * inv[2];
    float txmin = fminf(t1, t2), txmax = fmaxf(t1, t2);
    float tymin = fminf(t3, t4), tymax = fmaxf(t3, t4);
    float tzmin = fminf(t5, t6), tzmax = fmaxf(t5, t6);
    float tmin = fmaxf(fmaxf(txmin, tymin), tzmin);
    float tmax = fminf(fminf(txmax, tymax), tzmax);
    if (tmin > tmax) return 0;
    nrm[0] = (float)(txmin >= tymin && txmin >= tzmin);
    nrm[1] = (float)((nrm[0] == 0.0f) && tymin >= tzmin);
    nrm[2] = (float)(nrm[0] == 0.0f && nrm[1] == 0.0f);
    nrm[0] *= o_fsign(dir[0]);
    nrm[1] *= o_fsign(dir[1]);
    nrm[2] *= o_fsign(dir[2]);
    float t = (tmin > 0.0f) ? tmin : tmax;
    for (int k = 0; k < 3; ++k) inter[k] = pos[k] + t * dir[k];
    for (int k = 0; k < 3; ++k) {
        /* mc_fclip = clamp(x, lo, hi) */
        float hi = v->bottom_right[k] - 1.1920929e-07f;
        inter[k] = fminf(fmaxf(inter[k], v->top_left[k]), hi);
    }
    return tmax >= tmin && tmax >= 0.0f;
}

/* Launch: the Line (xopto/mcvox/mcsource/line.py:84-104) or GaussianBeam
 * (gaussianbeam.py:112-189) snippet, then mcsim_set_voxel_index_from_position
 * (mcvox.template.c:245-262). */
static void o_mcvox_launch(const o_vox *V, o_pkt *s, uint64_t *acc, int mm) {
    const xo_mcvox_cfg *cfg = V->cfg;
    const xo_voxel_cfg *vc = &cfg->voxels;
    if (cfg->source_kind == XO_SRC_LINE) {
        const xo_source_line *src = &cfg->line;
        s->w = 1.0f - src->reflectance;
        for (int k = 0; k < 3; ++k) {
            s->p[k] = src->position[k];
            s->d[k] = src->direction_sample[k];
        }
        o_detector_deposit(&cfg->specular, acc, s->p[0], s->p[1],
                           src->direction_reflected[0],
                           src->direction_reflected[1],
                           src->direction_reflected[2], src->reflectance, mm);
    } else {
        const xo_source_gaussian *src = &cfg->gaussian;
        float specular_r = 0.0f;
        float r = sqrtf(-2.0f * o_log(1.0f - o_rand(&s->x, s->a), mm));
        r = fminf(r, src->clip);
        float sin_fi, cos_fi;
        o_sincos_azimuth(o_rand(&s->x, s->a), &sin_fi, &cos_fi, mm);
        float ps[3] = {r * cos_fi * src->sigma[0], r * sin_fi * src->sigma[1],
                       0.0f};
        const float *T = src->transformation;
        float pm[3];
        for (int k = 0; k < 3; ++k)
            pm[k] = T[3 * k] * ps[0] + T[3 * k + 1] * ps[1] +
                    T[3 * k + 2] * ps[2];
        for (int k = 0; k < 3; ++k) pm[k] += src->position[k];
        float dir[3] = {src->direction[0], src->direction[1],
                        src->direction[2]};
        if (o_box_contains(vc, pm)) {
            dir[0] = -dir[0];
            dir[1] = -dir[1];
            dir[2] = -dir[2];
        }
        float inter[3], nrm[3];
        if (o_box_intersect(vc, pm, dir, inter, nrm)) {
            for (int k = 0; k < 3; ++k) {
                s->p[k] = inter[k];
                dir[k] = -dir[k];
                nrm[k] = -nrm[k];
            }
            float n_out = cfg->materials[0].n;
            int32_t vi[3];
            o_pos_to_index(vc, inter, vi);
            vi[0] = o_clip(vi[0], 0, vc->shape[0] - 1);
            vi[1] = o_clip(vi[1], 0, vc->shape[1] - 1);
            vi[2] = o_clip(vi[2], 0, vc->shape[2] - 1);
            int32_t mi = V->voxels[o_flat(vc, vi)];
            float n_in = cfg->materials[mi].n;
            float refr[3] = {dir[0], dir[1], dir[2]};
            o_refract_safe(dir, nrm, n_out, n_in, refr);
            for (int k = 0; k < 3; ++k) s->d[k] = refr[k];
            specular_r = o_reflectance(
                n_out, n_in, nrm[0] * dir[0] + nrm[1] * dir[1] + nrm[2] * dir[2],
                o_cos_critical(n_out, n_in));
            o_detector_deposit(&cfg->specular, acc, s->p[0], s->p[1], dir[0],
                               dir[1], dir[2], specular_r, mm);
        } else {
            for (int k = 0; k < 3; ++k) s->p[k] = vc->top_left[k];
            s->d[0] = 0.0f;
            s->d[1] = 0.0f;
            s->d[2] = 1.0f;
            specular_r = 1.0f;
        }
        s->w = 1.0f - specular_r;
    }
    o_pos_to_index(vc, s->p, s->vi);
    s->mi = V->voxels[o_flat(vc, s->vi)];
}

/* mcsim_intersect, mcvox.template.c:173-196 */
static float o_intersect(const xo_voxel_cfg *v, const o_pkt *s, float dist[3]) {
    for (int k = 0; k < 3; ++k) {
        float side = (float)((int32_t)(s->d[k] >= 0.0f) + s->vi[k]);
        dist[k] = v->top_left[k] + side * v->size[k] - s->p[k];
        dist[k] = (s->d[k] != 0.0f) ? dist[k] / s->d[k] : XO_INF_F;
    }
    return fminf(dist[0], fminf(dist[1], dist[2]));
}

/* mcsim_boundary, mcvox.template.c:275-388 */
static void o_mcvox_boundary(const o_vox *V, o_pkt *s, const float dist[3],
                             uint64_t *crossings) {
    const xo_mcvox_cfg *cfg = V->cfg;
    const xo_voxel_cfg *vc = &cfg->voxels;
    int32_t nrm[3];
    nrm[0] = (dist[0] <= dist[1]) && (dist[0] <= dist[2]);
    nrm[1] = nrm[0] ? 0 : ((dist[1] <= dist[0]) && (dist[1] <= dist[2]));
    nrm[2] = !(nrm[0] + nrm[1]);
    for (int k = 0; k < 3; ++k)
        if (s->d[k] < 0.0f) nrm[k] = -nrm[k];
    int32_t nv[3] = {s->vi[0] + nrm[0], s->vi[1] + nrm[1], s->vi[2] + nrm[2]};
    int escaping = !o_valid(vc, nv);
    int32_t next_mi = !escaping ? V->voxels[o_flat(vc, nv)] : 0;
    if (s->mi == next_mi && !escaping) {
        for (int k = 0; k < 3; ++k) s->vi[k] = nv[k];
        (*crossings)++;
        return;
    }
    float n1 = cfg->materials[s->mi].n;
    float n2 = cfg->materials[next_mi].n;
    if (n1 == n2) {
        for (int k = 0; k < 3; ++k) s->vi[k] = nv[k];
        s->mi = next_mi;
        (*crossings)++;
        return;
    }
    float cc = o_cos_critical(n1, n2);
    float cos1 = (float)nrm[0] * s->d[0] + (float)nrm[1] * s->d[1] +
                 (float)nrm[2] * s->d[2];
    float fn[3] = {(float)nrm[0], (float)nrm[1], (float)nrm[2]};
    if (cos1 > cc) {
        float R = o_reflectance(n1, n2, cos1, cc);
        if (R < o_rand(&s->x, s->a)) {
            o_refract(s->d, fn, n1, n2);
            for (int k = 0; k < 3; ++k) s->vi[k] = nv[k];
            s->mi = next_mi;
            (*crossings)++;
            return;
        }
    }
    o_reflect(s->d, fn);
}

/* mcsim_fluence_deposit_at, deposition mode (fluence.py:102-147) */
/* One packet: the while(!done) loop of mcvox.template.c:748-1012 with the
 * three stepping methods (microscopic Beer-Lambert :763-881, albedo
 * rejection :925-944, albedo weight :945-977). */
static void o_mcvox_packet(const o_vox *V, o_pkt *s, uint64_t *acc,
                           o_counts *cnt, int mm, const o_trace *tr,
                           uint64_t pkt) {
    const xo_mcvox_cfg *cfg = V->cfg;
    const xo_voxel_cfg *vc = &cfg->voxels;
    const float rmax2 = cfg->rmax * cfg->rmax;
    const float *sp = (cfg->source_kind == XO_SRC_LINE)
                          ? cfg->line.position
                          : cfg->gaussian.position;
    o_mcvox_launch(V, s, acc, mm);
    o_tstate ts = {pkt, 0u, XO_EVENT_LAUNCH, 0.0f};
    if (tr && (tr->cfg.options & XO_TRACE_START)) o_trace_event(tr, &ts, s);
    uint32_t pkt_steps = 0;
    for (;;) {
        const xo_material *M = &cfg->materials[s->mi];
        int done = 0;
        cnt->steps++;
        const int method = cfg->method;
        float step;
        if (method == XO_METHOD_MBL) /* mcvox.template.c:763-766 */
            step = (-o_log(o_rand(&s->x, s->a), mm)) / M->mus;
        else
            step = -o_log(o_rand(&s->x, s->a), mm) * M->inv_mut;
        step = fminf(step, XO_FLT_MAX);
        float dist[3];
        float d = o_intersect(vc, s, dist);
        float d_ok = fminf(d, step);
        s->p[0] = fmaf(d_ok, s->d[0], s->p[0]);
        s->p[1] = fmaf(d_ok, s->d[1], s->p[1]);
        s->p[2] = fmaf(d_ok, s->d[2], s->p[2]);
        /* optical path length, mcvox.template.c:792-798 */
        if (tr && tr->cfg.plon) ts.pl += M->n * d_ok;
        if (method == XO_METHOD_MBL) {
            /* absorption along the part of the step inside this voxel,
             * mcvox.template.c:800-818 */
            float mua = M->mua;
            float fraction = 1.0f - o_exp(-(mua * d_ok), mm);
            float deposit = fraction * s->w;
            s->w -= deposit;
            ts.ev |= XO_EVENT_ABSORPTION;
            if (cfg->fluence.enabled) {
                float back = (mua != 0.0f)
                                 ? d_ok - (-o_log(1.0f - o_rand(&s->x, s->a) * fraction,
                                                  mm)) / mua
                                 : 0.0f;
                float q[3] = {s->p[0] - back * s->d[0], s->p[1] - back * s->d[1],
                              s->p[2] - back * s->d[2]};
                o_fluence_deposit(&cfg->fluence, acc, q, deposit, mua, &cnt->deposits);
            }
        }
        if (d < step) {
            cnt->boundary++;
            const uint64_t crossed = cnt->crossings;
            o_mcvox_boundary(V, s, dist, &cnt->crossings);
            ts.ev |= XO_EVENT_BOUNDARY_HIT | (cnt->crossings != crossed
                                                  ? XO_EVENT_REFRACTION
                                                  : XO_EVENT_REFLECTION);
            if (!o_valid(vc, s->vi)) {
                if (s->vi[2] < 0)
                    o_detector_deposit(&cfg->top, acc, s->p[0], s->p[1],
                                       s->d[0], s->d[1], s->d[2], s->w, mm);
                else if (s->vi[2] >= vc->shape[2])
                    o_detector_deposit(&cfg->bottom, acc, s->p[0], s->p[1],
                                       s->d[0], s->d[1], s->d[2], s->w, mm);
                done = 1;
            }
        } else if (method == XO_METHOD_AR &&
                   o_rand(&s->x, s->a) < M->mua_inv_mut) {
            /* albedo rejection, mcvox.template.c:925-944 */
            float deposit = s->w;
            s->w -= deposit;
            ts.ev |= XO_EVENT_ABSORPTION;
            if (cfg->fluence.enabled)
                o_fluence_deposit(&cfg->fluence, acc, s->p, deposit, M->mua, &cnt->deposits);
            done = 1;
        } else {
            if (method == XO_METHOD_AW) {
                float deposit = s->w * M->mua_inv_mut;
                s->w -= deposit;
                ts.ev |= XO_EVENT_ABSORPTION;
                if (cfg->fluence.enabled)
                    o_fluence_deposit(&cfg->fluence, acc, s->p, deposit, M->mua,
                                      &cnt->deposits);
            }
            float xi_fi;
            float ct = o_pf_sample(cfg->pf_kind, &M->pf, V->lut, &s->x, s->a,
                                   &xi_fi, mm);
            o_scatter_direction(&s->d[0], &s->d[1], &s->d[2], ct, xi_fi, mm);
            ts.ev |= XO_EVENT_SCATTERING;
            if (method == XO_METHOD_AW &&
                o_roulette(s, cfg->w_min, cfg->lottery_chance, cfg->use_lottery))
                done = 1;
        }
        /* microscopic Beer-Lambert: roulette after the boundary or the
         * scattering, whether or not the packet left (mcvox.template.c:865-881) */
        if (method == XO_METHOD_MBL &&
            o_roulette(s, cfg->w_min, cfg->lottery_chance, cfg->use_lottery))
            done = 1;
        float ex = s->p[0] - sp[0], ey = s->p[1] - sp[1], ez = s->p[2] - sp[2];
        if (ex * ex + ey * ey + ez * ez > rmax2 || s->w <= 0.0f) {
            done = 1;
            ts.ev |= XO_EVENT_ESCAPED;
        }
        /* guard against packets that can never end (XO_MAX_PACKET_STEPS) */
        if (!done && ++pkt_steps >= XO_MAX_PACKET_STEPS) {
            cnt->truncated++;
            done = 1;
        }
        if (tr) o_trace_step(tr, &ts, s, done);
        if (done) return;
    }
}

static int mcvox_run(const xo_mcvox_cfg *cfg, const int32_t *voxels,
                     const float *lut, uint32_t lanes, uint64_t *rng_x,
                     const uint32_t *rng_a, uint64_t npackets, int math_mode,
                     int nthreads, uint64_t *accum, uint64_t naccum,
                     uint64_t *counts_out, const uint32_t *keyed_table,
                     const o_trace *tr);

int oracle_mcvox_run(const xo_mcvox_cfg *cfg, const int32_t *voxels,
                     const float *lut, uint32_t lanes, uint64_t *rng_x,
                     const uint32_t *rng_a, uint64_t npackets, int math_mode,
                     int nthreads, uint64_t *accum, uint64_t naccum,
                     uint64_t *counts_out, const uint32_t *keyed_table) {
    return mcvox_run(cfg, voxels, lut, lanes, rng_x, rng_a, npackets, math_mode,
                     nthreads, accum, naccum, counts_out, keyed_table, NULL);
}

/* oracle_mcvox_run with the photon packet trace (buffers as for
 * oracle_mcml_run_trace). */
int oracle_mcvox_run_trace(const xo_mcvox_cfg *cfg, const int32_t *voxels,
                           const float *lut, uint32_t lanes, uint64_t *rng_x,
                           const uint32_t *rng_a, uint64_t npackets,
                           int math_mode, int nthreads, uint64_t *accum,
                           uint64_t naccum, uint64_t *counts_out,
                           const uint32_t *keyed_table, const xo_trace *trace,
                           float *trace_data, int32_t *trace_n) {
    if (!trace || trace->max_events < 1) return -1;
    o_trace t = {*trace, trace_data, trace_n};
    memset(trace_data, 0, npackets * (uint64_t)trace->max_events *
                              XO_TRACE_ENTRY_LEN * sizeof(float));
    return mcvox_run(cfg, voxels, lut, lanes, rng_x, rng_a, npackets, math_mode,
                     nthreads, accum, naccum, counts_out, keyed_table, &t);
}

static int mcvox_run(const xo_mcvox_cfg *cfg, const int32_t *voxels,
                     const float *lut, uint32_t lanes, uint64_t *rng_x,
                     const uint32_t *rng_a, uint64_t npackets, int math_mode,
                     int nthreads, uint64_t *accum, uint64_t naccum,
                     uint64_t *counts_out, const uint32_t *keyed_table,
                     const o_trace *tr) {
    if (cfg->num_materials < 1 || cfg->num_materials > XO_MAX_MATERIALS ||
        lanes == 0)
        return -1;
    o_vox V = {cfg, voxels, lut};
    memset(accum, 0, naccum * sizeof(uint64_t));
    uint64_t st = 0, sb = 0, sc = 0, sd = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : st, sb, sc, sd)
    for (int64_t g = 0; g < (int64_t)lanes; ++g) {
        o_pkt s;
        o_counts c = {0, 0, 0, 0, 0};
        s.x = rng_x[g];
        s.a = rng_a[g];
        uint64_t n = o_lane_packets(npackets, lanes, (uint32_t)g);
        for (uint64_t k = 0; k < n; ++k) {
            if (keyed_table)
                xo_packet_state(cfg->packet_seed,
                                cfg->first_packet + (uint64_t)g + k * lanes,
                                keyed_table, &s.x, &s.a);
            o_mcvox_packet(&V, &s, accum, &c, math_mode, tr,
                           (uint64_t)g + k * lanes);
        }
        if (!keyed_table) rng_x[g] = s.x;
        st += c.steps;
        sb += c.boundary;
        sc += c.crossings;
        sd += c.deposits;
    }
    if (counts_out) {
        counts_out[0] = st;
        counts_out[1] = sb;
        counts_out[2] = sc;
        counts_out[3] = sd;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* batches of layered instances                                       */
/* ------------------------------------------------------------------ */
/* Every packet p of instance i runs the layered packet loop above with the
 * generator xo_packet_state(seed_i, p) (pyxopto_amd/csrc/xo_seedhash.h), so
 * the result is independent of scheduling; table holds the multipliers of
 * seed-table rows 1..499999. */
int oracle_batch_run(const xo_batch_instance *inst, uint32_t ninst,
                     const float *lut, const uint32_t *table, float w_min,
                     float lottery_chance, int use_lottery, int method,
                     int math_mode, int nthreads, uint64_t *accum,
                     uint64_t naccum, uint64_t *counts_out) {
    memset(accum, 0, naccum * sizeof(uint64_t));
    xo_mcml_cfg *cfgs = (xo_mcml_cfg *)calloc(ninst, sizeof(xo_mcml_cfg));
    if (!cfgs) return -1;
    for (uint32_t i = 0; i < ninst; ++i) {
        xo_mcml_cfg *c = &cfgs[i];
        if (inst[i].num_layers < 3 || inst[i].num_layers > XO_BATCH_MAX_LAYERS) {
            free(cfgs);
            return -1;
        }
        c->num_layers = inst[i].num_layers;
        c->pf_kind = inst[i].pf_kind;
        c->rmax = inst[i].rmax;
        c->w_min = w_min;
        c->lottery_chance = lottery_chance;
        c->use_lottery = use_lottery;
        c->method = method;
        c->source = inst[i].source;
        c->source_kind = inst[i].source_kind;
        c->fiber = inst[i].fiber;
        c->fiber_axis = inst[i].fiber_axis;
        c->top = inst[i].top;
        c->bottom = inst[i].bottom;
        c->specular = inst[i].specular;
        c->fluence = inst[i].fluence;
        memcpy(c->layers, inst[i].layers, sizeof(xo_layer) * inst[i].num_layers);
    }
    uint64_t st = 0, sb = 0, sd = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    for (uint32_t i = 0; i < ninst; ++i) {
        const int64_t np = (int64_t)inst[i].npackets;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : st, sb, sd)
        for (int64_t p = 0; p < np; ++p) {
            o_pkt s;
            o_counts c = {0, 0, 0, 0, 0};
            xo_packet_state(inst[i].seed, (uint64_t)p, table, &s.x, &s.a);
            o_mcml_packet(&cfgs[i], lut, &s, accum, &c, math_mode, NULL, 0);
            st += c.steps;
            sb += c.boundary;
            sd += c.deposits;
        }
    }
    free(cfgs);
    if (counts_out) {
        counts_out[0] = st;
        counts_out[1] = sb;
        counts_out[2] = 0;
        counts_out[3] = sd;
    }
    return 0;
}

/* Threads the oracle uses by default. */
int oracle_max_threads(void) { return omp_get_max_threads(); }

/* Evaluate the kernels' fixed-sequence math on n inputs (for the accuracy
 * tests): which 0 = xo_logf, 1 = xo_sincos_turns (sin in out, cos in out2),
 * 2 = xo_cbrtf. */
int oracle_math(int which, uint32_t n, const float *in, float *out,
                float *out2) {
    for (uint32_t i = 0; i < n; ++i) {
        if (which == 0) {
            out[i] = xo_logf(in[i]);
        } else if (which == 1) {
            xo_sincos_turns(in[i], &out[i], &out2[i]);
        } else if (which == 2) {
            out[i] = xo_cbrtf(in[i]);
        } else if (which == 3) {
            out[i] = xo_expf(in[i]);
        } else if (which == 4) {
            out[i] = xo_logf_unit(in[i]);
        } else {
            return -1;
        }
    }
    return 0;
}

/* n draws of mcsim_pf_sample_angles from one stream (phase-function tests). */
int oracle_pf_samples(int kind, const xo_pf *pf, const float *lut, uint64_t x,
                      uint32_t a, uint32_t n, int mm, float *cos_out,
                      float *fi_out) {
    for (uint32_t i = 0; i < n; ++i)
        cos_out[i] = o_pf_sample(kind, pf, lut, &x, a, &fi_out[i], mm);
    return 0;
}

/* scatter_direction on n directions (direction-update tests). */
int oracle_scatter(uint32_t n, const float *dir_in, const float *cos_theta,
                   const float *xi_fi, int mm, float *dir_out) {
    for (uint32_t i = 0; i < n; ++i) {
        float dx = dir_in[3 * i], dy = dir_in[3 * i + 1], dz = dir_in[3 * i + 2];
        o_scatter_direction(&dx, &dy, &dz, cos_theta[i], xi_fi[i], mm);
        dir_out[3 * i] = dx;
        dir_out[3 * i + 1] = dy;
        dir_out[3 * i + 2] = dz;
    }
    return 0;
}

/* reflectance() of mcbase.template.c:1244-1290 on n incidence cosines. */
int oracle_reflectance(float n1, float n2, uint32_t n, const float *cos1,
                       float *out) {
    float cc = o_cos_critical(n1, n2);
    for (uint32_t i = 0; i < n; ++i) out[i] = o_reflectance(n1, n2, cos1[i], cc);
    return 0;
}

/* ------------------------------------------------------------------ */
/* sampling volume                                                    */
/* ------------------------------------------------------------------ */
static inline uint32_t o_weight(float v) { return o_uint(v + 0.5f); }

static inline float o_distance(const float a[3], const float b[3]) {
    /* mc_distance_point3f, mcbase.template.c:914-916, 1061-1064 */
    float x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return sqrtf(x * x + y * y + z * z);
}

static inline void o_sv_deposit(const xo_sampling_volume *sv, uint64_t *acc,
                                const int32_t v[3], uint32_t w) {
    if (v[0] >= 0 && v[1] >= 0 && v[2] >= 0 && (uint32_t)v[0] < sv->shape[0] &&
        (uint32_t)v[1] < sv->shape[1] && (uint32_t)v[2] < sv->shape[2])
        acc[sv->offset + ((uint64_t)v[2] * sv->shape[1] + (uint64_t)v[1]) *
                             sv->shape[0] + (uint64_t)v[0]] += w;
}

/* The SamplingVolume kernel of mcsv.template.c:236-420, one packet after
 * the other: walk the path between consecutive trace events through the
 * voxel lattice (sva_intersect :95-115, sva_next_voxel :130-152) and give
 * each voxel round(terminal weight * path in the voxel * multiplier * k);
 * the terminal weight is field end_field of the last valid entry (the
 * reference reads field 7, :278). Packets without events are skipped.
 * counts_out: [walk steps, truncated walks]. */
int oracle_sampling_volume(const xo_sampling_volume *sv, int32_t max_events,
                           const float *data, const int32_t *n,
                           uint64_t npackets, uint64_t *accum, uint64_t naccum,
                           uint64_t *total_weight, uint64_t *counts_out) {
    if (max_events < 1 || (sv->end_field != 6 && sv->end_field != 7)) return -1;
    memset(accum, 0, naccum * sizeof(uint64_t));
    const float kf = (float)sv->k;
    uint64_t total = 0, walked = 0, truncated = 0;
    for (uint64_t p = 0; p < npackets; ++p) {
        int32_t nv = n[p] < max_events ? n[p] : max_events;
        if (nv <= 0) continue;
        const float *base = data + p * (uint64_t)max_events * XO_TRACE_ENTRY_LEN;
        float end_w = base[(uint64_t)(nv - 1) * XO_TRACE_ENTRY_LEN + sv->end_field];
        total += o_weight(end_w * kf);
        float p1[3], d1[3], p2[3], d2[3];
        const float *e = base;
        for (int k = 0; k < 3; ++k) {
            p1[k] = e[k];
            d1[k] = e[3 + k];
        }
        e = base + (uint64_t)(nv > 1 ? 1 : 0) * XO_TRACE_ENTRY_LEN;
        for (int k = 0; k < 3; ++k) {
            p2[k] = e[k];
            d2[k] = e[3 + k];
        }
        float d12 = o_distance(p1, p2);
        int32_t v[3];
        for (int k = 0; k < 3; ++k)
            v[k] = o_int((p1[k] - sv->top_left[k]) / sv->voxel_size[k]);
        float l_voxel = 0.0f;
        int32_t idx = 0;
        uint32_t steps = 0;
        for (;;) {
            float dist[3];
            for (int k = 0; k < 3; ++k) {
                float side = (float)((int64_t)(d1[k] >= 0.0f) + v[k]);
                dist[k] = sv->top_left[k] + side * sv->voxel_size[k] - p1[k];
                dist[k] = (d1[k] != 0.0f) ? dist[k] / d1[k] : XO_INF_F;
            }
            float d_voxel = fminf(dist[2], fminf(dist[0], dist[1]));
            float d_move = fminf(d_voxel, d12);
            l_voxel += d_move;
            int last = 0;
            if (d_voxel < d12) {
                o_sv_deposit(sv, accum, v,
                             o_weight(end_w * l_voxel * sv->multiplier * kf));
                for (int k = 0; k < 3; ++k) p1[k] += d1[k] * d_move;
                d12 -= d_move;
                int32_t nx = (dist[0] <= dist[1]) && (dist[0] <= dist[2]);
                int32_t ny = nx ? 0 : ((dist[1] <= dist[0]) && (dist[1] <= dist[2]));
                int32_t nz = !(nx + ny);
                v[0] += d1[0] < 0.0f ? -nx : nx;
                v[1] += d1[1] < 0.0f ? -ny : ny;
                v[2] += d1[2] < 0.0f ? -nz : nz;
                l_voxel = 0.0f;
            } else {
                for (int k = 0; k < 3; ++k) {
                    p1[k] = p2[k];
                    d1[k] = d2[k];
                }
                ++idx;
                if (idx < nv - 1) {
                    e = base + (uint64_t)(idx + 1) * XO_TRACE_ENTRY_LEN;
                    for (int k = 0; k < 3; ++k) {
                        p2[k] = e[k];
                        d2[k] = e[3 + k];
                    }
                    d12 = o_distance(p1, p2);
                } else {
                    o_sv_deposit(sv, accum, v,
                                 o_weight(end_w * l_voxel * sv->multiplier * kf));
                    last = 1;
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
    *total_weight = total;
    if (counts_out) {
        counts_out[0] = walked;
        counts_out[1] = truncated;
    }
    return 0;
}
