/*
 * xo_device.h — device building blocks shared by the transport kernels:
 * the MWC generator, detector and fluence deposits, phase-function sampling,
 * direction update and Fresnel physics. Each function states the reference
 * lines whose arithmetic it performs; the operation order (and every fused
 * multiply-add) is part of the contract, because the kernels are built with
 * -ffp-contract=off and must round like the host oracle.
 */
#ifndef XO_DEVICE_H
#define XO_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/xomc.h"
#include "xo_math.h"

namespace xo {

#define XO_DEV __device__ __forceinline__

constexpr float kCos0 = 1.0f;   /* FP_COS_0,  mcbase.template.h:512 */
constexpr float kCos90 = 0.0f;  /* FP_COS_90, mcbase.template.h:510 */
constexpr float kRmin = 1e-12f; /* FP_RMIN,   mcbase.template.h:476 */
constexpr float kInv2p32 = 2.3283064365386963e-10f; /* 2^-32 */

/* Layer as staged in LDS: the per-step fields in the first 16 bytes. */
struct alignas(16) LayerL {
    float top, bottom, inv_mut, mua_inv_mut;
    float n, cc_top, cc_bottom, g;
    float beta, la, lb, lc;
    uint32_t loff, lsize_m1;
    float mua, mus; /* microscopic Beer-Lambert steps */
};

/* Material as staged in LDS. */
struct alignas(16) MaterialL {
    float inv_mut, mua_inv_mut, n, g;
    float beta, la, lb, lc;
    uint32_t loff, lsize_m1;
    float mua, mus; /* microscopic Beer-Lambert steps */
};

/* Correctly rounded square root for x = +0 or 2^-96 <= x < 2^96, in 9
 * instructions: the hardware estimate s, then s - 1 ulp or s + 1 ulp chosen
 * from the signs of the residuals x - s' s (Markstein's correction, the
 * same operations in the same order as the compiler's general expansion,
 * whose rescaling of inputs below 2^-96 and special-value pass-through do
 * nothing in this range). Checked against sqrtf for every float of the
 * range. Callers: arguments 1 - c*c with |c| <= 1 (0 or >= 2^-24) and
 * squared norms of unit vectors. */
XO_DEV float sqrt_in_range(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float s_dn = __uint_as_float(__float_as_uint(s) - 1u);
    const float s_up = __uint_as_float(__float_as_uint(s) + 1u);
    const float r_dn = __builtin_fmaf(-s_dn, s, x);
    const float r_up = __builtin_fmaf(-s_up, s, x);
    float r = (0.0f >= r_dn) ? s_dn : s;
    r = (0.0f < r_up) ? s_up : r;
    return r;
}

/* Correctly rounded a / b for normal a, b with 2^-60 <= |a|, |b| <= 2^60:
 * reciprocal estimate, one Newton step on it, quotient, two residual
 * corrections (the compiler's general division sequence without the
 * operand scaling and the special-case fix-up, which leave such operands
 * unchanged). Checked against the division on 2^32 random operand pairs of
 * that range. Callers: 1 / |d| for a direction d of length about 1. */
XO_DEV float div_in_range(float a, float b) {
    float y = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y, 1.0f);
    y = __builtin_fmaf(e, y, y);
    float q = a * y;
    float r = __builtin_fmaf(-b, q, a);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(r, y, q);
}

/* a / k for k = sqrt(1 - c^2) with |c| < 1 (so 2^-12 < k <= 1) and |a| <= 2:
 * the range-restricted division when |a| >= 2^-60, the general one for
 * smaller |a| (rare; it keeps the sign of a zero). Same bits either way. */
XO_DEV float div_by_k(float a, float k) {
    if (__builtin_expect(fabsf(a) >= 0x1p-60f, 1)) return div_in_range(a, k);
    return a / k;
}

/* Arithmetic of the transport step, a template parameter of the kernels.
 *
 * MathExact: the fixed-sequence log, sin/cos, exp and cube root of
 * xo_math.h and correctly rounded division and square root (IEEE 754 fixes
 * their results), so a host build of the same operations rounds the same
 * way: the kernels then match the oracle bit for bit.
 *
 * MathFast: one hardware instruction each, v_log_f32 (log2), v_exp_f32
 * (exp2), v_sin_f32 / v_cos_f32 (argument in turns), v_sqrt_f32, v_rsq_f32
 * and v_rcp_f32 (division = multiplication by the reciprocal). Each is
 * within a few ulp (checked against float64 by the math probe); the
 * results are then statistically, not bitwise, those of the oracle. This
 * is the build the reference makes with -cl-fast-relaxed-math
 * (xopto/cl/cloptions.py:137-152), the option of its published benchmark
 * (test/performance.py:104). */
struct MathExact {
    static constexpr bool fast = false;
    /* ln(u) for u = +0 or a normal u <= 1 */
    XO_DEV static float log_unit(float u) { return xo_logf_unit(u); }
    XO_DEV static float log(float x) { return xo_logf(x); }
    XO_DEV static float exp(float x) { return xo_expf(x); }
    XO_DEV static void sincos_turns(float t, float *s, float *c) { xo_sincos_turns(t, s, c); }
    XO_DEV static float cbrt(float x) { return xo_cbrtf(x); }
    XO_DEV static float sqrt(float x) { return __builtin_sqrtf(x); }
    /* the range-restricted forms (see sqrt_in_range, div_in_range,
     * div_by_k for their ranges) */
    XO_DEV static float sqrt_r(float x) { return sqrt_in_range(x); }
    XO_DEV static float div_r(float a, float b) { return div_in_range(a, b); }
    XO_DEV static float div_k(float a, float k) { return div_by_k(a, k); }
    /* 1 / sqrt(x) for the squared norm x of a near-unit vector */
    XO_DEV static float rsqrt_r(float x) { return div_in_range(1.0f, sqrt_in_range(x)); }
    XO_DEV static float div(float a, float b) { return a / b; }
};

struct MathFast {
    static constexpr bool fast = true;
    XO_DEV static float log_unit(float u) {
        return __builtin_amdgcn_logf(u) * 0.693147180559945309f;
    }
    XO_DEV static float log(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
    XO_DEV static float exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
    XO_DEV static void sincos_turns(float t, float *s, float *c) {
        *s = __builtin_amdgcn_sinf(t);
        *c = __builtin_amdgcn_cosf(t);
    }
    /* sign(x) 2^(log2|x| / 3); +-0 stays +-0 (log2 0 = -inf, 2^-inf = 0) */
    XO_DEV static float cbrt(float x) {
        const float y = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(fabsf(x)) *
                                               0.333333333333333333f);
        return copysignf(y, x);
    }
    XO_DEV static float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    XO_DEV static float sqrt_r(float x) { return __builtin_amdgcn_sqrtf(x); }
    XO_DEV static float div_r(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    XO_DEV static float div_k(float a, float k) { return a * __builtin_amdgcn_rcpf(k); }
    XO_DEV static float rsqrt_r(float x) { return __builtin_amdgcn_rsqf(x); }
    XO_DEV static float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
};

/* fp_random_single (mcbase.template.c:1576-1594): one MWC step. The state
 * update is one v_mad_u64_u32; (float)(uint32)x / (float)0xFFFFFFFF divides
 * by 2^32 exactly, which is the multiply below. */
XO_DEV float rng_next(uint64_t &x, uint32_t a) {
    x = (uint64_t)(uint32_t)x * (uint64_t)a + (x >> 32);
    return (float)(uint32_t)x * kInv2p32;
}

/* convert_int with saturation (mc_int). */
XO_DEV int32_t f2i_sat(float v) {
    if (!(v == v)) return 0;
    if (v >= 2147483520.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}

XO_DEV uint32_t f2u_sat(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967040.0f) return UINT32_MAX;
    return (uint32_t)v;
}

/* Where detector bins live for this launch: LDS copies (flushed at the end
 * of the block) when every detector bin fits, else global memory. */
struct Bins {
    unsigned long long *lds;  /* nullptr if not privatized */
    unsigned long long *glob;
};

XO_DEV void bins_add(const Bins &b, uint32_t index, uint32_t w) {
    if (b.lds)
        atomicAdd(&b.lds[index], (unsigned long long)w);
    else
        atomicAdd(&b.glob[index], (unsigned long long)w);
}

/* Radial (radial.py:117-150) / Total (total.py:86-108) detector deposit.
 * accu_64_deposit_32 (mcbase.template.c:58-61) becomes one u64 atomic add:
 * the modular sum is identical. */
template <class M = MathExact>
XO_DEV void detector_deposit(const xo_detector &d, const Bins &bins, float px,
                             float py, float dx, float dy, float dz, float w) {
    if (d.kind == XO_DET_NONE) return;
    uint32_t index = 0;
    if (d.kind == XO_DET_RADIAL) {
        float ddx = px - d.position[0];
        float ddy = py - d.position[1];
        float r = M::sqrt(ddx * ddx + ddy * ddy);
        if (d.log_scale) r = M::log(fmaxf(r, kRmin));
        int32_t ri = f2i_sat((r - d.r_min) * d.inv_dr);
        ri = ri < 0 ? 0 : (ri > (int32_t)d.n - 1 ? (int32_t)d.n - 1 : ri);
        index = (uint32_t)ri;
    }
    float cosv = fabsf(dx * d.direction[0] + dy * d.direction[1] +
                       dz * d.direction[2]);
    float wi = (w * (float)XO_ACCU_K + 0.5f) * (float)(d.cos_min <= cosv);
    uint32_t ui = f2u_sat(wi);
    if (ui > 0u) bins_add(bins, d.offset + index, ui);
}

/* The deposit of n packets that all leave at the same point in the same
 * direction with the same weight: n times detector_deposit's integer, in
 * one atomic (the modular sum is the same). */
template <class M = MathExact>
XO_DEV void detector_deposit_n(const xo_detector &d, const Bins &bins, float px, float py,
                               float dx, float dy, float dz, float w, uint32_t n) {
    if (d.kind == XO_DET_NONE || n == 0u) return;
    uint32_t index = 0;
    if (d.kind == XO_DET_RADIAL) {
        float ddx = px - d.position[0];
        float ddy = py - d.position[1];
        float r = M::sqrt(ddx * ddx + ddy * ddy);
        if (d.log_scale) r = M::log(fmaxf(r, kRmin));
        int32_t ri = f2i_sat((r - d.r_min) * d.inv_dr);
        ri = ri < 0 ? 0 : (ri > (int32_t)d.n - 1 ? (int32_t)d.n - 1 : ri);
        index = (uint32_t)ri;
    }
    float cosv = fabsf(dx * d.direction[0] + dy * d.direction[1] + dz * d.direction[2]);
    float wi = (w * (float)XO_ACCU_K + 0.5f) * (float)(d.cos_min <= cosv);
    uint32_t ui = f2u_sat(wi);
    if (ui > 0u) {
        const unsigned long long v = (unsigned long long)ui * n;
        if (bins.lds)
            atomicAdd(&bins.lds[d.offset + index], v);
        else
            atomicAdd(&bins.glob[d.offset + index], v);
    }
}

/* Photon packet trace state of one lane: the packet's row, its logged
 * event count, this step's event flags and the optical path length. */
struct TraceState {
    uint64_t pkt;
    uint32_t count;
    uint32_t ev;
    float pl;
};

/* mcsim_trace_event (xopto/mcbase/mctrace.py:527-560): unless the event
 * mask rejects this step's events, write the packet state to entry
 * min(count, max_events - 1) of its row and count it. */
XO_DEV void trace_event(const xo_trace &T, float *data, TraceState &ts, float px, float py,
                        float pz, float dx, float dy, float dz, float w) {
    if (T.use_events && !(T.event_mask & ts.ev)) return;
    const uint32_t last = (uint32_t)T.max_events - 1u;
    const uint32_t slot = ts.count < last ? ts.count : last;
    float4 *dst = reinterpret_cast<float4 *>(
        data + (ts.pkt * (uint64_t)T.max_events + slot) * XO_TRACE_ENTRY_LEN);
    dst[0] = make_float4(px, py, pz, dx);
    dst[1] = make_float4(dy, dz, w, T.plon ? ts.pl : 0.0f);
    ts.count++;
}

/* A lane's one-entry cache of its last global fluence deposit (the
 * accumulator cache of mcbase.template.h:2780-2872, McUseFluenceCache):
 * consecutive deposits to the same bin add up in a register and reach
 * memory as one atomic. The cached integer is flushed before it could pass
 * 32 bits, so every bin receives the same integers in the same modular u64
 * sum as without the cache. */
struct FluenceCache {
    uint32_t index = 0xFFFFFFFFu;   /* accumulator index held, or none    */
    uint32_t w = 0u;                /* integer weight not yet added       */
};

XO_DEV void cache_add(FluenceCache &c, unsigned long long *accum, uint64_t index,
                      uint32_t ui) {
    if (index >= 0xFFFFFFFFull) {
        atomicAdd(&accum[index], (unsigned long long)ui);
        return;
    }
    if ((uint32_t)index == c.index && c.w <= 0xFFFFFFFFu - ui) {
        c.w += ui;
        return;
    }
    if (c.w) atomicAdd(&accum[c.index], (unsigned long long)c.w);
    c.index = (uint32_t)index;
    c.w = ui;
}

XO_DEV void cache_flush(FluenceCache &c, unsigned long long *accum) {
    if (c.w) atomicAdd(&accum[c.index], (unsigned long long)c.w);
    c.w = 0u;
    c.index = 0xFFFFFFFFu;
}

/* The weight a fluence accumulator adds for an absorbed weight w: w in
 * deposition mode, w / mua in fluence-rate mode (0 where mua is 0),
 * fluence.py:139-141. */
template <class M = MathExact> XO_DEV float fluence_weight(const xo_fluence &F, float w, float mua) {
    if (F.rate) w *= (mua != 0.0f) ? M::div(1.0f, mua) : 0.0f;
    return w;
}

/* Fluence deposit at a point: the Cartesian voxel of
 * mcsim_fluence_deposit_at (fluence.py:102-147) or the (r, z) cell of
 * fluencerz.py:112-160; mua is that of the medium the packet is in (used
 * in fluence-rate mode). Returns true if a non-zero integer was added. */
template <class M = MathExact>
XO_DEV bool fluence_deposit(const xo_fluence &F, unsigned long long *accum, float px,
                            float py, float pz, float w, float mua,
                            FluenceCache *cache = nullptr) {
    uint64_t index;
    if (F.enabled == XO_FLUENCE_XYZ) {
        const float fx = (px - F.top_left[0]) * F.inv_step[0];
        const float fy = (py - F.top_left[1]) * F.inv_step[1];
        const float fz = (pz - F.top_left[2]) * F.inv_step[2];
        if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx < (float)F.shape[0] &&
              fy < (float)F.shape[1] && fz < (float)F.shape[2]))
            return false;
        const uint64_t ix = f2u_sat(fx), iy = f2u_sat(fy), iz = f2u_sat(fz);
        index = (iz * F.shape[1] + iy) * F.shape[0] + ix;
    } else if (F.enabled == XO_FLUENCE_RZ) {
        const float ddx = px - F.top_left[0];
        const float ddy = py - F.top_left[1];
        const float r = M::sqrt(ddx * ddx + ddy * ddy);
        const float ddz = pz - F.top_left[2];
        const float fr = r * F.inv_step[0];
        const float fz = ddz * F.inv_step[2];
        if (!(fr >= 0.0f && fz >= 0.0f && fr < (float)F.shape[0] && fz < (float)F.shape[2]))
            return false;
        index = (uint64_t)f2u_sat(fz) * F.shape[0] + f2u_sat(fr);
    } else {
        return false;
    }
    const uint32_t ui = f2u_sat(fluence_weight<M>(F, w, mua) * (float)F.k + 0.5f);
    if (ui == 0u) return false;
    if (cache)
        cache_add(*cache, accum, F.offset + index, ui);
    else
        atomicAdd(&accum[F.offset + index], (unsigned long long)ui);
    return true;
}

/* The same LDS object behind a pointer the optimizer cannot see through:
 * loads through it stay where they are used instead of being hoisted out of
 * the transport loop into scalar registers, which the loop would spill and
 * reload lane by lane. For configuration read once per event (launch,
 * escape, deposit). */
template <class T> XO_DEV const T &lds_fresh(const T &ref) {
    typedef const T __attribute__((address_space(3))) *LdsPtr;
    LdsPtr q = (LdsPtr)&ref;
    asm volatile("" : "+v"(q));
    return *(const T *)q;
}

/* Event-time configuration of the single-simulation kernels, copied to
 * LDS once per block. */
struct LayerEventCfgL {
    xo_detector top, bottom, specular;
    xo_fluence fluence;
    xo_source_line line;
    xo_source_fiber fiber;
    xo_fiber_axis fiber_axis;
};

struct VoxEventCfgL {
    xo_detector top, bottom, specular;
    xo_fluence fluence;
    xo_source_line line;
    xo_source_gaussian gaussian;
};

/* fluence_deposit for a Cartesian grid whose bins in the box
 * [box, box + (bx, by, bz)) are held in LDS (sbox, x fastest): the same
 * index and integer as fluence_deposit, added to the box when the bin lies
 * in it, to accum through the lane's cache otherwise. */
template <class M = MathExact>
XO_DEV bool fluence_deposit_boxed(const xo_fluence &F, unsigned long long *accum,
                                  unsigned long long *sbox, const int32_t box[3], int32_t bx,
                                  int32_t by, int32_t bz, float px, float py, float pz,
                                  float w, float mua, FluenceCache &cache) {
    const float fx = (px - F.top_left[0]) * F.inv_step[0];
    const float fy = (py - F.top_left[1]) * F.inv_step[1];
    const float fz = (pz - F.top_left[2]) * F.inv_step[2];
    if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx < (float)F.shape[0] &&
          fy < (float)F.shape[1] && fz < (float)F.shape[2]))
        return false;
    const uint32_t ix = f2u_sat(fx), iy = f2u_sat(fy), iz = f2u_sat(fz);
    const uint32_t ui = f2u_sat(fluence_weight<M>(F, w, mua) * (float)F.k + 0.5f);
    if (ui == 0u) return false;
    const uint32_t rx = ix - (uint32_t)box[0], ry = iy - (uint32_t)box[1],
                   rz = iz - (uint32_t)box[2];
    if (rx < (uint32_t)bx && ry < (uint32_t)by && rz < (uint32_t)bz)
        atomicAdd(&sbox[(rz * (uint32_t)by + ry) * (uint32_t)bx + rx], (unsigned long long)ui);
    else
        cache_add(cache, accum, F.offset + ((uint64_t)iz * F.shape[1] + iy) * F.shape[0] + ix,
                  ui);
    return true;
}

/* Phase-function parameters of one layer/material, read from LDS. */
struct PfL {
    float g, beta, la, lb, lc;
    uint32_t loff, lsize_m1;
};

/* mcsim_pf_sample_angles: returns cos(theta), the azimuth as a fraction of
 * a turn in xi_fi. HG hg.py:80-96; MHg mhg.py:87-109; Lut lut.py:125-150
 * (index clamped to the table). */
template <int PF, class M = MathExact>
XO_DEV float pf_sample(const PfL &pf, const float *lut, uint64_t &x, uint32_t a,
                       float &xi_fi) {
    xi_fi = rng_next(x, a);
    if constexpr (PF == XO_PF_HG) {
        float g = pf.g;
        /* |g| <= 1 (the host clips it): the numerator is +0 or >= 2^-24, the
         * denominator 0 or in [2^-24, 2]; a 0 denominator gives NaN either
         * way, which the clamp below turns into the same cosine */
        float k = M::div_r(1.0f - g * g, 1.0f + g * (2.0f * rng_next(x, a) - 1.0f));
        float ct = M::div(1.0f + g * g - k * k, 2.0f * g);
        if (g == 0.0f) ct = 1.0f - 2.0f * rng_next(x, a);
        return fmaxf(fminf(ct, 1.0f), -1.0f);
    } else if constexpr (PF == XO_PF_MHG) {
        float g = pf.g;
        float ct;
        if (rng_next(x, a) <= pf.beta) {
            /* as for Hg: |g| <= 1 */
            float k = M::div_r(1.0f - g * g, 1.0f + g * (2.0f * rng_next(x, a) - 1.0f));
            ct = M::div(1.0f + g * g - k * k, 2.0f * g);
            if (g == 0.0f) ct = 1.0f - 2.0f * rng_next(x, a);
        } else {
            ct = M::cbrt(2.0f * rng_next(x, a) - 1.0f);
        }
        return fminf(fmaxf(ct, -1.0f), 1.0f);
    } else {
        float fsize = (float)pf.lsize_m1;
        float fi = (M::div(pf.la, rng_next(x, a) - pf.lc) - pf.lb + 1.0f) * fsize *
                   0.5f;
        fi = fminf(fmaxf(fi, 0.0f), fsize);
        float fl = floorf(fi);
        float d = fi - fl;
        uint32_t i0 = (uint32_t)f2i_sat(fl);
        uint32_t i1 = i0 + 1u < pf.lsize_m1 ? i0 + 1u : pf.lsize_m1;
        return lut[pf.loff + i0] * (1.0f - d) + lut[pf.loff + i1] * d;
    }
}

/* scatter_direction (mcbase.template.c:1522-1556) with the single-precision
 * renormalization (mc_normalize_fpv3, :975-981). */
template <class M = MathExact>
XO_DEV void scatter_direction(float &dx, float &dy, float &dz, float ct,
                              float xi_fi) {
    float sin_fi, cos_fi;
    float st = M::sqrt_r(1.0f - ct * ct);
    M::sincos_turns(xi_fi, &sin_fi, &cos_fi);
    float stcf = st * cos_fi;
    float stsf = st * sin_fi;
    float px = dx, py = dy, pz = dz;
    if (fabsf(pz) >= kCos0) {
        dx = stcf;
        dy = stsf;
        dz = copysignf(ct, pz * ct);
    } else {
        float k = M::sqrt_r(1.0f - pz * pz);
        dx = M::div_k(stcf * px * pz - stsf * py, k) + px * ct;
        dy = M::div_k(stcf * py * pz + stsf * px, k) + py * ct;
        dz = (-stcf) * k + pz * ct;
    }
    float inv = M::rsqrt_r(dx * dx + dy * dy + dz * dz);
    dx = dx * inv;
    dy = dy * inv;
    dz = dz * inv;
}

/* cos_critical (mcbase.template.c:1229-1231). */
template <class M = MathExact> XO_DEV float cos_critical(float n1, float n2) {
    return (n1 > n2) ? M::sqrt_r(1.0f - M::div(n2 * n2, n1 * n1)) : 0.0f;
}

/* reflectance (mcbase.template.c:1244-1290). */
template <class M = MathExact>
XO_DEV float reflectance(float n1, float n2, float cos1, float cc) {
    float R = 1.0f;
    cos1 = fabsf(cos1);
    if (n1 == n2) return 0.0f;
    if (cos1 > cc) {
        float n1_d_n2 = M::div(n1, n2);
        float sin1 = M::sqrt_r(1.0f - cos1 * cos1);
        if (cos1 >= kCos0) sin1 = 0.0f;
        float sin2 = fminf(1.0f, n1_d_n2 * sin1);
        float cos2 = M::sqrt_r(1.0f - sin2 * sin2);
        float n_cos1 = n1_d_n2 * cos1;
        float n_cos2 = n1_d_n2 * cos2;
        float Rs = M::div(n_cos1 - cos2, n_cos1 + cos2);
        Rs *= Rs;
        float Rp = M::div(n_cos2 - cos1, n_cos2 + cos1);
        Rp *= Rp;
        R = 0.5f * (Rp + Rs);
        if (cos1 <= kCos90 || sin2 == 1.0f) return 1.0f;
    }
    return R;
}

XO_DEV float fsign(float v) { return (v >= 0.0f) ? 1.0f : -1.0f; }

/* reflectance_cos2 (mcbase.template.c:1303-1346): Fresnel reflectance from
 * the cosine of the transmitted angle. */
template <class M = MathExact>
XO_DEV float reflectance_cos2(float n1, float n2, float cos2) {
    float R = 1.0f;
    cos2 = fabsf(cos2);
    if (n1 == n2) return 0.0f;
    float sin2 = M::sqrt(1.0f - cos2 * cos2);
    if (cos2 >= kCos0) sin2 = 0.0f;
    float sin1 = M::div(n2, n1) * sin2;
    if (sin1 < 1.0f) {
        float cos1 = M::sqrt(1.0f - sin1 * sin1);
        float n1_d_n2 = M::div(n1, n2);
        float n_cos1 = n1_d_n2 * cos1;
        float n_cos2 = n1_d_n2 * cos2;
        float Rs = M::div(n_cos1 - cos2, n_cos1 + cos2);
        Rs *= Rs;
        float Rp = M::div(n_cos2 - cos1, n_cos2 + cos1);
        Rp *= Rp;
        R = 0.5f * (Rp + Rs);
        if (cos1 <= kCos90 || sin2 == 1.0f) return 1.0f;
    }
    return R;
}

/* UniformFiberNI launch (xopto/mcml/mcsource/fiberni.py:237-290): a point
 * uniform over the core, a direction uniform in cos(theta) over the
 * acceptance cone, refracted into the sample (n_sample, layer 1), weight
 * 1 - R with R the Fresnel reflectance core -> sample. Returns R. */
template <class M = MathExact>
XO_DEV float fiber_ni_launch(const xo_source_fiber &src, float n_sample, uint64_t &x,
                             uint32_t a, float &px, float &py, float &pz, float &dx,
                             float &dy, float &dz, float &w) {
    float sin_fi, cos_fi;
    const float r = M::sqrt(rng_next(x, a)) * src.radius;
    M::sincos_turns(rng_next(x, a), &sin_fi, &cos_fi);
    px = src.position[0] + r * cos_fi;
    py = src.position[1] + r * sin_fi;
    pz = 0.0f;
    M::sincos_turns(rng_next(x, a), &sin_fi, &cos_fi);
    float cos_theta = 1.0f - rng_next(x, a) * (1.0f - src.cos_min);
    float sin_theta = M::sqrt(1.0f - cos_theta * cos_theta);
    sin_theta = M::div(sin_theta, n_sample);
    cos_theta = M::sqrt(1.0f - sin_theta * sin_theta);
    dx = cos_fi * sin_theta;
    dy = sin_fi * sin_theta;
    dz = cos_theta;
    const float R = reflectance_cos2<M>(src.n, n_sample, dz);
    w = 1.0f - R;
    return R;
}

/* Launch of UniformFiber (xopto/mcml/mcsource/fiber.py:265-330): a point
 * uniform over the core, rotated into the sample frame by the fiber axis and
 * projected along the fiber onto z = 0; cos(theta) uniform over the
 * acceptance cone in air, scaled to the core index, rotated likewise. The
 * reference tests the projected point's z (0) against the critical cosine
 * before refracting, so the direction enters the sample unrefracted; the
 * weight is 1 - the Fresnel reflectance at that direction. Returns that
 * reflectance (the specular part, deposited along the launch direction). */
template <class M = MathExact>
XO_DEV float fiber_launch(const xo_source_fiber &src, const xo_fiber_axis &ax, float n_sample,
                          uint64_t &x, uint32_t a, float &px, float &py, float &pz, float &dx,
                          float &dy, float &dz, float &w) {
    const float *T = ax.transformation;
    float sin_fi, cos_fi;
    const float r = M::sqrt(rng_next(x, a)) * src.radius;
    M::sincos_turns(rng_next(x, a), &sin_fi, &cos_fi);
    const float s0 = r * cos_fi, s1 = r * sin_fi, s2 = 0.0f;
    float m0 = T[0] * s0 + T[1] * s1 + T[2] * s2;
    float m1 = T[3] * s0 + T[4] * s1 + T[5] * s2;
    const float m2 = T[6] * s0 + T[7] * s1 + T[8] * s2;
    const float k = M::div(0.0f - m2, ax.direction[2]);
    m0 += k * ax.direction[0];
    m1 += k * ax.direction[1];
    px = src.position[0] + m0;
    py = src.position[1] + m1;
    pz = 0.0f;
    M::sincos_turns(rng_next(x, a), &sin_fi, &cos_fi);
    float cos_theta = 1.0f - rng_next(x, a) * (1.0f - src.cos_min);
    float sin_theta = M::sqrt(1.0f - cos_theta * cos_theta);
    sin_theta = M::div(sin_theta, src.n);
    cos_theta = M::sqrt(1.0f - sin_theta * sin_theta);
    const float q0 = cos_fi * sin_theta, q1 = sin_fi * sin_theta, q2 = cos_theta;
    dx = T[0] * q0 + T[1] * q1 + T[2] * q2;
    dy = T[3] * q0 + T[4] * q1 + T[5] * q2;
    dz = T[6] * q0 + T[7] * q1 + T[8] * q2;
    const float R = reflectance<M>(src.n, n_sample, dz, cos_critical<M>(src.n, n_sample));
    w = 1.0f - R;
    return R;
}

/* Direction of the specularly reflected part of a fiber packet: the
 * sample direction mirrored at the surface, (dx, dy, -dz), refracted from
 * the sample (n1) into the core (n2) across the normal (0, 0, -1)
 * (refract, mcbase.template.c:1443-1462). */
template <class M = MathExact>
XO_DEV void fiber_specular_direction(float dx, float dy, float dz, float n1, float n2,
                                     float &rx, float &ry, float &rz) {
    const float p0 = dx, p1 = dy, p2 = -dz;
    const float cos1 = p0 * 0.0f + p1 * 0.0f + p2 * -1.0f;
    const float n1_d_n2 = M::div(n1, n2);
    const float sin2_squared = n1_d_n2 * n1_d_n2 * (1.0f - cos1 * cos1);
    const float k = fsign(cos1) * (n1_d_n2 * fabsf(cos1) - M::sqrt(1.0f - sin2_squared));
    rx = n1_d_n2 * p0 - k * 0.0f;
    ry = n1_d_n2 * p1 - k * 0.0f;
    rz = n1_d_n2 * p2 - k * -1.0f;
}

/* Roulette (mcml.template.c:738-753); returns true if the packet ends. */
XO_DEV bool roulette(float &w, uint64_t &x, uint32_t a, float w_min,
                     float chance, int use_lottery) {
    if (w < w_min) {
        if (use_lottery) {
            if (rng_next(x, a) > chance) return true;
            w = w / chance;
        } else {
            return true;
        }
    }
    return false;
}

/* mcsim_boundary, mcml.template.c:80-203 (no surface layouts). */
template <class M = MathExact>
XO_DEV void mcml_boundary(const LayerL *sl, int32_t &layer, int32_t next,
                          float &dx, float &dy, float &dz, uint64_t &x,
                          uint32_t a) {
    const LayerL &cur = sl[layer];
    const LayerL &nxt = sl[layer + (dz >= 0.0f ? 1 : -1)];
    float cos_crit = (dz < 0.0f) ? cur.cc_top : cur.cc_bottom;
    float n1 = cur.n;
    float n2 = nxt.n;
    if (n1 == n2) {
        layer = next;
        return;
    }
    dz = -dz;
    float cos1 = fabsf(dz);
    if (cos1 > cos_crit) {
        if (cos1 > kCos0) {
            float R = M::div(n1 - n2, n1 + n2);
            if (R * R < rng_next(x, a)) {
                layer = next;
                dz = -dz;
            }
            return;
        }
        float n1_d_n2 = M::div(n1, n2);
        float sin1 = M::sqrt_r(1.0f - cos1 * cos1);
        if (cos1 >= kCos0) sin1 = 0.0f;
        float sin2 = fminf(1.0f, n1_d_n2 * sin1);
        float cos2 = M::sqrt_r(1.0f - sin2 * sin2);
        float n_cos1 = n1_d_n2 * cos1;
        float n_cos2 = n1_d_n2 * cos2;
        float Rs = M::div(n_cos1 - cos2, n_cos1 + cos2);
        Rs *= Rs;
        float Rp = M::div(n_cos2 - cos1, n_cos2 + cos1);
        Rp *= Rp;
        float R = 0.5f * (Rs + Rp);
        if (cos1 <= kCos90 || sin2 == 1.0f) R = 1.0f;
        if (R < rng_next(x, a)) {
            layer = next;
            dx *= n1_d_n2;
            dy *= n1_d_n2;
            dz = -copysignf(cos2, dz);
        }
    }
}

/* Wave-level helpers (64-lane wavefronts). */
XO_DEV uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

XO_DEV uint32_t prefix_count(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

XO_DEV uint64_t readlane_u64(uint64_t v, uint32_t lane) {
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)lane);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)lane);
    return ((uint64_t)hi << 32) | lo;
}

/* Dynamic schedule of the single-simulation kernels: a wave takes packet
 * indices from the launch's global counter in blocks of `claim` with one
 * atomic and hands them to its idle lanes in lane order, refilling only
 * once refill_min lanes (or all of its live lanes) wait. One counter shared
 * by every wave of the grid would otherwise see an atomic every couple of
 * transport steps per wave. Wave-uniform state. */
struct WaveDispenser {
    uint64_t next = 0, end = 0;
    bool empty = false;
};

/* Returns true (and the index in pkt) for a lane that gets a packet; sets
 * retire for a lane that wants one when the counter is exhausted; claimed
 * (wave-uniform) is the number of packets of the blocks taken. */
XO_DEV bool dispense(WaveDispenser &q, bool want, bool alive, unsigned long long *counter,
                     uint64_t npackets, uint32_t claim, uint32_t refill_min, uint64_t &pkt,
                     bool &retire, uint64_t &claimed) {
    claimed = 0;
    retire = false;
    const uint64_t m = __ballot(want);
    const uint32_t nwant = (uint32_t)__builtin_popcountll(m);
    if (nwant == 0) return false;
    const uint32_t nalive = (uint32_t)__builtin_popcountll(__ballot(alive));
    if (nwant < refill_min && nalive != 0) return false;
    const uint32_t rank = prefix_count(m);
    const uint32_t lane = lane_id();
    const uint32_t leader = (uint32_t)__builtin_ctzll(__ballot(true));
    uint32_t assigned = 0;
    bool got = false;
    while (assigned < nwant) {
        if (q.next >= q.end) {
            if (q.empty) break;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(counter, (unsigned long long)claim);
            base = readlane_u64(base, leader);
            if (base >= npackets) {
                q.empty = true;
                break;
            }
            q.next = base;
            q.end = base + claim < npackets ? base + claim : npackets;
            claimed += q.end - q.next;
        }
        const uint64_t avail = q.end - q.next;
        const uint32_t left = nwant - assigned;
        const uint32_t take = avail < (uint64_t)left ? (uint32_t)avail : left;
        if (want && rank >= assigned && rank < assigned + take) {
            got = true;
            pkt = q.next + (rank - assigned);
        }
        q.next += take;
        assigned += take;
    }
    if (want && !got && q.empty) retire = true;
    return got;
}

constexpr uint32_t kClaim = 128;    /* packets per counter atomic */
constexpr uint32_t kRefillMin = 4;  /* idle lanes that trigger a refill */

/* Packets a wave claims per counter atomic: kClaim when the run has at
 * least two packets per lane, one per lane otherwise, so that a run with
 * fewer packets than twice the lanes spreads them over every wave instead
 * of stacking two on each lane of the first waves. */
template <class Params> XO_DEV uint32_t claim_size(const Params &P) {
    return P.npackets >= 2ull * P.lanes ? kClaim : 64u;
}

} // namespace xo

#endif /* XO_DEVICE_H */
