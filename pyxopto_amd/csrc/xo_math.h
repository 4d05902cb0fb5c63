/*
 * xo_math.h — binary32 elementary functions with a fixed operation sequence.
 *
 * The transport kernels call these instead of the device math library so that
 * a host build of the same operations (gcc, -ffp-contract=off, explicit fmaf)
 * rounds identically: log, sin/cos and cbrt are built only from +, -, *, fma,
 * correctly rounded division/sqrt, bit casts and rint, which IEEE 754 fixes
 * bit for bit on every conforming machine.
 *
 * Methods (none taken from a vendor library):
 *  - xo_logf: S. L. Moshier, Cephes Math Library, logf (published): reduce to
 *    m in [sqrt(1/2), sqrt(2)), log(1+f) = f - f^2/2 + f^3 P(f), split ln 2.
 *  - xo_sincos_turns: sin/cos of 2*pi*t. Quadrant from rint(4t) (exact), the
 *    remainder r = t - q/4 is exact (Sterbenz), theta = 2*pi*r in [-pi/4, pi/4],
 *    Cephes sinf/cosf kernel polynomials on that interval.
 *  - xo_cbrtf: W. Kahan's exponent-division initial estimate followed by two
 *    Halley iterations in correction form y <- y - y (y^3 - a)/(2y^3 + a).
 *
 * Accuracy (checked against float64 references by the test suite): xo_logf
 * <= 2 ulp on [2^-32, 2], sin/cos <= 2e-7 absolute, xo_cbrtf <= 2 ulp.
 *
 * The reference draws the azimuth as fi = 2*pi*xi and calls sincos(fi)
 * (xopto/mcbase/mcpf/hg.py:80-96, mcbase.template.c:1522-1556); evaluating
 * sin/cos of 2*pi*xi through turns samples the same uniform azimuth.
 */
#ifndef XO_MATH_H
#define XO_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define XO_HD __host__ __device__ static inline
#else
#define XO_HD static inline
#endif

#if defined(__HIPCC__) || defined(__cplusplus)
#define XO_FMA(a, b, c) __builtin_fmaf((a), (b), (c))
#define XO_SQRT(x) __builtin_sqrtf(x)
#define XO_RINT(x) __builtin_rintf(x)
#else
#include <math.h>
#define XO_FMA(a, b, c) fmaf((a), (b), (c))
#define XO_SQRT(x) sqrtf(x)
#define XO_RINT(x) rintf(x)
#endif

#define XO_INF_F (__builtin_inff())
#define XO_NAN_F (__builtin_nanf(""))
#define XO_FLT_MAX 3.40282346638528859812e+38f

XO_HD uint32_t xo_f2u(float f) {
    union { float f; uint32_t u; } v;
    v.f = f;
    return v.u;
}

XO_HD float xo_u2f(uint32_t u) {
    union { float f; uint32_t u; } v;
    v.u = u;
    return v.f;
}

/* Natural logarithm. */
XO_HD float xo_logf(float x) {
    if (!(x > 0.0f)) return (x == 0.0f) ? -XO_INF_F : XO_NAN_F;
    if (x == XO_INF_F) return x;
    int32_t eadj = 0;
    if (x < 1.17549435e-38f) { /* subnormal: scale by 2^23 */
        x = x * 8388608.0f;
        eadj = -23;
    }
    uint32_t bits = xo_f2u(x);
    int32_t e = (int32_t)((bits >> 23) & 0xffu) - 126 + eadj;
    float m = xo_u2f((bits & 0x007fffffu) | 0x3f000000u); /* [0.5, 1) */
    float f;
    if (m < 0.707106781186547524f) {
        e -= 1;
        f = (m + m) - 1.0f;
    } else {
        f = m - 1.0f;
    }
    float z = f * f;
    float p = 7.0376836292e-2f;
    p = XO_FMA(p, f, -1.1514610310e-1f);
    p = XO_FMA(p, f, 1.1676998740e-1f);
    p = XO_FMA(p, f, -1.2420140846e-1f);
    p = XO_FMA(p, f, 1.4249322787e-1f);
    p = XO_FMA(p, f, -1.6668057665e-1f);
    p = XO_FMA(p, f, 2.0000714765e-1f);
    p = XO_FMA(p, f, -2.4999993993e-1f);
    p = XO_FMA(p, f, 3.3333331174e-1f);
    float y = (p * f) * z;
    float fe = (float)e;
    y = XO_FMA(fe, -2.12194440e-4f, y);
    y = XO_FMA(z, -0.5f, y);
    float r = f + y;
    return XO_FMA(fe, 0.693359375f, r);
}

/* xo_logf for x = +0 or a normal x <= 1, such as a uniform deviate
 * u * 2^-32 or 1 minus one: the same operations without the infinity and
 * subnormal branches, so the same result. */
XO_HD float xo_logf_unit(float x) {
    if (!(x > 0.0f)) return (x == 0.0f) ? -XO_INF_F : XO_NAN_F;
    uint32_t bits = xo_f2u(x);
    int32_t e = (int32_t)((bits >> 23) & 0xffu) - 126;
    float m = xo_u2f((bits & 0x007fffffu) | 0x3f000000u); /* [0.5, 1) */
    float f;
    if (m < 0.707106781186547524f) {
        e -= 1;
        f = (m + m) - 1.0f;
    } else {
        f = m - 1.0f;
    }
    float z = f * f;
    float p = 7.0376836292e-2f;
    p = XO_FMA(p, f, -1.1514610310e-1f);
    p = XO_FMA(p, f, 1.1676998740e-1f);
    p = XO_FMA(p, f, -1.2420140846e-1f);
    p = XO_FMA(p, f, 1.4249322787e-1f);
    p = XO_FMA(p, f, -1.6668057665e-1f);
    p = XO_FMA(p, f, 2.0000714765e-1f);
    p = XO_FMA(p, f, -2.4999993993e-1f);
    p = XO_FMA(p, f, 3.3333331174e-1f);
    float y = (p * f) * z;
    float fe = (float)e;
    y = XO_FMA(fe, -2.12194440e-4f, y);
    y = XO_FMA(z, -0.5f, y);
    float r = f + y;
    return XO_FMA(fe, 0.693359375f, r);
}

/* Exponential: n = round(x / ln 2), r = x - n ln 2 in two Cody-Waite
 * parts, Cephes' degree-5 polynomial (S. L. Moshier, Cephes Math Library,
 * expf.c) for e^r, scaled by 2^n through the exponent field (two steps
 * below 2^-126 so results stay correct down to the subnormals). */
XO_HD float xo_expf(float x) {
    if (x != x) return x;
    if (x > 88.7228391f) return XO_INF_F;
    if (x < -103.972084f) return 0.0f;
    float n = XO_RINT(x * 1.44269504088896341f);
    float r = XO_FMA(n, -0.693359375f, x);
    r = XO_FMA(n, 2.12194440e-4f, r);
    float p = 1.9875691500e-4f;
    p = XO_FMA(p, r, 1.3981999507e-3f);
    p = XO_FMA(p, r, 8.3334519073e-3f);
    p = XO_FMA(p, r, 4.1665795894e-2f);
    p = XO_FMA(p, r, 1.6666665459e-1f);
    p = XO_FMA(p, r, 5.0000001201e-1f);
    float y = XO_FMA(p, r * r, r) + 1.0f;
    int32_t ni = (int32_t)n;
    if (ni > 127) {
        y = y * 2.0f;
        ni -= 1;
    }
    if (ni < -126) {
        y = y * xo_u2f((uint32_t)(ni + 126 + 127) << 23);
        return y * 1.17549435e-38f; /* 2^-126 */
    }
    return y * xo_u2f((uint32_t)(ni + 127) << 23);
}

/* sin(2*pi*t) and cos(2*pi*t). */
XO_HD void xo_sincos_turns(float t, float *s_out, float *c_out) {
    float q = XO_RINT(4.0f * t);
    float r = XO_FMA(q, -0.25f, t);           /* exact */
    float th = r * 6.28318530717958647692f;   /* [-pi/4, pi/4] */
    float th2 = th * th;
    float ps = -1.9515295891e-4f;
    ps = XO_FMA(ps, th2, 8.3321608736e-3f);
    ps = XO_FMA(ps, th2, -1.6666654611e-1f);
    float s = XO_FMA(th * th2, ps, th);
    float pc = 2.443315711809948e-5f;
    pc = XO_FMA(pc, th2, -1.388731625493765e-3f);
    pc = XO_FMA(pc, th2, 4.166664568298827e-2f);
    float c = XO_FMA(th2 * th2, pc, XO_FMA(th2, -0.5f, 1.0f));
    int32_t iq = ((int32_t)q) & 3;
    float so, co;
    if (iq == 0) { so = s; co = c; }
    else if (iq == 1) { so = c; co = -s; }
    else if (iq == 2) { so = -s; co = -c; }
    else { so = -c; co = s; }
    *s_out = so;
    *c_out = co;
}

/* Real cube root. */
XO_HD float xo_cbrtf(float x) {
    if (x == 0.0f || !(x == x) || x == XO_INF_F || x == -XO_INF_F) return x;
    float a = (x < 0.0f) ? -x : x;
    int32_t eadj = 0;
    if (a < 1.17549435e-38f) { /* subnormal: scale by 2^24, cbrt scale 2^8 */
        a = a * 16777216.0f;
        eadj = 1;
    }
    float y = xo_u2f(xo_f2u(a) / 3u + 709921077u);
    for (int i = 0; i < 2; ++i) {
        float y2 = y * y;
        float res = XO_FMA(y2, y, -a);              /* y^3 - a */
        float den = XO_FMA(y2 + y2, y, a);          /* 2y^3 + a */
        y = XO_FMA(-y, res / den, y);
    }
    if (eadj) y = y * 0.00390625f;
    return (x < 0.0f) ? -y : y;
}

#endif /* XO_MATH_H */
