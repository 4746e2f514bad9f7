/*
 * gp_math.h - FP32/FP64 elementary functions built from IEEE +, -, *, /,
 * sqrt and exact bit operations only, shared by the HIP kernels and the host
 * oracle.  Compiled with -ffp-contract=off on both sides, every function gives
 * bit-identical results on gfx950 and on x86-64, which is what lets per-history
 * results of the GPU path be compared bit for bit with the CPU oracle.
 *
 * Methods (all textbook; coefficients are Taylor coefficients rounded to the
 * working precision):
 *   log:  x = 2^k m, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
 *         s = (m-1)/(m+1), odd series in s; k*ln2 added with a Cody-Waite
 *         two-part ln2.
 *   exp:  Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, degree-8 Taylor
 *         polynomial, scaling by 2^k through the exponent field.
 *   sin/cos(2 pi u): exact reduction of 4u to the nearest integer quadrant
 *         q and remainder f in [-1/2, 1/2]; Taylor polynomials of sin(pi f/2)
 *         and cos(pi f/2).
 *   sin/cos(x): Cody-Waite reduction by pi/2 in three parts, then Taylor.
 * Accuracy is stated and checked against FP64 libm in the tests.
 *
 * These replace the reference's use of __logf/__sinf/__cosf/expf/cbrtf
 * (e.g. gPET_kernals.cu:109,216-217,278,474-478), whose implementations are
 * not reproduced here.
 */
#ifndef GP_MATH_H
#define GP_MATH_H

#include <stdint.h>
#include <math.h>
#include <string.h>
#include "gp_philox.h"

GP_HD uint32_t gp_f2u(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
GP_HD float gp_u2f(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}
GP_HD uint64_t gp_d2u(double f)
{
    uint64_t u;
    memcpy(&u, &f, 8);
    return u;
}
GP_HD double gp_u2d(uint64_t u)
{
    double f;
    memcpy(&f, &u, 8);
    return f;
}

#define GP_LN2F_HI 0.693115234375f     /* ln 2 rounded to 12 bits          */
#define GP_LN2F_LO 3.19461833e-05f     /* ln 2 - GP_LN2F_HI                 */
#define GP_LOG2EF 1.44269502f
#define GP_SQRT2F 1.41421354f

/* Natural log, x > 0.  Max error about 1.5 ulp (tested). */
GP_HD float gp_logf(float x)
{
    uint32_t ix = gp_f2u(x);
    int k = 0;
    if (ix >= 0x7f800000u) return (ix == 0x7f800000u) ? x : (x - x) / (x - x); /* inf, nan, <0 */
    if (ix < 0x00800000u) {
        if (ix == 0u) return -INFINITY;
        x = x * 8388608.0f; /* 2^23 */
        ix = gp_f2u(x);
        k = -23;
    }
    k += (int)(ix >> 23) - 127;
    float m = gp_u2f((ix & 0x007fffffu) | 0x3f800000u);
    if (m > GP_SQRT2F) {
        m = m * 0.5f;
        k += 1;
    }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float p = z * (0.333333343f + z * (0.200000003f + z * (0.142857149f + z * 0.111111112f)));
    float t = s + s * p;
    float lm = t + t;
    float fk = (float)k;
    return fk * GP_LN2F_HI + (fk * GP_LN2F_LO + lm);
}

/* 2^k for integer k in [-149, 127] as a float, through two normal factors. */
GP_HD float gp_scale2f(float p, int k)
{
    int k1 = k / 2;
    int k2 = k - k1;
    return (p * gp_u2f((uint32_t)(k1 + 127) << 23)) * gp_u2f((uint32_t)(k2 + 127) << 23);
}

/* e^x.  Max error about 1 ulp for normal results (tested). */
GP_HD float gp_expf(float x)
{
    if (x != x) return x;
    if (x > 88.7228394f) return INFINITY;
    if (x < -103.972084f) return 0.0f;
    float kf = floorf(x * GP_LOG2EF + 0.5f);
    float r = (x - kf * GP_LN2F_HI) - kf * GP_LN2F_LO;
    float p = 2.48015876e-05f;
    p = 1.98412701e-04f + r * p;
    p = 1.38888892e-03f + r * p;
    p = 8.33333377e-03f + r * p;
    p = 4.16666679e-02f + r * p;
    p = 1.66666672e-01f + r * p;
    p = 0.5f + r * p;
    p = 1.0f + r * p;
    p = 1.0f + r * p;
    return gp_scale2f(p, (int)kf);
}

/* sin and cos of pi*f/2 for f in [-1/2, 1/2]. */
GP_HD void gp_sincos_halfpi(float f, float *s, float *c)
{
    float z = f * f;
    float ps = -3.59884323e-06f;
    ps = 1.60441178e-04f + z * ps;
    ps = -4.68175398e-03f + z * ps;
    ps = 7.96926245e-02f + z * ps;
    ps = -6.45964086e-01f + z * ps;
    ps = 1.57079637f + z * ps;
    float pc = 4.71087468e-07f;
    pc = -2.52020423e-05f + z * pc;
    pc = 9.19260256e-04f + z * pc;
    pc = -2.08634809e-02f + z * pc;
    pc = 2.53669500e-01f + z * pc;
    pc = -1.23370051f + z * pc;
    pc = 1.0f + z * pc;
    *s = f * ps;
    *c = pc;
}

/* sin(2 pi u), cos(2 pi u) for any finite u (used with u in (0, 1]). */
GP_HD void gp_sincos2pi(float u, float *s, float *c)
{
    float t = 4.0f * u;               /* exact */
    float q = floorf(t + 0.5f);
    float f = t - q;                  /* exact, in [-1/2, 1/2] */
    float sf, cf;
    gp_sincos_halfpi(f, &sf, &cf);
    int iq = ((int)q) & 3;
    if (iq == 0) { *s = sf; *c = cf; }
    else if (iq == 1) { *s = cf; *c = -sf; }
    else if (iq == 2) { *s = -sf; *c = -cf; }
    else { *s = -cf; *c = sf; }
}

/* sin x, cos x with Cody-Waite reduction; accurate for |x| < 6000. */
GP_HD void gp_sincosf(float x, float *s, float *c)
{
    float kf = floorf(x * 0.636619747f + 0.5f);
    float r = ((x - kf * 1.5703125f) - kf * 4.83751297e-04f) - kf * 7.54978995e-08f;
    float z = r * r;
    float ps = -2.50521079e-08f;
    ps = 2.75573188e-06f + z * ps;
    ps = -1.98412701e-04f + z * ps;
    ps = 8.33333377e-03f + z * ps;
    ps = -1.66666672e-01f + z * ps;
    float sr = r + r * (z * ps);
    float pc = 2.08767570e-09f;
    pc = -2.75573200e-07f + z * pc;
    pc = 2.48015876e-05f + z * pc;
    pc = -1.38888892e-03f + z * pc;
    pc = 4.16666679e-02f + z * pc;
    pc = -0.5f + z * pc;
    float cr = 1.0f + z * pc;
    int iq = ((int)kf) & 3;
    if (iq == 0) { *s = sr; *c = cr; }
    else if (iq == 1) { *s = cr; *c = -sr; }
    else if (iq == 2) { *s = -sr; *c = -cr; }
    else { *s = -cr; *c = sr; }
}

GP_HD float gp_cosf(float x)
{
    float s, c;
    gp_sincosf(x, &s, &c);
    return c;
}

/* atan2(y, x) in (-pi, pi]: reduce to a = min/max in [0, 1], then to
 * |t| <= tan(pi/8) with atan(a) = pi/4 + atan((a-1)/(a+1)); odd Taylor
 * series to t^17.  Error about 2 ulp (tested). */
GP_HD float gp_atan2f(float y, float x)
{
    float ax = fabsf(x), ay = fabsf(y);
    if (ax == 0.0f && ay == 0.0f) return 0.0f;
    bool swap = ay > ax;
    float a = swap ? ax / ay : ay / ax;
    float base = 0.0f, t = a;
    if (a > 0.414213562f) {
        t = (a - 1.0f) / (a + 1.0f);
        base = 0.785398163f;
    }
    float z = t * t;
    float p = 5.88235296e-02f;         /*  1/17 */
    p = -6.66666701e-02f + z * p;      /* -1/15 */
    p = 7.69230798e-02f + z * p;       /*  1/13 */
    p = -9.09090936e-02f + z * p;      /* -1/11 */
    p = 0.111111112f + z * p;          /*  1/9  */
    p = -0.142857149f + z * p;         /* -1/7  */
    p = 0.200000003f + z * p;          /*  1/5  */
    p = -0.333333343f + z * p;         /* -1/3  */
    float r = base + (t + t * (z * p));
    if (swap) r = 1.57079637f - r;
    if (x < 0.0f) r = 3.14159274f - r;
    return (y < 0.0f) ? -r : r;
}

/* Cube root for x >= 0: exp(log(x)/3) and one Newton step. */
GP_HD float gp_cbrtf(float x)
{
    if (x == 0.0f) return 0.0f;
    float y = gp_expf(gp_logf(x) / 3.0f);
    return y - (y * y * y - x) / (3.0f * (y * y));
}

/* Standard normal from two uniforms in (0, 1] (Box-Muller, cosine branch). */
GP_HD float gp_normal(float u1, float u2)
{
    float s, c;
    gp_sincos2pi(u2, &s, &c);
    return sqrtf(-2.0f * gp_logf(u1)) * c;
}

GP_HD float gp_rng_normal(gp_rng *r)
{
    float u1 = gp_rng_uniform(r);
    float u2 = gp_rng_uniform(r);
    return gp_normal(u1, u2);
}

/* ---- double precision (time sampling) ---------------------------------- */

#define GP_LN2D_HI 6.93147180369123816490e-01 /* ln 2 rounded to 32 bits */
#define GP_LN2D_LO 1.90821492927058770002e-10

/* 2 atanh(s) = log((1+s)/(1-s)) for |s| <= 0.1716 */
GP_HD double gp_atanh2_series(double s)
{
    double z = s * s;
    double p = 1.0 / 25.0;
    p = 1.0 / 23.0 + z * p;
    p = 1.0 / 21.0 + z * p;
    p = 1.0 / 19.0 + z * p;
    p = 1.0 / 17.0 + z * p;
    p = 1.0 / 15.0 + z * p;
    p = 1.0 / 13.0 + z * p;
    p = 1.0 / 11.0 + z * p;
    p = 1.0 / 9.0 + z * p;
    p = 1.0 / 7.0 + z * p;
    p = 1.0 / 5.0 + z * p;
    p = 1.0 / 3.0 + z * p;
    double t = s + s * (z * p);
    return t + t;
}

/* Natural log in FP64 for finite x > 0 (normal range). */
GP_HD double gp_log_d(double x)
{
    uint64_t ix = gp_d2u(x);
    if (x <= 0.0) return (x == 0.0) ? -INFINITY : (x - x) / (x - x);
    int k = (int)((ix >> 52) & 0x7ff) - 1023;
    double m = gp_u2d((ix & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        k += 1;
    }
    double f = m - 1.0;
    double lm = gp_atanh2_series(f / (2.0 + f));
    double dk = (double)k;
    return dk * GP_LN2D_HI + (dk * GP_LN2D_LO + lm);
}

/* log(1 + y) in FP64, y > -1; accurate for small |y|. */
GP_HD double gp_log1p_d(double y)
{
    if (y > -0.25 && y < 0.25) return gp_atanh2_series(y / (2.0 + y));
    return gp_log_d(1.0 + y);
}

#endif /* GP_MATH_H */
