/*
 * rk_crmath.h -- libm-independent transcendental functions used by the
 * oracle: x**(1.0/3.0) (Manning friction, rki_kernel_fused.F90:268),
 * sin and cos (tidal surface rki_boundary.F90:96-100, Coriolis
 * rki_kernel_fused.F90:212-213).
 *
 * The reference calls the platform libm (glibc, CUDA libdevice, ...), whose
 * last-ulp rounding differs between platforms: glibc's pow(x, 1/3) is not
 * correctly rounded for about 7e-4 of inputs and its sin for about 1.5e-3.
 * These functions evaluate in double-double arithmetic from IEEE basic
 * operations and fma only (x**(1/3) also binary32 ones for its seed and its
 * ln x term), so they are deterministic on every IEEE machine and round
 * correctly except in cases closer to a rounding midpoint than about 1e-26
 * relative (4e-24 for x**(1/3)).
 *
 * Methods (own implementation):
 *   inverse cube root seeded by exponent splitting and a float fit, two
 *   division-free Newton steps, one double-double residual correction;
 *   the exponent 1.0/3.0 is the double 1/3 - 2^-54/3, applied as the
 *   factor 1 - (2^-54/3) ln x, ln x from a float polynomial.
 *   sin/cos by Cody-Waite reduction with pi/2 in three doubles and Taylor
 *   series (through x^25 / x^26) in double-double.
 * TEST INFRASTRUCTURE: part of the oracle.
 */
#ifndef RK_CRMATH_H
#define RK_CRMATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct { double hi, lo; } crm_dd;

static inline double crm_bits2d(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
static inline uint64_t crm_d2bits(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }

static inline crm_dd crm_two_sum(double a, double b)
{
    crm_dd r;
    double s = a + b;
    double bb = s - a;
    r.hi = s;
    r.lo = (a - (s - bb)) + (b - bb);
    return r;
}

static inline crm_dd crm_fast_two_sum(double a, double b)
{
    crm_dd r;
    double s = a + b;
    r.hi = s;
    r.lo = b - (s - a);
    return r;
}

static inline crm_dd crm_dd_add(crm_dd x, crm_dd y)
{
    crm_dd s = crm_two_sum(x.hi, y.hi);
    double e = s.lo + (x.lo + y.lo);
    return crm_fast_two_sum(s.hi, e);
}

static inline crm_dd crm_dd_mul(crm_dd x, crm_dd y)
{
    double p = x.hi * y.hi;
    double e = fma(x.hi, y.hi, -p);
    e = e + (x.hi * y.lo + x.lo * y.hi);
    return crm_fast_two_sum(p, e);
}

/* x**(1.0/3.0) for 2^-900 <= x < 2^1000, rounded to nearest.
 *   x = m 2^e, m in [1, 2), e = 3q + r, y = m 2^r in [1, 8);
 *   y^(-1/3): degree-4 float fit of m^(-1/3) on [1, 2] (relative error
 *   < 4.6e-5) times 2^(-r/3), two division-free Newton steps in double;
 *   r0 = y rr^2 2^q, then one Newton step on the residual x - r0^3 in
 *   double-double;
 *   exponent 1/3 - 2^-54/3: times 1 - (2^-54/3) ln x, ln x = e ln 2 + ln m,
 *   ln m = f Q(f), f = m - 1, Q a degree-7 float polynomial (absolute
 *   error < 2e-7, moving the result by < 4e-24 relative). */
static inline double crm_pow13(double x)
{
    uint64_t u = crm_d2bits(x);
    uint32_t eb = (uint32_t)(u >> 52);   /* biased exponent, x > 0 */
    uint32_t w = eb + 2046u;             /* 3 (q + 1023) + r */
    uint32_t qq = w / 3u;                /* q + 1023 */
    uint32_t r = w - 3u * qq;            /* 0, 1, 2 */
    uint64_t mant = u & 0x000FFFFFFFFFFFFFull;
    double m = crm_bits2d(mant | 0x3FF0000000000000ull);                 /* [1, 2) */
    double y = crm_bits2d(mant | ((uint64_t)(1023u + r) << 52));         /* [1, 8) */
    float mf = (float)m;
    float s = 0x1.c73184p-6f;
    s = fmaf(s, mf, -0x1.bb1becp-3f);
    s = fmaf(s, mf, 0x1.5e8accp-1f);
    s = fmaf(s, mf, -0x1.2992e0p+0f);
    s = fmaf(s, mf, 0x1.aa913ep+0f);
    s = s * (r == 0 ? 1.0f : (r == 1 ? 0x1.965feap-1f : 0x1.428a30p-1f)); /* 2^(-r/3) */
    double rr = (double)s;
    for (int it = 0; it < 2; ++it) {
        double e = fma(-y, (rr * rr) * rr, 1.0);
        rr = fma(rr, e * (1.0 / 3.0), rr);
    }
    double r0 = ((y * rr) * rr) * crm_bits2d((uint64_t)qq << 52); /* ~x^(1/3) */
    /* residual x - r0^3 in double-double, Newton correction res/(3 r0^2) */
    double p = r0 * r0;
    double pl = fma(r0, r0, -p);
    double t3 = p * r0;
    double t3l = fma(p, r0, -t3);
    double res = (x - t3) - fma(pl, r0, t3l);
    double inv = ((rr * rr) * (1.0 / 3.0)) * crm_bits2d((uint64_t)(3069u - 2u * qq) << 52);
    double corr = res * inv;
    /* ln x, then the factor 1 - (2^-54/3) ln x */
    float f = (float)(m - 1.0);
    float lq = -0x1.124e52p-7f;
    lq = fmaf(lq, f, 0x1.644d52p-5f);
    lq = fmaf(lq, f, -0x1.b5a860p-4f);
    lq = fmaf(lq, f, 0x1.6a3d52p-3f);
    lq = fmaf(lq, f, -0x1.f53e4ep-3f);
    lq = fmaf(lq, f, 0x1.54b45ep-2f);
    lq = fmaf(lq, f, -0x1.fff898p-2f);
    lq = fmaf(lq, f, 0x1.fffff8p-1f);
    float lm = lq * f;
    double lnx = fma((double)((int)eb - 1023), 0x1.62e42fefa39efp-1, (double)lm);
    double t = lnx * -0x1.5555555555555p-56; /* -(2^-54 / 3) ln x */
    return r0 + fma(r0, t, corr);
}

/* 1/n! as double-double, n = 0..26 */
static const double CRM_INVFACT[27][2] = {
    {1.0, 0.0},
    {1.0, 0.0},
    {0.5, 0.0},
    {0x1.5555555555555p-3, 0x1.5555555555555p-57},
    {0x1.5555555555555p-5, 0x1.5555555555555p-59},
    {0x1.1111111111111p-7, 0x1.1111111111111p-63},
    {0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65},
    {0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73},
    {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76},
    {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73},
    {0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76},
    {0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80},
    {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83},
    {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87},
    {0x1.93974a8c07c9dp-37, 0x1.05d6f8a2efd1fp-92},
    {0x1.ae7f3e733b81fp-41, 0x1.1d8656b0ee8cbp-97},
    {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101},
    {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103},
    {0x1.6827863b97d97p-53, 0x1.eec01221a8b0bp-107},
    {0x1.2f49b46814157p-57, 0x1.2650f61dbdcb4p-112},
    {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120},
    {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120},
    {0x1.0ce396db7f853p-70, -0x1.aebcdbd20331cp-124},
    {0x1.761b41316381ap-75, -0x1.3423c7d91404fp-130},
    {0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135},
    {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139},
    {0x1.88e85fc6a4e5ap-89, -0x1.71c37ebd16540p-143},
};

/* pi/2 = PIO2_1 + PIO2_2 + PIO2_3 (each a double) */
#define CRM_PIO2_1 0x1.921fb54442d18p+0
#define CRM_PIO2_2 0x1.1a62633145c07p-54
#define CRM_PIO2_3 -0x1.f1976b7ed8fbcp-110
#define CRM_2_OVER_PI 0x1.45f306dc9c883p-1

/* x - k pi/2 in double-double, |k| < 2^30 */
static inline crm_dd crm_reduce(double x, double k)
{
    double p = k * CRM_PIO2_1;
    double pl = fma(k, CRM_PIO2_1, -p);
    crm_dd r = crm_two_sum(x, -p);
    r = crm_fast_two_sum(r.hi, r.lo - pl);
    double q = k * CRM_PIO2_2;
    double ql = fma(k, CRM_PIO2_2, -q);
    crm_dd mq; mq.hi = -q; mq.lo = -ql;
    r = crm_dd_add(r, mq);
    crm_dd mk; mk.hi = -(k * CRM_PIO2_3); mk.lo = 0.0;
    return crm_dd_add(r, mk);
}

/* sin (odd = 1) or cos (odd = 0) series of a reduced argument */
static inline crm_dd crm_series(crm_dd r, int odd)
{
    crm_dd r2 = crm_dd_mul(r, r);
    int n = odd ? 25 : 26;
    /* term n carries the sign (-1)^((n - odd) / 2) */
    crm_dd poly; poly.hi = CRM_INVFACT[n][0]; poly.lo = CRM_INVFACT[n][1];
    if (((n - odd) / 2) % 2 == 1) { poly.hi = -poly.hi; poly.lo = -poly.lo; }
    for (n = n - 2; n >= 0; n -= 2) {
        crm_dd c; c.hi = CRM_INVFACT[n][0]; c.lo = CRM_INVFACT[n][1];
        if (((n - odd) / 2) % 2 == 1) { c.hi = -c.hi; c.lo = -c.lo; }
        poly = crm_dd_add(c, crm_dd_mul(r2, poly));
    }
    return odd ? crm_dd_mul(r, poly) : poly;
}

static inline double crm_sincos(double x, int want_cos)
{
    double k = rint(x * CRM_2_OVER_PI);
    crm_dd r = crm_reduce(x, k);
    long long ki = (long long)k;
    int quad = (int)(((ki % 4) + 4) % 4);
    if (want_cos) quad = (quad + 1) % 4;
    /* sin(r + quad pi/2): 0 sin r, 1 cos r, 2 -sin r, 3 -cos r */
    crm_dd v = crm_series(r, (quad % 2 == 0) ? 1 : 0);
    double res = v.hi;
    return (quad >= 2) ? -res : res;
}

static inline double crm_sin(double x) { return crm_sincos(x, 0); }
static inline double crm_cos(double x) { return crm_sincos(x, 1); }

#endif
