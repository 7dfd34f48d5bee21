// crmath.hpp -- x**(1.0/3.0), sin and cos evaluated identically on the
// host and on gfx950, rounded to nearest.
//
// The reference evaluates h**(1.0/3.0) in Manning friction
// (rki_kernel_fused.F90:268), sin/cos for the tidal surface
// (rki_boundary.F90:96-100) and cos/sin for Coriolis
// (rki_kernel_fused.F90:212-213) with the platform libm, whose last-ulp
// rounding is platform specific (ROCm's ocml pow and glibc's pow disagree
// for a large fraction of inputs; glibc itself misrounds pow(x, 1/3) for
// ~7e-4 of inputs).  These routines use only IEEE +, -, *, /, fma and
// exact bit manipulation (pow13 also IEEE binary32 steps for its seed and
// its ln x term), in double-double where it matters, so
// they give the same bits on every IEEE machine and the correctly rounded
// result except within ~1e-26 relative of a rounding midpoint (~4e-24 for
// pow13).
//
// pow13: inverse cube root seeded by exponent splitting and a float fit,
// two division-free Newton steps, one double-double residual correction,
// then the factor exp(-(2^-54/3) ln x) because the double 1.0/3.0 is
// 1/3 - 2^-54/3 (ln x from a float polynomial; its error moves the result
// by < 4e-24 relative).
// sin/cos: Cody-Waite reduction by pi/2 held in three doubles, Taylor
// series to x^25 / x^26 in double-double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rk {
namespace crm {

struct dd {
    double hi, lo;
};

__host__ __device__ __forceinline__ double bits2d(uint64_t u) {
    union { uint64_t u; double d; } v;
    v.u = u;
    return v.d;
}
__host__ __device__ __forceinline__ uint64_t d2bits(double d) {
    union { uint64_t u; double d; } v;
    v.d = d;
    return v.u;
}

__host__ __device__ __forceinline__ dd two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return dd{s, (a - (s - bb)) + (b - bb)};
}
__host__ __device__ __forceinline__ dd fast_two_sum(double a, double b) {
    const double s = a + b;
    return dd{s, b - (s - a)};
}
__host__ __device__ __forceinline__ dd add(dd x, dd y) {
    const dd s = two_sum(x.hi, y.hi);
    const double e = s.lo + (x.lo + y.lo);
    return fast_two_sum(s.hi, e);
}
__host__ __device__ __forceinline__ dd mul(dd x, dd y) {
    const double p = x.hi * y.hi;
    double e = fma(x.hi, y.hi, -p);
    e = e + (x.hi * y.lo + x.lo * y.hi);
    return fast_two_sum(p, e);
}

// x**(1.0/3.0) for 2^-900 <= x < 2^1000 (below, the residual x - r0^3
// falls among subnormals and the last bit can be wrong).
//   x = m 2^e, m in [1, 2), e = 3q + r, y = m 2^r in [1, 8);
//   y^(-1/3): a degree-4 float fit of m^(-1/3) on [1, 2] (relative error
//   < 4.6e-5) times 2^(-r/3), then two division-free Newton steps in
//   double (error < 1e-16 before rounding);
//   r0 = y rr^2 2^q ~ x^(1/3), then one Newton step on the residual
//   x - r0^3 taken in double-double (error ~1e-31);
//   the double 1.0/3.0 is 1/3 - 2^-54/3: the factor exp(-(2^-54/3) ln x)
//   is applied as 1 - (2^-54/3) ln x, with ln x = e ln 2 + ln m and ln m
//   from a degree-8 float polynomial f Q(f), f = m - 1 (absolute error
//   < 2e-7, which moves the result by < 4e-24 relative).
// Float steps are IEEE binary32 operations (conversions round to nearest),
// so the sequence is as deterministic as the double steps.
__host__ __device__ __forceinline__ double pow13(double x) {
    const uint64_t u = d2bits(x);
    const uint32_t eb = (uint32_t)(u >> 52);   // biased exponent, x > 0
    const uint32_t w = eb + 2046u;             // 3 (q + 1023) + r
    const uint32_t qq = w / 3u;                // q + 1023
    const uint32_t r = w - 3u * qq;            // 0, 1, 2
    const uint64_t mant = u & 0x000FFFFFFFFFFFFFull;
    const double m = bits2d(mant | 0x3FF0000000000000ull);
    const double y = bits2d(mant | ((uint64_t)(1023u + r) << 52));
    const float mf = (float)m;
    float s = 0x1.c73184p-6f;
    s = fmaf(s, mf, -0x1.bb1becp-3f);
    s = fmaf(s, mf, 0x1.5e8accp-1f);
    s = fmaf(s, mf, -0x1.2992e0p+0f);
    s = fmaf(s, mf, 0x1.aa913ep+0f);
    s = s * (r == 0 ? 1.0f : (r == 1 ? 0x1.965feap-1f : 0x1.428a30p-1f));   // 2^(-r/3)
    double rr = (double)s;
    for (int it = 0; it < 2; ++it) {
        const double e = fma(-y, (rr * rr) * rr, 1.0);
        rr = fma(rr, e * (1.0 / 3.0), rr);
    }
    const double r0 = ((y * rr) * rr) * bits2d((uint64_t)qq << 52);
    const double p = r0 * r0;
    const double pl = fma(r0, r0, -p);
    const double t3 = p * r0;
    const double t3l = fma(p, r0, -t3);
    const double res = (x - t3) - fma(pl, r0, t3l);
    const double inv = ((rr * rr) * (1.0 / 3.0)) * bits2d((uint64_t)(3069u - 2u * qq) << 52);   // x^(-2/3) / 3
    const double corr = res * inv;
    const float f = (float)(m - 1.0);
    float lq = -0x1.124e52p-7f;
    lq = fmaf(lq, f, 0x1.644d52p-5f);
    lq = fmaf(lq, f, -0x1.b5a860p-4f);
    lq = fmaf(lq, f, 0x1.6a3d52p-3f);
    lq = fmaf(lq, f, -0x1.f53e4ep-3f);
    lq = fmaf(lq, f, 0x1.54b45ep-2f);
    lq = fmaf(lq, f, -0x1.fff898p-2f);
    lq = fmaf(lq, f, 0x1.fffff8p-1f);
    const float lm = lq * f;
    const double lnx = fma((double)((int)eb - 1023), 0x1.62e42fefa39efp-1, (double)lm);
    const double t = lnx * -0x1.5555555555555p-56;   // -(2^-54 / 3) ln x
    return r0 + fma(r0, t, corr);
}

// 1/n! as double-double, n = 0..26.
__host__ __device__ __forceinline__ dd inv_fact(int n) {
    switch (n) {
    case 0: return dd{1.0, 0.0};
    case 1: return dd{1.0, 0.0};
    case 2: return dd{0.5, 0.0};
    case 3: return dd{0x1.5555555555555p-3, 0x1.5555555555555p-57};
    case 4: return dd{0x1.5555555555555p-5, 0x1.5555555555555p-59};
    case 5: return dd{0x1.1111111111111p-7, 0x1.1111111111111p-63};
    case 6: return dd{0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65};
    case 7: return dd{0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73};
    case 8: return dd{0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76};
    case 9: return dd{0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73};
    case 10: return dd{0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76};
    case 11: return dd{0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80};
    case 12: return dd{0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83};
    case 13: return dd{0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87};
    case 14: return dd{0x1.93974a8c07c9dp-37, 0x1.05d6f8a2efd1fp-92};
    case 15: return dd{0x1.ae7f3e733b81fp-41, 0x1.1d8656b0ee8cbp-97};
    case 16: return dd{0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101};
    case 17: return dd{0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103};
    case 18: return dd{0x1.6827863b97d97p-53, 0x1.eec01221a8b0bp-107};
    case 19: return dd{0x1.2f49b46814157p-57, 0x1.2650f61dbdcb4p-112};
    case 20: return dd{0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120};
    case 21: return dd{0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120};
    case 22: return dd{0x1.0ce396db7f853p-70, -0x1.aebcdbd20331cp-124};
    case 23: return dd{0x1.761b41316381ap-75, -0x1.3423c7d91404fp-130};
    case 24: return dd{0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135};
    case 25: return dd{0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139};
    default: return dd{0x1.88e85fc6a4e5ap-89, -0x1.71c37ebd16540p-143};
    }
}

// x - k*pi/2 in double-double, |k| < 2^30.
__host__ __device__ __forceinline__ dd reduce(double x, double k) {
    const double P1 = 0x1.921fb54442d18p+0, P2 = 0x1.1a62633145c07p-54, P3 = -0x1.f1976b7ed8fbcp-110;
    const double p = k * P1;
    const double pl = fma(k, P1, -p);
    dd r = two_sum(x, -p);
    r = fast_two_sum(r.hi, r.lo - pl);
    const double q = k * P2;
    const double ql = fma(k, P2, -q);
    r = add(r, dd{-q, -ql});
    return add(r, dd{-(k * P3), 0.0});
}

// Taylor series of sin (odd = 1) or cos (odd = 0); term n has sign
// (-1)^((n - odd) / 2).
__host__ __device__ __forceinline__ dd series(dd r, int odd) {
    const dd r2 = mul(r, r);
    int n = odd ? 25 : 26;
    dd poly = inv_fact(n);
    if (((n - odd) / 2) % 2 == 1) { poly.hi = -poly.hi; poly.lo = -poly.lo; }
    for (n = n - 2; n >= 0; n -= 2) {
        dd c = inv_fact(n);
        if (((n - odd) / 2) % 2 == 1) { c.hi = -c.hi; c.lo = -c.lo; }
        poly = add(c, mul(r2, poly));
    }
    return odd ? mul(r, poly) : poly;
}

__host__ __device__ __forceinline__ double sincos_one(double x, int want_cos) {
    const double k = rint(x * 0x1.45f306dc9c883p-1);  // x * 2/pi
    const dd r = reduce(x, k);
    const long long ki = (long long)k;
    int quad = (int)(((ki % 4) + 4) % 4);
    if (want_cos) quad = (quad + 1) % 4;
    const dd v = series(r, (quad % 2 == 0) ? 1 : 0);
    return (quad >= 2) ? -v.hi : v.hi;
}

__host__ __device__ __forceinline__ double sin(double x) { return sincos_one(x, 0); }
__host__ __device__ __forceinline__ double cos(double x) { return sincos_one(x, 1); }

}  // namespace crm
}  // namespace rk
