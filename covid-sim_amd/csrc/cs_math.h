// cs_math.h -- device arithmetic for the infection step: logarithm,
// exponential, square root, counter-based random streams and the samplers
// built on them.  Every routine uses only IEEE-rounded +,-,*,/, fma and
// exact scaling, so a given input produces the same bits on any IEEE host
// or device; the step's random decisions are therefore reproducible.
//
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  A
// stream is fixed by the realisation seed (the Philox key) and a counter
// prefix {timestep | pass << 16, entity, sub}; the fourth counter word
// counts blocks.  Each block gives two uniforms
// ((hi:lo) >> 11 + 0.5) * 2^-53 in (0,1).  Streams replace CovidSim's
// per-thread L'Ecuyer generators (src/Rand.cpp:28-47): a decision's numbers
// depend only on what the decision is about, not on which thread makes it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cs {

__device__ __forceinline__ double bits_to_double(uint64_t b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ uint64_t double_to_bits(double d) { return (uint64_t)__double_as_longlong(d); }

__device__ __forceinline__ double exp2i(int k)  // 2^k, -1022 <= k <= 1023
{
    return bits_to_double((uint64_t)(k + 1023) << 52);
}

// correctly rounded square root: hardware estimate, then the neighbour whose
// square is closest to x (exact residuals through fma)
__device__ __forceinline__ double csqrt(double x)
{
    if (!(x > 0.0)) return x == 0.0 ? x : __builtin_nan("");
    double y = __builtin_sqrt(x);
    double lo = bits_to_double(double_to_bits(y) - 1), hi = bits_to_double(double_to_bits(y) + 1);
    double r = __builtin_fabs(__builtin_fma(-y, y, x));
    double rl = __builtin_fabs(__builtin_fma(-lo, lo, x));
    double rh = __builtin_fabs(__builtin_fma(-hi, hi, x));
    if (rl < r) { y = lo; r = rl; }
    if (rh < r) y = hi;
    return y;
}

__device__ __noinline__ double clog(double x)
{
    if (x != x) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    uint64_t b = double_to_bits(x);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) {
        x *= 18014398509481984.0;
        b = double_to_bits(x);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    b = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m = bits_to_double(b);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double p = 1.0 / 23.0;
    p = p * s2 + 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    double lm = 2.0 * s * p;
    double de = (double)e;
    return de * 6.93147180369123816490e-01 + (lm + de * 1.90821492927058770002e-10);
}

__device__ __noinline__ double cexp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return __builtin_inf();
    if (x < -745.1332191019412) return 0.0;
    double kf = __builtin_floor(x * 1.4426950408889634074 + 0.5);
    int k = (int)kf;
    double r = (x - kf * 6.93147180369123816490e-01) - kf * 1.90821492927058770002e-10;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    if (k > 1023) return p * exp2i(1023) * exp2i(k - 1023);
    if (k < -1021) return p * exp2i(k + 1000) * exp2i(-1000);
    return p * exp2i(k);
}

__device__ __forceinline__ double pow_int(double q, int32_t n)
{
    double r = 1.0, b = q;
    while (n > 0) {
        if (n & 1) r *= b;
        b *= b;
        n >>= 1;
    }
    return r;
}

__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t h0 = __umulhi(0xD2511F53u, c0), l0 = 0xD2511F53u * c0;
        uint32_t h1 = __umulhi(0xCD9E8D57u, c2), l1 = 0xCD9E8D57u * c2;
        uint32_t n0 = h1 ^ c1 ^ k0, n2 = h0 ^ c3 ^ k1;
        c0 = n0; c1 = l1; c2 = n2; c3 = l0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo)
{
    uint64_t v = (((uint64_t)hi << 32) | lo) >> 11;
    return ((double)v + 0.5) * (1.0 / 9007199254740992.0);
}

// Exact 64.64 fixed-point accumulator for non-negative doubles below 2^63:
// additions are exact integer adds, so a sum does not depend on the order
// of its terms; fx_to_double rounds the total once per component.
struct Fx { unsigned long long hi, lo; };
__host__ __device__ __forceinline__ Fx fx_from(double v)
{
    const double f = floor(v);
    Fx r;
    r.hi = (unsigned long long)f;
    r.lo = (unsigned long long)((v - f) * 18446744073709551616.0);  // 2^64
    return r;
}
__host__ __device__ __forceinline__ Fx fx_add(Fx a, Fx b)
{
    Fx r;
    r.lo = a.lo + b.lo;
    r.hi = a.hi + b.hi + (r.lo < a.lo ? 1ull : 0ull);
    return r;
}
__host__ __device__ __forceinline__ double fx_to_double(Fx a) { return (double)a.hi + (double)a.lo * 5.42101086242752217e-20; /* 2^-64 */ }

struct Stream {
    uint32_t k0, k1, c0, c1, c2, c3;
    double b0, b1;
    int idx;
    __device__ __forceinline__ Stream(uint64_t key, uint32_t a, uint32_t b, uint32_t c)
        : k0((uint32_t)key), k1((uint32_t)(key >> 32)), c0(a), c1(b), c2(c), c3(0), b0(0), b1(0), idx(2) {}
    __device__ __forceinline__ double next()
    {
        if (idx == 2) {
            uint32_t o[4];
            philox4x32_10(c0, c1, c2, c3, k0, k1, o);
            c3++;
            b0 = u53(o[0], o[1]);
            b1 = u53(o[2], o[3]);
            idx = 0;
        }
        return (idx++ == 0) ? b0 : b1;
    }
};

// pass identifiers (high half of counter word 0)
enum { PS_HH = 1, PS_PW_SAMPLE, PS_PW_ACCEPT, PS_PB_SAMPLE, PS_PB_ACCEPT, PS_SP_COUNT,
       PS_SP_DRAW, PS_INFECT, PS_INCUB, PS_CASE, PS_RECOVER, PS_HOLIDAY, PS_INIT, PS_SEED, PS_ITC,
       PS_HQ, PS_ISO, PS_PLACE, PS_POPGEN };
// pass identifiers inside the event order key
enum { KP_HH = 0, KP_PW = 1, KP_PB = 6, KP_SP = 11 };

__device__ __forceinline__ uint32_t cw0(unsigned ts, int pass) { return (uint32_t)ts | ((uint32_t)pass << 16); }

__device__ __forceinline__ double fc_corr(int32_t k)
{
    switch (k) {
    case 0: return 0.08106146679532726;
    case 1: return 0.04134069595540929;
    case 2: return 0.02767792568499834;
    case 3: return 0.02079067210376509;
    case 4: return 0.01664469118982119;
    case 5: return 0.01387612882307075;
    case 6: return 0.01189670994589177;
    case 7: return 0.01041126526197209;
    case 8: return 0.009255462182712733;
    case 9: return 0.008330563433362871;
    default: break;
    }
    double ik = 1.0 / (double)(k + 1);
    double ik2 = ik * ik;
    return (1.0 / 12.0 - (1.0 / 360.0 - ik2 / 1260.0) * ik2) * ik;
}

// BTRS's exact acceptance test, taken only when the squeeze fails: kept out
// of line so that its logarithms do not add to the caller's registers
__device__ __noinline__ bool btrs_accept(double v, double alpha, double a, double b, double us, double m, double r, int32_t n, double kf)
{
    double lv = clog(v * alpha / (a / (us * us) + b));
    double ub = (m + 0.5) * clog((m + 1.0) / (r * ((double)n - m + 1.0)))
              + ((double)n + 1.0) * clog(((double)n - m + 1.0) / ((double)n - kf + 1.0))
              + (kf + 0.5) * clog(r * ((double)n - kf + 1.0) / (kf + 1.0))
              + fc_corr((int32_t)m) + fc_corr(n - (int32_t)m)
              - fc_corr((int32_t)kf) - fc_corr(n - (int32_t)kf);
    return lv <= ub;
}

// Binomial(n, p): inversion when n*min(p,1-p) < 10, else BTRS (Hormann 1993)
__device__ __forceinline__ int32_t binomial(Stream& s, int32_t n, double pp)
{
    if (n <= 0 || pp <= 0.0) return 0;
    if (pp >= 1.0) return n;
    double p = pp < 0.5 ? pp : 1.0 - pp;
    double q = 1.0 - p;
    int32_t x;
    if ((double)n * p < 10.0) {
        double f = pow_int(q, n);
        double r = p / q;
        double u = s.next();
        x = 0;
        while (u > f) {
            u -= f;
            x++;
            if (x >= n) { x = n; break; }
            f *= r * (double)(n - x + 1) / (double)x;
        }
    } else {
        double spq = csqrt((double)n * p * q);
        double b = 1.15 + 2.53 * spq;
        double a = -0.0873 + 0.0248 * b + 0.01 * p;
        double c = (double)n * p + 0.5;
        double vr = 0.92 - 4.2 / b;
        double r = p / q;
        double alpha = (2.83 + 5.1 / b) * spq;
        double m = __builtin_floor((double)(n + 1) * p);
        for (;;) {
            double u = s.next() - 0.5;
            double v = s.next();
            double us = 0.5 - __builtin_fabs(u);
            double kf = __builtin_floor((2.0 * a / us + b) * u + c);
            if (kf < 0.0 || kf > (double)n) continue;
            if (us >= 0.07 && v <= vr) { x = (int32_t)kf; break; }
            if (btrs_accept(v, alpha, a, b, us, m, r, n, kf)) { x = (int32_t)kf; break; }
        }
    }
    return pp > 0.5 ? n - x : x;
}

__device__ __forceinline__ double log_fact(int32_t k)
{
    double kk = (double)k;
    return fc_corr(k) + (kk + 0.5) * clog(kk + 1.0) - (kk + 1.0) + 0.91893853320467274178;
}

// Poisson(mu): inversion for mu < 10, else PTRS (Hormann 1993)
__device__ __forceinline__ int32_t poisson(Stream& s, double mu)
{
    if (!(mu > 0.0)) return 0;
    if (mu < 10.0) {
        double pk = cexp(-mu);
        double u = s.next();
        int32_t x = 0;
        while (u > pk) {
            u -= pk;
            x++;
            if (x > 10000) break;
            pk *= mu / (double)x;
        }
        return x;
    }
    double slam = csqrt(mu);
    double loglam = clog(mu);
    double b = 0.931 + 2.53 * slam;
    double a = -0.059 + 0.02483 * b;
    double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        double u = s.next() - 0.5;
        double v = s.next();
        double us = 0.5 - __builtin_fabs(u);
        double kf = __builtin_floor((2.0 * a / us + b) * u + mu + 0.43);
        if (us >= 0.07 && v <= vr) return (int32_t)kf;
        if (kf < 0.0 || (us < 0.013 && v > us)) continue;
        if (clog(v) + clog(invalpha) - clog(a / (us * us) + b) <= -mu + kf * loglam - log_fact((int32_t)kf))
            return (int32_t)kf;
    }
}

__device__ __forceinline__ double normal(Stream& s)
{
    for (;;) {
        double a = 2.0 * s.next() - 1.0;
        double b = 2.0 * s.next() - 1.0;
        double q = a * a + b * b;
        if (q >= 1.0 || q == 0.0) continue;
        return a * csqrt(-2.0 * clog(q) / q);
    }
}

// gen_gamma_mt(beta = rate, alpha = shape): Marsaglia-Tsang with the
// acceptance test of src/Rand.cpp:1025-1086
__device__ __forceinline__ double gamma(Stream& s, double beta, double alpha)
{
    double al = alpha >= 1.0 ? alpha : alpha + 1.0;
    double d = al - 1.0 / 3.0;
    double c = 1.0 / (3.0 * csqrt(d));
    double u, z, v;
    for (;;) {
        u = s.next();
        z = normal(s);
        v = 1.0 + z * c;
        v = v * v * v;
        if (z <= -1.0 / c) continue;
        if (clog(u) >= 0.5 * z * z + d - d * v + d * clog(v)) continue;
        break;
    }
    double g = (beta != 1.0) ? (d * v) / beta : d * v;
    if (alpha < 1.0) g *= cexp(clog(s.next()) / alpha);
    return g;
}

}  // namespace cs
