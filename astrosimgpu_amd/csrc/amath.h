// amath: FP64 exp, log, sin and cos that compile to the same bits on the host
// (g++ or clang) and on gfx950 (hipcc).
//
// The reference calls the C library's exp/log/cos/sin (glibc 2.35 on x86-64,
// where the FMA-compiled variants are selected at run time). Its results
// decide spike times, so a device path that is to reproduce a host raster bit
// for bit needs the same function on both sides:
//
//  * asg_exp follows glibc 2.35's exp (Szabolcs Nagy's table-driven method:
//    x = k ln2/128 + r, exp(x) = 2^(k/128) exp(r), a degree-5 polynomial for
//    exp(r) - 1, 128-entry table of 2^(k/128) as scale + tail), evaluated with
//    the fused multiply-adds the FMA-compiled glibc variant performs. It
//    equals glibc's exp on every input tested (1.6e8 draws over [-745, 709]).
//  * asg_log follows glibc 2.35's log (same author: x = 2^k z, 128
//    subintervals of z with (1/c, log c) pairs, a degree-5 polynomial in
//    r = z/c - 1, and a degree-11 polynomial path for x within 2^-4 of 1),
//    again with the FMA variant's fusion pattern. Equal to glibc's log on all
//    inputs tested.
//  * asg_sin / asg_cos follow glibc 2.35's sin/cos (the IBM Accurate
//    Mathematical Library's table-driven method, s_sin.c), again with the
//    FMA variant's fusion pattern; the sin/cos table of k/128 is recomputed
//    from mpmath. Equal to glibc's cos on every Box-Muller angle tested
//    (5e7 draws of 2 pi u) and to glibc's sin on the inputs tested.
//
// Every function here is written with explicit fma() where one is intended
// and must be compiled with -ffp-contract=off everywhere else.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ASG_HD __host__ __device__ inline
#define ASG_TABLE static constexpr __device__ const
#else
#include <cmath>
#define ASG_HD inline
#define ASG_TABLE static constexpr const
#endif

#include "amath_tables.h"

namespace asg {

ASG_HD std::uint64_t as_u64(double x) { return __builtin_bit_cast(std::uint64_t, x); }
ASG_HD double as_f64(std::uint64_t u) { return __builtin_bit_cast(double, u); }
ASG_HD double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

// ---------------------------------------------------------------- exp ----

ASG_HD double exp_specialcase(double tmp, std::uint64_t sbits, std::uint64_t ki) {
    if ((ki & 0x80000000ULL) == 0) {
        // k > 0: the scale's exponent may have overflowed by <= 460.
        sbits -= 1009ULL << 52;
        const double scale = as_f64(sbits);
        return 0x1p1009 * fma_(scale, tmp, scale);
    }
    // k < 0: result near or in the subnormal range. The product scale*tmp
    // feeds two sums in different blocks, so it is not fused here.
    sbits += 1022ULL << 52;
    const double scale = as_f64(sbits);
    const double st = scale * tmp;
    double y = scale + st;
    if (y < 1.0) {
        double lo = scale - y + st;
        const double hi = 1.0 + y;
        lo = 1.0 - hi + y + lo;
        y = (hi + lo) - 1.0;
        if (y == 0.0) {
            y = 0.0;
        }
    }
    return 0x1p-1022 * y;
}

// exp's 2 KiB table as read by asg_exp: the table itself (indexed in place,
// on the host and the device), or a copy at a pointer (for instance in a
// workgroup's LDS, where lookups on a dependency chain wait less than on the
// global copy). Host code must index the table in place: the table is a
// device variable, and a pointer to it taken in host code points at its
// host-side stand-in, which holds zeros.
struct ExpTabDirect {
    ASG_HD std::uint64_t operator()(std::uint64_t i) const { return ASG_EXP_TAB[i]; }
};
struct ExpTabAt {
    const std::uint64_t* p;
    ASG_HD std::uint64_t operator()(std::uint64_t i) const { return p[i]; }
};

template <class Tab>
ASG_HD double asg_exp_with(double x, Tab tab) {
    constexpr double kInvLn2N = 0x1.71547652b82fep0 * 128.0;
    constexpr double kNegLn2hiN = -0x1.62e42fefa0000p-8;
    constexpr double kNegLn2loN = -0x1.cf79abc9e3b3ap-47;
    constexpr double kShift = 0x1.8p52;
    constexpr double C2 = 0x1.ffffffffffdbdp-2;
    constexpr double C3 = 0x1.555555555543cp-3;
    constexpr double C4 = 0x1.55555cf172b91p-5;
    constexpr double C5 = 0x1.1111167a4d017p-7;

    std::uint32_t abstop = static_cast<std::uint32_t>(as_u64(x) >> 52) & 0x7ff;
    constexpr std::uint32_t kTiny = 0x3c9;   // top12(0x1p-54)
    constexpr std::uint32_t kBig = 0x408;    // top12(512.0)
    constexpr std::uint32_t kHuge = 0x409;   // top12(1024.0)
    if (abstop - kTiny >= kBig - kTiny) {
        if (abstop - kTiny >= 0x80000000u) {
            return 1.0 + x;  // |x| < 2^-54, including 0
        }
        if (abstop >= kHuge) {
            if (as_u64(x) == 0xfff0000000000000ULL) {
                return 0.0;
            }
            if (abstop >= 0x7ff) {
                return 1.0 + x;
            }
            return (as_u64(x) >> 63) ? 0.0 : __builtin_inf();
        }
        abstop = 0;  // 512 <= |x| < 1024: handled by the special case below
    }
    double kd = fma_(kInvLn2N, x, kShift);
    const std::uint64_t ki = as_u64(kd);
    kd -= kShift;
    const double r = fma_(kd, kNegLn2loN, fma_(kd, kNegLn2hiN, x));
    const std::uint64_t idx = 2 * (ki % 128);
    const std::uint64_t top = ki << 45;
    const double tail = as_f64(tab(idx));
    const std::uint64_t sbits = tab(idx + 1) + top;
    const double r2 = r * r;
    const double tmp = fma_(r2 * r2, fma_(r, C5, C4), fma_(r2, fma_(r, C3, C2), tail + r));
    if (abstop == 0) {
        return exp_specialcase(tmp, sbits, ki);
    }
    const double scale = as_f64(sbits);
    return fma_(scale, tmp, scale);
}

ASG_HD double asg_exp(double x) { return asg_exp_with(x, ExpTabDirect{}); }

// ---------------------------------------------------------------- log ----

ASG_HD double asg_log(double x) {
    constexpr double A0 = -0x1.0000000000001p-1, A1 = 0x1.555555551305bp-2,
                     A2 = -0x1.fffffffeb459p-3, A3 = 0x1.999b324f10111p-3,
                     A4 = -0x1.55575e506c89fp-3;
    constexpr double B0 = -0x1p-1, B1 = 0x1.5555555555577p-2, B2 = -0x1.ffffffffffdcbp-3,
                     B3 = 0x1.999999995dd0cp-3, B4 = -0x1.55555556745a7p-3,
                     B5 = 0x1.24924a344de3p-3, B6 = -0x1.fffffa4423d65p-4,
                     B7 = 0x1.c7184282ad6cap-4, B8 = -0x1.999eb43b068ffp-4,
                     B9 = 0x1.78182f7afd085p-4, B10 = -0x1.5521375d145cdp-4;
    constexpr double kLn2hi = 0x1.62e42fefa3800p-1;
    constexpr double kLn2lo = 0x1.ef35793c76730p-45;
    constexpr std::uint64_t kOff = 0x3fe6000000000000ULL;
    constexpr std::uint64_t kLo = 0x3fee000000000000ULL;  // bits(1 - 0x1p-4)
    constexpr std::uint64_t kHi = 0x3ff1090000000000ULL;  // bits(1 + 0x1.09p-4)

    std::uint64_t ix = as_u64(x);
    if (ix - kLo < kHi - kLo) {
        if (ix == 0x3ff0000000000000ULL) {
            return 0.0;
        }
        const double r = x - 1.0;
        const double r2 = r * r;
        const double r3 = r * r2;
        const double p3 = fma_(r3, B10, fma_(r2, B9, fma_(r, B8, B7)));
        const double p2 = fma_(r3, p3, fma_(r2, B6, fma_(r, B5, B4)));
        const double p1 = fma_(r3, p2, fma_(r2, B3, fma_(r, B2, B1)));
        double w = r * 0x1p27;
        const double rhi = r + w - w;
        const double rlo = r - rhi;
        w = rhi * rhi * B0;
        const double hi = r + w;
        double lo = r - hi + w;
        lo = fma_(B0 * rlo, rhi + r, lo);
        const double y = fma_(r3, p1, lo);
        return y + hi;
    }
    const std::uint32_t top = static_cast<std::uint32_t>(ix >> 48);
    if (top - 0x0010u >= 0x7ff0u - 0x0010u) {
        if ((ix << 1) == 0) {
            return -__builtin_inf();
        }
        if (ix == 0x7ff0000000000000ULL) {
            return x;
        }
        if ((top & 0x8000u) || (top & 0x7ff0u) == 0x7ff0u) {
            return __builtin_nan("");
        }
        ix = as_u64(x * 0x1p52);  // subnormal: normalise
        ix -= 52ULL << 52;
    }
    const std::uint64_t tmp = ix - kOff;
    const int i = static_cast<int>((tmp >> 45) % 128);
    const int k = static_cast<int>(static_cast<std::int64_t>(tmp) >> 52);
    const std::uint64_t iz = ix - (tmp & (0xfffULL << 52));
    const double invc = ASG_LOG_TAB[2 * i];
    const double logc = ASG_LOG_TAB[2 * i + 1];
    const double z = as_f64(iz);
    const double r = fma_(z, invc, -1.0);
    const double kd = static_cast<double>(k);
    const double w = fma_(kd, kLn2hi, logc);
    const double hi = w + r;
    const double lo = fma_(kd, kLn2lo, (w - hi) + r);
    const double r2 = r * r;
    const double q = fma_(r2, fma_(r, A4, A3), fma_(r, A2, A1));
    const double t1 = fma_(r2, A0, lo);
    const double t2 = fma_(r * r2, q, t1);
    return t2 + hi;
}

// ------------------------------------------------------- sin and cos ----
//
// glibc 2.35's sin and cos for |x| < 105414350 (sysdeps/ieee754/dbl-64/
// s_sin.c, the IBM Accurate Mathematical Library method as simplified in
// glibc 2.28), restated:
//  * |x| < 2^-27 (cos) or 2^-26 (sin): the first Taylor term;
//  * |x| < 0.85546875: x = k/128 + r with k = nearest multiple (found by
//    adding 1.5 * 2^45, whose ulp is 1/128), then
//    sin/cos(x) from the double-double table of sin/cos(k/128) and short
//    polynomials in r (sn3, sn5; cs2, cs4, cs6); |x| < 0.126 uses a degree-11
//    odd Taylor-like polynomial (s1..s5) instead, with a first-order
//    correction for the argument's low part;
//  * |x| < 2.426265: the cofunction of pi/2 - |x|, carried as a double-double;
//  * larger: Cody-Waite reduction by multiples of pi/2 in four parts
//    (mp1, mp2, pp3, pp4) to a double-double argument and its quadrant.
// The x86-64 C library selects the variant compiled with -mfma at run time,
// in which GCC contracts each a*b + c of the C source into one fused
// multiply-add (each product whose every use is an add or subtract); the
// fma_() calls below reproduce those contractions. On 5e7 Box-Muller angles
// 2 pi u, u on the 53-bit draw grid, this equals the host C library's cos
// bit for bit. |x| >= 105414350 (glibc's Payne-Hanek range) is outside the
// domain here and returns NaN; the step only evaluates [0, 2 pi).

struct SinCosConst {
    static constexpr double sn3 = -1.66666666666664880952546298448555E-01;
    static constexpr double sn5 = 8.33333214285722277379541354343671E-03;
    static constexpr double cs2 = 4.99999999999999999999950396842453E-01;
    static constexpr double cs4 = -4.16666666666664434524222570944589E-02;
    static constexpr double cs6 = 1.38888874007937613028114285595617E-03;
    static constexpr double s1 = -0x1.5555555555555p-3;
    static constexpr double s2 = 0x1.1111111110ecep-7;
    static constexpr double s3 = -0x1.a01a019db08b8p-13;
    static constexpr double s4 = 0x1.71de27b9a7ed9p-19;
    static constexpr double s5 = -0x1.addffc2fcdf59p-26;
    static constexpr double big = 0x1.8p45;
    static constexpr double hp0 = 0x1.921fb54442d18p0;   // pi/2 high part
    static constexpr double hp1 = 0x1.1a62633145c07p-54;  // pi/2 - hp0
    static constexpr double mp1 = 0x1.921fb58p0;          // pi/2 in four parts
    static constexpr double mp2 = -0x1.dde973cp-27;
    static constexpr double pp3 = -0x1.cb3b398p-55;
    static constexpr double pp4 = -0x1.d747f23e32ed7p-83;
    static constexpr double hpinv = 0x1.45f306dc9c883p-1;  // 2/pi
    static constexpr double toint = 0x1.8p52;
};

// Nearest table point k/128 of |x| < 109.5/128 and the remainder |x| - k/128.
ASG_HD int sincos_index(double ax, double& r) {
    const double u = SinCosConst::big + ax;
    r = ax - (u - SinCosConst::big);
    return static_cast<int>(static_cast<std::uint32_t>(as_u64(u))) * 4;
}

// sin(a + da) for |a| < 0.126.
ASG_HD double sincos_taylor_sin(double xx, double a, double da) {
    using C = SinCosConst;
    const double p = fma_(fma_(fma_(fma_(C::s5, xx, C::s4), xx, C::s3), xx, C::s2), xx, C::s1);
    const double t = fma_(fma_(p, a, -(0.5 * da)), xx, da);
    return a + t;
}

// cos(x + dx), |x| < 0.85546875.
ASG_HD double sincos_do_cos(double x, double dx) {
    using C = SinCosConst;
    if (x < 0.0) {
        dx = -dx;
    }
    double r;
    const int k = sincos_index(__builtin_fabs(x), r);
    x = r + dx;
    const double xx = x * x;
    const double s = fma_(x * xx, fma_(xx, C::sn5, C::sn3), x);
    const double c = xx * fma_(xx, fma_(xx, C::cs6, C::cs4), C::cs2);
    const double sn = ASG_SINCOS_TAB[k], ssn = ASG_SINCOS_TAB[k + 1];
    const double cs = ASG_SINCOS_TAB[k + 2], ccs = ASG_SINCOS_TAB[k + 3];
    const double cor = fma_(-sn, s, fma_(-cs, c, fma_(-s, ssn, ccs)));
    return cs + cor;
}

// sin(x + dx), |x| < 0.85546875.
ASG_HD double sincos_do_sin(double x, double dx) {
    using C = SinCosConst;
    const double xold = x;
    if (__builtin_fabs(x) < 0.126) {
        return sincos_taylor_sin(x * x, x, dx);
    }
    if (x <= 0.0) {
        dx = -dx;
    }
    double r;
    const int k = sincos_index(__builtin_fabs(x), r);
    x = r;
    const double xx = x * x;
    const double s = x + fma_(x * xx, fma_(xx, C::sn5, C::sn3), dx);
    const double c = fma_(x, dx, xx * fma_(xx, fma_(xx, C::cs6, C::cs4), C::cs2));
    const double sn = ASG_SINCOS_TAB[k], ssn = ASG_SINCOS_TAB[k + 1];
    const double cs = ASG_SINCOS_TAB[k + 2], ccs = ASG_SINCOS_TAB[k + 3];
    const double cor = fma_(cs, s, fma_(-sn, c, fma_(s, ccs, ssn)));
    return __builtin_copysign(sn + cor, xold);
}

// x = n pi/2 + (a + da), |a + da| <= pi/4; returns n mod 4.
ASG_HD int sincos_reduce(double x, double& a, double& da) {
    using C = SinCosConst;
    const double t = fma_(x, C::hpinv, C::toint);
    const double xn = t - C::toint;
    const double y = fma_(-xn, C::mp2, fma_(-xn, C::mp1, x));
    const int n = static_cast<int>(as_u64(t) & 3);
    const double t2 = fma_(-xn, C::pp3, y);
    double db = fma_(-xn, C::pp3, y - t2);
    const double b = fma_(-xn, C::pp4, t2);
    db = db + fma_(-xn, C::pp4, t2 - b);
    a = b;
    da = db;
    return n;
}

// sin (n even) or cos (n odd) of a + da, negated for quadrants n & 2.
ASG_HD double sincos_quadrant(double a, double da, int n) {
    const double r = (n & 1) ? sincos_do_cos(a, da) : sincos_do_sin(a, da);
    return (n & 2) ? -r : r;
}

ASG_HD double asg_sin(double x) {
    using C = SinCosConst;
    const std::uint32_t k = static_cast<std::uint32_t>(as_u64(x) >> 32) & 0x7fffffffu;
    if (k < 0x3e500000u) {
        return x;
    }
    if (k < 0x3feb6000u) {
        return sincos_do_sin(x, 0.0);
    }
    if (k < 0x400368fdu) {
        return __builtin_copysign(sincos_do_cos(C::hp0 - __builtin_fabs(x), C::hp1), x);
    }
    if (k < 0x419921fbu) {
        double a, da;
        const int n = sincos_reduce(x, a, da);
        return sincos_quadrant(a, da, n);
    }
    return __builtin_nan("");
}

ASG_HD double asg_cos(double x) {
    using C = SinCosConst;
    const std::uint32_t k = static_cast<std::uint32_t>(as_u64(x) >> 32) & 0x7fffffffu;
    if (k < 0x3e400000u) {
        return 1.0;
    }
    if (k < 0x3feb6000u) {
        return sincos_do_cos(x, 0.0);
    }
    if (k < 0x400368fdu) {
        const double y = C::hp0 - __builtin_fabs(x);
        const double a = y + C::hp1;
        const double da = (y - a) + C::hp1;
        return sincos_do_sin(a, da);
    }
    if (k < 0x419921fbu) {
        double a, da;
        const int n = sincos_reduce(x, a, da);
        return sincos_quadrant(a, da, n + 1);
    }
    return __builtin_nan("");
}

ASG_HD double asg_sqrt(double x) { return __builtin_sqrt(x); }

}  // namespace asg
