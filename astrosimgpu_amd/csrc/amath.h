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
//  * asg_sin / asg_cos use fdlibm's kernels (Sun Microsystems, 1993:
//    k_sin.c, k_cos.c and the medium-range Cody-Waite reduction of
//    e_rem_pio2.c). They are accurate to < 1 ulp but are NOT glibc's sin/cos
//    (IBM Accurate Mathematical Library), so results that pass through them
//    (Box-Muller noise) are bit-identical host-vs-device but only
//    statistically identical to a glibc build.
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

ASG_HD double asg_exp(double x) {
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
    const double tail = as_f64(ASG_EXP_TAB[idx]);
    const std::uint64_t sbits = ASG_EXP_TAB[idx + 1] + top;
    const double r2 = r * r;
    const double tmp = fma_(r2 * r2, fma_(r, C5, C4), fma_(r2, fma_(r, C3, C2), tail + r));
    if (abstop == 0) {
        return exp_specialcase(tmp, sbits, ki);
    }
    const double scale = as_f64(sbits);
    return fma_(scale, tmp, scale);
}

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

ASG_HD double kernel_sin(double x, double y, int iy) {
    constexpr double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                     S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                     S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
    const double z = x * x;
    const double v = z * x;
    const double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
    if (iy == 0) {
        return x + v * (S1 + z * r);
    }
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

ASG_HD double kernel_cos(double x, double y) {
    constexpr double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                     C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                     C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
    const std::uint32_t ix = static_cast<std::uint32_t>(as_u64(x) >> 32) & 0x7fffffffu;
    if (ix < 0x3e400000u) {
        return 1.0;  // |x| < 2^-27
    }
    const double z = x * x;
    const double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
    if (ix < 0x3fd33333u) {
        return 1.0 - (0.5 * z - (z * r - x * y));
    }
    const double qx = ix > 0x3fe90000u
                          ? 0.28125
                          : as_f64(static_cast<std::uint64_t>(ix - 0x00200000u) << 32);
    const double hz = 0.5 * z - qx;
    const double a = 1.0 - qx;
    return a - (hz - (z * r - x * y));
}

// Reduces |x| < 2^19 * pi/2 to y0 + y1 in [-pi/4, pi/4]; returns the quadrant.
ASG_HD int rem_pio2_medium(double x, double& y0, double& y1) {
    constexpr double kInvPio2 = 6.36619772367581382433e-01;
    constexpr double kPio2_1 = 1.57079632673412561417e+00, kPio2_1t = 6.07710050650619224932e-11;
    constexpr double kPio2_2 = 6.07710050630396597660e-11, kPio2_2t = 2.02226624879595063154e-21;
    constexpr double kPio2_3 = 2.02226624871116645580e-21, kPio2_3t = 8.47842766036889956997e-32;
    const double t = x < 0.0 ? -x : x;
    const int n = static_cast<int>(t * kInvPio2 + 0.5);
    const double fn = static_cast<double>(n);
    double r = t - fn * kPio2_1;
    double w = fn * kPio2_1t;
    y0 = r - w;
    const int j = static_cast<int>((as_u64(t) >> 52) & 0x7ff);
    int i = j - static_cast<int>((as_u64(y0) >> 52) & 0x7ff);
    if (i > 16) {
        double tt = r;
        w = fn * kPio2_2;
        r = tt - w;
        w = fn * kPio2_2t - ((tt - r) - w);
        y0 = r - w;
        i = j - static_cast<int>((as_u64(y0) >> 52) & 0x7ff);
        if (i > 49) {
            tt = r;
            w = fn * kPio2_3;
            r = tt - w;
            w = fn * kPio2_3t - ((tt - r) - w);
            y0 = r - w;
        }
    }
    y1 = (r - y0) - w;
    if (x < 0.0) {
        y0 = -y0;
        y1 = -y1;
        return -n;
    }
    return n;
}

// Both need only |x| < 2^19 * pi/2 (the Box-Muller angle is in [0, 2 pi)).
ASG_HD double asg_sin(double x) {
    const std::uint32_t ix = static_cast<std::uint32_t>(as_u64(x) >> 32) & 0x7fffffffu;
    if (ix <= 0x3fe921fbu) {
        if (ix < 0x3e400000u) {
            return x;
        }
        return kernel_sin(x, 0.0, 0);
    }
    double y0, y1;
    const int n = rem_pio2_medium(x, y0, y1);
    switch (n & 3) {
        case 0: return kernel_sin(y0, y1, 1);
        case 1: return kernel_cos(y0, y1);
        case 2: return -kernel_sin(y0, y1, 1);
        default: return -kernel_cos(y0, y1);
    }
}

ASG_HD double asg_cos(double x) {
    const std::uint32_t ix = static_cast<std::uint32_t>(as_u64(x) >> 32) & 0x7fffffffu;
    if (ix <= 0x3fe921fbu) {
        return kernel_cos(x, 0.0);
    }
    double y0, y1;
    const int n = rem_pio2_medium(x, y0, y1);
    switch (n & 3) {
        case 0: return kernel_cos(y0, y1);
        case 1: return -kernel_sin(y0, y1, 1);
        case 2: return -kernel_cos(y0, y1);
        default: return kernel_sin(y0, y1, 1);
    }
}

ASG_HD double asg_sqrt(double x) { return __builtin_sqrt(x); }

}  // namespace asg
