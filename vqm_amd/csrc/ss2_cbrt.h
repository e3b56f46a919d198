// Cube roots of the XYB transform (prepare.comp:81-98 takes pow(m, 1/3) of
// each opsin mix), as host-and-device functions so that the CPU test that
// checks them runs this very code.
//
//   cbrt_fast  <= 0.84 ulp on [2^-126, 64); 16 FP32 operations.
//   cbrt_rn    the cube root correctly rounded to FP32 (round to nearest),
//              for every normal positive float: a third-order and a
//              second-order FP64 step from cbrt_fast's seed, rounded to
//              FP32, with two exact FP64 comparisons for the one input in
//              2^19 whose FP64 root lies too near a midpoint to round safely.
//   cbrt_rn_hw the same root on the device from a hardware log2/exp2 seed
//              and one FP64 step (the kernels' reference form).
//
// Every step of cbrt_fast and cbrt_rn is an IEEE conversion, add, multiply
// or fused multiply-add (written out), with contraction off, so the result
// does not depend on the compiler or on where it runs.
// tests/cpp/cbrt_exhaustive.hip checks cbrt_rn against an exact integer
// comparison for every float in [2^-126, 2^24). cbrt_rn_hw's seed comes from
// hardware instructions whose rounding no document fixes, so it is checked
// on the device instead: equal to cbrt_rn for every positive normal float.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#pragma clang fp contract(off)

namespace vqm::ss2 {

__host__ __device__ inline uint32_t cbrt_bits(float x) { return __builtin_bit_cast(uint32_t, x); }
__host__ __device__ inline float cbrt_float(uint32_t u) { return __builtin_bit_cast(float, u); }

// Bit-trick seed of x^(-1/3) (magic - bits/3, formed in float: one
// conversion each way around one fma), one third-order step
// r <- r (1 + h/3 + 2h^2/9 + 14h^3/81) with h = 1 - x r^3 (the series of
// (1 - h)^(-1/3); |h| <= 0.105 after the seed leaves r within 2.3e-5), then
// y = x r^2 with one Newton correction. x must be a positive normal float.
__host__ __device__ inline float cbrt_fast(float x) {
    constexpr float kThird = 0x1.555556p-2f;       // 1/3 rounded to FP32
    constexpr float kMagic = 1419910400.0f;        // 0x54a21d2a rounded to FP32
    constexpr float kC2 = 0x1.c71c72p-3f;          // 2/9
    constexpr float kC3 = 0x1.61f9aep-3f;          // 14/81
    const float seed = __builtin_fmaf(static_cast<float>(cbrt_bits(x)), -kThird, kMagic);
    float r = cbrt_float(static_cast<uint32_t>(seed));
    const float h = __builtin_fmaf(-x, (r * r) * r, 1.0f);
    const float p = __builtin_fmaf(h, __builtin_fmaf(h, kC3, kC2), kThird);
    r = __builtin_fmaf(r * h, p, r);
    const float rr = r * r;
    const float y = x * rr;
    const float e = __builtin_fmaf(-(y * y), y, x);
    return __builtin_fmaf(e, rr * kThird, y);
}

// Sign of x - m^3 for FP64 values x and m where m carries at most 26
// significant bits and m^3 lies within a factor of two of x: m^2 is exact
// in FP64, m^3 = p + e exactly with p = RN(m^2 m) and e = fma(m^2, m, -p)
// (the product's rounding error), and x - p is exact (Sterbenz), so
// x > m^3 exactly when x - p > e. Returns +1, 0 or -1.
__host__ __device__ inline int cube_cmp(double x, double m) {
    const double m2 = m * m;
    const double p = m2 * m;
    const double e = __builtin_fma(m2, m, -p);
    const double r = x - p;
    return r > e ? 1 : (r < e ? -1 : 0);
}

// The correctly rounded root from a float `y` within one ulp of it: the
// sign of x - y^3 names the side the true root t lies on, and one more
// comparison against the cube of the midpoint between y and its neighbour
// on that side decides whether t lies beyond it. A midpoint has 25
// significant bits and its cube is never a float, so there are no ties.
// Both midpoints and y are exact in FP64 and their cubes are within a
// factor of two of x.
__host__ __device__ inline float cbrt_round_exact(float x, float y) {
    const double xd = x, yd = y;
    const int side = cube_cmp(xd, yd);
    if (side == 0) return y;  // x is the cube of y
    const float n = cbrt_float(cbrt_bits(y) + (side > 0 ? 1u : 0xFFFFFFFFu));
    const double mid = (yd + static_cast<double>(n)) * 0.5;
    return cube_cmp(xd, mid) == side ? n : y;
}

// cbrt_rn (below) without its rare exact step: the FP64 root rounded to
// FP32, and in `near` whether that rounding is unsafe (then
// cbrt_round_exact(x, result) is the correctly rounded root). Kernels
// evaluate several roots this way and settle the near ones in one branch, so
// the roots' instruction chains stay in one basic block and interleave.
__host__ __device__ inline float cbrt_rn_near(float x, bool* near) {
    constexpr float kThird = 0x1.555556p-2f, kMagic = 1419910400.0f;
    constexpr double kThirdD = 1.0 / 3.0, kC2 = 2.0 / 9.0, kC3 = 14.0 / 81.0;
    const float seed = __builtin_fmaf(static_cast<float>(cbrt_bits(x)), -kThird, kMagic);
    const double xd = x;
    double r = cbrt_float(static_cast<uint32_t>(seed));
    double h = __builtin_fma(-xd, (r * r) * r, 1.0);
    r = __builtin_fma(r * h, __builtin_fma(h, __builtin_fma(h, kC3, kC2), kThirdD), r);
    h = __builtin_fma(-xd, (r * r) * r, 1.0);
    r = __builtin_fma(r * h, __builtin_fma(h, kC2, kThirdD), r);
    const double y = xd * (r * r);
    const float rounded = static_cast<float>(y);
    const uint32_t dropped = static_cast<uint32_t>(__builtin_bit_cast(uint64_t, y)) & 0x1FFFFFFFu;
    *near = dropped - (0x10000000u - 0x200u) <= 2 * 0x200u;  // within 2^9 units of a midpoint
    return rounded;
}

// The seed of cbrt_rn_near_hw: x^(-1/3) as 2^(-log2(x) / 3) from the
// hardware base-2 logarithm and exponential (v_log_f32, v_exp_f32; one
// transcendental instruction each).
__device__ inline float cbrt_seed_hw(float x) {
    constexpr float kThird = 0x1.555556p-2f;
    return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * -kThird);
}

// cbrt_rn_near for the device, from the hardware seed: |h| = |1 - x r^3|
// is at most 1.09e-5 (2^-16.5) over every positive normal float (measured
// by vqm_ss2_check_cbrt; the instruction set documents no bound), so one
// second-order FP64 step r <- r (1 + h (1/3 + h 2/9)) (dropped term
// 14/81 h^3 < 2^-52) and y = x r^2 leave y within about 2^-50 of the root,
// inside the 2^9-unit band (2^-44) of the midpoint test. 17 instructions,
// two of them transcendental, against cbrt_rn_near's 25. For every positive
// normal float, cbrt_rn_near_hw followed by cbrt_round_exact where `near`
// returns cbrt_rn's root (checked on the device for all 2^31 - 2^24 of them
// by vqm_ss2_check_cbrt; cbrt_rn is checked against exact arithmetic on the
// CPU).
__device__ inline float cbrt_rn_near_hw(float x, bool* near) {
    constexpr double kThirdD = 1.0 / 3.0, kC2 = 2.0 / 9.0;
    const double xd = x;
    double r = cbrt_seed_hw(x);
    const double h = __builtin_fma(-xd, (r * r) * r, 1.0);
    r = __builtin_fma(r * h, __builtin_fma(h, kC2, kThirdD), r);
    const double y = xd * (r * r);
    const float rounded = static_cast<float>(y);
    const uint32_t dropped = static_cast<uint32_t>(__builtin_bit_cast(uint64_t, y)) & 0x1FFFFFFFu;
    *near = dropped - (0x10000000u - 0x200u) <= 2 * 0x200u;  // within 2^9 units of a midpoint
    return rounded;
}

__device__ inline float cbrt_rn_hw(float x) {
    bool near;
    const float rounded = cbrt_rn_near_hw(x, &near);
    return near ? cbrt_round_exact(x, rounded) : rounded;
}

// The cube root of a positive normal float, correctly rounded. The FP32
// seed of cbrt_fast (x^(-1/3) within 3.7%), then in FP64 one third-order
// step r <- r (1 + h (1/3 + h (2/9 + h 14/81))), h = 1 - x r^3 (the error
// falls to |h| < 5e-5), and one second-order step r <- r (1 + h (1/3 +
// h 2/9)) (the dropped term 14/81 h^3 is below 2.2e-14), and y = x r^2 in
// FP64: within 2^-44.4 of the root, relatively (the largest error over every
// float in [2^-126, 2^24); the seed's bit pattern, and with it the error,
// repeats every three binades). Rounding y to FP32 gives the correctly
// rounded root unless y lies within 2^-44 of a midpoint between two floats:
// the 29 bits FP32 drops from y's significand (all in its low word) then sit
// within 2^9 units of 2^28 (one input in 2^19). Those take cbrt_round_exact.
__host__ __device__ inline float cbrt_rn(float x) {
    bool near;
    const float rounded = cbrt_rn_near(x, &near);
    return near ? cbrt_round_exact(x, rounded) : rounded;
}

}  // namespace vqm::ss2
