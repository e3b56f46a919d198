// CPU oracle for the SSIMULACRA 2.1 scoring path -- TEST INFRASTRUCTURE ONLY.
//
// A restatement of the reference's arithmetic (GreatValueCreamSoda/vqm) for
// the parity tests, the golden fixtures and the bench's CPU baseline. Nothing
// in the product library (vqm_amd/csrc) links, loads or calls this file; see
// ss2_oracle.h for the three arithmetic forms and the pins that check them.
//
// Build: oracle/Makefile (g++ -O3 -march=x86-64-v3 -ffp-contract=off). Every
// fused multiply-add below is an explicit std::fma so that the compiler adds
// no contraction of its own; the FP32 "dev" form then reproduces the HIP
// kernels' per-pixel values exactly.

#include "ss2_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr int kMaxScales = 6;      // ssimulacra2_internal.h:185
constexpr uint32_t kMinDim = 8;    // ssimulacra2_internal.h:186
constexpr int kRadius = 6;         // ceil(3.5 * 1.5), pipeline.cc:276-277
constexpr int kTaps = 13;

// The subjective weights, one row per (channel, scale) in the order
// ssim_1, ring_1, blur_1, ssim_4, ring_4, blur_4 (score.cc:29-87). A table of
// published constants: it has to be the same numbers.
constexpr double kWeights[18][6] = {
    {0.0, 0.0007376606707406586, 0.0, 0.0, 0.0007793481682867309, 0.0},
    {0.0, 0.0004371155730107379, 0.0, 1.1041726426657346, 0.00066284834129271,
     0.00015231632783718752},
    {0.0, 0.0016406437456599754, 0.0, 1.8422455520539298, 11.441172603757666, 0.0},
    {0.0007989109436015163, 0.000176816438078653, 0.0, 1.8787594979546387,
     10.94906990605142, 0.0},
    {0.0007289346991508072, 0.9677937080626833, 0.0, 0.00014003424285435884,
     0.9981766977854967, 0.00031949755934435053},
    {0.0004550992113792063, 0.0, 0.0, 0.0013648766163243398, 0.0, 0.0},
    {0.0, 0.0, 0.0, 7.466890328078848, 0.0, 17.445833984131262},
    {0.0006235601634041466, 0.0, 0.0, 6.683678146179332, 0.00037724407979611296,
     1.027889937768264},
    {225.20515300849274, 0.0, 0.0, 19.213238186143016, 0.0011401524586618361,
     0.001237755635509985},
    {176.39317598450694, 0.0, 0.0, 24.43300999870476, 0.28520802612117757,
     0.0004485436923833408},
    {0.0, 0.0, 0.0, 34.77906344483772, 44.835625328877896, 0.0},
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.0, 0.0008680556573291698, 0.0, 0.0, 0.0, 0.0},
    {0.0, 0.0005313191874358747, 0.0, 0.00016533814161379112, 0.0, 0.0},
    {0.0, 0.0, 0.0, 0.0004179171803251336, 0.0017290828234722833, 0.0},
    {0.0020827005846636437, 0.0, 0.0, 8.826982764996862, 23.19243343998926, 0.0},
    {95.1080498811086, 0.9863978034400682, 0.9834382792465353,
     0.0012286405048278493, 171.2667255897307, 0.9807858872435379},
    {0.0, 0.0, 0.0, 0.0005130064588990679, 0.0, 0.00010854057858411537},
};

// XYB constants of prepare.comp:40-54 (libjxl's opsin absorbance matrix).
constexpr double kM00 = 0.30, kM02 = 0.078, kM01 = 1.0 - 0.078 - 0.30;
constexpr double kM10 = 0.23, kM12 = 0.078, kM11 = 1.0 - 0.078 - 0.23;
constexpr double kM20 = 0.24342268924547819, kM21 = 0.20476744424496821;
constexpr double kM22 = 1.0 - 0.24342268924547819 - 0.20476744424496821;
constexpr double kOpsinBias = 0.0037930732552754493;
constexpr double kCbrtBias = 0.15595419704914093;
constexpr double kC2 = 0.0009;  // scale.comp:87

float bits_to_float(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
uint32_t float_to_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

uint64_t align4(uint64_t v) { return (v + 3u) / 4u * 4u; }

// Round-to-nearest-even float -> binary16 -> float.
float round_half(float v) {
    uint32_t bits = float_to_bits(v);
    const uint32_t sign = bits & 0x80000000u;
    const uint32_t abs = bits & 0x7FFFFFFFu;
    if (abs >= 0x7F800000u) return v;                  // inf/nan pass through
    if (abs >= 0x477FF000u) return bits_to_float(sign | 0x7F800000u);  // overflow
    const int32_t exp = static_cast<int32_t>(abs >> 23) - 127;
    if (exp < -25) return bits_to_float(sign);          // below half the smallest subnormal
    // Quantum of the binary16 value at this magnitude.
    const int32_t q_exp = exp < -14 ? -24 : exp - 10;
    const double quantum = std::ldexp(1.0, q_exp);
    const double mag = static_cast<double>(bits_to_float(abs));
    double scaled = mag / quantum;                      // exact (power of two)
    double fl = std::floor(scaled);
    const double rem = scaled - fl;
    if (rem > 0.5 || (rem == 0.5 && std::fmod(fl, 2.0) != 0.0)) fl += 1.0;
    const float out = static_cast<float>(fl * quantum);
    return sign ? -out : out;
}

// Round-to-nearest-even float -> bfloat16 -> float.
float round_bf16(float v) {
    uint32_t bits = float_to_bits(v);
    if ((bits & 0x7FFFFFFFu) > 0x7F800000u) return v;
    const uint32_t lsb = (bits >> 16) & 1u;
    bits += 0x7FFFu + lsb;
    bits &= 0xFFFF0000u;
    return bits_to_float(bits);
}

// Round-to-nearest-even float -> unsigned 16-bit fixed point q / 65536
// (clamped to 0..65535) -> float. XYB values of linear RGB in [0, 1] lie in
// [0, 0.94], so the step is a uniform 2^-16 and zero stays exactly zero.
float round_unorm16(float v) {
    if (!(v > 0.0f)) return 0.0f;
    const double scaled = static_cast<double>(v) * 65536.0;  // exact
    if (scaled >= 65535.0) return 65535.0f / 65536.0f;
    double fl = std::floor(scaled);
    const double rem = scaled - fl;
    if (rem > 0.5 || (rem == 0.5 && std::fmod(fl, 2.0) != 0.0)) fl += 1.0;
    return static_cast<float>(fl / 65536.0);
}

template <typename T> T fmaT(T a, T b, T c) { return std::fma(a, b, c); }

// Smallest normal float: the device clamps the XYB mixes to it instead of
// 0; its root 2^-42 is below half an ulp of the cube-root bias, so
// cbrt - bias rounds to -bias exactly as for a zero mix.
constexpr float kMinNormal = 0x1p-126f;

// Device cube root of max(x, 2^-126): a bit-trick seed for x^(-1/3),
// magic - bits/3 formed in float (uint -> float rounds to nearest, one fma
// with the magic 0x54a21d2a rounded to FP32, float -> uint truncates); one
// third-order step r <- r (1 + h (1/3 + h (2/9 + h 14/81))) with
// h = 1 - x r^3, the series of (1 - h)^(-1/3) to h^3; then y = x r^2 and one
// Newton correction of y. <= 0.84 ulp on [2^-126, 64); every operation is
// an IEEE conversion, add, mul or fma, so CPU and GPU agree bit for bit.
float cbrt_dev(float x) {
    x = std::max(x, kMinNormal);
    const float third = 0x1.555556p-2f, c2 = 0x1.c71c72p-3f, c3 = 0x1.61f9aep-3f;
    const float seed = std::fma(static_cast<float>(float_to_bits(x)), -third, 1419910400.0f);
    float r = bits_to_float(static_cast<uint32_t>(seed));
    const float h = std::fma(-x, (r * r) * r, 1.0f);
    const float p = std::fma(h, std::fma(h, c3, c2), third);
    r = std::fma(r * h, p, r);
    const float rr = r * r;
    const float y = x * rr;
    const float e = std::fma(-(y * y), y, x);
    return std::fma(e, rr * third, y);
}

// Device reciprocal of a positive normal float: the correctly rounded 1 / x
// (the kernels take the hardware reciprocal and one Newton step, which the
// library's device self-check compares with the IEEE quotient for every
// significand).
float rcp_dev(float x) { return 1.0f / x; }
double rcp_dev(double x) { return 1.0 / x; }  // the FP64 forms divide

// A finite positive float or a midpoint of two adjacent ones as M * 2^k
// with M an integer of at most 26 bits.
struct Dyadic {
    int64_t m;
    int k;
};

Dyadic dyadic_of(float f) {
    int e = 0;
    const double fr = std::frexp(static_cast<double>(f), &e);  // f = fr * 2^e, fr in [0.5, 1)
    return {static_cast<int64_t>(std::ldexp(fr, 24)), e - 24};
}

Dyadic midpoint_of(float lo, float hi) {
    const Dyadic a = dyadic_of(lo), b = dyadic_of(hi);
    const int k = std::min(a.k, b.k);
    return {(a.m << (a.k - k)) + (b.m << (b.k - k)), k - 1};
}

// Sign of m^3 - x, exactly, in 128-bit integers: m^3 = M^3 2^(3k) with
// M^3 < 2^78, x = X 2^j with X < 2^24.
int cube_vs(Dyadic m, float x) {
    const Dyadic xd = dyadic_of(x);
    const __int128 m3 = static_cast<__int128>(m.m) * m.m * m.m;
    const int s = 3 * m.k - xd.k;  // m^3 - x = 2^j (M^3 2^s - X)
    __int128 lhs = m3, rhs = xd.m;
    if (s >= 0) {
        if (s > 40) return 1;       // M^3 2^s >= 2^40 > X
        lhs <<= s;
    } else {
        if (-s > 100) return -1;    // M^3 < 2^78 <= X 2^-s
        rhs <<= -s;
    }
    return lhs > rhs ? 1 : (lhs < rhs ? -1 : 0);
}

// The cube root of x correctly rounded to FP32 (round to nearest): libm's
// FP64 root rounded to FP32 as the candidate, then moved to a neighbour
// while the true root lies beyond the midpoint between them, decided in
// exact integer arithmetic. 0 for x <= 0 (the reference clamps the opsin
// mixes at 0, prepare.comp:81-98, and pow(0, 1/3) = 0).
float cbrt_exact(float x) {
    if (!(x > 0.0f)) return 0.0f;
    const double d = std::cbrt(static_cast<double>(x));
    float c = static_cast<float>(d);
    // Fast path: when libm's root (within an FP64 ulp of the true root) lies
    // well away from the midpoint between c and its neighbour on d's side,
    // the true root rounds to c as well.
    const float side = d > static_cast<double>(c) ? std::nextafter(c, INFINITY) : std::nextafter(c, 0.0f);
    const double half = 0.5 * std::fabs(static_cast<double>(side) - static_cast<double>(c));
    if (std::fabs(std::fabs(d - static_cast<double>(c)) - half) > 0x1p-40 * d) return c;
    for (;;) {
        const float up = std::nextafter(c, INFINITY), down = std::nextafter(c, 0.0f);
        if (cube_vs(midpoint_of(c, up), x) < 0) {
            c = up;                 // the root lies above the upper midpoint
        } else if (down > 0.0f && cube_vs(midpoint_of(down, c), x) > 0) {
            c = down;               // below the lower midpoint
        } else {
            return c;
        }
    }
}

// sRGB decode of one u8 code: the decode's scale (input.cc:714-757, full
// range u8 -> 1/255 rounded to float, offset 0) and the IEC 61966-2-1 EOTF
// (color.glsl:135-137), FP32 operations with the power taken in double.
float srgb_f32(int code) {
    const float scale = static_cast<float>(1.0 / 255.0);
    const float v = static_cast<float>(code) * scale;
    if (v <= 0.04045f) return v / 12.92f;
    const float t = (v + 0.055f) / 1.055f;
    return static_cast<float>(std::pow(static_cast<double>(t), 2.4));
}

double srgb_f64(int code) {
    const double v = static_cast<double>(code) / 255.0;
    if (v <= 0.04045) return v / 12.92;
    return std::pow((v + 0.055) / 1.055, 2.4);
}

void xyb_f32(float r, float g, float b, int cbrt_kind, float out[3]) {
    const float m00 = static_cast<float>(kM00), m01 = static_cast<float>(kM01),
                m02 = static_cast<float>(kM02), m10 = static_cast<float>(kM10),
                m11 = static_cast<float>(kM11), m12 = static_cast<float>(kM12),
                m20 = static_cast<float>(kM20), m21 = static_cast<float>(kM21),
                m22 = static_cast<float>(kM22), bias = static_cast<float>(kOpsinBias);
    float mix0 = std::fma(m00, r, std::fma(m01, g, std::fma(m02, b, bias)));
    float mix1 = std::fma(m10, r, std::fma(m11, g, std::fma(m12, b, bias)));
    float mix2 = std::fma(m20, r, std::fma(m21, g, std::fma(m22, b, bias)));
    // The device clamps to the smallest normal (cbrt_dev), the reference's
    // order to 0; both give c = -bias for a mix at or below zero.
    mix0 = std::max(mix0, 0.0f);
    mix1 = std::max(mix1, 0.0f);
    mix2 = std::max(mix2, 0.0f);
    const float cbias = static_cast<float>(kCbrtBias);
    if (cbrt_kind == 1) {
        // The device form cancels the bias where it cancels exactly:
        // X = 7 (cbrt0 - cbrt1) + 0.42, Y = (cbrt0 + cbrt1)/2 + (0.01 - bias),
        // B = cbrt2 - (cbrt0 + cbrt1)/2 + 0.55 (fused as the kernels do).
        const float r0 = cbrt_dev(mix0), r1 = cbrt_dev(mix1), r2 = cbrt_dev(mix2);
        const float sum = r0 + r1;
        const float y_offset = 0.01f - cbias;
        out[0] = std::fma(r0 - r1, 7.0f, 0.42f);
        out[1] = std::fma(0.5f, sum, y_offset);
        out[2] = std::fma(-0.5f, sum, r2) + 0.55f;
        return;
    }
    const float c0 = cbrt_exact(mix0) - cbias;
    const float c1 = cbrt_exact(mix1) - cbias;
    const float c2 = cbrt_exact(mix2) - cbias;
    const float half_sum = 0.5f * (c0 + c1);
    out[0] = std::fma(0.5f * (c0 - c1), 14.0f, 0.42f);
    out[1] = half_sum + 0.01f;
    out[2] = (c2 - half_sum) + 0.55f;
}

void xyb_f64(double r, double g, double b, double out[3]) {
    double mix0 = kM00 * r + (kM01 * g + (kM02 * b + kOpsinBias));
    double mix1 = kM10 * r + (kM11 * g + (kM12 * b + kOpsinBias));
    double mix2 = kM20 * r + (kM21 * g + (kM22 * b + kOpsinBias));
    mix0 = std::max(mix0, 0.0);
    mix1 = std::max(mix1, 0.0);
    mix2 = std::max(mix2, 0.0);
    const double c0 = std::cbrt(mix0) - kCbrtBias;
    const double c1 = std::cbrt(mix1) - kCbrtBias;
    const double c2 = std::cbrt(mix2) - kCbrtBias;
    out[0] = (0.5 * (c0 - c1)) * 14.0 + 0.42;
    out[1] = 0.5 * (c0 + c1) + 0.01;
    out[2] = (c2 - 0.5 * (c0 + c1)) + 0.55;
}

template <typename T> void taps_as(T taps[kTaps]) {
    // pipeline.cc:274-291: exp(-0.5 (k/sigma)^2) in double, normalised, FP32.
    float f[kTaps];
    ss2o_gaussian_taps(f);
    for (int i = 0; i < kTaps; ++i) taps[i] = static_cast<T>(f[i]);
}

// The kernel of libjxl's recursive Gaussian (CreateRecursiveGaussian(1.5),
// Cloudinary ssimulacra2.cc:82-106; third-party, libjxl 0.8/0.9 as vendored
// by ssimulacra2 2.1), restated from its published algorithm: Charalampidis,
// "Recursive implementation of the Gaussian filter using truncated cosine
// functions", IEEE Trans. Signal Process. 64(14), 2016. The filter is
//     K[m] = sum_{k=1,3,5} beta_k cos(w_k m),  |m| <= N,  w_k = k pi / (2N),
// N = round(3.2795 sigma + 0.2546) (eq. 57) = 5 at sigma 1.5. Its three
// second-order recursions with zero input outside the row compute exactly
// this zero-padded FIR: with cos(w_k N) = 0 the window edges enter the
// recursion only through x[n+N-1] + x[n-N-1], and K[+-N] = 0, so 9 taps
// remain. beta is the cosine-series coefficient of the Gaussian,
// rho_k = exp(-sigma^2 w_k^2 / 2) / N (eq. 50), corrected by the smallest
// change that restores unit gain (sum K = 1, eq. 39) and variance
// (sum m^2 K = sigma^2, eq. 55): beta = rho + a p + b r with
// p_k = sum_m cos(w_k m), r_k = sum_m (N^2 - m^2) cos(w_k m) (eq. 37, 44);
// the paper's eq. (52)-(56) solve the same system. Returned in the 13-tap
// frame of the sampled kernel, centre at index 6, zeros at |m| = 5, 6.
void cosine_taps_f64(double taps[kTaps]) {
    constexpr double kPi = 3.14159265358979323846;
    const double sigma = 1.5;
    const int n = static_cast<int>(std::lround(3.2795 * sigma + 0.2546));
    double w[3], p[3] = {0, 0, 0}, r[3] = {0, 0, 0}, rho[3];
    for (int k = 0; k < 3; ++k) {
        w[k] = (2 * k + 1) * kPi / (2.0 * n);
        for (int m = -n; m <= n; ++m) {
            const double c = std::cos(w[k] * m);
            p[k] += c;
            r[k] += (static_cast<double>(n) * n - static_cast<double>(m) * m) * c;
        }
        rho[k] = std::exp(-0.5 * sigma * sigma * w[k] * w[k]) / n;
    }
    // Solve [p.p p.r; r.p r.r] [a b]^T = [1 - p.rho, N^2 - sigma^2 - r.rho].
    auto dot = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    const double m00 = dot(p, p), m01 = dot(p, r), m11 = dot(r, r);
    const double b0 = 1.0 - dot(p, rho), b1 = static_cast<double>(n) * n - sigma * sigma - dot(r, rho);
    const double det = m00 * m11 - m01 * m01;
    const double a = (b0 * m11 - m01 * b1) / det, b = (m00 * b1 - m01 * b0) / det;
    double beta[3];
    for (int k = 0; k < 3; ++k) beta[k] = rho[k] + a * p[k] + b * r[k];
    for (int i = 0; i < kTaps; ++i) {
        const int m = i - kRadius;
        double v = 0.0;
        if (m > -n && m < n)
            for (int k = 0; k < 3; ++k) v += beta[k] * std::cos(w[k] * m);
        taps[i] = v;
    }
}

// One 13-tap pass along rows (dir 0) or columns (dir 1) with zeros outside.
// The chain is acc = w0*x0, then acc = fma(wk, xk, acc) in ascending k: the
// tap order of scale.comp:187-199 and :243-256 and of the HIP kernels.
template <typename T>
void pass(const T* in, T* out, uint32_t w, uint32_t h, int dir, const T* wt) {
    if (dir == 0) {
        std::vector<T> row(w + 2 * kRadius, T(0));
        for (uint32_t y = 0; y < h; ++y) {
            std::copy(in + static_cast<size_t>(y) * w,
                      in + static_cast<size_t>(y + 1) * w, row.begin() + kRadius);
            T* o = out + static_cast<size_t>(y) * w;
            for (uint32_t x = 0; x < w; ++x) {
                const T* s = row.data() + x;
                T acc = wt[0] * s[0];
                for (int k = 1; k < kTaps; ++k) acc = fmaT(wt[k], s[k], acc);
                o[x] = acc;
            }
        }
    } else {
        std::vector<T> acc(w);
        for (uint32_t y = 0; y < h; ++y) {
            bool first = true;
            for (int k = 0; k < kTaps; ++k) {
                const int64_t sy = static_cast<int64_t>(y) + k - kRadius;
                const bool inside = sy >= 0 && sy < static_cast<int64_t>(h);
                const T* s = inside ? in + static_cast<size_t>(sy) * w : nullptr;
                if (first) {
                    for (uint32_t x = 0; x < w; ++x) acc[x] = wt[k] * (inside ? s[x] : T(0));
                    first = false;
                } else if (inside) {
                    for (uint32_t x = 0; x < w; ++x) acc[x] = fmaT(wt[k], s[x], acc[x]);
                }
                // fma(w, 0, acc) == acc for the finite, non-negative-zero-free
                // accumulators here, so rows outside the image are skipped.
            }
            std::copy(acc.begin(), acc.end(), out + static_cast<size_t>(y) * w);
        }
    }
}

template <typename T>
void blur(const T* in, T* out, uint32_t w, uint32_t h, int order, const T* wt) {
    std::vector<T> mid(static_cast<size_t>(w) * h);
    if (order == 0) {
        pass(in, mid.data(), w, h, 0, wt);
        pass(mid.data(), out, w, h, 1, wt);
    } else {
        pass(in, mid.data(), w, h, 1, wt);
        pass(mid.data(), out, w, h, 0, wt);
    }
}

struct Pyramid {
    int levels = 0;
    uint32_t w[kMaxScales]{}, h[kMaxScales]{};
};

// The six map sums of one (level, channel), scale.comp:305-333.
template <typename T>
void map_sums(const T* a, const T* b, uint32_t w, uint32_t h, int order,
              int sum_blocks, int sum_variances, const T* wt, double out[6]) {
    const size_t n = static_cast<size_t>(w) * h;
    std::vector<T> q(n), mu1(n), mu2(n), s11(n), s22(n), s12(n);
    blur(a, mu1.data(), w, h, order, wt);
    blur(b, mu2.data(), w, h, order, wt);
    if (sum_variances) {
        // The device form: (a - b)^2 as one blurred quantity (s11 holds it).
        // s11 + s22 = blur((a - b)^2) + 2 s12, so the SSIM' denominator is
        // num_s + v with v = blur((a - b)^2) - (mu1 - mu2)^2, and
        // den - num_m num_s = v + (mu1 - mu2)^2 num_s: no large-minus-large
        // cancellation beyond the one num_s already has.
        for (size_t i = 0; i < n; ++i) {
            const T dlt = a[i] - b[i];
            q[i] = dlt * dlt;
        }
        blur(q.data(), s11.data(), w, h, order, wt);
    } else {
        for (size_t i = 0; i < n; ++i) q[i] = a[i] * a[i];
        blur(q.data(), s11.data(), w, h, order, wt);
        for (size_t i = 0; i < n; ++i) q[i] = b[i] * b[i];
        blur(q.data(), s22.data(), w, h, order, wt);
    }
    for (size_t i = 0; i < n; ++i) q[i] = a[i] * b[i];
    blur(q.data(), s12.data(), w, h, order, wt);

    const T c2 = static_cast<T>(kC2);
    const T one = T(1), two = T(2), zero = T(0);
    for (int k = 0; k < 6; ++k) out[k] = 0.0;
    // sum_blocks: FP32 sums over blocks of one column by eight rows (rows
    // 8k .. 8k+7), in row order, then the block sums in FP64 -- the order of
    // the HIP kernels.
    std::vector<T> block(sum_blocks ? static_cast<size_t>(w) * 6 : 0, T(0));
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            const size_t i = static_cast<size_t>(y) * w + x;
            const T m1 = mu1[i], m2 = mu2[i];
            const T t = m1 - m2;
            const T num_s = fmaT(two, fmaT(-m1, m2, s12[i]), c2);
            const T det_dist = one + std::fabs(b[i] - m2);
            const T det_orig = one + std::fabs(a[i] - m1);
            T d, d1;
            if (sum_variances) {
                // The device form also shares one reciprocal between the two
                // quotients: q = 1 / (den det_orig), by rcp_dev.
                const T v = fmaT(-t, t, s11[i]);
                const T den = num_s + v;
                const T q = rcp_dev(den * det_orig);
                d = std::max((fmaT(t * t, num_s, v) * det_orig) * q, zero);
                d1 = ((det_dist - det_orig) * den) * q;
            } else {
                const T num_m = fmaT(-t, t, one);
                const T den = (fmaT(-m1, m1, s11[i]) + fmaT(-m2, m2, s22[i])) + c2;
                d = std::max(fmaT(-num_m, num_s, den) / den, zero);
                d1 = (det_dist - det_orig) / det_orig;
            }
            const T dd = d * d;
            const T ring = std::max(d1, zero);
            const T blurv = std::max(-d1, zero);
            const T rr = ring * ring, bb = blurv * blurv;
            const T v[6] = {d, dd * dd, ring, rr * rr, blurv, bb * bb};
            if (sum_blocks) {
                T* blk = block.data() + static_cast<size_t>(x) * 6;
                for (int k = 0; k < 6; ++k) blk[k] = blk[k] + v[k];
            } else {
                for (int k = 0; k < 6; ++k) out[k] += static_cast<double>(v[k]);
            }
        }
        if (sum_blocks && ((y & 7u) == 7u || y + 1 == h)) {
            for (size_t j = 0; j < block.size(); j += 6)
                for (int k = 0; k < 6; ++k) {
                    out[k] += static_cast<double>(block[j + k]);
                    block[j + k] = T(0);
                }
        }
    }
}

// prepare.comp:100-111,149-168: level l is the clamped 2x2 box of level
// l-1's linear RGB, summed ((((0 + a) + b) + c) + d) then scaled by 0.25, the
// samples in (0,0), (1,0), (0,1), (1,1) order.
template <typename T>
void downsample(std::vector<T> lin[3], uint32_t w, uint32_t h, uint32_t lw, uint32_t lh) {
    std::vector<T> next[3];
    for (int c = 0; c < 3; ++c) next[c].assign(static_cast<size_t>(lw) * lh, T(0));
    for (uint32_t y = 0; y < lh; ++y) {
        const uint32_t y0 = std::min(2 * y, h - 1), y1 = std::min(2 * y + 1, h - 1);
        for (uint32_t x = 0; x < lw; ++x) {
            const uint32_t x0 = std::min(2 * x, w - 1), x1 = std::min(2 * x + 1, w - 1);
            for (int c = 0; c < 3; ++c) {
                const T* s = lin[c].data();
                T sum = s[static_cast<size_t>(y0) * w + x0];
                sum = sum + s[static_cast<size_t>(y0) * w + x1];
                sum = sum + s[static_cast<size_t>(y1) * w + x0];
                sum = sum + s[static_cast<size_t>(y1) * w + x1];
                next[c][static_cast<size_t>(y) * lw + x] = sum * T(0.25);
            }
        }
    }
    for (int c = 0; c < 3; ++c) lin[c].swap(next[c]);
}

template <typename T>
int32_t pyramid(uint32_t width, uint32_t height, const uint8_t* image, T* out) {
    uint32_t sizes[12];
    const int32_t levels = ss2o_plan_scales(width, height, sizes);
    if (levels < 0) return -1;
    T lut[256];
    if constexpr (sizeof(T) == 4) {
        ss2o_srgb_lut_f32(reinterpret_cast<float*>(lut));
    } else {
        ss2o_srgb_lut_f64(reinterpret_cast<double*>(lut));
    }
    const uint64_t stride = align4(static_cast<uint64_t>(width) * height);
    std::vector<T> lin[3];
    const size_t n0 = static_cast<size_t>(width) * height;
    for (int c = 0; c < 3; ++c) {
        lin[c].resize(n0);
        for (size_t i = 0; i < n0; ++i) lin[c][i] = lut[image[c * stride + i]];
    }
    uint32_t w = width, h = height;
    for (int l = 0; l < levels; ++l) {
        if (l > 0) {
            downsample(lin, w, h, sizes[2 * l], sizes[2 * l + 1]);
            w = sizes[2 * l];
            h = sizes[2 * l + 1];
        }
        for (int c = 0; c < 3; ++c) out = std::copy(lin[c].begin(), lin[c].end(), out);
    }
    return levels;
}

// One image's stored XYB planes [level][channel] under `mode`: the linear
// pyramid (box averages rounded to plane_format first when round_linear),
// XYB of every level with a stored channel, stored planes rounded to
// plane_format at the levels in plane_round_levels. Pruned levels stay empty.
template <typename T>
void xyb_pyramid(const ss2o_mode& mode, uint32_t width, uint32_t height, int32_t levels,
                 const uint32_t sizes[12], const uint8_t* image, const T lut[256],
                 std::vector<T> xyb[kMaxScales][3]) {
    const uint64_t plane_stride = align4(static_cast<uint64_t>(width) * height);
    std::vector<T> lin[3];
    const size_t n0 = static_cast<size_t>(width) * height;
    for (int c = 0; c < 3; ++c) {
        lin[c].resize(n0);
        const uint8_t* p = image + c * plane_stride;
        for (size_t i = 0; i < n0; ++i) lin[c][i] = lut[p[i]];
    }
    uint32_t w = width, h = height;
    for (int l = 0; l < levels; ++l) {
        const uint32_t lw = sizes[2 * l], lh = sizes[2 * l + 1];
        if (l > 0) {
            downsample(lin, w, h, lw, lh);
            w = lw;
            h = lh;
            if constexpr (sizeof(T) == 4) {
                if (mode.round_linear && mode.plane_format != 0)
                    for (int c = 0; c < 3; ++c)
                        for (T& v : lin[c]) v = ss2o_round_plane(mode.plane_format, v);
            }
        }
        bool any = false;
        for (int c = 0; c < 3; ++c) any |= !ss2o_channel_pruned(c, l, levels);
        if (!any) continue;
        const size_t n = static_cast<size_t>(w) * h;
        for (int c = 0; c < 3; ++c) xyb[l][c].assign(n, T(0));
        const bool round_plane =
            sizeof(T) == 4 && mode.plane_format != 0 &&
            ((mode.plane_round_levels >> l) & 1u) != 0;
        for (size_t i = 0; i < n; ++i) {
            T v[3];
            if constexpr (sizeof(T) == 4) {
                xyb_f32(lin[0][i], lin[1][i], lin[2][i], mode.cbrt_kind,
                        reinterpret_cast<float*>(v));
                if (round_plane) {
                    for (int c = 0; c < 3; ++c)
                        v[c] = mode.plane_format == 1   ? round_half(v[c])
                               : mode.plane_format == 2 ? round_bf16(v[c])
                                                        : round_unorm16(v[c]);
                }
            } else {
                xyb_f64(lin[0][i], lin[1][i], lin[2][i], reinterpret_cast<double*>(v));
            }
            for (int c = 0; c < 3; ++c) xyb[l][c][i] = v[c];
        }
    }
}

template <typename T>
int32_t compute(const ss2o_mode& mode, uint32_t width, uint32_t height,
                const uint8_t* img[2], double norms[108], double sums_out[108]) {
    uint32_t sizes[12];
    const int32_t levels = ss2o_plan_scales(width, height, sizes);
    if (levels < 0) return -1;
    for (int k = 0; k < 108; ++k) norms[k] = 0.0;
    if (sums_out) for (int k = 0; k < 108; ++k) sums_out[k] = 0.0;

    T lut[256];
    if constexpr (sizeof(T) == 4) {
        ss2o_srgb_lut_f32(reinterpret_cast<float*>(lut));
    } else {
        ss2o_srgb_lut_f64(reinterpret_cast<double*>(lut));
    }
    T wt[kTaps];
    if (mode.blur_kind == 1) {
        double d[kTaps];
        cosine_taps_f64(d);
        for (int i = 0; i < kTaps; ++i) wt[i] = static_cast<T>(d[i]);
    } else if constexpr (sizeof(T) == 8) {
        // FP64: the taps before their FP32 rounding (pipeline.cc:279-285).
        double s = 0.0, d[kTaps];
        for (int i = 0; i < kTaps; ++i) {
            const double o = (static_cast<double>(i) - kRadius) / 1.5;
            d[i] = std::exp(-0.5 * o * o);
            s += d[i];
        }
        for (int i = 0; i < kTaps; ++i) wt[i] = static_cast<T>(d[i] / s);
    } else {
        taps_as(wt);
    }

    // XYB planes [image][level][channel].
    std::vector<T> xyb[2][kMaxScales][3];
    for (int im = 0; im < 2; ++im) xyb_pyramid(mode, width, height, levels, sizes, img[im], lut, xyb[im]);

    for (int l = 0; l < levels; ++l) {
        const uint32_t w = sizes[2 * l], h = sizes[2 * l + 1];
        const double one_per_pixel = 1.0 / (static_cast<double>(w) * static_cast<double>(h));
        for (int c = 0; c < 3; ++c) {
            if (ss2o_channel_pruned(c, l, levels)) continue;
            double s[6];
            map_sums<T>(xyb[0][l][c].data(), xyb[1][l][c].data(), w, h,
                        sizeof(T) == 8 ? 0 : mode.blur_order,
                        sizeof(T) == 8 ? 0 : mode.sum_blocks,
                        sizeof(T) == 8 ? 0 : mode.sum_variances, wt, s);
            // vulkan.cc:545-583: mean in double, fourth root of the 4-norms.
            double* out = norms + l * 18;
            out[c * 2 + 0] = one_per_pixel * s[0];
            out[c * 2 + 1] = std::sqrt(std::sqrt(one_per_pixel * s[1]));
            out[6 + c * 4 + 0] = one_per_pixel * s[2];
            out[6 + c * 4 + 1] = std::sqrt(std::sqrt(one_per_pixel * s[3]));
            out[6 + c * 4 + 2] = one_per_pixel * s[4];
            out[6 + c * 4 + 3] = std::sqrt(std::sqrt(one_per_pixel * s[5]));
            if (sums_out) {
                double* so = sums_out + l * 18;
                so[c * 2 + 0] = s[0];
                so[c * 2 + 1] = s[1];
                for (int k = 0; k < 4; ++k) so[6 + c * 4 + k] = s[2 + k];
            }
        }
    }
    return levels;
}

}  // namespace

extern "C" {

float ss2o_round_plane(int32_t plane_format, float v) {
    switch (plane_format) {
        case 1: return round_half(v);
        case 2: return round_bf16(v);
        case 3: return round_unorm16(v);
        default: return v;
    }
}

int32_t ss2o_plan_scales(uint32_t width, uint32_t height, uint32_t sizes[12]) {
    if (width < kMinDim || height < kMinDim) return -1;
    int32_t count = 0;
    uint32_t w = width, h = height;
    while (count < kMaxScales && w >= kMinDim && h >= kMinDim) {
        if (count != 0) {
            w = w / 2 + w % 2;
            h = h / 2 + h % 2;
        }
        sizes[2 * count] = w;
        sizes[2 * count + 1] = h;
        ++count;
    }
    return count;
}

void ss2o_gaussian_taps(float taps[13]) {
    const double sigma = 1.5;
    const int radius = static_cast<int>(std::ceil(3.5 * sigma));
    double w[13], sum = 0.0;
    for (int i = 0; i <= 2 * radius; ++i) {
        const double o = static_cast<double>(i) - radius;
        w[i] = std::exp(-0.5 * (o / sigma) * (o / sigma));
        sum += w[i];
    }
    for (int i = 0; i <= 2 * radius; ++i) taps[i] = static_cast<float>(w[i] / sum);
}

int32_t ss2o_channel_pruned(int32_t channel, int32_t scale, int32_t scales) {
    if (channel < 0 || channel >= 3 || scale < 0 || scale >= scales || scales > kMaxScales)
        return 0;
    const double* w = kWeights[channel * scales + scale];
    double m = w[0];
    for (int k = 1; k < 6; ++k) m = std::max(m, w[k]);
    return m <= 0.01 ? 1 : 0;
}

double ss2o_score(int32_t scales, const double norms[108], int32_t* status) {
    if (status) *status = 0;
    if (scales > kMaxScales || scales < 0) {
        if (status) *status = 1;
        return 0.0;
    }
    double error = 0.0;
    for (int c = 0; c < 3; ++c) {
        for (int s = 0; s < scales; ++s) {
            const double* w = kWeights[c * scales + s];
            const double* m = norms + s * 18;
            error += w[0] * std::fabs(m[c * 2 + 0]);
            error += w[1] * std::fabs(m[6 + c * 4 + 0]);
            error += w[2] * std::fabs(m[6 + c * 4 + 2]);
            error += w[3] * std::fabs(m[c * 2 + 1]);
            error += w[4] * std::fabs(m[6 + c * 4 + 1]);
            error += w[5] * std::fabs(m[6 + c * 4 + 3]);
        }
    }
    if (!std::isfinite(error)) {
        if (status) *status = 1;
        return 0.0;
    }
    const double scaled = error * 0.9562382616834844;
    const double mapped = 2.326765642916932 * scaled -
                          0.020884521182843837 * scaled * scaled +
                          6.248496625763138e-05 * scaled * scaled * scaled;
    if (!std::isfinite(mapped)) {
        if (status) *status = 1;
        return 0.0;
    }
    if (mapped > 0.0) return 100.0 - 10.0 * std::pow(mapped, 0.6276336467831387);
    return 100.0;
}

void ss2o_srgb_lut_f32(float lut[256]) {
    for (int i = 0; i < 256; ++i) lut[i] = srgb_f32(i);
}

void ss2o_srgb_lut_f64(double lut[256]) {
    for (int i = 0; i < 256; ++i) lut[i] = srgb_f64(i);
}

void ss2o_cosine_taps(double taps[13]) { cosine_taps_f64(taps); }

float ss2o_cbrt_dev(float x) { return cbrt_dev(x); }
float ss2o_cbrt_rn(float x) { return cbrt_exact(x); }
float ss2o_rcp_dev(float x) { return rcp_dev(x); }

void ss2o_xyb_f32(float r, float g, float b, int32_t cbrt_kind, float out[3]) {
    xyb_f32(r, g, b, cbrt_kind, out);
}

void ss2o_xyb_f64(double r, double g, double b, double out[3]) { xyb_f64(r, g, b, out); }

void ss2o_blur_f32(const float* in, float* out, uint32_t w, uint32_t h, int32_t order) {
    float wt[kTaps];
    taps_as(wt);
    blur(in, out, w, h, order, wt);
}

void ss2o_blur_f64(const double* in, double* out, uint32_t w, uint32_t h) {
    double wt[kTaps];
    taps_as(wt);
    blur(in, out, w, h, 0, wt);
}

int32_t ss2o_pyramid_f32(uint32_t width, uint32_t height, const uint8_t* image, float* out) {
    return pyramid(width, height, image, out);
}

int32_t ss2o_pyramid_f64(uint32_t width, uint32_t height, const uint8_t* image, double* out) {
    return pyramid(width, height, image, out);
}

int32_t ss2o_xyb_planes(const ss2o_mode* mode, uint32_t width, uint32_t height, const uint8_t* image,
                        float* out) {
    if (mode == nullptr || mode->arith != 1) return -1;
    uint32_t sizes[12];
    const int32_t levels = ss2o_plan_scales(width, height, sizes);
    if (levels < 0) return -1;
    float lut[256];
    ss2o_srgb_lut_f32(lut);
    std::vector<float> xyb[kMaxScales][3];
    xyb_pyramid(*mode, width, height, levels, sizes, image, lut, xyb);
    for (int l = 0; l < levels; ++l) {
        const size_t n = static_cast<size_t>(sizes[2 * l]) * sizes[2 * l + 1];
        for (int c = 0; c < 3; ++c) {
            if (xyb[l][c].empty()) {
                out = std::fill_n(out, n, 0.0f);
            } else {
                out = std::copy(xyb[l][c].begin(), xyb[l][c].end(), out);
            }
        }
    }
    return levels;
}

uint64_t ss2o_image_bytes(uint32_t width, uint32_t height) {
    return 3 * align4(static_cast<uint64_t>(width) * height);
}

int32_t ss2o_compute(const ss2o_mode* mode, uint32_t width, uint32_t height,
                     const uint8_t* image0, const uint8_t* image1, double norms[108],
                     double sums[108]) {
    const uint8_t* img[2] = {image0, image1};
    if (mode->arith == 0) return compute<double>(*mode, width, height, img, norms, sums);
    return compute<float>(*mode, width, height, img, norms, sums);
}

int32_t ss2o_compute_many(const ss2o_mode* mode, uint32_t width, uint32_t height,
                          uint32_t pairs, const uint8_t* const* images0,
                          const uint8_t* const* images1, double* norms,
                          uint32_t threads) {
    if (threads == 0) threads = 1;
    std::atomic<uint32_t> next{0};
    std::atomic<int32_t> result{0};
    auto worker = [&]() {
        for (;;) {
            const uint32_t i = next.fetch_add(1);
            if (i >= pairs) return;
            const int32_t r = ss2o_compute(mode, width, height, images0[i], images1[i],
                                           norms + static_cast<size_t>(i) * 108, nullptr);
            if (r < 0) result = -1; else if (result >= 0) result = r;
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    return result;
}

}  // extern "C"
