// SSIMULACRA 2.1 kernels for CDNA4 (gfx950), wave64.
//
//   prepare  u8 sRGB -> linear (256-entry table) -> XYB, plus the 2x2 box
//            pyramid of the linear planes, writing only the (level, channel)
//            planes whose maps survive pruning (prepare.comp:170-232).
//   scale    one wave per strip (128 or 64 columns) of one (level,
//            channel): the 13-tap blurs of the map's quantities, the SSIM'
//            and edge maps, and six map sums, FP32 over 8-row blocks then
//            FP64 (scale.comp:335-437). Rows stream through the wave: each
//            input row is staged in LDS, blurred horizontally (each lane owns
//            two output columns, or one), and the horizontal results enter a
//            13-row register ring from which the vertical blur of the row six
//            above is taken. No map ever reaches HBM.
//   reduce   FP64 sum of each norm's per-strip partials, 1-norm means and
//            4-norm fourth roots (reduce.comp:21-63, vulkan.cc:545-583).
//
// Two arithmetic forms (MapForm) share the kernels: the reference form
// (prepare.comp and scale.comp in their own operation order, correctly
// rounded cube roots, five blurred quantities) and the difference form (the
// SSIM' map from four blurred quantities, rearranged to avoid a cancellation).
//
// Arithmetic contract. Every FP32 operation below is an IEEE add, multiply,
// fused multiply-add (written explicitly), division, min/max, abs or an
// integer/float conversion, with contraction switched off, so the CPU
// oracle's FP32 forms (ref for the reference form, dev for the difference
// form) reproduce each pixel's values bit for bit. The blur runs horizontal
// pass first with taps in ascending order, as scale.comp:149-301 does.
//
// Memory access. Descriptors (pairs, plans, work items) are read through
// constant-address-space pointers so that their uniform fields become scalar
// loads; images and planes through global-address-space pointers.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>
#include <utility>

#include "ss2_cbrt.h"
#include "ss2_common.h"
#include "ss2_kernels.h"

#pragma clang fp contract(off)

namespace vqm::ss2 {
namespace {

template <typename T> using ConstAS = __attribute__((address_space(4))) const T;
template <typename T> using GlobalAS = __attribute__((address_space(1))) T;

// Builtin vector types for wide loads and stores through qualified pointers.
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ ConstAS<T>* as_const(const T* p) {
    return (ConstAS<T>*)p;
}
template <typename T>
__device__ __forceinline__ GlobalAS<T>* as_global(T* p) {
    return (GlobalAS<T>*)p;
}

// The tap tables SampledTaps and CosineTaps live in ss2_common.h.

// XYB constants of prepare.comp:40-54 rounded to FP32.
constexpr float kM00 = 0x1.333334p-2f, kM01 = 0x1.3e76c8p-1f, kM02 = 0x1.3f7ceep-4f;
constexpr float kM10 = 0x1.d70a3ep-3f, kM11 = 0x1.624dd2p-1f, kM12 = 0x1.3f7ceep-4f;
constexpr float kM20 = 0x1.f28798p-3f, kM21 = 0x1.a35d1ep-3f, kM22 = 0x1.1a86d2p-1f;
constexpr float kOpsinBias = 0x1.f12a6cp-9f;
constexpr float kCbrtBias = 0x1.3f64eap-3f;
constexpr float kC2 = 0.0009f;            // scale.comp:87

__device__ __forceinline__ float fma_(float a, float b, float c) {
    return __builtin_fmaf(a, b, c);
}

// Smallest normal float: the XYB mixes are clamped to it instead of 0, so the
// cube roots never see zero or a denormal. Its root, 2^-42, vanishes in every
// sum it enters (all are at least ~1e-2 away from it), as a zero root would:
// 2^-42 minus the cube-root bias rounds to minus the bias.
constexpr float kMinNormal = 0x1p-126f;

// 1 / x correctly rounded, for x a positive normal float in [2^-24, 2^24):
// the hardware reciprocal and one Newton step r += r (1 - x r), 3
// instructions. vqm_ss2_check_div checks it against the IEEE quotient for
// every significand and the hardware reciprocal's invariance under scaling
// by 2^e, e in [-24, 24) (the Newton step scales exactly).
__device__ __forceinline__ float rcp_rn(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    return fma_(fma_(-x, r, 1.0f), r, r);
}

// a / b correctly rounded (the IEEE quotient), for the map's quotients: b a
// positive normal float in [2^-24, 2^24) and a either 0.0f (never -0.0f) or
// |a| >= 2^-100, with |a / b| inside [2^-100, 2^100]. The hardware
// reciprocal, one Newton step (the reciprocal correctly rounded), the
// product, and one residual correction (Markstein's): 6 instructions, where
// the compiler's `a / b` takes 11 (two corrections, operand scaling and a
// special-value fix-up). vqm_ss2_check_div compares it with the IEEE
// quotient on the device: over every pair of significands (a, b in [1, 2),
// 2^46 pairs, about 30 s) it differs nowhere, and the hardware reciprocal of
// m 2^e is that of m scaled by 2^-e for every e in [-24, 24); the fused
// steps scale exactly, so the whole domain follows. The map's operands stay
// in it: denominators are at least c2 - 2^-22 or 1 and at most a few units,
// and numerators are 0.0f (an exact difference of equal values rounds to
// +0) or multiples of 2^-82.
__device__ __forceinline__ float div_rn(float a, float b) {
    const float r = rcp_rn(b);
    const float q = a * r;
    return fma_(fma_(-b, q, a), r, q);
}

// The two arithmetic forms of the path (MapForm):
//
//   Reference form (Ref = true): prepare.comp:81-98 and scale.comp:130-142,
//   305-333 in their own operation order. The cube roots are correctly
//   rounded (cbrt_rn_hw), the five quantities a, b, a*a, b*b and a*b are
//   blurred, and the SSIM' map subtracts the squared means from the blurred
//   second moments directly, with IEEE divisions.
//
//   Difference form (Ref = false): the cube-root bias cancelled where it
//   cancels exactly, the fast cube root (cbrt_fast), and the SSIM' map from
//   four blurred quantities a, b, (a - b)^2 and a*b (see scale_item).
//
// XYB of one linear sample, only the channels in `mask` (bit 0 X, 1 Y, 2 B).
// With c_i = cbrt(m_i) - bias, the reference computes X = (0.5 (c0 - c1)) 14
// + 0.42, Y = 0.5 (c0 + c1) + 0.01 and B = (c2 - 0.5 (c0 + c1)) + 0.55. The
// difference form takes them from the roots directly:
// X = 7 (cbrt0 - cbrt1) + 0.42, Y = (cbrt0 + cbrt1)/2 + (0.01 - bias) and
// B = cbrt2 - (cbrt0 + cbrt1)/2 + 0.55 (three subtractions fewer, no
// rounding of c_i).
constexpr float kYOffset = 0.01f - kCbrtBias;  // rounded once

template <bool Ref>
__device__ __forceinline__ void xyb(float r, float g, float b, uint32_t mask, float out[3]) {
    float m0 = fma_(kM00, r, fma_(kM01, g, fma_(kM02, b, kOpsinBias)));
    float m1 = fma_(kM10, r, fma_(kM11, g, fma_(kM12, b, kOpsinBias)));
    m0 = fmaxf(m0, kMinNormal);
    m1 = fmaxf(m1, kMinNormal);
    out[2] = 0.0f;
    if constexpr (Ref) {
        const float c0 = cbrt_rn_hw(m0) - kCbrtBias, c1 = cbrt_rn_hw(m1) - kCbrtBias;
        const float half_sum = 0.5f * (c0 + c1);
        out[0] = fma_(0.5f * (c0 - c1), 14.0f, 0.42f);
        out[1] = half_sum + 0.01f;
        if (mask & 4u) {
            float m2 = fma_(kM20, r, fma_(kM21, g, fma_(kM22, b, kOpsinBias)));
            m2 = fmaxf(m2, kMinNormal);
            out[2] = (cbrt_rn_hw(m2) - kCbrtBias - half_sum) + 0.55f;
        }
    } else {
        const float r0 = cbrt_fast(m0), r1 = cbrt_fast(m1);
        const float sum = r0 + r1;
        out[0] = fma_(r0 - r1, 7.0f, 0.42f);
        out[1] = fma_(0.5f, sum, kYOffset);
        if (mask & 4u) {
            float m2 = fma_(kM20, r, fma_(kM21, g, fma_(kM22, b, kOpsinBias)));
            m2 = fmaxf(m2, kMinNormal);
            out[2] = fma_(-0.5f, sum, cbrt_fast(m2)) + 0.55f;
        }
    }
}

// XYB of N pixels (px[j] = linear r, g, b), only the channels in `mask`. The
// reference form takes every cube root's fast part first and settles the
// rare near-midpoint roots in one branch for the group, so the roots'
// instruction chains share a basic block and interleave; the arithmetic of
// each pixel is xyb<true>'s.
template <bool Ref, int N>
__device__ __forceinline__ void xyb_group(const float (*px)[3], uint32_t mask, float (*out)[3]) {
    if constexpr (!Ref) {
#pragma unroll
        for (int j = 0; j < N; ++j) xyb<false>(px[j][0], px[j][1], px[j][2], mask, out[j]);
    } else {
        float m[N][3], c[N][3];
        uint32_t near_bits = 0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const float r = px[j][0], g = px[j][1], b = px[j][2];
            m[j][0] = fmaxf(fma_(kM00, r, fma_(kM01, g, fma_(kM02, b, kOpsinBias))), kMinNormal);
            m[j][1] = fmaxf(fma_(kM10, r, fma_(kM11, g, fma_(kM12, b, kOpsinBias))), kMinNormal);
            bool n0, n1;
            c[j][0] = cbrt_rn_near_hw(m[j][0], &n0);
            c[j][1] = cbrt_rn_near_hw(m[j][1], &n1);
            near_bits |= (static_cast<uint32_t>(n0) | static_cast<uint32_t>(n1) << 1) << (3 * j);
            m[j][2] = 1.0f;
            c[j][2] = 0.0f;
        }
        if (mask & 4u) {  // B: a block of its own (a uniform branch)
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const float r = px[j][0], g = px[j][1], b = px[j][2];
                m[j][2] = fmaxf(fma_(kM20, r, fma_(kM21, g, fma_(kM22, b, kOpsinBias))), kMinNormal);
                bool n2;
                c[j][2] = cbrt_rn_near_hw(m[j][2], &n2);
                near_bits |= static_cast<uint32_t>(n2) << (3 * j + 2);
            }
        }
        if (near_bits != 0u) {
#pragma unroll
            for (int k = 0; k < 3 * N; ++k)
                if ((near_bits >> k) & 1u) c[k / 3][k % 3] = cbrt_round_exact(m[k / 3][k % 3], c[k / 3][k % 3]);
        }
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const float c0 = c[j][0] - kCbrtBias, c1 = c[j][1] - kCbrtBias;
            const float half_sum = 0.5f * (c0 + c1);
            out[j][0] = fma_(0.5f * (c0 - c1), 14.0f, 0.42f);
            out[j][1] = half_sum + 0.01f;
            out[j][2] = (mask & 4u) ? (c[j][2] - kCbrtBias - half_sum) + 0.55f : 0.0f;
        }
    }
}

// ---- plane element formats ------------------------------------------------

// A format loads two adjacent elements as Raw bits (or one element) and
// unpacks them to FP32 only where they are consumed, so that the wait for a prefetched row lands
// one row later instead of right behind the load.
struct F32 {
    using E = float;
    using Raw = f32x2;
    static __device__ __forceinline__ Raw load2(GlobalAS<const E>* p) {
        return *reinterpret_cast<GlobalAS<const f32x2>*>(p);
    }
    static __device__ __forceinline__ float2 unpack(Raw v) { return make_float2(v.x, v.y); }
    static __device__ __forceinline__ float unpack1(E v) { return v; }
    static __device__ __forceinline__ void store1(GlobalAS<E>* p, float v) { *p = v; }
    static __device__ __forceinline__ void store2(GlobalAS<E>* p, float a, float b) {
        *reinterpret_cast<GlobalAS<f32x2>*>(p) = f32x2{a, b};
    }
    static __device__ __forceinline__ void store4(GlobalAS<E>* p, const float v[4]) {
        *reinterpret_cast<GlobalAS<f32x4>*>(p) = f32x4{v[0], v[1], v[2], v[3]};
    }
};

// FP32 -> FP16, round to nearest even, written out in integer operations so
// that every value (ties included) rounds exactly as IEEE 754 specifies and
// as the oracle's restatement does.
__device__ __forceinline__ uint16_t to_f16(float v) {
    const uint32_t x = __float_as_uint(v);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t a = x & 0x7FFFFFFFu;
    if (a >= 0x477FF000u)  // rounds to infinity (65520 and up), or NaN
        return static_cast<uint16_t>(sign | (a > 0x7F800000u ? 0x7E00u : 0x7C00u));
    if (a < 0x38800000u)   // below 2^-14: a multiple of 2^-24, exact product
        return static_cast<uint16_t>(sign | static_cast<uint32_t>(__builtin_rintf(__uint_as_float(a) * 0x1p24f)));
    uint32_t r = a - 0x38000000u;  // rebias the exponent from 127 to 15
    r += 0xFFFu + ((r >> 13) & 1u);
    return static_cast<uint16_t>(sign | (r >> 13));
}
__device__ __forceinline__ float from_f16(uint16_t v) {
    return static_cast<float>(__builtin_bit_cast(_Float16, v));
}
__device__ __forceinline__ uint16_t to_bf16(float v) {
    uint32_t bits = __float_as_uint(v);
    bits += 0x7FFFu + ((bits >> 16) & 1u);
    return static_cast<uint16_t>(bits >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t v) {
    return __uint_as_float(static_cast<uint32_t>(v) << 16);
}

template <uint16_t (*To)(float), float (*From)(uint16_t)>
struct Narrow {
    using E = uint16_t;
    using Raw = uint32_t;
    static __device__ __forceinline__ Raw load2(GlobalAS<const E>* p) {
        return *reinterpret_cast<GlobalAS<const uint32_t>*>(p);
    }
    static __device__ __forceinline__ float2 unpack(Raw v) {
        return make_float2(From(static_cast<uint16_t>(v & 0xFFFFu)),
                           From(static_cast<uint16_t>(v >> 16)));
    }
    static __device__ __forceinline__ float unpack1(E v) { return From(v); }
    static __device__ __forceinline__ void store1(GlobalAS<E>* p, float v) { *p = To(v); }
    static __device__ __forceinline__ void store2(GlobalAS<E>* p, float a, float b) {
        *reinterpret_cast<GlobalAS<uint32_t>*>(p) =
            static_cast<uint32_t>(To(a)) | (static_cast<uint32_t>(To(b)) << 16);
    }
    static __device__ __forceinline__ void store4(GlobalAS<E>* p, const float v[4]) {
        *reinterpret_cast<GlobalAS<u32x2>*>(p) =
            u32x2{static_cast<uint32_t>(To(v[0])) | (static_cast<uint32_t>(To(v[1])) << 16),
                  static_cast<uint32_t>(To(v[2])) | (static_cast<uint32_t>(To(v[3])) << 16)};
    }
};
// FP32 -> unsigned 16-bit fixed point q / 65536, round to nearest even,
// clamped to 0 .. 65535 (NaN and negatives to 0). The XYB values of linear
// RGB in [0, 1] lie in [0, 0.94]: the step is a uniform 2^-16 and zero, the
// plane borders' value, is exact. Scaling by 2^16 and back is exact.
__device__ __forceinline__ uint16_t to_unorm16(float v) {
    const float s = fminf(fmaxf(v * 65536.0f, 0.0f), 65535.0f);
    return static_cast<uint16_t>(__builtin_rintf(s));
}
__device__ __forceinline__ float from_unorm16(uint16_t v) {
    return static_cast<float>(v) * (1.0f / 65536.0f);
}

using F16 = Narrow<to_f16, from_f16>;
using BF16 = Narrow<to_bf16, from_bf16>;
using U16 = Narrow<to_unorm16, from_unorm16>;

// The element format of every stored plane, and RoundLinear: the linear box
// averages of levels >= 1 are rounded to the element format too, before
// their XYB and before the next level averages them -- the reference's
// half-plane build (prepare.comp:100-111, ssimu2_common.glsl:84-96), which
// FP16 planes reproduce.
template <typename FmtT, bool RoundLinear = false>
struct Layout {
    using Fmt = FmtT;
    static constexpr bool kRoundLinear = RoundLinear;
};
using LayoutF32 = Layout<F32>;
using LayoutF16 = Layout<F16, true>;
using LayoutBF16 = Layout<BF16>;
using LayoutU16 = Layout<U16>;

// A level's linear box average as the layout hands it on.
template <typename Lay>
__device__ __forceinline__ float linear_sample(float v) {
    if constexpr (Lay::kRoundLinear) {
        return from_f16(to_f16(v));
    } else {
        return v;
    }
}

// Element pointer of a plane: planes of one image start at `base` and
// `offset` is in bytes.
template <typename Fmt>
__device__ __forceinline__ GlobalAS<typename Fmt::E>* plane_at(GlobalAS<uint8_t>* base, uint64_t offset) {
    return reinterpret_cast<GlobalAS<typename Fmt::E>*>(base + offset);
}

// Workgroup b runs on XCD b % 8. Renumber so each XCD receives a contiguous
// run of logical work items: neighbouring strips and tiles, which share halo
// rows and input cache lines, then share one L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
    constexpr uint32_t kXcds = 8;
    const uint32_t xcd = b % kXcds, idx = b / kXcds;
    const uint32_t per = n / kXcds, rem = n % kXcds;
    return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

// ---- general input decode -----------------------------------------------------
//
// Images that are not 8-bit sRGB RGB are decoded per pixel from their
// SourceDesc, in the stage order of color.glsl:447-541: raw sample with the
// range/depth scale and offset, chroma upsampling, the matrix (or the
// constant-luminance or ICtCp op), alpha, the transfer curve, and the
// primaries matrix to linear BT.709. Powers and exponentials go through the
// hardware base-2 logarithm and exponential (as a shader's pow does), and
// the stages are held to the A/B tolerances of the input tests, not bit
// parity.

namespace decode {

// x^y for x >= 0 from v_log_f32 and v_exp_f32 (about one ulp each): relative
// error about |y log2 x| 2^-23 plus one ulp, some 150 instructions fewer than
// the library powf. Inputs below the smallest normal float are taken as it
// (their powers are 0 to FP32 precision for y >= 1; the curves with y < 1
// map them to 0 after their own offsets).
__device__ __forceinline__ float pow_hw(float x, float y) {
    if (x <= 0.0f) return 0.0f;
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(fmaxf(x, 0x1p-126f)));
}
// e^x and 10^x through v_exp_f32.
__device__ __forceinline__ float exp_hw(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ float exp10_hw(float x) { return __builtin_amdgcn_exp2f(x * 3.32192809488736235f); }

constexpr float kRec709Alpha = 1.09929682680944f;
constexpr float kRec709Beta = 0.018053968510807f;

__device__ float raw_sample(ConstAS<SourceDesc>* d, GlobalAS<const uint8_t>* src, uint32_t plane,
                            uint32_t x, uint32_t y) {
    const uint64_t index = static_cast<uint64_t>(y) * d->plane_w[plane] + x;
    GlobalAS<const uint8_t>* base = src + d->plane_off[plane];
    float v;
    switch (d->type) {
        case 0: v = static_cast<float>(base[index]); break;
        case 1: v = static_cast<float>(reinterpret_cast<GlobalAS<const uint16_t>*>(base)[index]); break;
        case 2: v = static_cast<float>(reinterpret_cast<GlobalAS<const _Float16>*>(base)[index]); break;
        default: v = reinterpret_cast<GlobalAS<const float>*>(base)[index]; break;
    }
    // Sample class: 2 alpha, 1 the chroma planes of YUV, 0 luma or RGB.
    const uint32_t cls = plane == 3 ? 2u : (plane != 0 && d->family == 2 ? 1u : 0u);
    return v * d->scale[cls] + d->offset[cls];
}

// BT.709 inverse OETF (the CL op and the xvYCC / BT.1361 extensions).
__device__ float rec709_inverse_oetf(float x) {
    x = fmaxf(x, 0.0f);
    return x < 4.5f * kRec709Beta ? x / 4.5f
                                  : pow_hw((x + (kRec709Alpha - 1.0f)) / kRec709Alpha, 1.0f / 0.45f);
}

// IEC 61966-2-1.
__device__ float srgb_eotf(float v) {
    return v <= 0.04045f ? v / 12.92f : pow_hw((v + 0.055f) / 1.055f, 2.4f);
}

// SMPTE ST 2084, normalised to 10000 cd/m^2.
__device__ float pq_eotf(float x) {
    if (x <= 0.0f) return 0.0f;
    const float p = pow_hw(x, 1.0f / 78.84375f);
    const float num = fmaxf(p - 0.8359375f, 0.0f);
    const float den = fmaxf(18.8515625f - 18.6875f * p, 1.17549435e-38f);
    return pow_hw(num / den, 1.0f / 0.1593017578125f);
}

// ARIB STD-B67 inverse OETF (scene light).
__device__ float hlg_inverse_oetf(float x) {
    x = fmaxf(x, 0.0f);
    return x <= 0.5f ? (x * x) * (1.0f / 3.0f) : (exp_hw((x - 0.55991073f) / 0.17883277f) + 0.28466892f) / 12.0f;
}

__device__ float curve(uint32_t id, float v, float scale) {
    switch (id) {
        case kCurveLinear: return v;
        case kCurvePow24: return v < 0.0f ? 0.0f : pow_hw(v, 2.4f);
        case kCurvePow22: return v < 0.0f ? 0.0f : pow_hw(v, 2.2f);
        case kCurvePow28: return v < 0.0f ? 0.0f : pow_hw(v, 2.8f);
        case kCurveSrgb: return srgb_eotf(v);
        case kCurveLog100: return v <= 0.0f ? 0.01f : exp10_hw(2.0f * (v - 1.0f));
        case kCurveLog316: return v <= 0.0f ? 0.00316227766f : exp10_hw(2.5f * (v - 1.0f));
        case kCurveXvycc:
            if (v >= 0.0f && v <= 1.0f) return pow_hw(v, 2.4f);
            return v < 0.0f ? -rec709_inverse_oetf(-v) : rec709_inverse_oetf(v);
        case kCurveBt1361:
            if (v >= 0.0f && v <= 1.0f) return pow_hw(v, 2.4f);
            return v >= 0.0f ? rec709_inverse_oetf(v) : -rec709_inverse_oetf(-4.0f * v) / 4.0f;
        case kCurvePq: return scale * pq_eotf(v);
        case kCurveSt428: return v < 0.0f ? 0.0f : (52.37f / 48.0f) * pow_hw(v, 2.6f);
        default: return v;
    }
}

__device__ float3 apply(ConstAS<float>* m, float3 v) {
    return make_float3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
                       m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

// Constant-luminance Y'CbCr to linear RGB through the BT.709 curve.
__device__ float3 cl_to_rgb(ConstAS<SourceDesc>* d, float3 e) {
    const float u = e.y, v = e.z;
    const float b_minus_y = u * 2.0f * (u < 0.0f ? d->cl[0] : d->cl[1]);
    const float r_minus_y = v * 2.0f * (v < 0.0f ? d->cl[2] : d->cl[3]);
    const float b = rec709_inverse_oetf(b_minus_y + e.x);
    const float r = rec709_inverse_oetf(r_minus_y + e.x);
    const float luma = rec709_inverse_oetf(e.x);
    return make_float3(r, (luma - d->kr * r - d->kb * b) / d->kg, b);
}

// Resampling kernels: triangle, Catmull-Rom (B = 0, C = 0.5), Lanczos-3.
__device__ float sinc(float x) {
    const float pix = 3.14159265358979323846f * x;
    return x == 0.0f ? 1.0f : sinf(pix) / pix;
}
__device__ float filter_weight(uint32_t filter, float x) {
    x = fabsf(x);
    if (filter == 0) return fmaxf(1.0f - x, 0.0f);
    if (filter == 1) {
        if (x < 1.0f) return 1.0f + x * x * (-2.5f + x * 1.5f);
        if (x < 2.0f) return 2.0f + x * (-4.0f + x * (2.5f - 0.5f * x));
        return 0.0f;
    }
    return x < 3.0f ? sinc(x) * sinc(x / 3.0f) : 0.0f;
}

// Rounds half up such that round(x - 1) == round(x) - 1.
__device__ float round_halfup(float x) { return x < 0.0f ? floorf(x + 0.5f) : floorf(x + 0.49999997f); }

// A tap position mirrored once at the plane edges, then clamped.
__device__ uint32_t mirror(float pos, uint32_t dim) {
    const float extent = static_cast<float>(dim);
    const float m = pos < 0.0f ? -pos : (pos >= extent ? 2.0f * extent - pos : pos);
    return static_cast<uint32_t>(fminf(fmaxf(m, 0.0f), extent - 1.0f));
}

constexpr int kMaxTaps = 6;

__device__ __forceinline__ int tap_count(uint32_t filter) { return filter == 0 ? 2 : filter == 1 ? 4 : 6; }

// Chroma position of output sample `out` on one axis: (out + 0.5) / ratio
// + shift.
__device__ __forceinline__ float chroma_pos(uint32_t out, uint32_t sub, float shift) {
    return (static_cast<float>(out) + 0.5f) * __builtin_ldexpf(1.0f, -static_cast<int>(sub)) + shift;
}

// The first tap's sample index before mirroring, round(pos - taps / 2): an
// integer-valued float, non-decreasing in pos.
__device__ __forceinline__ float first_tap(int taps, float pos) {
    return round_halfup(pos - 0.5f * static_cast<float>(taps));
}

// The weights of one axis for output sample `out`, taps centred half a
// sample past the first tap. Bilinear weights are 1 - t and t (their sum is
// exactly one); the others are normalised by their sum. Returns the tap
// count; *first is the first tap's unmirrored index. Bilinear (filter 0)
// callers set Bilinear, so the tap loops unroll and the weights stay in
// registers.
template <bool Bilinear>
__device__ __forceinline__ int axis_weights(uint32_t filter, uint32_t out, uint32_t sub, float shift,
                                            float (&w)[Bilinear ? 2 : kMaxTaps], int* first) {
    const int taps = Bilinear ? 2 : tap_count(filter);
    const float pos = chroma_pos(out, sub, shift);
    const float n0 = first_tap(taps, pos);
    const float base = n0 + 0.5f;
    if (Bilinear || filter == 0) {
        const float t = pos - base;
        w[0] = 1.0f - t;
        w[1] = t;
    } else {
        float total = 0.0f;
        for (int j = 0; j < taps; ++j) total += filter_weight(filter, base + static_cast<float>(j) - pos);
        for (int j = 0; j < taps; ++j) w[j] = filter_weight(filter, base + static_cast<float>(j) - pos) / total;
    }
    *first = static_cast<int>(n0);
    return taps;
}

// Sample index of unmirrored tap index n: the tap's centre n + 0.5 mirrored
// once at the plane edges, then clamped.
__device__ __forceinline__ uint32_t tap_index(int n, uint32_t dim) {
    return mirror(static_cast<float>(n) + 0.5f, dim);
}

// The chroma samples a tile's pixels read, staged in LDS: both chroma planes
// (scaled and offset) interleaved as (U, V), rows of `cols` entries from
// unmirrored tap index (x0, y0). Over every subsampling (1x, 2x, 4x per
// axis), siting and filter, a 64 x 32 tile needs at most 69 x 22 = 1518
// entries (Lanczos-3, 2x vertical subsampling only); the bound was found by
// enumerating tile positions on frames up to 3840 x 2160.
constexpr uint32_t kChromaWindow = 1664;
struct ChromaWindow {
    int x0, y0;
    uint32_t cols, rows;
};

// The window of pixels x_first .. x_last by y_first .. y_last; false if it
// does not fit in kChromaWindow entries.
__device__ bool chroma_window(ConstAS<SourceDesc>* d, uint32_t x_first, uint32_t x_last, uint32_t y_first,
                              uint32_t y_last, ChromaWindow* w) {
    const int taps = tap_count(d->filter);
    const int x0 = static_cast<int>(first_tap(taps, chroma_pos(x_first, d->sub_w, d->shift_x)));
    const int x1 = static_cast<int>(first_tap(taps, chroma_pos(x_last, d->sub_w, d->shift_x)));
    const int y0 = static_cast<int>(first_tap(taps, chroma_pos(y_first, d->sub_h, d->shift_y)));
    const int y1 = static_cast<int>(first_tap(taps, chroma_pos(y_last, d->sub_h, d->shift_y)));
    w->x0 = x0;
    w->y0 = y0;
    w->cols = static_cast<uint32_t>(x1 - x0 + taps);
    w->rows = static_cast<uint32_t>(y1 - y0 + taps);
    return w->cols * w->rows <= kChromaWindow;
}

// Block-wide: thread `tid` of 256 loads every 256th entry of the window.
__device__ void stage_chroma(ConstAS<SourceDesc>* d, GlobalAS<const uint8_t>* src, const ChromaWindow& w,
                             uint32_t tid, float2* window) {
    const uint32_t n = w.cols * w.rows;
    for (uint32_t i = tid; i < n; i += 256u) {
        const uint32_t r = i / w.cols, c = i - r * w.cols;
        const uint32_t ix = tap_index(w.x0 + static_cast<int>(c), d->plane_w[1]);
        const uint32_t iy = tap_index(w.y0 + static_cast<int>(r), d->plane_h[1]);
        window[i] = make_float2(raw_sample(d, src, 1, ix, iy), raw_sample(d, src, 2, ix, iy));
    }
}

// Both chroma samples of pixel (x, y): the two chroma planes share their
// size and siting, so one set of taps serves both. Rows are summed first,
// then weighted by the vertical taps. With a staged window the samples come
// from LDS (the same values, in the same order), else from the planes.
// Bilinear: the descriptor's filter is 0.
template <bool Bilinear>
__device__ __forceinline__ float2 chroma_samples(ConstAS<SourceDesc>* d, GlobalAS<const uint8_t>* src,
                                                 const float2* window, const ChromaWindow& cw, uint32_t x,
                                                 uint32_t y) {
    float wx[Bilinear ? 2 : kMaxTaps], wy[Bilinear ? 2 : kMaxTaps];
    int fx, fy;
    const int tx = axis_weights<Bilinear>(d->filter, x, d->sub_w, d->shift_x, wx, &fx);
    const int ty = axis_weights<Bilinear>(d->filter, y, d->sub_h, d->shift_y, wy, &fy);
    float sum_u = 0.0f, sum_v = 0.0f;
    if (window != nullptr) {
        // The first taps are non-decreasing in x and y, so they lie within
        // the window; the clamp only guards the LDS bounds.
        const uint32_t cx = min(static_cast<uint32_t>(fx - cw.x0), cw.cols - static_cast<uint32_t>(tx));
        const uint32_t cy = min(static_cast<uint32_t>(fy - cw.y0), cw.rows - static_cast<uint32_t>(ty));
        for (int j = 0; j < ty; ++j) {
            const float2* row = window + (cy + static_cast<uint32_t>(j)) * cw.cols + cx;
            float row_u = 0.0f, row_v = 0.0f;
            for (int k = 0; k < tx; ++k) {
                const float2 s = row[k];
                row_u += wx[k] * s.x;
                row_v += wx[k] * s.y;
            }
            sum_u += wy[j] * row_u;
            sum_v += wy[j] * row_v;
        }
    } else {
        for (int j = 0; j < ty; ++j) {
            const uint32_t iy = tap_index(fy + j, d->plane_h[1]);
            float row_u = 0.0f, row_v = 0.0f;
            for (int k = 0; k < tx; ++k) {
                const uint32_t ix = tap_index(fx + k, d->plane_w[1]);
                row_u += wx[k] * raw_sample(d, src, 1, ix, iy);
                row_v += wx[k] * raw_sample(d, src, 2, ix, iy);
            }
            sum_u += wy[j] * row_u;
            sum_v += wy[j] * row_v;
        }
    }
    return make_float2(sum_u, sum_v);
}

// Plane 0's samples (luma, R or grey) of pixels bx .. bx + 3 of row y,
// clamped to the image: one 4- or 8-byte load for byte and word samples on
// rows of a multiple of 4 samples, else one load a pixel.
__device__ __forceinline__ void first_plane4(ConstAS<SourceDesc>* d, GlobalAS<const uint8_t>* src, uint32_t bx,
                                             uint32_t y, uint32_t W, float (&out)[4]) {
    const uint32_t pw = d->plane_w[0];
    if (d->type <= 1u && (pw & 3u) == 0u && bx + 3u < W) {
        const uint64_t index = static_cast<uint64_t>(y) * pw + bx;
        GlobalAS<const uint8_t>* base = src + d->plane_off[0];
        const float scale = d->scale[0], offset = d->offset[0];
        if (d->type == 0u) {
            const uint32_t w = *reinterpret_cast<GlobalAS<const uint32_t>*>(base + index);
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) out[j] = static_cast<float>((w >> (8u * j)) & 255u) * scale + offset;
        } else {
            const u32x2 w = *reinterpret_cast<GlobalAS<const u32x2>*>(base + 2u * index);
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                const uint32_t word = (j < 2u ? w[0] : w[1]) >> (16u * (j & 1u));
                out[j] = static_cast<float>(word & 0xFFFFu) * scale + offset;
            }
        }
    } else {
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) out[j] = raw_sample(d, src, 0, min(bx + j, W - 1u), y);
    }
}

// Linear BT.709 RGB of pixel (x, y), whose plane 0 sample is y0. Inlined into
// its eight call sites (a called function measured 12% slower on 8-bit 4:2:0
// frames). `window` is the tile's staged chroma, or null to read the chroma
// planes directly.
__device__ __forceinline__ float3 pixel(ConstAS<SourceDesc>* d, GlobalAS<const uint8_t>* src, const float2* window,
                                       const ChromaWindow& cw, uint32_t x, uint32_t y, float y0) {
    const float alpha = d->alpha != 0 ? raw_sample(d, src, 3, x, y) : 1.0f;
    float3 e = make_float3(y0, y0, y0);
    if (d->family == 2 && (d->sub_w != 0 || d->sub_h != 0)) {
        const float2 uv = d->filter == 0 ? chroma_samples<true>(d, src, window, cw, x, y)
                                         : chroma_samples<false>(d, src, window, cw, x, y);
        e.y = uv.x;
        e.z = uv.y;
    } else if (d->family != 0) {
        e.y = raw_sample(d, src, 1, x, y);
        e.z = raw_sample(d, src, 2, x, y);
    }
    float3 rgb = d->mode == 1 ? cl_to_rgb(d, e) : (d->identity1 ? e : apply(d->m1, e));
    if (d->alpha == 2) {
        const float a = fminf(fmaxf(alpha, 0.0f), 1.0f);
        rgb = a == 0.0f ? make_float3(0.0f, 0.0f, 0.0f) : make_float3(rgb.x / a, rgb.y / a, rgb.z / a);
    }
    if (d->alpha != 0) {
        const float target = (1.0f - alpha) * d->background;
        rgb = make_float3(alpha * rgb.x + target, alpha * rgb.y + target, alpha * rgb.z + target);
    }
    float3 lin = rgb;
    if (d->mode != 1) {
        if (d->transfer == kCurveHlg) {
            // Per-channel inverse OETF, then the luminance OOTF and the peak.
            const float3 c = make_float3(hlg_inverse_oetf(rgb.x), hlg_inverse_oetf(rgb.y), hlg_inverse_oetf(rgb.z));
            const float luma = fmaxf(d->kr * c.x + d->kg * c.y + d->kb * c.z, 1.17549435e-38f);
            const float k = pow_hw(luma, 0.2f) * d->to_linear_scale;
            lin = make_float3(c.x * k, c.y * k, c.z * k);
        } else {
            lin = make_float3(curve(d->transfer, rgb.x, d->to_linear_scale),
                              curve(d->transfer, rgb.y, d->to_linear_scale),
                              curve(d->transfer, rgb.z, d->to_linear_scale));
        }
    }
    return d->identity2 ? lin : apply(d->m2, lin);
}

}  // namespace decode

// ---- prepare ----------------------------------------------------------------

// Level 1's tile is 32 x 16 samples; every later level is a quarter of the
// previous one. The ladder keeps a level's linear samples channel-planar.
constexpr uint32_t kLadder = (kPrepTileW / 2) * (kPrepTileH / 2);

template <typename Lay, bool Ref, uint32_t L>
__device__ __forceinline__ void prepare_level(ConstAS<Plan>* plan, GlobalAS<uint8_t>* planes,
                                              uint32_t X0, uint32_t Y0, uint32_t tid,
                                              float (*ladder)[3][kLadder]) {
    using Fmt = typename Lay::Fmt;
    // Level L from the ladder level L-1 published, with the clamp of level
    // L-1's last row and column applied in this tile's coordinates
    // (prepare.comp:146-168).
    constexpr uint32_t side_x = kPrepTileW >> L, side_y = kPrepTileH >> L;
    constexpr uint32_t prev_x = side_x * 2u;
    const float (*prev)[kLadder] = ladder[(L - 1u) & 1u];
    float (*next)[kLadder] = ladder[L & 1u];
    if (tid >= side_x * side_y) return;
    const uint32_t jx = tid % side_x, jy = tid / side_x;
    const uint32_t gx = (X0 >> L) + jx, gy = (Y0 >> L) + jy;
    float v[3] = {0.0f, 0.0f, 0.0f};
    if (gx < plan->w[L] && gy < plan->h[L]) {
        const uint32_t ox = X0 >> (L - 1u), oy = Y0 >> (L - 1u);
        const uint32_t cx0 = 2u * jx, cy0 = 2u * jy;
        // The block's right column is cx0 + 1 unless the level's last
        // column clamps it onto cx0 (odd widths); each row's pair is one
        // 8-byte read (cx0 is even, cx0 + 1 stays inside the ladder row), so
        // a wave reads whole contiguous rows without bank conflicts.
        const bool clamp_x = min(2u * gx + 1u, plan->w[L - 1] - 1u) - ox == cx0;
        const uint32_t cy1 = min(2u * gy + 1u, plan->h[L - 1] - 1u) - oy;
#pragma unroll
        for (uint32_t c = 0; c < 3; ++c) {
            const float2 top = *reinterpret_cast<const float2*>(&prev[c][cy0 * prev_x + cx0]);
            const float2 bottom = *reinterpret_cast<const float2*>(&prev[c][cy1 * prev_x + cx0]);
            float s = top.x;
            s = s + (clamp_x ? top.x : top.y);
            s = s + bottom.x;
            s = s + (clamp_x ? bottom.x : bottom.y);
            v[c] = linear_sample<Lay>(s * 0.25f);
        }
        const uint32_t mask = plan->mask[L];
        if (mask != 0u) {
            float o[3];
            xyb<Ref>(v[0], v[1], v[2], mask, o);
            const uint64_t at = static_cast<uint64_t>(gy + kPadTop) * plan->pitch[L] + gx + kPadLeft;
#pragma unroll
            for (uint32_t c = 0; c < 3; ++c)
                if (mask & (1u << c)) Fmt::store1(plane_at<Fmt>(planes, plan->plane_off[L][c]) + at, o[c]);
        }
    }
#pragma unroll
    for (uint32_t c = 0; c < 3; ++c) next[c][jy * side_x + jx] = v[c];
}

// Work item b's image, pair and tile origin (the launch's work list, or the
// arithmetic order of a uniform batch), either one read in xcd_remap order.
struct TileRef {
    ConstAS<PairDesc>* pdp;
    ConstAS<Plan>* plan;
    uint32_t image, X0, Y0;
};

__device__ __forceinline__ TileRef resolve_tile(uint32_t b, const PairDesc* __restrict__ pairs,
                                                const WorkRef* __restrict__ work, uint32_t blocks,
                                                uint32_t tiles_per_image) {
    uint32_t pair_idx, image, tile;
    if (work != nullptr) {
        ConstAS<WorkRef>* w = as_const(work) + xcd_remap(b, blocks);
        pair_idx = w->pair;
        image = w->index >> 31;
        tile = w->index & 0x7FFFFFFFu;
    } else {
        // One division: l = (pair * 2 + image) * tiles_per_image + tile.
        const uint32_t l = xcd_remap(b, blocks);
        const uint32_t q = l / tiles_per_image;
        tile = l - q * tiles_per_image;
        image = q & 1u;
        pair_idx = q >> 1;
    }
    TileRef t;
    t.pdp = as_const(pairs) + pair_idx;
    t.plan = as_const(t.pdp->plan);
    t.image = image;
    const uint32_t tiles_x = t.plan->prep_tiles_x;
    const uint32_t ty = tile / tiles_x;
    t.X0 = (tile - ty * tiles_x) * kPrepTileW;
    t.Y0 = ty * kPrepTileH;
    return t;
}

// The 8-bit table path's source bytes of this thread's 4 x 2 pixels: one
// 32-bit word per row and channel, when the row segment is whole and
// 4-aligned (valid); loaded one tile ahead so their latency overlaps the
// current tile's arithmetic.
struct TileWords {
    uint32_t w[2][3];
    bool valid;
};

__device__ __forceinline__ TileWords load_tile_words(const TileRef& t, const uint8_t* src0, const uint8_t* src1) {
    TileWords out;
    out.valid = false;
    const uint32_t tid = threadIdx.x;
    const uint32_t W = t.plan->width, H = t.plan->height;
    const uint32_t bx = t.X0 + 4u * (tid % 16u);
    if (W % 4u != 0u || bx + 3u >= W) return out;
    const uint8_t* src_ptr = t.image ? (src1 != nullptr ? src1 : t.pdp->src[1])
                                     : (src0 != nullptr ? src0 : t.pdp->src[0]);
    GlobalAS<const uint8_t>* src = as_global(src_ptr);
    const uint64_t plane_stride = t.plan->plane_stride;
#pragma unroll
    for (uint32_t r = 0; r < 2; ++r) {
        const uint32_t y = min(t.Y0 + 2u * (tid / 16u) + r, H - 1u);
#pragma unroll
        for (uint32_t c = 0; c < 3; ++c)
            out.w[r][c] = *reinterpret_cast<GlobalAS<const uint32_t>*>(src + c * plane_stride +
                                                                       static_cast<uint64_t>(y) * W + bx);
    }
    out.valid = true;
    return out;
}

// One prepare work item: tile `b` of the launch's work (a 64 x 32 base-pixel
// tile of one image of one pair). Every thread of the block calls it with the
// same `b`; the LUT's LDS stores are issued before the first call.
template <typename Lay, bool General, bool Ref>
__device__ __forceinline__ void prepare_tile(uint32_t b, const PairDesc* __restrict__ pairs,
                                             const WorkRef* __restrict__ work, uint32_t blocks,
                                             uint32_t tiles_per_image, const float* lut,
                                             float (*ladder)[3][kLadder], const uint8_t* src0,
                                             const uint8_t* src1, TileWords pre) {
    const uint32_t tid = threadIdx.x;
    const TileRef tr = resolve_tile(b, pairs, work, blocks, tiles_per_image);
    ConstAS<PairDesc>* pdp = tr.pdp;
    ConstAS<Plan>* plan = tr.plan;
    const uint32_t image = tr.image, X0 = tr.X0, Y0 = tr.Y0;
    const uint32_t W = plan->width, H = plan->height, levels = plan->levels;
    const uint32_t tx = tid % 16u, ty = tid / 16u;
    GlobalAS<uint8_t>* planes = as_global(static_cast<uint8_t*>(pdp->planes)) + (image ? plan->image_bytes : 0);
    // Single-pair launches name their inputs directly; batches use pd.src.
    const uint8_t* src_ptr = image ? (src1 != nullptr ? src1 : pdp->src[1])
                                   : (src0 != nullptr ? src0 : pdp->src[0]);
    GlobalAS<const uint8_t>* src = as_global(src_ptr);
    const uint64_t plane_stride = plan->plane_stride;
    // The decode table's load (issued before the first tile) overlaps the
    // descriptor loads above; this barrier publishes it.
    __syncthreads();

    // Level 0: this thread's 4 x 2 base pixels, coordinates clamped to the
    // image so the level-1 box below sees the clamped samples
    // (color.glsl:453-462).
    float lin[2][4][3];
    const uint32_t bx = X0 + 4u * tx;
    bool table = true;
    if constexpr (General) {
        // Formats other than 8-bit sRGB RGB decode per pixel (block-uniform).
        const SourceDesc* sdp = image ? pdp->source[1] : pdp->source[0];
        if (sdp != nullptr && as_const(sdp)->kind != 0u) {
            table = false;
            ConstAS<SourceDesc>* sd = as_const(sdp);
            // Subsampled chroma: the block stages the tile's chroma samples
            // in LDS once, instead of each pixel loading its taps from the
            // planes (8 loads a pixel for bilinear 4:2:0, 72 for Lanczos-3).
            __shared__ float2 chroma[decode::kChromaWindow];
            decode::ChromaWindow cw{0, 0, 0, 0};
            const float2* window = nullptr;
            if (sd->family == 2 && (sd->sub_w != 0 || sd->sub_h != 0) &&
                decode::chroma_window(sd, X0, min(X0 + kPrepTileW - 1u, W - 1u), Y0,
                                      min(Y0 + kPrepTileH - 1u, H - 1u), &cw)) {
                decode::stage_chroma(sd, src, cw, tid, chroma);
                window = chroma;
                __syncthreads();
            }
            for (uint32_t r = 0; r < 2; ++r) {
                const uint32_t y = min(Y0 + 2u * ty + r, H - 1u);
                float first[4];
                decode::first_plane4(sd, src, bx, y, W, first);
                for (uint32_t j = 0; j < 4; ++j) {
                    const float3 c = decode::pixel(sd, src, window, cw, min(bx + j, W - 1u), y, first[j]);
                    lin[r][j][0] = c.x;
                    lin[r][j][1] = c.y;
                    lin[r][j][2] = c.z;
                }
            }
        }
    }
    if (table) {
        const bool vec = (W % 4u == 0u) && (bx + 3u < W);
#pragma unroll
        for (uint32_t r = 0; r < 2; ++r) {
            const uint32_t y = min(Y0 + 2u * ty + r, H - 1u);
#pragma unroll
            for (uint32_t c = 0; c < 3; ++c) {
                GlobalAS<const uint8_t>* row = src + c * plane_stride + static_cast<uint64_t>(y) * W;
                if (vec) {
                    const uint32_t v = pre.valid ? pre.w[r][c] : *reinterpret_cast<GlobalAS<const uint32_t>*>(row + bx);
#pragma unroll
                    for (uint32_t j = 0; j < 4; ++j) lin[r][j][c] = lut[(v >> (8u * j)) & 255u];
                } else {
#pragma unroll
                    for (uint32_t j = 0; j < 4; ++j) lin[r][j][c] = lut[row[min(bx + j, W - 1u)]];
                }
            }
        }
    }
    const uint32_t mask0 = plan->mask[0];
    if (mask0 != 0u) {
        const uint32_t pitch0 = plan->pitch[0];
#pragma unroll
        for (uint32_t r = 0; r < 2; ++r) {
            const uint32_t y = Y0 + 2u * ty + r;
            if (y >= H) continue;
            float v[4][3];
            // Two pixels' roots at a time (see prepare_kernel's registers).
            xyb_group<Ref, 2>(lin[r], mask0, v);
            xyb_group<Ref, 2>(lin[r] + 2, mask0, v + 2);
            const uint64_t at = static_cast<uint64_t>(y + kPadTop) * pitch0 + bx + kPadLeft;
#pragma unroll
            for (uint32_t c = 0; c < 3; ++c) {
                if (!(mask0 & (1u << c))) continue;
                using Fmt = typename Lay::Fmt;
                auto* p = plane_at<Fmt>(planes, plan->plane_off[0][c]) + at;
                if (bx + 3u < W) {
                    const float q[4] = {v[0][c], v[1][c], v[2][c], v[3][c]};
                    Fmt::store4(p, q);
                } else {
#pragma unroll
                    for (uint32_t j = 0; j < 4; ++j)
                        if (bx + j < W) Fmt::store1(p + j, v[j][c]);
                }
            }
        }
    }
    if (levels < 2) return;

    // Level 1: two samples per thread, each the box of a 2x2 block in
    // (0,0), (1,0), (0,1), (1,1) order (prepare.comp:100-111).
    {
        const uint32_t W1 = plan->w[1], H1 = plan->h[1];
        const uint32_t gy = Y0 / 2u + ty;
        float v1[2][3];
#pragma unroll
        for (uint32_t c = 0; c < 3; ++c) {
#pragma unroll
            for (uint32_t jj = 0; jj < 2; ++jj) {
                float s = lin[0][2 * jj][c];
                s = s + lin[0][2 * jj + 1][c];
                s = s + lin[1][2 * jj][c];
                s = s + lin[1][2 * jj + 1][c];
                v1[jj][c] = linear_sample<Lay>(s * 0.25f);
            }
            *reinterpret_cast<float2*>(&ladder[1][c][ty * 32u + 2u * tx]) = make_float2(v1[0][c], v1[1][c]);
        }
        const uint32_t mask1 = plan->mask[1];
        if (mask1 != 0u && gy < H1) {
            const uint32_t gx = X0 / 2u + 2u * tx;
            float xo[2][3];
            xyb_group<Ref, 2>(v1, mask1, xo);
            const uint64_t at = static_cast<uint64_t>(gy + kPadTop) * plan->pitch[1] + gx + kPadLeft;
#pragma unroll
            for (uint32_t c = 0; c < 3; ++c) {
                if (!(mask1 & (1u << c))) continue;
                using Fmt = typename Lay::Fmt;
                auto* p = plane_at<Fmt>(planes, plan->plane_off[1][c]) + at;
                if (gx + 1u < W1) {
                    Fmt::store2(p, xo[0][c], xo[1][c]);
                } else if (gx < W1) {
                    Fmt::store1(p, xo[0][c]);
                }
            }
        }
    }

    // Levels 2 .. 5: each waits for the previous level's ladder.
    if (levels > 2) { __syncthreads(); prepare_level<Lay, Ref, 2>(plan, planes, X0, Y0, tid, ladder); }
    if (levels > 3) { __syncthreads(); prepare_level<Lay, Ref, 3>(plan, planes, X0, Y0, tid, ladder); }
    if (levels > 4) { __syncthreads(); prepare_level<Lay, Ref, 4>(plan, planes, X0, Y0, tid, ladder); }
    // Nothing reads level 5's samples but its own stored channels, and with
    // six scales the pruning leaves it none.
    if (levels > 5 && plan->mask[5] != 0u) {
        __syncthreads();
        prepare_level<Lay, Ref, 5>(plan, planes, X0, Y0, tid, ladder);
    }
}

// Work items blockIdx.x, blockIdx.x + gridDim.x, ... of `blocks`: one each
// when the grid covers the work, several per block for a persistent grid
// sized to the resident occupancy. The grid is a multiple of 8, so item b
// stays on XCD b % 8 and xcd_remap keeps neighbouring tiles on one L2.
// The table-decode instantiations are held to 80 VGPRs (six waves per SIMD)
// in the difference form and 96 (five) in the reference form, whose
// interleaved correctly rounded cube roots hold FP64 temporaries (level 0
// takes its four pixels' roots two pixels at a time: four at once need 117
// registers and spill at five waves; at six waves pairs spill too), both
// without spilling. The per-pixel decode of general formats is held to 96
// (five waves, 26 KB of LDS a block) without spilling vector registers; it
// would take 98 and four waves unconstrained.
template <typename Lay, bool General, bool Ref>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(General ? 5 : (Ref ? 5 : 6)))) void prepare_kernel(const PairDesc* __restrict__ pairs,
                                                      const WorkRef* __restrict__ work,
                                                      uint32_t blocks, uint32_t tiles_per_image,
                                                      const float* __restrict__ lut_g,
                                                      const uint8_t* src0, const uint8_t* src1) {
    __shared__ float lut[256];
    // Two ladders alternate between levels: level l publishes into l & 1.
    __shared__ __align__(16) float ladder[2][3][kLadder];
    lut[threadIdx.x] = as_global(lut_g)[threadIdx.x];  // published inside prepare_tile
    if constexpr (General) {
        TileWords none;
        none.valid = false;
        for (uint32_t b = blockIdx.x; b < blocks; b += gridDim.x) {
            prepare_tile<Lay, General, Ref>(b, pairs, work, blocks, tiles_per_image, lut, ladder, src0, src1, none);
            __syncthreads();  // the next tile reuses the ladder
        }
    } else {
        // The 8-bit table path loads each tile's source bytes at kernel
        // entry, and each next tile's before working on the current one.
        uint32_t b = blockIdx.x;
        if (b >= blocks) return;
        TileWords cur = load_tile_words(resolve_tile(b, pairs, work, blocks, tiles_per_image), src0, src1);
        for (; b < blocks; b += gridDim.x) {
            const uint32_t next = b + gridDim.x;
            TileWords ahead;
            ahead.valid = false;
            if (next < blocks)
                ahead = load_tile_words(resolve_tile(next, pairs, work, blocks, tiles_per_image), src0, src1);
            prepare_tile<Lay, General, Ref>(b, pairs, work, blocks, tiles_per_image, lut, ladder, src0, src1, cur);
            __syncthreads();  // the next tile reuses the ladder
            cur = ahead;
        }
    }
}

// ---- scale ------------------------------------------------------------------

constexpr int kRing = 8;  // raw rows kept in LDS: the map needs the row six back

// Blurred quantities and the products staged beside the raw rows, by form:
// the reference form blurs a, b, a*a, b*b and a*b (scale.comp:130-142); the
// difference form a, b, (a - b)^2 and a*b.
template <bool Ref> struct MapQuantities {
    static constexpr int kBlurred = Ref ? 5 : 4;
    static constexpr int kProducts = kBlurred - 2;
};

template <int... P, typename F>
__device__ __forceinline__ void unroll(std::integer_sequence<int, P...>, F&& f) {
    (f(std::integral_constant<int, P>{}), ...);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
    return v;
}

// Strip geometry of a scale work item: 64 lanes of Cols output columns
// each (64- or 128-column strips), the source row wider by the blur radius
// on both sides; the 12 halo columns are loaded by the first 12 / Cols lanes.
template <int Cols>
struct StripGeom {
    static_assert(Cols == 1 || Cols == 2, "one or two output columns per lane");
    static constexpr int kCols = 64 * Cols;
    static constexpr int kSrc = kCols + 2 * kRadius;
    static constexpr int kHaloLanes = 2 * kRadius / Cols;
};

// A lane's Cols adjacent plane samples of one row: loaded as raw bits,
// unpacked to FP32 where they are consumed.
template <typename Fmt, int Cols> struct LaneSamples;
template <typename Fmt> struct LaneSamples<Fmt, 2> {
    using Raw = typename Fmt::Raw;
    static __device__ __forceinline__ Raw load(GlobalAS<const typename Fmt::E>* p) { return Fmt::load2(p); }
    static __device__ __forceinline__ void unpack(Raw v, float out[2]) {
        const float2 t = Fmt::unpack(v);
        out[0] = t.x;
        out[1] = t.y;
    }
};
template <typename Fmt> struct LaneSamples<Fmt, 1> {
    using Raw = typename Fmt::E;
    static __device__ __forceinline__ Raw load(GlobalAS<const typename Fmt::E>* p) { return *p; }
    static __device__ __forceinline__ void unpack(Raw v, float out[1]) { out[0] = Fmt::unpack1(v); }
};

// Cols adjacent floats to and from LDS (one 8-byte access for two).
template <int Cols>
__device__ __forceinline__ void lds_put(float* p, const float* v) {
    if constexpr (Cols == 2) {
        *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
    } else {
        *p = v[0];
    }
}
template <int Cols>
__device__ __forceinline__ void lds_get(const float* p, float* v) {
    if constexpr (Cols == 2) {
        const float2 t = *reinterpret_cast<const float2*>(p);
        v[0] = t.x;
        v[1] = t.y;
    } else {
        v[0] = *p;
    }
}

// The products of one lane's samples of a and b that the form blurs, staged
// at source column j of each product row.
template <bool Ref, int Cols>
__device__ __forceinline__ void stage_products(float (*prod)[StripGeom<Cols>::kSrc], int j, const float* a,
                                               const float* b) {
    float p0[Cols], p1[Cols], p2[Cols];
    for (int c = 0; c < Cols; ++c) {
        if constexpr (Ref) {
            p0[c] = a[c] * a[c];
            p1[c] = b[c] * b[c];
            p2[c] = a[c] * b[c];
        } else {
            const float d = a[c] - b[c];
            p0[c] = d * d;
            p1[c] = a[c] * b[c];
        }
    }
    lds_put<Cols>(&prod[0][j], p0);
    lds_put<Cols>(&prod[1][j], p1);
    if constexpr (Ref) lds_put<Cols>(&prod[2][j], p2);
}

// The six map values of one pixel from its blurred quantities `bl` (in the
// form's order) and its raw samples a and b: d, d^4, ring, ring^4, blur,
// blur^4 (scale.comp:305-333).
template <bool Ref>
__device__ __forceinline__ void map_values(const float* bl, float a, float b, float v6[6]) {
    const float mu1 = bl[0], mu2 = bl[1];
    const float det_dist = 1.0f + fabsf(b - mu2);
    const float det_orig = 1.0f + fabsf(a - mu1);
    float d, d1;
    if constexpr (Ref) {
        // scale.comp:305-333 in its order (the multiply-adds fused):
        // num_m = 1 - (mu1 - mu2)^2, num_s = 2 (s12 - mu1 mu2) + c2,
        // den = (s11 - mu1^2) + (s22 - mu2^2) + c2,
        // d = max((den - num_m num_s) / den, 0), d1 = (det_dist - det_orig) / det_orig,
        // with correctly rounded divisions.
        const float t = mu1 - mu2;
        const float num_m = fma_(-t, t, 1.0f);
        const float num_s = fma_(2.0f, fma_(-mu1, mu2, bl[4]), kC2);
        const float den = (fma_(-mu1, mu1, bl[2]) + fma_(-mu2, mu2, bl[3])) + kC2;
        d = fmaxf(div_rn(fma_(-num_m, num_s, den), den), 0.0f);
        d1 = div_rn(det_dist - det_orig, det_orig);
    } else {
        // scale.comp:305-333 rearranged: with e the blurred (a - b)^2 and
        // t = mu1 - mu2, s11 + s22 - mu1^2 - mu2^2 = num_s - c2 + v where
        // v = e - t^2, so den = num_s + v and den - num_m num_s =
        // v + t^2 num_s. The small v replaces the difference of two large
        // second moments (closer to FP64 than the direct form), and
        // identical images give v = t = 0, d = 0 exactly. One reciprocal
        // serves both quotients: q = 1 / (den det_orig), with
        // den >= c2 - rounding > 0 and det_orig >= 1.
        const float t = mu1 - mu2;
        const float num_s = fma_(2.0f, fma_(-mu1, mu2, bl[3]), kC2);
        const float v = fma_(-t, t, bl[2]);
        const float den = num_s + v;
        const float q = rcp_rn(den * det_orig);
        d = fmaxf((fma_(t * t, num_s, v) * det_orig) * q, 0.0f);
        d1 = ((det_dist - det_orig) * den) * q;
    }
    const float dd = d * d;
    const float ring = fmaxf(d1, 0.0f);
    const float blur = fmaxf(-d1, 0.0f);
    const float rr = ring * ring, bb = blur * blur;
    v6[0] = d;
    v6[1] = dd * dd;
    v6[2] = ring;
    v6[3] = rr * rr;
    v6[4] = blur;
    v6[5] = bb * bb;
}

// One scale work item: the strip of `level`, `channel` starting at (x0, y0),
// with the plane elements in format Fmt and Cols output columns per lane.
template <typename Fmt, typename TapsT, bool Ref, int Cols>
__device__ __forceinline__ void scale_item(ConstAS<PairDesc>* pdp, ConstAS<Plan>* plan, uint32_t item_idx,
                                           uint32_t level, uint32_t channel, uint32_t x0, uint32_t y0,
                                           float (*ring_a)[StripGeom<Cols>::kSrc],
                                           float (*ring_b)[StripGeom<Cols>::kSrc],
                                           float (*prod)[StripGeom<Cols>::kSrc]) {
    using Geo = StripGeom<Cols>;
    using IO = LaneSamples<Fmt, Cols>;
    constexpr int kBlurred = MapQuantities<Ref>::kBlurred;
    const uint32_t W = plan->w[level], H = plan->h[level], pitch = plan->pitch[level];
    // The strip's output rows stop at the level's last row (y0 < H for every
    // item): rows past it would be blurred and then dropped, up to half of
    // a strip on the small levels.
    const int total = static_cast<int>(min(plan->strip_h, H - y0)) + 2 * kRadius;

    using E = typename Fmt::E;
    // Source row r (0 .. total-1) is image row y0 - 6 + r, padded row y0 + r;
    // source column j (0 .. kSrc-1) is image column x0 - 6 + j, padded
    // x0 + 2 + j.
    GlobalAS<const uint8_t>* base = as_global(static_cast<const uint8_t*>(pdp->planes));
    GlobalAS<const E>* pa = reinterpret_cast<GlobalAS<const E>*>(base + plan->plane_off[level][channel]) +
                            (static_cast<uint64_t>(y0) * pitch + x0 + 2u);
    GlobalAS<const E>* pb = reinterpret_cast<GlobalAS<const E>*>(
        reinterpret_cast<GlobalAS<const uint8_t>*>(pa) + plan->image_bytes);

    const int lane = static_cast<int>(threadIdx.x);
    const bool halo = lane < Geo::kHaloLanes;
    const int jm = Cols * lane, jh = Geo::kCols + Cols * lane;

    typename IO::Raw cur_a = IO::load(pa + jm), cur_b = IO::load(pb + jm);
    typename IO::Raw cur_ha{}, cur_hb{};
    if (halo) {
        cur_ha = IO::load(pa + jh);
        cur_hb = IO::load(pb + jh);
    }

    float win[kTaps][kBlurred][Cols];
    // Map sums: each lane adds each of its columns' values over a block of
    // eight rows (rows 8k .. 8k+7 of the level, in row order) in FP32, then
    // the block's sum into FP64. Strips start on rows that are multiples of
    // eight, so the blocks, and every FP32 sum, do not depend on the strip
    // geometry.
    float block[Cols][6];
    for (int c = 0; c < Cols; ++c)
        for (int k = 0; k < 6; ++k) block[c][k] = 0.0f;
    double sums[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const uint32_t col = x0 + static_cast<uint32_t>(jm);

    for (int base_row = 0; base_row < total; base_row += kTaps) {
        unroll(std::make_integer_sequence<int, kTaps>{}, [&](auto pc) {
            constexpr int P = decltype(pc)::value;
            const int i = base_row + P;
            if (i >= total) return;
            // Stage input row i: raw samples into the ring, products alongside.
            float* ra = ring_a[i & (kRing - 1)];
            float* rb = ring_b[i & (kRing - 1)];
            {
                float a[Cols], b[Cols];
                IO::unpack(cur_a, a);
                IO::unpack(cur_b, b);
                lds_put<Cols>(ra + jm, a);
                lds_put<Cols>(rb + jm, b);
                stage_products<Ref, Cols>(prod, jm, a, b);
            }
            if (halo) {
                float a[Cols], b[Cols];
                IO::unpack(cur_ha, a);
                IO::unpack(cur_hb, b);
                lds_put<Cols>(ra + jh, a);
                lds_put<Cols>(rb + jh, b);
                stage_products<Ref, Cols>(prod, jh, a, b);
            }
            // Prefetch input row i + 1.
            if (i + 1 < total) {
                const uint64_t off = static_cast<uint64_t>(i + 1) * pitch;
                cur_a = IO::load(pa + off + jm);
                cur_b = IO::load(pb + off + jm);
                if (halo) {
                    cur_ha = IO::load(pa + off + jh);
                    cur_hb = IO::load(pb + off + jh);
                }
            }
            __syncthreads();

            // Horizontal 13-tap pass of each blurred quantity at this lane's
            // output columns (source columns jm + 6 .. jm + 5 + Cols).
            for (int q = 0; q < kBlurred; ++q) {
                const float* s = q == 0 ? ra : q == 1 ? rb : prod[q - 2];
                float v[kTaps - 1 + Cols];
                for (int k = 0; k < kTaps - 1 + Cols; k += Cols) lds_get<Cols>(s + jm + k, v + k);
                for (int c = 0; c < Cols; ++c) {
                    float h = TapsT::w[0] * v[c];
                    for (int k = 1; k < kTaps; ++k) h = fma_(TapsT::w[k], v[k + c], h);
                    win[P][q][c] = h;
                }
            }

            // Vertical pass for output row i - 12 (image row y0 + i - 12),
            // oldest ring entry first: scale.comp:243-298's ascending taps.
            if (i >= 2 * kRadius) {
                auto vertical = [&](int c, float* out) {
                    for (int q = 0; q < kBlurred; ++q) {
                        float acc = TapsT::w[0] * win[(P + 1) % kTaps][q][c];
                        for (int k = 1; k < kTaps; ++k)
                            acc = fma_(TapsT::w[k], win[(P + 1 + k) % kTaps][q][c], acc);
                        out[q] = acc;
                    }
                };
                const uint32_t y = y0 + static_cast<uint32_t>(i - 2 * kRadius);
                const int raw_slot = (i - kRadius) & (kRing - 1);
                float av[Cols], bv[Cols];
                lds_get<Cols>(&ring_a[raw_slot][jm + kRadius], av);
                lds_get<Cols>(&ring_b[raw_slot][jm + kRadius], bv);
                if (y < H) {
                    if constexpr (Ref) {
                        // One column at a time: its five blurred values
                        // live only through its map.
                        for (int c = 0; c < Cols; ++c) {
                            float bl[kBlurred];
                            vertical(c, bl);
                            if (col + static_cast<uint32_t>(c) >= W) continue;
                            float v6[6];
                            map_values<true>(bl, av[c], bv[c], v6);
                            for (int k = 0; k < 6; ++k) block[c][k] = block[c][k] + v6[k];
                        }
                    } else {
                        float bl[Cols][kBlurred];
                        for (int c = 0; c < Cols; ++c) vertical(c, bl[c]);
                        for (int c = 0; c < Cols; ++c) {
                            if (col + static_cast<uint32_t>(c) >= W) continue;
                            float v6[6];
                            map_values<false>(bl[c], av[c], bv[c], v6);
                            for (int k = 0; k < 6; ++k) block[c][k] = block[c][k] + v6[k];
                        }
                    }
                    // A block ends on rows 7, 15, 23, ... and on the level's
                    // last row (a uniform branch: y is the wave's row).
                    if ((y & 7u) == 7u || y + 1u == H) {
                        for (int c = 0; c < Cols; ++c)
                            for (int k = 0; k < 6; ++k) {
                                sums[k] += static_cast<double>(block[c][k]);
                                block[c][k] = 0.0f;
                            }
                    }
                }
            }
        });
    }

    GlobalAS<double>* out = as_global(pdp->partials) + static_cast<uint64_t>(item_idx) * 6u;
    for (int k = 0; k < 6; ++k) {
        const double s = wave_sum(sums[k]);
        if (lane == 0) out[k] = s;
    }
}

// Work items blockIdx.x, blockIdx.x + gridDim.x, ... of `blocks`, as in
// prepare_kernel. Two columns per lane hold a 130-register ring in the
// reference form (two waves per SIMD) and a 104-register one in the
// difference form, held to 168 registers (three waves); one column per lane
// halves the ring (four waves).
template <int Cols, bool Ref>
constexpr int kScaleWaves = (Cols == 2 && !Ref) ? 3 : 1;

template <typename Lay, typename TapsT, bool Ref, int Cols>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kScaleWaves<Cols, Ref>))) void scale_kernel(const PairDesc* __restrict__ pairs,
                                                   const WorkRef* __restrict__ work,
                                                   uint32_t blocks, uint32_t items_per_pair) {
    constexpr int kSrc = StripGeom<Cols>::kSrc;
    __shared__ __align__(16) float ring_a[kRing][kSrc];
    __shared__ __align__(16) float ring_b[kRing][kSrc];
    __shared__ __align__(16) float prod[MapQuantities<Ref>::kProducts][kSrc];

    for (uint32_t b = blockIdx.x; b < blocks; b += gridDim.x) {
        uint32_t pair_idx, item_idx;
        if (work != nullptr) {
            ConstAS<WorkRef>* w = as_const(work) + xcd_remap(b, blocks);
            pair_idx = w->pair;
            item_idx = w->index;
        } else {
            const uint32_t l = xcd_remap(b, blocks);
            pair_idx = l / items_per_pair;
            item_idx = l % items_per_pair;
        }
        ConstAS<PairDesc>* pdp = as_const(pairs) + pair_idx;
        ConstAS<Plan>* plan = as_const(pdp->plan);
        ConstAS<ScaleItem>* item = as_const(pdp->items) + item_idx;
        const uint32_t level = item->level, channel = item->channel;
        const uint32_t x0 = item->x0, y0 = item->y0;
        scale_item<typename Lay::Fmt, TapsT, Ref, Cols>(pdp, plan, item_idx, level, channel, x0, y0, ring_a,
                                                        ring_b, prod);
        __syncthreads();  // the next item reuses the row ring
    }
}

// ---- reduce -----------------------------------------------------------------

// One block per pair. Wave w takes the (level, channel) ranges w, w+16, ...
// (18 at most, so nearly every range has a wave of its own and a single
// pair's reduction is as long as its longest range, not the sum of four);
// its lanes stride over the range's strips, each adding the six partial sums
// of its strips in FP64 (four strips' loads in flight at a time, added in
// strip order), and the wave then sums its lanes.
// Norm level*18 + k: k < 6 is channel k/2's SSIM' slot k%2; k >= 6 is
// channel (k-6)/4's edge slot 2 + (k-6)%4 (ring 1-, 4-norm, blur 1-, 4-norm).
constexpr uint32_t kReduceWaves = 16;

__global__ __launch_bounds__(64 * kReduceWaves) void reduce_kernel(const PairDesc* __restrict__ pairs) {
    ConstAS<PairDesc>* pdp = as_const(pairs) + blockIdx.x;
    ConstAS<Plan>* plan = as_const(pdp->plan);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    GlobalAS<const double>* partials = as_global(static_cast<const double*>(pdp->partials));
    GlobalAS<double>* norms = as_global(pdp->norms);
    for (uint32_t combo = wave; combo < static_cast<uint32_t>(kMaxScales * kChannels); combo += kReduceWaves) {
        const uint32_t level = combo / kChannels, c = combo % kChannels;
        const bool active = level < plan->levels && (plan->mask[level] & (1u << c)) != 0u;
        double sum[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (active) {
            const uint32_t begin = plan->item_begin[level][c], count = plan->item_count[level][c];
#pragma unroll 4
            for (uint32_t it = lane; it < count; it += 64u) {
                GlobalAS<const double>* p = partials + static_cast<uint64_t>(begin + it) * 6u;
#pragma unroll
                for (int k = 0; k < 6; ++k) sum[k] += p[k];
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) sum[k] = wave_sum(sum[k]);
        }
        if (lane != 0) continue;
        const double one_per_pixel =
            active ? 1.0 / (static_cast<double>(plan->w[level]) * static_cast<double>(plan->h[level])) : 0.0;
#pragma unroll
        for (uint32_t slot = 0; slot < 6; ++slot) {
            double value = one_per_pixel * sum[slot];
            if (slot & 1u) value = sqrt(sqrt(value));
            const uint32_t k = slot < 2u ? 2u * c + slot : 6u + 4u * c + (slot - 2u);
            norms[level * 18u + k] = value;
        }
    }
}

// A plain streaming copy (16-byte loads and stores): the reference point that
// scales the FETCH_SIZE / WRITE_SIZE counters to bytes.
__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src,
                                                   uint4* __restrict__ dst, uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

// Self-check of the reference form's cube root over the floats with bits in
// [first, end): cbrt_rn_hw (what the kernels take) against cbrt_rn (IEEE
// steps only). Counts the inputs where they differ and those that took the
// exact step, and keeps the smallest differing input and the largest
// |1 - x r^3| of the hardware seed r (as FP64 bits: non-negative doubles
// order as their bits). One atomic per wave and counter.
__global__ __launch_bounds__(256) void cbrt_check_kernel(uint32_t first, uint32_t end, CbrtCheck* out) {
    unsigned long long bad = 0, near = 0, h_bits = 0;
    uint32_t first_bad = 0xFFFFFFFFu;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t u = first + static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; u < end; u += stride) {
        const float x = __uint_as_float(static_cast<uint32_t>(u));
        bool n;
        const float rounded = cbrt_rn_near_hw(x, &n);
        const float got = n ? cbrt_round_exact(x, rounded) : rounded;
        near += n ? 1u : 0u;
        if (got != cbrt_rn(x)) {
            ++bad;
            first_bad = min(first_bad, static_cast<uint32_t>(u));
        }
        const double r = cbrt_seed_hw(x);
        const double h = fabs(fma_(-static_cast<double>(x), (r * r) * r, 1.0));
        h_bits = max(h_bits, static_cast<unsigned long long>(__double_as_longlong(h)));
    }
    for (int offset = 32; offset > 0; offset >>= 1) {
        bad += __shfl_xor(bad, offset, 64);
        near += __shfl_xor(near, offset, 64);
        h_bits = max(h_bits, __shfl_xor(h_bits, offset, 64));
        first_bad = min(first_bad, __shfl_xor(first_bad, offset, 64));
    }
    if (threadIdx.x % 64u == 0u) {
        atomicAdd(&out->mismatches, bad);
        atomicAdd(&out->near, near);
        atomicMax(&out->max_h_bits, h_bits);
        atomicMin(&out->first_mismatch, first_bad);
    }
}

// Self-check of div_rn over the significand pairs a = 1.ma, b = 1.mb with
// mb in [mb_first, mb_first + rows) and every ma, against the IEEE quotient;
// of rcp_rn(b) against 1 / b; and of the hardware reciprocal's exponent
// invariance for those b over 2^e, e in [-24, 24). One atomic per wave and
// counter.
__global__ __launch_bounds__(256) void div_check_kernel(uint32_t mb_first, uint32_t rows, DivCheck* out) {
    unsigned long long bad = 0, rcp_bad = 0;
    uint32_t first_bad = 0xFFFFFFFFu;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t lane0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t k = 0; k < rows; ++k) {
        const uint32_t mb = mb_first + k;
        const float b = __uint_as_float(0x3F800000u | mb);
        for (uint32_t ma = lane0; ma < (1u << 23); ma += stride) {
            const float a = __uint_as_float(0x3F800000u | ma);
            if (__float_as_uint(div_rn(a, b)) != __float_as_uint(a / b)) {
                ++bad;
                first_bad = min(first_bad, mb);
            }
        }
        if (lane0 < 48u) {
            const int e = static_cast<int>(lane0) - 24;
            const float want = __builtin_ldexpf(__builtin_amdgcn_rcpf(b), -e);
            if (__float_as_uint(__builtin_amdgcn_rcpf(__builtin_ldexpf(b, e))) != __float_as_uint(want)) ++rcp_bad;
        } else if (lane0 == 48u) {
            // The refined reciprocal is the correctly rounded 1 / b.
            if (__float_as_uint(rcp_rn(b)) != __float_as_uint(1.0f / b)) ++rcp_bad;
        }
    }
    for (int offset = 32; offset > 0; offset >>= 1) {
        bad += __shfl_xor(bad, offset, 64);
        rcp_bad += __shfl_xor(rcp_bad, offset, 64);
        first_bad = min(first_bad, __shfl_xor(first_bad, offset, 64));
    }
    if (threadIdx.x % 64u == 0u) {
        atomicAdd(&out->mismatches, bad);
        atomicAdd(&out->rcp_mismatches, rcp_bad);
        atomicMin(&out->first_mismatch, first_bad);
    }
}

// Blocks of `kernel` the device holds resident at once: occupancy per
// compute unit x compute units, rounded down to a multiple of 8 (at least
// 8). A persistent launch uses min(work items, this) blocks, each looping
// over its share; 0 when the runtime cannot say.
uint32_t ResidentGrid(const void* kernel, int block_threads) {
    int per_cu = 0, device = 0;
    hipDeviceProp_t prop{};
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block_threads, 0) != hipSuccess ||
        hipGetDevice(&device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
        per_cu <= 0 || prop.multiProcessorCount <= 0)
        return 0;
    const uint64_t resident = static_cast<uint64_t>(per_cu) * static_cast<uint64_t>(prop.multiProcessorCount);
    return static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(8, resident / 8 * 8), 0x7FFFFFF8u));
}

template <typename K>
uint32_t GridOf(K kernel, int block_threads, uint32_t blocks, bool persistent) {
    if (!persistent) return blocks;
    static const uint32_t resident = ResidentGrid(reinterpret_cast<const void*>(kernel), block_threads);
    return resident == 0 ? blocks : std::min(blocks, resident);
}

template <typename Lay, bool General, bool Ref>
hipError_t launch_prepare_g(const PairDesc* pairs, const WorkRef* work, uint32_t blocks, uint32_t tiles,
                            const float* lut, const uint8_t* src0, const uint8_t* src1, bool persistent,
                            hipStream_t stream) {
    const uint32_t grid = GridOf(&prepare_kernel<Lay, General, Ref>, 256, blocks, persistent);
    hipLaunchKernelGGL((prepare_kernel<Lay, General, Ref>), dim3(grid), dim3(256), 0, stream, pairs, work,
                       blocks, tiles, lut, src0, src1);
    return hipGetLastError();
}

template <typename Lay, bool Ref>
hipError_t launch_prepare_m(bool general, const PairDesc* pairs, const WorkRef* work, uint32_t blocks,
                            uint32_t tiles, const float* lut, const uint8_t* src0, const uint8_t* src1,
                            bool persistent, hipStream_t stream) {
    return general ? launch_prepare_g<Lay, true, Ref>(pairs, work, blocks, tiles, lut, src0, src1, persistent, stream)
                   : launch_prepare_g<Lay, false, Ref>(pairs, work, blocks, tiles, lut, src0, src1, persistent, stream);
}

template <typename Lay>
hipError_t launch_prepare_t(MapForm map, bool general, const PairDesc* pairs, const WorkRef* work,
                            uint32_t blocks, uint32_t tiles, const float* lut, const uint8_t* src0,
                            const uint8_t* src1, bool persistent, hipStream_t stream) {
    return map == kMapReference
               ? launch_prepare_m<Lay, true>(general, pairs, work, blocks, tiles, lut, src0, src1, persistent, stream)
               : launch_prepare_m<Lay, false>(general, pairs, work, blocks, tiles, lut, src0, src1, persistent, stream);
}

template <typename Lay, typename TapsT, bool Ref, int Cols>
hipError_t launch_scale_g(const PairDesc* pairs, const WorkRef* work, uint32_t blocks, uint32_t items,
                          bool persistent, hipStream_t stream) {
    const uint32_t grid = GridOf(&scale_kernel<Lay, TapsT, Ref, Cols>, 64, blocks, persistent);
    hipLaunchKernelGGL((scale_kernel<Lay, TapsT, Ref, Cols>), dim3(grid), dim3(64), 0, stream, pairs, work,
                       blocks, items);
    return hipGetLastError();
}

template <typename Lay, bool Ref, int Cols>
hipError_t launch_scale_m(uint32_t blur, const PairDesc* pairs, const WorkRef* work, uint32_t blocks,
                          uint32_t items, bool persistent, hipStream_t stream) {
    return blur == kBlurRecursive
               ? launch_scale_g<Lay, CosineTaps, Ref, Cols>(pairs, work, blocks, items, persistent, stream)
               : launch_scale_g<Lay, SampledTaps, Ref, Cols>(pairs, work, blocks, items, persistent, stream);
}

template <typename Lay, int Cols>
hipError_t launch_scale_c(uint32_t blur, MapForm map, const PairDesc* pairs, const WorkRef* work,
                          uint32_t blocks, uint32_t items, bool persistent, hipStream_t stream) {
    return map == kMapReference
               ? launch_scale_m<Lay, true, Cols>(blur, pairs, work, blocks, items, persistent, stream)
               : launch_scale_m<Lay, false, Cols>(blur, pairs, work, blocks, items, persistent, stream);
}

template <typename Lay>
hipError_t launch_scale_t(uint32_t strip_cols, uint32_t blur, MapForm map, const PairDesc* pairs,
                          const WorkRef* work, uint32_t blocks, uint32_t items, bool persistent,
                          hipStream_t stream) {
    return strip_cols == 64 ? launch_scale_c<Lay, 1>(blur, map, pairs, work, blocks, items, persistent, stream)
                            : launch_scale_c<Lay, 2>(blur, map, pairs, work, blocks, items, persistent, stream);
}

}  // namespace

hipError_t LaunchCopy(void* dst, const void* src, uint64_t bytes, hipStream_t stream) {
    const uint64_t n = bytes / 16;
    if (n == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(copy_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, stream,
                       static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
    return hipGetLastError();
}

hipError_t LaunchDivCheck(uint32_t mb_first, uint32_t rows, DivCheck* out, hipStream_t stream) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(div_check_kernel, dim3(2048), dim3(256), 0, stream, mb_first, rows, out);
    return hipGetLastError();
}

hipError_t LaunchCbrtCheck(uint32_t first, uint32_t end, CbrtCheck* out, hipStream_t stream) {
    if (end <= first) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((static_cast<uint64_t>(end - first) + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(cbrt_check_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, stream, first, end,
                       out);
    return hipGetLastError();
}

hipError_t LaunchPrepare(PlaneFormat format, MapForm map, bool general, const PairDesc* pairs,
                         const WorkRef* work, uint32_t blocks, uint32_t tiles_per_image, const float* lut,
                         const uint8_t* src0, const uint8_t* src1, hipStream_t stream, bool persistent) {
    if (blocks == 0) return hipSuccess;
    if (map > kMapDifference) return hipErrorInvalidValue;
    const uint32_t t = tiles_per_image;
    switch (format) {
        case kPlaneF32:
            return launch_prepare_t<LayoutF32>(map, general, pairs, work, blocks, t, lut, src0, src1, persistent,
                                               stream);
        case kPlaneF16:
            return launch_prepare_t<LayoutF16>(map, general, pairs, work, blocks, t, lut, src0, src1, persistent,
                                               stream);
        case kPlaneBF16:
            return launch_prepare_t<LayoutBF16>(map, general, pairs, work, blocks, t, lut, src0, src1, persistent,
                                                stream);
        case kPlaneU16:
            return launch_prepare_t<LayoutU16>(map, general, pairs, work, blocks, t, lut, src0, src1, persistent,
                                               stream);
    }
    return hipErrorInvalidValue;
}

hipError_t LaunchScale(PlaneFormat format, uint32_t blur, MapForm map, uint32_t strip_cols,
                       const PairDesc* pairs, const WorkRef* work, uint32_t blocks, uint32_t items_per_pair,
                       hipStream_t stream, bool persistent) {
    if (blocks == 0) return hipSuccess;
    if (blur > kBlurRecursive || map > kMapDifference || !ValidStripCols(strip_cols)) return hipErrorInvalidValue;
    const uint32_t n = items_per_pair, sc = strip_cols;
    switch (format) {
        case kPlaneF32: return launch_scale_t<LayoutF32>(sc, blur, map, pairs, work, blocks, n, persistent, stream);
        case kPlaneF16: return launch_scale_t<LayoutF16>(sc, blur, map, pairs, work, blocks, n, persistent, stream);
        case kPlaneBF16:
            return launch_scale_t<LayoutBF16>(sc, blur, map, pairs, work, blocks, n, persistent, stream);
        case kPlaneU16: return launch_scale_t<LayoutU16>(sc, blur, map, pairs, work, blocks, n, persistent, stream);
    }
    return hipErrorInvalidValue;
}

hipError_t LaunchReduce(const PairDesc* pairs, uint32_t pairs_count, hipStream_t stream) {
    if (pairs_count == 0) return hipSuccess;
    hipLaunchKernelGGL(reduce_kernel, dim3(pairs_count), dim3(64 * kReduceWaves), 0, stream, pairs);
    return hipGetLastError();
}

}  // namespace vqm::ss2
