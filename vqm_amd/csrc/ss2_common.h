// Shared host/device definitions of the SSIMULACRA 2.1 HIP path: the
// per-geometry plan, the per-pair descriptor the kernels read, and the
// constants of the metric.
//
// HBM layout of one pair's workspace (one "slot"):
//   planes   2 images x (sum over stored (level, channel) planes of
//            pitch_l * rows_l elements), every element FP32, FP16, BF16 or
//            unsigned 16-bit fixed point (q / 65536).
//            Every plane carries a zero border: element (x, y) of a level
//            lives at (y + kPadTop) * pitch + (x + kPadLeft), and the border
//            (never written after the workspace is cleared) is the zero
//            padding of the 13-tap blur (scale.comp:107-124), so the scale
//            kernel loads whole rows without bounds checks.
//   partials items x 6 doubles: one FP64 partial per scale work item.
//   norms    108 doubles: the finished norms of the pair.
#pragma once

#include <cstdint>

namespace vqm::ss2 {

inline constexpr int kMaxScales = 6;        // ssimulacra2_internal.h:185
inline constexpr int kChannels = 3;         // X, Y, B
inline constexpr int kNorms = 108;          // 18 per level
inline constexpr int kRadius = 6;           // ceil(3.5 * 1.5), pipeline.cc:276
inline constexpr int kTaps = 13;

// Scale kernel geometry: one 64-lane wave per work item, one or two output
// columns per lane (strips of 64 or 128 columns, Plan::strip_w). Planes are
// padded for the widest strip.
inline constexpr int kStripCols = 128;
inline constexpr uint32_t kDefaultStripCols = 128;
constexpr bool ValidStripCols(uint32_t cols) { return cols == 64 || cols == 128; }
// Strip heights are multiples of the map-sum block height (8 rows).
inline constexpr uint32_t kSumBlockRows = 8;
// Plane padding: 6 zero rows above, 8 zero columns on the left (so rows of
// four samples stay 16-byte aligned), and enough on the right and below to
// cover the last strip's halo.
inline constexpr int kPadTop = kRadius;
inline constexpr int kPadLeft = 8;

// Prepare kernel: a 256-thread block converts a 64 x 32 base-pixel tile and
// every level of its sub-pyramid (64 x 32 halves down to 2 x 1 at level 5).
inline constexpr int kPrepTileW = 64;
inline constexpr int kPrepTileH = 32;

enum PlaneFormat : uint32_t { kPlaneF32 = 0, kPlaneF16 = 1, kPlaneBF16 = 2, kPlaneU16 = 3 };

// Bytes of one stored plane element in `format`.
constexpr uint32_t ElementBytes(PlaneFormat format) { return format == kPlaneF32 ? 4u : 2u; }

struct ScaleItem {
    uint32_t level;
    uint32_t channel;
    uint32_t x0;  // first output column of the strip
    uint32_t y0;  // first output row of the strip
};

// Everything about one geometry the kernels need. Built on the host
// (ss2_plan.cc), copied to the device once per geometry.
struct Plan {
    uint32_t width;   // base geometry
    uint32_t height;
    uint32_t levels;  // 1..6
    uint32_t strip_h; // output rows per scale work item
    uint32_t strip_w; // output columns per scale work item: 64 or 128
    uint32_t w[kMaxScales], h[kMaxScales];
    uint32_t pitch[kMaxScales], rows[kMaxScales];  // padded plane geometry
    uint32_t mask[kMaxScales];                      // active channels, bit c
    uint64_t plane_off[kMaxScales][kChannels];      // bytes, ~0 = not stored
    uint64_t image_bytes;                           // plane bytes per image
    uint32_t item_begin[kMaxScales][kChannels];
    uint32_t item_count[kMaxScales][kChannels];
    uint32_t items;                                 // scale work items per pair
    uint32_t prep_tiles_x, prep_tiles_y;
    uint32_t plane_stride;                          // bytes per u8 input plane
};

// Decode curve ids of SourceDesc::transfer: linear; the BT.1886 2.4 power
// (BT.709, BT.601, BT.2020, SMPTE 240M); 2.2 and 2.8 powers (BT.470M/BG);
// IEC 61966-2-1 sRGB; the two log curves; xvYCC; BT.1361; SMPTE ST 2084 (PQ);
// SMPTE ST 428; ARIB STD-B67 (HLG).
// The blur of the maps: the reference's 13-tap sampled Gaussian
// (pipeline.cc:274-291), or the kernel Cloudinary's recursive Gaussian
// computes (ssimulacra2.cc:82-106).
enum BlurKind : uint32_t { kBlurSampled = 0, kBlurRecursive = 1 };

// The arithmetic form of the XYB transform and the SSIM' map.
//   kMapReference   prepare.comp:81-98 and scale.comp:130-142,305-333 in
//                   their own operation order: correctly rounded cube roots,
//                   five blurred quantities (a, b, a*a, b*b, a*b), the map's
//                   variances as s11 - mu1^2 and s22 - mu2^2, IEEE divisions.
//   kMapDifference  the same quantities rearranged: four blurred quantities
//                   (a, b, (a - b)^2, a*b), the variance sum as one small
//                   blurred difference, one reciprocal for both quotients, a
//                   fast cube root. Closer to FP64, a fifth less blur work.
enum MapForm : uint32_t { kMapReference = 0, kMapDifference = 1 };

// The 13 taps of MakeGaussian(1.5) (pipeline.cc:274-291), exact FP32 values.
struct SampledTaps {
    static constexpr float w[kTaps] = {
        0x1.763864p-14f, 0x1.0d891ap-10f, 0x1.f1e742p-8f, 0x1.26dd9ep-5f,
        0x1.bfdc88p-4f,  0x1.b42852p-3f,  0x1.1058fp-2f,  0x1.b42852p-3f,
        0x1.bfdc88p-4f,  0x1.26dd9ep-5f,  0x1.f1e742p-8f, 0x1.0d891ap-10f,
        0x1.763864p-14f};
};

// The recursive Gaussian of Cloudinary's ssimulacra2 (libjxl
// CreateRecursiveGaussian(1.5), ssimulacra2.cc:82-106) as the zero-padded FIR
// it computes: K[m] = sum_{k=1,3,5} beta_k cos(k pi m / 10), |m| <= 4, in the
// 13-tap frame (zeros at |m| = 5, 6 leave the FMA chain bit-identical to a
// 9-tap one). FP32 roundings of the FP64 kernel.
struct CosineTaps {
    static constexpr float w[kTaps] = {
        0.0f, 0.0f, 0x1.347d7p-7f, 0x1.2700c8p-5f, 0x1.bfd672p-4f, 0x1.b413e8p-3f, 0x1.0ef8d6p-2f,
        0x1.b413e8p-3f, 0x1.bfd672p-4f, 0x1.2700c8p-5f, 0x1.347d7p-7f, 0.0f, 0.0f};
};

enum TransferId : uint32_t {
    kCurveLinear = 0, kCurvePow24 = 1, kCurvePow22 = 2, kCurvePow28 = 3, kCurveSrgb = 4,
    kCurveLog100 = 5, kCurveLog316 = 6, kCurveXvycc = 7, kCurveBt1361 = 8, kCurvePq = 9,
    kCurveSt428 = 10, kCurveHlg = 11
};

// How the prepare kernel decodes one input image that is not 8-bit sRGB RGB
// (the general input surface of color.glsl:94-541 and input.cc:114-842):
// raw sample -> range/depth scale and offset -> chroma upsampling -> matrix
// (or constant-luminance / ICtCp op) -> alpha -> transfer -> primaries,
// ending in linear BT.709 RGB.
struct SourceDesc {
    uint32_t kind;            // 0: 8-bit sRGB RGB (table decode), 1: this descriptor
    float background;         // alpha blend target, in encoded RGB
    uint64_t plane_off[4];    // bytes from the image base: Y/R, U/G, V/B, alpha
    uint32_t plane_w[4], plane_h[4];
    uint32_t type;            // 0 byte, 1 word, 2 half, 3 float
    uint32_t family;          // 0 grey, 1 RGB, 2 YUV
    uint32_t alpha;           // 0 none, 1 straight, 2 premultiplied
    uint32_t sub_w, sub_h;    // log2 chroma subsampling
    uint32_t transfer;        // decode curve id, see TransferId
    uint32_t mode;            // 0 matrix, 1 constant luminance, 2 ICtCp
    uint32_t filter;          // chroma filter: 0 bilinear, 1 bicubic, 2 lanczos
    uint32_t identity1;       // matrix 1 is the identity
    uint32_t identity2;       // matrix 2 is the identity
    float scale[3], offset[3];  // luma/RGB, chroma, alpha sample classes
    float m1[9];              // YUV -> RGB, or ICtCp -> LMS
    float m2[9];              // gamut to BT.709 (after LMS -> RGB for ICtCp)
    float kr, kg, kb;         // luma weights of the CL op and the HLG OOTF
    float to_linear_scale;    // PQ and HLG peak over 100 cd/m^2
    float shift_x, shift_y;   // chroma siting offsets
    float cl[4];              // CL op constants nb, pb, nr, pr
};

// Scale work item i of a plan: the (level, channel) ranges follow one
// another in level-major order, and within a range the items run row-major
// over strip_h-row bands of strip_w-column strips. constexpr, so host code
// and device code share it.
constexpr ScaleItem ItemAt(const Plan& p, uint32_t i) {
    for (uint32_t l = 0; l < p.levels; ++l) {
        for (uint32_t c = 0; c < static_cast<uint32_t>(kChannels); ++c) {
            const uint32_t begin = p.item_begin[l][c], count = p.item_count[l][c];
            if (i >= begin && i - begin < count) {
                const uint32_t k = i - begin;
                const uint32_t strips = (p.w[l] + p.strip_w - 1) / p.strip_w;
                return ScaleItem{l, c, (k % strips) * p.strip_w, (k / strips) * p.strip_h};
            }
        }
    }
    return ScaleItem{0, 0, 0, 0};
}

// One pair as the kernels see it.
struct PairDesc {
    const uint8_t* src[2];     // packed planar images, original then distorted
    void* planes;              // 2 * image_bytes of planes
    double* partials;          // items * 6
    double* norms;             // 108
    const Plan* plan;          // device copy of the pair's plan
    const ScaleItem* items;    // device copy of the plan's item table
    const SourceDesc* source[2];  // null or kind 0: 8-bit sRGB RGB (table decode)
};

}  // namespace vqm::ss2
