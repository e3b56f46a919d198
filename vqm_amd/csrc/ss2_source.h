// The general input surface on the host: validation of every image format
// the reference accepts (input.cc:114-255, 576-714), and the per-image decode
// descriptor the prepare kernel reads (vulkan_source.cc:41-173). No GPU
// calls here.
#pragma once

#include <array>
#include <cstdint>
#include <string>

#include "ss2_common.h"
#include "vqm_ss2.h"

namespace vqm::ss2 {

// zimg's numeric enum values (include/vqm_image.h:24-96).
namespace fmt {
enum Type : uint8_t { kByte = 0, kWord = 1, kHalf = 2, kFloat = 3 };
enum Family : uint8_t { kGrey = 0, kRgb = 1, kYuv = 2 };
enum Alpha : uint8_t { kNoAlpha = 0, kStraight = 1, kPremultiplied = 2 };
enum Range : uint8_t { kLimited = 0, kFull = 1 };
enum Matrix : uint8_t {
    kMatRgb = 0, kMatBt709 = 1, kMatUnspecified = 2, kMatFcc = 4, kMatBt470Bg = 5,
    kMatSt170M = 6, kMatSt240M = 7, kMatYCgCo = 8, kMatBt2020Ncl = 9, kMatBt2020Cl = 10,
    kMatSt2085YDzDx = 11, kMatChromaNcl = 12, kMatChromaCl = 13, kMatIctCp = 14,
    kMatSt2128IptC2 = 15, kMatYCgCoRe = 16, kMatYCgCoRo = 17
};
enum Transfer : uint8_t {
    kTrBt709 = 1, kTrUnspecified = 2, kTrBt470M = 4, kTrBt470Bg = 5, kTrBt601 = 6,
    kTrSt240M = 7, kTrLinear = 8, kTrLog100 = 9, kTrLog316 = 10, kTrIec61966_2_4 = 11,
    kTrBt1361 = 12, kTrSrgb = 13, kTrBt2020_10 = 14, kTrBt2020_12 = 15, kTrSt2084 = 16,
    kTrSt428 = 17, kTrAribB67 = 18
};
enum Primaries : uint8_t {
    kPrBt709 = 1, kPrUnspecified = 2, kPrBt470M = 4, kPrBt470Bg = 5, kPrSt170M = 6,
    kPrSt240M = 7, kPrFilm = 8, kPrBt2020 = 9, kPrSt428 = 10, kPrSt431_2 = 11,
    kPrSt432_1 = 12, kPrEbu3213E = 22
};
enum Location : uint8_t { kLeft = 0, kCenter = 1, kTopLeft = 2, kTop = 3, kBottomLeft = 4, kBottom = 5 };
}  // namespace fmt

using Matrix3 = std::array<double, 9>;  // row-major

// The reference's ValidateImage (input.cc:152-222): enum membership, depth,
// subsampling, family/matrix/transfer/primaries combinations. Throws
// Error{VQM_INVALID_ARGUMENT} naming `name`.
void ValidateFormat(const vqm_image_format& format, uint32_t width, uint32_t height,
                    const std::string& name);

// 8-bit full-range opaque sRGB RGB with BT.709 primaries: the table-decoded
// fast path. Everything else goes through a SourceDesc.
bool IsSrgb8(const vqm_image_format& format);

// Byte offset and geometry of the four planes (input.cc:114-145).
struct PlaneGeometry {
    bool present = false;
    uint64_t offset = 0;
    uint32_t width = 0, height = 0;
};
std::array<PlaneGeometry, 4> PlaneLayout(const vqm_image_format& format, uint32_t width, uint32_t height);

// The per-sample scale and offset of a plane class (input.cc:714-757).
struct ScaleOffset {
    float scale = 1.0f, offset = 0.0f;
};
ScaleOffset PlaneScaleOffset(const vqm_image_format& format, bool chroma, bool alpha);

// The decode descriptor of a validated format (vulkan_source.cc:75-173).
SourceDesc BuildSourceDesc(const vqm_image_format& format, uint32_t width, uint32_t height,
                           uint32_t chroma_filter);

// Colorimetry in double, as the descriptor derives it (zimg's algebra,
// input.cc:311-842); exposed for tests.
Matrix3 NclToRgb(const vqm_image_format& format, bool* identity);
Matrix3 GamutToBt709(uint8_t primaries);
Matrix3 Inverse3(const Matrix3& m);
Matrix3 Multiply3(const Matrix3& a, const Matrix3& b);

}  // namespace vqm::ss2
