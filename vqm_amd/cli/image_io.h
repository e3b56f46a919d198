// On-disk inputs of the scoring command: PNG, JPEG and YUV4MPEG2 decoded to the
// packed-plane Image of include/ssimulacra2.h, with the colour declarations
// the reference's loader gives them (tools/image_io/io.h:17-60, io.cc).
//
// * PNG (8/16-bit, every colour type, Adam7): palettes and sub-byte grey are
//   expanded to 8 bits, tRNS becomes straight alpha, 16-bit samples become
//   host-order Word samples; sRGB transfer, BT.709 primaries.
// * YUV4MPEG2 (progressive): the C tags mono, 420, 420jpeg, 420mpeg2,
//   420paldv, 422, 444, 411 and 444alpha, with p9/p10/p12/p14/p16 depth
//   suffixes (little-endian words); limited range, BT.470BG matrix, BT.709
//   transfer and primaries, all overridable by the --input-* options.
// * JPEG (baseline, extended and progressive Huffman; 1 or 3 components):
//   decoded to the samples libjpeg's default set-up yields (islow IDCT,
//   fancy upsampling, JFIF YCbCr to RGB), as 8-bit RGB or grey with sRGB
//   transfer and BT.709 primaries (io.cc:277-355).
#pragma once

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <memory>
#include <optional>
#include <string>
#include <string_view>

#include "ssimulacra2.h"

namespace vqm::image_io {

// The first frame of a PNG, JPEG or Y4M file; nullopt with `error` set on failure.
std::optional<Image> LoadImage(std::string_view path, std::string& error);

// PNG from memory (the decoder LoadImage uses for PNG files).
std::optional<Image> DecodePng(const std::uint8_t* data, std::size_t size, std::string& error);

// JPEG from memory (the decoder LoadImage uses for JPEG files).
std::optional<Image> DecodeJpeg(const std::uint8_t* data, std::size_t size, std::string& error);

// Rational frame rate of a Y4M stream (F tag); 0/0 when absent or invalid.
struct FrameRate {
    std::uint32_t numerator = 0, denominator = 0;
};

// Frame-by-frame reader of a progressive YUV4MPEG2 file
// (io.h:37-60 SequenceReader, Y4M only).
class Y4mReader {
public:
    // Throws std::runtime_error naming the file and the reason. A sequence
    // reader also requires a valid frame rate and progressive frames
    // (io.cc:629-655); the still-image loader reads the first frame of any.
    explicit Y4mReader(const std::string& path, bool sequence = true);
    ~Y4mReader();
    Y4mReader(const Y4mReader&) = delete;
    Y4mReader& operator=(const Y4mReader&) = delete;

    // The next frame into `image`; false at the end of the stream. Throws
    // std::runtime_error on a malformed or truncated frame.
    bool Next(Image& image);
    [[nodiscard]] Size size() const { return size_; }
    [[nodiscard]] const ImageFormat& format() const { return format_; }
    [[nodiscard]] FrameRate rate() const { return rate_; }

private:
    std::ifstream in_;
    std::string path_;
    Size size_;
    ImageFormat format_;
    FrameRate rate_;
};

// True when the file starts with the YUV4MPEG2 signature.
bool IsY4m(const std::string& path);

// --input-matrix / --input-transfer / --input-primaries / --input-range /
// --chroma-filter values (io.cc:690-780): explicit overrides replace what
// the decoder declared, on both images.
struct InputColorOptions {
    std::optional<MatrixCoefficients> matrix;
    std::optional<TransferCharacteristics> transfer;
    std::optional<ColorPrimaries> primaries;
    std::optional<PixelRange> range;
    ChromaFilter chroma_filter = ChromaFilter::Bilinear;
    void ApplyTo(ImageFormat& format) const;
};

// Parses one option value; false with `accepted` listing the valid names.
bool ParseMatrix(std::string_view text, MatrixCoefficients& out, std::string& accepted);
bool ParseTransfer(std::string_view text, TransferCharacteristics& out, std::string& accepted);
bool ParsePrimaries(std::string_view text, ColorPrimaries& out, std::string& accepted);
bool ParseRange(std::string_view text, PixelRange& out, std::string& accepted);
bool ParseChromaFilter(std::string_view text, ChromaFilter& out, std::string& accepted);

}  // namespace vqm::image_io
