// PNG and YUV4MPEG2 loading for the scoring command (JPEG is in jpeg.cc). PNG follows the W3C
// PNG specification (chunks, zlib stream, the five row filters, Adam7) with
// zlib's inflate; the output matches what the reference's libpng set-up
// (tools/image_io/io.cc:156-238: expand, palette to RGB, tRNS to alpha,
// 16-bit swapped to host order) delivers. Y4M follows the yuv4mpeg2 header
// grammar the reference reads (io.cc:360-607).
#include "image_io.h"

#include <zlib.h>

#include <algorithm>
#include <array>
#include <charconv>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace vqm::image_io {
namespace {

constexpr std::uint8_t kPngSignature[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
constexpr char kY4mSignature[] = "YUV4MPEG2";
constexpr std::size_t kMaxHeaderLine = 1024;
// A PNG's decoded image is held in memory: refuse dimensions whose filtered
// rows alone would exceed this.
constexpr std::uint64_t kMaxPngBytes = 1ull << 34;

std::uint32_t ReadBe32(const std::uint8_t* p) {
    return (std::uint32_t{p[0]} << 24) | (std::uint32_t{p[1]} << 16) | (std::uint32_t{p[2]} << 8) | p[3];
}

std::size_t Align4(std::size_t n) { return (n + 3) / 4 * 4; }

// Byte offsets of the present planes of a packed image: planes in the order
// Y/R, U/G, V/B, alpha, each starting on a 4-byte boundary.
std::array<std::size_t, 4> PlaneOffsets(const ImageFormat& f, Size size, std::size_t sample_bytes,
                                        std::array<Size, 4>& dims) {
    std::array<std::size_t, 4> off{};
    std::size_t at = 0;
    for (int p = 0; p < 4; ++p) {
        const bool present = p == 0 || (p < 3 && f.family != ColorFamily::Grey) ||
                             (p == 3 && f.alpha != AlphaType::None);
        if (!present) {
            dims[p] = Size{0, 0};
            continue;
        }
        Size d = size;
        if ((p == 1 || p == 2) && f.family == ColorFamily::Yuv) {
            d.width >>= f.subsample_w;
            d.height >>= f.subsample_h;
        }
        dims[p] = d;
        off[p] = at;
        at += Align4(d.width * d.height * sample_bytes);
    }
    return off;
}

// ---- PNG -------------------------------------------------------------------------

struct PngHeader {
    std::uint32_t width = 0, height = 0;
    std::uint8_t depth = 0, color = 0, interlace = 0;
};

int PngChannels(std::uint8_t color) {
    switch (color) {
        case 0: return 1;  // grey
        case 2: return 3;  // RGB
        case 3: return 1;  // palette index
        case 4: return 2;  // grey + alpha
        case 6: return 4;  // RGBA
        default: return 0;
    }
}

bool PngDepthAllowed(std::uint8_t color, std::uint8_t depth) {
    switch (color) {
        case 0: return depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16;
        case 3: return depth == 1 || depth == 2 || depth == 4 || depth == 8;
        case 2: case 4: case 6: return depth == 8 || depth == 16;
        default: return false;
    }
}

std::uint8_t Paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return static_cast<std::uint8_t>(a);
    return static_cast<std::uint8_t>(pb <= pc ? b : c);
}

// Reverses the row filters of one (sub-)image in place: `rows` rows of
// 1 + row_bytes bytes each, filter type first. Returns false on an unknown
// filter type.
bool Unfilter(std::uint8_t* data, std::size_t rows, std::size_t row_bytes, std::size_t bpp) {
    std::vector<std::uint8_t> zero(row_bytes, 0);
    const std::uint8_t* prev = zero.data();
    for (std::size_t y = 0; y < rows; ++y) {
        std::uint8_t* line = data + y * (row_bytes + 1);
        const std::uint8_t type = line[0];
        std::uint8_t* cur = line + 1;
        switch (type) {
            case 0: break;
            case 1:
                for (std::size_t i = bpp; i < row_bytes; ++i) cur[i] = static_cast<std::uint8_t>(cur[i] + cur[i - bpp]);
                break;
            case 2:
                for (std::size_t i = 0; i < row_bytes; ++i) cur[i] = static_cast<std::uint8_t>(cur[i] + prev[i]);
                break;
            case 3:
                for (std::size_t i = 0; i < row_bytes; ++i) {
                    const int left = i >= bpp ? cur[i - bpp] : 0;
                    cur[i] = static_cast<std::uint8_t>(cur[i] + ((left + prev[i]) >> 1));
                }
                break;
            case 4:
                for (std::size_t i = 0; i < row_bytes; ++i) {
                    const int left = i >= bpp ? cur[i - bpp] : 0;
                    const int up_left = i >= bpp ? prev[i - bpp] : 0;
                    cur[i] = static_cast<std::uint8_t>(cur[i] + Paeth(left, prev[i], up_left));
                }
                break;
            default:
                return false;
        }
        prev = cur;
    }
    return true;
}

// Sample `index` of a row of `depth`-bit samples (MSB first below 8 bits,
// big-endian at 16).
std::uint32_t RowSample(const std::uint8_t* row, std::size_t index, std::uint8_t depth) {
    switch (depth) {
        case 16: return (std::uint32_t{row[2 * index]} << 8) | row[2 * index + 1];
        case 8: return row[index];
        default: {
            const std::size_t bit = index * depth;
            const unsigned shift = 8u - depth - static_cast<unsigned>(bit % 8);
            return (row[bit / 8] >> shift) & ((1u << depth) - 1u);
        }
    }
}

}  // namespace

std::optional<Image> DecodePng(const std::uint8_t* data, std::size_t size, std::string& error) {
    if (size < 8 || std::memcmp(data, kPngSignature, 8) != 0) {
        error = "not a PNG file";
        return std::nullopt;
    }
    PngHeader h;
    bool have_header = false, have_end = false;
    std::vector<std::uint8_t> idat;
    std::array<std::array<std::uint8_t, 3>, 256> palette{};
    std::size_t palette_size = 0;
    std::array<std::uint8_t, 256> palette_alpha;
    palette_alpha.fill(255);
    bool have_trns = false;
    std::array<std::uint16_t, 3> trns_key{};
    for (std::size_t at = 8; at < size && !have_end;) {
        if (size - at < 12) {
            error = "truncated PNG chunk";
            return std::nullopt;
        }
        const std::uint32_t length = ReadBe32(data + at);
        const std::uint8_t* type = data + at + 4;
        if (length > size - at - 12) {
            error = "truncated PNG chunk";
            return std::nullopt;
        }
        const std::uint8_t* body = type + 4;
        const bool critical = (type[0] & 0x20) == 0;
        const std::uint32_t crc = ReadBe32(body + length);
        if (critical && crc32(crc32(0L, Z_NULL, 0), type, 4 + length) != crc) {
            error = "PNG chunk CRC mismatch";
            return std::nullopt;
        }
        const std::string name(reinterpret_cast<const char*>(type), 4);
        if (!have_header && name != "IHDR") {
            error = "PNG does not start with IHDR";
            return std::nullopt;
        }
        if (name == "IHDR") {
            if (length != 13 || have_header) {
                error = "invalid PNG IHDR";
                return std::nullopt;
            }
            h.width = ReadBe32(body);
            h.height = ReadBe32(body + 4);
            h.depth = body[8];
            h.color = body[9];
            h.interlace = body[12];
            if (h.width == 0 || h.height == 0 || h.width > 0x7fffffffu || h.height > 0x7fffffffu ||
                !PngDepthAllowed(h.color, h.depth) || body[10] != 0 || body[11] != 0 || h.interlace > 1) {
                error = h.color != 0 && h.color != 2 && h.color != 3 && h.color != 4 && h.color != 6
                            ? "unsupported PNG channel count"
                            : "invalid PNG IHDR";
                return std::nullopt;
            }
            have_header = true;
        } else if (name == "PLTE") {
            if (length % 3 != 0 || length == 0 || length / 3 > 256) {
                error = "invalid PNG palette";
                return std::nullopt;
            }
            palette_size = length / 3;
            for (std::size_t i = 0; i < palette_size; ++i)
                for (int c = 0; c < 3; ++c) palette[i][c] = body[3 * i + c];
        } else if (name == "tRNS") {
            if (h.color == 3) {
                if (length > 256) {
                    error = "invalid PNG tRNS";
                    return std::nullopt;
                }
                for (std::uint32_t i = 0; i < length; ++i) palette_alpha[i] = body[i];
                have_trns = length > 0;
            } else if (h.color == 0 && length == 2) {
                trns_key[0] = static_cast<std::uint16_t>((body[0] << 8) | body[1]);
                have_trns = true;
            } else if (h.color == 2 && length == 6) {
                for (int c = 0; c < 3; ++c)
                    trns_key[c] = static_cast<std::uint16_t>((body[2 * c] << 8) | body[2 * c + 1]);
                have_trns = true;
            }  // tRNS is not allowed with an alpha channel: ignored, as libpng warns and skips
        } else if (name == "IDAT") {
            idat.insert(idat.end(), body, body + length);
        } else if (name == "IEND") {
            have_end = true;
        } else if (critical) {
            error = "unknown critical PNG chunk " + name;
            return std::nullopt;
        }
        at += 12 + std::size_t{length};
    }
    if (!have_header || idat.empty()) {
        error = "PNG has no image data";
        return std::nullopt;
    }
    if (h.color == 3 && palette_size == 0) {
        error = "PNG palette image has no PLTE";
        return std::nullopt;
    }

    // Passes: the whole image, or Adam7's seven sub-images.
    struct Pass {
        std::uint32_t x0, y0, dx, dy, w, h;
        std::size_t row_bytes, offset;
    };
    static constexpr std::uint32_t kAdam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                                   {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const int channels = PngChannels(h.color);
    const std::size_t bits_per_pixel = static_cast<std::size_t>(channels) * h.depth;
    const std::size_t bpp = std::max<std::size_t>(1, bits_per_pixel / 8);
    std::vector<Pass> passes;
    std::uint64_t total = 0;
    const int pass_count = h.interlace ? 7 : 1;
    for (int p = 0; p < pass_count; ++p) {
        Pass s{};
        if (h.interlace) {
            s.x0 = kAdam7[p][0], s.y0 = kAdam7[p][1], s.dx = kAdam7[p][2], s.dy = kAdam7[p][3];
        } else {
            s.x0 = 0, s.y0 = 0, s.dx = 1, s.dy = 1;
        }
        s.w = h.width > s.x0 ? (h.width - s.x0 + s.dx - 1) / s.dx : 0;
        s.h = h.height > s.y0 ? (h.height - s.y0 + s.dy - 1) / s.dy : 0;
        s.row_bytes = (static_cast<std::uint64_t>(s.w) * bits_per_pixel + 7) / 8;
        s.offset = static_cast<std::size_t>(total);
        if (s.w != 0 && s.h != 0) total += static_cast<std::uint64_t>(s.h) * (s.row_bytes + 1);
        passes.push_back(s);
    }
    if (total > kMaxPngBytes) {
        error = "PNG image is too large";
        return std::nullopt;
    }
    std::vector<std::uint8_t> raw(static_cast<std::size_t>(total));
    z_stream z{};
    if (inflateInit(&z) != Z_OK) {
        error = "zlib initialisation failed";
        return std::nullopt;
    }
    if (idat.size() > 0xffffffffu) {
        inflateEnd(&z);
        error = "PNG image data is too large";
        return std::nullopt;
    }
    z.next_in = idat.data();
    z.avail_in = static_cast<uInt>(idat.size());
    std::size_t produced = 0;
    int rc = Z_OK;
    // Output in pieces of at most 1 GiB (avail_out is 32-bit); data past the
    // image's last row is ignored.
    while (produced < raw.size() && rc == Z_OK) {
        const std::size_t chunk = std::min<std::size_t>(raw.size() - produced, std::size_t{1} << 30);
        z.next_out = raw.data() + produced;
        z.avail_out = static_cast<uInt>(chunk);
        rc = inflate(&z, Z_NO_FLUSH);
        produced += chunk - z.avail_out;
    }
    inflateEnd(&z);
    if ((rc != Z_OK && rc != Z_STREAM_END) || produced < raw.size()) {
        error = rc == Z_DATA_ERROR ? "corrupt PNG image data" : "truncated PNG image data";
        return std::nullopt;
    }
    for (const Pass& s : passes)
        if (s.w != 0 && s.h != 0 && !Unfilter(raw.data() + s.offset, s.h, s.row_bytes, bpp)) {
            error = "invalid PNG row filter";
            return std::nullopt;
        }

    // Output channels after expansion: palette -> RGB, tRNS -> alpha.
    const bool alpha = h.color == 4 || h.color == 6 || have_trns;
    const int colour_channels = (h.color == 0 || h.color == 4) ? 1 : 3;
    const bool wide = h.depth == 16;
    Image img;
    img.size = {h.width, h.height};
    img.format = ImageFormat{};
    img.format.type = wide ? PixelType::Word : PixelType::Byte;
    img.format.family = colour_channels == 1 ? ColorFamily::Grey : ColorFamily::Rgb;
    img.format.alpha = alpha ? AlphaType::Straight : AlphaType::None;
    // A grey plane has no matrix; BT.709 keeps it out of the RGB family's value.
    img.format.matrix = colour_channels == 1 ? MatrixCoefficients::Bt709 : MatrixCoefficients::Rgb;
    img.format.transfer = TransferCharacteristics::Srgb;
    img.format.primaries = ColorPrimaries::Bt709;
    img.format.depth = wide ? 16 : 0;
    const std::size_t sample_bytes = wide ? 2 : 1;
    std::array<Size, 4> dims{};
    const auto off = PlaneOffsets(img.format, img.size, sample_bytes, dims);
    img.data.assign(RequiredBytes(img.format, img.size), 0);
    const std::uint32_t max_value = wide ? 65535u : 255u;
    const std::uint32_t grey_scale = h.depth < 8 ? 255u / ((1u << h.depth) - 1u) : 1u;
    auto put = [&](int plane, std::size_t x, std::size_t y, std::uint32_t v) {
        std::uint8_t* base = img.data.data() + off[plane];
        const std::size_t i = y * h.width + x;
        if (wide) {
            const auto w = static_cast<std::uint16_t>(v);
            std::memcpy(base + 2 * i, &w, 2);  // host byte order
        } else {
            base[i] = static_cast<std::uint8_t>(v);
        }
    };
    for (const Pass& s : passes) {
        for (std::uint32_t py = 0; py < s.h; ++py) {
            const std::uint8_t* row = raw.data() + s.offset + py * (s.row_bytes + 1) + 1;
            const std::size_t y = s.y0 + static_cast<std::size_t>(py) * s.dy;
            for (std::uint32_t px = 0; px < s.w; ++px) {
                const std::size_t x = s.x0 + static_cast<std::size_t>(px) * s.dx;
                const std::size_t first = static_cast<std::size_t>(px) * channels;
                switch (h.color) {
                    case 0: {
                        const std::uint32_t g = RowSample(row, first, h.depth);
                        put(0, x, y, g * grey_scale);
                        if (alpha) put(3, x, y, g == trns_key[0] ? 0 : max_value);
                        break;
                    }
                    case 2: {
                        std::uint32_t c[3];
                        for (int k = 0; k < 3; ++k) put(k, x, y, c[k] = RowSample(row, first + k, h.depth));
                        if (alpha)
                            put(3, x, y, c[0] == trns_key[0] && c[1] == trns_key[1] && c[2] == trns_key[2] ? 0 : max_value);
                        break;
                    }
                    case 3: {
                        const std::uint32_t index = RowSample(row, first, h.depth);
                        if (index >= palette_size) {
                            error = "PNG palette index out of range";
                            return std::nullopt;
                        }
                        for (int k = 0; k < 3; ++k) put(k, x, y, palette[index][k]);
                        if (alpha) put(3, x, y, palette_alpha[index]);
                        break;
                    }
                    case 4:
                        put(0, x, y, RowSample(row, first, h.depth));
                        put(3, x, y, RowSample(row, first + 1, h.depth));
                        break;
                    case 6:
                        for (int k = 0; k < 3; ++k) put(k, x, y, RowSample(row, first + k, h.depth));
                        put(3, x, y, RowSample(row, first + 3, h.depth));
                        break;
                }
            }
        }
    }
    return img;
}

// ---- YUV4MPEG2 ---------------------------------------------------------------------
//
// A stream is "YUV4MPEG2" + parameters up to a newline, then frames, each a
// "FRAME" line (parameters allowed) followed by the planes. Parameters are a
// letter and a value: W and H the size, C the chroma layout, I the
// interlacing, F the frame rate; the rest (A, X) carry nothing the decoder
// needs.

namespace {

// What a C value declares: the family, the chroma subsampling and siting,
// and whether an alpha plane follows. A stream without C is 420jpeg.
struct ChromaLayout {
    std::string_view name;
    ColorFamily family;
    std::uint8_t sub_w, sub_h;
    ChromaLocation location;
    bool alpha;
};
constexpr ChromaLayout kChromaLayouts[] = {
    {"mono", ColorFamily::Grey, 0, 0, ChromaLocation::Left, false},
    {"420", ColorFamily::Yuv, 1, 1, ChromaLocation::Left, false},
    {"420mpeg2", ColorFamily::Yuv, 1, 1, ChromaLocation::Left, false},
    {"420jpeg", ColorFamily::Yuv, 1, 1, ChromaLocation::Center, false},
    {"420paldv", ColorFamily::Yuv, 1, 1, ChromaLocation::TopLeft, false},
    {"422", ColorFamily::Yuv, 1, 0, ChromaLocation::Left, false},
    {"444", ColorFamily::Yuv, 0, 0, ChromaLocation::Left, false},
    {"411", ColorFamily::Yuv, 2, 0, ChromaLocation::Left, false},
    {"444alpha", ColorFamily::Yuv, 0, 0, ChromaLocation::Left, true},
};
// High bit depths are written as a "p<depth>" suffix of the C value.
constexpr unsigned kSuffixDepths[] = {9, 10, 12, 14, 16};

// A whole decimal number, or nullopt.
std::optional<std::uint64_t> Decimal(std::string_view s) {
    std::uint64_t v = 0;
    const auto [end, ec] = std::from_chars(s.data(), s.data() + s.size(), v);
    if (s.empty() || ec != std::errc{} || end != s.data() + s.size()) return std::nullopt;
    return v;
}

// The layout a C value names and the depth its suffix gives (0: 8 bits).
const ChromaLayout* ResolveChroma(std::string_view value, std::uint8_t& depth) {
    depth = 0;
    for (const unsigned d : kSuffixDepths) {
        const std::string suffix = "p" + std::to_string(d);
        if (value.size() > suffix.size() && value.substr(value.size() - suffix.size()) == suffix) {
            value.remove_suffix(suffix.size());
            depth = static_cast<std::uint8_t>(d);
            break;
        }
    }
    const auto it = std::find_if(std::begin(kChromaLayouts), std::end(kChromaLayouts),
                                 [&](const ChromaLayout& l) { return l.name == value; });
    return it == std::end(kChromaLayouts) ? nullptr : &*it;
}

// F<num>:<den>, reduced; 0/0 unless both parts are positive 32-bit numbers.
FrameRate ParseRate(std::string_view value) {
    const std::size_t colon = value.find(':');
    if (colon == std::string_view::npos) return {};
    const auto num = Decimal(value.substr(0, colon)), den = Decimal(value.substr(colon + 1));
    constexpr std::uint64_t kMax = 0xffffffffu;
    if (!num || !den || *num == 0 || *den == 0 || *num > kMax || *den > kMax) return {};
    const std::uint64_t g = std::gcd(*num, *den);
    return {static_cast<std::uint32_t>(*num / g), static_cast<std::uint32_t>(*den / g)};
}

// One line without its newline; false at the end of the stream or when the
// line exceeds kMaxHeaderLine.
bool NextLine(std::istream& in, std::string& line) {
    std::array<char, kMaxHeaderLine + 1> buf{};
    in.getline(buf.data(), static_cast<std::streamsize>(buf.size()));
    const std::streamsize got = in.gcount();
    if (got == 0) return false;
    if (in.fail() && !in.eof()) return false;  // too long
    // getline counts the newline it consumed, not the terminator it stores.
    line.assign(buf.data(), std::strlen(buf.data()));
    return true;
}

}  // namespace

bool IsY4m(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    char magic[sizeof(kY4mSignature) - 1] = {};
    return in.read(magic, sizeof(magic)) && std::memcmp(magic, kY4mSignature, sizeof(magic)) == 0;
}

Y4mReader::Y4mReader(const std::string& path, bool sequence) : path_(path) {
    auto error = [&](const std::string& why) { return std::runtime_error(path + ": " + why); };
    in_.open(path, std::ios::binary);
    if (!in_) throw error("cannot open");
    std::string header;
    if (!NextLine(in_, header)) throw error("truncated y4m header");
    std::istringstream fields(header);
    std::string field;
    if (!(fields >> field) || field != kY4mSignature) throw error("not a YUV4MPEG2 file");

    std::optional<std::uint64_t> width, height;
    std::string chroma = "420jpeg";
    bool interlaced = false;
    while (fields >> field) {
        const char letter = field[0];
        const std::string_view value = std::string_view(field).substr(1);
        if (value.empty()) continue;
        if (letter == 'W') {
            if (!(width = Decimal(value))) throw error("invalid y4m width");
        } else if (letter == 'H') {
            if (!(height = Decimal(value))) throw error("invalid y4m height");
        } else if (letter == 'C') {
            chroma = value;
        } else if (letter == 'I') {
            interlaced = value != "p";
        } else if (letter == 'F') {
            rate_ = ParseRate(value);
        }
    }
    if (!width || !height || *width == 0 || *height == 0) throw error("y4m header has no frame size");
    if (sequence && rate_.numerator == 0) throw error("y4m requires a valid frame rate");
    if (sequence && interlaced) throw error("only progressive y4m is supported");
    std::uint8_t depth = 0;
    const ChromaLayout* layout = ResolveChroma(chroma, depth);
    if (layout == nullptr) throw error("unsupported y4m chroma tag");
    if (*width % (1u << layout->sub_w) != 0 || *height % (1u << layout->sub_h) != 0)
        throw error("y4m dimensions are not divisible by the chroma subsampling");

    size_ = {static_cast<std::size_t>(*width), static_cast<std::size_t>(*height)};
    format_ = ImageFormat{};
    format_.type = depth == 0 ? PixelType::Byte : PixelType::Word;
    format_.depth = depth;
    format_.family = layout->family;
    format_.alpha = layout->alpha ? AlphaType::Straight : AlphaType::None;
    format_.chroma_location = layout->location;
    format_.subsample_w = layout->sub_w;
    format_.subsample_h = layout->sub_h;
    // Y4M declares no colorimetry: limited range, the BT.470BG (BT.601)
    // matrix and BT.709 transfer and primaries, overridable by --input-*.
    format_.range = PixelRange::Limited;
    format_.matrix = MatrixCoefficients::Bt470Bg;
    format_.transfer = TransferCharacteristics::Bt709;
    format_.primaries = ColorPrimaries::Bt709;
}

Y4mReader::~Y4mReader() = default;

bool Y4mReader::Next(Image& image) {
    if (in_.peek() == std::char_traits<char>::eof()) {
        if (in_.bad()) throw std::runtime_error(path_ + ": could not read y4m frame");
        return false;
    }
    std::string marker;
    const std::string_view frame = "FRAME";
    if (!NextLine(in_, marker) || marker.compare(0, frame.size(), frame) != 0 ||
        (marker.size() > frame.size() && marker[frame.size()] != ' '))
        throw std::runtime_error(path_ + ": y4m frame header missing");
    image.size = size_;
    image.format = format_;
    image.data.assign(RequiredBytes(format_, size_), 0);
    // Samples are bytes, or little-endian words (as FFmpeg writes high-depth
    // y4m), which is the host order.
    const std::size_t sample = format_.type == PixelType::Word ? 2 : 1;
    std::array<Size, 4> dims{};
    const std::array<std::size_t, 4> at = PlaneOffsets(format_, size_, sample, dims);
    for (std::size_t p = 0; p < at.size(); ++p) {
        const std::size_t bytes = dims[p].width * dims[p].height * sample;
        if (bytes != 0 && !in_.read(reinterpret_cast<char*>(image.data.data() + at[p]),
                                    static_cast<std::streamsize>(bytes)))
            throw std::runtime_error(path_ + ": truncated y4m frame");
    }
    return true;
}

std::optional<Image> LoadImage(std::string_view path, std::string& error) {
    const std::string name(path);
    std::FILE* f = std::fopen(name.c_str(), "rb");
    if (f == nullptr) {
        error = "could not open file";
        return std::nullopt;
    }
    std::uint8_t head[9] = {};
    const std::size_t got = std::fread(head, 1, sizeof(head), f);
    if (got >= 9 && std::memcmp(head, kY4mSignature, 9) == 0) {
        std::fclose(f);
        try {
            Y4mReader reader(name, false);
            Image img;
            if (!reader.Next(img)) {
                error = "y4m file has no frame";
                return std::nullopt;
            }
            return img;
        } catch (const std::exception& e) {
            error = e.what();
            return std::nullopt;
        }
    }
    // io.cc:76-87: JPEG by its first two bytes, PNG by its signature.
    const bool jpeg = got >= 2 && head[0] == 0xff && head[1] == 0xd8;
    if (!jpeg && (got < 8 || std::memcmp(head, kPngSignature, 8) != 0)) {
        std::fclose(f);
        error = "unsupported image format (PNG, JPEG and YUV4MPEG2 are supported)";
        return std::nullopt;
    }
    std::vector<std::uint8_t> bytes;
    std::fseek(f, 0, SEEK_END);
    const long length = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (length <= 0) {
        std::fclose(f);
        error = "could not read file";
        return std::nullopt;
    }
    bytes.resize(static_cast<std::size_t>(length));
    const bool ok = std::fread(bytes.data(), 1, bytes.size(), f) == bytes.size();
    std::fclose(f);
    if (!ok) {
        error = "could not read file";
        return std::nullopt;
    }
    return jpeg ? DecodeJpeg(bytes.data(), bytes.size(), error) : DecodePng(bytes.data(), bytes.size(), error);
}

// ---- --input-* option values (io.cc:690-780) -----------------------------------------
//
// Each option takes one of a fixed list of names; a bad value is reported
// with the whole list, '|'-separated.

namespace {

template <typename T>
struct Named {
    std::string_view name;
    T value;
};

template <typename T, std::size_t N>
bool Choose(const std::array<Named<T>, N>& names, std::string_view text, T& out, std::string& accepted) {
    const auto hit = std::find_if(names.begin(), names.end(), [&](const Named<T>& n) { return n.name == text; });
    if (hit != names.end()) {
        out = hit->value;
        accepted.clear();
        return true;
    }
    std::ostringstream list;
    for (std::size_t i = 0; i < N; ++i) list << (i ? "|" : "") << names[i].name;
    accepted = list.str();
    return false;
}

using M = MatrixCoefficients;
using T = TransferCharacteristics;
using P = ColorPrimaries;
constexpr std::array<Named<M>, 12> kMatrixNames{{
    {"rgb", M::Rgb}, {"bt709", M::Bt709}, {"fcc", M::Fcc}, {"bt470bg", M::Bt470Bg}, {"st170m", M::St170M},
    {"st240m", M::St240M}, {"ycgco", M::YCgCo}, {"bt2020ncl", M::Bt2020Ncl}, {"bt2020cl", M::Bt2020Cl},
    {"chromaticity-derived-ncl", M::ChromaticityDerivedNcl},
    {"chromaticity-derived-cl", M::ChromaticityDerivedCl}, {"ictcp", M::Bt2100ICtCp},
}};
constexpr std::array<Named<T>, 16> kTransferNames{{
    {"bt709", T::Bt709}, {"bt601", T::Bt601}, {"bt470m", T::Bt470M}, {"bt470bg", T::Bt470Bg},
    {"st240m", T::St240M}, {"linear", T::Linear}, {"log100", T::Log100}, {"log316", T::Log316},
    {"xvycc", T::Iec61966_2_4}, {"bt1361", T::Bt1361}, {"srgb", T::Srgb}, {"bt2020-10", T::Bt2020_10},
    {"bt2020-12", T::Bt2020_12}, {"st2084", T::St2084}, {"st428", T::St428}, {"arib-b67", T::AribB67},
}};
constexpr std::array<Named<P>, 11> kPrimaryNames{{
    {"bt709", P::Bt709}, {"bt470m", P::Bt470M}, {"bt470bg", P::Bt470Bg}, {"st170m", P::St170M},
    {"st240m", P::St240M}, {"film", P::Film}, {"bt2020", P::Bt2020}, {"st428", P::St428},
    {"st431-2", P::St431_2}, {"st432-1", P::St432_1}, {"ebu3213e", P::Ebu3213E},
}};
constexpr std::array<Named<PixelRange>, 2> kRangeNames{{{"limited", PixelRange::Limited},
                                                        {"full", PixelRange::Full}}};
constexpr std::array<Named<ChromaFilter>, 3> kFilterNames{{{"bilinear", ChromaFilter::Bilinear},
                                                           {"bicubic", ChromaFilter::Bicubic},
                                                           {"lanczos", ChromaFilter::Lanczos}}};

}  // namespace

bool ParseMatrix(std::string_view t, MatrixCoefficients& o, std::string& a) { return Choose(kMatrixNames, t, o, a); }
bool ParseTransfer(std::string_view t, TransferCharacteristics& o, std::string& a) {
    return Choose(kTransferNames, t, o, a);
}
bool ParsePrimaries(std::string_view t, ColorPrimaries& o, std::string& a) { return Choose(kPrimaryNames, t, o, a); }
bool ParseRange(std::string_view t, PixelRange& o, std::string& a) { return Choose(kRangeNames, t, o, a); }
bool ParseChromaFilter(std::string_view t, ChromaFilter& o, std::string& a) { return Choose(kFilterNames, t, o, a); }

void InputColorOptions::ApplyTo(ImageFormat& format) const {
    if (matrix) format.matrix = *matrix;
    if (transfer) format.transfer = *transfer;
    if (primaries) format.primaries = *primaries;
    if (range) format.range = *range;
}

}  // namespace vqm::image_io
