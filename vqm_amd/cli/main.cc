// vqm-ss2: the SSIMULACRA 2.1 scoring command on the HIP path. It keeps the
// reference command's contract for this metric (cli/main.cc:42-53,108-180,
// cli/report.cc:221-248, cli/video.cc:276-340): two inputs, PNG or
// progressive YUV4MPEG2; one line with the score as %.8f; --raw first prints
// the 108 norms channel-major as %.12f, one line per frame; a sequence scores
// the mean of its frames, each against the worse of a black and a white
// background (ComputeWorstBackground).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "image_io.h"
#include "ssimulacra2.h"

namespace {

using namespace vqm;

struct Options {
    bool raw = false;
    bool half_planes = false;
    std::optional<PlaneFormat> planes;
    Blur blur = Blur::Sampled;
    MapForm map_form = MapForm::Default;
    image_io::InputColorOptions input;
    std::string original, distorted;
};

const char* kUsage =
    "vqm-ss2 - SSIMULACRA 2.1 on AMD Instinct (HIP)\n"
    "Usage: %s [--metric ssimulacra2] [--raw] [--half-planes] [--planes f32|f16|bf16|u16]\n"
    "          [--chroma-filter bilinear|bicubic|lanczos] [--blur sampled|recursive]\n"
    "          [--map reference|difference]\n"
    "          [--input-matrix M] [--input-transfer T] [--input-primaries P] [--input-range R]\n"
    "          original distorted\n"
    "\n"
    "Inputs may be PNG (8/16-bit), JPEG or progressive YUV4MPEG2 (all frames are scored\n"
    "and averaged; both inputs need the same frame count and frame rate).\n"
    "Y4M chroma tags: Cmono, C420, C420jpeg, C420mpeg2, C420paldv, C422, C444, C411,\n"
    "C444alpha, with p9, p10, p12, p14 or p16 for high bit depths. PNG is sRGB with\n"
    "BT.709 primaries; Y4M defaults to the bt470bg matrix, bt709 transfer and\n"
    "primaries, and limited range. The --input-* flags override both images.\n"
    "--raw prints the 108 norms, channel-major, before the score.\n"
    "--half-planes stores intermediate planes as FP16 (--planes f16); --planes u16\n"
    "is the mixed form (planes in 16-bit fixed point at every scale, FP32 arithmetic).\n"
    "--map reference (the default, except with --planes u16) computes the SSIM' map\n"
    "in the reference's operation order; --map difference rearranges it to track FP64\n"
    "more closely (not in the reference command).\n"
    "--blur recursive uses the kernel of Cloudinary ssimulacra2's recursive Gaussian\n"
    "instead of the 13-tap sampled Gaussian (not in the reference command).\n"
    "--input-matrix     rgb|bt709|fcc|bt470bg|st170m|st240m|ycgco|bt2020ncl|\n"
    "                   bt2020cl|chromaticity-derived-ncl|chromaticity-derived-cl|ictcp\n"
    "--input-transfer   bt709|bt601|bt470m|bt470bg|st240m|linear|log100|log316|xvycc|\n"
    "                   bt1361|srgb|bt2020-10|bt2020-12|st2084|st428|arib-b67\n"
    "--input-primaries  bt709|bt470m|bt470bg|st170m|st240m|film|bt2020|st428|st431-2|\n"
    "                   st432-1|ebu3213e\n"
    "--input-range      limited|full\n"
    "\n"
    "SSIMULACRA 2 returns a score in -inf..100, higher is better.\n";

struct ParseError {
    std::string message;
};

// Returns nullopt after printing help.
std::optional<Options> Parse(int argc, char** argv) {
    Options o;
    std::vector<std::string_view> positional;
    for (int i = 1; i < argc; ++i) {
        const std::string_view a = argv[i];
        auto value = [&]() -> std::string_view {
            if (i + 1 >= argc) throw ParseError{std::string(a) + " requires a value"};
            return argv[++i];
        };
        std::string accepted;
        auto bad = [&](std::string_view v) {
            return ParseError{"invalid " + std::string(a) + " '" + std::string(v) + "' (expected " + accepted + ")"};
        };
        if (a == "-h" || a == "--help") {
            std::printf(kUsage, argv[0]);
            return std::nullopt;
        } else if (a == "--raw") {
            o.raw = true;
        } else if (a == "--half-planes") {
            o.half_planes = true;
        } else if (a == "--metric") {
            const std::string_view v = value();
            if (v != "ssimulacra2")
                throw ParseError{"metric '" + std::string(v) + "' is not provided by vqm-ss2 (ssimulacra2 only)"};
        } else if (a == "--planes") {
            const std::string_view v = value();
            if (v == "f32") o.planes = PlaneFormat::F32;
            else if (v == "f16") o.planes = PlaneFormat::F16;
            else if (v == "bf16") o.planes = PlaneFormat::BF16;
            else if (v == "u16") o.planes = PlaneFormat::U16;
            else throw ParseError{"invalid --planes '" + std::string(v) + "' (expected f32|f16|bf16|u16)"};
        } else if (a == "--blur") {
            const std::string_view v = value();
            if (v == "sampled") o.blur = Blur::Sampled;
            else if (v == "recursive") o.blur = Blur::Recursive;
            else throw ParseError{"invalid --blur '" + std::string(v) + "' (expected sampled|recursive)"};
        } else if (a == "--map") {
            const std::string_view v = value();
            if (v == "reference") o.map_form = MapForm::Reference;
            else if (v == "difference") o.map_form = MapForm::Difference;
            else throw ParseError{"invalid --map '" + std::string(v) + "' (expected reference|difference)"};
        } else if (a == "--chroma-filter") {
            const std::string_view v = value();
            if (!image_io::ParseChromaFilter(v, o.input.chroma_filter, accepted)) throw bad(v);
        } else if (a == "--input-matrix") {
            const std::string_view v = value();
            MatrixCoefficients m{};
            if (!image_io::ParseMatrix(v, m, accepted)) throw bad(v);
            o.input.matrix = m;
        } else if (a == "--input-transfer") {
            const std::string_view v = value();
            TransferCharacteristics t{};
            if (!image_io::ParseTransfer(v, t, accepted)) throw bad(v);
            o.input.transfer = t;
        } else if (a == "--input-primaries") {
            const std::string_view v = value();
            ColorPrimaries p{};
            if (!image_io::ParsePrimaries(v, p, accepted)) throw bad(v);
            o.input.primaries = p;
        } else if (a == "--input-range") {
            const std::string_view v = value();
            PixelRange r{};
            if (!image_io::ParseRange(v, r, accepted)) throw bad(v);
            o.input.range = r;
        } else if (a.size() > 1 && a[0] == '-') {
            throw ParseError{"unknown option " + std::string(a)};
        } else {
            positional.push_back(a);
        }
    }
    if (positional.size() != 2) throw ParseError{"expected two inputs: original distorted"};
    if (o.half_planes && o.planes && *o.planes != PlaneFormat::F16)
        throw ParseError{"--half-planes and --planes disagree"};
    o.original = positional[0];
    o.distorted = positional[1];
    return o;
}

Ssimulacra2Config MakeConfig(const Options& o, Size size) {
    Ssimulacra2Config c;
    c.size = size;
    c.half_planes = o.half_planes;
    if (o.planes) c.plane_format = *o.planes;
    c.chroma_filter = o.input.chroma_filter;
    c.blur = o.blur;
    c.map_form = o.map_form;
    return c;
}

// report.cc:221-248: for each channel, scale and norm: ssim, ring, blur.
void PrintRawNorms(const Ssimulacra2Result& r) {
    for (std::size_t c = 0; c < 3; ++c)
        for (const Ssimulacra2Scale& s : r.scales)
            for (std::size_t n = 0; n < 2; ++n)
                std::printf("%.12f,%.12f,%.12f,", s.avg_ssim[c * 2 + n], s.avg_edgediff[c * 4 + n],
                            s.avg_edgediff[c * 4 + 2 + n]);
    std::printf("\n");
}

// One input as a frame sequence: a PNG is one frame, a Y4M all of its own.
class Frames {
public:
    explicit Frames(const std::string& path) {
        if (image_io::IsY4m(path)) {
            y4m_ = std::make_unique<image_io::Y4mReader>(path);
            size_ = y4m_->size();
        } else {
            std::string error;
            auto img = image_io::LoadImage(path, error);
            if (!img) throw std::runtime_error("Could not load image: " + path + "\n" + error);
            size_ = img->size;
            still_ = std::move(img);
        }
    }
    bool Next(Image& out) {
        if (y4m_) return y4m_->Next(out);
        if (!still_) return false;
        out = std::move(*still_);
        still_.reset();
        return true;
    }
    [[nodiscard]] Size size() const { return size_; }
    [[nodiscard]] bool sequence() const { return y4m_ != nullptr; }
    [[nodiscard]] image_io::FrameRate rate() const { return y4m_ ? y4m_->rate() : image_io::FrameRate{}; }

private:
    std::unique_ptr<image_io::Y4mReader> y4m_;
    std::optional<Image> still_;
    Size size_;
};

int Run(const Options& o) {
    Frames original(o.original), distorted(o.distorted);
    if (original.size() != distorted.size()) {
        std::fprintf(stderr, "Image size mismatch\n");
        return 1;
    }
    const auto ra = original.rate(), rb = distorted.rate();
    if (original.sequence() && distorted.sequence() &&
        (ra.numerator != rb.numerator || ra.denominator != rb.denominator)) {
        std::fprintf(stderr, "Frame rate mismatch\n");
        return 1;
    }
    Ssimulacra2Pipeline pipeline(MakeConfig(o, original.size()));
    double sum = 0.0;
    std::size_t frames = 0;
    Image a, b;
    for (;;) {
        const bool has_a = original.Next(a), has_b = distorted.Next(b);
        if (has_a != has_b) {
            std::fprintf(stderr, "Frame count mismatch\n");
            return 1;
        }
        if (!has_a) break;
        o.input.ApplyTo(a.format);
        o.input.ApplyTo(b.format);
        const Ssimulacra2Result r = ComputeWorstBackground(pipeline, a, b);
        if (o.raw) PrintRawNorms(r);
        sum += r.Score();
        ++frames;
    }
    if (frames == 0) {
        std::fprintf(stderr, "No frames to score\n");
        return 1;
    }
    std::printf("%.8f\n", sum / static_cast<double>(frames));
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    std::optional<Options> options;
    try {
        options = Parse(argc, argv);
    } catch (const ParseError& e) {
        std::fprintf(stderr, "%s\n", e.message.c_str());
        std::fprintf(stderr, kUsage, argv[0]);
        return 2;
    }
    if (!options) return 0;
    try {
        return Run(*options);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
}
