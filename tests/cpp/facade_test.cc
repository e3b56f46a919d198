// The C++ facade (include/ssimulacra2.h) against the reference's own C++
// test expectations (tests/metrics/ssimulacra2/pipeline_test.cc,
// score_test.cc, pool_test.cc). Without arguments only the host-side
// contract runs; with --gpu the pipeline tests run on the device.

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>

#include "ssimulacra2.h"

static int failures = 0;
#define CHECK(cond)                                                          \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                      \
        }                                                                    \
    } while (0)
#define CHECK_THROWS_AS(expr, E)                  \
    do {                                          \
        bool thrown = false;                      \
        try {                                     \
            (void)(expr);                         \
        } catch (const E&) {                      \
            thrown = true;                        \
        } catch (...) {                           \
        }                                         \
        CHECK(thrown && #expr " throws " #E);     \
    } while (0)

// tests/support.h:93-103,259-289.
static vqm::Image Gradient(int variant) {
    constexpr std::size_t side = 32;
    vqm::Image img;
    img.size = {side, side};
    img.format = vqm::ImageFormat::Srgb8();
    img.data.assign(vqm::RequiredBytes(img.format, img.size), 0);
    for (std::size_t y = 0; y < side; ++y)
        for (std::size_t x = 0; x < side; ++x) {
            int r = static_cast<int>(x * 255 / (side - 1));
            int g = static_cast<int>(y * 255 / (side - 1));
            int b = static_cast<int>(((x * 7 + y * 11) * 13) & 255);
            if (variant == 1 && ((x + y) & 1) == 0) r = std::min(r + 16, 255);
            if (variant == 2) r &= 0xF8, g &= 0xF8, b &= 0xF8;
            const std::size_t i = y * side + x;
            img.data[i] = static_cast<std::uint8_t>(r);
            img.data[side * side + i] = static_cast<std::uint8_t>(g);
            img.data[2 * side * side + i] = static_cast<std::uint8_t>(b);
        }
    return img;
}

// 8-bit limited-range BT.709 4:2:0, centre-sited.
static vqm::ImageFormat Yuv420() {
    vqm::ImageFormat f;
    f.type = vqm::PixelType::Byte;
    f.family = vqm::ColorFamily::Yuv;
    f.range = vqm::PixelRange::Limited;
    f.matrix = vqm::MatrixCoefficients::Bt709;
    f.transfer = vqm::TransferCharacteristics::Bt709;
    f.chroma_location = vqm::ChromaLocation::Center;
    f.subsample_w = 1;
    f.subsample_h = 1;
    return f;
}

// A 4:2:0 frame from an 8-bit RGB gradient: its red plane as luma, chroma
// from its green and blue planes, decimated.
static vqm::Image Yuv420From(const vqm::Image& rgb) {
    const std::size_t side = rgb.size.width, half = side / 2;
    vqm::Image img;
    img.size = rgb.size;
    img.format = Yuv420();
    img.data.assign(vqm::RequiredBytes(img.format, img.size), 0);
    for (std::size_t i = 0; i < side * side; ++i) img.data[i] = static_cast<std::uint8_t>(16 + rgb.data[i] * 219 / 255);
    for (std::size_t y = 0; y < half; ++y)
        for (std::size_t x = 0; x < half; ++x) {
            const std::size_t src = 2 * y * side + 2 * x;
            img.data[side * side + y * half + x] = static_cast<std::uint8_t>(16 + rgb.data[side * side + src] * 224 / 255);
            img.data[side * side + half * half + y * half + x] =
                static_cast<std::uint8_t>(16 + rgb.data[2 * side * side + src] * 224 / 255);
        }
    return img;
}

static void HostContract() {
    // pipeline_test.cc:318-328: the geometry is validated before device work.
    bool named = false;
    try {
        vqm::Ssimulacra2Pipeline bad(vqm::Ssimulacra2Config{{4, 8}});
    } catch (const std::invalid_argument& e) {
        named = std::string(e.what()).find("8x8") != std::string::npos;
    }
    CHECK(named);
    // score_test.cc:11-46.
    vqm::Ssimulacra2Result empty;
    CHECK(empty.Score() == 100.0);
    vqm::Ssimulacra2Result r;
    r.scales.resize(6);
    CHECK(r.Score() == 100.0);
    for (double bad : {std::numeric_limits<double>::infinity(), std::nan("")}) {
        r.scales[3].avg_edgediff[5] = bad;
        CHECK_THROWS_AS(r.Score(), std::invalid_argument);
        r.scales[3].avg_edgediff[5] = 0.0;
    }
    r.scales[0].avg_ssim[3] = 1e200;
    CHECK_THROWS_AS(r.Score(), std::invalid_argument);
    r.scales[0].avg_ssim[3] = 100.0;
    CHECK(std::isfinite(r.Score()) && r.Score() < 0.0);
    r.scales.resize(7);
    CHECK_THROWS_AS(r.Score(), std::invalid_argument);
    CHECK(vqm::RequiredBytes(vqm::ImageFormat::Srgb8(), {13, 9}) == 360);
    CHECK(vqm::RequiredBytes(Yuv420(), {8, 4}) == 48);  // input_format_test.cc:122-177
    // vqm_memory.h:63-124.
    const vqm::PipelineMemoryUsage plan = vqm::GetMemoryRequirements(vqm::Ssimulacra2Config{{1920, 1080}});
    CHECK(plan.resident_bytes == 2ull * 3 * 1920 * 1080);
    CHECK(vqm::MaxWorkersForBudget(plan, 3 * plan.total_device_bytes) == 3);
    CHECK_THROWS_AS(vqm::MaxWorkersForBudget(vqm::PipelineMemoryUsage{}, 1), std::invalid_argument);
    CHECK_THROWS_AS(vqm::GetMemoryRequirements(vqm::Ssimulacra2Config{{4, 8}}), std::invalid_argument);
}

static void DeviceContract() {
    const vqm::Image orig = Gradient(0), dist_b = Gradient(1), dist_c = Gradient(2);
    vqm::Ssimulacra2Pipeline pipe(vqm::Ssimulacra2Config{{32, 32}});
    // pipeline_test.cc:146-174 goldens, :120-144 identity.
    const double b = pipe.Compute(orig, dist_b).Score();
    const double c = pipe.Compute(orig, dist_c).Score();
    CHECK(std::fabs(b - 93.17505696) <= 5e-3);
    CHECK(std::fabs(c - 92.12553218) <= 5e-3);
    CHECK(pipe.Compute(orig, orig).Score() == 100.0);
    // Resident handles and in-flight misuse (pipeline_test.cc:380-500).
    const auto ra = pipe.Upload(orig);
    const auto rb = pipe.Upload(dist_b);
    CHECK(pipe.Compute(ra, rb).Score() == b);
    const auto token = pipe.Submit(orig, dist_b);
    CHECK_THROWS_AS(pipe.Submit(orig, orig), std::logic_error);
    CHECK_THROWS_AS(pipe.Upload(orig), std::logic_error);
    CHECK_THROWS_AS(pipe.Release(ra), std::logic_error);
    CHECK(pipe.Finish(token).Score() == b);
    CHECK_THROWS_AS(pipe.Finish(token), std::invalid_argument);
    pipe.Release(ra);
    CHECK_THROWS_AS(pipe.Compute(ra, rb), std::invalid_argument);
    const auto rc = pipe.Upload(orig);  // the released slot is free again
    CHECK(rc.slot != ra.slot);
    CHECK_THROWS_AS(pipe.Upload(orig), std::invalid_argument);  // both slots taken
    // Pool ordering (pool_test.cc:27-62).
    vqm::Ssimulacra2Pool pool(vqm::Ssimulacra2Config{{32, 32}}, 2);
    const auto scores = pool.ScoreMany({{orig, dist_b}, {orig, dist_c}, {orig, orig}});
    CHECK(scores.size() == 3 && scores[0] == b && scores[1] == c && scores[2] == 100.0);
    const auto usage = pipe.GetMemoryUsage();
    CHECK(usage.resident_bytes == 2 * 3 * 1024);
    // A general format with a non-default chroma filter (input_format_chroma_test.cc).
    vqm::Ssimulacra2Config lanczos{{32, 32}};
    lanczos.chroma_filter = vqm::ChromaFilter::Lanczos;
    vqm::Ssimulacra2Pipeline yuv_pipe(lanczos);
    const vqm::Image ya = Yuv420From(orig), yb = Yuv420From(dist_b);
    CHECK(yuv_pipe.Compute(ya, ya).Score() == 100.0);
    const double yuv_score = yuv_pipe.Compute(ya, yb).Score();
    CHECK(yuv_score > 50.0 && yuv_score < 100.0);
}

int main(int argc, char** argv) {
    try {
        HostContract();
        if (argc > 1 && std::strcmp(argv[1], "--gpu") == 0) DeviceContract();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "unexpected exception: %s\n", e.what());
        return 2;
    }
    if (failures) {
        std::fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    std::printf("facade ok\n");
    return 0;
}
