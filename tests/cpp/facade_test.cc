// The C++ facade (include/ssimulacra2.h) against the reference's own C++
// test expectations (tests/metrics/ssimulacra2/pipeline_test.cc,
// score_test.cc, pool_test.cc). Without arguments only the host-side
// contract runs; with --gpu the pipeline tests run on the device.

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <span>
#include <string>
#include <vector>

#include "ssimulacra2.h"
#include "vqm_memory.h"
#include "vqm_metrics.h"

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

// A view of `img` whose rows sit `pad` bytes apart beyond their length, in
// buffers of its own (`store` keeps them alive).
static vqm::ImageView PaddedView(const vqm::Image& img, std::size_t pad, std::vector<std::vector<std::byte>>& store) {
    vqm::ImageView view;
    view.size = img.size;
    view.format = img.format;
    const bool yuv = img.format.family == vqm::ColorFamily::Yuv;
    std::size_t offset = 0;
    for (std::size_t p = 0; p < 3; ++p) {
        const std::size_t w = (yuv && p > 0) ? img.size.width >> img.format.subsample_w : img.size.width;
        const std::size_t h = (yuv && p > 0) ? img.size.height >> img.format.subsample_h : img.size.height;
        std::vector<std::byte> plane((w + pad) * h, std::byte{0xA5});
        for (std::size_t y = 0; y < h; ++y)
            for (std::size_t x = 0; x < w; ++x) plane[y * (w + pad) + x] = std::byte{img.data[offset + y * w + x]};
        offset = (offset + w * h + 3) / 4 * 4;
        store.push_back(std::move(plane));
        view.planes[p] = {store.back().data(), w + pad};
    }
    return view;
}

static vqm::MetricsConfig Srgb8Metrics(vqm::Size size) {
    vqm::MetricsConfig c;
    c.size = size;
    c.original_format = vqm::ImageFormat::Srgb8();
    c.distorted_format = c.original_format;
    c.ssimulacra2 = vqm::Ssimulacra2Config{size};
    return c;
}

// vqm_views.h:32-36 and tests/core/metrics_test.cc:98-240: everything that
// needs no device.
static void ViewAndMetricsHostContract() {
    const vqm::Image rgb = Gradient(1), yuv = Yuv420From(rgb);
    for (const vqm::Image* img : {&rgb, &yuv}) {
        std::vector<std::vector<std::byte>> store;
        const vqm::Image copy = vqm::CopyImage(PaddedView(*img, 7, store));
        CHECK(copy.size == img->size && copy.format == img->format && copy.data == img->data);
        vqm::ImageView short_stride = PaddedView(*img, 0, store);
        short_stride.planes[1].stride_bytes -= 1;
        CHECK_THROWS_AS(vqm::CopyImage(short_stride), std::invalid_argument);
        vqm::ImageView missing = PaddedView(*img, 0, store);
        missing.planes[2].data = nullptr;
        CHECK_THROWS_AS(vqm::CopyImage(missing), std::invalid_argument);
        vqm::ImageView huge = PaddedView(*img, 0, store);
        huge.planes[0].stride_bytes = std::numeric_limits<std::size_t>::max() / 2;
        CHECK_THROWS_AS(vqm::CopyImage(huge), std::overflow_error);
    }
    vqm::ImageView empty;
    CHECK_THROWS_AS(vqm::CopyImage(empty), std::invalid_argument);

    // Selection and configuration checks come before any device work.
    CHECK_THROWS_AS(vqm::MetricPipelines(vqm::MetricsConfig{}), std::invalid_argument);
    CHECK_THROWS_AS(vqm::Metrics(vqm::MetricsConfig{}), std::invalid_argument);
    vqm::MetricsConfig config = Srgb8Metrics({64, 48});
    config.background = 1.5f;
    CHECK_THROWS_AS(vqm::Metrics(config), std::invalid_argument);
    config.background = std::nanf("");
    CHECK_THROWS_AS(vqm::GetMemoryRequirements(config), std::invalid_argument);
    config.background = 0.5f;
    config.original_format.transfer = vqm::TransferCharacteristics::Unspecified;
    bool named = false;
    try {
        (void)vqm::GetMemoryRequirements(config);
    } catch (const std::invalid_argument& e) {
        named = std::string(e.what()).rfind("original format: ", 0) == 0;
    }
    CHECK(named);
    config.original_format = vqm::ImageFormat::Srgb8();

    // The metric's own geometry, resident slots and pool are overridden.
    config.ssimulacra2->size = {8, 8};
    config.ssimulacra2->gpu_resident_images = 5;
    config.ssimulacra2->input_bytes = 1 << 20;
    const vqm::MetricsMemoryRequirements req = vqm::GetMemoryRequirements(config);
    vqm::Ssimulacra2Config single{{64, 48}};
    single.gpu_resident_images = 0;
    const vqm::MetricMemoryRequirements expected = vqm::GetMemoryRequirements(single);
    CHECK(req.spatial.has_value() && !req.scratch.has_value());
    CHECK(req.spatial->direct.supported && !req.spatial->staged.supported);
    CHECK(req.spatial->direct.arena.allocation_bytes == expected.direct.arena.allocation_bytes);
    CHECK(req.spatial->direct.staging.allocation_bytes == expected.direct.staging.allocation_bytes);

    // metrics_test.cc:187-240: per-heap worker counts.
    const std::uint64_t dev = req.spatial->direct.arena.allocation_bytes;
    const std::uint64_t host = req.spatial->direct.staging.allocation_bytes;
    const std::array<std::uint64_t, 2> roomy{5 * dev + 1, 9 * host};
    CHECK(vqm::MaxWorkersForBudgetWorstCase(req, roomy) == 5);
    const std::array<std::uint64_t, 2> host_bound{5 * dev, 2 * host + host / 2};
    CHECK(vqm::MaxWorkersForBudgetWorstCase(req, host_bound) == 2);
    const std::array<std::uint64_t, 1> one_heap{dev};
    CHECK_THROWS_AS(vqm::MaxWorkersForBudgetWorstCase(req, one_heap), std::invalid_argument);
    CHECK_THROWS_AS(vqm::MaxWorkersForBudgetWorstCase(vqm::MetricsMemoryRequirements{}, roomy),
                    std::invalid_argument);
    vqm::MetricsMemoryRequirements huge_req = req;
    huge_req.scratch = vqm::MetricMemoryRequirements{};
    huge_req.scratch->direct.supported = true;
    huge_req.scratch->direct.arena = {1, std::numeric_limits<std::uint64_t>::max(), 0, 0};
    CHECK_THROWS_AS(vqm::MaxWorkersForBudgetWorstCase(huge_req, roomy), std::overflow_error);
    vqm::MetricsConfig staged = Srgb8Metrics({64, 48});
    staged.memory_policy = vqm::MemoryPolicy::Staged;  // reaches the metric
    CHECK(!vqm::GetMemoryRequirements(staged).spatial->direct.supported);
    CHECK_THROWS_AS(vqm::MaxWorkersForBudgetWorstCase(vqm::GetMemoryRequirements(staged), roomy),
                    std::invalid_argument);
}

// Caller code written for the reference's headers, unchanged: the config's
// designated and positional initialisers in the reference's field order
// (ssimulacra2.h:35-64), the memory planning types (vqm_memory.h:26-135,
// vqm_memory_types.h:14-42), the pool's span overloads (vqm_pool.h:65-80)
// and the Image and view format defaults (vqm_image.h:134-141).
static void ReferenceCallerCode() {
    const vqm::Ssimulacra2Config designated{.size = {1920, 1080},
                                            .gpu_resident_images = 2,
                                            .half_planes = false,
                                            .chroma_filter = vqm::ChromaFilter::Bilinear,
                                            .input_bytes = 0,
                                            .streaming_inputs = true,
                                            .memory_policy = vqm::MemoryPolicy::Automatic};
    const vqm::Ssimulacra2Config positional{{1920, 1080}, 2, false, vqm::ChromaFilter::Bicubic, 0, true,
                                            vqm::MemoryPolicy::Direct};
    const vqm::MetricMemoryRequirements plan = vqm::GetMemoryRequirements(designated);
    CHECK(plan.direct.supported && !plan.staged.supported);
    CHECK(plan.direct.arena.heap == vqm::kDeviceHeap && plan.direct.staging.heap == vqm::kPinnedHostHeap);
    CHECK(plan.direct.arena.buffer_bytes == plan.direct.arena.allocation_bytes);
    CHECK(vqm::GetMemoryRequirements(positional).direct.arena.allocation_bytes ==
          plan.direct.arena.allocation_bytes);
    CHECK(plan.heaps.size() == 2 && plan.heaps[0].device_local && !plan.heaps[1].device_local);
    const std::vector<std::uint64_t> budgets{3 * plan.direct.arena.allocation_bytes + 7,
                                             10 * plan.direct.staging.allocation_bytes};
    CHECK(vqm::MaxWorkersForBudget(plan.direct, budgets) == 3);
    const std::array<vqm::MemoryScenario, 2> two{plan.direct, plan.direct};
    CHECK(vqm::MaxWorkersForBudget(std::span<const vqm::MemoryScenario>(two), budgets) == 1);
    const std::array<vqm::MetricMemoryRequirements, 1> worst{plan};
    CHECK(vqm::MaxWorkersForBudgetWorstCase(std::span<const vqm::MetricMemoryRequirements>(worst), budgets) == 3);
    CHECK_THROWS_AS(vqm::MaxWorkersForBudget(plan.staged, budgets), std::invalid_argument);
    const std::array<std::uint64_t, 1> one_heap{budgets[0]};
    CHECK_THROWS_AS(vqm::MaxWorkersForBudget(plan.direct, one_heap), std::invalid_argument);
    if (const auto available = vqm::AvailableHeapBudgets(plan)) CHECK(available->size() == 2);
    vqm::Ssimulacra2Config strict = designated;
    strict.memory_policy = vqm::MemoryPolicy::Staged;
    CHECK(!vqm::GetMemoryRequirements(strict).direct.supported);
    CHECK_THROWS_AS(vqm::Ssimulacra2Pipeline{strict}, std::runtime_error);
    const vqm::Image image;
    const vqm::DeviceImageView view;
    CHECK(image.format == vqm::ImageFormat{} && image.format.type == vqm::PixelType::Float);
    CHECK(view.format == vqm::ImageFormat{});
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
    // vqm_memory.h:62-124.
    vqm::Ssimulacra2Config no_slots{{1920, 1080}};
    no_slots.gpu_resident_images = 0;
    const std::uint64_t with = vqm::GetMemoryRequirements(vqm::Ssimulacra2Config{{1920, 1080}}).direct.arena.allocation_bytes;
    const std::uint64_t without = vqm::GetMemoryRequirements(no_slots).direct.arena.allocation_bytes;
    CHECK(with - without == 2ull * 3 * 1920 * 1080);  // two resident 8-bit RGB slots
    CHECK_THROWS_AS(vqm::MaxWorkersForBudget(vqm::MemoryScenario{}, std::vector<std::uint64_t>{1, 1}),
                    std::invalid_argument);
    CHECK_THROWS_AS(vqm::GetMemoryRequirements(vqm::Ssimulacra2Config{{4, 8}}), std::invalid_argument);
    ViewAndMetricsHostContract();
    ReferenceCallerCode();
}

// tests/core/metrics_test.cc:241-575, the SSIMULACRA 2 parts: every input
// form scores what the pipeline scores, frames are numbered, rejected pairs
// consume no frame, and the sequence states follow the reference.
static void MetricsDeviceContract() {
    const vqm::Image orig = Gradient(0), dist = Gradient(1);
    vqm::Ssimulacra2Pipeline pipe(vqm::Ssimulacra2Config{{32, 32}});
    const double want = pipe.Compute(orig, dist).Score();
    std::vector<std::vector<std::byte>> store;
    const vqm::ImageView va = PaddedView(orig, 5, store), vb = PaddedView(dist, 3, store);
    CHECK(pipe.Compute(va, vb).Score() == want);
    CHECK(vqm::Score(pipe, va, vb) == want);

    vqm::Metrics metrics(Srgb8Metrics({32, 32}));
    auto first = metrics.Compute(orig, dist);
    CHECK(first && first->frame_index == 0 && first->ssimulacra2 && *first->ssimulacra2 == want);
    vqm::Image wrong = dist;
    wrong.size.width -= 1;
    CHECK_THROWS_AS(metrics.Compute(orig, wrong), std::invalid_argument);
    wrong = dist;
    wrong.format.transfer = vqm::TransferCharacteristics::Bt709;
    CHECK_THROWS_AS(metrics.Compute(orig, wrong), std::invalid_argument);
    wrong = dist;
    wrong.data.pop_back();
    CHECK_THROWS_AS(metrics.Compute(orig, wrong), std::invalid_argument);
    auto second = metrics.Compute(va, vb);
    CHECK(second && second->frame_index == 1 && *second->ssimulacra2 == want);
    const auto a = std::as_bytes(std::span(orig.data)), b = std::as_bytes(std::span(dist.data));
    CHECK_THROWS_AS(metrics.Compute(a.first(a.size() - 1), b), std::invalid_argument);
    CHECK_THROWS_AS(metrics.Compute(a, b.first(b.size() - 1)), std::invalid_argument);
    auto third = metrics.Compute(a, b);
    CHECK(third && third->frame_index == 2 && *third->ssimulacra2 == want);
    CHECK_THROWS_AS(metrics.ComputeImage(orig, dist), std::logic_error);  // sequence active
    CHECK(!metrics.Flush());
    CHECK_THROWS_AS(metrics.Compute(orig, dist), std::logic_error);  // sequence finished
    const vqm::MetricsFinalResult fin = metrics.Finalize();
    CHECK(!fin.last_frame && !fin.penultimate_frame && fin.frames == 3);
    metrics.Reset();
    const vqm::MetricsResult still = metrics.ComputeImage(a, b);
    CHECK(still.frame_index == 0 && still.ssimulacra2 && *still.ssimulacra2 == want);
    const vqm::MetricsResult again = metrics.ComputeImage(orig, dist);  // ComputeImage resets
    CHECK(again.frame_index == 0 && *again.ssimulacra2 == want);
    const auto usage = metrics.GetMemoryUsage();
    CHECK(usage.spatial && !usage.scratch);
    CHECK(usage.spatial->arena.allocation_bytes ==
          vqm::GetMemoryRequirements(Srgb8Metrics({32, 32})).spatial->direct.arena.allocation_bytes);

    // Device inputs, and streaming disabled: pixel overloads are refused
    // before any work, device views still score.
    vqm::MetricsConfig device_only = Srgb8Metrics({32, 32});
    device_only.streaming_inputs = false;
    vqm::Metrics dm(device_only);
    CHECK_THROWS_AS(dm.Compute(orig, dist), std::invalid_argument);
    void* d[2] = {nullptr, nullptr};
    for (int i = 0; i < 2; ++i) {
        const vqm::Image& src = i ? dist : orig;
        CHECK(vqm_device_malloc(src.data.size(), &d[i]) == VQM_OK);
        CHECK(vqm_memcpy_htod(d[i], src.data.data(), src.data.size()) == VQM_OK);
    }
    const vqm::DeviceImageView da{d[0], {32, 32}, vqm::ImageFormat::Srgb8()};
    const vqm::DeviceImageView db{d[1], {32, 32}, vqm::ImageFormat::Srgb8()};
    auto dev = dm.Compute(da, db);
    CHECK(dev && dev->frame_index == 0 && *dev->ssimulacra2 == want);
    CHECK_THROWS_AS(dm.Compute(da, vqm::DeviceImageView{nullptr, {32, 32}, vqm::ImageFormat::Srgb8()}),
                    std::invalid_argument);
    auto dev2 = dm.Compute(da, db);
    CHECK(dev2 && dev2->frame_index == 1);
    for (void* p : d) CHECK(vqm_device_free(p) == VQM_OK);

    // MetricPipelines: the getter's pipeline survives a move and scores as
    // a standalone one.
    vqm::MetricPipelines pipelines(Srgb8Metrics({32, 32}));
    vqm::Ssimulacra2Pipeline* held = pipelines.Ssimulacra2();
    vqm::MetricPipelines moved = std::move(pipelines);
    CHECK(moved.Ssimulacra2() == held && held != nullptr);
    CHECK(held->Compute(orig, dist).Score() == want);
    CHECK(moved.GetMemoryUsage().spatial.has_value());

    // The top-level background reaches the metric: with alpha it changes
    // the score, and equals the pipeline's at that background.
    vqm::Image ao = orig, ad = dist;
    for (vqm::Image* img : {&ao, &ad}) {
        img->format.alpha = vqm::AlphaType::Straight;
        img->data.resize(vqm::RequiredBytes(img->format, img->size));
        for (std::size_t i = 0; i < 32 * 32; ++i) img->data[3 * 32 * 32 + i] = static_cast<std::uint8_t>(i * 7);
    }
    vqm::MetricsConfig alpha = Srgb8Metrics({32, 32});
    alpha.original_format = ao.format;
    alpha.distorted_format = ad.format;
    alpha.background = 0.2f;
    vqm::Metrics am(alpha);
    const double at02 = pipe.Compute(ao, ad, 0.2f).Score(), at08 = pipe.Compute(ao, ad, 0.8f).Score();
    CHECK(at02 != at08);
    CHECK(*am.ComputeImage(ao, ad).ssimulacra2 == at02);
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
    const std::vector<vqm::ImagePair> pairs{{orig, dist_b}, {orig, dist_c}, {orig, orig}};
    const auto scores = pool.ScoreMany(pairs);
    CHECK(scores.size() == 3 && scores[0] == b && scores[1] == c && scores[2] == 100.0);
    std::vector<std::vector<std::byte>> store;
    const vqm::ImageView vo = PaddedView(orig, 3, store), vb = PaddedView(dist_b, 0, store);
    const std::array<vqm::ImageViewPair, 2> view_pairs{{{vo, vb}, {vo, vo}}};
    const auto view_scores = pool.ScoreMany(view_pairs);
    CHECK(view_scores[0] == b && view_scores[1] == 100.0);
    CHECK(pool.Workers() == 2);
    // ssimulacra2.h:204-205 as reference callers read it.
    const auto usage = pipe.GetMemoryUsage();
    CHECK(usage.arena.buffer_bytes > 0 && usage.arena.heap == vqm::kDeviceHeap && usage.uses_staging);
    CHECK(usage.staging.allocation_bytes > 0 && usage.auxiliary.empty());
    CHECK(pipe.GetMemoryRegions().resident_bytes == 2 * 3 * 1024);
    // A general format with a non-default chroma filter (input_format_chroma_test.cc).
    vqm::Ssimulacra2Config lanczos{{32, 32}};
    lanczos.chroma_filter = vqm::ChromaFilter::Lanczos;
    vqm::Ssimulacra2Pipeline yuv_pipe(lanczos);
    const vqm::Image ya = Yuv420From(orig), yb = Yuv420From(dist_b);
    CHECK(yuv_pipe.Compute(ya, ya).Score() == 100.0);
    const double yuv_score = yuv_pipe.Compute(ya, yb).Score();
    CHECK(yuv_score > 50.0 && yuv_score < 100.0);
    MetricsDeviceContract();
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
