// Host-side logic of the SSIMULACRA 2.1 path (see ss2_host.h).

#include "ss2_host.h"

#include "ss2_source.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <limits>

namespace vqm::ss2 {

void Fail(int status, std::string message) { throw Error{status, std::move(message)}; }

namespace {

// Subjective weights per (channel, scale) row, in the order ssim_1, ring_1,
// blur_1, ssim_4, ring_4, blur_4 (score.cc:29-87). Published constants of the
// metric: the same numbers, typed from the table.
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

// Score mapping constants, score.cc:18-24.
constexpr double kScoreScale = 0.9562382616834844;
constexpr double kScoreLinear = 2.326765642916932;
constexpr double kScoreQuadratic = 0.020884521182843837;
constexpr double kScoreCubic = 6.248496625763138e-05;
constexpr double kScoreExponent = 0.6276336467831387;
constexpr double kWeightPruning = 0.01;  // score.cc:97

constexpr uint32_t kMinDimension = 8;

uint64_t CheckedMul(uint64_t a, uint64_t b) {
    if (b != 0 && a > std::numeric_limits<uint64_t>::max() / b)
        Fail(VQM_OVERFLOW, "image layout exceeds 64-bit sizes");
    return a * b;
}

uint64_t Align(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

// Bytes per sample of a validated pixel type (Byte, Word, Half, Float).
uint64_t SampleBytes(const vqm_image_format& f) { return f.type == 0 ? 1 : f.type == 3 ? 4 : 2; }

}  // namespace

ScalePlan PlanScales(uint32_t width, uint32_t height) {
    if (width < kMinDimension || height < kMinDimension)
        Fail(VQM_INVALID_ARGUMENT, "minimum image size is 8x8 pixels");
    ScalePlan plan;
    uint32_t w = width, h = height;
    // The previous level decides whether the next is emitted, so the last
    // level may be below 8 pixels.
    while (plan.count < static_cast<uint32_t>(kMaxScales) && w >= kMinDimension &&
           h >= kMinDimension) {
        if (plan.count != 0) {
            w = w / 2 + w % 2;
            h = h / 2 + h % 2;
        }
        plan.w[plan.count] = w;
        plan.h[plan.count] = h;
        ++plan.count;
    }
    return plan;
}

std::array<float, kTaps> GaussianTaps() {
    constexpr double sigma = 1.5;
    const int radius = static_cast<int>(std::ceil(3.5 * sigma));
    std::array<double, kTaps> w{};
    double sum = 0.0;
    for (int i = 0; i < kTaps; ++i) {
        const double offset = static_cast<double>(i - radius) / sigma;
        w[i] = std::exp(-0.5 * offset * offset);
        sum += w[i];
    }
    std::array<float, kTaps> taps{};
    for (int i = 0; i < kTaps; ++i) taps[i] = static_cast<float>(w[i] / sum);
    return taps;
}

bool ChannelPruned(uint32_t channel, uint32_t scale, uint32_t scales) {
    if (channel >= static_cast<uint32_t>(kChannels) || scale >= scales ||
        scales > static_cast<uint32_t>(kMaxScales))
        return false;
    // The table is read as one flat run of `scales` rows per channel.
    const double* w = kWeights[channel * scales + scale];
    return *std::max_element(w, w + 6) <= kWeightPruning;
}

double Score(uint32_t scales, const double* norms) {
    if (scales > static_cast<uint32_t>(kMaxScales))
        Fail(VQM_INVALID_ARGUMENT, "Ssimulacra2Result has too many scales");
    double error = 0.0;
    for (uint32_t c = 0; c < 3; ++c) {
        for (uint32_t s = 0; s < scales; ++s) {
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
    if (!std::isfinite(error))
        Fail(VQM_INVALID_ARGUMENT, "Ssimulacra2Result contains nonfinite or overflowing norms");
    const double s = error * kScoreScale;
    const double mapped = kScoreLinear * s - kScoreQuadratic * s * s + kScoreCubic * s * s * s;
    if (!std::isfinite(mapped)) Fail(VQM_INVALID_ARGUMENT, "Ssimulacra2Result score mapping overflowed");
    return mapped > 0.0 ? 100.0 - 10.0 * std::pow(mapped, kScoreExponent) : 100.0;
}

std::array<float, 256> SrgbTable() {
    std::array<float, 256> table{};
    const float scale = static_cast<float>(1.0 / 255.0);
    for (int code = 0; code < 256; ++code) {
        const float v = static_cast<float>(code) * scale;
        if (v <= 0.04045f) {
            table[code] = v / 12.92f;
        } else {
            const float t = (v + 0.055f) / 1.055f;
            table[code] = static_cast<float>(std::pow(static_cast<double>(t), 2.4));
        }
    }
    return table;
}

uint64_t RequiredBytes(const vqm_image_format& format, uint32_t width, uint32_t height) {
    uint64_t sample = 0;
    switch (format.type) {
        case 0: sample = 1; break;
        case 1: case 2: sample = 2; break;
        case 3: sample = 4; break;
        default: Fail(VQM_INVALID_ARGUMENT, "invalid pixel type");
    }
    if (format.subsample_w > 2 || format.subsample_h > 2)
        Fail(VQM_INVALID_ARGUMENT, "subsampling greater than 4x is not supported");
    const bool chroma = format.family != 0;
    const uint64_t plane = CheckedMul(CheckedMul(width, height), sample);
    const uint64_t chroma_plane =
        CheckedMul(CheckedMul(width >> format.subsample_w, height >> format.subsample_h), sample);
    uint64_t total = Align(plane, 4);
    if (chroma) total = Align(total + chroma_plane, 4), total = Align(total + chroma_plane, 4);
    if (format.alpha != 0) total = Align(total + plane, 4);
    return total;
}

void CheckImage(const vqm_image_desc& image, uint32_t width, uint32_t height, const char* name) {
    const std::string who(name);
    if (image.width != width || image.height != height)
        Fail(VQM_INVALID_ARGUMENT, who + ": image size does not match the pipeline geometry");
    ValidateFormat(image.format, width, height, who);
    const uint64_t expected = RequiredBytes(image.format, width, height);
    if (image.bytes != expected)
        Fail(VQM_INVALID_ARGUMENT, who + ": data size does not match the image format (expected " +
                                       std::to_string(expected) + " bytes)");
}

uint64_t CheckView(const vqm_host_view& view) {
    if (view.width == 0 || view.height == 0) Fail(VQM_INVALID_ARGUMENT, "image view: size must be non-zero");
    ValidateFormat(view.format, view.width, view.height, "image view");
    const uint64_t sample = SampleBytes(view.format);
    const auto planes = PlaneLayout(view.format, view.width, view.height);
    constexpr uint64_t kMaxOffset = static_cast<uint64_t>(std::numeric_limits<std::ptrdiff_t>::max());
    for (int p = 0; p < 4; ++p) {
        if (!planes[p].present) continue;
        const uint64_t row = CheckedMul(planes[p].width, sample);
        const uint64_t stride = view.strides[p];
        if (view.planes[p] == nullptr || stride < row)
            Fail(VQM_INVALID_ARGUMENT, "image view: missing plane or short stride");
        // The last row's end, (height - 1) * stride + row, must be a
        // representable pointer difference before any row address is formed.
        if (row > kMaxOffset || (planes[p].height > 1 && stride > (kMaxOffset - row) / (planes[p].height - 1)))
            Fail(VQM_OVERFLOW, "image view: plane stride is too large");
    }
    return RequiredBytes(view.format, view.width, view.height);
}

void PackView(const vqm_host_view& view, uint8_t* dst) {
    const uint64_t sample = SampleBytes(view.format);
    const auto planes = PlaneLayout(view.format, view.width, view.height);
    for (int p = 0; p < 4; ++p) {
        if (!planes[p].present) continue;
        const uint64_t row = planes[p].width * sample, rows = planes[p].height;
        const auto* src = static_cast<const uint8_t*>(view.planes[p]);
        uint8_t* out = dst + planes[p].offset;
        if (view.strides[p] == row) {
            std::memcpy(out, src, row * rows);
        } else {
            for (uint64_t y = 0; y < rows; ++y) std::memcpy(out + y * row, src + y * view.strides[p], row);
        }
        const uint64_t end = planes[p].offset + row * rows;
        std::memset(dst + end, 0, Align(end, 4) - end);
    }
}

Plan FillPlan(uint32_t width, uint32_t height, PlaneFormat format, uint32_t strip_rows, uint32_t strip_cols) {
    const ScalePlan scales = PlanScales(width, height);
    if (strip_rows == 0) strip_rows = 64;
    if (strip_cols == 0) strip_cols = kDefaultStripCols;
    if (strip_rows % kSumBlockRows != 0 || strip_rows > 1024)
        Fail(VQM_INVALID_ARGUMENT, "strip_rows must be a multiple of 8 up to 1024");
    if (!ValidStripCols(strip_cols)) Fail(VQM_INVALID_ARGUMENT, "strip_cols must be 64 or 128");
    if (width > (1u << 20) || height > (1u << 20))
        Fail(VQM_INVALID_ARGUMENT, "image dimensions above 1048576 are not supported");
    Plan p{};
    p.width = width;
    p.height = height;
    p.levels = scales.count;
    p.strip_h = strip_rows;
    p.strip_w = strip_cols;
    p.plane_stride = static_cast<uint32_t>(Align(static_cast<uint64_t>(width) * height, 4));
    uint64_t cursor = 0;  // bytes
    for (uint32_t l = 0; l < static_cast<uint32_t>(kMaxScales); ++l) {
        for (int c = 0; c < kChannels; ++c) p.plane_off[l][c] = ~0ull;
        if (l >= scales.count) continue;
        p.w[l] = scales.w[l];
        p.h[l] = scales.h[l];
        const uint32_t strips = (p.w[l] + kStripCols - 1) / kStripCols;
        const uint32_t bands = (p.h[l] + strip_rows - 1) / strip_rows;
        p.pitch[l] = static_cast<uint32_t>(Align(static_cast<uint64_t>(strips) * kStripCols + 14, 4));
        p.rows[l] = bands * strip_rows + 2 * kRadius;
        for (uint32_t c = 0; c < static_cast<uint32_t>(kChannels); ++c) {
            if (ChannelPruned(c, l, scales.count)) continue;
            p.mask[l] |= 1u << c;
            p.plane_off[l][c] = cursor;
            // 256-byte alignment keeps every row start and vector store aligned.
            cursor = Align(cursor + CheckedMul(CheckedMul(p.pitch[l], p.rows[l]), ElementBytes(format)), 256);
        }
    }
    p.image_bytes = cursor;
    // Scale work: per level and active channel, bands of strip_rows rows by
    // strip_cols-column strips, row-major so that consecutive items are
    // neighbours (ItemAt enumerates them).
    uint32_t items = 0;
    for (uint32_t l = 0; l < scales.count; ++l) {
        for (uint32_t c = 0; c < static_cast<uint32_t>(kChannels); ++c) {
            p.item_begin[l][c] = items;
            if (!(p.mask[l] & (1u << c))) continue;
            const uint32_t strips = (p.w[l] + strip_cols - 1) / strip_cols;
            const uint32_t bands = (p.h[l] + strip_rows - 1) / strip_rows;
            p.item_count[l][c] = strips * bands;
            items += strips * bands;
        }
    }
    p.items = items;
    p.prep_tiles_x = (width + kPrepTileW - 1) / kPrepTileW;
    p.prep_tiles_y = (height + kPrepTileH - 1) / kPrepTileH;
    return p;
}

uint32_t DefaultStripRows(uint32_t width, uint32_t height, PlaneFormat format, uint32_t strip_cols,
                          uint64_t pairs, uint32_t cus) {
    const uint64_t per_cu = std::max<uint64_t>(cus, 1);
    auto items = [&](uint32_t rows) {
        return static_cast<uint64_t>(FillPlan(width, height, format, rows, strip_cols).items) * pairs;
    };
    if (items(192) >= 32 * per_cu) return 192;
    if (items(128) >= 8 * per_cu) return 128;
    for (uint32_t rows : {64u, 32u})
        if (items(rows) >= 4 * per_cu) return rows;
    return 16;
}

uint32_t SinglePairStripRows(uint32_t width, uint32_t height, PlaneFormat format, uint32_t strip_rows,
                             uint32_t strip_cols) {
    constexpr uint32_t kWholeDeviceCUs = 256;
    return strip_rows != 0 ? strip_rows : DefaultStripRows(width, height, format, strip_cols, 1, kWholeDeviceCUs);
}

GeometryPlan BuildPlan(uint32_t width, uint32_t height, PlaneFormat format, uint32_t strip_rows,
                       uint32_t strip_cols) {
    GeometryPlan g;
    g.plan = FillPlan(width, height, format, strip_rows, strip_cols);
    g.format = format;
    g.plane_bytes_per_image = g.plan.image_bytes;
    g.prep_tiles = g.plan.prep_tiles_x * g.plan.prep_tiles_y;
    g.items.resize(g.plan.items);
    for (uint32_t i = 0; i < g.plan.items; ++i) g.items[i] = ItemAt(g.plan, i);
    return g;
}

WorkspaceLayout LayoutWorkspace(const GeometryPlan& plan) {
    WorkspaceLayout l;
    l.planes = Align(2 * plan.plane_bytes_per_image, 256);
    l.partials = Align(static_cast<uint64_t>(std::max(plan.plan.items, 1u)) * 6 * sizeof(double), 256);
    l.norms = Align(kNorms * sizeof(double), 256);
    l.total = l.planes + l.partials + l.norms;
    return l;
}

uint64_t AlgorithmicBytes(const GeometryPlan& plan) {
    const Plan& p = plan.plan;
    uint64_t bytes = 0;  // stored plane bytes of one image, unpadded
    for (uint32_t l = 0; l < p.levels; ++l)
        for (int c = 0; c < kChannels; ++c)
            if (p.mask[l] & (1u << c))
                bytes += static_cast<uint64_t>(p.w[l]) * p.h[l] * ElementBytes(plan.format);
    const uint64_t input = 3ull * p.width * p.height;
    return 2 * input + 2 * (2 * bytes);
}

RangeAllocator::RangeAllocator(uint64_t capacity) : capacity_(capacity / kAlign * kAlign) {
    if (capacity_ > 0) free_[0] = capacity_;
}

uint64_t RangeAllocator::Alloc(uint64_t bytes) {
    bytes = Align(std::max<uint64_t>(bytes, 1), kAlign);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        const uint64_t offset = it->first, length = it->second;
        free_.erase(it);
        if (length > bytes) free_[offset + bytes] = length - bytes;
        return offset;
    }
    return kNone;
}

void RangeAllocator::Free(uint64_t offset, uint64_t bytes) {
    bytes = Align(std::max<uint64_t>(bytes, 1), kAlign);
    auto next = free_.lower_bound(offset);
    if (next != free_.end() && offset + bytes == next->first) {  // merge with the range after
        bytes += next->second;
        next = free_.erase(next);
    }
    if (next != free_.begin()) {  // merge with the range before
        auto prev = std::prev(next);
        if (prev->first + prev->second == offset) {
            prev->second += bytes;
            return;
        }
    }
    free_[offset] = bytes;
}

uint64_t RangeAllocator::free_bytes() const {
    uint64_t total = 0;
    for (const auto& [offset, length] : free_) total += length;
    return total;
}

}  // namespace vqm::ss2
