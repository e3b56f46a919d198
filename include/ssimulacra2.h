// ssimulacra2.h -- C++20 facade of the MI355X-native SSIMULACRA 2.1 path.
//
// Header-only: rebuilds the reference's public API (GreatValueCreamSoda/vqm
// include/ssimulacra2.h:17-215, include/vqm_image.h:17-166,
// include/vqm_pool.h:51-86, include/vqm_alpha.h:16-27) on the C ABI of
// libvqm_hip.so (vqm_ss2.h). Names, argument meaning and exceptions follow
// the reference:
//   std::invalid_argument  bad geometry, size or format mismatch, stale
//                          handle or token, full pool
//   std::logic_error       Submit/Upload/Release/AllocateImage/MapImage while
//                          a pair is in flight
//   std::runtime_error     no usable device, HIP failure, allocation failure
// The Vulkan-specific overloads are replaced by DeviceImageView (a device
// pointer to packed planes) and an optional HIP stream to wait on.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vqm_memory_types.h"
#include "vqm_ss2.h"

namespace vqm {

struct Size {
    std::size_t width = 0;
    std::size_t height = 0;
    [[nodiscard]] constexpr bool operator==(const Size&) const noexcept = default;
};

// zimg numeric values, as in the reference (vqm_image.h:25-96).
enum class PixelType : std::uint8_t { Byte = 0, Word = 1, Half = 2, Float = 3 };
enum class PixelRange : std::uint8_t { Limited = 0, Full = 1 };
enum class ColorFamily : std::uint8_t { Grey = 0, Rgb = 1, Yuv = 2 };
enum class AlphaType : std::uint8_t { None = 0, Straight = 1, Premultiplied = 2 };
enum class ChromaLocation : std::uint8_t { Left = 0, Center = 1, TopLeft = 2, Top = 3, BottomLeft = 4, Bottom = 5 };
enum class MatrixCoefficients : std::uint8_t {
    Rgb = 0, Bt709 = 1, Unspecified = 2, Fcc = 4, Bt470Bg = 5, St170M = 6, St240M = 7, YCgCo = 8,
    Bt2020Ncl = 9, Bt2020Cl = 10, St2085YDzDx = 11, ChromaticityDerivedNcl = 12,
    ChromaticityDerivedCl = 13, Bt2100ICtCp = 14, St2128IptC2 = 15, YCgCoRe = 16, YCgCoRo = 17
};
enum class TransferCharacteristics : std::uint8_t {
    Bt709 = 1, Unspecified = 2, Bt470M = 4, Bt470Bg = 5, Bt601 = 6, St240M = 7, Linear = 8,
    Log100 = 9, Log316 = 10, Iec61966_2_4 = 11, Bt1361 = 12, Srgb = 13, Bt2020_10 = 14,
    Bt2020_12 = 15, St2084 = 16, St428 = 17, AribB67 = 18
};
enum class ColorPrimaries : std::uint8_t {
    Bt709 = 1, Unspecified = 2, Bt470M = 4, Bt470Bg = 5, St170M = 6, St240M = 7, Film = 8,
    Bt2020 = 9, St428 = 10, St431_2 = 11, St432_1 = 12, Ebu3213E = 22
};
// Upsampling kernel of subsampled chroma planes (vqm_image.h:94-99).
enum class ChromaFilter : std::uint8_t { Bilinear = 0, Bicubic = 1, Lanczos = 2 };
// The blur of the maps: the reference's sampled Gaussian, or the kernel of
// Cloudinary ssimulacra2's recursive Gaussian (VQM_BLUR_*; no reference field).
enum class Blur : std::uint32_t { Sampled = VQM_BLUR_SAMPLED, Recursive = VQM_BLUR_RECURSIVE };
enum class PlaneFormat : std::uint32_t {
    F32 = VQM_PLANES_F32, F16 = VQM_PLANES_F16, BF16 = VQM_PLANES_BF16, U16 = VQM_PLANES_U16
};
// The arithmetic of the XYB transform and the SSIM' map (VQM_MAP_*; no
// reference field): Reference is the reference's own operation order,
// Difference the rearranged map that tracks FP64 more closely, Default the
// reference order except with U16 planes.
enum class MapForm : std::uint32_t {
    Default = VQM_MAP_DEFAULT, Reference = VQM_MAP_REFERENCE, Difference = VQM_MAP_DIFFERENCE
};


// vqm_image.h:100-120. The reference's default is Float sRGB RGB;
// ImageFormat::Srgb8() is the 8-bit form the table-driven fast decode serves,
// every other accepted format takes the general per-pixel decode.
struct ImageFormat {
    PixelType type = PixelType::Float;
    ColorFamily family = ColorFamily::Rgb;
    AlphaType alpha = AlphaType::None;
    PixelRange range = PixelRange::Full;
    MatrixCoefficients matrix = MatrixCoefficients::Rgb;
    TransferCharacteristics transfer = TransferCharacteristics::Srgb;
    ColorPrimaries primaries = ColorPrimaries::Bt709;
    ChromaLocation chroma_location = ChromaLocation::Left;
    std::uint8_t subsample_w = 0;
    std::uint8_t subsample_h = 0;
    std::uint8_t depth = 0;

    // SameImageFormat (input.h:21-30): every field equal.
    [[nodiscard]] constexpr bool operator==(const ImageFormat&) const noexcept = default;

    static ImageFormat Srgb8() {
        ImageFormat f;
        f.type = PixelType::Byte;
        return f;
    }
    [[nodiscard]] vqm_image_format ToC() const {
        vqm_image_format f{};
        f.type = static_cast<std::uint8_t>(type);
        f.family = static_cast<std::uint8_t>(family);
        f.alpha = static_cast<std::uint8_t>(alpha);
        f.range = static_cast<std::uint8_t>(range);
        f.matrix = static_cast<std::uint8_t>(matrix);
        f.transfer = static_cast<std::uint8_t>(transfer);
        f.primaries = static_cast<std::uint8_t>(primaries);
        f.chroma_location = static_cast<std::uint8_t>(chroma_location);
        f.subsample_w = subsample_w;
        f.subsample_h = subsample_h;
        f.depth = depth;
        return f;
    }
};

// vqm_image.h:134-141: packed planes, each starting on a 4-byte boundary.
// The format defaults to ImageFormat{} (Float sRGB RGB), as in the reference.
struct Image {
    Size size;
    ImageFormat format{};
    std::vector<std::uint8_t> data;
    [[nodiscard]] bool HasAlpha() const noexcept { return format.alpha != AlphaType::None; }
};

struct UploadedImage {  // vqm_image.h:151-154
    std::uint64_t slot = 0;
    Size size;
};

struct ImageBuffer {  // vqm_image.h:162-166
    std::uint64_t slot = 0;
    Size size;
    std::size_t bytes = 0;
};

// vqm_views.h:18-30: a non-owning planar host image. Each present plane
// (Y/R, U/G, V/B, alpha) has its own address and a row stride of at least
// one row; padding at row ends is skipped. Absent planes are ignored.
struct PlaneView {
    const std::byte* data = nullptr;
    std::size_t stride_bytes = 0;
};
struct ImageView {
    Size size;
    ImageFormat format;
    std::array<PlaneView, 4> planes{};
};

// Caller-owned device image (replaces VulkanBufferView, vqm_vulkan.h:74-101):
// must stay alive and unmodified until Finish returns. Like the reference's
// view, its format defaults to ImageFormat{}.
struct DeviceImageView {
    const void* data = nullptr;
    Size size;
    ImageFormat format{};
};

namespace detail {

[[noreturn]] inline void Throw(int status) {
    const std::string msg = vqm_last_error();
    switch (status) {
        case VQM_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case VQM_LOGIC_ERROR: throw std::logic_error(msg);
        case VQM_OVERFLOW: throw std::overflow_error(msg);
        default: throw std::runtime_error(msg);
    }
}

inline void Check(int status) {
    if (status != VQM_OK) Throw(status);
}

inline std::uint32_t Dim(std::size_t v) {
    if (v > 0xFFFFFFFFull) throw std::invalid_argument("image dimension exceeds 32 bits");
    return static_cast<std::uint32_t>(v);
}

inline vqm_image_desc Desc(const Image& image) {
    return vqm_image_desc{Dim(image.size.width), Dim(image.size.height), image.format.ToC(),
                          image.data.size()};
}

inline vqm_host_view HostView(const ImageView& view) {
    vqm_host_view v{};
    v.width = Dim(view.size.width);
    v.height = Dim(view.size.height);
    v.format = view.format.ToC();
    for (std::size_t p = 0; p < 4; ++p) {
        v.planes[p] = view.planes[p].data;
        v.strides[p] = view.planes[p].stride_bytes;
    }
    return v;
}

}  // namespace detail

// RequiredBytes, vqm_image.h:143-146.
[[nodiscard]] inline std::size_t RequiredBytes(const ImageFormat& format, Size size) {
    std::uint64_t bytes = 0;
    const auto f = format.ToC();
    detail::Check(vqm_required_bytes(&f, detail::Dim(size.width), detail::Dim(size.height), &bytes));
    return static_cast<std::size_t>(bytes);
}

// CopyImage, vqm_views.h:32-36: packs a host view into an owning Image (a
// CPU copy). std::invalid_argument for an invalid image or plane,
// std::overflow_error when a plane's geometry or stride is not addressable.
[[nodiscard]] inline Image CopyImage(const ImageView& view) {
    const vqm_host_view v = detail::HostView(view);
    std::uint64_t bytes = 0;
    detail::Check(vqm_copy_image(&v, nullptr, &bytes));
    Image image{view.size, view.format, {}};
    image.data.resize(static_cast<std::size_t>(bytes));
    detail::Check(vqm_copy_image(&v, image.data.data(), &bytes));
    return image;
}

// ssimulacra2.h:17-25.
struct Ssimulacra2Scale {
    static constexpr std::size_t kSsimNormCount = 6;
    static constexpr std::size_t kEdgeNormCount = 12;
    std::array<double, kSsimNormCount> avg_ssim{};
    std::array<double, kEdgeNormCount> avg_edgediff{};
};

// ssimulacra2.h:27-33.
struct Ssimulacra2Result {
    std::vector<Ssimulacra2Scale> scales;

    static Ssimulacra2Result FromNorms(std::uint32_t count, const double* norms) {
        Ssimulacra2Result r;
        r.scales.resize(count);
        for (std::uint32_t s = 0; s < count; ++s) {
            for (std::size_t k = 0; k < 6; ++k) r.scales[s].avg_ssim[k] = norms[s * 18 + k];
            for (std::size_t k = 0; k < 12; ++k) r.scales[s].avg_edgediff[k] = norms[s * 18 + 6 + k];
        }
        return r;
    }
    // Weighted sum of up to 108 norms mapped to a score of at most 100;
    // std::invalid_argument for more than six scales or nonfinite norms.
    [[nodiscard]] double Score() const {
        if (scales.size() > 6) throw std::invalid_argument("Ssimulacra2Result has too many scales");
        std::array<double, 108> flat{};
        for (std::size_t s = 0; s < scales.size(); ++s) {
            for (std::size_t k = 0; k < 6; ++k) flat[s * 18 + k] = scales[s].avg_ssim[k];
            for (std::size_t k = 0; k < 12; ++k) flat[s * 18 + 6 + k] = scales[s].avg_edgediff[k];
        }
        double score = 0.0;
        detail::Check(vqm_ss2_score(static_cast<std::uint32_t>(scales.size()), flat.data(), &score));
        return score;
    }
};

// ssimulacra2.h:35-64: the reference's fields in the reference's order, so
// designated and positional initialisers written for it compile unchanged;
// the fields without a reference counterpart follow them.
struct Ssimulacra2Config {
    Size size;
    std::size_t gpu_resident_images = 2;
    bool half_planes = false;             // FP16 planes
    ChromaFilter chroma_filter = ChromaFilter::Bilinear;
    std::size_t input_bytes = 0;          // caller-owned image pool
    bool streaming_inputs = true;
    MemoryPolicy memory_policy = MemoryPolicy::Automatic;
    // No reference fields below.
    PlaneFormat plane_format = PlaneFormat::F32;  // overridden to F16 by half_planes
    std::uint32_t strip_rows = 0;         // scale work-item height, a multiple of 8 (0 = chosen per geometry)
    bool graph = false;  // replay one captured HIP graph per pair
    Blur blur = Blur::Sampled;
    MapForm map_form = MapForm::Default;
    std::uint32_t strip_cols = 0;         // scale work-item width, 64 or 128 (0 = 128)

    [[nodiscard]] vqm_ss2_config ToC() const {
        if (memory_policy == MemoryPolicy::Staged)
            throw std::runtime_error(
                "MemoryPolicy::Staged is unavailable: pipelines hold their arena in device memory");
        vqm_ss2_config c{};
        vqm_ss2_config_init(&c, detail::Dim(size.width), detail::Dim(size.height));
        if (gpu_resident_images > 0xFFFFFFFFull)
            throw std::invalid_argument("gpu_resident_images is limited to 4096 slots");
        c.resident_images = static_cast<std::uint32_t>(gpu_resident_images);
        c.plane_format = half_planes ? static_cast<std::uint32_t>(VQM_PLANES_F16) : static_cast<std::uint32_t>(plane_format);
        c.streaming_inputs = streaming_inputs ? 1u : 0u;
        c.strip_rows = strip_rows;
        c.input_bytes = input_bytes;
        c.graph = graph ? 1u : 0u;
        c.chroma_filter = static_cast<std::uint32_t>(chroma_filter);
        c.blur = static_cast<std::uint32_t>(blur);
        c.map_form = static_cast<std::uint32_t>(map_form);
        c.strip_cols = strip_cols;
        return c;
    }
};

namespace detail {

// A pipeline's vqm_ss2_memory as the reference's PipelineMemoryUsage: the
// device memory is the arena (heap 0), the pinned staging and read-back
// buffers the staging (heap 1).
inline PipelineMemoryUsage Usage(const vqm_ss2_memory& m) {
    PipelineMemoryUsage u;
    u.arena = MemoryAllocation{m.total_device_bytes, m.total_device_bytes, kDeviceHeap, kDeviceHeap};
    u.staging = MemoryAllocation{m.pinned_host_bytes, m.pinned_host_bytes, kPinnedHostHeap, kPinnedHostHeap};
    u.uses_staging = m.streaming_bytes > 0;
    return u;
}

}  // namespace detail

class Metrics;

// ssimulacra2.h:66-215. One geometry, one submission in flight, not
// thread-safe; use one pipeline per thread.
class Ssimulacra2Pipeline {
public:
    explicit Ssimulacra2Pipeline(Ssimulacra2Config config) : size_(config.size) {
        const vqm_ss2_config c = config.ToC();
        detail::Check(vqm_ss2_create(&c, &p_));
    }
    ~Ssimulacra2Pipeline() { vqm_ss2_destroy(p_); }
    Ssimulacra2Pipeline(Ssimulacra2Pipeline&& o) noexcept
        : p_(std::exchange(o.p_, nullptr)), size_(o.size_) {}
    Ssimulacra2Pipeline& operator=(Ssimulacra2Pipeline&& o) noexcept {
        if (this != &o) {
            vqm_ss2_destroy(p_);
            p_ = std::exchange(o.p_, nullptr);
            size_ = o.size_;
        }
        return *this;
    }
    Ssimulacra2Pipeline(const Ssimulacra2Pipeline&) = delete;
    Ssimulacra2Pipeline& operator=(const Ssimulacra2Pipeline&) = delete;

    [[nodiscard]] Ssimulacra2Result Compute(const Image& a, const Image& b, float background = 0.5f) {
        return Finish(Submit(a, b, background));
    }
    // Both views are packed straight into the pinned staging (vqm_views.h).
    [[nodiscard]] Ssimulacra2Result Compute(const ImageView& a, const ImageView& b, float background = 0.5f) {
        return Finish(Submit(a, b, background));
    }
    [[nodiscard]] Ssimulacra2Result Compute(const UploadedImage& a, const UploadedImage& b,
                                            float background = 0.5f) {
        return Finish(Submit(a, b, background));
    }
    [[nodiscard]] Ssimulacra2Result Compute(const UploadedImage& a, const Image& b,
                                            float background = 0.5f) {
        return Finish(Submit(a, b, background));
    }
    [[nodiscard]] Ssimulacra2Result Compute(const ImageBuffer& a, const ImageBuffer& b,
                                            float background = 0.5f) {
        return Finish(Submit(a, b, background));
    }
    [[nodiscard]] Ssimulacra2Result Compute(const DeviceImageView& a, const DeviceImageView& b,
                                            float background = 0.5f) {
        return Finish(Submit(a, b, background));
    }

    [[nodiscard]] UploadedImage Upload(const Image& image) {
        const vqm_image_desc d = detail::Desc(image);
        std::uint64_t handle = 0;
        detail::Check(vqm_ss2_upload(p_, &d, image.data.data(), &handle));
        return UploadedImage{handle, image.size};
    }
    void Release(UploadedImage image) { detail::Check(vqm_ss2_release(p_, image.slot)); }

    [[nodiscard]] std::uint64_t Submit(const Image& a, const Image& b, float background = 0.5f) {
        const vqm_image_desc da = detail::Desc(a), db = detail::Desc(b);
        std::uint64_t token = 0;
        detail::Check(vqm_ss2_submit_host(p_, &da, a.data.data(), &db, b.data.data(), background, &token));
        return token;
    }
    [[nodiscard]] std::uint64_t Submit(const ImageView& a, const ImageView& b, float background = 0.5f) {
        const vqm_host_view va = detail::HostView(a), vb = detail::HostView(b);
        std::uint64_t token = 0;
        detail::Check(vqm_ss2_submit_views(p_, &va, &vb, background, &token));
        return token;
    }
    [[nodiscard]] std::uint64_t Submit(const UploadedImage& a, const UploadedImage& b,
                                       float background = 0.5f) {
        std::uint64_t token = 0;
        detail::Check(vqm_ss2_submit_resident(p_, a.slot, b.slot, background, &token));
        return token;
    }
    [[nodiscard]] std::uint64_t Submit(const UploadedImage& a, const Image& b, float background = 0.5f) {
        const vqm_image_desc db = detail::Desc(b);
        std::uint64_t token = 0;
        detail::Check(vqm_ss2_submit_resident_host(p_, a.slot, &db, b.data.data(), background, &token));
        return token;
    }
    [[nodiscard]] std::uint64_t Submit(const ImageBuffer& a, const ImageBuffer& b,
                                       float background = 0.5f) {
        std::uint64_t token = 0;
        detail::Check(vqm_ss2_submit_buffers(p_, a.slot, b.slot, background, &token));
        return token;
    }
    // Optional `wait_stream` (a hipStream_t): the pair starts after the work
    // already queued there (the VulkanSubmitSync wait, ssimulacra2.h:196-201).
    [[nodiscard]] std::uint64_t Submit(const DeviceImageView& a, const DeviceImageView& b,
                                       float background = 0.5f, void* wait_stream = nullptr) {
        const vqm_device_image da = View(a), db = View(b);
        std::uint64_t token = 0;
        detail::Check(vqm_ss2_submit_device(p_, &da, &db, wait_stream, background, &token));
        return token;
    }

    [[nodiscard]] Ssimulacra2Result Finish(std::uint64_t token) {
        std::array<double, 108> norms{};
        std::uint32_t scales = 0;
        detail::Check(vqm_ss2_finish(p_, token, &scales, norms.data()));
        return Ssimulacra2Result::FromNorms(scales, norms.data());
    }

    // Caller-owned device memory carved from input_bytes; MapImage returns a
    // device pointer (fill it with hipMemcpy or a kernel before Submit).
    [[nodiscard]] ImageBuffer AllocateImage(const ImageFormat& format) {
        const vqm_image_format f = format.ToC();
        std::uint64_t handle = 0, bytes = 0;
        detail::Check(vqm_ss2_allocate_image(p_, &f, &handle, &bytes));
        return ImageBuffer{handle, size_, static_cast<std::size_t>(bytes)};
    }
    [[nodiscard]] void* MapImage(ImageBuffer buffer) {
        void* ptr = nullptr;
        detail::Check(vqm_ss2_map_image(p_, buffer.slot, &ptr));
        return ptr;
    }
    void Release(ImageBuffer buffer) { detail::Check(vqm_ss2_release_image(p_, buffer.slot)); }

    // ssimulacra2.h:204-205: the allocations the pipeline holds.
    [[nodiscard]] PipelineMemoryUsage GetMemoryUsage() const {
        vqm_ss2_memory m{};
        detail::Check(vqm_ss2_memory_usage(p_, &m));
        return detail::Usage(m);
    }
    // The same by region (workspace, resident slots, streaming buffers,
    // input pool, pinned host); no reference counterpart.
    [[nodiscard]] vqm_ss2_memory GetMemoryRegions() const {
        vqm_ss2_memory m{};
        detail::Check(vqm_ss2_memory_usage(p_, &m));
        return m;
    }

private:
    friend class Metrics;  // scores tightly packed byte spans without a copy
    [[nodiscard]] std::uint64_t SubmitPacked(const vqm_image_desc& a, const void* da, const vqm_image_desc& b,
                                             const void* db, float background) {
        std::uint64_t token = 0;
        detail::Check(vqm_ss2_submit_host(p_, &a, da, &b, db, background, &token));
        return token;
    }
    vqm_device_image View(const DeviceImageView& v) const {
        return vqm_device_image{v.data, detail::Dim(v.size.width), detail::Dim(v.size.height),
                                v.format.ToC(), RequiredBytes(v.format, v.size)};
    }
    vqm_ss2_pipeline* p_ = nullptr;
    Size size_;
};

// vqm_alpha.h:16-27: opaque inputs score once; with alpha, the worse of the
// dark (0.1) and bright (0.9) background.
template <typename Frame>
[[nodiscard]] Ssimulacra2Result ComputeWorstBackground(Ssimulacra2Pipeline& pipeline,
                                                       const Frame& original, const Frame& distorted) {
    if (original.format.alpha == AlphaType::None && distorted.format.alpha == AlphaType::None)
        return pipeline.Compute(original, distorted);
    auto dark = pipeline.Compute(original, distorted, 0.1f);
    auto bright = pipeline.Compute(original, distorted, 0.9f);
    return dark.Score() <= bright.Score() ? dark : bright;
}

// Score, vqm_views.h:38-46 (image_view.cc:97-105): one pair of views at
// one background; the two-background alpha policy is not applied.
[[nodiscard]] inline double Score(Ssimulacra2Pipeline& pipeline, const ImageView& original,
                                  const ImageView& distorted, float background = 0.5f) {
    if (original.size != distorted.size) throw std::invalid_argument("image views must have the same size");
    return pipeline.Compute(original, distorted, background).Score();
}

// vqm_pool.h:20-33: non-owning input pairs; both images must stay alive
// and unchanged until ComputeMany/ScoreMany returns.
struct ImagePair {
    const Image& original;
    const Image& distorted;
};
struct ImageViewPair {
    const ImageView& original;
    const ImageView& distorted;
};
// Replaces VulkanBufferPair: caller-owned device images.
struct DeviceImagePair {
    const DeviceImageView& original;
    const DeviceImageView& distorted;
};

// vqm_pool.h:35-86 / pool.cc:24-116: `workers` pipelines of one geometry
// driven round-robin from the calling thread, one submission in flight
// each, results in input order. Not thread-safe. If a pair fails, the
// submissions already queued are drained before the error is rethrown; if
// the drain itself fails, the pipelines are destroyed and every later call
// throws std::logic_error (the pool must then be destroyed).
class Ssimulacra2Pool {
public:
    Ssimulacra2Pool(Ssimulacra2Config config, std::size_t workers) : workers_(workers) {
        if (workers == 0) throw std::invalid_argument("pool needs at least one worker");
        pipes_.reserve(workers);
        for (std::size_t i = 0; i < workers; ++i) pipes_.emplace_back(config);
    }
    [[nodiscard]] std::size_t Workers() const noexcept { return workers_; }

    [[nodiscard]] std::vector<Ssimulacra2Result> ComputeMany(std::span<const ImagePair> pairs,
                                                             float background = 0.5f) {
        return Run(pairs, background);
    }
    [[nodiscard]] std::vector<Ssimulacra2Result> ComputeMany(std::span<const ImageViewPair> pairs,
                                                             float background = 0.5f) {
        return Run(pairs, background);
    }
    [[nodiscard]] std::vector<Ssimulacra2Result> ComputeMany(std::span<const DeviceImagePair> pairs,
                                                             float background = 0.5f) {
        return Run(pairs, background);
    }
    [[nodiscard]] std::vector<double> ScoreMany(std::span<const ImagePair> pairs, float background = 0.5f) {
        return Scores(ComputeMany(pairs, background));
    }
    [[nodiscard]] std::vector<double> ScoreMany(std::span<const ImageViewPair> pairs, float background = 0.5f) {
        return Scores(ComputeMany(pairs, background));
    }
    [[nodiscard]] std::vector<double> ScoreMany(std::span<const DeviceImagePair> pairs,
                                                float background = 0.5f) {
        return Scores(ComputeMany(pairs, background));
    }

private:
    static std::vector<double> Scores(const std::vector<Ssimulacra2Result>& results) {
        std::vector<double> scores;
        scores.reserve(results.size());
        for (const auto& r : results) scores.push_back(r.Score());
        return scores;
    }

    template <typename Pair>
    std::vector<Ssimulacra2Result> Run(std::span<const Pair> pairs, float background) {
        if (poisoned_) throw std::logic_error("pool cannot be reused after a failed drain");
        const std::size_t n = pipes_.size();
        std::vector<std::uint64_t> tokens(n, 0);
        std::vector<std::size_t> owner(n, 0);
        std::vector<Ssimulacra2Result> out(pairs.size());
        try {
            for (std::size_t i = 0; i < pairs.size(); ++i) {
                const std::size_t slot = i % n;
                if (tokens[slot] != 0) out[owner[slot]] = pipes_[slot].Finish(std::exchange(tokens[slot], 0));
                tokens[slot] = pipes_[slot].Submit(pairs[i].original, pairs[i].distorted, background);
                owner[slot] = i;
            }
            for (std::size_t slot = 0; slot < n; ++slot)
                if (tokens[slot] != 0) out[owner[slot]] = pipes_[slot].Finish(std::exchange(tokens[slot], 0));
        } catch (...) {
            DrainAfterFailure(tokens);
            throw;
        }
        return out;
    }

    // Finishes every queued submission; when one of those waits fails, the
    // pipelines (whose destructors wait for their device work) are dropped
    // and the pool is poisoned (pool.cc:94-112).
    void DrainAfterFailure(std::vector<std::uint64_t>& tokens) noexcept {
        bool failed = false;
        for (std::size_t slot = 0; slot < tokens.size(); ++slot) {
            if (tokens[slot] == 0) continue;
            try {
                (void)pipes_[slot].Finish(std::exchange(tokens[slot], 0));
            } catch (...) {
                failed = true;
            }
        }
        if (failed) {
            pipes_.clear();
            poisoned_ = true;
        }
    }

    std::size_t workers_ = 0;
    std::vector<Ssimulacra2Pipeline> pipes_;
    bool poisoned_ = false;
};

}  // namespace vqm
