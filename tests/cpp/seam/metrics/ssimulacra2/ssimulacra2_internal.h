// A stand-in for the reference's backend seam, for compiling the adapter in
// INTEGRATION.md section 1 on its own: the types the seam's signatures name
// (include/vqm_image.h, vqm_views.h, vqm_vulkan.h, vqm_memory_types.h,
// ssimulacra2.h of GreatValueCreamSoda/vqm) reduced to the fields the
// adapter reads, and internal::Ssimulacra2GpuPipeline with the pure virtual
// functions of src/metrics/ssimulacra2/ssimulacra2_internal.h:102-152. Test
// infrastructure; no product code includes it.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <optional>
#include <span>
#include <vector>

namespace vqm {

struct Size {
    std::size_t width = 0;
    std::size_t height = 0;
};

enum class PixelType : std::uint8_t { Byte = 0, Word = 1, Half = 2, Float = 3 };
enum class PixelRange : std::uint8_t { Limited = 0, Full = 1 };
enum class ColorFamily : std::uint8_t { Grey = 0, Rgb = 1, Yuv = 2 };
enum class AlphaType : std::uint8_t { None = 0, Straight = 1, Premultiplied = 2 };
enum class ChromaLocation : std::uint8_t { Left = 0, Center = 1, TopLeft = 2, Top = 3, BottomLeft = 4, Bottom = 5 };
enum class MatrixCoefficients : std::uint8_t { Rgb = 0, Bt709 = 1 };
enum class TransferCharacteristics : std::uint8_t { Bt709 = 1, Srgb = 13 };
enum class ColorPrimaries : std::uint8_t { Bt709 = 1 };
enum class ChromaFilter : std::uint8_t { Bilinear, Bicubic, Lanczos };
enum class MemoryPolicy : std::uint8_t { Automatic, Direct, Staged };

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
};

struct Image {
    Size size;
    ImageFormat format;
    std::vector<std::uint8_t> data;
};

struct PlaneView {
    const std::byte* data = nullptr;
    std::size_t stride_bytes = 0;
};

struct ImageView {
    Size size;
    ImageFormat format;
    std::array<PlaneView, 4> planes{};
};

struct ImageBuffer {
    std::uint64_t slot = 0;
    Size size;
    std::size_t bytes = 0;
};

struct MemoryAllocation {
    std::uint64_t buffer_bytes = 0;
    std::uint64_t allocation_bytes = 0;
    std::uint32_t memory_type = 0;
    std::uint32_t heap = 0;
};

struct PipelineMemoryUsage {
    bool uses_staging = false;
    MemoryAllocation arena;
    MemoryAllocation staging;
    std::vector<MemoryAllocation> auxiliary = {};
};

struct Ssimulacra2Scale {
    std::array<double, 6> avg_ssim{};
    std::array<double, 12> avg_edgediff{};
};

struct Ssimulacra2Result {
    std::vector<Ssimulacra2Scale> scales;
};

struct Ssimulacra2Config {
    Size size;
    std::size_t gpu_resident_images = 2;
    bool half_planes = false;
    ChromaFilter chroma_filter = ChromaFilter::Bilinear;
    std::size_t input_bytes = 0;
    bool streaming_inputs = true;
    MemoryPolicy memory_policy = MemoryPolicy::Automatic;
};

using VkBuffer = struct VkBuffer_T*;
using VkSemaphore = struct VkSemaphore_T*;

struct VulkanPlaneLayout {
    std::uint64_t offset_bytes = 0;
    std::uint64_t stride_bytes = 0;
};

struct VulkanBufferView {
    VkBuffer buffer = nullptr;
    std::uint64_t offset = 0;
    std::uint64_t bytes = 0;
    Size size;
    ImageFormat format;
    std::array<VulkanPlaneLayout, 4> planes{};
};

struct VulkanSemaphoreWait {
    VkSemaphore semaphore = nullptr;
    std::uint64_t value = 0;
};

struct VulkanSubmitSync {
    std::span<const VulkanSemaphoreWait> waits{};
    std::span<const VkSemaphore> signals{};
};

}  // namespace vqm

namespace vqm::internal {

class Ssimulacra2GpuPipeline {
public:
    virtual ~Ssimulacra2GpuPipeline() = default;
    virtual void RequireIdle() const = 0;
    virtual std::optional<std::uint32_t> Upload(const Image& image, std::size_t capacity) = 0;
    virtual void Release(std::uint32_t slot) = 0;
    [[nodiscard]] virtual std::uint64_t Submit(const Image& original, const Image& distorted, float background) = 0;
    [[nodiscard]] virtual std::uint64_t Submit(const ImageView& original, const ImageView& distorted,
                                               float background) = 0;
    [[nodiscard]] virtual std::uint64_t Submit(std::uint32_t original, std::uint32_t distorted, float background) = 0;
    [[nodiscard]] virtual std::uint64_t Submit(std::uint32_t original, const Image& distorted, float background) = 0;
    [[nodiscard]] virtual Ssimulacra2Result Finish(std::uint64_t token) = 0;
    [[nodiscard]] virtual PipelineMemoryUsage GetMemoryUsage() const = 0;
    [[nodiscard]] virtual ImageBuffer AllocateImage(const ImageFormat& format) = 0;
    [[nodiscard]] virtual std::byte* MapImage(ImageBuffer buffer) = 0;
    virtual void Release(ImageBuffer buffer) = 0;
    [[nodiscard]] virtual std::uint64_t Submit(const ImageBuffer& original, const ImageBuffer& distorted,
                                               float background) = 0;
    [[nodiscard]] virtual std::uint64_t Submit(const VulkanBufferView& original, const VulkanBufferView& distorted,
                                               float background) = 0;
    [[nodiscard]] virtual std::uint64_t Submit(const VulkanBufferView& original, const VulkanBufferView& distorted,
                                               const VulkanSubmitSync& sync, float background) = 0;
};

}  // namespace vqm::internal
