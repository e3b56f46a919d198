// Host-side logic of the SSIMULACRA 2.1 path: scale plan, taps, pruning,
// score mapping, input validation, decode table and the per-geometry work
// plan the kernels consume. No GPU calls here.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "ss2_common.h"
#include "vqm_ss2.h"

namespace vqm::ss2 {

// Thrown by host logic; the C ABI maps each to its status code.
struct Error {
    int status;
    std::string message;
};

[[noreturn]] void Fail(int status, std::string message);

struct ScalePlan {
    uint32_t count = 0;
    std::array<uint32_t, kMaxScales> w{}, h{};
};

// pipeline.cc:28-48.
ScalePlan PlanScales(uint32_t width, uint32_t height);

// pipeline.cc:274-291: 13 FP32 taps of exp(-0.5 (k/1.5)^2) normalised in FP64.
std::array<float, kTaps> GaussianTaps();

// score.cc:103-114.
bool ChannelPruned(uint32_t channel, uint32_t scale, uint32_t scales);

// score.cc:118-163. Throws Error{VQM_INVALID_ARGUMENT} on more than six
// scales, nonfinite norms or an overflowing mapping.
double Score(uint32_t scales, const double* norms);

// The u8 sRGB decode table the prepare kernel reads: FP32 decode scale
// (input.cc:714-757) and the IEC 61966-2-1 EOTF (color.glsl:135-137).
std::array<float, 256> SrgbTable();

// vqm_image.h:122-133 and input.cc:114-145,244-255: packed planes, each
// starting on a 4-byte boundary.
uint64_t RequiredBytes(const vqm_image_format& format, uint32_t width, uint32_t height);

// input.cc:152-240: the image matches the geometry, its format is one the
// reference accepts, and its byte length is RequiredBytes. Throws
// Error{VQM_INVALID_ARGUMENT}.
void CheckImage(const vqm_image_desc& image, uint32_t width, uint32_t height,
                const char* name);

// image_view.cc:16-45 (DescribeImageView): a non-empty size, a format the
// reference accepts, a pointer and a stride of at least one row for every
// present plane, and a last row whose end is addressable. Returns the
// packed byte length. Throws Error{VQM_INVALID_ARGUMENT} or
// Error{VQM_OVERFLOW}.
uint64_t CheckView(const vqm_host_view& view);

// image_view.cc:47-77 (PackImageView): copies a checked view into the
// packed layout at `dst` (CheckView's length), zeroing the at most three
// bytes that pad each plane to a 4-byte boundary.
void PackView(const vqm_host_view& view, uint8_t* dst);

// Per-geometry plan: level sizes, padded plane geometry, plane offsets per
// image, the scale work items and their per-(level, channel) ranges.
struct GeometryPlan {
    Plan plan{};
    std::vector<ScaleItem> items;
    uint64_t plane_bytes_per_image = 0;  // padded plane storage
    PlaneFormat format = kPlaneF32;
    uint32_t prep_tiles = 0;             // prepare blocks per image
};

// strip_rows: output rows per scale work item, a multiple of 8 up to 1024
// (0 = 64; launches pick theirs with DefaultStripRows); strip_cols: its
// width, 64 or 128 columns (0 = 128).
GeometryPlan BuildPlan(uint32_t width, uint32_t height, PlaneFormat format,
                       uint32_t strip_rows, uint32_t strip_cols = 0);

// The plan alone, without the item table (ItemAt enumerates the items):
// nothing is allocated.
Plan FillPlan(uint32_t width, uint32_t height, PlaneFormat format, uint32_t strip_rows,
              uint32_t strip_cols = 0);

// The default strip height of a scale launch over `pairs` pairs of this
// geometry on a device of `cus` compute units: 192 rows when their work
// items fill every wave slot of the device (8 per CU) four times over, else
// 128 when theirs fill every slot once, else the tallest of 64 and 32 rows
// whose items give each SIMD one (4 per CU), else 16. Taller strips blur
// each strip's 12 halo rows over more output rows; shorter ones keep a
// narrow launch (one pair, a small batch) wide.
uint32_t DefaultStripRows(uint32_t width, uint32_t height, PlaneFormat format, uint32_t strip_cols,
                          uint64_t pairs, uint32_t cus);

// The strip height of a single-pair pipeline (and of the allocation-free
// entry point): `strip_rows` when set, else DefaultStripRows for one pair on
// a whole MI355X (256 CUs), a function of the geometry alone so that the
// memory plan, which needs no device, predicts it (16 rows at 1080p, 64 at
// 4K).
uint32_t SinglePairStripRows(uint32_t width, uint32_t height, PlaneFormat format, uint32_t strip_rows,
                             uint32_t strip_cols);

// Bytes of one pair's workspace: planes of both images, partials, norms,
// each rounded to 256 bytes.
struct WorkspaceLayout {
    uint64_t planes = 0, partials = 0, norms = 0, total = 0;
};
WorkspaceLayout LayoutWorkspace(const GeometryPlan& plan);

// Offset bookkeeping of a device arena: ranges of a fixed capacity handed
// out in 256-byte multiples by first fit over an address-ordered free list,
// neighbours merged on release.
class RangeAllocator {
public:
    static constexpr uint64_t kAlign = 256;
    explicit RangeAllocator(uint64_t capacity);
    // Offset of a new range of `bytes`, or ~0 when no free range fits.
    uint64_t Alloc(uint64_t bytes);
    void Free(uint64_t offset, uint64_t bytes);
    uint64_t capacity() const { return capacity_; }
    uint64_t free_bytes() const;
    size_t free_ranges() const { return free_.size(); }
    static constexpr uint64_t kNone = ~0ull;

private:
    uint64_t capacity_;
    std::map<uint64_t, uint64_t> free_;  // offset -> length
};

// Algorithmic HBM bytes of one pair: both u8 inputs read once, every stored
// plane element written once and read once per image.
uint64_t AlgorithmicBytes(const GeometryPlan& plan);

}  // namespace vqm::ss2
