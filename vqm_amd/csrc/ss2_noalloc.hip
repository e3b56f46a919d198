// Allocation-free scoring of one pair of caller-owned 8-bit sRGB RGB device
// images (vqm_ss2_score_device_noalloc). Everything the kernels need beyond
// the inputs and the output norms -- the pair descriptor, the plan, the work
// items, the decode table, the partial sums and both images' plane pyramids
// -- lives in one file-scope device array, `scratch`, of
// kNoAllocScratchBytes. No device memory is allocated per call: the
// descriptors travel as kernel arguments and a setup kernel writes them into
// the scratch, together with the work items and the cleared planes. The
// kernels it launches use no private (stack) memory, so the runtime grows
// no scratch of its own for them either; the general per-pixel decode does
// (its filter tables), which is why other input formats are refused.
//
// Every call on a device shares that device's scratch, so calls that overlap
// on one device race; the setup kernel rewrites everything a call reads
// before the scoring kernels run.

#include <hip/hip_runtime.h>

#include <array>
#include <cstring>
#include <mutex>

#include "ss2_host.h"
#include "ss2_kernels.h"
#include "ss2_noalloc.h"
#include "ss2_runtime.h"
#include "ss2_source.h"  // IsSrgb8

namespace vqm::ss2 {
namespace {

// Fixed regions of the scratch (bytes); the work items, the partials and
// the planes follow, each 256-byte aligned.
constexpr uint64_t kPairAt = 0;
constexpr uint64_t kPlanAt = 256;
constexpr uint64_t kLutAt = 1024;
constexpr uint64_t kItemsAt = 2048;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

// Per device, the address of its copy of `scratch` (hipGetSymbolAddress),
// looked up once with that device current.
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

uint64_t Align256(uint64_t v) { return (v + 255) / 256 * 256; }

// Everything the setup kernel writes, passed by value.
struct SetupArgs {
    Plan plan;
    PairDesc pair;
    float lut[256];
};

// Writes the descriptors (block 0), the work items and zeros over both
// images' planes (every block, grid-stride). The planes' borders are the
// blur's zero padding and must read zero; their interiors are overwritten
// by the prepare kernel.
__global__ __launch_bounds__(256) void setup_kernel(SetupArgs args, uint8_t* base, uint64_t items_at,
                                                    uint64_t planes_at, uint64_t plane_words) {
    const uint32_t tid = threadIdx.x;
    if (blockIdx.x == 0) {
        auto copy = [&](const void* src, uint64_t at, uint32_t bytes) {
            const uint32_t* s = static_cast<const uint32_t*>(src);
            uint32_t* d = reinterpret_cast<uint32_t*>(base + at);
            for (uint32_t k = tid; k < bytes / 4; k += blockDim.x) d[k] = s[k];
        };
        copy(&args.pair, kPairAt, sizeof(PairDesc));
        copy(&args.plan, kPlanAt, sizeof(Plan));
        copy(args.lut, kLutAt, sizeof(args.lut));
    }
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    const uint64_t first = static_cast<uint64_t>(blockIdx.x) * blockDim.x + tid;
    ScaleItem* items = reinterpret_cast<ScaleItem*>(base + items_at);
    for (uint64_t i = first; i < args.plan.items; i += stride) items[i] = ItemAt(args.plan, static_cast<uint32_t>(i));
    uint4* planes = reinterpret_cast<uint4*>(base + planes_at);
    for (uint64_t i = first; i < plane_words; i += stride) planes[i] = make_uint4(0u, 0u, 0u, 0u);
}

static_assert(sizeof(PairDesc) <= kPlanAt - kPairAt, "pair descriptor region");
static_assert(sizeof(Plan) <= kLutAt - kPlanAt, "plan region");
static_assert(256 * sizeof(float) <= kItemsAt - kLutAt, "decode table region");

// The calling thread's current device's scratch address.
uint8_t* DeviceScratch() {
    int device = 0;
    Check(hipGetDevice(&device), "hipGetDevice");
    if (device < 0 || device >= kMaxDevices) Fail(VQM_RUNTIME_ERROR, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_scratch[device] == nullptr) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || ptr == nullptr) {
            (void)hipGetLastError();
            Fail(VQM_RUNTIME_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        g_scratch[device] = ptr;
    }
    return static_cast<uint8_t*>(g_scratch[device]);
}

}  // namespace

void InitNoAllocScratch() { (void)DeviceScratch(); }

uint32_t ScoreDeviceNoAlloc(const vqm_ss2_config& config, const vqm_device_image& a,
                            const vqm_device_image& b, hipStream_t stream, double* norms) {
    const PlaneFormat format = ToFormat(config.plane_format);
    const MapForm map = ToMapForm(config.map_form, format);
    if (config.blur > kBlurRecursive) Fail(VQM_INVALID_ARGUMENT, "blur must be sampled (0) or recursive (1)");
    if (config.chroma_filter > 2) Fail(VQM_INVALID_ARGUMENT, "chroma_filter must be bilinear, bicubic or lanczos");
    const vqm_image_desc da{a.width, a.height, a.format, a.bytes};
    const vqm_image_desc db{b.width, b.height, b.format, b.bytes};
    CheckImage(da, config.width, config.height, "original");
    CheckImage(db, config.width, config.height, "distorted");
    if (a.data == nullptr || b.data == nullptr) Fail(VQM_INVALID_ARGUMENT, "device image data is null");
    if (norms == nullptr) Fail(VQM_INVALID_ARGUMENT, "norms is null");
    if (!IsSrgb8(a.format) || !IsSrgb8(b.format))
        Fail(VQM_INVALID_ARGUMENT,
             "the allocation-free entry point takes 8-bit full-range sRGB RGB images only "
             "(score other formats through a pipeline)");

    // The strip height a pipeline of this geometry takes, so both give the
    // same sums.
    SetupArgs args{};
    args.plan = FillPlan(config.width, config.height, format,
                         SinglePairStripRows(config.width, config.height, format, config.strip_rows,
                                             config.strip_cols),
                         config.strip_cols);
    const Plan& p = args.plan;
    const uint64_t partials_at = Align256(kItemsAt + static_cast<uint64_t>(p.items) * sizeof(ScaleItem));
    const uint64_t planes_at = Align256(partials_at + static_cast<uint64_t>(p.items) * 6 * sizeof(double));
    const uint64_t need = planes_at + 2 * p.image_bytes;
    if (need > kNoAllocScratchBytes)
        Fail(VQM_INVALID_ARGUMENT,
             "a " + std::to_string(config.width) + "x" + std::to_string(config.height) +
                 " pair in this plane format needs " + std::to_string(need) +
                 " bytes of scratch; the allocation-free entry point holds " +
                 std::to_string(kNoAllocScratchBytes) + " (use a pipeline for larger pairs)");

    uint8_t* base = DeviceScratch();
    const std::array<float, 256> lut = SrgbTable();
    std::memcpy(args.lut, lut.data(), sizeof(args.lut));
    PairDesc& pd = args.pair;
    pd.src[0] = static_cast<const uint8_t*>(a.data);
    pd.src[1] = static_cast<const uint8_t*>(b.data);
    pd.planes = base + planes_at;
    pd.partials = reinterpret_cast<double*>(base + partials_at);
    pd.norms = norms;
    pd.plan = reinterpret_cast<const Plan*>(base + kPlanAt);
    pd.items = reinterpret_cast<const ScaleItem*>(base + kItemsAt);

    const uint64_t plane_words = 2 * p.image_bytes / 16;  // image_bytes is a multiple of 256
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(plane_words / 256, 1), 2048));
    hipLaunchKernelGGL(setup_kernel, dim3(blocks), dim3(256), 0, stream, args, base, kItemsAt, planes_at,
                       plane_words);
    Check(hipGetLastError(), "launch setup");
    const PairDesc* dpair = reinterpret_cast<const PairDesc*>(base + kPairAt);
    const float* dlut = reinterpret_cast<const float*>(base + kLutAt);
    const uint32_t tiles = p.prep_tiles_x * p.prep_tiles_y;
    Check(LaunchPrepare(format, map, false, dpair, nullptr, 2 * tiles, tiles, dlut, nullptr, nullptr, stream),
          "launch prepare");
    Check(LaunchScale(format, config.blur, map, p.strip_w, dpair, nullptr, p.items, p.items, stream),
          "launch scale");
    Check(LaunchReduce(dpair, 1, stream), "launch reduce");
    return p.levels;
}

}  // namespace vqm::ss2
