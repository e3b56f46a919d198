// HIP runtime of the SSIMULACRA 2.1 path: device context, device/pinned
// buffers, the single-geometry pipeline (the backend behind
// internal::Ssimulacra2GpuPipeline, ssimulacra2_internal.h:102-152) and the
// multi-geometry batch engine.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <tuple>
#include <utility>
#include <vector>

#include "ss2_common.h"
#include "ss2_host.h"
#include "ss2_kernels.h"
#include "vqm_ss2.h"

namespace vqm::ss2 {

void Check(hipError_t e, const char* what);

// Process-unique 64-bit ids for tokens and handles (handles.h:21-32): never
// zero, never reused.
uint64_t NextId();

class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(uint64_t bytes);
    ~DeviceBuffer();
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(std::exchange(o.ptr_, nullptr)), bytes_(std::exchange(o.bytes_, 0)) {}
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void* get() const { return ptr_; }
    template <typename T> T* as() const { return static_cast<T*>(ptr_); }
    uint64_t bytes() const { return bytes_; }

private:
    void* ptr_ = nullptr;
    uint64_t bytes_ = 0;
};

class PinnedBuffer {
public:
    PinnedBuffer() = default;
    explicit PinnedBuffer(uint64_t bytes);
    ~PinnedBuffer();
    PinnedBuffer(PinnedBuffer&& o) noexcept : ptr_(std::exchange(o.ptr_, nullptr)), bytes_(std::exchange(o.bytes_, 0)) {}
    PinnedBuffer& operator=(PinnedBuffer&& o) noexcept;
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    void* get() const { return ptr_; }
    template <typename T> T* as() const { return static_cast<T*>(ptr_); }
    uint64_t bytes() const { return bytes_; }

private:
    void* ptr_ = nullptr;
    uint64_t bytes_ = 0;
};

// One HIP stream at default priority and its events.
class Stream {
public:
    Stream();
    ~Stream();
    Stream(const Stream&) = delete;
    Stream& operator=(const Stream&) = delete;
    hipStream_t get() const { return stream_; }

private:
    hipStream_t stream_ = nullptr;
};

class Event {
public:
    Event();
    ~Event();
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;
    hipEvent_t get() const { return event_; }

private:
    hipEvent_t event_ = nullptr;
};

// Milliseconds between two recorded events (0 when unavailable).
float Elapsed(const Event& a, const Event& b);

// The one decoded input format: 8-bit full-range sRGB RGB, BT.709, opaque.
vqm_image_format Srgb8Format();

// VQM_PLANES_* to PlaneFormat; INVALID_ARGUMENT otherwise.
PlaneFormat ToFormat(uint32_t plane_format);

// VQM_MAP_* to the kernels' MapForm for planes of `format`: the default is
// the reference form, except for the fixed-point (U16) planes, which hold
// the score bound only in the difference form. INVALID_ARGUMENT otherwise.
MapForm ToMapForm(uint32_t map_form, PlaneFormat format);

// Per-process device state: the current device must be a gfx950; the sRGB
// decode table lives in its memory.
struct DeviceContext {
    int device = 0;
    std::string arch;
    int cus = 0;
    const float* lut = nullptr;
    static DeviceContext& Get();
};

// A geometry's plan on the host and its device copies.
struct DevicePlan {
    GeometryPlan host;
    WorkspaceLayout layout;
    DeviceBuffer plan;   // one Plan
    DeviceBuffer items;  // ScaleItem[items]
    DevicePlan(uint32_t width, uint32_t height, PlaneFormat format, uint32_t strip_rows, uint32_t strip_cols);
};

// One pair's device workspace: planes (zero borders), partials, norms and,
// for pairs whose inputs arrive from the host, room for both input images.
// It owns its memory, or uses `bytes` at `memory` carved from an arena.
struct Workspace {
    DeviceBuffer owned;
    uint8_t* base = nullptr;
    uint64_t bytes = 0;
    void* planes = nullptr;
    double* partials = nullptr;
    double* norms = nullptr;
    uint8_t* inputs = nullptr;  // 2 * input_image_bytes, or null
    Workspace(const WorkspaceLayout& layout, uint64_t input_image_bytes = 0, void* memory = nullptr);
    // Bytes a workspace of this layout needs.
    static uint64_t Bytes(const WorkspaceLayout& layout, uint64_t input_image_bytes);
};

// One device allocation carved into ranges by a RangeAllocator. Not
// thread-safe.
class Arena {
public:
    explicit Arena(uint64_t bytes) : memory_(bytes), ranges_(bytes) {}
    // A range of `bytes`, or null when no free range is large enough.
    void* Alloc(uint64_t bytes) {
        const uint64_t offset = ranges_.Alloc(bytes);
        return offset == RangeAllocator::kNone ? nullptr : memory_.as<uint8_t>() + offset;
    }
    void Free(void* p, uint64_t bytes) {
        ranges_.Free(static_cast<uint64_t>(static_cast<uint8_t*>(p) - memory_.as<uint8_t>()), bytes);
    }

private:
    DeviceBuffer memory_;
    RangeAllocator ranges_;
};

// Per-geometry workspace pools under one device-byte budget. A geometry's
// plan is built on first use and workspaces are reused. With a budget, the
// workspaces are carved from one arena of that size (allocated on first
// use), and idle workspaces of other geometries are handed back to it to
// make room, so admission under pressure never frees or allocates device
// memory. Without a budget each workspace is its own allocation and nothing
// is freed. Not thread-safe.
class WorkspacePools {
public:
    WorkspacePools(PlaneFormat format, uint32_t strip_rows, uint32_t strip_cols, uint64_t budget);
    const DevicePlan& PlanFor(uint32_t width, uint32_t height);
    bool HasPlan(uint32_t width, uint32_t height) const {
        return plans_.count(std::make_pair(width, height)) != 0;
    }
    // The strip height and width of a geometry whose plan does not exist yet.
    void SetStripShape(uint32_t width, uint32_t height, uint32_t strip_rows, uint32_t strip_cols) {
        shape_for_[std::make_pair(width, height)] = std::make_pair(strip_rows, strip_cols);
    }
    bool HasShape(uint32_t width, uint32_t height) const {
        return shape_for_.count(std::make_pair(width, height)) != 0;
    }
    // Device bytes one workspace of this geometry holds.
    uint64_t Need(uint32_t width, uint32_t height, bool inputs);
    // A workspace, or null when it cannot fit the budget until busy
    // workspaces come back.
    std::unique_ptr<Workspace> Acquire(uint32_t width, uint32_t height, bool inputs);
    void Return(uint32_t width, uint32_t height, bool inputs, std::unique_ptr<Workspace> ws);
    uint64_t bytes() const { return bytes_; }
    uint64_t peak() const { return peak_; }
    uint64_t budget() const { return budget_; }

private:
    struct Key {
        uint32_t width, height;
        bool inputs;
        bool operator<(const Key& o) const {
            return std::tie(width, height, inputs) < std::tie(o.width, o.height, o.inputs);
        }
    };
    struct Pool {
        std::vector<std::unique_ptr<Workspace>> idle;
        uint64_t each = 0;
    };
    void Drop(std::unique_ptr<Workspace> ws);
    PlaneFormat format_;
    uint32_t strip_rows_, strip_cols_;
    uint64_t budget_;
    std::unique_ptr<Arena> arena_;
    std::map<std::pair<uint32_t, uint32_t>, std::unique_ptr<DevicePlan>> plans_;
    std::map<std::pair<uint32_t, uint32_t>, std::pair<uint32_t, uint32_t>> shape_for_;
    std::map<Key, Pool> pools_;
    uint64_t bytes_ = 0, peak_ = 0;
};

// One pair of a launch: its plan, workspace, device inputs and, when the
// inputs are on the host, the host images to copy into the workspace first.
struct LaunchEntry {
    const DevicePlan* plan = nullptr;
    Workspace* ws = nullptr;
    const uint8_t* src[2] = {nullptr, nullptr};
    const void* host[2] = {nullptr, nullptr};
    uint64_t image_bytes = 0;
    // Device decode descriptors (null: 8-bit sRGB RGB); `general` when one
    // of them is of kind 1.
    const SourceDesc* source[2] = {nullptr, nullptr};
    bool general = false;
};

// A stream with the descriptor, work-list and norm buffers of one launch of
// the three kernels over up to `max_pairs` pairs of any geometries. A lane
// holds one launch at a time.
class Lane {
public:
    static constexpr uint32_t kMaxChunks = 16;
    // Instantiated graphs a lane keeps (one per launch shape).
    static constexpr size_t kMaxGraphs = 32;
    // Launches of at most this many pairs qualify for graph replay in the
    // automatic mode (where the launch overhead is a visible share).
    static constexpr uint32_t kAutoGraphPairs = 8;
    // graphs: 0 eager; 1 replay every launch that qualifies from captured
    // HIP graphs (see Launch); 2 only those of up to kAutoGraphPairs pairs.
    explicit Lane(uint32_t max_pairs, uint32_t graphs = 0);
    ~Lane();
    // Queues the uploads, the prepare, scale and reduce launches and the norm
    // read-back, then returns. With chunks > 1 the pairs are split into that
    // many runs of consecutive pairs: the prepare launches of the runs follow
    // one another on the lane's first stream, and each run's scale launch
    // follows on a second stream as soon as its prepare is done, so the
    // scale of one run overlaps the prepare of the next.
    // `blur` is the scale kernel's BlurKind, `map` both kernels' MapForm;
    // `persistent` sizes both kernels' grids to the resident occupancy
    // (LaunchPrepare, LaunchScale). With graphs on, a launch of one
    // geometry, one run, 8-bit sRGB device inputs is replayed from a HIP
    // graph captured the first time its shape (plan, pair count, format,
    // map form, blur, persistent) is seen: the descriptor upload from the
    // pinned descriptors written for this launch, the three kernels and the
    // norm read-back, as one submission. It holds no memset node.
    void Launch(PlaneFormat format, MapForm map, const float* lut, const LaunchEntry* entries,
                uint32_t count, uint32_t chunks = 1, uint32_t blur = kBlurSampled,
                bool persistent = false);
    bool Done() const;
    void Wait();
    // Forgets a launch that failed part-way (its streams already synchronized).
    void Abandon() { pending_ = false; }
    void Synchronize();
    // Norms of entry i of the finished launch (pinned host memory).
    const double* Norms(uint32_t i) const { return h_norms_.as<double>() + static_cast<uint64_t>(i) * kNorms; }
    // Milliseconds the prepare, scale and reduce launches of the finished
    // launch ran, each summed over its chunks (chunks may overlap). A
    // replayed graph reports its whole time as the first stage.
    std::array<float, 3> StageTimes() const;
    // Whether the last launch was a graph replay.
    bool replayed() const { return replayed_; }
    hipStream_t stream() const { return stream_.get(); }
    uint32_t max_pairs() const { return max_pairs_; }

private:
    uint32_t max_pairs_;
    DeviceBuffer pairs_, work_prep_, work_scale_, norms_;
    PinnedBuffer h_pairs_, h_work_prep_, h_work_scale_, h_norms_;
    Stream stream_, scale_stream_;
    // Per chunk: prepare start/end, scale start/end; then reduce start/end
    // and the end of the read-back.
    std::vector<std::unique_ptr<Event>> ev_;
    uint32_t chunks_ = 1;
    bool pending_ = false;
    struct GraphKey {
        const DevicePlan* plan;
        uint32_t count, format, map, blur;
        bool persistent;
        bool operator<(const GraphKey& o) const {
            return std::tie(plan, count, format, map, blur, persistent) <
                   std::tie(o.plan, o.count, o.format, o.map, o.blur, o.persistent);
        }
    };
    uint32_t graphs_ = 0;
    bool replayed_ = false;
    std::map<GraphKey, hipGraphExec_t> graph_cache_;
    hipGraphExec_t CaptureGraph(PlaneFormat fmt, MapForm map, const float* lut, const DevicePlan& plan,
                                uint32_t count, uint32_t blur, bool persistent);
    Event& Ev(uint32_t i) { return *ev_[i]; }
    const Event& Ev(uint32_t i) const { return *ev_[i]; }
    uint32_t DoneEvent() const { return 4 * chunks_ + 2; }
};

class Pipeline {
public:
    explicit Pipeline(const vqm_ss2_config& config);
    ~Pipeline();

    uint64_t Upload(const vqm_image_desc& image, const void* data);
    void Release(uint64_t handle);
    // Every submission takes the alpha blend target `background` (encoded
    // RGB; opaque inputs ignore it).
    uint64_t SubmitHost(const vqm_image_desc& a, const void* da, const vqm_image_desc& b,
                        const void* db, float background);
    uint64_t SubmitResident(uint64_t a, uint64_t b, float background);
    // vulkan_runtime_inputs.cc:156-183: both views are checked, then packed
    // straight into the pinned staging.
    uint64_t SubmitViews(const vqm_host_view& a, const vqm_host_view& b, float background);
    uint64_t SubmitResidentHost(uint64_t a, const vqm_image_desc& b, const void* db, float background);
    uint64_t SubmitDevice(const vqm_device_image& a, const vqm_device_image& b,
                          hipStream_t wait_stream, float background);
    uint64_t AllocateImage(const vqm_image_format& format, uint64_t* bytes);
    void* MapImage(uint64_t buffer);
    void ReleaseImage(uint64_t buffer);
    uint64_t SubmitBuffers(uint64_t a, uint64_t b, float background);
    uint32_t Finish(uint64_t token, double* norms);
    vqm_ss2_memory Memory() const;
    std::array<float, 3> StageTimes() const { return stage_ms_; }

private:
    struct Slot {
        uint64_t handle = 0;  // 0 = free
        DeviceBuffer data;    // grows to the largest image it has held
        vqm_image_format format{};
    };
    struct InputBuffer {
        uint64_t handle = 0;
        uint64_t offset = 0;
        uint64_t bytes = 0;
        vqm_image_format format{};
    };
    // Decode descriptors of the pair's images: kind 0 for 8-bit sRGB RGB,
    // else built from the format; uploaded by the next launch when needed.
    void SetSources(const vqm_image_format& a, const vqm_image_format& b, float background);
    // Room for two host images of `bytes` each in the streaming buffers.
    void EnsureStreaming(uint64_t bytes);
    void RequireIdle() const;
    const Slot& RequireSlot(uint64_t handle, const char* name) const;
    const InputBuffer& RequireBuffer(uint64_t handle, const char* name) const;
    // reuse_original: the original's planes are already in the workspace
    // (the resident reference cache), so only the distorted image is prepared.
    uint64_t Launch(const uint8_t* src0, const uint8_t* src1, bool reuse_original = false);
    void BuildGraph();

    vqm_ss2_config config_;
    MapForm map_;
    DeviceContext& ctx_;
    DevicePlan plan_;
    Workspace ws_;
    DeviceBuffer pair_;    // one PairDesc
    PinnedBuffer h_pair_;  // its host copy; the graph uploads it per pair
    hipGraphExec_t graph_ = nullptr;
    bool graph_launched_ = false;
    std::vector<Slot> slots_;
    DeviceBuffer stream_inputs_;   // two images (streaming_inputs)
    PinnedBuffer staging_;         // two images, pinned
    uint64_t stream_image_bytes_ = 0;  // capacity per image of the two above
    DeviceBuffer sources_;         // two SourceDesc, read through PairDesc::source
    PinnedBuffer h_sources_;
    bool general_ = false;         // this pair has an image of kind 1
    bool sources_pending_ = false; // h_sources_ not yet copied to sources_
    PinnedBuffer host_norms_;
    DeviceBuffer input_pool_;      // caller-owned images (input_bytes)
    std::vector<InputBuffer> inputs_;
    uint64_t image_bytes_ = 0;
    Stream stream_;
    Event ev_[5];
    uint64_t in_flight_ = 0;
    bool broken_ = false;  // a wait failed: every later call is refused
    std::array<float, 3> stage_ms_{};
    // Resident reference cache (vulkan.cc:238-241,716-735): the handle of
    // the resident original whose planes the workspace holds, or 0.
    uint64_t cached_original_ = 0;
    float cached_background_ = 0.0f;
    DeviceBuffer distorted_work_;  // prepare work list of the distorted image only
    uint32_t prepared_images_ = 0; // images the last submission prepared

public:
    // Images the last submission prepared: 1 when the resident original was
    // reused from the previous submission, else 2.
    uint32_t PreparedImages() const { return prepared_images_; }
};

class Batch {
public:
    explicit Batch(const vqm_ss2_batch_config& config);
    ~Batch();
    uint64_t Submit(const vqm_ss2_device_pair* pairs, uint32_t count);
    void Finish(uint64_t token, uint32_t* scales, double* norms);
    std::array<float, 3> StageTimes() const { return stage_ms_; }
    uint64_t PoolBytes() const { return pools_.bytes(); }
    // Fixes the strip shape of a geometry not used yet (a shape tuned
    // earlier, or by another process).
    void SetShape(uint32_t width, uint32_t height, uint32_t strip_rows, uint32_t strip_cols);
    uint32_t DefaultStripRows(uint32_t width, uint32_t height, uint32_t pairs) const;

private:
    struct Taken {
        uint32_t width, height;
        std::unique_ptr<Workspace> ws;
    };
    vqm_ss2_batch_config config_;
    MapForm map_;
    DeviceContext& ctx_;
    WorkspacePools pools_;
    Lane lane_;
    std::vector<Taken> busy_;
    std::vector<uint32_t> levels_;
    uint64_t in_flight_ = 0;
    std::array<float, 3> stage_ms_{};
};

}  // namespace vqm::ss2
