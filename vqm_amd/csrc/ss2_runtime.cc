// HIP runtime of the SSIMULACRA 2.1 path (see ss2_runtime.h).

#include "ss2_runtime.h"

#include "ss2_hostcopy.h"
#include "ss2_source.h"
#include "ss2_tune.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>

namespace vqm::ss2 {

void Check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        Fail(VQM_RUNTIME_ERROR, std::string("hip: ") + what + ": " + hipGetErrorString(e));
}

uint64_t NextId() {
    static std::atomic<uint64_t> counter{0};
    return ++counter;
}

vqm_image_format Srgb8Format() {
    vqm_image_format f{};
    f.type = 0;
    f.family = 1;
    f.alpha = 0;
    f.range = 1;
    f.matrix = 0;
    f.transfer = 13;
    f.primaries = 1;
    return f;
}

PlaneFormat ToFormat(uint32_t f) {
    if (f > 3) Fail(VQM_INVALID_ARGUMENT, "plane_format must be VQM_PLANES_F32, _F16, _BF16 or _U16");
    return static_cast<PlaneFormat>(f);
}

MapForm ToMapForm(uint32_t map_form, PlaneFormat format) {
    switch (map_form) {
        case VQM_MAP_DEFAULT: return format == kPlaneU16 ? kMapDifference : kMapReference;
        case VQM_MAP_REFERENCE: return kMapReference;
        case VQM_MAP_DIFFERENCE: return kMapDifference;
    }
    Fail(VQM_INVALID_ARGUMENT, "map_form must be VQM_MAP_DEFAULT, _REFERENCE or _DIFFERENCE");
}

float Elapsed(const Event& a, const Event& b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a.get(), b.get()) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0f;
    }
    return ms;
}

// ---- buffers --------------------------------------------------------------

DeviceBuffer::DeviceBuffer(uint64_t bytes) : bytes_(bytes) {
    if (bytes == 0) return;
    if (hipMalloc(&ptr_, bytes) != hipSuccess) {
        ptr_ = nullptr;
        (void)hipGetLastError();
        Fail(VQM_RUNTIME_ERROR, "hip: device allocation of " + std::to_string(bytes) + " bytes failed");
    }
}

DeviceBuffer::~DeviceBuffer() {
    if (ptr_) (void)hipFree(ptr_);
}

DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = std::exchange(o.ptr_, nullptr);
        bytes_ = std::exchange(o.bytes_, 0);
    }
    return *this;
}

PinnedBuffer::PinnedBuffer(uint64_t bytes) : bytes_(bytes) {
    if (bytes == 0) return;
    if (hipHostMalloc(&ptr_, bytes, hipHostMallocDefault) != hipSuccess) {
        ptr_ = nullptr;
        (void)hipGetLastError();
        Fail(VQM_RUNTIME_ERROR, "hip: pinned host allocation of " + std::to_string(bytes) + " bytes failed");
    }
}

PinnedBuffer::~PinnedBuffer() {
    if (ptr_) (void)hipHostFree(ptr_);
}

PinnedBuffer& PinnedBuffer::operator=(PinnedBuffer&& o) noexcept {
    if (this != &o) {
        if (ptr_) (void)hipHostFree(ptr_);
        ptr_ = std::exchange(o.ptr_, nullptr);
        bytes_ = std::exchange(o.bytes_, 0);
    }
    return *this;
}

Stream::Stream() { Check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate"); }
Stream::~Stream() {
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
}

Event::Event() { Check(hipEventCreate(&event_), "hipEventCreate"); }
Event::~Event() {
    if (event_) (void)hipEventDestroy(event_);
}

// ---- device context ---------------------------------------------------------

DeviceContext& DeviceContext::Get() {
    static std::mutex mutex;
    static std::unique_ptr<DeviceContext> ctx;
    static DeviceBuffer lut;
    std::lock_guard<std::mutex> lock(mutex);
    if (ctx) return *ctx;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        Fail(VQM_RUNTIME_ERROR, "hip: no usable compute device");
    }
    auto c = std::make_unique<DeviceContext>();
    Check(hipGetDevice(&c->device), "hipGetDevice");
    hipDeviceProp_t prop{};
    Check(hipGetDeviceProperties(&prop, c->device), "hipGetDeviceProperties");
    c->arch = prop.gcnArchName;
    c->cus = prop.multiProcessorCount;
    if (c->arch.rfind("gfx950", 0) != 0)
        Fail(VQM_RUNTIME_ERROR, "hip: this library is built for gfx950 only, device is " + c->arch);
    const std::array<float, 256> table = SrgbTable();
    lut = DeviceBuffer(sizeof(table));
    Check(hipMemcpy(lut.get(), table.data(), sizeof(table), hipMemcpyHostToDevice), "upload decode table");
    c->lut = lut.as<float>();
    ctx = std::move(c);
    return *ctx;
}

DevicePlan::DevicePlan(uint32_t width, uint32_t height, PlaneFormat format, uint32_t strip_rows,
                       uint32_t strip_cols)
    : host(BuildPlan(width, height, format, strip_rows, strip_cols)),
      layout(LayoutWorkspace(host)),
      plan(sizeof(Plan)),
      items(std::max<uint64_t>(host.items.size(), 1) * sizeof(ScaleItem)) {
    Check(hipMemcpy(plan.get(), &host.plan, sizeof(Plan), hipMemcpyHostToDevice), "upload plan");
    if (!host.items.empty())
        Check(hipMemcpy(items.get(), host.items.data(), host.items.size() * sizeof(ScaleItem),
                        hipMemcpyHostToDevice),
              "upload work items");
}

namespace {
constexpr uint64_t kAlign = 256;
uint64_t AlignUp(uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
uint64_t InputArea(uint64_t input_image_bytes) { return 2 * AlignUp(input_image_bytes); }
}  // namespace

uint64_t Workspace::Bytes(const WorkspaceLayout& layout, uint64_t input_image_bytes) {
    return layout.total + InputArea(input_image_bytes);
}

Workspace::Workspace(const WorkspaceLayout& layout, uint64_t input_image_bytes, void* memory)
    : bytes(Bytes(layout, input_image_bytes)) {
    if (memory == nullptr) {
        owned = DeviceBuffer(bytes);
        memory = owned.get();
    }
    base = static_cast<uint8_t*>(memory);
    // The plane borders must read as zero for every later pair: they are the
    // blur's zero padding and nothing writes them.
    Check(hipMemset(base, 0, layout.total), "clear workspace");
    // The clear runs on the null stream; the pipelines' streams do not wait
    // for it, so finish it before any kernel can touch the planes.
    Check(hipStreamSynchronize(nullptr), "clear workspace");
    planes = base;
    partials = reinterpret_cast<double*>(base + layout.planes);
    norms = reinterpret_cast<double*>(base + layout.planes + layout.partials);
    if (input_image_bytes > 0) inputs = base + layout.total;
}

// ---- workspace pools ------------------------------------------------------------

WorkspacePools::WorkspacePools(PlaneFormat format, uint32_t strip_rows, uint32_t strip_cols, uint64_t budget)
    : format_(format), strip_rows_(strip_rows), strip_cols_(strip_cols), budget_(budget) {}

void WorkspacePools::Drop(std::unique_ptr<Workspace> ws) {
    bytes_ -= ws->bytes;
    if (ws->owned.bytes() == 0) arena_->Free(ws->base, ws->bytes);
}

const DevicePlan& WorkspacePools::PlanFor(uint32_t width, uint32_t height) {
    auto& slot = plans_[std::make_pair(width, height)];
    if (!slot) {
        const auto shape = shape_for_.find(std::make_pair(width, height));
        const bool set = shape != shape_for_.end();
        try {
            slot = std::make_unique<DevicePlan>(width, height, format_, set ? shape->second.first : strip_rows_,
                                                set ? shape->second.second : strip_cols_);
        } catch (...) {
            plans_.erase(std::make_pair(width, height));
            throw;
        }
    }
    return *slot;
}

uint64_t WorkspacePools::Need(uint32_t width, uint32_t height, bool inputs) {
    const DevicePlan& plan = PlanFor(width, height);
    uint64_t need = plan.layout.total;
    if (inputs) need += InputArea(RequiredBytes(Srgb8Format(), width, height));
    return need;
}

std::unique_ptr<Workspace> WorkspacePools::Acquire(uint32_t width, uint32_t height, bool inputs) {
    Pool& pool = pools_[Key{width, height, inputs}];
    if (!pool.idle.empty()) {
        std::unique_ptr<Workspace> ws = std::move(pool.idle.back());
        pool.idle.pop_back();
        return ws;
    }
    const DevicePlan& plan = PlanFor(width, height);
    const uint64_t input_bytes = inputs ? RequiredBytes(Srgb8Format(), width, height) : 0;
    const uint64_t need = Workspace::Bytes(plan.layout, input_bytes);
    void* memory = nullptr;
    if (budget_ != 0) {
        if (need > budget_) return nullptr;
        if (!arena_) arena_ = std::make_unique<Arena>(budget_);
        memory = arena_->Alloc(need);
        // Hand idle workspaces of other geometries back, largest first, until
        // a range fits.
        while (memory == nullptr) {
            Pool* victim = nullptr;
            for (auto& [key, p] : pools_)
                if (!p.idle.empty() && (victim == nullptr || p.each > victim->each)) victim = &p;
            if (victim == nullptr) return nullptr;
            Drop(std::move(victim->idle.back()));
            victim->idle.pop_back();
            memory = arena_->Alloc(need);
        }
    }
    std::unique_ptr<Workspace> ws;
    try {
        ws = std::make_unique<Workspace>(plan.layout, input_bytes, memory);
    } catch (...) {
        if (memory) arena_->Free(memory, need);
        throw;
    }
    pool.each = need;
    bytes_ += need;
    peak_ = std::max(peak_, bytes_);
    return ws;
}

void WorkspacePools::Return(uint32_t width, uint32_t height, bool inputs, std::unique_ptr<Workspace> ws) {
    if (ws) pools_[Key{width, height, inputs}].idle.push_back(std::move(ws));
}

// ---- lane -----------------------------------------------------------------------

Lane::Lane(uint32_t max_pairs, uint32_t graphs)
    : max_pairs_(max_pairs),
      pairs_(max_pairs * sizeof(PairDesc)),
      norms_(static_cast<uint64_t>(max_pairs) * kNorms * sizeof(double)),
      h_pairs_(max_pairs * sizeof(PairDesc)),
      h_norms_(static_cast<uint64_t>(max_pairs) * kNorms * sizeof(double)),
      graphs_(graphs) {
    for (uint32_t i = 0; i < 4 * kMaxChunks + 3; ++i) ev_.push_back(std::make_unique<Event>());
}

Lane::~Lane() {
    if (pending_) (void)hipEventSynchronize(Ev(DoneEvent()).get());
    for (auto& [key, exec] : graph_cache_) (void)hipGraphExecDestroy(exec);
}

// Captures one launch shape on the lane's first stream: the upload of the
// pinned pair descriptors (rewritten before every replay: workspaces and
// inputs change), the prepare, scale and reduce kernels over `count` pairs
// of `plan`, and the read-back of their norms.
hipGraphExec_t Lane::CaptureGraph(PlaneFormat fmt, MapForm map, const float* lut, const DevicePlan& plan,
                                  uint32_t count, uint32_t blur, bool persistent) {
    const hipStream_t s = stream_.get();
    const GeometryPlan& g = plan.host;
    const PairDesc* dpairs = pairs_.as<PairDesc>();
    const uint64_t prep = 2ull * g.prep_tiles * count, scale = static_cast<uint64_t>(g.plan.items) * count;
    Check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin graph capture");
    hipError_t e = hipMemcpyAsync(pairs_.get(), h_pairs_.get(), count * sizeof(PairDesc), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = LaunchPrepare(fmt, map, false, dpairs, nullptr, static_cast<uint32_t>(prep), g.prep_tiles, lut, nullptr,
                          nullptr, s, persistent);
    if (e == hipSuccess)
        e = LaunchScale(fmt, blur, map, g.plan.strip_w, dpairs, nullptr, static_cast<uint32_t>(scale), g.plan.items, s,
                        persistent);
    if (e == hipSuccess) e = LaunchReduce(dpairs, count, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h_norms_.get(), norms_.get(), static_cast<uint64_t>(count) * kNorms * sizeof(double),
                           hipMemcpyDeviceToHost, s);
    hipGraph_t graph = nullptr;
    const hipError_t end = hipStreamEndCapture(s, &graph);
    Check(e, "capture a batch launch");
    Check(end, "end graph capture");
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    Check(e, "instantiate a batch launch's graph");
    return exec;
}

void Lane::Launch(PlaneFormat fmt, MapForm map, const float* lut, const LaunchEntry* entries,
                  uint32_t count, uint32_t chunks, uint32_t blur, bool persistent) {
    if (pending_) Fail(VQM_LOGIC_ERROR, "lane: the previous launch has not finished");
    if (count == 0 || count > max_pairs_) Fail(VQM_INVALID_ARGUMENT, "lane: bad pair count");
    chunks = std::max(1u, std::min({chunks, count, kMaxChunks}));
    auto* hp = h_pairs_.as<PairDesc>();
    bool uniform = true, general = false;
    uint64_t prep_work = 0, scale_work = 0;
    std::vector<uint64_t> prep_at(count + 1, 0), scale_at(count + 1, 0);  // work-list offsets
    for (uint32_t i = 0; i < count; ++i) {
        const LaunchEntry& e = entries[i];
        PairDesc& pd = hp[i];
        for (int k = 0; k < 2; ++k)
            pd.src[k] = e.host[k] ? e.ws->inputs + k * ((e.image_bytes + 255) / 256 * 256) : e.src[k];
        pd.planes = e.ws->planes;
        pd.partials = e.ws->partials;
        pd.norms = norms_.as<double>() + static_cast<uint64_t>(i) * kNorms;
        pd.plan = e.plan->plan.as<Plan>();
        pd.items = e.plan->items.as<ScaleItem>();
        pd.source[0] = e.source[0];
        pd.source[1] = e.source[1];
        general = general || e.general;
        uniform = uniform && e.plan == entries[0].plan;
        prep_work += 2ull * e.plan->host.prep_tiles;
        scale_work += e.plan->host.plan.items;
        prep_at[i + 1] = prep_work;
        scale_at[i + 1] = scale_work;
    }
    if (prep_work > 0x7FFFFFFFull || scale_work > 0x7FFFFFFFull)
        Fail(VQM_INVALID_ARGUMENT, "the submission has too many work items");

    const hipStream_t s = stream_.get();
    bool host_inputs = false;
    for (uint32_t i = 0; i < count; ++i) host_inputs = host_inputs || entries[i].host[0] || entries[i].host[1];
    replayed_ = (graphs_ == 1 || (graphs_ == 2 && count <= kAutoGraphPairs)) && uniform && chunks == 1 &&
                !general && !host_inputs;
    if (replayed_) {
        const GraphKey key{entries[0].plan, count, static_cast<uint32_t>(fmt), static_cast<uint32_t>(map), blur,
                           persistent};
        auto it = graph_cache_.find(key);
        if (it == graph_cache_.end()) {
            if (graph_cache_.size() >= kMaxGraphs) {
                for (auto& [k, exec] : graph_cache_) (void)hipGraphExecDestroy(exec);
                graph_cache_.clear();
            }
            it = graph_cache_.emplace(key, CaptureGraph(fmt, map, lut, *entries[0].plan, count, blur, persistent)).first;
        }
        chunks_ = 1;
        Check(hipEventRecord(Ev(0).get(), s), "hipEventRecord");
        Check(hipGraphLaunch(it->second, s), "launch a batch graph");
        Check(hipEventRecord(Ev(DoneEvent()).get(), s), "hipEventRecord");
        pending_ = true;
        return;
    }
    const hipStream_t t = chunks > 1 ? scale_stream_.get() : s;
    for (uint32_t i = 0; i < count; ++i)
        for (int k = 0; k < 2; ++k)
            if (entries[i].host[k])
                Check(hipMemcpyAsync(const_cast<uint8_t*>(hp[i].src[k]), entries[i].host[k],
                                     entries[i].image_bytes, hipMemcpyHostToDevice, s),
                      "upload input image");
    Check(hipMemcpyAsync(pairs_.get(), hp, count * sizeof(PairDesc), hipMemcpyHostToDevice, s),
          "upload pair descriptors");
    const WorkRef* wprep = nullptr;
    const WorkRef* wscale = nullptr;
    auto chunk_begin = [&](uint32_t c) { return static_cast<uint32_t>(uint64_t(c) * count / chunks); };
    constexpr uint32_t kStripWidths[2] = {64, 128};
    std::vector<uint64_t> scale_seg(2 * chunks + 1, 0);  // work-list runs by chunk and strip width
    if (!uniform) {
        if (h_work_prep_.bytes() < prep_work * sizeof(WorkRef)) {
            h_work_prep_ = PinnedBuffer(prep_work * sizeof(WorkRef));
            work_prep_ = DeviceBuffer(prep_work * sizeof(WorkRef));
        }
        if (h_work_scale_.bytes() < scale_work * sizeof(WorkRef)) {
            h_work_scale_ = PinnedBuffer(scale_work * sizeof(WorkRef));
            work_scale_ = DeviceBuffer(scale_work * sizeof(WorkRef));
        }
        auto* wp = h_work_prep_.as<WorkRef>();
        auto* ws = h_work_scale_.as<WorkRef>();
        uint64_t np = 0, ns = 0;
        for (uint32_t i = 0; i < count; ++i) {
            const GeometryPlan& g = entries[i].plan->host;
            for (uint32_t image = 0; image < 2; ++image)
                for (uint32_t k = 0; k < g.prep_tiles; ++k) wp[np++] = WorkRef{i, k | (image << 31)};
        }
        // Scale work, per chunk, in one run per strip width (each run is one
        // launch of that width's kernel).
        for (uint32_t c = 0; c < chunks; ++c) {
            for (uint32_t g = 0; g < 2; ++g) {
                scale_seg[2 * c + g] = ns;
                for (uint32_t i = chunk_begin(c); i < chunk_begin(c + 1); ++i) {
                    const Plan& p = entries[i].plan->host.plan;
                    if (p.strip_w != kStripWidths[g]) continue;
                    for (uint32_t k = 0; k < p.items; ++k) ws[ns++] = WorkRef{i, k};
                }
            }
        }
        scale_seg[2 * chunks] = ns;
        Check(hipMemcpyAsync(work_prep_.get(), wp, np * sizeof(WorkRef), hipMemcpyHostToDevice, s),
              "upload prepare work");
        Check(hipMemcpyAsync(work_scale_.get(), ws, ns * sizeof(WorkRef), hipMemcpyHostToDevice, s),
              "upload scale work");
        wprep = work_prep_.as<WorkRef>();
        wscale = work_scale_.as<WorkRef>();
    }
    const GeometryPlan& g0 = entries[0].plan->host;
    const PairDesc* dpairs = pairs_.as<PairDesc>();
    chunks_ = chunks;
    for (uint32_t c = 0; c < chunks; ++c) {
        const uint32_t i0 = chunk_begin(c), i1 = chunk_begin(c + 1);
        // Uniform runs index pairs from their first descriptor; work lists
        // hold absolute pair indices and are cut at the run's offsets.
        const PairDesc* base = uniform ? dpairs + i0 : dpairs;
        const WorkRef* wp = uniform ? nullptr : wprep + prep_at[i0];
        const auto np = static_cast<uint32_t>(prep_at[i1] - prep_at[i0]);
        Check(hipEventRecord(Ev(4 * c).get(), s), "hipEventRecord");
        Check(LaunchPrepare(fmt, map, general, base, wp, np, g0.prep_tiles, lut, nullptr, nullptr, s, persistent),
              "launch prepare");
        Check(hipEventRecord(Ev(4 * c + 1).get(), s), "hipEventRecord");
    }
    for (uint32_t c = 0; c < chunks; ++c) {
        const uint32_t i0 = chunk_begin(c), i1 = chunk_begin(c + 1);
        if (t != s) Check(hipStreamWaitEvent(t, Ev(4 * c + 1).get(), 0), "order scale after prepare");
        Check(hipEventRecord(Ev(4 * c + 2).get(), t), "hipEventRecord");
        if (uniform) {
            const auto ns = static_cast<uint32_t>(scale_at[i1] - scale_at[i0]);
            Check(LaunchScale(fmt, blur, map, g0.plan.strip_w, dpairs + i0, nullptr, ns, g0.plan.items, t,
                              persistent),
                  "launch scale");
        } else {
            for (uint32_t g = 0; g < 2; ++g) {
                const auto ns = static_cast<uint32_t>(scale_seg[2 * c + g + 1] - scale_seg[2 * c + g]);
                if (ns == 0) continue;
                Check(LaunchScale(fmt, blur, map, kStripWidths[g], dpairs, wscale + scale_seg[2 * c + g], ns,
                                  g0.plan.items, t, persistent),
                      "launch scale");
            }
        }
        Check(hipEventRecord(Ev(4 * c + 3).get(), t), "hipEventRecord");
    }
    const uint32_t r = 4 * chunks;
    Check(hipEventRecord(Ev(r).get(), t), "hipEventRecord");
    Check(LaunchReduce(dpairs, count, t), "launch reduce");
    Check(hipEventRecord(Ev(r + 1).get(), t), "hipEventRecord");
    Check(hipMemcpyAsync(h_norms_.get(), norms_.get(), static_cast<uint64_t>(count) * kNorms * sizeof(double),
                         hipMemcpyDeviceToHost, t),
          "read back norms");
    Check(hipEventRecord(Ev(r + 2).get(), t), "hipEventRecord");
    pending_ = true;
}

bool Lane::Done() const {
    if (!pending_) return true;
    const hipError_t e = hipEventQuery(Ev(DoneEvent()).get());
    if (e == hipErrorNotReady) return false;
    Check(e, "query lane");
    return true;
}

void Lane::Wait() {
    if (!pending_) return;
    pending_ = false;
    Check(hipEventSynchronize(Ev(DoneEvent()).get()), "wait for launch");
}

void Lane::Synchronize() {
    (void)hipStreamSynchronize(stream_.get());
    (void)hipStreamSynchronize(scale_stream_.get());
}

std::array<float, 3> Lane::StageTimes() const {
    std::array<float, 3> t{};
    if (replayed_) {
        t[0] = Elapsed(Ev(0), Ev(DoneEvent()));
        return t;
    }
    for (uint32_t c = 0; c < chunks_; ++c) {
        t[0] += Elapsed(Ev(4 * c), Ev(4 * c + 1));
        t[1] += Elapsed(Ev(4 * c + 2), Ev(4 * c + 3));
    }
    t[2] = Elapsed(Ev(4 * chunks_), Ev(4 * chunks_ + 1));
    return t;
}

// ---- pipeline -----------------------------------------------------------------

namespace {
constexpr uint32_t kMaxResident = 4096;
vqm_image_format Srgb8() { return Srgb8Format(); }
}  // namespace

// Geometry and option checks run before any device work, so an impossible
// size names the minimum rather than failing on the device
// (pipeline_test.cc:318-328).
static const vqm_ss2_config& Validated(const vqm_ss2_config& config) {
    PlanScales(config.width, config.height);
    ToFormat(config.plane_format);
    if (config.resident_images > kMaxResident)
        Fail(VQM_INVALID_ARGUMENT, "gpu_resident_images is limited to 4096 slots");
    if (config.chroma_filter > 2)
        Fail(VQM_INVALID_ARGUMENT, "chroma_filter must be bilinear, bicubic or lanczos");
    if (config.blur > kBlurRecursive) Fail(VQM_INVALID_ARGUMENT, "blur must be sampled (0) or recursive (1)");
    ToMapForm(config.map_form, static_cast<PlaneFormat>(config.plane_format));
    return config;
}

Pipeline::Pipeline(const vqm_ss2_config& config)
    : config_(Validated(config)),
      map_(ToMapForm(config.map_form, static_cast<PlaneFormat>(config.plane_format))),
      ctx_(DeviceContext::Get()),
      plan_(config.width, config.height, static_cast<PlaneFormat>(config.plane_format),
            SinglePairStripRows(config.width, config.height, static_cast<PlaneFormat>(config.plane_format),
                                config.strip_rows, config.strip_cols),
            config.strip_cols),
      ws_(plan_.layout),
      pair_(sizeof(PairDesc)),
      sources_(2 * sizeof(SourceDesc)),
      h_sources_(2 * sizeof(SourceDesc)) {
    const vqm_image_format srgb = Srgb8();
    image_bytes_ = RequiredBytes(srgb, config.width, config.height);
    std::memset(h_sources_.get(), 0, 2 * sizeof(SourceDesc));
    Check(hipMemcpy(sources_.get(), h_sources_.get(), 2 * sizeof(SourceDesc), hipMemcpyHostToDevice),
          "upload source descriptors");
    PairDesc pd{};
    pd.planes = ws_.planes;
    pd.partials = ws_.partials;
    pd.norms = ws_.norms;
    pd.plan = plan_.plan.as<Plan>();
    pd.items = plan_.items.as<ScaleItem>();
    pd.source[0] = sources_.as<SourceDesc>();
    pd.source[1] = sources_.as<SourceDesc>() + 1;
    Check(hipMemcpy(pair_.get(), &pd, sizeof(pd), hipMemcpyHostToDevice), "upload pair descriptor");
    if (config.graph) {
        h_pair_ = PinnedBuffer(sizeof(PairDesc));
        std::memcpy(h_pair_.get(), &pd, sizeof(pd));
    }
    // Slots and streaming buffers start at the size of an 8-bit RGB image and
    // grow to the largest image they are handed.
    slots_.resize(config.resident_images);
    for (Slot& s : slots_) s.data = DeviceBuffer(image_bytes_);
    if (config.streaming_inputs) EnsureStreaming(image_bytes_);
    host_norms_ = PinnedBuffer(kNorms * sizeof(double));
    if (config.input_bytes > 0) input_pool_ = DeviceBuffer(config.input_bytes);
    if (config.resident_images > 0) {
        // The prepare work of the distorted image alone, for pairs whose
        // resident original is still in the workspace.
        std::vector<WorkRef> work(plan_.host.prep_tiles);
        for (uint32_t t = 0; t < plan_.host.prep_tiles; ++t) work[t] = WorkRef{0, t | (1u << 31)};
        distorted_work_ = DeviceBuffer(work.size() * sizeof(WorkRef));
        Check(hipMemcpy(distorted_work_.get(), work.data(), work.size() * sizeof(WorkRef),
                        hipMemcpyHostToDevice),
              "upload distorted work list");
    }
}

Pipeline::~Pipeline() {
    if (in_flight_ != 0) (void)hipEventSynchronize(ev_[4].get());
    if (graph_) (void)hipGraphExecDestroy(graph_);
}

void Pipeline::EnsureStreaming(uint64_t bytes) {
    if (bytes <= stream_image_bytes_) return;
    const uint64_t each = (bytes + 255) / 256 * 256;
    stream_inputs_ = DeviceBuffer(2 * each);
    staging_ = PinnedBuffer(2 * each);
    stream_image_bytes_ = each;
}

void Pipeline::SetSources(const vqm_image_format& a, const vqm_image_format& b, float background) {
    auto* h = h_sources_.as<SourceDesc>();
    const vqm_image_format* f[2] = {&a, &b};
    bool general = false;
    for (int k = 0; k < 2; ++k) {
        if (IsSrgb8(*f[k])) {
            h[k] = SourceDesc{};
        } else {
            h[k] = BuildSourceDesc(*f[k], config_.width, config_.height, config_.chroma_filter);
            h[k].background = background;
            general = true;
        }
    }
    general_ = general;
    sources_pending_ = general;
}

// Captures, on the pipeline's stream, the upload of the pinned pair
// descriptor (its input pointers change per pair), the three kernels (the
// prepare kernel reads its inputs through the descriptor) and the norm
// read-back. The graph holds no memset node: the plane borders are cleared
// once, when the workspace is created. Graphs cover 8-bit sRGB RGB pairs.
void Pipeline::BuildGraph() {
    const hipStream_t s = stream_.get();
    const PairDesc* pd = pair_.as<PairDesc>();
    const PlaneFormat fmt = static_cast<PlaneFormat>(config_.plane_format);
    const GeometryPlan& g = plan_.host;
    Check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin graph capture");
    hipError_t e = hipMemcpyAsync(pair_.get(), h_pair_.get(), sizeof(PairDesc), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = LaunchPrepare(fmt, map_, false, pd, nullptr, 2 * g.prep_tiles, g.prep_tiles, ctx_.lut, nullptr, nullptr, s);
    if (e == hipSuccess)
        e = LaunchScale(fmt, config_.blur, map_, g.plan.strip_w, pd, nullptr, g.plan.items, g.plan.items, s);
    if (e == hipSuccess) e = LaunchReduce(pd, 1, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(host_norms_.get(), ws_.norms, kNorms * sizeof(double), hipMemcpyDeviceToHost, s);
    hipGraph_t graph = nullptr;
    const hipError_t end = hipStreamEndCapture(s, &graph);
    Check(e, "capture the pair's launches");
    Check(end, "end graph capture");
    e = hipGraphInstantiate(&graph_, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) graph_ = nullptr;
    Check(e, "instantiate the pair's graph");
}

void Pipeline::RequireIdle() const {
    if (broken_)
        Fail(VQM_RUNTIME_ERROR,
             "a submission of this pipeline failed on the device; its buffers are in an unknown state, "
             "so the pipeline must be destroyed");
    if (in_flight_ != 0)
        Fail(VQM_LOGIC_ERROR, "a pair is in flight on this pipeline: call Finish first");
}

const Pipeline::Slot& Pipeline::RequireSlot(uint64_t handle, const char* name) const {
    if (handle != 0)
        for (const Slot& s : slots_)
            if (s.handle == handle) return s;
    Fail(VQM_INVALID_ARGUMENT, std::string(name) + ": stale or unknown resident image handle");
}

const Pipeline::InputBuffer& Pipeline::RequireBuffer(uint64_t handle, const char* name) const {
    if (handle != 0)
        for (const InputBuffer& b : inputs_)
            if (b.handle == handle) return b;
    Fail(VQM_INVALID_ARGUMENT, std::string(name) + ": stale or unknown image buffer");
}

uint64_t Pipeline::Upload(const vqm_image_desc& image, const void* data) {
    RequireIdle();
    CheckImage(image, config_.width, config_.height, "image");
    if (data == nullptr) Fail(VQM_INVALID_ARGUMENT, "image: data is null");
    for (Slot& s : slots_) {
        if (s.handle != 0) continue;
        if (s.data.bytes() < image.bytes) s.data = DeviceBuffer(image.bytes);
        Check(hipMemcpy(s.data.get(), data, image.bytes, hipMemcpyHostToDevice), "upload image");
        s.format = image.format;
        s.handle = NextId();
        return s.handle;
    }
    Fail(VQM_INVALID_ARGUMENT, "every resident image slot is in use");
}

void Pipeline::Release(uint64_t handle) {
    RequireIdle();
    if (handle != 0 && handle == cached_original_) cached_original_ = 0;
    for (Slot& s : slots_)
        if (handle != 0 && s.handle == handle) s.handle = 0;
}

uint64_t Pipeline::Launch(const uint8_t* src0, const uint8_t* src1, bool reuse_original) {
    const hipStream_t s = stream_.get();
    const PairDesc* pd = pair_.as<PairDesc>();
    const PlaneFormat fmt = static_cast<PlaneFormat>(config_.plane_format);
    const GeometryPlan& g = plan_.host;
    prepared_images_ = reuse_original ? 1 : 2;
    if (config_.graph && !reuse_original && !general_) {
        if (graph_ == nullptr) BuildGraph();
        auto* hp = h_pair_.as<PairDesc>();
        hp->src[0] = src0;
        hp->src[1] = src1;
        Check(hipEventRecord(ev_[0].get(), s), "hipEventRecord");
        Check(hipGraphLaunch(graph_, s), "launch the pair's graph");
        Check(hipEventRecord(ev_[4].get(), s), "hipEventRecord");
        graph_launched_ = true;
        in_flight_ = NextId();
        return in_flight_;
    }
    graph_launched_ = false;
    if (sources_pending_) {
        Check(hipMemcpyAsync(sources_.get(), h_sources_.get(), 2 * sizeof(SourceDesc), hipMemcpyHostToDevice, s),
              "upload source descriptors");
        sources_pending_ = false;
    }
    Check(hipEventRecord(ev_[0].get(), s), "hipEventRecord");
    if (reuse_original)
        Check(LaunchPrepare(fmt, map_, general_, pd, distorted_work_.as<WorkRef>(), g.prep_tiles, g.prep_tiles,
                            ctx_.lut, src0, src1, s),
              "launch prepare");
    else
        Check(LaunchPrepare(fmt, map_, general_, pd, nullptr, 2 * g.prep_tiles, g.prep_tiles, ctx_.lut, src0, src1,
                            s),
              "launch prepare");
    Check(hipEventRecord(ev_[1].get(), s), "hipEventRecord");
    Check(LaunchScale(fmt, config_.blur, map_, g.plan.strip_w, pd, nullptr, g.plan.items, g.plan.items, s),
          "launch scale");
    Check(hipEventRecord(ev_[2].get(), s), "hipEventRecord");
    Check(LaunchReduce(pd, 1, s), "launch reduce");
    Check(hipEventRecord(ev_[3].get(), s), "hipEventRecord");
    Check(hipMemcpyAsync(host_norms_.get(), ws_.norms, kNorms * sizeof(double),
                         hipMemcpyDeviceToHost, s),
          "read back norms");
    Check(hipEventRecord(ev_[4].get(), s), "hipEventRecord");
    in_flight_ = NextId();
    return in_flight_;
}

uint64_t Pipeline::SubmitHost(const vqm_image_desc& a, const void* da, const vqm_image_desc& b,
                              const void* db, float background) {
    RequireIdle();
    if (!config_.streaming_inputs)
        Fail(VQM_INVALID_ARGUMENT,
             "streaming_inputs is false, so the pixel overloads are unavailable; use Upload() "
             "and Compute() with resident handles");
    CheckImage(a, config_.width, config_.height, "original");
    CheckImage(b, config_.width, config_.height, "distorted");
    if (da == nullptr || db == nullptr) Fail(VQM_INVALID_ARGUMENT, "image data is null");
    EnsureStreaming(std::max(a.bytes, b.bytes));
    // Each image goes up as soon as it is staged, so the original's upload
    // overlaps the distorted image's copy; the copies use several threads.
    auto* staging = staging_.as<uint8_t>();
    auto* inputs = stream_inputs_.as<uint8_t>();
    HostCopy(staging, da, a.bytes);
    Check(hipMemcpyAsync(inputs, staging, a.bytes, hipMemcpyHostToDevice, stream_.get()), "upload original");
    HostCopy(staging + stream_image_bytes_, db, b.bytes);
    Check(hipMemcpyAsync(inputs + stream_image_bytes_, staging + stream_image_bytes_, b.bytes,
                         hipMemcpyHostToDevice, stream_.get()),
          "upload distorted");
    const auto* base = stream_inputs_.as<uint8_t>();
    SetSources(a.format, b.format, background);
    cached_original_ = 0;  // the workspace now holds a host image's planes
    return Launch(base, base + stream_image_bytes_);
}

uint64_t Pipeline::SubmitViews(const vqm_host_view& a, const vqm_host_view& b, float background) {
    RequireIdle();
    if (!config_.streaming_inputs)
        Fail(VQM_INVALID_ARGUMENT, "streaming_inputs is false, so the pixel overloads are unavailable");
    if (a.width != config_.width || a.height != config_.height || b.width != config_.width ||
        b.height != config_.height)
        Fail(VQM_INVALID_ARGUMENT, "image view geometry does not match the pipeline");
    // Both sides are validated before either is written to the staging.
    const uint64_t bytes_a = CheckView(a), bytes_b = CheckView(b);
    EnsureStreaming(std::max(bytes_a, bytes_b));
    auto* staging = staging_.as<uint8_t>();
    PackView(a, staging);
    PackView(b, staging + stream_image_bytes_);
    Check(hipMemcpyAsync(stream_inputs_.get(), staging, 2 * stream_image_bytes_, hipMemcpyHostToDevice,
                         stream_.get()),
          "upload pair");
    const auto* base = stream_inputs_.as<uint8_t>();
    SetSources(a.format, b.format, background);
    cached_original_ = 0;
    return Launch(base, base + stream_image_bytes_);
}

uint64_t Pipeline::SubmitResident(uint64_t a, uint64_t b, float background) {
    RequireIdle();
    const Slot& sa = RequireSlot(a, "original");
    const Slot& sb = RequireSlot(b, "distorted");
    // vulkan.cc:716-735: the original's planes are reused when it is the
    // previous submission's resident original with the same background.
    const bool reuse = a == cached_original_ && background == cached_background_;
    cached_original_ = 0;
    SetSources(sa.format, sb.format, background);
    const uint64_t token = Launch(sa.data.as<uint8_t>(), sb.data.as<uint8_t>(), reuse);
    cached_original_ = a;
    cached_background_ = background;
    return token;
}

uint64_t Pipeline::SubmitResidentHost(uint64_t a, const vqm_image_desc& b, const void* db, float background) {
    RequireIdle();
    if (!config_.streaming_inputs)
        Fail(VQM_INVALID_ARGUMENT, "streaming_inputs is false, so the pixel overloads are unavailable");
    const Slot& sa = RequireSlot(a, "original");
    CheckImage(b, config_.width, config_.height, "distorted");
    if (db == nullptr) Fail(VQM_INVALID_ARGUMENT, "image data is null");
    EnsureStreaming(b.bytes);
    auto* staging = staging_.as<uint8_t>() + stream_image_bytes_;
    HostCopy(staging, db, b.bytes);
    auto* dst = stream_inputs_.as<uint8_t>() + stream_image_bytes_;
    Check(hipMemcpyAsync(dst, staging, b.bytes, hipMemcpyHostToDevice, stream_.get()), "upload distorted");
    const bool reuse = a == cached_original_ && background == cached_background_;
    cached_original_ = 0;
    SetSources(sa.format, b.format, background);
    const uint64_t token = Launch(sa.data.as<uint8_t>(), dst, reuse);
    cached_original_ = a;
    cached_background_ = background;
    return token;
}

uint64_t Pipeline::SubmitDevice(const vqm_device_image& a, const vqm_device_image& b,
                                hipStream_t wait_stream, float background) {
    RequireIdle();
    const vqm_image_desc da{a.width, a.height, a.format, a.bytes};
    const vqm_image_desc db{b.width, b.height, b.format, b.bytes};
    CheckImage(da, config_.width, config_.height, "original");
    CheckImage(db, config_.width, config_.height, "distorted");
    if (a.data == nullptr || b.data == nullptr) Fail(VQM_INVALID_ARGUMENT, "device image data is null");
    if (wait_stream != nullptr) {
        Event ready;
        Check(hipEventRecord(ready.get(), wait_stream), "record caller stream");
        Check(hipStreamWaitEvent(stream_.get(), ready.get(), 0), "wait caller stream");
    }
    SetSources(a.format, b.format, background);
    cached_original_ = 0;  // caller-owned memory may change between submissions
    return Launch(static_cast<const uint8_t*>(a.data), static_cast<const uint8_t*>(b.data));
}

uint64_t Pipeline::AllocateImage(const vqm_image_format& format, uint64_t* bytes) {
    RequireIdle();
    if (input_pool_.bytes() == 0)
        Fail(VQM_INVALID_ARGUMENT, "no caller-owned image buffers: Ssimulacra2Config::input_bytes is 0");
    ValidateFormat(format, config_.width, config_.height, "image buffer");
    const uint64_t need = RequiredBytes(format, config_.width, config_.height);
    // First fit over the pool, 256-byte aligned, in address order.
    std::vector<InputBuffer> sorted = inputs_;
    std::sort(sorted.begin(), sorted.end(),
              [](const InputBuffer& x, const InputBuffer& y) { return x.offset < y.offset; });
    uint64_t cursor = 0;
    for (const InputBuffer& b : sorted) {
        if (cursor + need <= b.offset) break;
        cursor = (b.offset + b.bytes + 255) / 256 * 256;
    }
    if (cursor + need > input_pool_.bytes())
        Fail(VQM_INVALID_ARGUMENT, "the caller-owned image pool is full");
    InputBuffer ib{NextId(), cursor, need, format};
    inputs_.push_back(ib);
    if (bytes) *bytes = need;
    return ib.handle;
}

void* Pipeline::MapImage(uint64_t buffer) {
    RequireIdle();
    return input_pool_.as<uint8_t>() + RequireBuffer(buffer, "image buffer").offset;
}

void Pipeline::ReleaseImage(uint64_t buffer) {
    RequireIdle();
    inputs_.erase(std::remove_if(inputs_.begin(), inputs_.end(),
                                 [&](const InputBuffer& b) { return buffer != 0 && b.handle == buffer; }),
                  inputs_.end());
}

uint64_t Pipeline::SubmitBuffers(uint64_t a, uint64_t b, float background) {
    RequireIdle();
    const InputBuffer& ba = RequireBuffer(a, "original");
    const InputBuffer& bb = RequireBuffer(b, "distorted");
    const auto* base = input_pool_.as<uint8_t>();
    SetSources(ba.format, bb.format, background);
    cached_original_ = 0;
    return Launch(base + ba.offset, base + bb.offset);
}

uint32_t Pipeline::Finish(uint64_t token, double* norms) {
    if (token == 0 || token != in_flight_)
        Fail(VQM_INVALID_ARGUMENT, "unknown or already finished submission token");
    const hipError_t waited = hipEventSynchronize(ev_[4].get());
    in_flight_ = 0;
    if (waited != hipSuccess) {
        // The pair's work may still touch the staging and the workspace:
        // no later submission may reuse them.
        broken_ = true;
        Check(waited, "wait for pair");
    }
    if (graph_launched_)
        stage_ms_ = {Elapsed(ev_[0], ev_[4]), 0.0f, 0.0f};
    else
        stage_ms_ = {Elapsed(ev_[0], ev_[1]), Elapsed(ev_[1], ev_[2]), Elapsed(ev_[2], ev_[3])};
    std::memcpy(norms, host_norms_.get(), kNorms * sizeof(double));
    return plan_.host.plan.levels;
}

vqm_ss2_memory Pipeline::Memory() const {
    vqm_ss2_memory m{};
    m.workspace_bytes = ws_.bytes + pair_.bytes() + sources_.bytes() + plan_.plan.bytes() + plan_.items.bytes();
    for (const Slot& s : slots_) m.resident_bytes += s.data.bytes();
    m.streaming_bytes = stream_inputs_.bytes();
    m.input_pool_bytes = input_pool_.bytes();
    m.pinned_host_bytes = staging_.bytes() + host_norms_.bytes();
    m.total_device_bytes = m.workspace_bytes + m.resident_bytes + m.streaming_bytes + m.input_pool_bytes;
    return m;
}

// ---- batch --------------------------------------------------------------------

namespace {
const vqm_ss2_batch_config& ValidatedBatch(const vqm_ss2_batch_config& config) {
    ToFormat(config.plane_format);
    if (config.max_pairs == 0) Fail(VQM_INVALID_ARGUMENT, "max_pairs must be at least 1");
    if (config.chunks > Lane::kMaxChunks) Fail(VQM_INVALID_ARGUMENT, "chunks must be at most 16");
    if (config.blur > kBlurRecursive) Fail(VQM_INVALID_ARGUMENT, "blur must be sampled (0) or recursive (1)");
    if (config.persistent > 1) Fail(VQM_INVALID_ARGUMENT, "persistent must be 0 or 1");
    if (config.graph > 2) Fail(VQM_INVALID_ARGUMENT, "graph must be 0, 1 or 2");
    ToMapForm(config.map_form, static_cast<PlaneFormat>(config.plane_format));
    // Validates strip_rows and strip_cols.
    BuildPlan(8, 8, static_cast<PlaneFormat>(config.plane_format), config.strip_rows, config.strip_cols);
    return config;
}
}  // namespace

// A batch's default strip height for a geometry, chosen when the geometry
// is first submitted with `pairs` pairs of it (DefaultStripRows). Measured on
// 300 1080p pairs, 192 rows run the scale launch 2.5% faster than 128 (as
// fast as 256, with less padding); one 1080p pair takes 6.5k pairs/s at 64
// rows and 3.6k at 192.
uint32_t Batch::DefaultStripRows(uint32_t width, uint32_t height, uint32_t pairs) const {
    return ss2::DefaultStripRows(width, height, static_cast<PlaneFormat>(config_.plane_format), config_.strip_cols,
                                 pairs, static_cast<uint32_t>(std::max(ctx_.cus, 1)));
}

Batch::Batch(const vqm_ss2_batch_config& config)
    : config_(ValidatedBatch(config)),
      map_(ToMapForm(config.map_form, static_cast<PlaneFormat>(config.plane_format))),
      ctx_(DeviceContext::Get()),
      pools_(static_cast<PlaneFormat>(config.plane_format), config.strip_rows != 0 ? config.strip_rows : 64,
             config.strip_cols, config.hbm_budget),
      lane_(config.max_pairs, config.graph) {}

Batch::~Batch() = default;

void Batch::SetShape(uint32_t width, uint32_t height, uint32_t strip_rows, uint32_t strip_cols) {
    BuildPlan(width, height, static_cast<PlaneFormat>(config_.plane_format), strip_rows, strip_cols);  // validates
    if (pools_.HasPlan(width, height))
        Fail(VQM_LOGIC_ERROR, "the strip shape of a geometry already used cannot change");
    pools_.SetStripShape(width, height,
                         strip_rows != 0 ? strip_rows : (config_.strip_rows != 0 ? config_.strip_rows : 128),
                         strip_cols != 0 ? strip_cols : config_.strip_cols);
}

uint64_t Batch::Submit(const vqm_ss2_device_pair* pairs, uint32_t count) {
    if (in_flight_ != 0) Fail(VQM_LOGIC_ERROR, "a batch is in flight: call Finish first");
    if (count == 0 || count > config_.max_pairs)
        Fail(VQM_INVALID_ARGUMENT, "pair count must be between 1 and max_pairs");
    // Plans first: every geometry is validated before anything is acquired.
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> per_geometry;
    for (uint32_t i = 0; i < count; ++i) {
        if (pairs[i].original == nullptr || pairs[i].distorted == nullptr)
            Fail(VQM_INVALID_ARGUMENT, "pair " + std::to_string(i) + ": image data is null");
        PlanScales(pairs[i].width, pairs[i].height);
        ++per_geometry[std::make_pair(pairs[i].width, pairs[i].height)];
    }
    uint32_t chunks = config_.chunks;
    if (config_.autotune) {
        // Each new geometry's strip shape comes from the tuner, timed for
        // this batch's count of it, unless SetShape pinned one; the chunk
        // count follows the geometry with the most pairs.
        uint32_t most = 0;
        for (const auto& [geometry, n] : per_geometry) {
            const auto [w, h] = geometry;
            const TuneChoice c = Tuner::Get().Choose(w, h, static_cast<PlaneFormat>(config_.plane_format), map_, n);
            if (!pools_.HasPlan(w, h) && !pools_.HasShape(w, h)) pools_.SetStripShape(w, h, c.strip_rows, c.strip_cols);
            if (n > most) {
                most = n;
                chunks = c.chunks;
            }
        }
    }
    for (const auto& [geometry, n] : per_geometry) {
        const auto [w, h] = geometry;
        if (config_.strip_rows == 0 && !pools_.HasPlan(w, h) && !pools_.HasShape(w, h))
            pools_.SetStripShape(w, h, DefaultStripRows(w, h, n), config_.strip_cols);
        pools_.PlanFor(w, h);
    }
    std::vector<Taken> taken;
    std::vector<LaunchEntry> entries(count);
    try {
        for (uint32_t i = 0; i < count; ++i) {
            auto ws = pools_.Acquire(pairs[i].width, pairs[i].height, false);
            if (!ws) Fail(VQM_INVALID_ARGUMENT, "the submission does not fit the HBM budget");
            LaunchEntry& e = entries[i];
            e.plan = &pools_.PlanFor(pairs[i].width, pairs[i].height);
            e.ws = ws.get();
            e.src[0] = static_cast<const uint8_t*>(pairs[i].original);
            e.src[1] = static_cast<const uint8_t*>(pairs[i].distorted);
            taken.push_back(Taken{pairs[i].width, pairs[i].height, std::move(ws)});
        }
        lane_.Launch(static_cast<PlaneFormat>(config_.plane_format), map_, ctx_.lut, entries.data(), count,
                     chunks, config_.blur, config_.persistent != 0);
    } catch (...) {
        for (Taken& t : taken) pools_.Return(t.width, t.height, false, std::move(t.ws));
        throw;
    }
    levels_.resize(count);
    for (uint32_t i = 0; i < count; ++i) levels_[i] = entries[i].plan->host.plan.levels;
    busy_ = std::move(taken);
    in_flight_ = NextId();
    return in_flight_;
}

void Batch::Finish(uint64_t token, uint32_t* scales, double* norms) {
    if (token == 0 || token != in_flight_)
        Fail(VQM_INVALID_ARGUMENT, "unknown or already finished submission token");
    in_flight_ = 0;
    auto give_back = [&] {
        for (Taken& t : busy_) pools_.Return(t.width, t.height, false, std::move(t.ws));
        busy_.clear();
    };
    try {
        lane_.Wait();
    } catch (...) {
        // The launch's work may still touch its workspaces: whatever was
        // queued ends before they go back to the pools.
        lane_.Synchronize();
        lane_.Abandon();
        give_back();
        throw;
    }
    give_back();
    stage_ms_ = lane_.StageTimes();
    const auto count = static_cast<uint32_t>(levels_.size());
    if (norms) std::memcpy(norms, lane_.Norms(0), static_cast<uint64_t>(count) * kNorms * sizeof(double));
    if (scales)
        for (uint32_t i = 0; i < count; ++i) scales[i] = levels_[i];
}

}  // namespace vqm::ss2
