// C ABI of libvqm_hip.so (include/vqm_ss2.h): status codes, thread-local
// error messages, and the opaque handles over the C++ runtime.

#include <sys/sysinfo.h>

#include <algorithm>

#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "ss2_host.h"
#include "ss2_kernels.h"
#include "ss2_noalloc.h"
#include "ss2_runtime.h"
#include "ss2_server.h"
#include "ss2_source.h"
#include "ss2_tune.h"
#include "vqm_ss2.h"

struct vqm_ss2_pipeline {
    vqm::ss2::Pipeline impl;
    explicit vqm_ss2_pipeline(const vqm_ss2_config& c) : impl(c) {}
};

struct vqm_ss2_batch {
    vqm::ss2::Batch impl;
    explicit vqm_ss2_batch(const vqm_ss2_batch_config& c) : impl(c) {}
};

struct vqm_ss2_server {
    vqm::ss2::Server impl;
    explicit vqm_ss2_server(const vqm_ss2_server_config& c) : impl(c) {}
};

namespace {

thread_local std::string g_last_error;

int SetError(int status, const std::string& message) {
    g_last_error = message;
    return status;
}

template <typename F>
int Guard(F&& body) {
    try {
        body();
        g_last_error.clear();
        return VQM_OK;
    } catch (const vqm::ss2::Error& e) {
        return SetError(e.status, e.message);
    } catch (const std::bad_alloc&) {
        return SetError(VQM_RUNTIME_ERROR, "host allocation failed");
    } catch (const std::exception& e) {
        return SetError(VQM_RUNTIME_ERROR, e.what());
    } catch (...) {
        return SetError(VQM_RUNTIME_ERROR, "unknown error");
    }
}

void Need(const void* p, const char* what) {
    if (p == nullptr) vqm::ss2::Fail(VQM_INVALID_ARGUMENT, std::string(what) + " is null");
}

}  // namespace

extern "C" {

const char* vqm_last_error(void) { return g_last_error.c_str(); }

void vqm_ss2_config_init(vqm_ss2_config* config, uint32_t width, uint32_t height) {
    if (config == nullptr) return;
    std::memset(config, 0, sizeof(*config));
    config->width = width;
    config->height = height;
    config->resident_images = 2;
    config->plane_format = VQM_PLANES_F32;
    config->streaming_inputs = 1;
    config->graph = 0;
}

int vqm_memory_heaps(vqm_memory_heap heaps[2]) {
    if (heaps == nullptr) return SetError(VQM_INVALID_ARGUMENT, "heaps is null");
    std::memset(heaps, 0, 2 * sizeof(vqm_memory_heap));
    int count = 0;
    size_t free_bytes = 0, total_bytes = 0;
    if (hipGetDeviceCount(&count) == hipSuccess && count > 0 &&
        hipMemGetInfo(&free_bytes, &total_bytes) == hipSuccess) {
        heaps[0].size_bytes = total_bytes;
        heaps[0].device_local = 1;
        heaps[0].budget_available = 1;
        heaps[0].budget_bytes = total_bytes;
        heaps[0].usage_bytes = total_bytes - std::min(free_bytes, total_bytes);
    } else {
        (void)hipGetLastError();
        heaps[0].device_local = 1;
    }
    struct sysinfo si {};
    if (sysinfo(&si) == 0) {
        const uint64_t unit = si.mem_unit ? si.mem_unit : 1;
        const uint64_t total = static_cast<uint64_t>(si.totalram) * unit;
        const uint64_t avail = static_cast<uint64_t>(si.freeram + si.bufferram) * unit;
        heaps[1].size_bytes = total;
        heaps[1].budget_available = 1;
        heaps[1].budget_bytes = total;
        heaps[1].usage_bytes = total - std::min(avail, total);
    }
    g_last_error.clear();
    return VQM_OK;
}

void vqm_image_format_srgb8(vqm_image_format* f) {
    if (f == nullptr) return;
    std::memset(f, 0, sizeof(*f));
    f->type = 0;
    f->family = 1;
    f->alpha = 0;
    f->range = 1;
    f->matrix = 0;
    f->transfer = 13;
    f->primaries = 1;
}

int vqm_required_bytes(const vqm_image_format* format, uint32_t width, uint32_t height,
                       uint64_t* bytes) {
    return Guard([&] {
        Need(format, "format");
        Need(bytes, "bytes");
        *bytes = vqm::ss2::RequiredBytes(*format, width, height);
    });
}

int vqm_copy_image(const vqm_host_view* view, void* out, uint64_t* bytes) {
    return Guard([&] {
        Need(view, "view");
        Need(bytes, "bytes");
        const uint64_t need = vqm::ss2::CheckView(*view);
        if (out != nullptr && *bytes != need)
            vqm::ss2::Fail(VQM_INVALID_ARGUMENT, "image view: the output buffer must hold exactly " +
                                                     std::to_string(need) + " bytes");
        if (out != nullptr) vqm::ss2::PackView(*view, static_cast<uint8_t*>(out));
        *bytes = need;
    });
}

int vqm_validate_image_format(const vqm_image_format* format, uint32_t width, uint32_t height) {
    return Guard([&] {
        Need(format, "format");
        vqm::ss2::ValidateFormat(*format, width, height, "image");
    });
}

int vqm_plane_scale_offset(const vqm_image_format* format, int chroma, int alpha, float* scale,
                           float* offset) {
    return Guard([&] {
        Need(format, "format");
        Need(scale, "scale");
        Need(offset, "offset");
        const vqm::ss2::ScaleOffset so = vqm::ss2::PlaneScaleOffset(*format, chroma != 0, alpha != 0);
        *scale = so.scale;
        *offset = so.offset;
    });
}

int vqm_ss2_plan_scales(uint32_t width, uint32_t height, uint32_t sizes[12], uint32_t* count) {
    return Guard([&] {
        Need(sizes, "sizes");
        Need(count, "count");
        const vqm::ss2::ScalePlan p = vqm::ss2::PlanScales(width, height);
        *count = p.count;
        for (uint32_t i = 0; i < p.count; ++i) {
            sizes[2 * i] = p.w[i];
            sizes[2 * i + 1] = p.h[i];
        }
    });
}

void vqm_ss2_gaussian_taps(float taps[13]) {
    if (taps == nullptr) return;
    const auto t = vqm::ss2::GaussianTaps();
    std::memcpy(taps, t.data(), sizeof(float) * 13);
}

int vqm_ss2_blur_taps(uint32_t blur, float taps[13]) {
    return Guard([&] {
        if (taps == nullptr) vqm::ss2::Fail(VQM_INVALID_ARGUMENT, "taps is null");
        if (blur > vqm::ss2::kBlurRecursive)
            vqm::ss2::Fail(VQM_INVALID_ARGUMENT, "blur must be sampled (0) or recursive (1)");
        const float* w = blur == vqm::ss2::kBlurRecursive ? vqm::ss2::CosineTaps::w : vqm::ss2::SampledTaps::w;
        std::memcpy(taps, w, sizeof(float) * 13);
    });
}

int vqm_ss2_channel_pruned(uint32_t channel, uint32_t scale, uint32_t scales) {
    return vqm::ss2::ChannelPruned(channel, scale, scales) ? 1 : 0;
}

int vqm_ss2_score(uint32_t scales, const double norms[108], double* score) {
    return Guard([&] {
        Need(score, "score");
        if (scales > 0) Need(norms, "norms");
        *score = vqm::ss2::Score(scales, norms);
    });
}

void vqm_ss2_srgb_table(float table[256]) {
    if (table == nullptr) return;
    const auto t = vqm::ss2::SrgbTable();
    std::memcpy(table, t.data(), sizeof(float) * 256);
}

int vqm_ss2_memory_plan(const vqm_ss2_config* config, vqm_ss2_memory* plan) {
    return Guard([&] {
        Need(config, "config");
        Need(plan, "plan");
        if (config->plane_format > 3) vqm::ss2::Fail(VQM_INVALID_ARGUMENT, "invalid plane_format");
        const auto fmt = static_cast<vqm::ss2::PlaneFormat>(config->plane_format);
        const auto g = vqm::ss2::BuildPlan(
            config->width, config->height, fmt,
            vqm::ss2::SinglePairStripRows(config->width, config->height, fmt, config->strip_rows, config->strip_cols),
            config->strip_cols);
        const auto layout = vqm::ss2::LayoutWorkspace(g);
        vqm_image_format f;
        vqm_image_format_srgb8(&f);
        const uint64_t image = vqm::ss2::RequiredBytes(f, config->width, config->height);
        std::memset(plan, 0, sizeof(*plan));
        plan->workspace_bytes = layout.total + sizeof(vqm::ss2::PairDesc) + 2 * sizeof(vqm::ss2::SourceDesc) +
                                sizeof(vqm::ss2::Plan) + g.items.size() * sizeof(vqm::ss2::ScaleItem);
        // Slots and streaming buffers as created, for 8-bit RGB images; they
        // grow when a larger format arrives.
        const uint64_t staged = (image + 255) / 256 * 256;
        plan->resident_bytes = static_cast<uint64_t>(config->resident_images) * image;
        plan->streaming_bytes = config->streaming_inputs ? 2 * staged : 0;
        plan->input_pool_bytes = config->input_bytes;
        plan->pinned_host_bytes = (config->streaming_inputs ? 2 * staged : 0) + 108 * sizeof(double);
        plan->total_device_bytes = plan->workspace_bytes + plan->resident_bytes +
                                   plan->streaming_bytes + plan->input_pool_bytes;
    });
}

int vqm_ss2_algorithmic_bytes(uint32_t width, uint32_t height, uint32_t plane_format,
                              uint64_t* bytes) {
    return Guard([&] {
        Need(bytes, "bytes");
        if (plane_format > 3) vqm::ss2::Fail(VQM_INVALID_ARGUMENT, "invalid plane_format");
        const auto g = vqm::ss2::BuildPlan(width, height,
                                           static_cast<vqm::ss2::PlaneFormat>(plane_format), 0);
        *bytes = vqm::ss2::AlgorithmicBytes(g);
    });
}

int vqm_ss2_create(const vqm_ss2_config* config, vqm_ss2_pipeline** out) {
    return Guard([&] {
        Need(config, "config");
        Need(out, "out");
        *out = nullptr;
        *out = new vqm_ss2_pipeline(*config);
    });
}

void vqm_ss2_destroy(vqm_ss2_pipeline* pipeline) { delete pipeline; }

int vqm_ss2_upload(vqm_ss2_pipeline* p, const vqm_image_desc* image, const void* data,
                   uint64_t* handle) {
    return Guard([&] {
        Need(p, "pipeline");
        Need(image, "image");
        Need(handle, "handle");
        *handle = p->impl.Upload(*image, data);
    });
}

int vqm_ss2_release(vqm_ss2_pipeline* p, uint64_t handle) {
    return Guard([&] {
        Need(p, "pipeline");
        p->impl.Release(handle);
    });
}

int vqm_ss2_submit_host(vqm_ss2_pipeline* p, const vqm_image_desc* a, const void* da,
                        const vqm_image_desc* b, const void* db, float background,
                        uint64_t* token) {
    return Guard([&] {
        Need(p, "pipeline");
        Need(a, "original");
        Need(b, "distorted");
        Need(token, "token");
        *token = p->impl.SubmitHost(*a, da, *b, db, background);
    });
}

int vqm_ss2_submit_views(vqm_ss2_pipeline* p, const vqm_host_view* a, const vqm_host_view* b,
                         float background, uint64_t* token) {
    return Guard([&] {
        Need(p, "pipeline");
        Need(a, "original");
        Need(b, "distorted");
        Need(token, "token");
        *token = p->impl.SubmitViews(*a, *b, background);
    });
}

int vqm_ss2_submit_resident(vqm_ss2_pipeline* p, uint64_t a, uint64_t b, float background,
                            uint64_t* token) {
    return Guard([&] {
        Need(p, "pipeline");
        Need(token, "token");
        *token = p->impl.SubmitResident(a, b, background);
    });
}

int vqm_ss2_submit_resident_host(vqm_ss2_pipeline* p, uint64_t a, const vqm_image_desc* b,
                                 const void* db, float background, uint64_t* token) {
    return Guard([&] {
        Need(p, "pipeline");
        Need(b, "distorted");
        Need(token, "token");
        *token = p->impl.SubmitResidentHost(a, *b, db, background);
    });
}

int vqm_ss2_submit_device(vqm_ss2_pipeline* p, const vqm_device_image* a,
                          const vqm_device_image* b, void* wait_stream, float background,
                          uint64_t* token) {
    return Guard([&] {
        Need(p, "pipeline");
        Need(a, "original");
        Need(b, "distorted");
        Need(token, "token");
        *token = p->impl.SubmitDevice(*a, *b, static_cast<hipStream_t>(wait_stream), background);
    });
}

int vqm_ss2_score_device_noalloc(const vqm_ss2_config* config, const vqm_device_image* original,
                                 const vqm_device_image* distorted, void* stream, double* norms_device,
                                 uint32_t* scales) {
    return Guard([&] {
        Need(config, "config");
        Need(original, "original");
        Need(distorted, "distorted");
        Need(norms_device, "norms");
        const uint32_t n = vqm::ss2::ScoreDeviceNoAlloc(*config, *original, *distorted,
                                                        static_cast<hipStream_t>(stream), norms_device);
        if (scales) *scales = n;
    });
}

int vqm_ss2_finish(vqm_ss2_pipeline* p, uint64_t token, uint32_t* scales, double norms[108]) {
    return Guard([&] {
        Need(p, "pipeline");
        Need(norms, "norms");
        const uint32_t n = p->impl.Finish(token, norms);
        if (scales) *scales = n;
    });
}

int vqm_ss2_memory_usage(const vqm_ss2_pipeline* p, vqm_ss2_memory* usage) {
    return Guard([&] {
        Need(p, "pipeline");
        Need(usage, "usage");
        *usage = p->impl.Memory();
    });
}

int vqm_ss2_allocate_image(vqm_ss2_pipeline* p, const vqm_image_format* format, uint64_t* buffer,
                           uint64_t* bytes) {
    return Guard([&] {
        Need(p, "pipeline");
        Need(format, "format");
        Need(buffer, "buffer");
        *buffer = p->impl.AllocateImage(*format, bytes);
    });
}

int vqm_ss2_map_image(vqm_ss2_pipeline* p, uint64_t buffer, void** device_ptr) {
    return Guard([&] {
        Need(p, "pipeline");
        Need(device_ptr, "device_ptr");
        *device_ptr = p->impl.MapImage(buffer);
    });
}

int vqm_ss2_release_image(vqm_ss2_pipeline* p, uint64_t buffer) {
    return Guard([&] {
        Need(p, "pipeline");
        p->impl.ReleaseImage(buffer);
    });
}

int vqm_ss2_submit_buffers(vqm_ss2_pipeline* p, uint64_t a, uint64_t b, float background,
                           uint64_t* token) {
    return Guard([&] {
        Need(p, "pipeline");
        Need(token, "token");
        *token = p->impl.SubmitBuffers(a, b, background);
    });
}

int vqm_ss2_stage_times(const vqm_ss2_pipeline* p, float ms[3]) {
    return Guard([&] {
        Need(p, "pipeline");
        Need(ms, "ms");
        const auto t = p->impl.StageTimes();
        for (int i = 0; i < 3; ++i) ms[i] = t[i];
    });
}

int vqm_ss2_prepared_images(const vqm_ss2_pipeline* p, uint32_t* images) {
    return Guard([&] {
        Need(p, "pipeline");
        Need(images, "images");
        *images = p->impl.PreparedImages();
    });
}

int vqm_ss2_batch_create(const vqm_ss2_batch_config* config, vqm_ss2_batch** out) {
    return Guard([&] {
        Need(config, "config");
        Need(out, "out");
        *out = nullptr;
        *out = new vqm_ss2_batch(*config);
    });
}

void vqm_ss2_batch_destroy(vqm_ss2_batch* batch) { delete batch; }

int vqm_ss2_batch_submit(vqm_ss2_batch* batch, const vqm_ss2_device_pair* pairs, uint32_t count,
                         uint64_t* token) {
    return Guard([&] {
        Need(batch, "batch");
        Need(pairs, "pairs");
        Need(token, "token");
        *token = batch->impl.Submit(pairs, count);
    });
}

int vqm_ss2_batch_finish(vqm_ss2_batch* batch, uint64_t token, uint32_t* scales, double* norms) {
    return Guard([&] {
        Need(batch, "batch");
        batch->impl.Finish(token, scales, norms);
    });
}

int vqm_ss2_batch_stage_times(const vqm_ss2_batch* batch, float ms[3]) {
    return Guard([&] {
        Need(batch, "batch");
        Need(ms, "ms");
        const auto t = batch->impl.StageTimes();
        for (int i = 0; i < 3; ++i) ms[i] = t[i];
    });
}

int vqm_ss2_batch_pool_bytes(const vqm_ss2_batch* batch, uint64_t* bytes) {
    return Guard([&] {
        Need(batch, "batch");
        Need(bytes, "bytes");
        *bytes = batch->impl.PoolBytes();
    });
}

int vqm_ss2_batch_set_shape(vqm_ss2_batch* batch, uint32_t width, uint32_t height, uint32_t strip_rows,
                            uint32_t strip_cols) {
    return Guard([&] {
        Need(batch, "batch");
        batch->impl.SetShape(width, height, strip_rows, strip_cols);
    });
}

int vqm_ss2_autotune(uint32_t width, uint32_t height, uint32_t plane_format, uint32_t map_form,
                     uint32_t pairs, vqm_ss2_tune_result* result) {
    return Guard([&] {
        Need(result, "result");
        const vqm::ss2::PlaneFormat format = vqm::ss2::ToFormat(plane_format);
        const vqm::ss2::TuneChoice c = vqm::ss2::Tuner::Get().Choose(
            width, height, format, vqm::ss2::ToMapForm(map_form, format), pairs);
        result->strip_rows = c.strip_rows;
        result->strip_cols = c.strip_cols;
        result->chunks = c.chunks;
        result->candidates = c.candidates;
        result->cached = c.cached ? 1u : 0u;
        result->ms_per_launch = c.ms_per_launch;
        result->default_ms = c.default_ms;
    });
}

void vqm_ss2_server_config_init(vqm_ss2_server_config* config) {
    if (config == nullptr) return;
    std::memset(config, 0, sizeof(*config));
    config->plane_format = VQM_PLANES_F32;
    config->lanes = 3;
    config->max_batch = 64;
}

int vqm_ss2_server_create(const vqm_ss2_server_config* config, vqm_ss2_server** out) {
    return Guard([&] {
        Need(config, "config");
        Need(out, "out");
        *out = nullptr;
        *out = new vqm_ss2_server(*config);
    });
}

void vqm_ss2_server_destroy(vqm_ss2_server* server) { delete server; }

int vqm_ss2_server_submit_device(vqm_ss2_server* server, const vqm_ss2_device_pair* pair,
                                 uint64_t* ticket) {
    return Guard([&] {
        Need(server, "server");
        Need(pair, "pair");
        Need(ticket, "ticket");
        *ticket = server->impl.SubmitDevice(*pair);
    });
}

int vqm_ss2_server_submit_host(vqm_ss2_server* server, const vqm_image_desc* original,
                               const void* original_data, const vqm_image_desc* distorted,
                               const void* distorted_data, uint64_t* ticket) {
    return Guard([&] {
        Need(server, "server");
        Need(original, "original");
        Need(distorted, "distorted");
        Need(ticket, "ticket");
        *ticket = server->impl.SubmitHost(*original, original_data, *distorted, distorted_data);
    });
}

int vqm_ss2_server_wait(vqm_ss2_server* server, uint64_t ticket, vqm_ss2_request_result* result) {
    std::string message;
    int status = VQM_OK;
    const int call = Guard([&] {
        Need(server, "server");
        status = server->impl.Wait(ticket, result, &message);
    });
    if (call != VQM_OK) return call;
    return status == VQM_OK ? VQM_OK : SetError(status, message);
}

int vqm_ss2_server_get_stats(vqm_ss2_server* server, vqm_ss2_server_stats* stats) {
    return Guard([&] {
        Need(server, "server");
        Need(stats, "stats");
        *stats = server->impl.Stats();
    });
}

uint64_t vqm_now_ns(void) { return vqm::ss2::NowNs(); }

int vqm_ss2_server_replay(vqm_ss2_server* server, const vqm_ss2_replay_request* requests,
                          uint32_t count, vqm_ss2_request_result* results) {
    std::string message;
    int status = VQM_OK;
    const int call = Guard([&] {
        Need(server, "server");
        Need(requests, "requests");
        Need(results, "results");
        status = vqm::ss2::Replay(server->impl, requests, count, results, &message);
    });
    if (call != VQM_OK) return call;
    return status == VQM_OK ? VQM_OK : SetError(status, message);
}

int vqm_device_count(int* count) {
    return Guard([&] {
        Need(count, "count");
        *count = 0;
        if (hipGetDeviceCount(count) != hipSuccess) {
            (void)hipGetLastError();
            *count = 0;
        }
    });
}

int vqm_device_info(char name[256], char arch[64], int* cus, int* lds_per_block,
                    uint64_t* hbm_bytes) {
    return Guard([&] {
        int dev = 0;
        vqm::ss2::Check(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t prop{};
        vqm::ss2::Check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        if (name) std::snprintf(name, 256, "%s", prop.name);
        if (arch) std::snprintf(arch, 64, "%s", prop.gcnArchName);
        if (cus) *cus = prop.multiProcessorCount;
        if (lds_per_block) *lds_per_block = static_cast<int>(prop.sharedMemPerBlock);
        if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    });
}

int vqm_device_malloc(uint64_t bytes, void** ptr) {
    return Guard([&] {
        Need(ptr, "ptr");
        *ptr = nullptr;
        vqm::ss2::Check(hipMalloc(ptr, bytes), "hipMalloc");
    });
}

int vqm_device_free(void* ptr) {
    return Guard([&] { vqm::ss2::Check(hipFree(ptr), "hipFree"); });
}

int vqm_host_malloc(uint64_t bytes, void** ptr) {
    return Guard([&] {
        Need(ptr, "ptr");
        *ptr = nullptr;
        vqm::ss2::Check(hipHostMalloc(ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
    });
}

int vqm_host_free(void* ptr) {
    return Guard([&] { vqm::ss2::Check(hipHostFree(ptr), "hipHostFree"); });
}

int vqm_memcpy_htod(void* dst, const void* src, uint64_t bytes) {
    return Guard([&] { vqm::ss2::Check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy"); });
}

int vqm_memcpy_dtoh(void* dst, const void* src, uint64_t bytes) {
    return Guard([&] { vqm::ss2::Check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy"); });
}

int vqm_device_copy(void* dst, const void* src, uint64_t bytes) {
    return Guard([&] {
        if (bytes % 16 != 0) vqm::ss2::Fail(VQM_INVALID_ARGUMENT, "copy size must be a multiple of 16");
        vqm::ss2::Check(vqm::ss2::LaunchCopy(dst, src, bytes, nullptr), "copy kernel");
        vqm::ss2::Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    });
}

int vqm_ss2_check_cbrt(uint32_t first_bits, uint32_t end_bits, vqm_ss2_cbrt_check* result) {
    return Guard([&] {
        Need(result, "result");
        if (first_bits < 0x00800000u || end_bits > 0x7F800000u || first_bits >= end_bits)
            vqm::ss2::Fail(VQM_INVALID_ARGUMENT,
                           "the range must lie in the positive normal floats [0x00800000, 0x7F800000) and not be empty");
        vqm::ss2::DeviceBuffer counters(sizeof(vqm::ss2::CbrtCheck));
        vqm::ss2::CbrtCheck c{};
        c.first_mismatch = 0xFFFFFFFFu;
        vqm::ss2::Check(hipMemcpy(counters.get(), &c, sizeof(c), hipMemcpyHostToDevice), "hipMemcpy");
        vqm::ss2::Check(vqm::ss2::LaunchCbrtCheck(first_bits, end_bits, counters.as<vqm::ss2::CbrtCheck>(), nullptr),
                        "cube-root check kernel");
        vqm::ss2::Check(hipMemcpy(&c, counters.get(), sizeof(c), hipMemcpyDeviceToHost), "hipMemcpy");
        result->mismatches = c.mismatches;
        result->exact_steps = c.near;
        std::memcpy(&result->max_seed_error, &c.max_h_bits, sizeof(double));
        result->first_mismatch = c.first_mismatch;
    });
}

int vqm_ss2_check_div(uint32_t first_significand, uint32_t end_significand, vqm_ss2_div_check* result) {
    return Guard([&] {
        Need(result, "result");
        if (first_significand >= end_significand || end_significand > (1u << 23))
            vqm::ss2::Fail(VQM_INVALID_ARGUMENT, "the significand range must lie in [0, 2^23) and not be empty");
        vqm::ss2::DeviceBuffer counters(sizeof(vqm::ss2::DivCheck));
        vqm::ss2::DivCheck c{};
        c.first_mismatch = 0xFFFFFFFFu;
        vqm::ss2::Check(hipMemcpy(counters.get(), &c, sizeof(c), hipMemcpyHostToDevice), "hipMemcpy");
        // Launches of at most 4096 rows (about 0.25 s each), waited for one
        // by one.
        for (uint32_t mb = first_significand; mb < end_significand; mb += 4096u) {
            const uint32_t rows = std::min(4096u, end_significand - mb);
            vqm::ss2::Check(vqm::ss2::LaunchDivCheck(mb, rows, counters.as<vqm::ss2::DivCheck>(), nullptr),
                            "division check kernel");
            vqm::ss2::Check(hipDeviceSynchronize(), "division check");
        }
        vqm::ss2::Check(hipMemcpy(&c, counters.get(), sizeof(c), hipMemcpyDeviceToHost), "hipMemcpy");
        result->mismatches = c.mismatches;
        result->reciprocal_mismatches = c.rcp_mismatches;
        result->first_mismatch = c.first_mismatch;
    });
}

int vqm_device_synchronize(void) {
    return Guard([&] { vqm::ss2::Check(hipDeviceSynchronize(), "hipDeviceSynchronize"); });
}

}  // extern "C"
