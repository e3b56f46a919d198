// vqm_metrics.h -- the SSIMULACRA 2 slice of the unified metrics API.
//
// Header-only, over the Ssimulacra2Pipeline facade (ssimulacra2.h). It keeps
// the reference's interface for this metric (GreatValueCreamSoda/vqm
// include/vqm_metrics.h:35-215, src/metrics/metrics.cc:23-398,
// src/metrics/metric_pipelines.cc:29-128,187-279): a MetricsConfig selects
// the metric by holding its configuration; Metrics scores an ordered
// sequence of frame pairs, numbering them from 0; Flush closes the sequence,
// Finalize reports the frame count, Reset starts a new one; ComputeImage
// scores one independent pair. SSIMULACRA 2 is the only metric this library
// computes, so a result is never held back (the reference's lookahead belongs
// to VMAF) and the reference's members for the other metrics are absent.
//
// Errors, as in the reference:
//   std::invalid_argument  no metric selected, background outside [0, 1],
//                          invalid format, a pair whose size or format
//                          differs from the configuration, a packed span of
//                          the wrong length, streaming inputs disabled. These
//                          are raised before any work and leave the sequence
//                          usable.
//   std::logic_error       a call after a failed one or after Flush (Reset
//                          first), ComputeImage inside an active sequence
//   std::overflow_error    the frame counter is exhausted
// A failure inside the scoring itself marks the sequence failed: call Reset.
//
// Memory: the reference shares one scratch arena among several selected
// metrics. With one metric there is nothing to share, so `scratch` is always
// empty and `spatial` holds the pipeline's own allocations, the same figure
// GetMemoryRequirements plans. Budgets have two heaps here: heap 0 is device
// memory (HBM), heap 1 pinned host memory (vqm_memory_types.h).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ssimulacra2.h"
#include "vqm_memory.h"

namespace vqm {

// vqm_metrics.h:35-47. size, streaming_inputs and background govern the
// selected metric: its own size, streaming_inputs, resident slot count and
// input_bytes pool are overridden (metric_pipelines.cc:81-93).
struct MetricsConfig {
    Size size;
    ImageFormat original_format{};
    ImageFormat distorted_format{};
    bool streaming_inputs = true;
    MemoryPolicy memory_policy = MemoryPolicy::Automatic;
    float background = 0.5f;
    std::optional<Ssimulacra2Config> ssimulacra2 = std::nullopt;
};

// vqm_metrics.h:50-61: one frame's scores.
struct MetricsResult {
    std::uint64_t frame_index = 0;
    std::optional<double> ssimulacra2;
};

// vqm_metrics.h:68-75.
struct MetricsMemoryUsage {
    std::optional<PipelineMemoryUsage> spatial;
    std::optional<PipelineMemoryUsage> scratch;
};

// vqm_metrics.h:80-87.
struct MetricsMemoryRequirements {
    std::optional<MetricMemoryRequirements> spatial;
    std::optional<MetricMemoryRequirements> scratch;
};

// vqm_metrics.h:147-152 (no CVVDP clip score here).
struct MetricsFinalResult {
    std::optional<MetricsResult> last_frame;
    std::uint64_t frames = 0;
    std::optional<MetricsResult> penultimate_frame;
};

namespace detail {

// ValidateImage (input.cc:152-222) with the reference's "<name>: <reason>".
inline void ValidateFormatNamed(const ImageFormat& format, Size size, const char* name) {
    const vqm_image_format f = format.ToC();
    const int status = vqm_validate_image_format(&f, Dim(size.width), Dim(size.height));
    if (status == VQM_OK) return;
    std::string reason = vqm_last_error();
    const std::string generic = "image: ";
    if (reason.compare(0, generic.size(), generic) == 0) reason.erase(0, generic.size());
    throw std::invalid_argument(std::string(name) + ": " + reason);
}

// SelectConfigs (metric_pipelines.cc:29-128): the SSIMULACRA 2 configuration
// the collection builds, after the top-level fields override its own.
inline Ssimulacra2Config SelectSsimulacra2(const MetricsConfig& config) {
    if (!config.ssimulacra2) throw std::invalid_argument("metrics: select at least one metric");
    if (!std::isfinite(config.background) || config.background < 0.0f || config.background > 1.0f)
        throw std::invalid_argument("metrics: background must be between 0 and 1");
    ValidateFormatNamed(config.original_format, config.size, "original format");
    ValidateFormatNamed(config.distorted_format, config.size, "distorted format");
    Ssimulacra2Config metric = *config.ssimulacra2;
    metric.size = config.size;
    metric.streaming_inputs = config.streaming_inputs;
    metric.memory_policy = config.memory_policy;
    metric.gpu_resident_images = 0;
    metric.input_bytes = 0;
    return metric;
}

}  // namespace detail

// GetMemoryRequirements(MetricsConfig) (metric_pipelines.cc:187-196): planned
// without touching the device; throws as the constructors would.
[[nodiscard]] inline MetricsMemoryRequirements GetMemoryRequirements(const MetricsConfig& config) {
    MetricsMemoryRequirements r;
    r.spatial = GetMemoryRequirements(detail::SelectSsimulacra2(config));
    return r;
}

// MaxWorkersForBudgetWorstCase (metric_pipelines.cc:197-212): the
// selected metrics' requirements (and the shared scratch, never present
// here) charged together per worker against the per-heap budgets (heap 0
// device memory, heap 1 pinned host memory). Arithmetic, not an allocation
// guarantee.
[[nodiscard]] inline std::size_t MaxWorkersForBudgetWorstCase(const MetricsMemoryRequirements& requirements,
                                                              std::span<const std::uint64_t> heap_budgets) {
    std::vector<MetricMemoryRequirements> parts;
    if (requirements.scratch) parts.push_back(*requirements.scratch);
    if (requirements.spatial) parts.push_back(*requirements.spatial);
    return MaxWorkersForBudgetWorstCase(std::span<const MetricMemoryRequirements>(parts), heap_budgets);
}

// vqm_metrics.h:97-119: owns the selected SSIMULACRA 2 pipeline, configured
// as Metrics configures its own. The pointer the getter returns stays valid
// across moves of the collection. One collection is one worker.
class MetricPipelines {
public:
    explicit MetricPipelines(MetricsConfig config)
        : state_(std::make_unique<State>(State{std::move(config), nullptr})) {
        state_->ssimu = std::make_unique<Ssimulacra2Pipeline>(detail::SelectSsimulacra2(state_->config));
    }
    MetricPipelines(MetricPipelines&&) noexcept = default;
    MetricPipelines& operator=(MetricPipelines&&) noexcept = default;
    MetricPipelines(const MetricPipelines&) = delete;
    MetricPipelines& operator=(const MetricPipelines&) = delete;

    [[nodiscard]] Ssimulacra2Pipeline* Ssimulacra2() noexcept { return state_->ssimu.get(); }
    [[nodiscard]] MetricsMemoryUsage GetMemoryUsage() const {
        MetricsMemoryUsage usage;
        usage.spatial = state_->ssimu->GetMemoryUsage();
        return usage;
    }

private:
    struct State {
        MetricsConfig config;
        std::unique_ptr<Ssimulacra2Pipeline> ssimu;
    };
    std::unique_ptr<State> state_;
};

// vqm_metrics.h:154-215 (metrics.cc:78-398). One Metrics is one worker and
// is not thread-safe. The DeviceImageView overloads replace the
// VulkanBufferView ones; `wait_stream` (a hipStream_t) replaces the waits of
// VulkanSubmitSync: the pair's work starts after the work already queued
// there. Each call returns with the score, so there is no signal to schedule.
class Metrics {
public:
    explicit Metrics(MetricsConfig config) : state_(std::make_unique<State>(std::move(config))) {}
    Metrics(Metrics&&) noexcept = default;
    Metrics& operator=(Metrics&&) noexcept = default;
    Metrics(const Metrics&) = delete;
    Metrics& operator=(const Metrics&) = delete;

    [[nodiscard]] std::optional<MetricsResult> Compute(const Image& original, const Image& distorted) {
        return state_->Run(original, distorted);
    }
    [[nodiscard]] std::optional<MetricsResult> Compute(const ImageView& original, const ImageView& distorted) {
        return state_->Run(original, distorted);
    }
    // Tightly packed planes of the configured formats; each span must be
    // exactly RequiredBytes(format, size) long.
    [[nodiscard]] std::optional<MetricsResult> Compute(std::span<const std::byte> original,
                                                       std::span<const std::byte> distorted) {
        return state_->Run(state_->PackedInput(original, true), state_->PackedInput(distorted, false));
    }
    [[nodiscard]] std::optional<MetricsResult> Compute(const DeviceImageView& original,
                                                       const DeviceImageView& distorted,
                                                       void* wait_stream = nullptr) {
        return state_->Run(Device{original, wait_stream}, Device{distorted, wait_stream});
    }

    // Scores one independent pair: requires a new or reset sequence and
    // resets it after returning.
    [[nodiscard]] MetricsResult ComputeImage(const Image& original, const Image& distorted) {
        return state_->RunImage(original, distorted);
    }
    [[nodiscard]] MetricsResult ComputeImage(const ImageView& original, const ImageView& distorted) {
        return state_->RunImage(original, distorted);
    }
    [[nodiscard]] MetricsResult ComputeImage(std::span<const std::byte> original,
                                             std::span<const std::byte> distorted) {
        return state_->RunImage(state_->PackedInput(original, true), state_->PackedInput(distorted, false));
    }
    [[nodiscard]] MetricsResult ComputeImage(const DeviceImageView& original, const DeviceImageView& distorted,
                                             void* wait_stream = nullptr) {
        return state_->RunImage(Device{original, wait_stream}, Device{distorted, wait_stream});
    }

    // Closes the input; returns the remaining result, of which there is
    // never one here.
    [[nodiscard]] std::optional<MetricsResult> Flush() { return state_->Flush(); }
    // metrics.cc:368-375: two flushes; `frames` counts the frames consumed.
    [[nodiscard]] MetricsFinalResult Finalize() {
        auto tail = state_->Flush();
        auto last = state_->Flush();
        if (!last) return {std::move(tail), state_->next_index, std::nullopt};
        return {std::move(last), state_->next_index, std::move(tail)};
    }
    void Reset() { state_->Reset(); }
    [[nodiscard]] MetricsMemoryUsage GetMemoryUsage() const {
        MetricsMemoryUsage usage;
        usage.spatial = state_->ssimu.GetMemoryUsage();
        return usage;
    }

private:
    struct Device {
        DeviceImageView view;
        void* wait_stream;
    };
    struct Packed {
        vqm_image_desc desc;
        const void* data;
    };

    struct State {
        explicit State(MetricsConfig c) : config(std::move(c)), ssimu(detail::SelectSsimulacra2(config)) {}

        // metrics.cc:82-88.
        void RequireOpen() const {
            if (failed) throw std::logic_error("metrics: previous call failed; call Reset");
            if (finished) throw std::logic_error("metrics: sequence finished; call Reset");
        }

        // ValidatePairFormat (metrics.cc:26-43).
        void CheckPair(Size a_size, const ImageFormat& a, Size b_size, const ImageFormat& b) const {
            if (a_size != config.size || b_size != config.size)
                throw std::invalid_argument("metrics: image size differs from configured size");
            if (a != config.original_format || b != config.distorted_format)
                throw std::invalid_argument("metrics: image format differs from configured format");
        }
        void CheckStreaming() const {
            if (!config.streaming_inputs) throw std::invalid_argument("metrics: streaming inputs are disabled");
        }

        // PackedView (metrics.cc:60-75): the span's length is checked
        // before anything else is.
        Packed PackedInput(std::span<const std::byte> bytes, bool original) const {
            const ImageFormat& format = original ? config.original_format : config.distorted_format;
            if (bytes.size() != RequiredBytes(format, config.size))
                throw std::invalid_argument("metrics: packed input has incorrect size");
            return {vqm_image_desc{detail::Dim(config.size.width), detail::Dim(config.size.height), format.ToC(),
                                   bytes.size()},
                    bytes.data()};
        }

        // The storage checks run before the try of Run (metrics.cc:160-176),
        // so a rejected pair leaves the sequence usable.
        void Validate(const Image& a, const Image& b) const {
            CheckPair(a.size, a.format, b.size, b.format);
            CheckImage(a, "original");
            CheckImage(b, "distorted");
            CheckStreaming();
        }
        void Validate(const ImageView& a, const ImageView& b) const {
            CheckPair(a.size, a.format, b.size, b.format);
            for (const ImageView* v : {&a, &b}) {
                const vqm_host_view hv = detail::HostView(*v);
                std::uint64_t bytes = 0;
                detail::Check(vqm_copy_image(&hv, nullptr, &bytes));
            }
            CheckStreaming();
        }
        void Validate(const Packed&, const Packed&) const { CheckStreaming(); }
        void Validate(const Device& a, const Device& b) const {
            CheckPair(a.view.size, a.view.format, b.view.size, b.view.format);
            detail::ValidateFormatNamed(a.view.format, a.view.size, "original");
            detail::ValidateFormatNamed(b.view.format, b.view.size, "distorted");
            if (a.view.data == nullptr || b.view.data == nullptr)
                throw std::invalid_argument("metrics: invalid device input");
        }
        // CheckImage (input.cc:224-240).
        void CheckImage(const Image& image, const char* name) const {
            detail::ValidateFormatNamed(image.format, image.size, name);
            const std::size_t expected = RequiredBytes(image.format, image.size);
            if (image.data.size() != expected)
                throw std::invalid_argument(std::string(name) +
                                            ": data size does not match the image format (expected " +
                                            std::to_string(expected) + " bytes)");
        }

        double Score(const Image& a, const Image& b) { return ssimu.Compute(a, b, config.background).Score(); }
        double Score(const ImageView& a, const ImageView& b) {
            return ssimu.Compute(a, b, config.background).Score();
        }
        double Score(const Packed& a, const Packed& b) {
            return ssimu.Finish(ssimu.SubmitPacked(a.desc, a.data, b.desc, b.data, config.background)).Score();
        }
        double Score(const Device& a, const Device& b) {
            return ssimu.Finish(ssimu.Submit(a.view, b.view, config.background, a.wait_stream)).Score();
        }

        // Run (metrics.cc:155-262).
        template <class Input>
        std::optional<MetricsResult> Run(const Input& a, const Input& b) {
            RequireOpen();
            if (next_index == std::numeric_limits<std::uint64_t>::max())
                throw std::overflow_error("metrics: too many frames");
            Validate(a, b);
            try {
                MetricsResult value;
                value.frame_index = next_index;
                value.ssimulacra2 = Score(a, b);
                ++next_index;
                return value;
            } catch (...) {
                failed = true;
                throw;
            }
        }

        // RunImage (metrics.cc:264-285).
        template <class Input>
        MetricsResult RunImage(const Input& a, const Input& b) {
            RequireOpen();
            if (next_index != 0) throw std::logic_error("metrics: reset the active sequence before ComputeImage");
            auto first = Run(a, b);
            auto last = Flush();
            if (first && last) throw std::logic_error("metrics: unexpected extra still-image result");
            if (!first && !last) throw std::logic_error("metrics: missing still-image result");
            MetricsResult result = first ? std::move(*first) : std::move(*last);
            Reset();
            return result;
        }

        // Flush (metrics.cc:287-308): nothing is ever pending.
        std::optional<MetricsResult> Flush() {
            if (failed) throw std::logic_error("metrics: previous call failed; call Reset");
            finished = true;
            return std::nullopt;
        }

        void Reset() {
            next_index = 0;
            failed = false;
            finished = false;
        }

        MetricsConfig config;
        Ssimulacra2Pipeline ssimu;
        std::uint64_t next_index = 0;
        bool failed = false;
        bool finished = false;
    };
    std::unique_ptr<State> state_;
};

}  // namespace vqm
