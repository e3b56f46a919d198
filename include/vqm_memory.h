// vqm_memory.h -- memory planning for the MI355X-native SSIMULACRA 2.1 path,
// in the shapes of the reference's include/vqm_memory.h:26-135.
//
// Header-only over the C ABI (vqm_ss2.h). Heaps: 0 device memory (HBM),
// 1 pinned host memory (vqm_memory_types.h). A pipeline's one allocation
// path is the direct one (device memory, with pinned staging for host
// inputs): `direct` is supported unless the policy is Staged, and `staged`
// never is.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <optional>
#include <span>
#include <stdexcept>
#include <vector>

#include "ssimulacra2.h"
#include "vqm_memory_types.h"

namespace vqm {

// vqm_memory.h:26-31: one allocation path.
struct MemoryScenario {
    bool supported = false;
    MemoryAllocation arena;
    MemoryAllocation staging;
    std::vector<MemoryAllocation> auxiliary = {};
};

// vqm_memory.h:33-42: a heap's capacity and, when known, its budget.
struct MemoryHeapInfo {
    std::uint64_t size_bytes = 0;
    bool device_local = false;
    bool budget_available = false;
    std::uint64_t budget_bytes = 0;
    std::uint64_t usage_bytes = 0;
};

// vqm_memory.h:48-53.
struct MetricMemoryRequirements {
    MemoryScenario direct;
    MemoryScenario staged;
    std::vector<MemoryHeapInfo> heaps;
    MetricScratchRequirements scratch{};
};

namespace detail {

inline MemoryAllocation Allocation(std::uint64_t bytes, std::uint32_t heap) {
    return MemoryAllocation{bytes, bytes, heap, heap};
}

// The heaps of the current device and host (vqm_memory_heaps); zero sizes
// and no budget where there is no device.
inline std::vector<MemoryHeapInfo> Heaps() {
    vqm_memory_heap h[2]{};
    std::vector<MemoryHeapInfo> out;
    if (vqm_memory_heaps(h) != VQM_OK) return out;
    for (const auto& x : h)
        out.push_back(MemoryHeapInfo{x.size_bytes, x.device_local != 0, x.budget_available != 0,
                                     x.budget_bytes, x.usage_bytes});
    return out;
}

// AddBytes (vulkan_memory_query.cc:233-238).
inline void AddBytes(std::uint64_t& total, std::uint64_t value) {
    if (value > std::numeric_limits<std::uint64_t>::max() - total)
        throw std::overflow_error("memory: heap allocation size overflow");
    total += value;
}

// A supported scenario's bytes, added per heap (vulkan_memory_query.cc:240-257).
inline void Charge(std::vector<std::uint64_t>& per_heap, const MemoryScenario& scenario) {
    if (!scenario.supported) throw std::invalid_argument("memory: unsupported or empty scenario");
    auto add = [&](const MemoryAllocation& a) {
        if (a.allocation_bytes == 0) return;
        if (a.heap >= per_heap.size()) throw std::invalid_argument("memory: missing heap budget");
        AddBytes(per_heap[a.heap], a.allocation_bytes);
    };
    add(scenario.arena);
    add(scenario.staging);
    for (const auto& a : scenario.auxiliary) add(a);
}

// Workers that fit: the smallest budget / charge over charged heaps
// (vulkan_memory_query.cc:259-269).
inline std::size_t Workers(std::span<const std::uint64_t> charge, std::span<const std::uint64_t> budgets) {
    std::uint64_t n = std::numeric_limits<std::uint64_t>::max();
    for (std::size_t i = 0; i < charge.size(); ++i)
        if (charge[i] != 0) n = std::min(n, budgets[i] / charge[i]);
    return static_cast<std::size_t>(std::min<std::uint64_t>(n, std::numeric_limits<std::size_t>::max()));
}

}  // namespace detail

// vqm_memory.h:55-60 (vulkan_memory_query.cc:298-311): the remaining bytes
// per heap from the budget snapshot, or nullopt when a heap has none.
[[nodiscard]] inline std::optional<std::vector<std::uint64_t>> AvailableHeapBudgets(
    const MetricMemoryRequirements& requirements) {
    std::vector<std::uint64_t> budgets;
    for (const auto& heap : requirements.heaps) {
        if (!heap.budget_available) return std::nullopt;
        budgets.push_back(heap.budget_bytes > heap.usage_bytes ? heap.budget_bytes - heap.usage_bytes : 0);
    }
    return budgets;
}

// vqm_memory.h:62-70: the bytes a pipeline with `config` will hold, planned
// before construction without allocating. Throws std::invalid_argument for
// a geometry or configuration the pipeline would refuse.
[[nodiscard]] inline MetricMemoryRequirements GetMemoryRequirements(const Ssimulacra2Config& config) {
    const MemoryPolicy policy = config.memory_policy;
    Ssimulacra2Config direct_config = config;
    direct_config.memory_policy = MemoryPolicy::Automatic;
    const vqm_ss2_config c = direct_config.ToC();
    vqm_ss2_memory m{};
    detail::Check(vqm_ss2_memory_plan(&c, &m));
    MetricMemoryRequirements r;
    r.direct.supported = policy != MemoryPolicy::Staged;
    r.direct.arena = detail::Allocation(m.total_device_bytes, kDeviceHeap);
    r.direct.staging = detail::Allocation(m.pinned_host_bytes, kPinnedHostHeap);
    r.heaps = detail::Heaps();
    return r;
}

// vqm_memory.h:110-124 (vulkan_memory_query.cc:375-393): how many
// independent pipelines fit the per-heap budgets. Throws
// std::invalid_argument for an unsupported scenario or a missing heap
// budget, std::overflow_error on overflow. Arithmetic, not a guarantee.
[[nodiscard]] inline std::size_t MaxWorkersForBudget(std::span<const MemoryScenario> per_worker,
                                                     std::span<const std::uint64_t> heap_budgets) {
    if (per_worker.empty()) throw std::invalid_argument("memory: no scenarios");
    std::vector<std::uint64_t> charge(heap_budgets.size(), 0);
    for (const auto& s : per_worker) detail::Charge(charge, s);
    return detail::Workers(charge, heap_budgets);
}
[[nodiscard]] inline std::size_t MaxWorkersForBudget(const MemoryScenario& scenario,
                                                     std::span<const std::uint64_t> heap_budgets) {
    return MaxWorkersForBudget(std::span<const MemoryScenario>(&scenario, 1), heap_budgets);
}

// vqm_memory.h:126-133 (vulkan_memory_query.cc:395-424): each worker's
// metrics charged at their largest supported alternative.
[[nodiscard]] inline std::size_t MaxWorkersForBudgetWorstCase(
    std::span<const MetricMemoryRequirements> per_worker, std::span<const std::uint64_t> heap_budgets) {
    if (per_worker.empty()) throw std::invalid_argument("memory: no requirements");
    std::vector<std::uint64_t> total(heap_budgets.size(), 0);
    for (const auto& req : per_worker) {
        std::vector<std::uint64_t> largest(heap_budgets.size(), 0);
        bool supported = false;
        for (const MemoryScenario* s : {&req.direct, &req.staged}) {
            if (!s->supported) continue;
            supported = true;
            std::vector<std::uint64_t> charge(heap_budgets.size(), 0);
            detail::Charge(charge, *s);
            for (std::size_t i = 0; i < charge.size(); ++i) largest[i] = std::max(largest[i], charge[i]);
        }
        if (!supported) throw std::invalid_argument("memory: unsupported requirements");
        for (std::size_t i = 0; i < total.size(); ++i) detail::AddBytes(total[i], largest[i]);
    }
    return detail::Workers(total, heap_budgets);
}

}  // namespace vqm
