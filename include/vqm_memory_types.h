// vqm_memory_types.h -- allocation records of the MI355X-native SSIMULACRA 2.1
// path, in the shapes of the reference's include/vqm_memory_types.h:14-38.
//
// Two memory heaps describe every allocation here: heap 0 is device memory
// in HBM (memory type 0, hipMalloc: workspaces, resident slots, streaming
// input buffers, the caller-owned image pool), heap 1 pinned host memory
// (memory type 1, hipHostMalloc: the staging of host inputs and the norm
// read-back). The arena is the pipeline's device memory, the staging its
// pinned host memory; nothing is auxiliary.
#pragma once

#include <cstdint>
#include <vector>

namespace vqm {

// vqm_memory_types.h:14. Device memory is the only arena path: Automatic
// and Direct select it; Staged (a host-visible staging arena in front of
// device memory) does not exist here, and a pipeline or plan configured
// with it fails as the reference fails a strict policy whose path is
// unavailable.
enum class MemoryPolicy : std::uint8_t { Automatic, Direct, Staged };

// The heap indices and memory types of the two kinds of memory above.
inline constexpr std::uint32_t kDeviceHeap = 0;
inline constexpr std::uint32_t kPinnedHostHeap = 1;

// vqm_memory_types.h:16-23: one allocation. buffer_bytes is what the
// library uses, allocation_bytes what it holds; an absent allocation has
// zero sizes.
struct MemoryAllocation {
    std::uint64_t buffer_bytes = 0;
    std::uint64_t allocation_bytes = 0;
    std::uint32_t memory_type = 0;
    std::uint32_t heap = 0;
};

// vqm_memory_types.h:27-35: what one pipeline holds. uses_staging: host
// inputs pass through pinned staging (streaming inputs are enabled).
struct PipelineMemoryUsage {
    bool uses_staging = false;
    MemoryAllocation arena;
    MemoryAllocation staging;
    std::vector<MemoryAllocation> auxiliary = {};
};

// vqm_memory_types.h:37-42: scratch shared between several metrics. One
// metric shares nothing, so requirements carry an empty one.
struct MetricScratchRequirements {
    std::uint64_t arena_bytes = 0;
    std::uint64_t fallback_staging_bytes = 0;
    std::vector<std::uint64_t> persistent_bytes = {};
    bool reusable = false;
};

}  // namespace vqm
