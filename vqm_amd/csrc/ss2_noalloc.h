// Allocation-free scoring of one pair of caller-owned 8-bit sRGB RGB device
// images over a fixed, file-scope device scratch array (ss2_noalloc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "vqm_ss2.h"

namespace vqm::ss2 {

// Size of the scratch array each device holds for the allocation-free entry
// point: the workspace of any pair up to 4096x2160 in every plane format.
inline constexpr uint64_t kNoAllocScratchBytes = 128ull << 20;

// Looks up the current device's scratch address (once per device).
void InitNoAllocScratch();

// Queues the scoring of (a, b) on `stream` with the kernels of
// vqm_ss2_submit_device, reading the plan, descriptors and workspace from
// the scratch; the 108 norms land in `norms` (device memory). Returns the
// level count. Throws Error{VQM_INVALID_ARGUMENT} for a mismatched image or
// a workspace larger than the scratch.
uint32_t ScoreDeviceNoAlloc(const vqm_ss2_config& config, const vqm_device_image& a,
                            const vqm_device_image& b, hipStream_t stream, double* norms);

}  // namespace vqm::ss2
