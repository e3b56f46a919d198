// Host-side launch API of the SSIMULACRA 2.1 kernels (ss2_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ss2_common.h"

namespace vqm::ss2 {

// A flat work list entry: pair index in the batch and the local work index
// (prepare: tile | image << 31; scale: item). Used when a batch mixes
// geometries; a batch of one geometry indexes blocks arithmetically.
struct WorkRef {
    uint32_t pair;
    uint32_t index;
};

// sRGB decode table and taps live in device memory owned by the caller.
// All launches are asynchronous on `stream`.

// Builds every stored plane of both images of each pair.
// Uniform geometry: blocks = pairs * 2 * tiles_per_image, work = nullptr.
// src0/src1, when not null, replace every pair's inputs (single-pair use).
// general: some image of the launch has a SourceDesc of kind 1 (formats
// other than 8-bit sRGB RGB); the table-only kernel ignores descriptors.
// `persistent`: launch only as many blocks as stay resident at once, each
// looping over the work items (the grid otherwise covers them, one each).
// `map` picks the arithmetic form of the XYB transform (MapForm).
hipError_t LaunchPrepare(PlaneFormat format, MapForm map, bool general, const PairDesc* pairs,
                         const WorkRef* work, uint32_t blocks, uint32_t tiles_per_image, const float* lut,
                         const uint8_t* src0, const uint8_t* src1, hipStream_t stream,
                         bool persistent = false);

// One 64-lane wave per scale work item (level, channel, strip of 64 or 128
// columns). Uniform geometry: blocks = pairs * items_per_pair, work =
// nullptr. `blur` picks the kernel: kBlurSampled (13-tap sampled Gaussian)
// or kBlurRecursive (the recursive Gaussian's 9-tap truncated-cosine
// kernel); `map` the SSIM' map's form, which must be the one the planes were
// prepared with; `strip_cols` the strip width of every plan in the launch
// (Plan::strip_w).
hipError_t LaunchScale(PlaneFormat format, uint32_t blur, MapForm map, uint32_t strip_cols,
                       const PairDesc* pairs, const WorkRef* work, uint32_t blocks, uint32_t items_per_pair,
                       hipStream_t stream, bool persistent = false);

// Sums each (level, channel) norm's partials in FP64 and writes the 108 norms.
hipError_t LaunchReduce(const PairDesc* pairs, uint32_t pairs_count, hipStream_t stream);

// Device-to-device copy of `bytes` (a multiple of 16) by a plain kernel.
hipError_t LaunchCopy(void* dst, const void* src, uint64_t bytes, hipStream_t stream);

// Counters of the cube-root self-check; `first_mismatch` starts at
// 0xFFFFFFFF and the others at 0.
struct CbrtCheck {
    unsigned long long mismatches;
    unsigned long long near;
    unsigned long long max_h_bits;
    uint32_t first_mismatch;
    uint32_t pad;
};

// Compares the reference form's cube root (hardware-seeded) with the one of
// IEEE steps only for every float with bits in [first, end), adding to `out`.
hipError_t LaunchCbrtCheck(uint32_t first, uint32_t end, CbrtCheck* out, hipStream_t stream);

// Counters of the division self-check; `first_mismatch` starts at
// 0xFFFFFFFF and the others at 0.
struct DivCheck {
    unsigned long long mismatches;
    unsigned long long rcp_mismatches;
    uint32_t first_mismatch;
    uint32_t pad;
};

// Compares the map's division with the IEEE quotient for b significands
// [mb_first, mb_first + rows) against every a significand, and the hardware
// reciprocal's exponent invariance for those b, adding to `out`.
hipError_t LaunchDivCheck(uint32_t mb_first, uint32_t rows, DivCheck* out, hipStream_t stream);

}  // namespace vqm::ss2
