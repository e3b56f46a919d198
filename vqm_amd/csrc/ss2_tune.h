// Launch-shape autotuner: picks the scale strip height and width and the
// batch chunk count for a geometry, plane format and batch size by timing
// every candidate on the device. The strip width also picks the scale
// kernel's occupancy (one output column per lane holds half the register
// ring of two). No shape changes a pixel's arithmetic or an FP32 block sum;
// only the order of the FP64 strip partials follows the strips, so the
// choice is free to follow the measurement. Choices are cached per (width,
// height, format, map form, batch-size class) for the process.
#pragma once

#include <cstdint>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "ss2_common.h"

namespace vqm::ss2 {

struct TuneChoice {
    uint32_t strip_rows = 64;
    uint32_t strip_cols = kDefaultStripCols;
    uint32_t chunks = 1;
    float ms_per_launch = 0.0f;  // best candidate, milliseconds per launch
    float default_ms = 0.0f;     // the default shape (64 rows, 128 columns, 1 chunk)
    uint32_t candidates = 0;     // shapes timed
    bool cached = false;
};

class Tuner {
public:
    static Tuner& Get();
    // Strip heights, widths and chunk counts tried.
    static constexpr uint32_t kStripRows[] = {16, 32, 64, 128, 192, 256};
    static constexpr uint32_t kStripWidths[] = {64, 128};
    static constexpr uint32_t kChunks[] = {1, 2, 4, 8};
    // The batch-size class a pair count falls in: the next power of two,
    // capped at 512.
    static uint32_t SizeClass(uint32_t pairs);
    TuneChoice Choose(uint32_t width, uint32_t height, PlaneFormat format, MapForm map, uint32_t pairs);

private:
    std::mutex mu_;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>, TuneChoice> cache_;
};

}  // namespace vqm::ss2
