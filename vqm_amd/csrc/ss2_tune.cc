// Launch-shape autotuner (see ss2_tune.h).

#include "ss2_tune.h"

#include <algorithm>
#include <chrono>
#include <limits>

#include "ss2_host.h"
#include "ss2_runtime.h"

namespace vqm::ss2 {

namespace {
// Workspace memory the tuner may hold at once; larger classes are timed on
// fewer pairs and compared per pair.
constexpr uint64_t kTuneBytes = 16ull << 30;
constexpr int kRepeats = 3;
}  // namespace

Tuner& Tuner::Get() {
    static Tuner tuner;
    return tuner;
}

uint32_t Tuner::SizeClass(uint32_t pairs) {
    uint32_t c = 1;
    while (c < pairs && c < 512) c *= 2;
    return c;
}

TuneChoice Tuner::Choose(uint32_t width, uint32_t height, PlaneFormat format, MapForm map, uint32_t pairs) {
    PlanScales(width, height);
    ToFormat(format);
    const uint32_t cls = SizeClass(std::max(pairs, 1u));
    const auto key = std::make_tuple(width, height, static_cast<uint32_t>(format), static_cast<uint32_t>(map), cls);
    std::lock_guard<std::mutex> lock(mu_);
    if (auto it = cache_.find(key); it != cache_.end()) {
        TuneChoice c = it->second;
        c.cached = true;
        return c;
    }
    DeviceContext& ctx = DeviceContext::Get();
    // Timing does not depend on pixel values: every candidate scores the
    // same zero-filled pair.
    const uint64_t image = (RequiredBytes(Srgb8Format(), width, height) + 255) / 256 * 256;
    DeviceBuffer inputs(2 * image);
    Check(hipMemset(inputs.get(), 0, 2 * image), "clear tuning inputs");
    Check(hipDeviceSynchronize(), "clear tuning inputs");
    const uint8_t* src[2] = {inputs.as<uint8_t>(), inputs.as<uint8_t>() + image};

    TuneChoice best;
    float best_per_pair = std::numeric_limits<float>::infinity();
    for (uint32_t strip : kStripRows) {
        for (uint32_t cols : kStripWidths) {
            WorkspacePools pools(format, strip, cols, 0);
            const DevicePlan& plan = pools.PlanFor(width, height);
            const uint64_t each = pools.Need(width, height, false);
            const uint32_t n =
                static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(cls, kTuneBytes / each)));
            Lane lane(n);
            std::vector<std::unique_ptr<Workspace>> ws;
            std::vector<LaunchEntry> entries(n);
            for (uint32_t i = 0; i < n; ++i) {
                ws.push_back(pools.Acquire(width, height, false));
                entries[i].plan = &plan;
                entries[i].ws = ws.back().get();
                entries[i].src[0] = src[0];
                entries[i].src[1] = src[1];
            }
            for (uint32_t chunks : kChunks) {
                if (chunks > n) continue;
                lane.Launch(format, map, ctx.lut, entries.data(), n, chunks);  // warm-up
                lane.Wait();
                float ms = std::numeric_limits<float>::infinity();
                for (int r = 0; r < kRepeats; ++r) {
                    const auto t0 = std::chrono::steady_clock::now();
                    lane.Launch(format, map, ctx.lut, entries.data(), n, chunks);
                    lane.Wait();
                    const auto t1 = std::chrono::steady_clock::now();
                    ms = std::min(ms, std::chrono::duration<float, std::milli>(t1 - t0).count());
                }
                const float per_pair = ms / static_cast<float>(n);
                ++best.candidates;
                if (strip == 64 && cols == kDefaultStripCols && chunks == 1)
                    best.default_ms = per_pair * static_cast<float>(cls);
                if (per_pair < best_per_pair) {
                    best_per_pair = per_pair;
                    best.strip_rows = strip;
                    best.strip_cols = cols;
                    best.chunks = chunks;
                }
            }
        }
    }
    best.ms_per_launch = best_per_pair * static_cast<float>(cls);
    cache_[key] = best;
    return best;
}

}  // namespace vqm::ss2
