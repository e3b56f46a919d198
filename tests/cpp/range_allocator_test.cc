// RangeAllocator (the workspace arena's bookkeeping) against a byte-map
// model: random allocations and releases never overlap, stay inside the
// capacity, are 256-byte aligned, and releasing everything leaves one free
// range of the whole capacity. First fit is checked against the model's own
// lowest fitting gap.

#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "ss2_host.h"

using vqm::ss2::RangeAllocator;

static int failures = 0;
#define CHECK(cond)                                                          \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            if (++failures > 20) std::exit(1);                               \
        }                                                                    \
    } while (0)

int main() {
    constexpr uint64_t kUnits = 4096;  // capacity in 256-byte units
    RangeAllocator a(kUnits * 256 + 100);  // the tail below one unit is unusable
    CHECK(a.capacity() == kUnits * 256);
    std::vector<int> owner(kUnits, -1);
    struct Live { uint64_t offset, bytes; int id; };
    std::vector<Live> live;
    std::mt19937_64 rng(7);
    int next_id = 0;
    for (int step = 0; step < 200000; ++step) {
        const bool alloc = live.empty() || (rng() % 100) < 55;
        if (alloc) {
            const uint64_t bytes = 1 + rng() % (rng() % 4 == 0 ? 64 * 256 : 8 * 256);
            const uint64_t units = (bytes + 255) / 256;
            // The model's first fit: lowest run of `units` free units.
            uint64_t expect = RangeAllocator::kNone, run = 0;
            for (uint64_t u = 0; u < kUnits; ++u) {
                run = owner[u] < 0 ? run + 1 : 0;
                if (run == units) {
                    expect = (u + 1 - units) * 256;
                    break;
                }
            }
            const uint64_t off = a.Alloc(bytes);
            CHECK(off == expect);
            if (off == RangeAllocator::kNone) continue;
            CHECK(off % 256 == 0 && off + units * 256 <= a.capacity());
            for (uint64_t u = off / 256; u < off / 256 + units; ++u) {
                CHECK(owner[u] < 0);
                owner[u] = next_id;
            }
            live.push_back({off, bytes, next_id++});
        } else {
            const size_t k = rng() % live.size();
            const Live l = live[k];
            live[k] = live.back();
            live.pop_back();
            a.Free(l.offset, l.bytes);
            for (uint64_t u = l.offset / 256; u < l.offset / 256 + (l.bytes + 255) / 256; ++u) {
                CHECK(owner[u] == l.id);
                owner[u] = -1;
            }
        }
        uint64_t free_units = 0;
        for (int o : owner) free_units += o < 0;
        CHECK(a.free_bytes() == free_units * 256);
    }
    for (const Live& l : live) a.Free(l.offset, l.bytes);
    CHECK(a.free_bytes() == a.capacity());
    CHECK(a.free_ranges() == 1);
    CHECK(a.Alloc(a.capacity()) == 0);
    CHECK(a.Alloc(1) == RangeAllocator::kNone);
    if (failures) {
        std::fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    std::printf("range allocator ok\n");
    return 0;
}
