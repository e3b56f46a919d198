// Host-only check of the work plans the kernels consume (vqm_amd/csrc/
// ss2_host.cc): for every geometry, strip height and strip width, every
// address the scale and prepare kernels compute lies inside its plane,
// planes do not overlap, the work items cover each (level, channel) exactly
// once, and strips start on the map-sum block rows.
//
// Built and run by tests/test_plan_bounds.py (g++, no GPU).

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "ss2_common.h"
#include "ss2_host.h"

using namespace vqm::ss2;

static int failures = 0;
#define EXPECT(cond, ...)                                      \
    do {                                                       \
        if (!(cond)) {                                         \
            if (failures < 20) {                               \
                std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
                std::fprintf(stderr, __VA_ARGS__);             \
                std::fprintf(stderr, "\n");                    \
            }                                                  \
            ++failures;                                        \
        }                                                      \
    } while (0)

static void check(uint32_t width, uint32_t height, PlaneFormat fmt, uint32_t strip_rows, uint32_t strip_cols) {
    const GeometryPlan g = BuildPlan(width, height, fmt, strip_rows, strip_cols);
    const Plan& p = g.plan;
    const uint32_t sh = p.strip_h, sw = p.strip_w;
    EXPECT(sw == (strip_cols ? strip_cols : 128u), "strip width %u", sw);
    EXPECT(sh % kSumBlockRows == 0, "strip height %u", sh);
    // Planes: aligned, disjoint, inside image_bytes (offsets in bytes, the
    // element size per level).
    uint64_t prev_end = 0;
    for (uint32_t l = 0; l < p.levels; ++l) {
        const uint32_t e = ElementBytes(fmt);
        EXPECT(e == (fmt == kPlaneF32 ? 4u : 2u), "element size");
        EXPECT(p.pitch[l] % 4 == 0, "pitch %u not a multiple of 4", p.pitch[l]);
        const uint32_t strips = (p.w[l] + kStripCols - 1) / kStripCols;
        EXPECT(p.pitch[l] >= strips * kStripCols + 14, "pitch too small at level %u", l);
        EXPECT(p.rows[l] >= (p.h[l] + sh - 1) / sh * sh + 2 * kRadius, "rows too small");
        for (int c = 0; c < 3; ++c) {
            const bool active = p.mask[l] & (1u << c);
            EXPECT(active == !ChannelPruned(c, l, p.levels), "mask mismatch");
            if (!active) {
                EXPECT(p.plane_off[l][c] == ~0ull, "pruned plane has storage");
                continue;
            }
            EXPECT(p.plane_off[l][c] % 256 == 0, "plane offset not 256-byte aligned");
            EXPECT(p.plane_off[l][c] >= prev_end, "planes overlap");
            prev_end = p.plane_off[l][c] + static_cast<uint64_t>(p.pitch[l]) * p.rows[l] * e;
        }
    }
    EXPECT(prev_end <= p.image_bytes, "planes beyond image_bytes");
    EXPECT(p.image_bytes % 256 == 0, "image planes not 256-byte aligned");
    EXPECT(g.plane_bytes_per_image == p.image_bytes, "plane bytes");

    // Scale items: reads of rows y0 .. y0+sh+11 and columns x0+2 .. x0+sw+13
    // of the padded plane; the map covers each output pixel once; strips
    // start on even columns and on rows that are multiples of 8.
    std::vector<std::vector<std::vector<uint8_t>>> cover(p.levels);
    for (uint32_t l = 0; l < p.levels; ++l) cover[l].assign(3, std::vector<uint8_t>(static_cast<size_t>(p.w[l]) * p.h[l], 0));
    for (uint32_t i = 0; i < p.items; ++i) {
        const ScaleItem& it = g.items[i];
        EXPECT(it.level < p.levels && it.channel < 3, "item out of range");
        EXPECT(p.mask[it.level] & (1u << it.channel), "item on a pruned plane");
        EXPECT(i >= p.item_begin[it.level][it.channel] &&
                   i < p.item_begin[it.level][it.channel] + p.item_count[it.level][it.channel],
               "item outside its range");
        const uint64_t last_row = it.y0 + sh + 2 * kRadius - 1;
        const uint64_t last_col = it.x0 + 2 + sw + 2 * kRadius - 1;
        EXPECT(last_row < p.rows[it.level], "scale reads past the last row");
        EXPECT(last_col < p.pitch[it.level], "scale reads past the pitch");
        EXPECT(it.x0 % 2 == 0 && it.y0 % kSumBlockRows == 0, "strip origin off the block grid");
        for (uint32_t y = it.y0; y < it.y0 + sh && y < p.h[it.level]; ++y)
            for (uint32_t x = it.x0; x < it.x0 + sw && x < p.w[it.level]; ++x)
                cover[it.level][it.channel][static_cast<size_t>(y) * p.w[it.level] + x]++;
    }
    for (uint32_t l = 0; l < p.levels; ++l)
        for (int c = 0; c < 3; ++c) {
            if (!(p.mask[l] & (1u << c))) continue;
            for (uint8_t v : cover[l][c]) EXPECT(v == 1, "pixel covered %u times", v);
        }

    // Prepare: tiles cover the base image; level-l stores of in-plane pixels
    // stay inside the plane's interior; the tile nests levels up to 5.
    EXPECT(p.prep_tiles_x * kPrepTileW >= width && p.prep_tiles_y * kPrepTileH >= height, "tiles");
    for (uint32_t l = 0; l < p.levels; ++l) {
        EXPECT((kPrepTileW >> l) >= 1 && (kPrepTileH >> l) >= 1, "tile does not nest level %u", l);
        const uint64_t at = static_cast<uint64_t>(p.h[l] - 1 + kPadTop) * p.pitch[l] + p.w[l] - 1 + kPadLeft;
        EXPECT(at < static_cast<uint64_t>(p.pitch[l]) * p.rows[l], "prepare store out of plane");
        if (l > 0) {
            EXPECT(p.w[l] == (p.w[l - 1] + 1) / 2 && p.h[l] == (p.h[l - 1] + 1) / 2, "level sizes");
        }
    }
    EXPECT(p.plane_stride == (static_cast<uint64_t>(width) * height + 3) / 4 * 4, "plane stride");
}

int main() {
    const PlaneFormat fmts[] = {kPlaneF32, kPlaneF16, kPlaneBF16, kPlaneU16};
    std::vector<std::pair<uint32_t, uint32_t>> sizes = {
        {8, 8}, {13, 9}, {32, 32}, {64, 64}, {65, 65}, {129, 97}, {512, 512}, {1920, 1080},
        {3840, 2160}, {640, 360}, {1280, 720}, {2560, 1440}, {127, 8}, {8, 300}, {255, 257}};
    for (uint32_t w = 8; w < 80; w += 7)
        for (uint32_t h = 8; h < 80; h += 9) sizes.push_back({w, h});
    for (auto [w, h] : sizes)
        for (PlaneFormat f : fmts)
            for (uint32_t sh : {0u, 8u, 16u, 64u, 104u, 128u, 256u})
                for (uint32_t sw : {0u, 64u, 128u}) check(w, h, f, sh, sw);
    // Invalid geometry and strip heights are rejected.
    bool threw = false;
    try {
        BuildPlan(7, 8, kPlaneF32, 0);
    } catch (const Error& e) {
        threw = e.status == VQM_INVALID_ARGUMENT;
    }
    EXPECT(threw, "7x8 accepted");
    threw = false;
    try {
        BuildPlan(64, 64, kPlaneF32, 6);
    } catch (const Error& e) {
        threw = e.status == VQM_INVALID_ARGUMENT;
    }
    EXPECT(threw, "strip_rows 6 accepted");
    threw = false;
    try {
        BuildPlan(64, 64, kPlaneF32, 4);
    } catch (const Error& e) {
        threw = e.status == VQM_INVALID_ARGUMENT;
    }
    EXPECT(threw, "strip_rows 4 accepted");
    threw = false;
    try {
        BuildPlan(64, 64, kPlaneF32, 64, 96);
    } catch (const Error& e) {
        threw = e.status == VQM_INVALID_ARGUMENT;
    }
    EXPECT(threw, "strip_cols 96 accepted");

    // Default strip heights on a 256-CU device: 1080p has 163 / 254 / 463 /
    // 893 / 1769 work items per pair at 192 / 128 / 64 / 32 / 16 rows, 4K
    // 1752 at 64 rows.
    struct Case {
        uint32_t w, h;
        uint64_t pairs;
        uint32_t rows;
    };
    for (const Case& c : {Case{1920, 1080, 300, 192}, Case{1920, 1080, 16, 128}, Case{1920, 1080, 64, 192}, Case{1920, 1080, 4, 64},
                          Case{1920, 1080, 2, 32}, Case{1920, 1080, 1, 16}, Case{3840, 2160, 1, 64},
                          Case{3840, 2160, 50, 192}, Case{64, 64, 1, 16}}) {
        const uint32_t got = DefaultStripRows(c.w, c.h, kPlaneF32, 0, c.pairs, 256);
        EXPECT(got == c.rows, "%ux%u x %llu pairs: %u rows, want %u", c.w, c.h,
               static_cast<unsigned long long>(c.pairs), got, c.rows);
    }
    EXPECT(SinglePairStripRows(1920, 1080, kPlaneF32, 0, 0) == 16, "1080p pipeline default");
    EXPECT(SinglePairStripRows(3840, 2160, kPlaneF32, 0, 0) == 64, "4K pipeline default");
    EXPECT(SinglePairStripRows(1920, 1080, kPlaneF32, 40, 0) == 40, "an explicit height is kept");
    if (failures) {
        std::fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    std::printf("plan bounds ok over %zu geometries\n", sizes.size());
    return 0;
}
