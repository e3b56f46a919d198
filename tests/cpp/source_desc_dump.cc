// Prints the decode descriptor the prepare kernel receives for one image
// format, as one JSON object, so the host colorimetry can be checked against
// an independent FP64 restatement without a GPU.
//
// usage: source_desc_dump type family alpha range matrix transfer primaries
//                          location sub_w sub_h depth width height filter
#include <cstdio>
#include <cstdlib>

#include "ss2_host.h"
#include "ss2_source.h"

using namespace vqm::ss2;

static void Array(const char* name, const float* v, int n) {
    std::printf("\"%s\": [", name);
    for (int i = 0; i < n; ++i) std::printf("%s%.9g", i ? ", " : "", v[i]);
    std::printf("], ");
}

int main(int argc, char** argv) {
    if (argc != 15) {
        std::fprintf(stderr, "expected 14 arguments\n");
        return 2;
    }
    int a[14];
    for (int i = 0; i < 14; ++i) a[i] = std::atoi(argv[i + 1]);
    vqm_image_format f{};
    f.type = a[0], f.family = a[1], f.alpha = a[2], f.range = a[3], f.matrix = a[4];
    f.transfer = a[5], f.primaries = a[6], f.chroma_location = a[7];
    f.subsample_w = a[8], f.subsample_h = a[9], f.depth = a[10];
    try {
        ValidateFormat(f, a[11], a[12], "image");
        const SourceDesc d = BuildSourceDesc(f, a[11], a[12], a[13]);
        std::printf("{\"kind\": %u, \"type\": %u, \"family\": %u, \"alpha\": %u, \"sub_w\": %u, "
                    "\"sub_h\": %u, \"transfer\": %u, \"mode\": %u, \"filter\": %u, "
                    "\"identity1\": %u, \"identity2\": %u, ",
                    d.kind, d.type, d.family, d.alpha, d.sub_w, d.sub_h, d.transfer, d.mode,
                    d.filter, d.identity1, d.identity2);
        std::printf("\"plane_off\": [%llu, %llu, %llu, %llu], ",
                    (unsigned long long)d.plane_off[0], (unsigned long long)d.plane_off[1],
                    (unsigned long long)d.plane_off[2], (unsigned long long)d.plane_off[3]);
        std::printf("\"plane_w\": [%u, %u, %u, %u], \"plane_h\": [%u, %u, %u, %u], ", d.plane_w[0],
                    d.plane_w[1], d.plane_w[2], d.plane_w[3], d.plane_h[0], d.plane_h[1],
                    d.plane_h[2], d.plane_h[3]);
        Array("scale", d.scale, 3);
        Array("offset", d.offset, 3);
        Array("m1", d.m1, 9);
        Array("m2", d.m2, 9);
        const float k[4] = {d.kr, d.kg, d.kb, d.to_linear_scale};
        Array("k", k, 4);
        const float shift[2] = {d.shift_x, d.shift_y};
        Array("shift", shift, 2);
        Array("cl", d.cl, 4);
        std::printf("\"srgb8\": %d}\n", IsSrgb8(f) ? 1 : 0);
    } catch (const Error& e) {
        std::printf("{\"error\": \"%s\"}\n", e.message.c_str());
    }
    return 0;
}
