// Decodes one file with the scoring command's loaders and prints what it
// got as one JSON object per frame (sizes, format fields, the packed plane
// bytes as hex), or {"error": ...}. usage: image_io_dump [--sequence] path
#include <cstdio>
#include <cstring>
#include <string>

#include "image_io.h"

static void Dump(const vqm::Image& img) {
    const auto& f = img.format;
    std::printf("{\"width\": %zu, \"height\": %zu, \"type\": %d, \"family\": %d, \"alpha\": %d, "
                "\"range\": %d, \"matrix\": %d, \"transfer\": %d, \"primaries\": %d, \"location\": %d, "
                "\"sub_w\": %d, \"sub_h\": %d, \"depth\": %d, \"data\": \"",
                img.size.width, img.size.height, int(f.type), int(f.family), int(f.alpha), int(f.range),
                int(f.matrix), int(f.transfer), int(f.primaries), int(f.chroma_location), int(f.subsample_w),
                int(f.subsample_h), int(f.depth));
    for (unsigned char b : img.data) std::printf("%02x", b);
    std::printf("\"}\n");
}

int main(int argc, char** argv) {
    const bool sequence = argc == 3 && std::strcmp(argv[1], "--sequence") == 0;
    const std::string path = argv[argc - 1];
    if (!sequence) {
        std::string error;
        const auto img = vqm::image_io::LoadImage(path, error);
        if (!img) {
            std::printf("{\"error\": \"%s\"}\n", error.c_str());
            return 0;
        }
        Dump(*img);
        return 0;
    }
    try {
        vqm::image_io::Y4mReader reader(path);
        std::printf("{\"rate\": [%u, %u]}\n", reader.rate().numerator, reader.rate().denominator);
        vqm::Image img;
        while (reader.Next(img)) Dump(img);
        std::printf("{\"end\": true}\n");
    } catch (const std::exception& e) {
        std::printf("{\"error\": \"%s\"}\n", e.what());
    }
    return 0;
}
