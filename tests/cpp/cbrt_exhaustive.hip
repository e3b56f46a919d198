// Exhaustive check of the XYB cube roots on the CPU.
//
//   cbrt_exhaustive <oracle .so> <first float bits> <end float bits>
//
// For every float x with bits in [first, end) (positive normal floats):
//   * vqm::ss2::cbrt_rn(x) (vqm_amd/csrc/ss2_cbrt.h, the device function
//     compiled for the host; every step is an IEEE operation, so the host
//     result is the device result) is the correctly rounded cube root:
//     the true root lies strictly between the cubes' midpoints on either
//     side of the result, decided in exact 128-bit integer arithmetic;
//   * the oracle's ss2o_cbrt_rn(x) returns the same float;
//   * vqm::ss2::cbrt_fast(x) is within 0.84 ulp (measured in FP64 against
//     the correctly rounded root, as a bound check of the fast root).
// Prints the counts and exits 1 on any failure.
#include <dlfcn.h>

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "ss2_cbrt.h"

namespace {

uint32_t bits_of(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
float float_of(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// Positive normal float as (significand with the hidden bit, exponent of its
// last bit).
void split(float f, uint64_t* sig, int* exp) {
    const uint32_t u = bits_of(f);
    *sig = (u & 0x7FFFFFu) | 0x800000u;
    *exp = static_cast<int>(u >> 23) - 150;
}

// Sign of ((a + b) / 2)^3 - x for adjacent positive normal floats a < b.
int mid_cube_sign(float a, float b, float x) {
    uint64_t sa, sb, sx;
    int ea, eb, ex;
    split(a, &sa, &ea);
    split(b, &sb, &eb);
    split(x, &sx, &ex);
    // a + b = (sa 2^(ea - e) + sb 2^(eb - e)) 2^e with e = min(ea, eb);
    // the midpoint is that over 2: M 2^(e - 1).
    const int e = ea < eb ? ea : eb;
    const unsigned __int128 m = (static_cast<unsigned __int128>(sa) << (ea - e)) +
                                (static_cast<unsigned __int128>(sb) << (eb - e));
    const unsigned __int128 m3 = m * m * m;  // < 2^78
    const int shift = 3 * (e - 1) - ex;      // mid^3 - x = 2^ex (m3 2^shift - sx)
    unsigned __int128 lhs = m3, rhs = sx;
    if (shift >= 0) {
        if (shift > 45) return 1;
        lhs <<= shift;
    } else {
        if (-shift > 100) return -1;
        rhs <<= -shift;
    }
    return lhs > rhs ? 1 : (lhs < rhs ? -1 : 0);
}

bool correctly_rounded(float x, float y) {
    if (!(y > 0.0f) || !std::isfinite(y)) return false;
    const float up = float_of(bits_of(y) + 1), down = float_of(bits_of(y) - 1);
    // root < upper midpoint  <=>  x < mid_up^3; root > lower midpoint likewise.
    if (mid_cube_sign(y, up, x) <= 0) return false;
    if (bits_of(y) > 0x00800000u && mid_cube_sign(down, y, x) >= 0) return false;
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s <oracle.so> <first bits> <end bits>\n", argv[0]);
        return 2;
    }
    void* h = dlopen(argv[1], RTLD_NOW);
    if (h == nullptr) {
        std::fprintf(stderr, "dlopen: %s\n", dlerror());
        return 2;
    }
    using Fn = float (*)(float);
    const Fn oracle_rn = reinterpret_cast<Fn>(dlsym(h, "ss2o_cbrt_rn"));
    if (oracle_rn == nullptr) {
        std::fprintf(stderr, "ss2o_cbrt_rn missing\n");
        return 2;
    }
    const uint32_t first = static_cast<uint32_t>(std::strtoul(argv[2], nullptr, 0));
    const uint32_t end = static_cast<uint32_t>(std::strtoul(argv[3], nullptr, 0));
    if (first < 0x00800000u || end > 0x7F800000u || first >= end) {
        std::fprintf(stderr, "range must lie within the positive normal floats\n");
        return 2;
    }

    unsigned threads = std::thread::hardware_concurrency();
    if (threads == 0) threads = 1;
    if (threads > 16) threads = 16;
    constexpr uint32_t kChunk = 1u << 20;
    std::atomic<uint32_t> next{first};
    std::atomic<uint64_t> bad_rn{0}, bad_oracle{0}, checked{0};
    std::atomic<uint32_t> first_bad{0};
    std::vector<double> worst(threads, 0.0);
    auto work = [&](unsigned t) {
        double w = 0.0;
        for (;;) {
            const uint32_t lo = next.fetch_add(kChunk);
            if (lo >= end) break;
            const uint32_t hi = end - lo < kChunk ? end : lo + kChunk;
            uint64_t n = 0;
            for (uint32_t u = lo; u < hi; ++u) {
                const float x = float_of(u);
                const float y = vqm::ss2::cbrt_rn(x);
                if (!correctly_rounded(x, y)) {
                    if (bad_rn.fetch_add(1) == 0) first_bad = u;
                }
                if (oracle_rn(x) != y) bad_oracle.fetch_add(1);
                // cbrt_fast's error in ulps of the rounded root.
                const float f = vqm::ss2::cbrt_fast(x);
                const double ulp = std::ldexp(1.0, std::ilogb(y) - 23);
                const double err = std::fabs(static_cast<double>(f) - std::cbrt(static_cast<double>(x))) / ulp;
                if (err > w) w = err;
                ++n;
            }
            checked.fetch_add(n);
        }
        worst[t] = w;
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t) pool.emplace_back(work, t);
    for (auto& t : pool) t.join();
    double w = 0.0;
    for (double v : worst) w = v > w ? v : w;
    std::printf("checked %llu floats in [%a, %a): cbrt_rn wrong %llu, oracle differs %llu, "
                "cbrt_fast max error %.4f ulp\n",
                static_cast<unsigned long long>(checked.load()), static_cast<double>(float_of(first)),
                static_cast<double>(float_of(end)), static_cast<unsigned long long>(bad_rn.load()),
                static_cast<unsigned long long>(bad_oracle.load()), w);
    if (bad_rn.load() != 0) std::printf("first wrong input: %a\n", static_cast<double>(float_of(first_bad.load())));
    return bad_rn.load() == 0 && bad_oracle.load() == 0 ? 0 : 1;
}
