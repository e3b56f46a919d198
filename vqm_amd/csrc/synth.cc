// Deterministic synthetic image pairs for the SSIMULACRA 2 workloads.
//
// Host-only C++ (no GPU). The reference has no generator of its own, so the
// bench and the tests share this one; every value is a pure function of
// (config, pair, width, height, strength) so any machine regenerates the
// same bytes.
//
// Reference image, per channel c independently:
//   v = 128 + sum_{k<8} 16 sin(2 pi (x cos t_k + y sin t_k) / P_k + f_k) + U(-6, 6)
//   t_k ~ U[0, pi), f_k ~ U[0, 2 pi), P_k log-uniform in [8, pmax] px,
//   rounded half away from zero, clamped to 0..255.
// Distorted image: per channel and 8x8 block aligned at (0, 0) (edge blocks
// use their own pixels), p' = p - m + step * round(m / step) with m the block
// mean, then + sigma * N(0, 1) by Box-Muller, rounded and clamped.
//
// Random streams: xoshiro256** seeded by SplitMix64 from a hash of
// (config, pair, stream, channel, row); stream 0 = sinusoid parameters,
// 1 = uniform content noise (per row), 2 = Gaussian distortion noise (per
// row), 3 = per-pair draws of the batch/serving workloads. Rows are
// independent streams, so the images are generated in parallel and stay
// identical for any thread count.
//
// Images are packed planar u8 R, G, B, each plane starting on a 4-byte
// boundary (the Image layout of the reference, vqm_image.h:122-133).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t combine(uint64_t h, uint64_t v) {
    uint64_t s = h ^ (v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2));
    return splitmix64(s);
}

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        for (auto& w : s) w = splitmix64(seed);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    // Uniform in [0, 1) with 53 random bits.
    double u01() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

uint64_t stream_seed(uint32_t cfg, uint32_t pair, uint32_t stream, uint32_t channel,
                     uint32_t row) {
    uint64_t h = 0x5353494D55324152ull;  // "SSIMU2AR"
    h = combine(h, cfg);
    h = combine(h, pair);
    h = combine(h, stream);
    h = combine(h, channel);
    h = combine(h, row);
    return h;
}

double round_half_away(double v) {
    return v >= 0.0 ? std::floor(v + 0.5) : -std::floor(-v + 0.5);
}

uint8_t to_u8(double v) {
    const double r = round_half_away(v);
    return static_cast<uint8_t>(std::min(255.0, std::max(0.0, r)));
}

uint64_t plane_stride(uint32_t w, uint32_t h) {
    return (static_cast<uint64_t>(w) * h + 3u) / 4u * 4u;
}

template <typename F>
void parallel_rows(uint32_t rows, uint32_t threads, F&& body) {
    if (threads <= 1 || rows < 64) {
        for (uint32_t y = 0; y < rows; ++y) body(y);
        return;
    }
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; ++t) {
        pool.emplace_back([&, t]() {
            for (uint32_t y = t; y < rows; y += threads) body(y);
        });
    }
    for (auto& th : pool) th.join();
}

void reference_channel(uint32_t cfg, uint32_t pair, uint32_t c, uint32_t w, uint32_t h,
                       double pmax, uint8_t* out, uint32_t threads) {
    Xoshiro params(stream_seed(cfg, pair, 0, c, 0));
    double ct[8], st[8], ph[8], per[8];
    const double lo = std::log(8.0), hi = std::log(pmax);
    for (int k = 0; k < 8; ++k) {
        const double theta = kPi * params.u01();
        ph[k] = 2.0 * kPi * params.u01();
        per[k] = std::exp(lo + params.u01() * (hi - lo));
        ct[k] = std::cos(theta);
        st[k] = std::sin(theta);
    }
    // sin(A(x) + B(y)) = sin A cos B + cos A sin B, A and B tabulated.
    std::vector<double> sa(8 * static_cast<size_t>(w)), ca(8 * static_cast<size_t>(w));
    for (int k = 0; k < 8; ++k) {
        for (uint32_t x = 0; x < w; ++x) {
            const double a = 2.0 * kPi * static_cast<double>(x) * ct[k] / per[k];
            sa[k * w + x] = std::sin(a);
            ca[k * w + x] = std::cos(a);
        }
    }
    parallel_rows(h, threads, [&](uint32_t y) {
        double sb[8], cb[8];
        for (int k = 0; k < 8; ++k) {
            const double b = 2.0 * kPi * static_cast<double>(y) * st[k] / per[k] + ph[k];
            sb[k] = std::sin(b);
            cb[k] = std::cos(b);
        }
        Xoshiro noise(stream_seed(cfg, pair, 1, c, y));
        uint8_t* row = out + static_cast<size_t>(y) * w;
        for (uint32_t x = 0; x < w; ++x) {
            double s = 0.0;
            for (int k = 0; k < 8; ++k)
                s += 16.0 * (sa[k * w + x] * cb[k] + ca[k * w + x] * sb[k]);
            const double v = 128.0 + s + (12.0 * noise.u01() - 6.0);
            row[x] = to_u8(v);
        }
    });
}

void distort_channel(uint32_t cfg, uint32_t pair, uint32_t c, uint32_t w, uint32_t h,
                     const uint8_t* ref, int step, double sigma, uint8_t* out,
                     uint32_t threads) {
    // Block means of the 8x8 grid.
    const uint32_t bw = (w + 7) / 8, bh = (h + 7) / 8;
    std::vector<double> shift(static_cast<size_t>(bw) * bh);
    for (uint32_t by = 0; by < bh; ++by) {
        for (uint32_t bx = 0; bx < bw; ++bx) {
            double sum = 0.0;
            uint32_t count = 0;
            for (uint32_t y = by * 8; y < std::min(h, by * 8 + 8); ++y)
                for (uint32_t x = bx * 8; x < std::min(w, bx * 8 + 8); ++x) {
                    sum += ref[static_cast<size_t>(y) * w + x];
                    ++count;
                }
            const double m = sum / count;
            shift[static_cast<size_t>(by) * bw + bx] =
                step > 0 ? static_cast<double>(step) * round_half_away(m / step) - m : 0.0;
        }
    }
    parallel_rows(h, threads, [&](uint32_t y) {
        Xoshiro noise(stream_seed(cfg, pair, 2, c, y));
        const uint8_t* src = ref + static_cast<size_t>(y) * w;
        uint8_t* row = out + static_cast<size_t>(y) * w;
        double spare = 0.0;
        bool have = false;
        for (uint32_t x = 0; x < w; ++x) {
            double z;
            if (have) {
                z = spare;
                have = false;
            } else {
                const double u1 = 1.0 - noise.u01();  // (0, 1]
                const double u2 = noise.u01();
                const double r = std::sqrt(-2.0 * std::log(u1));
                z = r * std::cos(2.0 * kPi * u2);
                spare = r * std::sin(2.0 * kPi * u2);
                have = true;
            }
            const double p = src[x] + shift[static_cast<size_t>(y / 8) * bw + x / 8];
            row[x] = to_u8(p + sigma * z);
        }
    });
}

}  // namespace

extern "C" {

uint64_t vqm_synth_image_bytes(uint32_t width, uint32_t height) {
    return 3 * plane_stride(width, height);
}

// One reference image (packed planar RGB, vqm_synth_image_bytes long).
void vqm_synth_reference(uint32_t cfg, uint32_t pair, uint32_t width, uint32_t height,
                         double pmax, uint8_t* out, uint32_t threads) {
    const uint64_t stride = plane_stride(width, height);
    std::memset(out, 0, 3 * stride);
    for (uint32_t c = 0; c < 3; ++c)
        reference_channel(cfg, pair, c, width, height, pmax, out + c * stride, threads);
}

// The distorted partner of `ref` at one strength (block step, noise sigma).
void vqm_synth_distort(uint32_t cfg, uint32_t pair, uint32_t width, uint32_t height,
                       const uint8_t* ref, int32_t step, double sigma, uint8_t* out,
                       uint32_t threads) {
    const uint64_t stride = plane_stride(width, height);
    std::memset(out, 0, 3 * stride);
    for (uint32_t c = 0; c < 3; ++c)
        distort_channel(cfg, pair, c, width, height, ref + c * stride, step, sigma,
                        out + c * stride, threads);
}

// Uniform [0, 1) draw k of the per-pair stream 3 (size and strength choices
// of the batch and serving workloads).
double vqm_synth_draw(uint32_t cfg, uint32_t pair, uint32_t k) {
    Xoshiro r(stream_seed(cfg, pair, 3, 0, 0));
    double v = 0.0;
    for (uint32_t i = 0; i <= k; ++i) v = r.u01();
    return v;
}

// FNV-1a over a byte range, for recording fixture hashes.
uint64_t vqm_synth_hash(const uint8_t* data, uint64_t bytes) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (uint64_t i = 0; i < bytes; ++i) {
        h ^= data[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

}  // extern "C"
