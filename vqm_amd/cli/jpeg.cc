// JPEG decoding for the scoring command. The reference reads JPEG through
// libjpeg with its default decompression set-up (tools/image_io/io.cc:
// 277-355): 1 or 3 components, RGB or grey output, islow IDCT, fancy
// upsampling, the JFIF YCbCr to RGB conversion, and an error for entropy
// data that runs into a marker. This decoder is written from ITU-T T.81
// (baseline and extended sequential Huffman, progressive Huffman with
// spectral selection and successive approximation, restart intervals) and
// reproduces that set-up's integer arithmetic, so it yields the same 8-bit
// samples libjpeg-turbo does:
//   * the islow inverse DCT: the Loeffler-Ligtenberg-Moschytz factorisation
//     with 13-bit constants, two passes with 2 extra bits between them,
//     rounding shifts, and the post-IDCT range table (index masked to 10
//     bits);
//   * upsampling: the triangle filters for 2:1 horizontal (3/4 and 1/4 of
//     the nearer and farther sample, biases 1 and 2), 2:1 vertical and 2x2
//     (column sums 3 this + 1 neighbour, biases 8 and 7 over 16), edge
//     samples replicated, rows above the image taken from its first row and
//     rows below from its last; box replication for other integral factors
//     and for 2:1 horizontal factors on widths of at most two samples;
//   * YCbCr to RGB with 16-bit fixed-point tables rounded to nearest;
//   * the colour space of three components: JFIF marker means YCbCr, an
//     Adobe marker's transform 0 means RGB and 1 YCbCr, otherwise component
//     ids 'R','G','B' mean RGB and anything else YCbCr.
// Arithmetic-coded, lossless, hierarchical and 12-bit files are refused.
#include "image_io.h"

#include <algorithm>
#include <array>
#include <cstring>
#include <string>
#include <vector>

namespace vqm::image_io {
namespace {

struct JpegFailure {
    std::string message;
};

[[noreturn]] void Fail(std::string message) { throw JpegFailure{std::move(message)}; }

constexpr int kMaxDimension = 65500;

// Natural (row-major) position of the k-th coefficient in zigzag order
// (T.81 figure A.6), with 16 extra entries so that a corrupt run past the
// block's end lands on the last coefficient instead of outside it.
constexpr std::array<int, 80> kNatural = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33,
    40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36,
    29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54,
    47, 55, 62, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

// ---- entropy decoding ---------------------------------------------------------

// A canonical Huffman table (T.81 annex C), decoded with a 9-bit lookahead
// and the per-length code ranges of figure F.16 beyond it.
struct Huffman {
    bool defined = false;
    std::array<std::uint8_t, 256> values{};
    std::array<std::int32_t, 18> maxcode{};
    std::array<std::int32_t, 17> valptr{}, mincode{};
    std::array<std::uint16_t, 512> fast{};  // (length << 8) | value, 0 = longer code
};

void BuildHuffman(Huffman& h, const std::uint8_t counts[16], const std::uint8_t* values, int n, bool dc) {
    h = Huffman{};
    h.defined = true;
    if (n < 0 || n > 256) Fail("Bogus Huffman table definition");
    std::copy(values, values + n, h.values.begin());
    // First pass: the canonical code ranges per length. Every length's codes
    // must fit in `len` bits with room for one more (no code is all ones),
    // checked before any code is used as an index.
    std::int32_t code = 0;
    int k = 0;
    for (int len = 1; len <= 16; ++len) {
        const int c = counts[len - 1];
        h.valptr[len] = k;
        h.mincode[len] = code;
        h.maxcode[len] = c ? code + c - 1 : -1;
        code += c;
        k += c;
        if (code >= (1 << len)) Fail("Bogus Huffman table definition");
        code <<= 1;
    }
    if (k != n) Fail("Bogus Huffman table definition");
    h.maxcode[17] = 0x7FFFFFFF;  // sentinel
    // Second pass: the 9-bit lookahead table, from ranges already checked
    // (code << (9 - len) | j stays below 512 because code < 2^len).
    for (int len = 1; len <= 9; ++len) {
        const int shift = 9 - len;
        for (std::int32_t cd = h.mincode[len]; cd <= h.maxcode[len]; ++cd) {
            const int v = values[h.valptr[len] + (cd - h.mincode[len])];
            for (int j = 0; j < (1 << shift); ++j)
                h.fast[static_cast<std::size_t>((cd << shift) | j)] = static_cast<std::uint16_t>((len << 8) | v);
        }
    }
    if (dc)
        for (int i = 0; i < n; ++i)
            if (values[i] > 15) Fail("Bogus Huffman table definition");
}

// Entropy-coded bytes of one scan: 0xFF 0x00 is a data byte 0xFF, fill
// bytes 0xFF before a marker are skipped, and the first other marker ends
// the data. Reading a bit past the end is the reference's error
// (io.cc:264-273 turns libjpeg's JWRN_HIT_MARKER into a failure).
class BitReader {
public:
    BitReader(const std::uint8_t* p, const std::uint8_t* end) : p_(p), end_(end) {}

    std::uint32_t Peek(int n) {
        if (bits_ < n) Fill();
        return static_cast<std::uint32_t>(buf_ >> (64 - n));
    }
    void Skip(int n) {
        if (bits_ < n) Fill();
        if (bits_ < n) Fail("Corrupt JPEG data: premature end of data segment");
        buf_ <<= n;
        bits_ -= n;
    }
    std::uint32_t Get(int n) {
        if (n == 0) return 0;
        const std::uint32_t v = Peek(n);
        Skip(n);
        return v;
    }
    int Decode(const Huffman& h) {
        const std::uint32_t look = Peek(9);
        const std::uint16_t f = h.fast[look];
        if (f != 0) {
            Skip(f >> 8);
            return f & 0xFF;
        }
        std::int32_t code = static_cast<std::int32_t>(Peek(9));
        int len = 9;
        for (;;) {
            if (code <= h.maxcode[len] && h.maxcode[len] >= 0) break;
            ++len;
            if (len > 16) Fail("Corrupt JPEG data: bad Huffman code");
            code = static_cast<std::int32_t>(Peek(len));
        }
        Skip(len);
        return h.values[static_cast<std::size_t>(h.valptr[len] + code - h.mincode[len])];
    }
    // Restart (T.81 F.1.2.3, RSTm): drop the buffered bits, then the next
    // marker must be RSTm with the expected m.
    void Restart(int expected) {
        buf_ = 0;
        bits_ = 0;
        if (!marker_) {
            while (p_ < end_ && !(p_[0] == 0xFF && p_ + 1 < end_ && p_[1] != 0x00 && p_[1] != 0xFF)) ++p_;
            marker_ = p_;
        }
        if (marker_ + 1 >= end_ || marker_[1] != 0xD0 + expected) Fail("Corrupt JPEG data: bad restart marker");
        p_ = marker_ + 2;
        marker_ = nullptr;
    }
    // Where parsing resumes after the scan: the marker that ended it, or the
    // first unread byte.
    const std::uint8_t* Position() const { return marker_ ? marker_ : p_; }

private:
    void Fill() {
        while (bits_ <= 56 && !marker_ && p_ < end_) {
            std::uint8_t b = *p_;
            if (b == 0xFF) {
                const std::uint8_t* q = p_ + 1;
                while (q < end_ && *q == 0xFF) ++q;
                if (q >= end_) {
                    marker_ = p_;
                    break;
                }
                if (*q != 0x00) {
                    marker_ = q - 1;
                    break;
                }
                p_ = q + 1;
            } else {
                ++p_;
            }
            buf_ |= static_cast<std::uint64_t>(b) << (56 - bits_);
            bits_ += 8;
        }
    }
    const std::uint8_t* p_;
    const std::uint8_t* end_;
    const std::uint8_t* marker_ = nullptr;
    std::uint64_t buf_ = 0;
    int bits_ = 0;
};

int Extend(std::uint32_t v, int s) {
    return static_cast<int>(v) < (1 << (s - 1)) ? static_cast<int>(v) - (1 << s) + 1 : static_cast<int>(v);
}

// ---- frame model -------------------------------------------------------------

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int dc_table = 0, ac_table = 0;
    std::array<std::uint16_t, 64> quant{};  // natural order, latched at its first scan
    bool latched = false;
    int width = 0, height = 0;              // downsampled_width / _height
    int blocks_w = 0, blocks_h = 0;         // blocks holding samples
    int alloc_w = 0;                        // blocks per row in storage (whole MCUs)
    std::vector<std::int16_t> coef;         // 64 per block, natural order
    int dc_pred = 0;
    std::int16_t* Block(int bx, int by) { return coef.data() + (static_cast<std::size_t>(by) * alloc_w + bx) * 64; }
};

struct Frame {
    bool progressive = false;
    int width = 0, height = 0, hmax = 1, vmax = 1, mcus_x = 0, mcus_y = 0;
    std::vector<Component> comps;
};

std::uint16_t Be16(const std::uint8_t* p) { return static_cast<std::uint16_t>((p[0] << 8) | p[1]); }

// ---- inverse DCT (islow) -----------------------------------------------------

constexpr int kConstBits = 13, kPass1Bits = 2;
constexpr std::int64_t F0_298631336 = 2446, F0_390180644 = 3196, F0_541196100 = 4433, F0_765366865 = 6270,
                       F0_899976223 = 7373, F1_175875602 = 9633, F1_501321110 = 12299, F1_847759065 = 15137,
                       F1_961570560 = 16069, F2_053119869 = 16819, F2_562915447 = 20995, F3_072711026 = 25172;

std::int64_t Descale(std::int64_t x, int n) { return (x + (std::int64_t{1} << (n - 1))) >> n; }

// The post-IDCT range table: index x & 1023 of a value x centred on 0 gives
// x + 128 for x in [-128, 127], 255 above and 0 below (wrapping only for
// values no valid stream produces).
std::uint8_t RangeLimit(std::int64_t x) {
    const int m = static_cast<int>(x & 1023);
    if (m < 128) return static_cast<std::uint8_t>(m + 128);
    if (m < 512) return 255;
    if (m < 896) return 0;
    return static_cast<std::uint8_t>(m - 896);
}

// The shared odd-part and even-part butterflies of one 1-D pass: inputs
// in[0..7] in natural order, outputs before the final descale.
void IdctPass(const std::int64_t in[8], std::int64_t out[8]) {
    std::int64_t z2 = in[2], z3 = in[6];
    std::int64_t z1 = (z2 + z3) * F0_541196100;
    const std::int64_t t2 = z1 + z3 * -F1_847759065;
    const std::int64_t t3 = z1 + z2 * F0_765366865;
    const std::int64_t t0 = (in[0] + in[4]) * (std::int64_t{1} << kConstBits);
    const std::int64_t t1 = (in[0] - in[4]) * (std::int64_t{1} << kConstBits);
    const std::int64_t t10 = t0 + t3, t13 = t0 - t3, t11 = t1 + t2, t12 = t1 - t2;
    std::int64_t o0 = in[7], o1 = in[5], o2 = in[3], o3 = in[1];
    z1 = o0 + o3;
    z2 = o1 + o2;
    z3 = o0 + o2;
    std::int64_t z4 = o1 + o3;
    const std::int64_t z5 = (z3 + z4) * F1_175875602;
    o0 *= F0_298631336;
    o1 *= F2_053119869;
    o2 *= F3_072711026;
    o3 *= F1_501321110;
    z1 *= -F0_899976223;
    z2 *= -F2_562915447;
    z3 *= -F1_961570560;
    z4 *= -F0_390180644;
    z3 += z5;
    z4 += z5;
    o0 += z1 + z3;
    o1 += z2 + z4;
    o2 += z2 + z3;
    o3 += z1 + z4;
    out[0] = t10 + o3;
    out[7] = t10 - o3;
    out[1] = t11 + o2;
    out[6] = t11 - o2;
    out[2] = t12 + o1;
    out[5] = t12 - o1;
    out[3] = t13 + o0;
    out[4] = t13 - o0;
}

void IdctBlock(const std::int16_t* coef, const std::uint16_t* q, std::uint8_t* out, std::size_t stride) {
    std::int64_t ws[64];
    for (int c = 0; c < 8; ++c) {
        bool ac_zero = true;
        for (int r = 1; r < 8; ++r) ac_zero &= coef[r * 8 + c] == 0;
        if (ac_zero) {
            const std::int64_t dc = static_cast<std::int64_t>(coef[c]) * q[c] * (1 << kPass1Bits);
            for (int r = 0; r < 8; ++r) ws[r * 8 + c] = dc;
            continue;
        }
        std::int64_t in[8], o[8];
        for (int r = 0; r < 8; ++r) in[r] = static_cast<std::int64_t>(coef[r * 8 + c]) * q[r * 8 + c];
        IdctPass(in, o);
        for (int r = 0; r < 8; ++r) ws[r * 8 + c] = Descale(o[r], kConstBits - kPass1Bits);
    }
    for (int r = 0; r < 8; ++r) {
        const std::int64_t* w = ws + r * 8;
        std::uint8_t* dst = out + r * stride;
        if (w[1] == 0 && w[2] == 0 && w[3] == 0 && w[4] == 0 && w[5] == 0 && w[6] == 0 && w[7] == 0) {
            const std::uint8_t v = RangeLimit(Descale(w[0], kPass1Bits + 3));
            std::fill(dst, dst + 8, v);
            continue;
        }
        std::int64_t o[8];
        IdctPass(w, o);
        for (int x = 0; x < 8; ++x) dst[x] = RangeLimit(Descale(o[x], kConstBits + kPass1Bits + 3));
    }
}

// ---- scans -------------------------------------------------------------------

struct Tables {
    std::array<Huffman, 4> dc, ac;
    std::array<std::array<std::uint16_t, 64>, 4> quant{};  // natural order
    std::array<bool, 4> quant_defined{};
    int restart_interval = 0;
};

struct Scan {
    std::vector<Component*> comps;
    int ss = 0, se = 63, ah = 0, al = 0;
};

class ScanDecoder {
public:
    ScanDecoder(Frame& f, Tables& t, Scan& s, BitReader& br) : f_(f), t_(t), s_(s), br_(br) {}

    void Run() {
        for (Component* c : s_.comps) c->dc_pred = 0;
        const bool single = s_.comps.size() == 1;
        const int mx = single ? s_.comps[0]->blocks_w : f_.mcus_x;
        const int my = single ? s_.comps[0]->blocks_h : f_.mcus_y;
        const long total = static_cast<long>(mx) * my;
        int restart = 0;
        for (long m = 0; m < total; ++m) {
            if (t_.restart_interval && m > 0 && m % t_.restart_interval == 0) {
                br_.Restart(restart);
                restart = (restart + 1) & 7;
                for (Component* c : s_.comps) c->dc_pred = 0;
                eobrun_ = 0;
            }
            const int x = static_cast<int>(m % mx), y = static_cast<int>(m / mx);
            if (single) {
                Block(*s_.comps[0], s_.comps[0]->Block(x, y));
            } else {
                for (Component* c : s_.comps)
                    for (int by = 0; by < c->v; ++by)
                        for (int bx = 0; bx < c->h; ++bx) Block(*c, c->Block(x * c->h + bx, y * c->v + by));
            }
        }
    }

private:
    void Block(Component& c, std::int16_t* b) {
        if (!f_.progressive) return Sequential(c, b);
        if (s_.ss == 0) return s_.ah == 0 ? DcFirst(c, b) : DcRefine(b);
        return s_.ah == 0 ? AcFirst(c, b) : AcRefine(c, b);
    }
    void Sequential(Component& c, std::int16_t* b) {
        const int s = br_.Decode(t_.dc[c.dc_table]);
        const int diff = s ? Extend(br_.Get(s), s) : 0;
        c.dc_pred += diff;
        b[0] = static_cast<std::int16_t>(c.dc_pred);
        const Huffman& ac = t_.ac[c.ac_table];
        for (int k = 1; k < 64; ++k) {
            const int rs = br_.Decode(ac);
            const int r = rs >> 4, sz = rs & 15;
            if (sz) {
                k += r;
                b[kNatural[k]] = static_cast<std::int16_t>(Extend(br_.Get(sz), sz));
            } else {
                if (r != 15) break;
                k += 15;
            }
        }
    }
    void DcFirst(Component& c, std::int16_t* b) {
        const int s = br_.Decode(t_.dc[c.dc_table]);
        const int diff = s ? Extend(br_.Get(s), s) : 0;
        c.dc_pred += diff;
        b[0] = static_cast<std::int16_t>(static_cast<int>(static_cast<unsigned>(c.dc_pred) << s_.al));
    }
    void DcRefine(std::int16_t* b) {
        if (br_.Get(1)) b[0] = static_cast<std::int16_t>(b[0] | (1 << s_.al));
    }
    void AcFirst(Component& c, std::int16_t* b) {
        if (eobrun_ > 0) {
            --eobrun_;
            return;
        }
        const Huffman& ac = t_.ac[c.ac_table];
        for (int k = s_.ss; k <= s_.se; ++k) {
            const int rs = br_.Decode(ac);
            const int r = rs >> 4, s = rs & 15;
            if (s) {
                k += r;
                const int v = Extend(br_.Get(s), s);
                b[kNatural[k]] = static_cast<std::int16_t>(static_cast<int>(static_cast<unsigned>(v) << s_.al));
            } else if (r == 15) {
                k += 15;
            } else {
                eobrun_ = (1 << r) - 1;
                if (r) eobrun_ += static_cast<int>(br_.Get(r));
                break;
            }
        }
    }
    // T.81 G.1.2.3: one correction bit per already-nonzero coefficient on
    // the way, newly nonzero coefficients +-2^Al.
    void Correct(std::int16_t* coef) {
        const int p1 = 1 << s_.al;
        if (br_.Get(1) && (*coef & p1) == 0) *coef = static_cast<std::int16_t>(*coef >= 0 ? *coef + p1 : *coef - p1);
    }
    void AcRefine(Component& c, std::int16_t* b) {
        const int p1 = 1 << s_.al;
        int k = s_.ss;
        if (eobrun_ == 0) {
            const Huffman& ac = t_.ac[c.ac_table];
            for (; k <= s_.se; ++k) {
                const int rs = br_.Decode(ac);
                int r = rs >> 4;
                const int s = rs & 15;
                int value = 0;
                if (s) {
                    value = br_.Get(1) ? p1 : -p1;
                } else if (r != 15) {
                    eobrun_ = 1 << r;
                    if (r) eobrun_ += static_cast<int>(br_.Get(r));
                    break;
                }
                do {
                    std::int16_t* coef = b + kNatural[k];
                    if (*coef != 0) {
                        Correct(coef);
                    } else {
                        if (--r < 0) break;
                    }
                    ++k;
                } while (k <= s_.se);
                if (value) b[kNatural[k]] = static_cast<std::int16_t>(value);
            }
        }
        if (eobrun_ > 0) {
            for (; k <= s_.se; ++k) {
                std::int16_t* coef = b + kNatural[k];
                if (*coef != 0) Correct(coef);
            }
            --eobrun_;
        }
    }

    Frame& f_;
    Tables& t_;
    Scan& s_;
    BitReader& br_;
    int eobrun_ = 0;
};

// ---- upsampling and colour -----------------------------------------------------

// One component's samples at full resolution (hmax/h x vmax/v times its
// own), `out_w` x `out_h` of them.
std::vector<std::uint8_t> Upsample(const Component& c, const std::vector<std::uint8_t>& plane, int hmax, int vmax,
                                   int out_w, int out_h) {
    const int stride = c.blocks_w * 8;
    const int fx = hmax / c.h, fy = vmax / c.v;
    const int dw = c.width, dh = c.height;
    auto at = [&](int x, int y) -> int { return plane[static_cast<std::size_t>(y) * stride + x]; };
    std::vector<std::uint8_t> out(static_cast<std::size_t>(out_w) * out_h);
    auto put = [&](int x, int y, int v) {
        if (x < out_w && y < out_h) out[static_cast<std::size_t>(y) * out_w + x] = static_cast<std::uint8_t>(v);
    };
    if (fx == 1 && fy == 1) {
        for (int y = 0; y < out_h; ++y)
            for (int x = 0; x < out_w; ++x) put(x, y, at(x, y));
    } else if (fx == 2 && fy == 1 && dw > 2) {
        for (int y = 0; y < out_h; ++y) {
            put(0, y, at(0, y));
            put(1, y, (at(0, y) * 3 + at(1, y) + 2) >> 2);
            for (int x = 1; x < dw - 1; ++x) {
                const int v3 = at(x, y) * 3;
                put(2 * x, y, (v3 + at(x - 1, y) + 1) >> 2);
                put(2 * x + 1, y, (v3 + at(x + 1, y) + 2) >> 2);
            }
            put(2 * dw - 2, y, (at(dw - 1, y) * 3 + at(dw - 2, y) + 1) >> 2);
            put(2 * dw - 1, y, at(dw - 1, y));
        }
    } else if (fx == 1 && fy == 2) {
        for (int i = 0; i < dh; ++i) {
            const int above = i == 0 ? 0 : i - 1, below = std::min(i + 1, dh - 1);
            for (int x = 0; x < out_w; ++x) {
                const int v3 = at(x, i) * 3;
                put(x, 2 * i, (v3 + at(x, above) + 1) >> 2);
                put(x, 2 * i + 1, (v3 + at(x, below) + 2) >> 2);
            }
        }
    } else if (fx == 2 && fy == 2 && dw > 2) {
        for (int i = 0; i < dh; ++i) {
            for (int half = 0; half < 2; ++half) {
                const int n = half == 0 ? (i == 0 ? 0 : i - 1) : std::min(i + 1, dh - 1);
                const int y = 2 * i + half;
                auto col = [&](int x) { return at(x, i) * 3 + at(x, n); };
                int last = col(0), cur = last, next = col(1);
                put(0, y, (cur * 4 + 8) >> 4);
                put(1, y, (cur * 3 + next + 7) >> 4);
                last = cur;
                cur = next;
                for (int x = 1; x < dw - 1; ++x) {
                    next = col(x + 1);
                    put(2 * x, y, (cur * 3 + last + 8) >> 4);
                    put(2 * x + 1, y, (cur * 3 + next + 7) >> 4);
                    last = cur;
                    cur = next;
                }
                put(2 * dw - 2, y, (cur * 3 + last + 8) >> 4);
                put(2 * dw - 1, y, (cur * 4 + 7) >> 4);
            }
        }
    } else {
        for (int y = 0; y < out_h; ++y)
            for (int x = 0; x < out_w; ++x) put(x, y, at(x / fx, y / fy));
    }
    return out;
}

// jdcolor-style YCbCr -> RGB: 16-bit fixed point, nearest-integer tables.
struct YccTables {
    std::array<int, 256> cr_r{}, cb_b{};
    std::array<std::int64_t, 256> cr_g{}, cb_g{};
    YccTables() {
        constexpr int kBits = 16;
        constexpr std::int64_t kHalf = std::int64_t{1} << (kBits - 1);
        auto fix = [](double x) { return static_cast<std::int64_t>(x * (1 << kBits) + 0.5); };
        for (int i = 0; i < 256; ++i) {
            const std::int64_t x = i - 128;
            cr_r[i] = static_cast<int>((fix(1.40200) * x + kHalf) >> kBits);
            cb_b[i] = static_cast<int>((fix(1.77200) * x + kHalf) >> kBits);
            cr_g[i] = -fix(0.71414) * x;
            cb_g[i] = -fix(0.34414) * x + kHalf;
        }
    }
};

std::uint8_t Clamp255(int v) { return static_cast<std::uint8_t>(std::clamp(v, 0, 255)); }

// ---- markers -------------------------------------------------------------------

Image Decode(const std::uint8_t* data, std::size_t size) {
    if (size < 2 || data[0] != 0xFF || data[1] != 0xD8) Fail("Not a JPEG file");
    Frame f;
    Tables t;
    bool have_frame = false, jfif = false, adobe = false, done = false;
    int adobe_transform = -1;
    std::size_t pos = 2;
    while (!done) {
        // Next marker: skip anything up to an 0xFF, then fill bytes.
        while (pos < size && data[pos] != 0xFF) ++pos;
        while (pos < size && data[pos] == 0xFF) ++pos;
        if (pos >= size) {
            if (!have_frame) Fail("Premature end of JPEG file");
            break;
        }
        const int marker = data[pos++];
        if (marker == 0xD9) break;  // EOI
        if (marker == 0xD8 || (marker >= 0xD0 && marker <= 0xD7) || marker == 0x01) continue;
        if (pos + 2 > size) Fail("Premature end of JPEG file");
        const std::size_t len = Be16(data + pos);
        if (len < 2 || pos + len > size) Fail("Premature end of JPEG file");
        const std::uint8_t* body = data + pos + 2;
        const std::size_t n = len - 2;
        pos += len;
        switch (marker) {
            case 0xC0: case 0xC1: case 0xC2: {
                if (have_frame) Fail("Invalid JPEG file structure: two SOF markers");
                have_frame = true;
                f.progressive = marker == 0xC2;
                if (n < 6) Fail("Bogus marker length");
                if (body[0] != 8) Fail("Unsupported JPEG data precision " + std::to_string(body[0]));
                f.height = Be16(body + 1);
                f.width = Be16(body + 3);
                const int nc = body[5];
                if (f.height == 0 || f.width == 0) Fail("Empty JPEG image (DNL not supported)");
                if (f.height > kMaxDimension || f.width > kMaxDimension)
                    Fail("Maximum supported image dimension is 65500 pixels");
                if (nc != 1 && nc != 3) Fail("unsupported number of components in JPEG");
                if (n < 6 + 3 * static_cast<std::size_t>(nc)) Fail("Bogus marker length");
                for (int i = 0; i < nc; ++i) {
                    Component c;
                    c.id = body[6 + 3 * i];
                    c.h = body[7 + 3 * i] >> 4;
                    c.v = body[7 + 3 * i] & 15;
                    c.tq = body[8 + 3 * i];
                    if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) Fail("Bogus sampling factors");
                    f.comps.push_back(c);
                }
                for (const Component& c : f.comps) {
                    f.hmax = std::max(f.hmax, c.h);
                    f.vmax = std::max(f.vmax, c.v);
                }
                f.mcus_x = (f.width + 8 * f.hmax - 1) / (8 * f.hmax);
                f.mcus_y = (f.height + 8 * f.vmax - 1) / (8 * f.vmax);
                for (Component& c : f.comps) {
                    if (f.hmax % c.h || f.vmax % c.v) Fail("Fractional sampling not implemented yet");
                    c.width = static_cast<int>((static_cast<long>(f.width) * c.h + f.hmax - 1) / f.hmax);
                    c.height = static_cast<int>((static_cast<long>(f.height) * c.v + f.vmax - 1) / f.vmax);
                    c.blocks_w = (c.width + 7) / 8;
                    c.blocks_h = (c.height + 7) / 8;
                    c.alloc_w = f.mcus_x * c.h;
                    c.coef.assign(static_cast<std::size_t>(c.alloc_w) * f.mcus_y * c.v * 64, 0);
                }
                break;
            }
            case 0xC3: case 0xC5: case 0xC6: case 0xC7: case 0xC9: case 0xCA: case 0xCB: case 0xCD: case 0xCE:
            case 0xCF:
                Fail("Unsupported JPEG process: SOF type 0x" + std::string(1, "0123456789abcdef"[marker >> 4]) +
                     std::string(1, "0123456789abcdef"[marker & 15]));
            case 0xC4: {  // DHT
                std::size_t at = 0;
                while (at < n) {
                    if (at + 17 > n) Fail("Bogus DHT marker length");
                    const int tc = body[at] >> 4, th = body[at] & 15;
                    if (tc > 1 || th > 3) Fail("Bogus DHT index");
                    int count = 0;
                    for (int i = 0; i < 16; ++i) count += body[at + 1 + i];
                    if (count > 256 || at + 17 + count > n) Fail("Bogus Huffman table definition");
                    BuildHuffman(tc ? t.ac[th] : t.dc[th], body + at + 1, body + at + 17, count, tc == 0);
                    at += 17 + count;
                }
                break;
            }
            case 0xDB: {  // DQT
                std::size_t at = 0;
                while (at < n) {
                    const int pq = body[at] >> 4, tq = body[at] & 15;
                    if (tq > 3 || pq > 1) Fail("Bogus DQT index");
                    const std::size_t need = 1 + 64 * (pq + 1);
                    if (at + need > n) Fail("Bogus marker length");
                    for (int k = 0; k < 64; ++k)
                        t.quant[tq][kNatural[k]] = pq ? Be16(body + at + 1 + 2 * k) : body[at + 1 + k];
                    t.quant_defined[tq] = true;
                    at += need;
                }
                break;
            }
            case 0xDD:  // DRI
                if (n != 2) Fail("Bogus marker length");
                t.restart_interval = Be16(body);
                break;
            case 0xE0:  // APP0: JFIF (io.cc relies on libjpeg's examine_app0)
                if (n >= 14 && std::memcmp(body, "JFIF\0", 5) == 0) jfif = true;
                break;
            case 0xEE:  // APP14: Adobe transform flag
                if (n >= 12 && std::memcmp(body, "Adobe", 5) == 0) {
                    adobe = true;
                    adobe_transform = body[11];
                }
                break;
            case 0xDA: {  // SOS
                if (!have_frame) Fail("Invalid JPEG file structure: SOS before SOF");
                if (n < 1) Fail("Bogus marker length");
                const int ns = body[0];
                if (ns < 1 || ns > 4 || n != 4 + 2 * static_cast<std::size_t>(ns)) Fail("Bogus marker length");
                Scan s;
                for (int i = 0; i < ns; ++i) {
                    const int id = body[1 + 2 * i];
                    Component* c = nullptr;
                    for (Component& k : f.comps)
                        if (k.id == id) c = &k;
                    if (c == nullptr) Fail("Invalid component ID in SOS");
                    c->dc_table = body[2 + 2 * i] >> 4;
                    c->ac_table = body[2 + 2 * i] & 15;
                    if (c->dc_table > 3 || c->ac_table > 3) Fail("Bogus DHT index");
                    s.comps.push_back(c);
                }
                s.ss = body[1 + 2 * ns];
                s.se = body[2 + 2 * ns];
                s.ah = body[3 + 2 * ns] >> 4;
                s.al = body[3 + 2 * ns] & 15;
                if (f.progressive) {
                    const bool bad = s.ss > s.se || s.se > 63 || s.al > 13 || (s.ss == 0 && s.se != 0) ||
                                     (s.ss != 0 && ns != 1) || (s.ah != 0 && s.al != s.ah - 1);
                    if (bad) Fail("Invalid progressive parameters");
                } else if (s.ss != 0 || s.se != 63 || s.ah != 0 || s.al != 0) {
                    // Sequential scans ignore these fields (libjpeg only warns).
                    s.ss = 0;
                    s.se = 63;
                    s.ah = s.al = 0;
                }
                int blocks = 0;
                for (Component* c : s.comps) {
                    blocks += c->h * c->v;
                    if (!c->latched) {
                        if (!t.quant_defined[c->tq]) Fail("Quantization table 0x0" + std::to_string(c->tq) +
                                                          " was not defined");
                        c->quant = t.quant[c->tq];
                        c->latched = true;
                    }
                    const bool needs_dc = !f.progressive || (s.ss == 0 && s.ah == 0);
                    const bool needs_ac = !f.progressive ? true : s.ss != 0;
                    if (needs_dc && !t.dc[c->dc_table].defined) Fail("Huffman table 0x0" +
                                                                     std::to_string(c->dc_table) + " was not defined");
                    if (needs_ac && !t.ac[c->ac_table].defined) Fail("Huffman table 0x0" +
                                                                     std::to_string(c->ac_table) + " was not defined");
                }
                if (ns > 1 && blocks > 10) Fail("Sampling factors too large for interleaved scan");
                BitReader br(data + pos, data + size);
                ScanDecoder(f, t, s, br).Run();
                pos = static_cast<std::size_t>(br.Position() - data);
                break;
            }
            default:
                break;  // APPn, COM and others carry nothing the output depends on
        }
    }
    if (!have_frame) Fail("JPEG datastream contains no image");
    // A component no scan reached keeps zero coefficients and a zero
    // quantisation table: its samples are all 128, as libjpeg outputs them.

    // IDCT every block that holds samples.
    std::vector<std::vector<std::uint8_t>> planes;
    for (Component& c : f.comps) {
        const std::size_t stride = static_cast<std::size_t>(c.blocks_w) * 8;
        std::vector<std::uint8_t> plane(stride * c.blocks_h * 8);
        for (int by = 0; by < c.blocks_h; ++by)
            for (int bx = 0; bx < c.blocks_w; ++bx)
                IdctBlock(c.Block(bx, by), c.quant.data(), plane.data() + by * 8 * stride + bx * 8, stride);
        planes.push_back(std::move(plane));
    }

    Image img;
    img.size = {static_cast<std::size_t>(f.width), static_cast<std::size_t>(f.height)};
    img.format = ImageFormat{};
    img.format.type = PixelType::Byte;
    img.format.family = f.comps.size() == 1 ? ColorFamily::Grey : ColorFamily::Rgb;
    img.format.matrix = f.comps.size() == 1 ? MatrixCoefficients::Bt709 : MatrixCoefficients::Rgb;
    img.format.transfer = TransferCharacteristics::Srgb;
    img.format.primaries = ColorPrimaries::Bt709;
    img.data.assign(RequiredBytes(img.format, img.size), 0);
    const std::size_t px = static_cast<std::size_t>(f.width) * f.height;
    const std::size_t plane_bytes = (px + 3) / 4 * 4;
    std::vector<std::vector<std::uint8_t>> full;
    for (std::size_t i = 0; i < f.comps.size(); ++i)
        full.push_back(Upsample(f.comps[i], planes[i], f.hmax, f.vmax, f.width, f.height));
    if (f.comps.size() == 1) {
        std::copy(full[0].begin(), full[0].end(), img.data.begin());
        return img;
    }
    bool ycc = true;
    if (jfif) {
        ycc = true;
    } else if (adobe) {
        ycc = adobe_transform != 0;
    } else {
        ycc = !(f.comps[0].id == 'R' && f.comps[1].id == 'G' && f.comps[2].id == 'B');
    }
    std::uint8_t* r = img.data.data();
    std::uint8_t* g = r + plane_bytes;
    std::uint8_t* b = g + plane_bytes;
    if (!ycc) {
        std::copy(full[0].begin(), full[0].end(), r);
        std::copy(full[1].begin(), full[1].end(), g);
        std::copy(full[2].begin(), full[2].end(), b);
        return img;
    }
    static const YccTables tab;
    for (std::size_t i = 0; i < px; ++i) {
        const int y = full[0][i], cb = full[1][i], cr = full[2][i];
        r[i] = Clamp255(y + tab.cr_r[cr]);
        g[i] = Clamp255(y + static_cast<int>((tab.cb_g[cb] + tab.cr_g[cr]) >> 16));
        b[i] = Clamp255(y + tab.cb_b[cb]);
    }
    return img;
}

}  // namespace

std::optional<Image> DecodeJpeg(const std::uint8_t* data, std::size_t size, std::string& error) {
    try {
        return Decode(data, size);
    } catch (const JpegFailure& e) {
        error = e.message;
    } catch (const std::bad_alloc&) {
        error = "Insufficient memory";
    }
    return std::nullopt;
}

}  // namespace vqm::image_io
