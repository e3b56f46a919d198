// The general input surface on the host (see ss2_source.h). The colorimetry
// follows the published definitions the reference takes from zimg: luma
// weights of BT.601/709/2020, FCC and SMPTE 240M; chromaticities and white
// points of H.273's primaries; RGB->XYZ from primaries and white; the BT.2100
// LMS and ICtCp matrices.

#include "ss2_source.h"

#include <cmath>

#include "ss2_host.h"

namespace vqm::ss2 {

namespace {

using Vec3 = std::array<double, 3>;
using Xy = std::array<double, 2>;

constexpr double kRec709Alpha = 1.09929682680944;
constexpr double kRec709Beta = 0.018053968510807;

uint32_t TypeDepth(uint8_t type) { return type == fmt::kByte ? 8u : type == fmt::kFloat ? 32u : 16u; }

uint32_t EffectiveDepth(const vqm_image_format& f) { return f.depth != 0 ? f.depth : TypeDepth(f.type); }

bool KnownMatrix(uint8_t m) { return m <= 17 && m != 3; }

bool KnownPrimaries(uint8_t p) { return (p >= 1 && p <= 12 && p != 3) || p == fmt::kPrEbu3213E; }

bool KnownTransfer(uint8_t t) { return t >= 1 && t <= 18 && t != 3; }

bool IsBt709Curve(uint8_t t) {
    return t == fmt::kTrBt709 || t == fmt::kTrBt601 || t == fmt::kTrBt2020_10 || t == fmt::kTrBt2020_12;
}

double Rec709Oetf(double x) {
    x = std::max(x, 0.0);
    return x < kRec709Beta ? x * 4.5 : kRec709Alpha * std::pow(x, 0.45) - (kRec709Alpha - 1.0);
}

Matrix3 Identity3() { return {1, 0, 0, 0, 1, 0, 0, 0, 1}; }

Vec3 Apply(const Matrix3& m, const Vec3& v) {
    return {m[0] * v[0] + m[1] * v[1] + m[2] * v[2], m[3] * v[0] + m[4] * v[1] + m[5] * v[2],
            m[6] * v[0] + m[7] * v[1] + m[8] * v[2]};
}

double Dot(const Vec3& a, const Vec3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

Vec3 Cross(const Vec3& a, const Vec3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}

// Chromaticity (x, y) to XYZ at Y = 1.
Vec3 XyzOf(const Xy& xy) { return {xy[0] / xy[1], 1.0, (1.0 - xy[0] - xy[1]) / xy[1]}; }

// H.273 primaries (R, G, B chromaticities).
std::array<Xy, 3> PrimariesXy(uint8_t p) {
    switch (p) {
        case fmt::kPrBt470M: return {{{0.670, 0.330}, {0.210, 0.710}, {0.140, 0.080}}};
        case fmt::kPrBt470Bg: return {{{0.640, 0.330}, {0.290, 0.600}, {0.150, 0.060}}};
        case fmt::kPrSt170M:
        case fmt::kPrSt240M: return {{{0.630, 0.340}, {0.310, 0.595}, {0.155, 0.070}}};
        case fmt::kPrBt709: return {{{0.640, 0.330}, {0.300, 0.600}, {0.150, 0.060}}};
        case fmt::kPrFilm: return {{{0.681, 0.319}, {0.243, 0.692}, {0.145, 0.049}}};
        case fmt::kPrBt2020: return {{{0.708, 0.292}, {0.170, 0.797}, {0.131, 0.046}}};
        case fmt::kPrSt431_2:
        case fmt::kPrSt432_1: return {{{0.680, 0.320}, {0.265, 0.690}, {0.150, 0.060}}};
        case fmt::kPrEbu3213E: return {{{0.630, 0.340}, {0.295, 0.605}, {0.155, 0.077}}};
        default: break;
    }
    Fail(VQM_INVALID_ARGUMENT, "color primaries do not define chromaticities");
}

// White points: illuminant C for BT.470M and film, E for XYZ (ST 428), the
// DCI white for ST 431-2, D65 otherwise.
Vec3 WhiteXyz(uint8_t p) {
    if (p == fmt::kPrBt470M || p == fmt::kPrFilm) return XyzOf({0.31, 0.316});
    if (p == fmt::kPrSt428) return XyzOf({1.0 / 3.0, 1.0 / 3.0});
    if (p == fmt::kPrSt431_2) return XyzOf({0.314, 0.351});
    return XyzOf({0.3127, 0.3290});
}

// Linear RGB of these primaries -> XYZ: the primaries' XYZ as columns, each
// scaled so that RGB (1, 1, 1) lands on the white point.
Matrix3 RgbToXyz(uint8_t p) {
    if (p == fmt::kPrSt428) return Identity3();
    const auto xy = PrimariesXy(p);
    const Vec3 r = XyzOf(xy[0]), g = XyzOf(xy[1]), b = XyzOf(xy[2]);
    const Matrix3 m = {r[0], g[0], b[0], r[1], g[1], b[1], r[2], g[2], b[2]};
    const Vec3 s = Apply(Inverse3(m), WhiteXyz(p));
    Matrix3 out{};
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col) out[row * 3 + col] = m[row * 3 + col] * s[col];
    return out;
}

// Published luma weights (Kr, Kb) of the YUV matrices.
void KrKbOf(uint8_t matrix, double& kr, double& kb) {
    switch (matrix) {
        case fmt::kMatFcc: kr = 0.3, kb = 0.11; return;
        case fmt::kMatSt240M: kr = 0.212, kb = 0.087; return;
        case fmt::kMatBt470Bg:
        case fmt::kMatSt170M: kr = 0.299, kb = 0.114; return;
        case fmt::kMatBt709: kr = 0.2126, kb = 0.0722; return;
        case fmt::kMatBt2020Ncl:
        case fmt::kMatBt2020Cl: kr = 0.2627, kb = 0.0593; return;
        default: break;
    }
    Fail(VQM_INVALID_ARGUMENT, "matrix coefficients have no luma weights");
}

// Luma weights derived from chromaticities: the Y row of RgbToXyz, i.e.
// white . (g x b) and white . (r x g) over the primaries' determinant. The
// BT.709 and BT.2020 primaries keep their published weights.
void KrKbOfPrimaries(uint8_t p, double& kr, double& kb) {
    if (p == fmt::kPrBt709) return KrKbOf(fmt::kMatBt709, kr, kb);
    if (p == fmt::kPrBt2020) return KrKbOf(fmt::kMatBt2020Ncl, kr, kb);
    const auto xy = PrimariesXy(p);
    const Vec3 r = XyzOf(xy[0]), g = XyzOf(xy[1]), b = XyzOf(xy[2]);
    const Vec3 w = WhiteXyz(p);
    const Vec3 xs = {r[0], g[0], b[0]}, ys = {r[1], g[1], b[1]}, zs = {r[2], g[2], b[2]};
    const double det = Dot(xs, Cross(ys, zs));
    kr = Dot(w, Cross(g, b)) / det;
    kb = Dot(w, Cross(r, g)) / det;
}

// RGB -> Y, (B - Y) / (2 - 2 Kb), (R - Y) / (2 - 2 Kr).
Matrix3 RgbToYuv(double kr, double kb) {
    const double kg = 1.0 - kr - kb;
    const double u = 1.0 / (2.0 - 2.0 * kb), v = 1.0 / (2.0 - 2.0 * kr);
    return {kr, kg, kb, -kr * u, -kg * u, (1.0 - kb) * u, (1.0 - kr) * v, -kg * v, -kb * v};
}

Matrix3 RgbToYuvOf(uint8_t matrix) {
    if (matrix == fmt::kMatYCgCo) return {0.25, 0.5, 0.25, -0.25, 0.5, -0.25, 0.5, 0.0, -0.5};
    double kr = 0, kb = 0;
    KrKbOf(matrix, kr, kb);
    return RgbToYuv(kr, kb);
}

// BT.2100: LMS -> ICtCp for the PQ and HLG variants; the LMS matrix of RGB
// (BT.2020 primaries).
Matrix3 LmsToIctCp(uint8_t transfer) {
    if (transfer == fmt::kTrAribB67)
        return {0.5, 0.5, 0.0, 3625.0 / 4096, -7465.0 / 4096, 3840.0 / 4096,
                9500.0 / 4096, -9212.0 / 4096, -288.0 / 4096};
    return {0.5, 0.5, 0.0, 6610.0 / 4096, -13613.0 / 4096, 7003.0 / 4096,
            17933.0 / 4096, -17390.0 / 4096, -543.0 / 4096};
}

Matrix3 RgbToLms() {
    return {1688.0 / 4096, 2146.0 / 4096, 262.0 / 4096, 683.0 / 4096, 2951.0 / 4096,
            462.0 / 4096, 99.0 / 4096, 309.0 / 4096, 3688.0 / 4096};
}

uint32_t Mode(const vqm_image_format& f) {
    if (f.family == fmt::kGrey) return 0;
    if (f.matrix == fmt::kMatBt2020Cl || f.matrix == fmt::kMatChromaCl) return 1;
    if (f.matrix == fmt::kMatIctCp) return 2;
    return 0;
}

void Curve(uint8_t transfer, uint32_t& id, float& scale) {
    scale = 1.0f;
    switch (transfer) {
        case fmt::kTrLinear: id = kCurveLinear; return;
        case fmt::kTrBt709:
        case fmt::kTrBt601:
        case fmt::kTrBt2020_10:
        case fmt::kTrBt2020_12:
        case fmt::kTrSt240M: id = kCurvePow24; return;
        case fmt::kTrBt470M: id = kCurvePow22; return;
        case fmt::kTrBt470Bg: id = kCurvePow28; return;
        case fmt::kTrSrgb: id = kCurveSrgb; return;
        case fmt::kTrLog100: id = kCurveLog100; return;
        case fmt::kTrLog316: id = kCurveLog316; return;
        case fmt::kTrIec61966_2_4: id = kCurveXvycc; return;
        case fmt::kTrBt1361: id = kCurveBt1361; return;
        case fmt::kTrSt2084: id = kCurvePq, scale = 10000.0f / 100.0f; return;
        case fmt::kTrSt428: id = kCurveSt428; return;
        case fmt::kTrAribB67: id = kCurveHlg, scale = 1000.0f / 100.0f; return;
        default: break;
    }
    Fail(VQM_INVALID_ARGUMENT, "unsupported transfer characteristics");
}

}  // namespace

Matrix3 Inverse3(const Matrix3& m) {
    const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    return {c00 / det, (m[2] * m[7] - m[1] * m[8]) / det, (m[1] * m[5] - m[2] * m[4]) / det,
            c01 / det, (m[0] * m[8] - m[2] * m[6]) / det, (m[2] * m[3] - m[0] * m[5]) / det,
            c02 / det, (m[1] * m[6] - m[0] * m[7]) / det, (m[0] * m[4] - m[1] * m[3]) / det};
}

Matrix3 Multiply3(const Matrix3& a, const Matrix3& b) {
    Matrix3 r{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[i * 3 + j] = a[i * 3 + 0] * b[0 * 3 + j] + a[i * 3 + 1] * b[1 * 3 + j] + a[i * 3 + 2] * b[2 * 3 + j];
    return r;
}

Matrix3 GamutToBt709(uint8_t primaries) {
    return Multiply3(Inverse3(RgbToXyz(fmt::kPrBt709)), RgbToXyz(primaries));
}

Matrix3 NclToRgb(const vqm_image_format& f, bool* identity) {
    *identity = true;
    if (f.family == fmt::kGrey || f.matrix == fmt::kMatRgb || f.matrix == fmt::kMatUnspecified)
        return Identity3();
    *identity = false;
    if (f.matrix == fmt::kMatChromaNcl) {
        double kr = 0, kb = 0;
        KrKbOfPrimaries(f.primaries, kr, kb);
        return Inverse3(RgbToYuv(kr, kb));
    }
    return Inverse3(RgbToYuvOf(f.matrix));
}

void ValidateFormat(const vqm_image_format& f, uint32_t width, uint32_t height, const std::string& name) {
    const auto fail = [&](const char* message) { Fail(VQM_INVALID_ARGUMENT, name + ": " + message); };
    if (f.type > fmt::kFloat) fail("invalid pixel type");
    if (f.range > fmt::kFull) fail("invalid pixel range");
    if (f.alpha > fmt::kPremultiplied) fail("invalid alpha type");
    if (f.chroma_location > fmt::kBottom) fail("invalid chroma location");
    if (!KnownMatrix(f.matrix)) fail("invalid matrix coefficients");
    if (!KnownPrimaries(f.primaries)) fail("invalid color primaries");
    if (f.depth > TypeDepth(f.type)) fail("bit depth exceeds the pixel type");
    if (f.subsample_w > 2 || f.subsample_h > 2) fail("subsampling greater than 4x is not supported");
    if (width % (1u << f.subsample_w) != 0 || height % (1u << f.subsample_h) != 0)
        fail("image dimensions must be divisible by the subsampling");
    if (f.range == fmt::kLimited && EffectiveDepth(f) < 8) fail("limited range requires a bit depth of at least 8");
    const bool subsampled = f.subsample_w != 0 || f.subsample_h != 0;
    switch (f.family) {
        case fmt::kGrey:
            if (subsampled) fail("subsampled greyscale images are not supported");
            if (f.matrix == fmt::kMatRgb) fail("greyscale images cannot declare RGB matrix coefficients");
            break;
        case fmt::kRgb:
            if (subsampled) fail("subsampled RGB images are not supported");
            if (f.matrix != fmt::kMatRgb && f.matrix != fmt::kMatUnspecified)
                fail("RGB images cannot declare YUV matrix coefficients");
            break;
        case fmt::kYuv:
            if (f.matrix == fmt::kMatRgb) fail("YUV images cannot declare RGB matrix coefficients");
            if (f.matrix == fmt::kMatUnspecified) fail("matrix coefficients are required for YUV input");
            break;
        default: fail("invalid color family");
    }
    if (f.transfer == fmt::kTrUnspecified) fail("transfer characteristics are required");
    if (!KnownTransfer(f.transfer)) fail("unsupported transfer characteristics");
    if (f.family != fmt::kGrey && f.primaries == fmt::kPrUnspecified) fail("color primaries are required");
    if (f.family == fmt::kGrey) return;
    switch (f.matrix) {
        case fmt::kMatBt2020Cl:
        case fmt::kMatChromaCl:
            if (!IsBt709Curve(f.transfer))
                fail("constant-luminance matrix coefficients require the BT.709 transfer");
            break;
        case fmt::kMatIctCp:
            if (!(f.transfer == fmt::kTrSt2084 || f.transfer == fmt::kTrAribB67) || f.primaries != fmt::kPrBt2020)
                fail("ICtCp requires the ST 2084 or ARIB B67 transfer and BT.2020 primaries");
            break;
        case fmt::kMatSt2085YDzDx:
        case fmt::kMatSt2128IptC2:
        case fmt::kMatYCgCoRe:
        case fmt::kMatYCgCoRo: fail("unsupported matrix coefficients");
        default: break;
    }
    if (f.primaries == fmt::kPrSt428) {
        const bool derived = f.matrix == fmt::kMatChromaNcl || f.matrix == fmt::kMatChromaCl;
        if (derived || f.transfer == fmt::kTrAribB67)
            fail("XYZ color primaries do not define the chromaticity this conversion needs");
    }
}

bool IsSrgb8(const vqm_image_format& f) {
    return f.type == fmt::kByte && f.family == fmt::kRgb && f.alpha == fmt::kNoAlpha && f.range == fmt::kFull &&
           (f.matrix == fmt::kMatRgb || f.matrix == fmt::kMatUnspecified) && f.transfer == fmt::kTrSrgb &&
           f.primaries == fmt::kPrBt709 && f.subsample_w == 0 && f.subsample_h == 0 &&
           (f.depth == 0 || f.depth == 8);
}

std::array<PlaneGeometry, 4> PlaneLayout(const vqm_image_format& f, uint32_t width, uint32_t height) {
    const uint64_t sample = f.type == fmt::kByte ? 1 : f.type == fmt::kFloat ? 4 : 2;
    std::array<PlaneGeometry, 4> planes{};
    uint64_t cursor = 0;
    for (int p = 0; p < 4; ++p) {
        const bool chroma = p == 1 || p == 2;
        const bool present = p == 0 || (chroma && f.family != fmt::kGrey) || (p == 3 && f.alpha != fmt::kNoAlpha);
        if (!present) continue;
        PlaneGeometry& g = planes[p];
        g.present = true;
        g.width = chroma ? width >> f.subsample_w : width;
        g.height = chroma ? height >> f.subsample_h : height;
        g.offset = cursor;
        cursor = (cursor + static_cast<uint64_t>(g.width) * g.height * sample + 3) / 4 * 4;
    }
    return planes;
}

ScaleOffset PlaneScaleOffset(const vqm_image_format& f, bool chroma, bool alpha) {
    if (f.type == fmt::kHalf || f.type == fmt::kFloat) return {1.0f, 0.0f};
    const uint32_t depth = EffectiveDepth(f);
    if (depth < 1 || depth > TypeDepth(f.type)) Fail(VQM_INVALID_ARGUMENT, "invalid bit depth for pixel type");
    if (f.range == fmt::kLimited && depth < 8)
        Fail(VQM_INVALID_ARGUMENT, "limited range requires a bit depth of at least 8");
    const bool full = alpha || f.range == fmt::kFull;
    const bool ycgco = f.matrix == fmt::kMatYCgCo;
    const double range = full ? static_cast<double>((1u << depth) - 1u)
                              : static_cast<double>((chroma && !ycgco ? 224u : 219u) << (depth - 8));
    // Chroma is centred on the code midpoint in both ranges.
    const double offset = alpha ? 0.0 : chroma ? static_cast<double>(1u << (depth - 1))
                                              : full ? 0.0 : static_cast<double>(16u << (depth - 8));
    return {static_cast<float>(1.0 / range), static_cast<float>(-offset / range)};
}

SourceDesc BuildSourceDesc(const vqm_image_format& f, uint32_t width, uint32_t height, uint32_t chroma_filter) {
    if (chroma_filter > 2) Fail(VQM_INVALID_ARGUMENT, "chroma_filter must be bilinear, bicubic or lanczos");
    SourceDesc d{};
    d.kind = 1;
    const auto planes = PlaneLayout(f, width, height);
    for (int p = 0; p < 4; ++p) {
        d.plane_off[p] = planes[p].offset;
        d.plane_w[p] = planes[p].width;
        d.plane_h[p] = planes[p].height;
    }
    d.type = f.type;
    d.family = f.family;
    d.alpha = f.alpha;
    d.sub_w = f.subsample_w;
    d.sub_h = f.subsample_h;
    d.mode = Mode(f);
    d.filter = chroma_filter;
    Curve(f.transfer, d.transfer, d.to_linear_scale);
    const ScaleOffset luma = PlaneScaleOffset(f, false, false);
    const ScaleOffset chroma = PlaneScaleOffset(f, f.family == fmt::kYuv, false);
    const ScaleOffset alpha = PlaneScaleOffset(f, false, true);
    d.scale[0] = luma.scale, d.offset[0] = luma.offset;
    d.scale[1] = chroma.scale, d.offset[1] = chroma.offset;
    d.scale[2] = alpha.scale, d.offset[2] = alpha.offset;

    Matrix3 m1 = Identity3();
    d.identity1 = 1;
    if (d.mode == 2) {
        m1 = Inverse3(LmsToIctCp(f.transfer));
        d.identity1 = 0;
    } else if (d.mode == 0) {
        bool identity = true;
        m1 = NclToRgb(f, &identity);
        d.identity1 = identity ? 1 : 0;
    }
    Matrix3 m2 = Identity3();
    d.identity2 = 1;
    if (f.family != fmt::kGrey && f.primaries != fmt::kPrBt709) {
        m2 = GamutToBt709(f.primaries);
        if (d.mode == 2) m2 = Multiply3(m2, Inverse3(RgbToLms()));
        d.identity2 = 0;
    } else if (d.mode == 2) {
        m2 = Inverse3(RgbToLms());
        d.identity2 = 0;
    }
    for (int i = 0; i < 9; ++i) {
        d.m1[i] = static_cast<float>(m1[i]);
        d.m2[i] = static_cast<float>(m2[i]);
    }

    // Luma weights: the CL op takes them from its matrix, the HLG OOTF from
    // the primaries; nothing else reads them.
    double kr = 0.0, kg = 1.0, kb = 0.0;
    if (f.family != fmt::kGrey) {
        if (d.mode == 1) {
            if (f.matrix == fmt::kMatChromaCl)
                KrKbOfPrimaries(f.primaries, kr, kb);
            else
                KrKbOf(f.matrix, kr, kb);
            kg = 1.0 - kr - kb;
        } else if (f.transfer == fmt::kTrAribB67) {
            KrKbOfPrimaries(f.primaries, kr, kb);
            kg = 1.0 - kr - kb;
        }
    }
    d.kr = static_cast<float>(kr), d.kg = static_cast<float>(kg), d.kb = static_cast<float>(kb);
    if (d.mode == 1) {
        d.cl[0] = static_cast<float>(Rec709Oetf(1.0 - kb));
        d.cl[1] = static_cast<float>(1.0 - Rec709Oetf(kb));
        d.cl[2] = static_cast<float>(Rec709Oetf(1.0 - kr));
        d.cl[3] = static_cast<float>(1.0 - Rec709Oetf(kr));
    }

    // Chroma siting: the offset of the chroma sample from the centre of the
    // luma samples it covers, in chroma samples; zero on an axis that is not
    // subsampled.
    if (f.family == fmt::kYuv && (f.subsample_w != 0 || f.subsample_h != 0)) {
        const double left = 0.5 - 0.5 / static_cast<double>(1u << f.subsample_w);
        const double top = 0.5 - 0.5 / static_cast<double>(1u << f.subsample_h);
        double sx = 0.0, sy = 0.0;
        switch (f.chroma_location) {
            case fmt::kLeft: sx = left; break;
            case fmt::kTopLeft: sx = left, sy = top; break;
            case fmt::kTop: sy = top; break;
            case fmt::kBottomLeft: sx = left, sy = -top; break;
            case fmt::kBottom: sy = -top; break;
            default: break;
        }
        d.shift_x = static_cast<float>(sx);
        d.shift_y = static_cast<float>(sy);
    }
    return d;
}

}  // namespace vqm::ss2
