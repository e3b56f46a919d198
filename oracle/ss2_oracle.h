/*
 * CPU oracle for the SSIMULACRA 2.1 scoring path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker the parity tests compare the HIP path against.
 * It is never linked into, loaded by or called from the product library
 * (vqm_amd/csrc); only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * It restates the reference's algorithm (GreatValueCreamSoda/vqm,
 * src/metrics/ssimulacra2/ and its shaders) in three arithmetic forms:
 *
 *   FP64      everything in double: the "true" value of the metric as the
 *             reference defines it (13-tap zero-padded FIR, pruned channels).
 *   FP32 ref  the reference shaders' FP32 operation order (horizontal blur
 *             first, scale.comp:149-301), correctly rounded cube root.
 *   FP32 dev  the HIP kernels' FP32 operations: the reference's blur order
 *             (horizontal pass first), the kernels' own reproducible cube
 *             root, (a - b)^2 blurred in place of a^2 and b^2 (the SSIM'
 *             map needs only s11 + s22 = blur((a - b)^2) + 2 s12), optional
 *             FP16/BF16/fixed-point rounding of stored planes. The HIP path
 *             matches this form per pixel bit for bit.
 *
 * Both FP32 forms sum each map in FP32 over 1 x 8 pixel blocks, then in FP64
 * (sum_blocks), as the HIP kernels do: the kernels then match every FP32
 * block sum, and only the order of the FP64 sums differs.
 *
 * Parity pins (checked in tests/test_oracle.py):
 *   - end to end: the reference's 32x32 goldens 93.17505696 / 92.12553218
 *     (tests/metrics/ssimulacra2/pipeline_test.cc:146-174);
 *   - identity scores exactly 100 (pipeline_test.cc:120-144);
 *   - scale counts (pipeline_test.cc:126-128, 767-776);
 *   - pruning mask (docs/ssimulacra2.md:96-104, pipeline_test.cc:163-169);
 *   - score mapping contract (tests/metrics/ssimulacra2/score_test.cc);
 *   - Gaussian taps, sRGB decode, XYB and box filter against closed forms;
 *   - blur against a brute-force 2D convolution on tiny images.
 */
#ifndef VQM_SS2_ORACLE_H
#define VQM_SS2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic form of one oracle run. */
typedef struct ss2o_mode {
    int32_t arith;        /* 0 = FP64, 1 = FP32 */
    int32_t blur_order;   /* 0 = horizontal then vertical, 1 = vertical then horizontal */
    int32_t cbrt_kind;    /* 0 = libm cbrt (correctly rounded to FP32), 1 = device Newton */
    int32_t plane_format; /* FP32 arith only: 0 = FP32, 1 = FP16, 2 = BF16, 3 = unorm16 stored planes */
    uint32_t plane_round_levels; /* levels whose stored planes use plane_format */
    int32_t sum_blocks;   /* FP32 only: 1 = each map's values summed in FP32 over blocks of
                             one column by eight rows (rows 8k .. 8k+7, in row order), then
                             the block sums in FP64 (the HIP kernels' order); 0 = FP64 per
                             pixel */
    int32_t round_linear; /* 1 = the linear box averages of levels >= 1 are rounded to
                             plane_format as well, as the reference's half-plane build
                             does (prepare.comp:100-111, ssimu2_common.glsl:84-96) */
    int32_t blur_kind;    /* 0 = the 13-tap sampled Gaussian (pipeline.cc:274-291);
                             1 = the 9-tap truncated-cosine kernel that libjxl's recursive
                             Gaussian computes (Cloudinary ssimulacra2.cc:82-106) */
    int32_t sum_variances; /* FP32 only: 1 = the device form of the SSIM' map: (a - b)^2
                              blurred as one quantity e in place of a^2 and b^2, then
                              v = e - (mu1 - mu2)^2, den = num_s + v and
                              d = max(((mu1 - mu2)^2 num_s + v) / den, 0), algebraically
                              scale.comp:305-333; 0 = s11 and s22 blurred apart and the
                              map in the reference's order (scale.comp:130-142,305-333) */
} ss2o_mode;

/* One stored plane value as plane_format keeps it: 1 FP16 and 2 BF16 (round
 * to nearest even), 3 unsigned 16-bit fixed point q / 65536 (round to
 * nearest even, clamped to 0..65535, NaN to 0); 0 returns v. */
float ss2o_round_plane(int32_t plane_format, float v);

/* The truncated-cosine kernel of the recursive Gaussian at sigma 1.5, in FP64,
 * in the 13-tap frame (centre at index 6, zeros at |m| >= 5). */
void ss2o_cosine_taps(double taps[13]);

/* Scale plan: returns the level count (1..6) and writes width,height pairs,
 * or -1 below 8x8. pipeline.cc:28-48. */
int32_t ss2o_plan_scales(uint32_t width, uint32_t height, uint32_t sizes[12]);

/* The 13 taps of the sigma 1.5 sampled Gaussian. pipeline.cc:274-291. */
void ss2o_gaussian_taps(float taps[13]);

/* 1 if every one of the six weights of (channel, scale) is <= 0.01 at a
 * pyramid of `scales` levels. score.cc:103-114. */
int32_t ss2o_channel_pruned(int32_t channel, int32_t scale, int32_t scales);

/* Weighted sum and score mapping. score.cc:118-163. status: 0 ok,
 * 1 invalid (too many scales, nonfinite norms, mapping overflow). */
double ss2o_score(int32_t scales, const double norms[108], int32_t* status);

/* u8 sRGB code -> linear FP32 / FP64 (color.glsl:94-113,135-137). */
void ss2o_srgb_lut_f32(float lut[256]);
void ss2o_srgb_lut_f64(double lut[256]);

/* The HIP kernels' fast cube root (bit-exact restatement). */
float ss2o_cbrt_dev(float x);

/* The cube root correctly rounded to FP32 (round to nearest): the FP32 ref
 * form's root. 0 for x <= 0. */
float ss2o_cbrt_rn(float x);

/* The kernels' reciprocal of a positive normal float: the correctly rounded
 * 1 / x. */
float ss2o_rcp_dev(float x);

/* linear RGB -> XYB in FP32 (fused multiply-adds as the kernels issue them)
 * and FP64. prepare.comp:81-98. */
void ss2o_xyb_f32(float r, float g, float b, int32_t cbrt_kind, float out[3]);
void ss2o_xyb_f64(double r, double g, double b, double out[3]);

/* Blur one zero-padded plane with the 13-tap Gaussian, in the given order,
 * FP32 (arith 1) or FP64 (arith 0); in/out are float or double arrays. */
void ss2o_blur_f32(const float* in, float* out, uint32_t w, uint32_t h, int32_t order);
void ss2o_blur_f64(const double* in, double* out, uint32_t w, uint32_t h);

/* The linear RGB pyramid of one packed u8 image: level by level, channel
 * by channel, w_l*h_l values each, FP32 (device and FP32 forms) or FP64.
 * prepare.comp:100-111,149-168. Returns the level count or -1. */
/* The stored XYB planes of one packed u8 image under an FP32 `mode` (arith 1):
 * level by level, channel X, Y, B, each w_l * h_l floats; levels where every
 * channel is pruned are written as zeros. Returns the level count, -1 below
 * 8x8 or for an FP64 mode. */
int32_t ss2o_xyb_planes(const ss2o_mode* mode, uint32_t width, uint32_t height, const uint8_t* image,
                        float* out);
int32_t ss2o_pyramid_f32(uint32_t width, uint32_t height, const uint8_t* image, float* out);
int32_t ss2o_pyramid_f64(uint32_t width, uint32_t height, const uint8_t* image, double* out);

/* Bytes of one packed u8 RGB image (planes 4-byte aligned). vqm_image.h:122-133,
 * input.cc:114-145,244-255. */
uint64_t ss2o_image_bytes(uint32_t width, uint32_t height);

/* Score one pair of packed planar u8 sRGB RGB images. norms: 108 doubles,
 * index level*18 + k, k 0..5 = avg_ssim[c*2+n], 6..17 = avg_edgediff[c*4+m];
 * pruned and absent levels are 0. sums (may be NULL): the raw per-level map
 * sums in the same layout. Returns the level count or -1. */
int32_t ss2o_compute(const ss2o_mode* mode, uint32_t width, uint32_t height,
                     const uint8_t* image0, const uint8_t* image1,
                     double norms[108], double sums[108]);

/* Several pairs of one geometry on `threads` host threads, one pair per
 * thread at a time; images are arrays of pointers. Used as the CPU baseline. */
int32_t ss2o_compute_many(const ss2o_mode* mode, uint32_t width, uint32_t height,
                          uint32_t pairs, const uint8_t* const* images0,
                          const uint8_t* const* images1, double* norms,
                          uint32_t threads);

#ifdef __cplusplus
}
#endif

#endif
