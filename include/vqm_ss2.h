/*
 * vqm_ss2.h -- C ABI of the MI355X-native SSIMULACRA 2.1 scoring path
 * (libvqm_hip.so). Plain pointers and sizes only; every entry point that can
 * fail returns a status code and leaves a thread-local message for
 * vqm_last_error().
 *
 * This is the operator boundary of the reference's backend seam:
 * internal::Ssimulacra2GpuPipeline (src/metrics/ssimulacra2/
 * ssimulacra2_internal.h:102-152), created by CreateVulkanSsimulacra2Pipeline
 * (src/metrics/ssimulacra2/vulkan.h:19-30, vulkan.cc:751-774). The C++ facade
 * in include/ssimulacra2.h rebuilds the reference's public API
 * (include/ssimulacra2.h:17-215) on top of it, mapping status codes to the
 * reference's exceptions:
 *   VQM_INVALID_ARGUMENT -> std::invalid_argument
 *   VQM_LOGIC_ERROR      -> std::logic_error   (call while a pair is in flight)
 *   VQM_RUNTIME_ERROR    -> std::runtime_error (no device, HIP failure, allocation)
 *   VQM_OVERFLOW         -> std::overflow_error (layout does not fit)
 *
 * Inputs are planar images in the reference's Image layout
 * (include/vqm_image.h:122-133): planes Y/R, U/G, V/B and alpha, each starting
 * on a 4-byte boundary, in any format the reference accepts (byte, word,
 * half or float samples; grey, RGB or YUV with 4:4:4 to 4:1:0 chroma; limited
 * or full range; its matrix coefficients, transfer characteristics and
 * primaries; straight or premultiplied alpha). Pipelines decode every format;
 * the batch engine and the server take 8-bit full-range sRGB RGB. Norm
 * arrays hold 108 doubles, index level * 18 + k with k = 0..5 the
 * Ssimulacra2Scale::avg_ssim entries
 * [channel * 2 + norm] and k = 6..17 the avg_edgediff entries
 * [channel * 4 + n] (include/ssimulacra2.h:17-25); pruned (channel, level)
 * pairs and levels beyond `scales` are 0.0.
 *
 * Threading (include/ssimulacra2.h:66-70, 135-139): a pipeline is not
 * thread-safe and holds one submission in flight; separate pipelines on
 * separate threads run concurrently, each on its own HIP stream.
 */
#ifndef VQM_SS2_H
#define VQM_SS2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VQM_OK = 0,
    VQM_INVALID_ARGUMENT = 1,
    VQM_LOGIC_ERROR = 2,
    VQM_RUNTIME_ERROR = 3,
    VQM_OVERFLOW = 4
};

/* Stored plane precision; arithmetic is FP32 in every format.
 *   F32   FP32 planes.
 *   F16   FP16 planes: the reference's half_planes (include/ssimulacra2.h:
 *         41-46), with its rounding of each level's linear box average to
 *         FP16 as well (ssimu2_common.glsl:84-96); scores move by up to
 *         ~0.06 against FP64.
 *   BF16  BF16 planes: 8 mantissa bits are too few for the metric (scores
 *         move by points); kept for comparison only.
 *   U16   the mixed form: planes stored as unsigned 16-bit fixed point
 *         (q / 65536, exact zero, uniform 2^-16 step over the XYB range
 *         [0, 0.94]), arithmetic in FP32, sums in FP64. Scores stay within
 *         ~2e-4 of FP64 (FP32: ~1.5e-4) with half the FP32 plane bytes. An
 *         addition with no reference counterpart. */
enum { VQM_PLANES_F32 = 0, VQM_PLANES_F16 = 1, VQM_PLANES_BF16 = 2, VQM_PLANES_U16 = 3 };

/* The blur of the maps (vqm_ss2_config.blur, vqm_ss2_batch_config.blur):
 *   SAMPLED    the reference's 13-tap sampled Gaussian, sigma 1.5
 *              (pipeline.cc:274-291); the default.
 *   RECURSIVE  the filter of Cloudinary's ssimulacra2 2.1, libjxl's recursive
 *              Gaussian (ssimulacra2.cc:82-106; Charalampidis 2016), applied
 *              as the zero-padded 9-tap FIR its recursions compute. The
 *              reference dropped it (docs/divergence.md:22); no reference
 *              counterpart. */
enum { VQM_BLUR_SAMPLED = 0, VQM_BLUR_RECURSIVE = 1 };

/* The arithmetic of the XYB transform and the SSIM' map (the map_form
 * fields below):
 *   REFERENCE   the reference's own operation order: prepare.comp:81-98
 *               with correctly rounded cube roots, and scale.comp:130-142,
 *               305-333: a, b, a*a, b*b and a*b blurred, the variances taken
 *               as s11 - mu1^2 and s22 - mu2^2, IEEE divisions. Matches the
 *               FP32 restatement of the reference bit for bit per pixel.
 *   DIFFERENCE  the same map rearranged: a, b, (a - b)^2 and a*b blurred,
 *               s11 + s22 - mu1^2 - mu2^2 formed as blur((a - b)^2) -
 *               (mu1 - mu2)^2 + 2 (s12 - mu1 mu2), so no difference of two
 *               large second moments is rounded; one reciprocal for both
 *               quotients, a 0.84-ulp cube root. Closer to FP64 than the
 *               reference's order (scores within ~1.5e-4 of FP64) and a
 *               fifth less blur work, but on ill-conditioned pairs (the
 *               reference's 32x32 gradient pair C) it moves away from the
 *               reference's FP32 result by up to ~3e-3. No reference
 *               counterpart.
 *   DEFAULT     REFERENCE, except with VQM_PLANES_U16 planes (the mixed
 *               form), whose 16-bit fixed-point rounding holds the 1e-3
 *               score bound only with DIFFERENCE. */
enum { VQM_MAP_DEFAULT = 0, VQM_MAP_REFERENCE = 1, VQM_MAP_DIFFERENCE = 2 };

/* ImageFormat, include/vqm_image.h:100-120; numeric enum values are zimg's
 * as in the reference (vqm_image.h:25-96). Validation follows the
 * reference's ValidateImage (input.cc:152-222): an unknown or unsupported
 * combination is VQM_INVALID_ARGUMENT. 8-bit full-range opaque sRGB RGB with
 * BT.709 primaries decodes through a table; every other format through a
 * per-image descriptor. */
typedef struct vqm_image_format {
    uint8_t type;
    uint8_t family;
    uint8_t alpha;
    uint8_t range;
    uint8_t matrix;
    uint8_t transfer;
    uint8_t primaries;
    uint8_t chroma_location;
    uint8_t subsample_w;
    uint8_t subsample_h;
    uint8_t depth;
    uint8_t reserved[5];
} vqm_image_format;

/* A host image: geometry, format and the byte length of its data
 * (Image::data.size(), include/vqm_image.h:134-141). */
typedef struct vqm_image_desc {
    uint32_t width;
    uint32_t height;
    vqm_image_format format;
    uint64_t bytes;
} vqm_image_desc;

/* A non-owning planar host image, ImageView/PlaneView
 * (include/vqm_views.h:18-30): the planes in Image order (Y/R, U/G, V/B,
 * alpha), each at its own address with its own row stride in bytes (at
 * least the row's length; padding at row ends is skipped). Planes the
 * format does not have are ignored. */
typedef struct vqm_host_view {
    uint32_t width;
    uint32_t height;
    vqm_image_format format;
    const void* planes[4];
    uint64_t strides[4];
} vqm_host_view;

/* A device-resident image owned by the caller: replaces VulkanBufferView
 * (include/vqm_vulkan.h:74-101). `data` is a device pointer to the packed
 * planes and must stay alive and unmodified until vqm_ss2_finish returns. */
typedef struct vqm_device_image {
    const void* data;
    uint32_t width;
    uint32_t height;
    vqm_image_format format;
    uint64_t bytes;
} vqm_device_image;

/* Ssimulacra2Config, include/ssimulacra2.h:35-64. */
typedef struct vqm_ss2_config {
    uint32_t width;
    uint32_t height;
    uint32_t resident_images;   /* gpu_resident_images, default 2 */
    uint32_t plane_format;      /* VQM_PLANES_*; half_planes = VQM_PLANES_F16 */
    uint32_t streaming_inputs;  /* default 1 */
    uint32_t strip_rows;        /* output rows per scale work item, a multiple
                                   of 8 up to 1024; 0 = chosen for the geometry
                                   as a one-pair batch's on a whole MI355X
                                   (below): 16 rows at 1080p, 64 at 4K */
    uint64_t input_bytes;       /* caller-owned image pool (AllocateImage), default 0 */
    uint32_t graph;             /* 1: each pair replays one HIP graph of the
                                   descriptor upload, the three kernels and the
                                   norm read-back, captured on first use; 0
                                   (default): eager launches, which measured
                                   slightly faster. No reference counterpart. */
    uint32_t chroma_filter;     /* subsampled chroma upsampling: 0 bilinear
                                   (default), 1 bicubic, 2 lanczos
                                   (Ssimulacra2Config::chroma_filter) */
    uint32_t blur;              /* VQM_BLUR_SAMPLED (default): the reference's
                                   13-tap sampled Gaussian; VQM_BLUR_RECURSIVE:
                                   the kernel of Cloudinary ssimulacra2's
                                   recursive Gaussian. No reference field. */
    uint32_t map_form;          /* VQM_MAP_*, default VQM_MAP_DEFAULT. No
                                   reference field. */
    uint32_t strip_cols;        /* output columns per scale work item: 64 or
                                   128; 0 = default (128). No reference
                                   field. */
} vqm_ss2_config;

/* PipelineMemoryUsage, include/vqm_memory_types.h:29-38: device bytes the
 * pipeline holds, by region. */
typedef struct vqm_ss2_memory {
    uint64_t workspace_bytes;   /* planes + partials + norms */
    uint64_t resident_bytes;    /* resident image slots */
    uint64_t streaming_bytes;   /* streaming input buffers on the device */
    uint64_t input_pool_bytes;  /* caller-owned image pool */
    uint64_t pinned_host_bytes; /* pinned staging on the host */
    uint64_t total_device_bytes;
} vqm_ss2_memory;

typedef struct vqm_ss2_pipeline vqm_ss2_pipeline;

/* ---- host-only helpers (no GPU) ---------------------------------------- */

/* Thread-local message of the last failing call on this thread. */
const char* vqm_last_error(void);

/* Defaults of Ssimulacra2Config for a geometry. */
void vqm_ss2_config_init(vqm_ss2_config* config, uint32_t width, uint32_t height);

/* The default image format: Byte, Rgb, opaque, full range, sRGB, BT.709. */
void vqm_image_format_srgb8(vqm_image_format* format);

/* RequiredBytes, include/vqm_image.h:143-146 (input.cc:244-255). */
int vqm_required_bytes(const vqm_image_format* format, uint32_t width, uint32_t height,
                       uint64_t* bytes);

/* ValidateImage (input.cc:152-222): VQM_OK when the reference accepts
 * `format` at this geometry, else INVALID_ARGUMENT naming the reason. */
int vqm_validate_image_format(const vqm_image_format* format, uint32_t width, uint32_t height);

/* CopyImage (include/vqm_views.h:32-36, image_view.cc:16-95): validates a
 * view and packs it into the Image layout at `out`. On entry `*bytes` is
 * the capacity of `out`; on return it is the packed length (RequiredBytes
 * of the view's format and size). With `out` NULL the view is only
 * validated and its length reported. INVALID_ARGUMENT for an empty size,
 * an invalid format, a missing plane, a short stride or a capacity other
 * than the packed length; OVERFLOW when a plane's last row is not
 * addressable. A CPU copy. */
int vqm_copy_image(const vqm_host_view* view, void* out, uint64_t* bytes);

/* PlaneScaleOffset (input.cc:714-757): the decode scale and offset of a
 * sample class (luma/RGB, chroma when `chroma`, alpha when `alpha`). */
int vqm_plane_scale_offset(const vqm_image_format* format, int chroma, int alpha, float* scale,
                           float* offset);

/* PlanSsimulacra2Scales, pipeline.cc:28-48: writes `count` and width,height
 * pairs of each level. INVALID_ARGUMENT below 8x8. */
int vqm_ss2_plan_scales(uint32_t width, uint32_t height, uint32_t sizes[12],
                        uint32_t* count);

/* MakeGaussian(1.5), pipeline.cc:274-291. */
void vqm_ss2_gaussian_taps(float taps[13]);

/* The 13 FP32 taps the scale kernel applies for `blur` (VQM_BLUR_*), centre
 * at index 6; the recursive kernel's outer four are zero. */
int vqm_ss2_blur_taps(uint32_t blur, float taps[13]);

/* Ssimulacra2ChannelPruned, score.cc:103-114: 1 when pruned. */
int vqm_ss2_channel_pruned(uint32_t channel, uint32_t scale, uint32_t scales);

/* Ssimulacra2Result::Score, score.cc:118-163. Writes the score to *score;
 * INVALID_ARGUMENT for more than six scales, nonfinite norms or overflow. */
int vqm_ss2_score(uint32_t scales, const double norms[108], double* score);

/* The u8 sRGB decode table the kernels use (color.glsl:94-113,135-137). */
void vqm_ss2_srgb_table(float table[256]);

/* Device bytes a pipeline with `config` allocates (layout.cc:106-325). */
int vqm_ss2_memory_plan(const vqm_ss2_config* config, vqm_ss2_memory* plan);

/* MemoryHeapInfo (include/vqm_memory.h:33-42) of the two heaps allocations
 * come from: heap 0 device memory (HBM, device-local; budget = total, usage
 * = total - free from hipMemGetInfo), heap 1 host memory for pinned buffers
 * (budget = total, usage = total - available RAM). Without a usable device
 * heap 0 reports zeros and no budget. Never fails. */
typedef struct vqm_memory_heap {
    uint64_t size_bytes;
    uint32_t device_local;
    uint32_t budget_available;
    uint64_t budget_bytes;
    uint64_t usage_bytes;
} vqm_memory_heap;
int vqm_memory_heaps(vqm_memory_heap heaps[2]);

/* Algorithmic HBM bytes of one resident pair of this geometry: both u8
 * inputs read, every stored plane element written and read once. */
int vqm_ss2_algorithmic_bytes(uint32_t width, uint32_t height, uint32_t plane_format,
                              uint64_t* bytes);

/* ---- the pipeline (Ssimulacra2Pipeline, include/ssimulacra2.h:83-215) ---- */

/* Ssimulacra2Pipeline(Ssimulacra2Config), pipeline.cc:76-106.
 * INVALID_ARGUMENT below 8x8 or for a bad config; RUNTIME_ERROR without a
 * usable gfx950 device or on allocation failure. */
int vqm_ss2_create(const vqm_ss2_config* config, vqm_ss2_pipeline** out);

/* Waits for in-flight work and frees everything; never fails. */
void vqm_ss2_destroy(vqm_ss2_pipeline* pipeline);

/* Upload (include/ssimulacra2.h:115-118, pipeline.cc:108-114): copies a host
 * image into a resident slot. INVALID_ARGUMENT on a mismatched image or when
 * every slot is in use; LOGIC_ERROR while a pair is in flight. The handle is
 * process-unique and never reused. */
int vqm_ss2_upload(vqm_ss2_pipeline* pipeline, const vqm_image_desc* image,
                   const void* data, uint64_t* handle);

/* Release (include/ssimulacra2.h:119-121): unknown handles are ignored;
 * LOGIC_ERROR while a pair is in flight. */
int vqm_ss2_release(vqm_ss2_pipeline* pipeline, uint64_t handle);

/* Submit(Image, Image) (include/ssimulacra2.h:145-147): streaming; the host
 * data is copied before this returns. `background` is accepted for API
 * parity (the alpha blend target; opaque inputs ignore it). */
int vqm_ss2_submit_host(vqm_ss2_pipeline* pipeline, const vqm_image_desc* original,
                        const void* original_data, const vqm_image_desc* distorted,
                        const void* distorted_data, float background, uint64_t* token);

/* Submit(ImageView, ImageView) (include/ssimulacra2.h:148-151,
 * vulkan_runtime_inputs.cc:156-183): both views are validated, then packed
 * straight into the pipeline's pinned staging (no intermediate copy) before
 * this returns. INVALID_ARGUMENT when a view's size differs from the
 * pipeline's or as vqm_copy_image. */
int vqm_ss2_submit_views(vqm_ss2_pipeline* pipeline, const vqm_host_view* original,
                         const vqm_host_view* distorted, float background, uint64_t* token);

/* Submit(UploadedImage, UploadedImage) (include/ssimulacra2.h:152-155).
 * INVALID_ARGUMENT for a stale or unknown handle. */
int vqm_ss2_submit_resident(vqm_ss2_pipeline* pipeline, uint64_t original,
                            uint64_t distorted, float background, uint64_t* token);

/* Submit(UploadedImage, Image) (include/ssimulacra2.h:156-159). */
int vqm_ss2_submit_resident_host(vqm_ss2_pipeline* pipeline, uint64_t original,
                                 const vqm_image_desc* distorted, const void* distorted_data,
                                 float background, uint64_t* token);

/* Submit(VulkanBufferView, VulkanBufferView[, VulkanSubmitSync])
 * (include/ssimulacra2.h:186-201): caller-owned device images. When
 * `wait_stream` is not NULL (a hipStream_t), the pair's work is ordered
 * after everything already queued on it, the HIP analogue of the wait
 * semaphores of VulkanSubmitSync. */
int vqm_ss2_submit_device(vqm_ss2_pipeline* pipeline, const vqm_device_image* original,
                          const vqm_device_image* distorted, void* wait_stream,
                          float background, uint64_t* token);

/* Finish (include/ssimulacra2.h:203, vulkan.cc:644-650): waits for the token
 * and writes the level count and 108 norms. INVALID_ARGUMENT for an unknown
 * or already finished token. */
int vqm_ss2_finish(vqm_ss2_pipeline* pipeline, uint64_t token, uint32_t* scales,
                   double norms[108]);

/* Allocation-free scoring of one pair: no counterpart in the reference. For
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc. The variant
 * of vqm_ss2_create + vqm_ss2_submit_device + vqm_ss2_finish that runs the
 * same kernels (so its norms are those of a pipeline with this config, bit
 * for bit) with their workspace in a fixed device array of 128 MiB that the
 * library holds on each device it runs on until the process ends: pairs up
 * to 4096x2160 in every plane format. `config` gives the geometry,
 * plane_format, strip_rows, blur and map_form (the other fields are
 * ignored); both images are caller-owned 8-bit sRGB RGB device images of
 * that geometry. The work is queued on `stream` (a hipStream_t, NULL for the
 * default stream) and the call returns; the 108 norms land in
 * `norms_device` (device memory, 108 doubles) when the stream reaches them,
 * and `scales` (host, may be NULL) receives the level count at once. No
 * hipMalloc, hipFree or other allocation happens on any call. Every call on
 * a device shares its array: calls that overlap on one device (several
 * streams or threads) race and must be serialized by the caller.
 * INVALID_ARGUMENT for a mismatched or non-8-bit-sRGB image or a pair whose
 * workspace exceeds the array. */
int vqm_ss2_score_device_noalloc(const vqm_ss2_config* config, const vqm_device_image* original,
                                 const vqm_device_image* distorted, void* stream, double* norms_device,
                                 uint32_t* scales);

/* GetMemoryUsage (include/ssimulacra2.h:204-205). */
int vqm_ss2_memory_usage(const vqm_ss2_pipeline* pipeline, vqm_ss2_memory* usage);

/* AllocateImage / MapImage / Release(ImageBuffer) (include/ssimulacra2.h:
 * 161-184) over a device pool of config.input_bytes. MapImage returns a
 * device pointer: fill it with hipMemcpy or a kernel before the Submit that
 * scores it. INVALID_ARGUMENT when the pool is zero or full; LOGIC_ERROR
 * while a pair is in flight. */
int vqm_ss2_allocate_image(vqm_ss2_pipeline* pipeline, const vqm_image_format* format,
                           uint64_t* buffer, uint64_t* bytes);
int vqm_ss2_map_image(vqm_ss2_pipeline* pipeline, uint64_t buffer, void** device_ptr);
int vqm_ss2_release_image(vqm_ss2_pipeline* pipeline, uint64_t buffer);
int vqm_ss2_submit_buffers(vqm_ss2_pipeline* pipeline, uint64_t original, uint64_t distorted,
                           float background, uint64_t* token);

/* Stage times of the last finished pair in milliseconds (prepare, scale,
 * reduce), measured with HIP events on the pipeline's stream. The reference's
 * VQM_STAGE_TIMING stages (src/timing.h:18-21). */
int vqm_ss2_stage_times(const vqm_ss2_pipeline* pipeline, float ms[3]);

/* The resident reference cache (src/metrics/ssimulacra2/vulkan.cc:5-11,
 * 238-241, 716-735): a submission whose original is the same resident
 * handle as the previous submission's original reuses that image's planes
 * and prepares the distorted image only. Uploads get new handles and
 * Release, or any submission with a non-resident original, ends the reuse.
 * Writes how many images the last submission prepared (1 or 2). */
int vqm_ss2_prepared_images(const vqm_ss2_pipeline* pipeline, uint32_t* images);

/* ---- batched scoring (no counterpart in the reference: it overlaps pairs
 * only by holding several pipelines, vqm_pool.h:35-50) -------------------- */

typedef struct vqm_ss2_batch vqm_ss2_batch;

typedef struct vqm_ss2_batch_config {
    uint32_t plane_format;   /* VQM_PLANES_* */
    uint32_t strip_rows;     /* multiple of 8; 0 = chosen per geometry when it
                                is first submitted: 192 rows when the launch
                                has four work items per wave slot of the
                                device, else 128 when it has one, else 64 or
                                32, the tallest that gives one per SIMD, else
                                16 (servers: 64) */
    uint32_t max_pairs;      /* pairs per submission */
    uint32_t chunks;         /* 0 or 1: prepare, then scale, on one stream; 2..16:
                                runs of consecutive pairs whose scale launches
                                (second stream) overlap the next run's prepare */
    uint64_t hbm_budget;     /* device bytes the workspace pool may hold, 0 = unlimited */
    uint32_t autotune;       /* 1: strip_rows, strip_cols and chunks come from
                                vqm_ss2_autotune for each geometry the first
                                time it is seen */
    uint32_t blur;           /* VQM_BLUR_SAMPLED or VQM_BLUR_RECURSIVE */
    uint32_t persistent;     /* 1: each kernel launches only the blocks that stay
                                resident at once (occupancy x compute units), each
                                looping over the batch's tiles and strips in a
                                fixed stride; 0: one block per tile or strip */
    uint32_t map_form;       /* VQM_MAP_* */
    uint32_t strip_cols;     /* 64 or 128; 0 = default (128) */
    uint32_t graph;          /* 1: a submission of one geometry in one run
                                (chunks 1), 8-bit sRGB device inputs, is
                                replayed from a HIP graph captured the first
                                time its shape (geometry, pair count) is seen;
                                its stage times then report the whole graph
                                as the prepare stage. 2: the same for
                                submissions of up to 8 pairs, where the launch
                                overhead shows (+16% for one 1080p pair, +6%
                                for four, nothing from 16 on; the Python
                                binding's default). 0: eager. */
} vqm_ss2_batch_config;

/* One pair of caller-owned device images (packed u8 sRGB RGB). */
typedef struct vqm_ss2_device_pair {
    const void* original;
    const void* distorted;
    uint32_t width;
    uint32_t height;
} vqm_ss2_device_pair;

int vqm_ss2_batch_create(const vqm_ss2_batch_config* config, vqm_ss2_batch** out);
void vqm_ss2_batch_destroy(vqm_ss2_batch* batch);

/* Scores `count` pairs of any geometries (each >= 8x8) in one launch of each
 * kernel. Workspaces come from a per-geometry device pool that grows on
 * demand within hbm_budget (INVALID_ARGUMENT when a submission cannot fit).
 * One submission may be in flight (LOGIC_ERROR otherwise). */
int vqm_ss2_batch_submit(vqm_ss2_batch* batch, const vqm_ss2_device_pair* pairs,
                         uint32_t count, uint64_t* token);

/* Waits; writes scales[count] and norms[count * 108]. */
int vqm_ss2_batch_finish(vqm_ss2_batch* batch, uint64_t token, uint32_t* scales,
                         double* norms);

/* Stage times of the last finished submission (prepare, scale, reduce), ms. */
int vqm_ss2_batch_stage_times(const vqm_ss2_batch* batch, float ms[3]);

/* Device bytes held by the batch's workspace pool. */
int vqm_ss2_batch_pool_bytes(const vqm_ss2_batch* batch, uint64_t* bytes);

/* Fixes the scale strip shape (rows: a multiple of 8, 0 = the batch's
 * default; cols: 64 or 128, 0 = the batch's default) of one geometry before
 * its first submission, for example a shape vqm_ss2_autotune chose in an
 * earlier process; it takes precedence over `autotune`. LOGIC_ERROR once
 * the geometry has been submitted. One launch may mix strip widths (one
 * scale kernel launch per width). */
int vqm_ss2_batch_set_shape(vqm_ss2_batch* batch, uint32_t width, uint32_t height, uint32_t strip_rows,
                            uint32_t strip_cols);

/* ---- launch-shape autotuner (no reference counterpart) ------------------ */

typedef struct vqm_ss2_tune_result {
    uint32_t strip_rows;     /* scale work-item height chosen */
    uint32_t chunks;         /* batch chunk count chosen */
    uint32_t candidates;     /* shapes timed */
    uint32_t cached;         /* 1 when the choice came from the process cache */
    float ms_per_launch;     /* measured time of the chosen shape */
    float default_ms;        /* measured time of 64 rows, 128 columns, 1 chunk */
    uint32_t strip_cols;     /* scale work-item width chosen */
} vqm_ss2_tune_result;

/* Times every (strip height 16/32/64/128/192/256) x (strip width 64/128) x
 * (chunks 1/2/4/8) shape for a launch of `pairs` pairs of this geometry and
 * format (rounded up to a power of two, at most 512) on the device and
 * returns the fastest; the choice is cached for the process. The strip width
 * sets the scale kernel's registers per lane and so its waves per SIMD (two
 * columns per lane: 2 waves in the reference form, 3 in the difference
 * form; one column: more). Shapes change only the order in which the FP64
 * partial sums of the strips are added (every pixel's map value and every
 * FP32 block sum is the same), so results agree to about 1e-15 relative.
 * Holds up to 16 GiB of temporary device memory while it runs. */
int vqm_ss2_autotune(uint32_t width, uint32_t height, uint32_t plane_format, uint32_t map_form,
                     uint32_t pairs, vqm_ss2_tune_result* result);

/* ---- serving (no counterpart in the reference: its closest caller is the
 * worker pool of Ssimulacra2Pool, vqm_pool.h:35-86, pool.cc:24-116, and
 * the bench driver's workers, bench/driver/ssimulacra2.cc:175-285) --------
 *
 * A request server with continuous batching. Requests of any geometry
 * (>= 8x8) are queued as they arrive. A dispatcher thread forms a batch from
 * the queue head whenever one of `lanes` HIP streams is free, admitting
 * requests in arrival order while their workspaces fit `hbm_budget`, and
 * launches it; batches on different lanes overlap on the device. A
 * completion thread hands results back as batches finish and returns their
 * workspaces to the per-geometry pools. Each request's result does not
 * depend on what it was batched with. All entry points are thread-safe. */

typedef struct vqm_ss2_server vqm_ss2_server;

typedef struct vqm_ss2_server_config {
    uint32_t plane_format;   /* VQM_PLANES_* */
    uint32_t lanes;          /* batches in flight at once, 1..16; 0 = 3 */
    uint32_t max_batch;      /* requests per launch, 1..4096; 0 = 64 */
    uint32_t strip_rows;     /* multiple of 8; 0 = default */
    uint64_t hbm_budget;     /* device bytes for workspaces and host-input
                                staging; 0 = unlimited */
    uint32_t blur;           /* VQM_BLUR_SAMPLED (default) or VQM_BLUR_RECURSIVE */
    uint32_t map_form;       /* VQM_MAP_* */
    uint32_t strip_cols;     /* 64 or 128; 0 = default (128) */
    uint32_t graph;          /* 1: batches of one geometry with device inputs
                                are replayed from HIP graphs, one per (geometry,
                                batch size) a lane has seen (at most 32 a lane) */
} vqm_ss2_server_config;

/* The outcome of one request. Times are vqm_now_ns() readings. */
typedef struct vqm_ss2_request_result {
    int32_t status;          /* VQM_OK or the request's error */
    uint32_t scales;
    uint32_t batch_size;     /* requests in the launch that scored it */
    uint32_t lane;
    uint64_t submit_ns;      /* accepted by vqm_ss2_server_submit_* */
    uint64_t launch_ns;      /* its batch was queued on the device */
    uint64_t done_ns;        /* its batch's completion was observed */
    double norms[108];
} vqm_ss2_request_result;

typedef struct vqm_ss2_server_stats {
    uint64_t submitted;
    uint64_t completed;      /* including failed */
    uint64_t failed;
    uint64_t batches;
    uint64_t batched_requests; /* sum of batch sizes */
    uint64_t max_batch;
    uint64_t pool_bytes;     /* workspace bytes held now */
    uint64_t peak_pool_bytes;
    uint32_t max_lanes_busy; /* most batches in flight at once */
    uint32_t queued;         /* requests waiting now */
} vqm_ss2_server_stats;

void vqm_ss2_server_config_init(vqm_ss2_server_config* config);
/* INVALID_ARGUMENT for a bad config; RUNTIME_ERROR without a device. */
int vqm_ss2_server_create(const vqm_ss2_server_config* config, vqm_ss2_server** out);
/* Scores every queued request, then stops; results not collected are lost. */
void vqm_ss2_server_destroy(vqm_ss2_server* server);

/* Queues a pair of caller-owned device images; they must stay alive and
 * unmodified until the request completes. INVALID_ARGUMENT below 8x8 or for
 * null data. */
int vqm_ss2_server_submit_device(vqm_ss2_server* server, const vqm_ss2_device_pair* pair,
                                 uint64_t* ticket);

/* Queues a pair of host images (packed u8 sRGB RGB); they are copied to the
 * device when the request's batch launches, so they must stay readable until
 * the request completes. Pinned memory (vqm_host_malloc) makes the copy
 * asynchronous. */
int vqm_ss2_server_submit_host(vqm_ss2_server* server, const vqm_image_desc* original,
                               const void* original_data, const vqm_image_desc* distorted,
                               const void* distorted_data, uint64_t* ticket);

/* Blocks until the request completes and consumes its ticket. Returns the
 * request's own status (INVALID_ARGUMENT when it cannot fit the budget even
 * alone; RUNTIME_ERROR on a device failure), or INVALID_ARGUMENT for an
 * unknown ticket. `result` is filled in either case when the ticket is
 * known. */
int vqm_ss2_server_wait(vqm_ss2_server* server, uint64_t ticket, vqm_ss2_request_result* result);

int vqm_ss2_server_get_stats(vqm_ss2_server* server, vqm_ss2_server_stats* stats);

/* The clock of the request times (steady, nanoseconds). */
uint64_t vqm_now_ns(void);

/* One request of an open-loop replay: device or host images. */
typedef struct vqm_ss2_replay_request {
    const void* original;
    const void* distorted;
    uint32_t width;
    uint32_t height;
    uint32_t host;           /* 1: host images (vqm_ss2_server_submit_host) */
    uint32_t reserved;
    uint64_t arrival_ns;     /* offset from the start of the replay */
} vqm_ss2_replay_request;

/* Submits `count` requests at their arrival offsets (open loop: a request is
 * submitted at its arrival time whether or not earlier ones are done) from
 * the calling thread, collects them on a second thread, and writes one
 * result per request, with submit_ns replaced by the scheduled arrival time
 * so that done_ns - submit_ns is the latency the request saw. Returns the
 * first failing request's status, else VQM_OK. */
int vqm_ss2_server_replay(vqm_ss2_server* server, const vqm_ss2_replay_request* requests,
                          uint32_t count, vqm_ss2_request_result* results);

/* ---- device plumbing for callers without a HIP toolchain ---------------- */

int vqm_device_count(int* count);
/* Name, gfx architecture name, CU count, LDS bytes per block, HBM bytes. */
int vqm_device_info(char name[256], char arch[64], int* cus, int* lds_per_block,
                    uint64_t* hbm_bytes);
int vqm_device_malloc(uint64_t bytes, void** ptr);
int vqm_device_free(void* ptr);
/* Pinned (page-locked) host memory, for host inputs copied asynchronously. */
int vqm_host_malloc(uint64_t bytes, void** ptr);
int vqm_host_free(void* ptr);
int vqm_memcpy_htod(void* dst, const void* src, uint64_t bytes);
int vqm_memcpy_dtoh(void* dst, const void* src, uint64_t bytes);
int vqm_device_synchronize(void);
/* Copies `bytes` (a multiple of 16) between device buffers with a plain
 * streaming kernel and waits: the reference point for HBM traffic counters. */
int vqm_device_copy(void* dst, const void* src, uint64_t bytes);

/* ---- self-check (no reference counterpart) ------------------------------ */

/* Result of vqm_ss2_check_cbrt. */
typedef struct vqm_ss2_cbrt_check {
    uint64_t mismatches;     /* inputs whose two roots differ */
    uint64_t exact_steps;    /* inputs that took the exact midpoint comparison */
    double max_seed_error;   /* largest |1 - x r^3| of the hardware seed r */
    uint32_t first_mismatch; /* bits of the smallest differing input, or 0xFFFFFFFF */
    uint32_t reserved;
} vqm_ss2_cbrt_check;

/* Checks, on the current device, the correctly rounded cube root the
 * reference form's XYB takes (seeded by the hardware log2 and exp2, whose
 * rounding no document fixes) against the same root computed in IEEE FP64
 * steps only (which the CPU tests check against exact arithmetic), for
 * every float whose bits lie in [first_bits, end_bits), a range of positive
 * normal floats (0x00800000 to 0x7F800000 covers them all). Waits for the
 * device. */
int vqm_ss2_check_cbrt(uint32_t first_bits, uint32_t end_bits, vqm_ss2_cbrt_check* result);

/* Result of vqm_ss2_check_div. */
typedef struct vqm_ss2_div_check {
    uint64_t mismatches;             /* (a, b) pairs whose quotients differ */
    uint64_t reciprocal_mismatches;  /* b whose refined reciprocal is not the
                                        correctly rounded 1/b, and (b, e) where
                                        rcp(b 2^e) != rcp(b) 2^-e */
    uint32_t first_mismatch;         /* smallest b significand that differs, or 0xFFFFFFFF */
    uint32_t reserved;
} vqm_ss2_div_check;

/* Checks, on the current device, the division the reference form's SSIM'
 * map takes (hardware reciprocal, one Newton step, one correction) against
 * the IEEE quotient: for b = 1.mb with mb in [first_significand,
 * end_significand) and every a = 1.ma; that the refined reciprocal (the
 * difference form's) is the correctly rounded 1/b; and the hardware
 * reciprocal's invariance under scaling b by 2^e, e in [-24, 24).
 * [0, 2^23) covers every pair of significands (about 30 s). Waits for the
 * device. */
int vqm_ss2_check_div(uint32_t first_significand, uint32_t end_significand, vqm_ss2_div_check* result);

#ifdef __cplusplus
}
#endif

#endif /* VQM_SS2_H */
