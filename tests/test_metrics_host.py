"""Host views and the metrics API with no GPU: CopyImage packs strided views
exactly like an independent numpy packing (vqm_views.h:32-36,
image_view.cc:16-95), and the Metrics/MetricPipelines configuration, memory
planning and worker budgets follow tests/core/metrics_test.cc:98-240 and
metric_pipelines.cc:29-212."""
from __future__ import annotations

import math

import numpy as np
import pytest

import vqm_amd
from vqm_amd import (AlphaType, ColorFamily, ImageFormat, ImageView, InvalidArgument, MatrixCoefficients,
                     MetricsConfig, MetricsMemoryRequirements, PipelineMemoryUsage, PixelRange, PixelType,
                     Size, Ssimulacra2Config, TransferCharacteristics)

DTYPES = {PixelType.Byte: np.uint8, PixelType.Word: np.uint16, PixelType.Half: np.float16,
          PixelType.Float: np.float32}


def plane_shapes(fmt: ImageFormat, size: Size):
    """The present planes' (height, width), in Image order."""
    shapes = [(size.height, size.width)]
    if fmt.family != ColorFamily.Grey:
        chroma = (size.height >> fmt.subsample_h, size.width >> fmt.subsample_w)
        shapes += [chroma, chroma]
    else:
        shapes += [None, None]
    shapes.append((size.height, size.width) if fmt.alpha != AlphaType.NoAlpha else None)
    return shapes


def padded_view(fmt: ImageFormat, size: Size, pad: int, seed: int):
    """A view over random planes, each a slice of a wider array, and the
    packed bytes built independently: planes back to back, each starting on
    a 4-byte boundary, the padding bytes zero."""
    rng = np.random.default_rng(seed)
    dtype = DTYPES[fmt.type]
    planes, packed = [], bytearray()
    for shape in plane_shapes(fmt, size):
        if shape is None:
            planes.append(None)
            continue
        h, w = shape
        wide = (rng.random((h, w + pad)) * 1000).astype(dtype)
        plane = wide[:, :w]
        planes.append(plane)
        packed += np.ascontiguousarray(plane).tobytes()
        packed += b"\0" * (-len(packed) % 4)
    return ImageView(size, fmt, planes), bytes(packed)


def yuv_format(sub_w, sub_h, type_=PixelType.Byte, depth=0, alpha=AlphaType.NoAlpha) -> ImageFormat:
    return ImageFormat(type=type_, family=ColorFamily.Yuv, range=PixelRange.Limited,
                       matrix=MatrixCoefficients.Bt709, transfer=TransferCharacteristics.Bt709,
                       subsample_w=sub_w, subsample_h=sub_h, depth=depth, alpha=alpha)


FORMATS = [
    ImageFormat.srgb8(),
    ImageFormat(),  # the reference default: Float sRGB
    ImageFormat(type=PixelType.Half, alpha=AlphaType.Straight),
    yuv_format(1, 1),
    yuv_format(1, 0, PixelType.Word, depth=10),
    yuv_format(2, 2, PixelType.Word, depth=12, alpha=AlphaType.Premultiplied),
    ImageFormat(type=PixelType.Word, family=ColorFamily.Grey, matrix=MatrixCoefficients.Bt709,
                transfer=TransferCharacteristics.Bt709, depth=16),
]


@pytest.mark.parametrize("fmt", FORMATS, ids=range(len(FORMATS)))
@pytest.mark.parametrize("size,pad", [((13, 9), 0), ((36, 20), 3), ((64, 8), 64)])
def test_copy_image_packs_like_numpy(fmt, size, pad):
    size = Size(*size)
    if size.width % (1 << fmt.subsample_w) or size.height % (1 << fmt.subsample_h):
        pytest.skip("geometry not divisible by the subsampling")
    view, want = padded_view(fmt, size, pad, seed=size.width * 31 + pad)
    assert vqm_amd.view_bytes(view) == vqm_amd.required_bytes(fmt, size) == len(want)
    img = vqm_amd.CopyImage(view)
    assert img.size == size and img.format == fmt
    assert bytes(img.data) == want


def test_copy_image_rejections():
    """image_view.cc:16-45: empty size, invalid format, missing plane, short
    stride (invalid_argument); an unaddressable stride (overflow_error)."""
    fmt = yuv_format(1, 1)
    size = Size(16, 8)
    view, _ = padded_view(fmt, size, 2, seed=1)
    with pytest.raises(InvalidArgument, match="size must be non-zero"):
        vqm_amd.CopyImage(ImageView(Size(0, 8), fmt, view.planes))
    with pytest.raises(InvalidArgument, match="image view: "):
        vqm_amd.CopyImage(ImageView(size, yuv_format(1, 1, alpha=5), view.planes))
    missing = list(view.planes)
    missing[2] = None
    with pytest.raises(InvalidArgument, match="missing plane or short stride"):
        vqm_amd.CopyImage(ImageView(size, fmt, missing))
    # A plane narrower than its row: stride shorter than the row's length.
    short = list(view.planes)
    short[1] = np.zeros((4, 6), np.uint8)
    with pytest.raises(InvalidArgument, match="missing plane or short stride"):
        vqm_amd.CopyImage(ImageView(size, fmt, short))
    # A stride whose last row ends past the pointer range: the C ABI sees
    # the stride only, so build the raw view by hand.
    raw = view.to_c()
    raw.strides[0] = (1 << 62)
    from vqm_amd import _lib as L
    import ctypes
    n = L.u64(0)
    assert L.lib().vqm_copy_image(ctypes.byref(raw), None, ctypes.byref(n)) == L.VQM_OVERFLOW
    # Exactly the packed length is required of an output buffer.
    raw = view.to_c()
    n = L.u64(vqm_amd.required_bytes(fmt, size) + 1)
    out = (ctypes.c_uint8 * n.value)()
    assert L.lib().vqm_copy_image(ctypes.byref(raw), out, ctypes.byref(n)) == L.VQM_INVALID_ARGUMENT
    with pytest.raises(InvalidArgument):
        ImageView(size, fmt, [np.zeros((8, 16, 2), np.uint8)]).to_c()


def metrics_config(size=Size(64, 48), **kw) -> MetricsConfig:
    return MetricsConfig(size=size, original_format=ImageFormat.srgb8(), distorted_format=ImageFormat.srgb8(),
                         ssimulacra2=Ssimulacra2Config(size=size), **kw)


def test_metrics_selection_is_checked_before_device_work():
    """metric_pipelines.cc:29-43: a metric must be selected; the background
    lies in [0, 1]; both formats are validated and named."""
    with pytest.raises(InvalidArgument, match="select at least one metric"):
        vqm_amd.Metrics(MetricsConfig())
    with pytest.raises(InvalidArgument, match="select at least one metric"):
        vqm_amd.MetricPipelines(MetricsConfig())
    for bg in (-0.1, 1.5, math.nan, math.inf):
        with pytest.raises(InvalidArgument, match="background must be between 0 and 1"):
            vqm_amd.metrics.GetMemoryRequirements(metrics_config(background=bg))
    bad = metrics_config()
    bad.distorted_format = ImageFormat(transfer=TransferCharacteristics.Unspecified)
    with pytest.raises(InvalidArgument, match="^distorted format: "):
        vqm_amd.metrics.GetMemoryRequirements(bad)


def test_metrics_memory_plan_overrides_the_metric():
    """metric_pipelines.cc:81-93: size, streaming_inputs, resident slots and
    the input pool of the metric's own configuration are overridden; a single
    metric has no shared scratch."""
    config = metrics_config(streaming_inputs=False)
    config.ssimulacra2 = Ssimulacra2Config(size=Size(8, 8), gpu_resident_images=7, input_bytes=1 << 20)
    req = vqm_amd.metrics.GetMemoryRequirements(config)
    assert req.scratch is None
    want = vqm_amd.GetMemoryRequirements(Ssimulacra2Config(size=Size(64, 48), gpu_resident_images=0,
                                                           streaming_inputs=False))
    assert req.spatial.direct == want.direct
    regions = vqm_amd.memory_plan(Ssimulacra2Config(size=Size(64, 48), gpu_resident_images=0,
                                                    streaming_inputs=False))
    assert regions.resident_bytes == 0 and regions.input_pool_bytes == 0 and regions.streaming_bytes == 0
    # memory_policy is the top-level one (vqm_metrics.h:40).
    staged = metrics_config(memory_policy=vqm_amd.MemoryPolicy.Staged)
    assert not vqm_amd.metrics.GetMemoryRequirements(staged).spatial.direct.supported


def test_worst_case_worker_budget():
    """vulkan_memory_query.cc:259-269,395-424: per-heap sums, the tightest
    heap decides, a charged heap needs a budget, sums must not overflow."""
    req = vqm_amd.metrics.GetMemoryRequirements(metrics_config())
    dev, host = req.spatial.direct.arena.allocation_bytes, req.spatial.direct.staging.allocation_bytes
    assert dev > 0 and host > 0
    W = vqm_amd.MaxWorkersForBudgetWorstCase
    assert W(req, [4 * dev, 100 * host]) == 4
    assert W(req, [4 * dev, 3 * host + 1]) == 3
    assert W(req, [dev - 1, 100 * host]) == 0
    with pytest.raises(InvalidArgument, match="missing heap budget"):
        W(req, [10 * dev])
    with pytest.raises(InvalidArgument, match="no requirements"):
        W(MetricsMemoryRequirements(), [1, 1])
    both = MetricsMemoryRequirements(spatial=req.spatial, scratch=req.spatial)
    assert W(both, [4 * dev, 100 * host]) == 2
    huge_scenario = vqm_amd.MemoryScenario(True, vqm_amd.MemoryAllocation(1, (1 << 64) - 1, 0, 0))
    huge = MetricsMemoryRequirements(spatial=req.spatial,
                                     scratch=vqm_amd.MetricMemoryRequirements(direct=huge_scenario))
    with pytest.raises(OverflowError):
        W(huge, [1, 1])
