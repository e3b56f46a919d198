"""The metrics API and host views on the GPU (tests/core/metrics_test.cc:241-575,
the SSIMULACRA 2 parts; vqm_views.h:26-46): every input form scores what the
pipeline scores, bit for bit; the pipeline's score is the oracle's device
form; frames are numbered; rejected pairs consume no frame; the sequence
states and the shared background follow the reference."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import gradient_image, random_pair
import vqm_amd
from vqm_amd import (AlphaType, DeviceBuffer, DeviceImage, Image, ImageFormat, ImageView, InvalidArgument,
                     LogicError, MetricsConfig, Size, Ssimulacra2Config, Ssimulacra2Pipeline)
from test_metrics_host import padded_view, yuv_format

pytestmark = pytest.mark.gpu


def srgb8_view(data: np.ndarray, size: Size, pad: int) -> ImageView:
    """A padded view over the planes of a packed 8-bit sRGB image."""
    n = size.width * size.height
    stride = (n + 3) // 4 * 4
    planes = []
    for p in range(3):
        wide = np.full((size.height, size.width + pad), 0xA5, np.uint8)
        wide[:, :size.width] = data[p * stride:p * stride + n].reshape(size.height, size.width)
        planes.append(wide[:, :size.width])
    return ImageView(size, ImageFormat.srgb8(), planes)


def config(size: Size, fmt=None, **kw) -> MetricsConfig:
    fmt = fmt or ImageFormat.srgb8()
    return MetricsConfig(size=size, original_format=fmt, distorted_format=fmt,
                         ssimulacra2=Ssimulacra2Config(size=size), **kw)


def test_views_score_like_images(oracle_lib):
    """Submit(ImageView) packs into the staging: the same norms as the packed
    Image, for 8-bit sRGB (checked against the oracle's FP32 ref form, the
    default map form's) and a
    10-bit 4:2:0 YUV general format."""
    w, h = 97, 61
    a, b = random_pair(w, h, seed=5)
    size = Size(w, h)
    with Ssimulacra2Pipeline(Ssimulacra2Config(size=size)) as p:
        want = p.Compute(Image(size, ImageFormat.srgb8(), a), Image(size, ImageFormat.srgb8(), b))
        got = p.Compute(srgb8_view(a, size, 3), srgb8_view(b, size, 17))
        assert np.array_equal(got.norms(), want.norms())
        assert vqm_amd.Score(p, srgb8_view(a, size, 0), srgb8_view(b, size, 1)) == want.Score()
    n, dev, _ = oracle_lib.compute(oracle_lib.F32_REF, w, h, a, b)
    assert np.allclose(want.norms(), dev, rtol=1e-10, atol=0)

    fmt = yuv_format(1, 1, vqm_amd.PixelType.Word, depth=10)
    size = Size(64, 48)
    va, _ = padded_view(fmt, size, 5, seed=3)
    vb, _ = padded_view(fmt, size, 0, seed=4)
    for v in (va, vb):  # keep the samples within 10 bits
        for plane in v.planes:
            if plane is not None:
                plane &= 0x3FF
    with Ssimulacra2Pipeline(Ssimulacra2Config(size=size)) as p:
        want = p.Compute(vqm_amd.CopyImage(va), vqm_amd.CopyImage(vb))
        got = p.Compute(va, vb)
    assert np.array_equal(got.norms(), want.norms())
    assert want.Score() < 100.0


def test_metrics_sequence_and_input_forms():
    size = Size(32, 32)
    orig, dist = gradient_image(0), gradient_image(1)
    ia, ib = Image(size, ImageFormat.srgb8(), orig), Image(size, ImageFormat.srgb8(), dist)
    with Ssimulacra2Pipeline(Ssimulacra2Config(size=size)) as p:
        want = p.Compute(ia, ib).Score()
    m = vqm_amd.Metrics(config(size))
    r0 = m.Compute(ia, ib)
    assert r0.frame_index == 0 and r0.ssimulacra2 == want
    # metrics_test.cc:555-575: rejected pairs consume no frame.
    with pytest.raises(InvalidArgument, match="size differs"):
        m.Compute(ia, Image(Size(31, 32), ImageFormat.srgb8(), dist[:-96]))
    with pytest.raises(InvalidArgument, match="format differs"):
        m.Compute(ia, Image(size, ImageFormat(type=vqm_amd.PixelType.Byte,
                                              transfer=vqm_amd.TransferCharacteristics.Bt709), dist))
    with pytest.raises(InvalidArgument, match="data size"):
        m.Compute(ia, Image(size, ImageFormat.srgb8(), dist[:-1]))
    r1 = m.Compute(srgb8_view(orig, size, 2), srgb8_view(dist, size, 9))
    assert r1.frame_index == 1 and r1.ssimulacra2 == want
    # metrics_test.cc:439-475: packed spans of exactly RequiredBytes.
    with pytest.raises(InvalidArgument, match="packed input has incorrect size"):
        m.Compute(orig.tobytes()[:-1], dist.tobytes())
    with pytest.raises(InvalidArgument, match="packed input has incorrect size"):
        m.Compute(orig, dist[:-1])
    r2 = m.Compute(orig.tobytes(), memoryview(bytearray(dist.tobytes())))
    assert r2.frame_index == 2 and r2.ssimulacra2 == want
    r3 = m.Compute(orig, dist)
    assert r3.frame_index == 3 and r3.ssimulacra2 == want
    with pytest.raises(LogicError, match="reset the active sequence"):
        m.ComputeImage(orig, dist)
    assert m.Flush() is None
    with pytest.raises(LogicError, match="sequence finished"):
        m.Compute(ia, ib)
    fin = m.Finalize()
    assert fin.last_frame is None and fin.penultimate_frame is None and fin.frames == 4
    assert m.Finalize().frames == 4
    m.Reset()
    still = m.ComputeImage(orig, dist)
    again = m.ComputeImage(ia, ib)  # ComputeImage resets after itself
    assert still.frame_index == again.frame_index == 0
    assert still.ssimulacra2 == again.ssimulacra2 == want
    usage = m.GetMemoryUsage()
    assert usage.scratch is None
    plan = vqm_amd.metrics.GetMemoryRequirements(config(size)).spatial
    assert usage.spatial.arena == plan.direct.arena and usage.spatial.staging == plan.direct.staging
    m.close()


def test_metrics_device_inputs_without_streaming():
    """streaming_inputs=False refuses host pixels before any work and still
    scores device images."""
    size = Size(32, 32)
    orig, dist = gradient_image(0), gradient_image(2)
    with Ssimulacra2Pipeline(Ssimulacra2Config(size=size)) as p:
        want = p.Compute(Image(size, ImageFormat.srgb8(), orig), Image(size, ImageFormat.srgb8(), dist)).Score()
    m = vqm_amd.Metrics(config(size, streaming_inputs=False))
    assert not m.GetMemoryUsage().spatial.uses_staging
    with pytest.raises(InvalidArgument, match="streaming inputs are disabled"):
        m.Compute(orig, dist)
    bufs = [DeviceBuffer(orig.nbytes), DeviceBuffer(dist.nbytes)]
    bufs[0].upload(orig)
    bufs[1].upload(dist)
    da, db = DeviceImage(bufs[0].ptr, size, ImageFormat.srgb8()), DeviceImage(bufs[1].ptr, size, ImageFormat.srgb8())
    r = m.Compute(da, db)
    assert r.frame_index == 0 and r.ssimulacra2 == want
    with pytest.raises(InvalidArgument, match="invalid device input"):
        m.Compute(da, DeviceImage(0, size, ImageFormat.srgb8()))
    assert m.Compute(da, db).frame_index == 1
    m.close()
    for buf in bufs:
        buf.free()


def test_metrics_background_and_pipelines():
    """metrics_test.cc:508-553: the top-level background is the metric's;
    MetricPipelines hands out the pipeline it configured."""
    size = Size(32, 32)
    fmt = ImageFormat(type=vqm_amd.PixelType.Byte, alpha=AlphaType.Straight)
    n = size.width * size.height
    alpha = (np.arange(n) * 7 % 256).astype(np.uint8)
    a = np.concatenate([gradient_image(0), alpha])
    b = np.concatenate([gradient_image(1), alpha])
    ia, ib = Image(size, fmt, a), Image(size, fmt, b)
    with Ssimulacra2Pipeline(Ssimulacra2Config(size=size)) as p:
        at02, at08 = p.Compute(ia, ib, 0.2).Score(), p.Compute(ia, ib, 0.8).Score()
    assert at02 != at08
    m = vqm_amd.Metrics(config(size, fmt, background=0.2))
    assert m.ComputeImage(ia, ib).ssimulacra2 == at02
    m.close()
    pipes = vqm_amd.MetricPipelines(config(size, fmt, background=0.8))
    assert pipes.Ssimulacra2().Compute(ia, ib, 0.8).Score() == at08
    assert pipes.Ssimulacra2().GetMemoryRegions().resident_bytes == 0
    pipes.close()
