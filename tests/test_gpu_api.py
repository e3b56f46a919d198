"""The reference API's contract on the HIP backend (pipeline_test.cc,
pool_test.cc): input kinds agree bit for bit, in-flight misuse raises
LogicError, stale handles and tokens raise InvalidArgument."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import gradient_image, ramp_image
import vqm_amd
from vqm_amd import (DeviceBuffer, DeviceImage, Image, ImageBuffer, ImageFormat, InvalidArgument,
                     LogicError, Size, Ssimulacra2Config, Ssimulacra2Pipeline, Ssimulacra2Pool,
                     UploadedImage)

pytestmark = pytest.mark.gpu

S = Size(32, 32)


def img(data) -> Image:
    return Image(S, ImageFormat.srgb8(), data)


def same(a, b) -> bool:
    return np.array_equal(a.norms(), b.norms()) and len(a.scales) == len(b.scales)


def test_every_input_kind_agrees():
    """Resident, streaming, mixed, caller-owned buffers and device images give
    identical norms (pipeline_test.cc:451-462, 722-759)."""
    orig, dist = gradient_image(0), gradient_image(1)
    cfg = Ssimulacra2Config(size=S, input_bytes=2 * 3 * 1024 + 512)
    p = Ssimulacra2Pipeline(cfg)
    streamed = p.Compute(img(orig), img(dist))
    a, b = p.Upload(img(orig)), p.Upload(img(dist))
    assert same(p.Compute(a, b), streamed)
    assert same(p.Compute(a, img(dist)), streamed)
    ba, bb = p.AllocateImage(ImageFormat.srgb8()), p.AllocateImage(ImageFormat.srgb8())
    assert ba.bytes == bb.bytes == 3 * 1024
    vqm_amd.memcpy_htod(p.MapImage(ba), orig)
    vqm_amd.memcpy_htod(p.MapImage(bb), dist)
    assert same(p.Compute(ba, bb), streamed)
    da, db = DeviceBuffer(3 * 1024), DeviceBuffer(3 * 1024)
    da.upload(orig)
    db.upload(dist)
    assert same(p.Compute(DeviceImage(da.ptr, S, ImageFormat.srgb8()), DeviceImage(db.ptr, S, ImageFormat.srgb8())), streamed)
    assert streamed.Score() == pytest.approx(93.17505696, abs=5e-3)


def test_in_flight_misuse_raises_logic_error():
    """pipeline_test.cc:380-450: one submission in flight; uploads, releases,
    allocations and second submits raise until Finish."""
    orig, dist = gradient_image(0), gradient_image(1)
    p = Ssimulacra2Pipeline(Ssimulacra2Config(size=S, input_bytes=8192))
    resident = p.Upload(img(orig))
    resident_dist = p.Upload(img(dist))
    expected = p.Compute(img(orig), img(dist)).Score()
    token = p.Submit(img(orig), img(dist))
    invalid = UploadedImage()
    invalid_buf = ImageBuffer()
    with pytest.raises(LogicError):
        p.Submit(img(orig), img(orig))
    with pytest.raises(LogicError):
        p.Upload(img(orig))
    with pytest.raises(LogicError):
        p.Release(resident)
    with pytest.raises(LogicError):
        p.Release(invalid)
    with pytest.raises(LogicError):
        p.Submit(resident, resident_dist)
    with pytest.raises(LogicError):
        p.Submit(invalid, resident)
    with pytest.raises(LogicError):
        p.AllocateImage(ImageFormat.srgb8())
    with pytest.raises(LogicError):
        p.MapImage(invalid_buf)
    with pytest.raises(LogicError):
        p.Release(invalid_buf)
    assert p.Finish(token).Score() == pytest.approx(expected, abs=1e-12)
    with pytest.raises(InvalidArgument):
        p.Finish(token)
    with pytest.raises(InvalidArgument):
        p.Submit(invalid, resident)
    with pytest.raises(InvalidArgument):
        p.Submit(invalid_buf, invalid_buf)
    p.Release(invalid)          # unknown handles are ignored when idle
    p.Release(invalid_buf)
    p.Release(resident)
    p.Release(resident_dist)


def test_input_validation():
    """pipeline_test.cc:296-333 and :380-395."""
    orig = gradient_image(0)
    p = Ssimulacra2Pipeline(Ssimulacra2Config(size=S))
    with pytest.raises(InvalidArgument):
        p.Compute(img(orig), Image(Size(16, 32), ImageFormat.srgb8(), orig))
    with pytest.raises(InvalidArgument):
        p.Compute(img(orig[:10]), img(orig))
    with pytest.raises(InvalidArgument):
        p.Compute(img(orig), img(np.concatenate([orig, orig])))
    bad = ImageFormat.srgb8()
    bad.subsample_w = 1
    with pytest.raises(InvalidArgument):
        p.Compute(Image(S, bad, orig), img(orig))
    # Formats the reference refuses (input_format_chroma_test.cc:430-510).
    for fmt in (ImageFormat(family=2, matrix=0), ImageFormat(type=0, depth=12),
                ImageFormat(type=0, depth=7, range=0), ImageFormat(family=0, subsample_h=1, matrix=1),
                ImageFormat(transfer=2), ImageFormat(family=2, matrix=10, primaries=1)):
        with pytest.raises(InvalidArgument):
            p.Upload(Image(S, fmt, np.zeros(vqm_amd.required_bytes(fmt, S), np.uint8)))
    # Formats it accepts upload, each slot growing to its image.
    slots = []
    for fmt in (ImageFormat(), ImageFormat(type=0, alpha=1), ImageFormat(type=0, transfer=16),
                ImageFormat(type=0, range=0), ImageFormat(type=0, family=2, matrix=1)):
        slots.append(p.Upload(Image(S, fmt, np.zeros(vqm_amd.required_bytes(fmt, S), np.uint8))))
        p.Release(slots[-1])
    # Nothing in flight after the failures: a valid pair still scores.
    assert p.Compute(img(orig), img(orig)).Score() == 100.0


def test_resident_slots_and_stale_handles():
    """pipeline_test.cc:463-500: a released handle never aliases the image
    that reuses its slot; a full pool refuses uploads."""
    orig, dist = gradient_image(0), gradient_image(1)
    p = Ssimulacra2Pipeline(Ssimulacra2Config(size=S, gpu_resident_images=1))
    a = p.Upload(img(orig))
    with pytest.raises(InvalidArgument):
        p.Upload(img(dist))
    p.Release(a)
    b = p.Upload(img(dist))
    assert b.slot != a.slot
    with pytest.raises(InvalidArgument):
        p.Compute(a, b)
    assert p.Compute(b, b).Score() == 100.0


def test_streaming_inputs_disabled():
    """pipeline_test.cc:276-294."""
    orig, dist = gradient_image(0), gradient_image(1)
    expected = Ssimulacra2Pipeline(Ssimulacra2Config(size=S)).Compute(img(orig), img(dist)).Score()
    p = Ssimulacra2Pipeline(Ssimulacra2Config(size=S, streaming_inputs=False))
    with pytest.raises(InvalidArgument):
        p.Compute(img(orig), img(dist))
    a, b = p.Upload(img(orig)), p.Upload(img(dist))
    assert p.Compute(a, b).Score() == pytest.approx(expected, abs=1e-12)
    with pytest.raises(InvalidArgument):
        p.Compute(a, img(dist))


def test_caller_owned_pool_limits():
    """pipeline_test.cc:180-215: two buffers fit, a third is refused; no pool
    at all refuses too."""
    p = Ssimulacra2Pipeline(Ssimulacra2Config(size=S, input_bytes=2 * 3 * 1024))
    c, d = p.AllocateImage(ImageFormat.srgb8()), p.AllocateImage(ImageFormat.srgb8())
    with pytest.raises(InvalidArgument):
        p.AllocateImage(ImageFormat.srgb8())
    p.Release(c)
    e = p.AllocateImage(ImageFormat.srgb8())
    assert e.slot not in (c.slot, d.slot)
    with pytest.raises(InvalidArgument):
        p.MapImage(c)
    none = Ssimulacra2Pipeline(Ssimulacra2Config(size=S))
    with pytest.raises(InvalidArgument):
        none.AllocateImage(ImageFormat.srgb8())


def test_memory_usage_matches_plan():
    cfg = Ssimulacra2Config(size=Size(640, 360))
    p = Ssimulacra2Pipeline(cfg)
    used = p.GetMemoryUsage()
    plan = vqm_amd.GetMemoryRequirements(cfg)
    assert used.arena == plan.direct.arena and used.staging == plan.direct.staging
    assert used.uses_staging and used.auxiliary == []
    assert p.GetMemoryRegions().resident_bytes == 2 * 3 * 640 * 360


def test_pool_matches_single_pipeline():
    """pool_test.cc:27-62: ordered results equal a single pipeline's."""
    pairs = [(img(gradient_image(0)), img(gradient_image(v))) for v in (1, 2, 0, 1, 2)]
    single = Ssimulacra2Pipeline(Ssimulacra2Config(size=S))
    want = [single.Compute(a, b) for a, b in pairs]
    pool = Ssimulacra2Pool(Ssimulacra2Config(size=S), 3)
    got = pool.ComputeMany(pairs)
    assert pool.Workers() == 3
    for g, w in zip(got, want):
        assert same(g, w)
    assert pool.ScoreMany(pairs) == [w.Score() for w in want]


def test_pool_drains_on_invalid_pair():
    """pool_test.cc:64-80: an invalid pair mid-batch raises after draining,
    and the pool keeps working."""
    good = (img(gradient_image(0)), img(gradient_image(1)))
    bad = (img(gradient_image(0)), Image(S, ImageFormat.srgb8(), gradient_image(0)[:5]))
    pool = Ssimulacra2Pool(Ssimulacra2Config(size=S), 2)
    with pytest.raises(InvalidArgument):
        pool.ComputeMany([good, good, bad, good])
    assert len(pool.ComputeMany([good])) == 1


def test_stage_times_are_measured():
    """Eager launches time each stage; a graph replay times the whole pair."""
    a, b = ramp_image(256, 256), ramp_image(256, 256, seed=5)
    p = Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(256, 256), graph=False))
    p.Compute(Image(Size(256, 256), ImageFormat.srgb8(), a), Image(Size(256, 256), ImageFormat.srgb8(), b))
    t = p.StageTimes()
    assert len(t) == 3 and all(v > 0.0 for v in t)
    g = Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(256, 256), graph=True))
    g.Compute(Image(Size(256, 256), ImageFormat.srgb8(), a), Image(Size(256, 256), ImageFormat.srgb8(), b))
    t = g.StageTimes()
    assert t[0] > 0.0 and t[1] == 0.0 and t[2] == 0.0


@pytest.mark.parametrize("graph", [False, True])
def test_resident_reference_is_prepared_once(graph):
    """vulkan.cc:5-11,716-735: a resident original scored against several
    distorted images is prepared once; results equal fresh pipelines bit for
    bit, and a release, a new upload or a non-resident original ends the
    reuse."""
    S = Size(320, 200)
    a, b1 = ramp_image(320, 200, seed=1), ramp_image(320, 200, seed=2)
    b2, c = ramp_image(320, 200, seed=3), ramp_image(320, 200, seed=4)
    img = lambda d: Image(S, ImageFormat.srgb8(), d)  # noqa: E731

    def fresh(x, y):
        q = Ssimulacra2Pipeline(Ssimulacra2Config(size=S))
        r = q.Compute(img(x), img(y)).norms()
        q.close()
        return r

    p = Ssimulacra2Pipeline(Ssimulacra2Config(size=S, gpu_resident_images=4, graph=graph))
    ra, rb1, rb2 = p.Upload(img(a)), p.Upload(img(b1)), p.Upload(img(b2))
    assert np.array_equal(p.Compute(ra, rb1).norms(), fresh(a, b1))
    assert p.PreparedImages() == 2
    assert np.array_equal(p.Compute(ra, rb2).norms(), fresh(a, b2))
    assert p.PreparedImages() == 1
    assert np.array_equal(p.Compute(ra, img(b1)).norms(), fresh(a, b1))   # host distorted
    assert p.PreparedImages() == 1
    assert np.array_equal(p.Compute(img(c), img(b2)).norms(), fresh(c, b2))  # host original
    assert p.PreparedImages() == 2
    assert np.array_equal(p.Compute(ra, rb2).norms(), fresh(a, b2))
    assert p.PreparedImages() == 2
    p.Release(ra)
    rc = p.Upload(img(c))  # may land in the released slot: a new handle
    assert np.array_equal(p.Compute(rc, rb1).norms(), fresh(c, b1))
    assert p.PreparedImages() == 2
    assert np.array_equal(p.Compute(rc, rb2).norms(), fresh(c, b2))
    assert p.PreparedImages() == 1
