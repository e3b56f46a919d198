"""The allocation-free entry point (vqm_ss2_score_device_noalloc) through
the C ABI against the entry point it varies (vqm_ss2_create +
vqm_ss2_submit_device + vqm_ss2_finish): the same norms bit for bit, on
calls in a row whose inputs shrink (so the second call reuses a scratch the
first one filled), in both map forms and a 16-bit plane format; sizes the
scratch cannot hold and other input formats are refused before any work."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import vqm_amd
from vqm_amd import DeviceBuffer, DeviceImage, ImageFormat, Size, Ssimulacra2Config, synth
from vqm_amd import _lib as L

pytestmark = pytest.mark.gpu


def upload_pair(w, h, seed):
    a, b = synth.pair(1, seed, w, h, synth.STRENGTHS[seed % 3])
    da, db = DeviceBuffer(a.nbytes), DeviceBuffer(b.nbytes)
    da.upload(a)
    db.upload(b)
    return da, db


def noalloc(cfg: Ssimulacra2Config, da, db, out: DeviceBuffer):
    w, h = cfg.size.width, cfg.size.height
    c = cfg.to_c()
    ia = DeviceImage(da.ptr, Size(w, h), ImageFormat.srgb8()).to_c()
    ib = DeviceImage(db.ptr, Size(w, h), ImageFormat.srgb8()).to_c()
    scales = ctypes.c_uint32()
    status = L.lib().vqm_ss2_score_device_noalloc(ctypes.byref(c), ctypes.byref(ia), ctypes.byref(ib), None,
                                                  ctypes.c_void_p(out.ptr), ctypes.byref(scales))
    return status, scales.value


def pipeline(cfg: Ssimulacra2Config, da, db):
    w, h = cfg.size.width, cfg.size.height
    with vqm_amd.Ssimulacra2Pipeline(cfg) as p:
        r = p.Compute(DeviceImage(da.ptr, Size(w, h), ImageFormat.srgb8()), DeviceImage(db.ptr, Size(w, h), ImageFormat.srgb8()))
    return len(r.scales), r.norms()


@pytest.mark.parametrize("plane_format,map_form", [(vqm_amd.PLANES_F32, vqm_amd.MapForm.Reference),
                                                   (vqm_amd.PLANES_F32, vqm_amd.MapForm.Difference),
                                                   (vqm_amd.PLANES_U16, vqm_amd.MapForm.Default)])
def test_noalloc_matches_the_pipeline(plane_format, map_form):
    out = DeviceBuffer(108 * 8)
    for (w, h), seed in (((1920, 1080), 0), ((640, 360), 1), ((1920, 1080), 2), ((97, 61), 3)):
        print(f"noalloc {w}x{h} planes {plane_format} map {int(map_form)}", flush=True)
        cfg = Ssimulacra2Config(size=Size(w, h), plane_format=plane_format, map_form=map_form)
        da, db = upload_pair(w, h, seed)
        status, scales = noalloc(cfg, da, db, out)
        assert status == 0, L.lib().vqm_last_error()
        assert L.lib().vqm_device_synchronize() == 0
        got = out.download().view(np.float64)
        n, want = pipeline(cfg, da, db)
        assert scales == n
        np.testing.assert_array_equal(got, want)
        da.free()
        db.free()
    out.free()


def test_noalloc_refusals():
    out = DeviceBuffer(108 * 8)
    da, db = upload_pair(64, 64, 0)
    # A pair whose workspace exceeds the scratch: refused on the host.
    big = Ssimulacra2Config(size=Size(8192, 4320))
    c = big.to_c()
    ia = DeviceImage(da.ptr, Size(8192, 4320), ImageFormat.srgb8()).to_c()
    status = L.lib().vqm_ss2_score_device_noalloc(ctypes.byref(c), ctypes.byref(ia), ctypes.byref(ia), None,
                                                  ctypes.c_void_p(out.ptr), None)
    assert status == 1  # VQM_INVALID_ARGUMENT
    assert b"scratch" in L.lib().vqm_last_error()
    # Another input format, and a geometry mismatch.
    cfg = Ssimulacra2Config(size=Size(64, 64)).to_c()
    fmt = ImageFormat.srgb8()
    fmt.type = vqm_amd.PixelType.Word
    fmt.depth = 10
    ia = DeviceImage(da.ptr, Size(64, 64), fmt).to_c()
    ib = DeviceImage(db.ptr, Size(64, 64), ImageFormat.srgb8()).to_c()
    assert L.lib().vqm_ss2_score_device_noalloc(ctypes.byref(cfg), ctypes.byref(ia), ctypes.byref(ib), None,
                                                ctypes.c_void_p(out.ptr), None) == 1
    ia = DeviceImage(da.ptr, Size(32, 64), ImageFormat.srgb8()).to_c()
    assert L.lib().vqm_ss2_score_device_noalloc(ctypes.byref(cfg), ctypes.byref(ia), ctypes.byref(ib), None,
                                                ctypes.c_void_p(out.ptr), None) == 1
    da.free()
    db.free()
    out.free()
