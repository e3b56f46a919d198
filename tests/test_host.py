"""Host logic of the product library through its C ABI, with no GPU: the
library loads, exports every declared symbol, and its host-side functions
(scale plan, taps, pruning, score mapping, decode table, validation, memory
plan) agree with the oracle and the reference's contracts."""
from __future__ import annotations

import ctypes
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import vqm_amd
from vqm_amd import _lib as L
from vqm_amd import (ImageFormat, InvalidArgument, PixelType, Size, Ssimulacra2Config,
                     Ssimulacra2Result, Ssimulacra2Scale)


def declared_functions(header: str):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vqm_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    declared = declared_functions("vqm_ss2.h")
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(lib, name), f"libvqm_hip.so does not export {name}"
    assert set(declared) == set(L.SIGNATURES), "ctypes signatures and the header disagree"


def test_batch_graph_modes_are_validated_before_any_device_work():
    """graph is 0 (eager), 1 (every qualifying submission) or 2 (submissions
    of up to 8 pairs); other values and other Python spellings are refused
    before the device is touched."""
    for graph in (3, 7):
        cfg = L.BatchConfigC()
        cfg.max_pairs = 4
        cfg.graph = graph
        h = ctypes.c_void_p()
        assert L.lib().vqm_ss2_batch_create(ctypes.byref(cfg), ctypes.byref(h)) == L.VQM_INVALID_ARGUMENT
        assert b"graph must be 0, 1 or 2" in L.lib().vqm_last_error()
    with pytest.raises(InvalidArgument, match="graph must be"):
        vqm_amd.Ssimulacra2Batch(max_pairs=4, graph="always")


def test_div_check_rejects_ranges_outside_the_significands():
    r = L.DivCheckC()
    for first, end in [(0, 0), (5, 5), (10, 3), (0, (1 << 23) + 1)]:
        assert L.lib().vqm_ss2_check_div(first, end, ctypes.byref(r)) == L.VQM_INVALID_ARGUMENT
        assert b"significand range" in L.lib().vqm_last_error()


def test_cbrt_check_rejects_ranges_outside_the_normal_floats():
    """The device cube-root check validates its range before touching the GPU."""
    r = L.CbrtCheckC()
    for first, end in [(0, 0x00900000), (0x00800000, 0x7F800001), (0x3F800000, 0x3F800000),
                       (0x40000000, 0x3F800000)]:
        assert L.lib().vqm_ss2_check_cbrt(first, end, ctypes.byref(r)) == L.VQM_INVALID_ARGUMENT
        assert b"positive normal floats" in L.lib().vqm_last_error()
    assert L.lib().vqm_ss2_check_cbrt(0x3F800000, 0x40000000, None) == L.VQM_INVALID_ARGUMENT


def test_scale_plan_matches_oracle(oracle_lib):
    for size in [(8, 8), (13, 9), (32, 32), (64, 64), (65, 65), (129, 97), (1920, 1080),
                 (3840, 2160), (640, 360), (1280, 720), (2560, 1440), (512, 512), (4096, 17)]:
        got = [(s.width, s.height) for s in vqm_amd.plan_scales(Size(*size))]
        assert got == oracle_lib.plan_scales(*size)
    with pytest.raises(InvalidArgument, match="8x8"):
        vqm_amd.plan_scales(Size(4, 8))


def test_taps_and_pruning_match_oracle(oracle_lib):
    assert np.array_equal(vqm_amd.gaussian_taps(), oracle_lib.gaussian_taps())
    for scales in range(1, 7):
        for s in range(scales):
            for c in range(3):
                assert vqm_amd.channel_pruned(c, s, scales) == oracle_lib.channel_pruned(c, s, scales)
    assert not vqm_amd.channel_pruned(0, 6, 6)  # out of range is never pruned
    assert not vqm_amd.channel_pruned(3, 0, 6)


def test_decode_table_is_the_oracles_bit_for_bit(oracle_lib):
    """The table the kernels read equals the oracle's FP32 decode exactly."""
    assert np.array_equal(vqm_amd.srgb_table(), oracle_lib.srgb_lut_f32())


def test_score_matches_oracle(oracle_lib):
    rng = np.random.default_rng(5)
    for scales in range(0, 7):
        norms = np.zeros(108)
        norms[:scales * 18] = rng.random(scales * 18) * 0.05
        r = Ssimulacra2Result.from_norms(scales, norms)
        assert r.Score() == oracle_lib.score(scales, norms)


def test_score_contract():
    """score_test.cc:11-46 through the product's Ssimulacra2Result."""
    assert Ssimulacra2Result().Score() == 100.0
    r = Ssimulacra2Result([Ssimulacra2Scale() for _ in range(6)])
    assert r.Score() == 100.0
    for bad in (math.inf, -math.inf, math.nan):
        for s in range(6):
            for n in range(6):
                r.scales[s].avg_ssim[n] = bad
                with pytest.raises(InvalidArgument):
                    r.Score()
                r.scales[s].avg_ssim[n] = 0.0
            for n in range(12):
                r.scales[s].avg_edgediff[n] = bad
                with pytest.raises(InvalidArgument):
                    r.Score()
                r.scales[s].avg_edgediff[n] = 0.0
    r.scales[0].avg_ssim[3] = np.finfo(np.float64).max
    with pytest.raises(InvalidArgument):
        r.Score()
    r.scales[0].avg_ssim[3] = 1e200
    with pytest.raises(InvalidArgument):
        r.Score()
    r.scales[0].avg_ssim[3] = 100.0
    assert math.isfinite(r.Score()) and r.Score() < 0.0
    with pytest.raises(InvalidArgument):
        Ssimulacra2Result([Ssimulacra2Scale() for _ in range(7)]).Score()


def test_required_bytes():
    """input.cc:114-145,244-255: planes start on 4-byte boundaries."""
    f = ImageFormat.srgb8()
    assert vqm_amd.required_bytes(f, Size(32, 32)) == 3 * 1024
    assert vqm_amd.required_bytes(f, Size(13, 9)) == 3 * 120  # 117 -> 120
    assert vqm_amd.required_bytes(f, Size(1920, 1080)) == 3 * 1920 * 1080
    assert vqm_amd.required_bytes(ImageFormat(), Size(3, 3)) == 3 * 36  # Float samples
    grey = ImageFormat(type=PixelType.Byte, family=0, matrix=1)
    assert vqm_amd.required_bytes(grey, Size(5, 5)) == 28
    with pytest.raises(InvalidArgument):
        vqm_amd.required_bytes(ImageFormat(type=7), Size(8, 8))


def test_config_validation_before_device_work():
    """pipeline_test.cc:318-328: an impossible geometry names the minimum
    before any device work, so it holds with or without a GPU."""
    for size in (Size(4, 8), Size(8, 4), Size(0, 0)):
        with pytest.raises(InvalidArgument, match="8x8"):
            vqm_amd.Ssimulacra2Pipeline(Ssimulacra2Config(size=size))
    with pytest.raises(InvalidArgument):
        vqm_amd.Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(32, 32), plane_format=5))
    with pytest.raises(InvalidArgument):
        vqm_amd.Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(32, 32), gpu_resident_images=5000))


def test_memory_plan_and_algorithmic_bytes():
    """The pruned plane set at six scales has 1.684 N active channel pixels;
    resident pair traffic is 6N + 4 * plane bytes (32.94 B/px FP32,
    19.47 B/px FP16 at 1080p)."""
    n = 1920 * 1080
    b32 = vqm_amd.algorithmic_bytes(Size(1920, 1080), L.PLANES_F32)
    b16 = vqm_amd.algorithmic_bytes(Size(1920, 1080), L.PLANES_F16)
    assert b32 / n == pytest.approx(32.94, abs=0.01)
    assert b16 / n == pytest.approx(19.47, abs=0.01)
    assert b32 == pytest.approx(68.30e6, rel=1e-3)
    # The mixed form: 2-byte fixed-point elements at every level, like FP16.
    sizes = vqm_amd.plan_scales(Size(1920, 1080))
    stored = [sum(not vqm_amd.channel_pruned(c, l, len(sizes)) for c in range(3)) for l in range(len(sizes))]
    mixed_planes = sum(k * s.width * s.height * 2 for k, s in zip(stored, sizes))
    assert vqm_amd.algorithmic_bytes(Size(1920, 1080), L.PLANES_U16) == 6 * n + 4 * mixed_planes == b16
    m = vqm_amd.memory_plan(Ssimulacra2Config(size=Size(1920, 1080)))
    assert m.resident_bytes == 2 * 3 * n
    assert m.streaming_bytes == 2 * 3 * n
    # Padded planes of both images: at least the 1.684 N active pixels each.
    assert m.workspace_bytes >= 2 * 4 * 1.684 * n
    assert m.total_device_bytes == m.workspace_bytes + m.resident_bytes + m.streaming_bytes
    half = vqm_amd.memory_plan(Ssimulacra2Config(size=Size(1920, 1080), half_planes=True))
    assert half.workspace_bytes < 0.6 * m.workspace_bytes
    # vqm_memory.h:26-135 under the reference's names and shapes: the device
    # bytes are the direct scenario's arena (heap 0), the pinned host bytes
    # its staging (heap 1); there is no staged scenario.
    req = vqm_amd.GetMemoryRequirements(Ssimulacra2Config(size=Size(1920, 1080)))
    assert req.direct.supported and not req.staged.supported
    assert req.direct.arena.allocation_bytes == m.total_device_bytes and req.direct.arena.heap == 0
    assert req.direct.staging.allocation_bytes == m.pinned_host_bytes and req.direct.staging.heap == 1
    assert len(req.heaps) == 2 and req.heaps[0].device_local and not req.heaps[1].device_local
    dev, host = m.total_device_bytes, m.pinned_host_bytes
    assert vqm_amd.MaxWorkersForBudget(req.direct, [10 * dev + 1, 100 * host]) == 10
    assert vqm_amd.MaxWorkersForBudget(req.direct, [dev - 1, 100 * host]) == 0
    assert vqm_amd.MaxWorkersForBudget([req.direct, req.direct], [10 * dev, 100 * host]) == 5
    with pytest.raises(vqm_amd.InvalidArgument, match="missing heap budget"):
        vqm_amd.MaxWorkersForBudget(req.direct, [10 * dev])
    with pytest.raises(vqm_amd.InvalidArgument, match="unsupported"):
        vqm_amd.MaxWorkersForBudget(req.staged, [1 << 40, 1 << 40])
    assert vqm_amd.MaxWorkersForBudgetWorstCase([req], [3 * dev, 100 * host]) == 3
    staged = vqm_amd.GetMemoryRequirements(Ssimulacra2Config(size=Size(1920, 1080),
                                                             memory_policy=vqm_amd.MemoryPolicy.Staged))
    assert not staged.direct.supported
    with pytest.raises(vqm_amd.InvalidArgument, match="unsupported requirements"):
        vqm_amd.MaxWorkersForBudgetWorstCase([staged], [1 << 40, 1 << 40])
    with pytest.raises(vqm_amd.DeviceError):
        vqm_amd.Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(64, 64), memory_policy=vqm_amd.MemoryPolicy.Staged))
    with pytest.raises(vqm_amd.InvalidArgument):
        vqm_amd.GetMemoryRequirements(Ssimulacra2Config(size=Size(4, 8)))


def test_last_error_is_thread_local_and_cleared():
    lib = L.lib()
    sizes = (L.u32 * 12)()
    count = L.u32()
    assert lib.vqm_ss2_plan_scales(4, 4, sizes, ctypes.byref(count)) == L.VQM_INVALID_ARGUMENT
    assert b"8x8" in lib.vqm_last_error()
    assert lib.vqm_ss2_plan_scales(8, 8, sizes, ctypes.byref(count)) == L.VQM_OK
    assert lib.vqm_last_error() == b""


def test_server_config_validation_before_device_work():
    """The server rejects a bad configuration before touching the device."""
    for kwargs in ({"lanes": 17}, {"max_batch": 5000}, {"plane_format": 4}, {"strip_rows": 6}, {"blur": 2}):
        with pytest.raises(InvalidArgument):
            vqm_amd.Ssimulacra2Server(**kwargs)
    with pytest.raises(InvalidArgument):
        vqm_amd.Ssimulacra2Batch(max_pairs=1, blur=2)
    with pytest.raises(InvalidArgument):
        vqm_amd.Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(32, 32), blur=2))
    cfg = L.ServerConfigC()
    L.lib().vqm_ss2_server_config_init(ctypes.byref(cfg))
    assert (cfg.plane_format, cfg.lanes, cfg.max_batch, cfg.strip_rows, cfg.hbm_budget) == (0, 3, 64, 0, 0)
    t0 = vqm_amd.now_ns()
    assert vqm_amd.now_ns() >= t0 > 0


class _FakePipeline:
    """Stands in for a pipeline: Submit fails for pair `bad`, Finish fails
    when `finish_fails`."""

    def __init__(self, bad=None, finish_fails=False):
        self.bad, self.finish_fails, self.pending, self.closed = bad, finish_fails, None, False

    def Submit(self, a, b, background=0.5):
        if a == self.bad:
            raise vqm_amd.InvalidArgument("pair refused")
        self.pending = a
        return a + 1

    def Finish(self, token):
        self.pending = None
        if self.finish_fails:
            raise vqm_amd.DeviceError("wait failed")
        return token

    def close(self):
        self.closed = True


def _fake_pool(pipes):
    pool = vqm_amd.Ssimulacra2Pool.__new__(vqm_amd.Ssimulacra2Pool)
    pool._pipes, pool._workers, pool._poisoned = pipes, len(pipes), False
    return pool


def test_pool_drains_and_poisons_like_the_reference():
    """pool.cc:47-49,94-112: a failing pair drains the queued submissions
    and rethrows; when a drain wait fails too, the pipelines are dropped and
    every later call raises LogicError."""
    ok = _fake_pool([_FakePipeline(bad=2), _FakePipeline()])
    with pytest.raises(vqm_amd.InvalidArgument, match="pair refused"):
        ok.ComputeMany([(0, 0), (1, 1), (2, 2), (3, 3)])
    assert all(p.pending is None for p in ok._pipes)  # drained
    assert ok.ComputeMany([(4, 4), (5, 5)]) == [5, 6]  # still usable
    bad = _fake_pool([_FakePipeline(bad=2), _FakePipeline(finish_fails=True)])
    pipes = list(bad._pipes)
    with pytest.raises(vqm_amd.InvalidArgument, match="pair refused"):
        bad.ComputeMany([(0, 0), (1, 1), (2, 2)])
    assert all(p.closed for p in pipes)
    with pytest.raises(vqm_amd.LogicError, match="failed drain"):
        bad.ComputeMany([(0, 0)])
