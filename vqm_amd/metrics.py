"""The SSIMULACRA 2 slice of vqm's unified metrics API, over Ssimulacra2Pipeline.

Follows include/vqm_metrics.h:35-215, src/metrics/metrics.cc:23-398 and
src/metrics/metric_pipelines.cc:29-128,187-279 of the reference (the same
contract as the C++ header include/vqm_metrics.h of this library):

* a MetricsConfig selects the metric by holding its configuration; its size,
  streaming_inputs and background govern the metric, whose own size,
  streaming_inputs, resident slot count and input_bytes pool are overridden;
* Metrics scores an ordered sequence of frame pairs numbered from 0; Flush
  closes the sequence, Finalize reports the frame count, Reset starts a new
  one; ComputeImage scores one independent pair;
* checks of the pair (size, format, packed length, streaming) raise
  InvalidArgument before any work and leave the sequence usable; a failure
  inside scoring marks it failed (LogicError until Reset).

SSIMULACRA 2 is the only metric here, so a result is never held back and
the members for the reference's other metrics are absent. With one metric
there is no shared scratch arena: `scratch` stays None and `spatial` holds
the pipeline's own allocations. Budgets have two heaps: 0 is device memory
(HBM), 1 pinned host memory.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import Optional, Sequence

import numpy as np

from .ssimulacra2 import (DeviceImage, Image, ImageFormat, ImageView, InvalidArgument, LogicError,
                          MemoryPolicy, MetricMemoryRequirements, PipelineMemoryUsage, Size,
                          Ssimulacra2Config, Ssimulacra2Pipeline,
                          GetMemoryRequirements as _PipelineRequirements,
                          MaxWorkersForBudgetWorstCase as _WorstCase, required_bytes,
                          validate_image_format, view_bytes)

_U64_MAX = (1 << 64) - 1


@dataclass
class MetricsConfig:
    """vqm_metrics.h:35-47."""
    size: Size = Size()
    original_format: ImageFormat = field(default_factory=ImageFormat)
    distorted_format: ImageFormat = field(default_factory=ImageFormat)
    streaming_inputs: bool = True
    memory_policy: int = MemoryPolicy.Automatic
    background: float = 0.5
    ssimulacra2: Optional[Ssimulacra2Config] = None


@dataclass
class MetricsResult:
    """vqm_metrics.h:50-61: one frame's scores."""
    frame_index: int = 0
    ssimulacra2: Optional[float] = None


@dataclass
class MetricsMemoryUsage:
    """vqm_metrics.h:68-75."""
    spatial: Optional[PipelineMemoryUsage] = None
    scratch: Optional[PipelineMemoryUsage] = None


@dataclass
class MetricsMemoryRequirements:
    """vqm_metrics.h:80-87."""
    spatial: Optional[MetricMemoryRequirements] = None
    scratch: Optional[MetricMemoryRequirements] = None


@dataclass
class MetricsFinalResult:
    """vqm_metrics.h:147-152 (no CVVDP clip score here)."""
    last_frame: Optional[MetricsResult] = None
    frames: int = 0
    penultimate_frame: Optional[MetricsResult] = None


def _validate_named(fmt: ImageFormat, size: Size, name: str) -> None:
    """ValidateImage (input.cc:152-222) with the reference's '<name>: <reason>'."""
    try:
        validate_image_format(fmt, size)
    except InvalidArgument as e:
        reason = str(e)
        if reason.startswith("image: "):
            reason = reason[len("image: "):]
        raise InvalidArgument(f"{name}: {reason}") from None


def _select(config: MetricsConfig) -> Ssimulacra2Config:
    """SelectConfigs (metric_pipelines.cc:29-128): the SSIMULACRA 2
    configuration the collection builds."""
    if config.ssimulacra2 is None:
        raise InvalidArgument("metrics: select at least one metric")
    bg = float(config.background)
    if not math.isfinite(bg) or bg < 0.0 or bg > 1.0:
        raise InvalidArgument("metrics: background must be between 0 and 1")
    _validate_named(config.original_format, config.size, "original format")
    _validate_named(config.distorted_format, config.size, "distorted format")
    return replace(config.ssimulacra2, size=config.size, streaming_inputs=config.streaming_inputs,
                   memory_policy=config.memory_policy, gpu_resident_images=0, input_bytes=0)


def GetMemoryRequirements(config: MetricsConfig) -> MetricsMemoryRequirements:
    """metric_pipelines.cc:187-196: planned without touching the device."""
    return MetricsMemoryRequirements(spatial=_PipelineRequirements(_select(config)))


def MaxWorkersForBudgetWorstCase(requirements, heap_budgets: Sequence[int]) -> int:
    """metric_pipelines.cc:197-212 and vqm_memory.h:126-133: workers that fit
    the per-heap budgets (0 device, 1 pinned host), each charged for its
    requirements at their largest supported alternative. `requirements` is a
    MetricsMemoryRequirements or a sequence of MetricMemoryRequirements."""
    if isinstance(requirements, MetricsMemoryRequirements):
        requirements = [r for r in (requirements.scratch, requirements.spatial) if r is not None]
    return _WorstCase(requirements, heap_budgets)


class MetricPipelines:
    """vqm_metrics.h:97-119: owns the selected SSIMULACRA 2 pipeline,
    configured as Metrics configures its own. One collection is one worker."""

    def __init__(self, config: MetricsConfig):
        self._config = config
        self._ssimu = Ssimulacra2Pipeline(_select(config))

    def Ssimulacra2(self) -> Ssimulacra2Pipeline:
        return self._ssimu

    def GetMemoryUsage(self) -> MetricsMemoryUsage:
        return MetricsMemoryUsage(spatial=self._ssimu.GetMemoryUsage())

    def close(self) -> None:
        self._ssimu.close()


def _is_packed(x) -> bool:
    return isinstance(x, (bytes, bytearray, memoryview)) or (isinstance(x, np.ndarray) and x.dtype == np.uint8
                                                              and x.ndim == 1)


class Metrics:
    """vqm_metrics.h:154-215 (metrics.cc:78-398). Inputs: Image, ImageView,
    tightly packed bytes of the configured formats (bytes, bytearray,
    memoryview or a 1-D uint8 array of exactly RequiredBytes), or DeviceImage
    (the VulkanBufferView analogue; `wait_stream`, a hipStream_t, replaces the
    waits of VulkanSubmitSync). One Metrics is one worker, not thread-safe."""

    def __init__(self, config: MetricsConfig):
        self._config = config
        self._ssimu = Ssimulacra2Pipeline(_select(config))
        self._next = 0
        self._failed = False
        self._finished = False

    def close(self) -> None:
        self._ssimu.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- input checks, before the try of Run (metrics.cc:155-176)
    def _packed(self, data, original: bool) -> Image:
        fmt = self._config.original_format if original else self._config.distorted_format
        n = data.nbytes if isinstance(data, (np.ndarray, memoryview)) else len(data)
        if n != required_bytes(fmt, self._config.size):
            raise InvalidArgument("metrics: packed input has incorrect size")
        if isinstance(data, memoryview):
            data = np.frombuffer(data, np.uint8)
        return Image(self._config.size, fmt, data)

    def _validate(self, a, b) -> None:
        c = self._config
        if a.size != c.size or b.size != c.size:
            raise InvalidArgument("metrics: image size differs from configured size")
        if a.format != c.original_format or b.format != c.distorted_format:
            raise InvalidArgument("metrics: image format differs from configured format")
        if isinstance(a, Image):
            for img, name in ((a, "original"), (b, "distorted")):
                _validate_named(img.format, img.size, name)
                expected = required_bytes(img.format, img.size)
                got = img.data.nbytes if isinstance(img.data, np.ndarray) else len(img.data)
                if got != expected:
                    raise InvalidArgument(f"{name}: data size does not match the image format "
                                          f"(expected {expected} bytes)")
        elif isinstance(a, ImageView):
            view_bytes(a)
            view_bytes(b)
        elif isinstance(a, DeviceImage):
            _validate_named(a.format, a.size, "original")
            _validate_named(b.format, b.size, "distorted")
            if not a.ptr or not b.ptr:
                raise InvalidArgument("metrics: invalid device input")
        if not isinstance(a, DeviceImage) and not c.streaming_inputs:
            raise InvalidArgument("metrics: streaming inputs are disabled")

    def _inputs(self, original, distorted):
        if _is_packed(original) and _is_packed(distorted):
            return self._packed(original, True), self._packed(distorted, False)
        kinds = (Image, ImageView, DeviceImage)
        if not any(isinstance(original, k) and isinstance(distorted, k) for k in kinds):
            raise InvalidArgument("metrics: unsupported pair of input kinds")
        return original, distorted

    def _require_open(self) -> None:
        if self._failed:
            raise LogicError("metrics: previous call failed; call Reset")
        if self._finished:
            raise LogicError("metrics: sequence finished; call Reset")

    def _run(self, a, b, wait_stream: int) -> Optional[MetricsResult]:
        self._require_open()
        if self._next == _U64_MAX:
            raise OverflowError("metrics: too many frames")
        self._validate(a, b)
        try:
            bg = float(self._config.background)
            if isinstance(a, DeviceImage):
                score = self._ssimu.Finish(self._ssimu.Submit(a, b, bg, wait_stream)).Score()
            else:
                score = self._ssimu.Compute(a, b, bg).Score()
            result = MetricsResult(self._next, score)
            self._next += 1
            return result
        except BaseException:
            self._failed = True
            raise

    def Compute(self, original, distorted, wait_stream: int = 0) -> Optional[MetricsResult]:
        a, b = self._inputs(original, distorted)
        return self._run(a, b, wait_stream)

    def ComputeImage(self, original, distorted, wait_stream: int = 0) -> MetricsResult:
        """metrics.cc:264-285: one independent pair; requires a new or reset
        sequence and resets it afterwards."""
        a, b = self._inputs(original, distorted)
        self._require_open()
        if self._next != 0:
            raise LogicError("metrics: reset the active sequence before ComputeImage")
        first = self._run(a, b, wait_stream)
        last = self.Flush()
        if first is not None and last is not None:
            raise LogicError("metrics: unexpected extra still-image result")
        if first is None and last is None:
            raise LogicError("metrics: missing still-image result")
        self.Reset()
        return first if first is not None else last

    def Flush(self) -> Optional[MetricsResult]:
        """metrics.cc:287-308: closes the input; nothing is ever pending."""
        if self._failed:
            raise LogicError("metrics: previous call failed; call Reset")
        self._finished = True
        return None

    def Finalize(self) -> MetricsFinalResult:
        """metrics.cc:368-375."""
        tail = self.Flush()
        last = self.Flush()
        if last is None:
            return MetricsFinalResult(tail, self._next, None)
        return MetricsFinalResult(last, self._next, tail)

    def Reset(self) -> None:
        self._next = 0
        self._failed = False
        self._finished = False

    def GetMemoryUsage(self) -> MetricsMemoryUsage:
        return MetricsMemoryUsage(spatial=self._ssimu.GetMemoryUsage())
