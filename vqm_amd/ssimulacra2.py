"""Python mirror of vqm's SSIMULACRA 2 API over the HIP C ABI.

Names, argument meaning and error behaviour follow the reference's C++ public
API (include/ssimulacra2.h:17-215, include/vqm_image.h:17-166,
include/vqm_pool.h:51-86, include/vqm_alpha.h:16-27) so code and tests read
like the reference's own:

    std::invalid_argument  -> InvalidArgument (a ValueError)
    std::logic_error       -> LogicError
    std::runtime_error     -> DeviceError
    std::overflow_error    -> OverflowError

Every scoring call runs on the GPU through libvqm_hip.so; there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import enum
import threading
from dataclasses import dataclass, field, replace
from typing import List, Optional, Sequence, Union

import numpy as np

from . import _lib as L


class InvalidArgument(ValueError):
    """std::invalid_argument of the reference API."""


class LogicError(RuntimeError):
    """std::logic_error: a call made while a pair is in flight."""


class DeviceError(RuntimeError):
    """std::runtime_error: no usable device, HIP failure or allocation failure."""


def _check(status: int) -> None:
    if status == L.VQM_OK:
        return
    msg = (L.lib().vqm_last_error() or b"").decode("utf-8", "replace")
    if status == L.VQM_INVALID_ARGUMENT:
        raise InvalidArgument(msg)
    if status == L.VQM_LOGIC_ERROR:
        raise LogicError(msg)
    if status == L.VQM_OVERFLOW:
        raise OverflowError(msg)
    raise DeviceError(msg)


# ---- image model (vqm_image.h) ---------------------------------------------

class PixelType(enum.IntEnum):
    Byte = 0
    Word = 1
    Half = 2
    Float = 3


class PixelRange(enum.IntEnum):
    Limited = 0
    Full = 1


class ColorFamily(enum.IntEnum):
    Grey = 0
    Rgb = 1
    Yuv = 2


class AlphaType(enum.IntEnum):
    NoAlpha = 0
    Straight = 1
    Premultiplied = 2


class MatrixCoefficients(enum.IntEnum):
    Rgb = 0
    Bt709 = 1
    Unspecified = 2
    Fcc = 4
    Bt470Bg = 5
    St170M = 6
    St240M = 7
    YCgCo = 8
    Bt2020Ncl = 9
    Bt2020Cl = 10
    St2085YDzDx = 11
    ChromaticityDerivedNcl = 12
    ChromaticityDerivedCl = 13
    Bt2100ICtCp = 14
    St2128IptC2 = 15
    YCgCoRe = 16
    YCgCoRo = 17


class TransferCharacteristics(enum.IntEnum):
    Bt709 = 1
    Unspecified = 2
    Bt470M = 4
    Bt470Bg = 5
    Bt601 = 6
    St240M = 7
    Linear = 8
    Log100 = 9
    Log316 = 10
    Iec61966_2_4 = 11
    Bt1361 = 12
    Srgb = 13
    Bt2020_10 = 14
    Bt2020_12 = 15
    St2084 = 16
    St428 = 17
    AribB67 = 18


class ColorPrimaries(enum.IntEnum):
    Bt709 = 1
    Unspecified = 2
    Bt470M = 4
    Bt470Bg = 5
    St170M = 6
    St240M = 7
    Film = 8
    Bt2020 = 9
    St428 = 10
    St431_2 = 11
    St432_1 = 12
    Ebu3213E = 22


@dataclass(frozen=True)
class Size:
    width: int = 0
    height: int = 0


@dataclass
class ImageFormat:
    """vqm_image.h:100-120. The default is the reference's (Float sRGB RGB);
    ImageFormat.srgb8() is the 8-bit sRGB form the table-driven fast decode
    serves, every other accepted format takes the general per-pixel decode."""
    type: int = PixelType.Float
    family: int = ColorFamily.Rgb
    alpha: int = AlphaType.NoAlpha
    range: int = PixelRange.Full
    matrix: int = MatrixCoefficients.Rgb
    transfer: int = TransferCharacteristics.Srgb
    primaries: int = ColorPrimaries.Bt709
    chroma_location: int = 0
    subsample_w: int = 0
    subsample_h: int = 0
    depth: int = 0

    @staticmethod
    def srgb8() -> "ImageFormat":
        return ImageFormat(type=PixelType.Byte)

    def to_c(self) -> L.ImageFormatC:
        f = L.ImageFormatC()
        for name in ("type", "family", "alpha", "range", "matrix", "transfer", "primaries",
                     "chroma_location", "subsample_w", "subsample_h", "depth"):
            value = int(getattr(self, name))
            if not 0 <= value <= 255:
                raise InvalidArgument(f"ImageFormat.{name} out of range")
            setattr(f, name, value)
        return f


class ChromaLocation(enum.IntEnum):
    Left = 0
    Center = 1
    TopLeft = 2
    Top = 3
    BottomLeft = 4
    Bottom = 5


class ChromaFilter(enum.IntEnum):
    Bilinear = 0
    Bicubic = 1
    Lanczos = 2


class Blur(enum.IntEnum):
    """The blur of the maps: the reference's 13-tap sampled Gaussian, or the
    kernel of Cloudinary ssimulacra2's recursive Gaussian (no reference
    counterpart; include/vqm_ss2.h VQM_BLUR_*)."""
    Sampled = 0
    Recursive = 1


class MapForm(enum.IntEnum):
    """The arithmetic of the XYB transform and the SSIM' map
    (include/vqm_ss2.h VQM_MAP_*): Reference is the reference's own
    operation order (prepare.comp:81-98, scale.comp:130-142,305-333) with
    correctly rounded cube roots; Difference blurs (a - b)^2 in place of a^2
    and b^2 (closer to FP64, a fifth less blur work, no reference
    counterpart); Default is Reference except with 16-bit fixed-point (U16)
    planes."""
    Default = 0
    Reference = 1
    Difference = 2


def blur_taps(blur: int = Blur.Sampled) -> np.ndarray:
    """The 13 FP32 taps the scale kernel applies for `blur`."""
    t = (L.f32 * 13)()
    _check(L.lib().vqm_ss2_blur_taps(int(blur), t))
    return np.array(t, np.float32)


def validate_image_format(fmt: ImageFormat, size: Size) -> None:
    """ValidateImage (input.cc:152-222): raises InvalidArgument naming the reason."""
    _check(L.lib().vqm_validate_image_format(ctypes.byref(fmt.to_c()), size.width, size.height))


def plane_scale_offset(fmt: ImageFormat, chroma: bool = False, alpha: bool = False):
    """PlaneScaleOffset (input.cc:714-757): (scale, offset) as float32."""
    scale, offset = L.f32(), L.f32()
    _check(L.lib().vqm_plane_scale_offset(ctypes.byref(fmt.to_c()), int(chroma), int(alpha),
                                          ctypes.byref(scale), ctypes.byref(offset)))
    return scale.value, offset.value


def required_bytes(fmt: ImageFormat, size: Size) -> int:
    """RequiredBytes, vqm_image.h:143-146."""
    out = L.u64()
    _check(L.lib().vqm_required_bytes(ctypes.byref(fmt.to_c()), size.width, size.height,
                                      ctypes.byref(out)))
    return out.value


@dataclass
class Image:
    """vqm_image.h:134-141: packed planes, each on a 4-byte boundary. The
    format defaults to ImageFormat() (Float sRGB RGB), as in the reference."""
    size: Size
    format: ImageFormat = field(default_factory=ImageFormat)
    data: Union[bytes, bytearray, np.ndarray] = b""

    @staticmethod
    def from_planes(r: np.ndarray, g: np.ndarray, b: np.ndarray) -> "Image":
        """An 8-bit sRGB image from three HxW uint8 planes."""
        h, w = r.shape
        stride = (w * h + 3) // 4 * 4
        buf = np.zeros(3 * stride, np.uint8)
        for i, p in enumerate((r, g, b)):
            buf[i * stride:i * stride + w * h] = np.ascontiguousarray(p, np.uint8).reshape(-1)
        return Image(Size(w, h), ImageFormat.srgb8(), buf)

    def _buffer(self):
        data = self.data
        if isinstance(data, np.ndarray):
            arr = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
            return arr, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes
        if isinstance(data, bytearray):
            arr = (ctypes.c_uint8 * len(data)).from_buffer(data)
            return arr, ctypes.cast(arr, ctypes.c_void_p), len(data)
        arr = ctypes.create_string_buffer(bytes(data), len(data)) if data else None
        return arr, (ctypes.cast(arr, ctypes.c_void_p) if arr is not None else None), len(data)

    def desc(self) -> L.ImageDescC:
        d = L.ImageDescC()
        d.width, d.height = self.size.width, self.size.height
        d.format = self.format.to_c()
        d.bytes = len(self.data) if not isinstance(self.data, np.ndarray) else self.data.nbytes
        return d


@dataclass(frozen=True)
class UploadedImage:
    """vqm_image.h:151-154: a resident slot handle (0 = invalid)."""
    slot: int = 0
    size: Size = Size()


@dataclass(frozen=True)
class ImageBuffer:
    """vqm_image.h:162-166: caller-owned device image memory (0 = invalid)."""
    slot: int = 0
    size: Size = Size()
    bytes: int = 0


@dataclass(frozen=True)
class DeviceImage:
    """A caller-owned device image (replaces VulkanBufferView): a device
    pointer to packed planes, kept alive until Finish. Like the reference's
    view, its format defaults to ImageFormat()."""
    ptr: int
    size: Size
    format: ImageFormat = field(default_factory=ImageFormat)
    bytes: int = 0

    def to_c(self) -> L.DeviceImageC:
        d = L.DeviceImageC()
        d.data = self.ptr
        d.width, d.height = self.size.width, self.size.height
        d.format = self.format.to_c()
        d.bytes = self.bytes or required_bytes(self.format, self.size)
        return d


@dataclass
class ImageView:
    """vqm_views.h:18-30: a non-owning planar host image. `planes` holds the
    present planes in Image order (Y/R, U/G, V/B, alpha) as 2-D numpy arrays
    of rows, each with its own row stride (a slice of a wider array is
    fine); samples within a row must be adjacent. Absent planes may be None
    or left out."""
    size: Size
    format: ImageFormat = field(default_factory=ImageFormat)
    planes: Sequence[Optional[np.ndarray]] = ()

    def to_c(self) -> L.HostViewC:
        if len(self.planes) > 4:
            raise InvalidArgument("image view: at most four planes")
        v = L.HostViewC()
        v.width, v.height = self.size.width, self.size.height
        v.format = self.format.to_c()
        for i, plane in enumerate(self.planes):
            if plane is None:
                continue
            if plane.ndim != 2 or plane.strides[1] != plane.itemsize or plane.strides[0] < 0:
                raise InvalidArgument("image view: a plane must be a 2-D array of rows with adjacent samples")
            v.planes[i] = plane.ctypes.data
            v.strides[i] = plane.strides[0]
        return v


def view_bytes(view: ImageView) -> int:
    """DescribeImageView (image_view.cc:16-45): validates a view and returns
    its packed length. InvalidArgument for an invalid image or plane,
    OverflowError when a plane's geometry or stride is not addressable."""
    n = L.u64(0)
    _check(L.lib().vqm_copy_image(ctypes.byref(view.to_c()), None, ctypes.byref(n)))
    return n.value


def CopyImage(view: ImageView) -> Image:
    """vqm_views.h:32-36: packs a host view into an owning Image (a CPU
    copy); raises as view_bytes."""
    v = view.to_c()
    n = L.u64(view_bytes(view))
    out = np.empty(n.value, np.uint8)
    _check(L.lib().vqm_copy_image(ctypes.byref(v), out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n)))
    return Image(view.size, view.format, out)


# ---- results (ssimulacra2.h:17-33) -----------------------------------------

@dataclass
class Ssimulacra2Scale:
    avg_ssim: List[float] = field(default_factory=lambda: [0.0] * 6)
    avg_edgediff: List[float] = field(default_factory=lambda: [0.0] * 12)


@dataclass
class Ssimulacra2Result:
    scales: List[Ssimulacra2Scale] = field(default_factory=list)

    @staticmethod
    def from_norms(count: int, norms: Sequence[float]) -> "Ssimulacra2Result":
        return Ssimulacra2Result([
            Ssimulacra2Scale(list(norms[s * 18:s * 18 + 6]), list(norms[s * 18 + 6:s * 18 + 18]))
            for s in range(count)])

    def norms(self) -> np.ndarray:
        """The 108-entry flat layout of include/vqm_ss2.h."""
        out = np.zeros(108)
        for s, sc in enumerate(self.scales[:6]):
            out[s * 18:s * 18 + 6] = sc.avg_ssim
            out[s * 18 + 6:s * 18 + 18] = sc.avg_edgediff
        return out

    def Score(self) -> float:
        """score.cc:118-163; InvalidArgument for >6 scales or nonfinite norms."""
        if len(self.scales) > 6:
            raise InvalidArgument("Ssimulacra2Result has too many scales")
        flat = self.norms()
        score = L.f64()
        _check(L.lib().vqm_ss2_score(len(self.scales), flat.ctypes.data_as(ctypes.POINTER(L.f64)),
                                     ctypes.byref(score)))
        return score.value


def plan_scales(size: Size) -> List[Size]:
    """PlanSsimulacra2Scales (pipeline.cc:28-48)."""
    sizes = (L.u32 * 12)()
    count = L.u32()
    _check(L.lib().vqm_ss2_plan_scales(size.width, size.height, sizes, ctypes.byref(count)))
    return [Size(sizes[2 * i], sizes[2 * i + 1]) for i in range(count.value)]


def channel_pruned(channel: int, scale: int, scales: int) -> bool:
    """Ssimulacra2ChannelPruned (score.cc:103-114)."""
    return bool(L.lib().vqm_ss2_channel_pruned(channel, scale, scales))


def gaussian_taps() -> np.ndarray:
    t = (L.f32 * 13)()
    L.lib().vqm_ss2_gaussian_taps(t)
    return np.array(t, np.float32)


def srgb_table() -> np.ndarray:
    t = (L.f32 * 256)()
    L.lib().vqm_ss2_srgb_table(t)
    return np.array(t, np.float32)


# ---- config and memory --------------------------------------------------------

@dataclass
class Ssimulacra2Config:
    """ssimulacra2.h:35-64: the reference's fields in its order, then those
    without a reference counterpart. half_planes selects FP16 planes;
    plane_format (0 FP32, 1 FP16, 2 BF16, 3 fixed point) overrides it."""
    size: Size = Size()
    gpu_resident_images: int = 2
    half_planes: bool = False
    chroma_filter: int = 0  # ChromaFilter: 0 bilinear, 1 bicubic, 2 lanczos
    input_bytes: int = 0
    streaming_inputs: bool = True
    memory_policy: int = 0  # MemoryPolicy: 0 automatic, 1 direct, 2 staged (unavailable)
    plane_format: Optional[int] = None
    strip_rows: int = 0  # scale work-item height, a multiple of 8 (0 = chosen per geometry)
    graph: bool = False  # replay one captured HIP graph per pair (no reference field)
    blur: int = 0  # Blur: 0 sampled (the reference), 1 recursive (no reference field)
    map_form: int = 0  # MapForm: 0 default, 1 reference order, 2 difference (no reference field)
    strip_cols: int = 0  # scale work-item width, 64 or 128 (0 = 128; no reference field)

    def to_c(self) -> L.ConfigC:
        c = L.ConfigC()
        L.lib().vqm_ss2_config_init(ctypes.byref(c), self.size.width, self.size.height)
        c.graph = 1 if self.graph else 0
        c.chroma_filter = int(self.chroma_filter)
        c.blur = int(self.blur)
        c.map_form = int(self.map_form)
        c.width, c.height = self.size.width, self.size.height
        c.resident_images = self.gpu_resident_images
        fmt = self.plane_format
        if fmt is None:
            fmt = L.PLANES_F16 if self.half_planes else L.PLANES_F32
        c.plane_format = fmt
        c.streaming_inputs = 1 if self.streaming_inputs else 0
        c.strip_rows = self.strip_rows
        c.strip_cols = self.strip_cols
        c.input_bytes = self.input_bytes
        return c


class MemoryPolicy(enum.IntEnum):
    """vqm_memory_types.h:14. Device memory is the only arena path: Automatic
    and Direct select it; Staged does not exist here (a pipeline configured
    with it raises DeviceError, its plan marks every scenario unsupported)."""
    Automatic = 0
    Direct = 1
    Staged = 2


DEVICE_HEAP, PINNED_HOST_HEAP = 0, 1  # heap index = memory type


@dataclass
class MemoryAllocation:
    """vqm_memory_types.h:16-23."""
    buffer_bytes: int = 0
    allocation_bytes: int = 0
    memory_type: int = 0
    heap: int = 0


@dataclass
class PipelineMemoryUsage:
    """vqm_memory_types.h:27-35: the device memory a pipeline holds is its
    arena (heap 0), its pinned host buffers the staging (heap 1)."""
    uses_staging: bool = False
    arena: MemoryAllocation = field(default_factory=MemoryAllocation)
    staging: MemoryAllocation = field(default_factory=MemoryAllocation)
    auxiliary: List[MemoryAllocation] = field(default_factory=list)

    @staticmethod
    def from_regions(r: "PipelineMemoryRegions") -> "PipelineMemoryUsage":
        return PipelineMemoryUsage(
            uses_staging=r.streaming_bytes > 0,
            arena=MemoryAllocation(r.total_device_bytes, r.total_device_bytes, DEVICE_HEAP, DEVICE_HEAP),
            staging=MemoryAllocation(r.pinned_host_bytes, r.pinned_host_bytes, PINNED_HOST_HEAP,
                                     PINNED_HOST_HEAP))


@dataclass
class PipelineMemoryRegions:
    """The same bytes by region (vqm_ss2_memory; no reference counterpart)."""
    workspace_bytes: int = 0
    resident_bytes: int = 0
    streaming_bytes: int = 0
    input_pool_bytes: int = 0
    pinned_host_bytes: int = 0
    total_device_bytes: int = 0

    @staticmethod
    def from_c(m: L.MemoryC) -> "PipelineMemoryRegions":
        return PipelineMemoryRegions(m.workspace_bytes, m.resident_bytes, m.streaming_bytes,
                                     m.input_pool_bytes, m.pinned_host_bytes, m.total_device_bytes)


@dataclass
class MemoryScenario:
    """vqm_memory.h:26-31."""
    supported: bool = False
    arena: MemoryAllocation = field(default_factory=MemoryAllocation)
    staging: MemoryAllocation = field(default_factory=MemoryAllocation)
    auxiliary: List[MemoryAllocation] = field(default_factory=list)


@dataclass
class MemoryHeapInfo:
    """vqm_memory.h:33-42."""
    size_bytes: int = 0
    device_local: bool = False
    budget_available: bool = False
    budget_bytes: int = 0
    usage_bytes: int = 0


@dataclass
class MetricScratchRequirements:
    """vqm_memory_types.h:37-42 (one metric shares nothing)."""
    arena_bytes: int = 0
    fallback_staging_bytes: int = 0
    persistent_bytes: List[int] = field(default_factory=list)
    reusable: bool = False


@dataclass
class MetricMemoryRequirements:
    """vqm_memory.h:48-53."""
    direct: MemoryScenario = field(default_factory=MemoryScenario)
    staged: MemoryScenario = field(default_factory=MemoryScenario)
    heaps: List[MemoryHeapInfo] = field(default_factory=list)
    scratch: MetricScratchRequirements = field(default_factory=MetricScratchRequirements)


_U64_MAX = (1 << 64) - 1


def memory_plan(config: Ssimulacra2Config) -> PipelineMemoryRegions:
    """Bytes a pipeline with `config` allocates, by region (layout.cc:106-325)."""
    m = L.MemoryC()
    cfg = replace(config, memory_policy=MemoryPolicy.Automatic)
    _check(L.lib().vqm_ss2_memory_plan(ctypes.byref(cfg.to_c()), ctypes.byref(m)))
    return PipelineMemoryRegions.from_c(m)


def memory_heaps() -> List[MemoryHeapInfo]:
    """The two heaps (vqm_memory_heaps): 0 device memory, 1 pinned host."""
    h = (L.MemoryHeapC * 2)()
    _check(L.lib().vqm_memory_heaps(h))
    return [MemoryHeapInfo(x.size_bytes, bool(x.device_local), bool(x.budget_available), x.budget_bytes,
                           x.usage_bytes) for x in h]


def GetMemoryRequirements(config: Ssimulacra2Config) -> MetricMemoryRequirements:
    """vqm_memory.h:62-70: the allocations a pipeline with `config` will hold,
    planned without allocating; InvalidArgument for a configuration the
    pipeline would refuse."""
    r = memory_plan(config)
    usage = PipelineMemoryUsage.from_regions(r)
    direct = MemoryScenario(config.memory_policy != MemoryPolicy.Staged, usage.arena, usage.staging)
    return MetricMemoryRequirements(direct=direct, heaps=memory_heaps())


def AvailableHeapBudgets(requirements: MetricMemoryRequirements) -> Optional[List[int]]:
    """vqm_memory.h:55-60: remaining bytes per heap, or None without budgets."""
    out = []
    for h in requirements.heaps:
        if not h.budget_available:
            return None
        out.append(max(h.budget_bytes - h.usage_bytes, 0))
    return out


def _charge(per_heap: List[int], scenario: MemoryScenario) -> None:
    if not scenario.supported:
        raise InvalidArgument("memory: unsupported or empty scenario")
    for a in [scenario.arena, scenario.staging] + list(scenario.auxiliary):
        if a.allocation_bytes == 0:
            continue
        if a.heap >= len(per_heap):
            raise InvalidArgument("memory: missing heap budget")
        if a.allocation_bytes > _U64_MAX - per_heap[a.heap]:
            raise OverflowError("memory: heap allocation size overflow")
        per_heap[a.heap] += a.allocation_bytes


def _workers(charge: Sequence[int], budgets: Sequence[int]) -> int:
    n = _U64_MAX
    for c, b in zip(charge, budgets):
        if c:
            n = min(n, int(b) // c)
    return n


def MaxWorkersForBudget(per_worker, heap_budgets: Sequence[int]) -> int:
    """vqm_memory.h:110-124: independent pipelines (one MemoryScenario, or a
    sequence of them per worker) that fit the per-heap budgets; arithmetic
    only. InvalidArgument for an unsupported scenario or a missing heap."""
    scenarios = [per_worker] if isinstance(per_worker, MemoryScenario) else list(per_worker)
    if not scenarios:
        raise InvalidArgument("memory: no scenarios")
    charge = [0] * len(heap_budgets)
    for s in scenarios:
        _charge(charge, s)
    return _workers(charge, heap_budgets)


def MaxWorkersForBudgetWorstCase(per_worker: Sequence[MetricMemoryRequirements],
                                 heap_budgets: Sequence[int]) -> int:
    """vqm_memory.h:126-133: each worker's metrics charged at their largest
    supported alternative."""
    per_worker = list(per_worker)
    if not per_worker:
        raise InvalidArgument("memory: no requirements")
    total = [0] * len(heap_budgets)
    for req in per_worker:
        largest = [0] * len(heap_budgets)
        supported = False
        for s in (req.direct, req.staged):
            if not s.supported:
                continue
            supported = True
            charge = [0] * len(heap_budgets)
            _charge(charge, s)
            largest = [max(x, y) for x, y in zip(largest, charge)]
        if not supported:
            raise InvalidArgument("memory: unsupported requirements")
        for i, v in enumerate(largest):
            if v > _U64_MAX - total[i]:
                raise OverflowError("memory: heap allocation size overflow")
            total[i] += v
    return _workers(total, heap_budgets)


def algorithmic_bytes(size: Size, plane_format: int = L.PLANES_F32) -> int:
    out = L.u64()
    _check(L.lib().vqm_ss2_algorithmic_bytes(size.width, size.height, plane_format,
                                             ctypes.byref(out)))
    return out.value


# ---- the pipeline (ssimulacra2.h:83-215) ----------------------------------------

class Ssimulacra2Pipeline:
    """One geometry, one submission in flight, not thread-safe (use one per
    thread). All scoring runs on the GPU."""

    def __init__(self, config: Ssimulacra2Config):
        if isinstance(config, Size):
            config = Ssimulacra2Config(size=config)
        if config.memory_policy == MemoryPolicy.Staged:
            raise DeviceError("MemoryPolicy.Staged is unavailable: pipelines hold their arena in device memory")
        self.config = config
        handle = ctypes.c_void_p()
        _check(L.lib().vqm_ss2_create(ctypes.byref(config.to_c()), ctypes.byref(handle)))
        self._h = handle
        self._keep = None

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().vqm_ss2_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def size(self) -> Size:
        return self.config.size

    # -- resident images
    def Upload(self, image: Image) -> UploadedImage:
        buf, ptr, _ = image._buffer()
        handle = L.u64()
        _check(L.lib().vqm_ss2_upload(self._h, ctypes.byref(image.desc()), ptr, ctypes.byref(handle)))
        return UploadedImage(handle.value, image.size)

    def Release(self, handle: Union[UploadedImage, ImageBuffer]) -> None:
        if isinstance(handle, ImageBuffer):
            _check(L.lib().vqm_ss2_release_image(self._h, handle.slot))
        else:
            _check(L.lib().vqm_ss2_release(self._h, handle.slot))

    # -- caller-owned device buffers
    def AllocateImage(self, fmt: ImageFormat) -> ImageBuffer:
        handle, nbytes = L.u64(), L.u64()
        _check(L.lib().vqm_ss2_allocate_image(self._h, ctypes.byref(fmt.to_c()),
                                              ctypes.byref(handle), ctypes.byref(nbytes)))
        return ImageBuffer(handle.value, self.size, nbytes.value)

    def MapImage(self, buffer: ImageBuffer) -> int:
        """Device pointer of the buffer (fill it with memcpy_htod)."""
        ptr = ctypes.c_void_p()
        _check(L.lib().vqm_ss2_map_image(self._h, buffer.slot, ctypes.byref(ptr)))
        return ptr.value

    # -- split submission
    def Submit(self, original, distorted, background: float = 0.5, wait_stream: int = 0) -> int:
        token = L.u64()
        lib = L.lib()
        if isinstance(original, Image) and isinstance(distorted, Image):
            ba, pa, _ = original._buffer()
            bb, pb, _ = distorted._buffer()
            self._keep = (ba, bb)
            _check(lib.vqm_ss2_submit_host(self._h, ctypes.byref(original.desc()), pa,
                                           ctypes.byref(distorted.desc()), pb, background,
                                           ctypes.byref(token)))
        elif isinstance(original, ImageView) and isinstance(distorted, ImageView):
            va, vb = original.to_c(), distorted.to_c()
            self._keep = (original, distorted)
            _check(lib.vqm_ss2_submit_views(self._h, ctypes.byref(va), ctypes.byref(vb), background,
                                            ctypes.byref(token)))
        elif isinstance(original, UploadedImage) and isinstance(distorted, UploadedImage):
            _check(lib.vqm_ss2_submit_resident(self._h, original.slot, distorted.slot, background,
                                               ctypes.byref(token)))
        elif isinstance(original, UploadedImage) and isinstance(distorted, Image):
            bb, pb, _ = distorted._buffer()
            self._keep = (bb,)
            _check(lib.vqm_ss2_submit_resident_host(self._h, original.slot,
                                                    ctypes.byref(distorted.desc()), pb, background,
                                                    ctypes.byref(token)))
        elif isinstance(original, ImageBuffer) and isinstance(distorted, ImageBuffer):
            _check(lib.vqm_ss2_submit_buffers(self._h, original.slot, distorted.slot, background,
                                              ctypes.byref(token)))
        elif isinstance(original, DeviceImage) and isinstance(distorted, DeviceImage):
            _check(lib.vqm_ss2_submit_device(self._h, ctypes.byref(original.to_c()),
                                             ctypes.byref(distorted.to_c()),
                                             ctypes.c_void_p(wait_stream or None), background,
                                             ctypes.byref(token)))
        else:
            raise InvalidArgument("unsupported pair of input kinds")
        return token.value

    def Finish(self, token: int) -> Ssimulacra2Result:
        norms = (L.f64 * 108)()
        scales = L.u32()
        _check(L.lib().vqm_ss2_finish(self._h, token, ctypes.byref(scales), norms))
        self._keep = None
        return Ssimulacra2Result.from_norms(scales.value, list(norms))

    def Compute(self, original, distorted, background: float = 0.5) -> Ssimulacra2Result:
        return self.Finish(self.Submit(original, distorted, background))

    def GetMemoryUsage(self) -> PipelineMemoryUsage:
        """ssimulacra2.h:204-205."""
        return PipelineMemoryUsage.from_regions(self.GetMemoryRegions())

    def GetMemoryRegions(self) -> PipelineMemoryRegions:
        """The same bytes by region (no reference counterpart)."""
        m = L.MemoryC()
        _check(L.lib().vqm_ss2_memory_usage(self._h, ctypes.byref(m)))
        return PipelineMemoryRegions.from_c(m)

    def PreparedImages(self) -> int:
        """Images the last submission prepared: 1 when its resident original
        was the previous submission's and its planes were reused."""
        n = L.u32()
        _check(L.lib().vqm_ss2_prepared_images(self._h, ctypes.byref(n)))
        return n.value

    def StageTimes(self) -> List[float]:
        """Prepare, scale and reduce times of the last finished pair, in ms."""
        ms = (L.f32 * 3)()
        _check(L.lib().vqm_ss2_stage_times(self._h, ms))
        return list(ms)


def Score(pipeline: Ssimulacra2Pipeline, original: ImageView, distorted: ImageView,
          background: float = 0.5) -> float:
    """vqm_views.h:38-46: one pair of views at one background (the
    two-background alpha policy is not applied)."""
    if original.size != distorted.size:
        raise InvalidArgument("image views must have the same size")
    return pipeline.Compute(original, distorted, background).Score()


def ComputeWorstBackground(pipeline: Ssimulacra2Pipeline, original, distorted) -> Ssimulacra2Result:
    """vqm_alpha.h:16-27: opaque inputs score once; with alpha, the worse of
    the dark (0.1) and bright (0.9) backgrounds."""
    fa = getattr(original, "format", None)
    fb = getattr(distorted, "format", None)
    if (fa is None or fa.alpha == AlphaType.NoAlpha) and (fb is None or fb.alpha == AlphaType.NoAlpha):
        return pipeline.Compute(original, distorted)
    dark = pipeline.Compute(original, distorted, 0.1)
    bright = pipeline.Compute(original, distorted, 0.9)
    return dark if dark.Score() <= bright.Score() else bright


class Ssimulacra2Pool:
    """vqm_pool.h:51-86 / pool.cc:24-116: N pipelines round-robin, results in
    input order; after a failure the pending pairs are drained and the error
    re-raised. When a drain itself fails, the pipelines are dropped and the
    pool is poisoned: later calls raise LogicError (pool.cc:47-49,94-112)."""

    def __init__(self, config: Ssimulacra2Config, workers: int):
        if workers < 1:
            raise InvalidArgument("a pool needs at least one worker")
        self._pipes = [Ssimulacra2Pipeline(config) for _ in range(workers)]
        self._workers = workers
        self._poisoned = False

    def Workers(self) -> int:
        return self._workers

    def close(self) -> None:
        for p in self._pipes:
            p.close()
        self._pipes = []

    def _drain_after_failure(self, tokens) -> None:
        failed = False
        for slot, token in enumerate(tokens):
            if token is None:
                continue
            try:
                self._pipes[slot].Finish(token)
            except Exception:
                failed = True
            tokens[slot] = None
        if failed:
            self.close()
            self._poisoned = True

    def ComputeMany(self, pairs, background: float = 0.5) -> List[Ssimulacra2Result]:
        if self._poisoned:
            raise LogicError("pool cannot be reused after a failed drain")
        if not self._pipes:
            raise LogicError("the pool has been closed")
        n = len(self._pipes)
        tokens: List[Optional[int]] = [None] * n
        owner: List[int] = [-1] * n
        out: List[Optional[Ssimulacra2Result]] = [None] * len(pairs)
        try:
            for i, (a, b) in enumerate(pairs):
                slot = i % n
                if tokens[slot] is not None:
                    out[owner[slot]] = self._pipes[slot].Finish(tokens[slot])
                    tokens[slot] = None
                tokens[slot] = self._pipes[slot].Submit(a, b, background)
                owner[slot] = i
            for slot in range(n):
                if tokens[slot] is not None:
                    out[owner[slot]] = self._pipes[slot].Finish(tokens[slot])
                    tokens[slot] = None
        except Exception:
            self._drain_after_failure(tokens)
            raise
        return out  # type: ignore[return-value]

    def ScoreMany(self, pairs, background: float = 0.5) -> List[float]:
        return [r.Score() for r in self.ComputeMany(pairs, background)]


# ---- device memory plumbing ----------------------------------------------------

class DeviceBuffer:
    """Device memory through the C ABI (hipMalloc/hipMemcpy)."""

    def __init__(self, nbytes: int):
        ptr = ctypes.c_void_p()
        _check(L.lib().vqm_device_malloc(nbytes, ctypes.byref(ptr)))
        self.ptr = ptr.value
        self.nbytes = nbytes

    def upload(self, data, offset: int = 0) -> None:
        arr = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
        if offset + arr.nbytes > self.nbytes:
            raise InvalidArgument("upload beyond the buffer")
        _check(L.lib().vqm_memcpy_htod(ctypes.c_void_p(self.ptr + offset),
                                       arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes))

    def download(self, nbytes: Optional[int] = None, offset: int = 0) -> np.ndarray:
        n = self.nbytes - offset if nbytes is None else nbytes
        out = np.empty(n, np.uint8)
        _check(L.lib().vqm_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p),
                                       ctypes.c_void_p(self.ptr + offset), n))
        return out

    def free(self) -> None:
        if self.ptr:
            L.lib().vqm_device_free(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def memcpy_htod(dst: int, data) -> None:
    arr = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
    _check(L.lib().vqm_memcpy_htod(ctypes.c_void_p(dst), arr.ctypes.data_as(ctypes.c_void_p),
                                   arr.nbytes))


def device_count() -> int:
    n = ctypes.c_int()
    _check(L.lib().vqm_device_count(ctypes.byref(n)))
    return n.value


def device_info() -> dict:
    name = ctypes.create_string_buffer(256)
    arch = ctypes.create_string_buffer(64)
    cus, lds = ctypes.c_int(), ctypes.c_int()
    hbm = L.u64()
    _check(L.lib().vqm_device_info(name, arch, ctypes.byref(cus), ctypes.byref(lds),
                                   ctypes.byref(hbm)))
    return {"name": name.value.decode(), "arch": arch.value.decode(), "cus": cus.value,
            "lds_per_block": lds.value, "hbm_bytes": hbm.value}


# ---- batched scoring ------------------------------------------------------------

class Ssimulacra2Batch:
    """Many pairs of any geometries per launch (no reference counterpart).
    Pairs are device-resident (DeviceImage pointers). `graph`: True replays
    every qualifying submission from a captured HIP graph, False none, and
    "auto" (the default) those of up to 8 pairs, where the launch overhead
    is a visible share of the time."""

    def __init__(self, max_pairs: int, plane_format: int = L.PLANES_F32, hbm_budget: int = 0,
                 strip_rows: int = 0, chunks: int = 1, autotune: bool = False, blur: int = 0,
                 persistent: bool = False, map_form: int = 0, strip_cols: int = 0,
                 graph: Union[bool, str] = "auto"):
        cfg = L.BatchConfigC()
        cfg.strip_cols = int(strip_cols)
        if graph == "auto":
            cfg.graph = 2
        elif isinstance(graph, bool):
            cfg.graph = 1 if graph else 0
        else:
            raise InvalidArgument('graph must be True, False or "auto"')
        cfg.blur = int(blur)
        cfg.map_form = int(map_form)
        cfg.persistent = 1 if persistent else 0
        cfg.plane_format = plane_format
        cfg.strip_rows = strip_rows
        cfg.chunks = chunks
        cfg.autotune = 1 if autotune else 0
        cfg.max_pairs = max_pairs
        cfg.hbm_budget = hbm_budget
        h = ctypes.c_void_p()
        _check(L.lib().vqm_ss2_batch_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.max_pairs = max_pairs
        self._count = 0

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().vqm_ss2_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def SetShape(self, size: Size, strip_rows: int = 0, strip_cols: int = 0) -> None:
        """Fixes the scale strip shape of one geometry before its first
        submission (vqm_ss2_batch_set_shape)."""
        _check(L.lib().vqm_ss2_batch_set_shape(self._h, size.width, size.height, int(strip_rows),
                                               int(strip_cols)))

    def Submit(self, pairs: Sequence[tuple]) -> int:
        """pairs: (original_ptr, distorted_ptr, width, height) tuples."""
        arr = (L.DevicePairC * len(pairs))()
        for i, (a, b, w, h) in enumerate(pairs):
            arr[i].original, arr[i].distorted, arr[i].width, arr[i].height = a, b, w, h
        token = L.u64()
        _check(L.lib().vqm_ss2_batch_submit(self._h, arr, len(pairs), ctypes.byref(token)))
        self._count = len(pairs)
        return token.value

    def Finish(self, token: int, raw: bool = False):
        n = self._count
        norms = np.zeros(n * 108)
        scales = np.zeros(n, np.uint32)
        _check(L.lib().vqm_ss2_batch_finish(self._h, token,
                                            scales.ctypes.data_as(ctypes.POINTER(L.u32)),
                                            norms.ctypes.data_as(ctypes.POINTER(L.f64))))
        if raw:
            return scales, norms.reshape(n, 108)
        return [Ssimulacra2Result.from_norms(int(scales[i]), norms[i * 108:(i + 1) * 108])
                for i in range(n)]

    def StageTimes(self) -> List[float]:
        ms = (L.f32 * 3)()
        _check(L.lib().vqm_ss2_batch_stage_times(self._h, ms))
        return list(ms)

    def PoolBytes(self) -> int:
        b = L.u64()
        _check(L.lib().vqm_ss2_batch_pool_bytes(self._h, ctypes.byref(b)))
        return b.value


class PinnedBuffer:
    """Page-locked host memory through the C ABI (hipHostMalloc), viewed as a
    numpy uint8 array: host inputs in it are copied to the device
    asynchronously."""

    def __init__(self, nbytes: int):
        ptr = ctypes.c_void_p()
        _check(L.lib().vqm_host_malloc(nbytes, ctypes.byref(ptr)))
        self.ptr = ptr.value
        self.nbytes = nbytes
        self.array = np.ctypeslib.as_array((ctypes.c_uint8 * nbytes).from_address(self.ptr))

    def free(self) -> None:
        if self.ptr:
            self.array = None
            L.lib().vqm_host_free(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def autotune(size: Size, plane_format: int = L.PLANES_F32, pairs: int = 1, map_form: int = 0) -> dict:
    """The launch shape (strip height, chunk count) measured fastest for a
    batch of `pairs` pairs of this size, plane format and map form
    (vqm_ss2_autotune)."""
    r = L.TuneResultC()
    _check(L.lib().vqm_ss2_autotune(size.width, size.height, plane_format, int(map_form), pairs,
                                    ctypes.byref(r)))
    return {name: getattr(r, name) for name, _ in L.TuneResultC._fields_}


# ---- serving ----------------------------------------------------------------------

@dataclass
class RequestResult:
    """Outcome of one served request. Times are vqm_now_ns() readings; for a
    replay, submit_ns is the scheduled arrival time."""
    status: int
    scales: int
    batch_size: int
    lane: int
    submit_ns: int
    launch_ns: int
    done_ns: int
    norms: np.ndarray

    @staticmethod
    def from_c(r: "L.RequestResultC") -> "RequestResult":
        return RequestResult(r.status, r.scales, r.batch_size, r.lane, r.submit_ns, r.launch_ns,
                             r.done_ns, np.ctypeslib.as_array(r.norms).copy())

    @property
    def latency_ms(self) -> float:
        return (self.done_ns - self.submit_ns) * 1e-6

    def result(self) -> Ssimulacra2Result:
        return Ssimulacra2Result.from_norms(self.scales, self.norms)


def now_ns() -> int:
    return int(L.lib().vqm_now_ns())


class Ssimulacra2Server:
    """Request server with continuous batching (no reference counterpart; the
    reference overlaps pairs through Ssimulacra2Pool workers, vqm_pool.h:35-86).

    Requests of any geometry are queued; batches form from the queue head
    whenever one of `lanes` HIP streams is free, admitted in arrival order
    against `hbm_budget`, and overlap on the device. Thread-safe."""

    def __init__(self, plane_format: int = L.PLANES_F32, lanes: int = 3, max_batch: int = 64,
                 hbm_budget: int = 0, strip_rows: int = 0, blur: int = 0, map_form: int = 0,
                 strip_cols: int = 0, graph: bool = False):
        cfg = L.ServerConfigC()
        L.lib().vqm_ss2_server_config_init(ctypes.byref(cfg))
        cfg.strip_cols = int(strip_cols)
        cfg.graph = 1 if graph else 0
        cfg.blur = int(blur)
        cfg.map_form = int(map_form)
        cfg.plane_format = plane_format
        cfg.lanes = lanes
        cfg.max_batch = max_batch
        cfg.hbm_budget = hbm_budget
        cfg.strip_rows = strip_rows
        h = ctypes.c_void_p()
        self._keep = {}
        _check(L.lib().vqm_ss2_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().vqm_ss2_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def SubmitDevice(self, original: int, distorted: int, width: int, height: int) -> int:
        pair = L.DevicePairC(original, distorted, width, height)
        ticket = L.u64()
        _check(L.lib().vqm_ss2_server_submit_device(self._h, ctypes.byref(pair), ctypes.byref(ticket)))
        return ticket.value

    def SubmitHost(self, original: Image, distorted: Image) -> int:
        """The server copies the host images when the request's batch
        launches; this object keeps them alive until Wait."""
        keep, ptrs = [], []
        for img in (original, distorted):
            arr, ptr, _ = img._buffer()
            keep.append(arr)
            ptrs.append(ptr)
        da, db = original.desc(), distorted.desc()
        ticket = L.u64()
        _check(L.lib().vqm_ss2_server_submit_host(self._h, ctypes.byref(da), ptrs[0],
                                                  ctypes.byref(db), ptrs[1], ctypes.byref(ticket)))
        self._keep[ticket.value] = (original, distorted, keep)
        return ticket.value

    def Wait(self, ticket: int) -> RequestResult:
        out = L.RequestResultC()
        try:
            _check(L.lib().vqm_ss2_server_wait(self._h, ticket, ctypes.byref(out)))
        finally:
            self._keep.pop(ticket, None)
        return RequestResult.from_c(out)

    def Stats(self) -> dict:
        s = L.ServerStatsC()
        _check(L.lib().vqm_ss2_server_get_stats(self._h, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in L.ServerStatsC._fields_}

    def Replay(self, requests: Sequence[tuple], check: bool = True) -> List[RequestResult]:
        """Open-loop replay. requests: (original_ptr, distorted_ptr, width,
        height, host, arrival_ns) tuples; host=True means host pointers."""
        n = len(requests)
        arr = (L.ReplayRequestC * n)()
        for i, (a, b, w, h, host, t) in enumerate(requests):
            arr[i].original, arr[i].distorted = a, b
            arr[i].width, arr[i].height, arr[i].host, arr[i].arrival_ns = w, h, int(bool(host)), int(t)
        out = (L.RequestResultC * n)()
        status = L.lib().vqm_ss2_server_replay(self._h, arr, n, out)
        if check:
            _check(status)
        return [RequestResult.from_c(out[i]) for i in range(n)]
