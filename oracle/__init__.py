"""Python loader of the CPU oracle (libss2_oracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker; the product path (vqm_amd) never does.
See ss2_oracle.h for the three arithmetic forms and what pins each one.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libss2_oracle.so")


class Mode(ctypes.Structure):
    _fields_ = [("arith", ctypes.c_int32), ("blur_order", ctypes.c_int32),
                ("cbrt_kind", ctypes.c_int32), ("plane_format", ctypes.c_int32),
                ("plane_round_levels", ctypes.c_uint32), ("sum_blocks", ctypes.c_int32),
                ("round_linear", ctypes.c_int32), ("blur_kind", ctypes.c_int32),
                ("sum_variances", ctypes.c_int32)]


# The three forms (ss2_oracle.h).
F64 = Mode(0, 0, 0, 0, 0, 0)
F32_REF = Mode(1, 0, 0, 0, 0, 1)
F32_DEV = Mode(1, 0, 1, 0, 0, 1, 0, 0, 1)


def f32_dev_planes(plane_format: int, levels_mask: int = 0x3F) -> Mode:
    """The device form with the stored planes of the levels in `levels_mask`
    rounded to FP16 (1), BF16 (2) or unsigned 16-bit fixed point (3)."""
    return Mode(1, 0, 1, plane_format, levels_mask, 1, 0, 0, 1)


def f32_dev_for(plane_format: int) -> Mode:
    """The device form of a library plane format: FP32 (0), FP16 (1), BF16 (2)
    or unsigned 16-bit fixed point (3, the mixed form) planes at every level.
    FP16 is the reference's half-plane build: its linear box averages are
    rounded to FP16 as well."""
    if plane_format == 0:
        return F32_DEV
    mode = f32_dev_planes(plane_format, 0x3F)
    mode.round_linear = 1 if plane_format == 1 else 0
    return mode


# The library's map forms (include/vqm_ss2.h VQM_MAP_*).
MAP_DEFAULT, MAP_REFERENCE, MAP_DIFFERENCE = 0, 1, 2


def form_for(plane_format: int = 0, map_form: int = MAP_DEFAULT, blur: int = 0) -> Mode:
    """The oracle form the HIP path reproduces bit for bit per pixel under a
    library configuration: the reference form (the FP32 ref form, with the
    stored planes rounded to `plane_format`) or the difference form (the
    FP32 dev form); the default map form is the reference form except for
    16-bit fixed-point planes (plane_format 3)."""
    if map_form == MAP_DEFAULT:
        map_form = MAP_DIFFERENCE if plane_format == 3 else MAP_REFERENCE
    if map_form == MAP_DIFFERENCE:
        mode = Mode.from_buffer_copy(f32_dev_for(plane_format))
    else:
        mode = Mode(1, 0, 0, plane_format, 0x3F if plane_format else 0, 1,
                    1 if plane_format == 1 else 0, 0, 0)
    mode.blur_kind = blur
    return mode


def build() -> None:
    subprocess.run(["make", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        u8p, f32p, f64p, u32p = P(ctypes.c_uint8), P(ctypes.c_float), P(ctypes.c_double), P(ctypes.c_uint32)
        sig = {
            "ss2o_round_plane": (ctypes.c_float, [ctypes.c_int32, ctypes.c_float]),
            "ss2o_plan_scales": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, u32p]),
            "ss2o_gaussian_taps": (None, [f32p]),
            "ss2o_channel_pruned": (ctypes.c_int32, [ctypes.c_int32] * 3),
            "ss2o_score": (ctypes.c_double, [ctypes.c_int32, f64p, P(ctypes.c_int32)]),
            "ss2o_srgb_lut_f32": (None, [f32p]),
            "ss2o_srgb_lut_f64": (None, [f64p]),
            "ss2o_cbrt_dev": (ctypes.c_float, [ctypes.c_float]),
            "ss2o_cbrt_rn": (ctypes.c_float, [ctypes.c_float]),
            "ss2o_rcp_dev": (ctypes.c_float, [ctypes.c_float]),
            "ss2o_xyb_f32": (None, [ctypes.c_float] * 3 + [ctypes.c_int32, f32p]),
            "ss2o_xyb_f64": (None, [ctypes.c_double] * 3 + [f64p]),
            "ss2o_blur_f32": (None, [f32p, f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32]),
            "ss2o_blur_f64": (None, [f64p, f64p, ctypes.c_uint32, ctypes.c_uint32]),
            "ss2o_pyramid_f32": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, u8p, f32p]),
            "ss2o_cosine_taps": (None, [f64p]),
            "ss2o_xyb_planes": (ctypes.c_int32, [ctypes.POINTER(Mode), ctypes.c_uint32, ctypes.c_uint32, u8p,
                                                 f32p]),
            "ss2o_pyramid_f64": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, u8p, f64p]),
            "ss2o_image_bytes": (ctypes.c_uint64, [ctypes.c_uint32, ctypes.c_uint32]),
            "ss2o_compute": (ctypes.c_int32, [P(Mode), ctypes.c_uint32, ctypes.c_uint32, u8p, u8p,
                                              f64p, f64p]),
            "ss2o_compute_many": (ctypes.c_int32, [P(Mode), ctypes.c_uint32, ctypes.c_uint32,
                                                   ctypes.c_uint32, P(u8p), P(u8p), f64p,
                                                   ctypes.c_uint32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(h, name)
            fn.restype, fn.argtypes = res, args
        _lib = h
    return _lib


def _u8(a: np.ndarray):
    return np.ascontiguousarray(a, np.uint8).ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def compute(mode: Mode, width: int, height: int, image0: np.ndarray,
            image1: np.ndarray) -> Tuple[int, np.ndarray, np.ndarray]:
    """(levels, norms[108], sums[108]) of one pair of packed u8 images."""
    a = np.ascontiguousarray(image0, np.uint8)
    b = np.ascontiguousarray(image1, np.uint8)
    need = image_bytes(width, height)
    if a.nbytes != need or b.nbytes != need:
        raise ValueError("image byte length does not match the geometry")
    norms = np.zeros(108)
    sums = np.zeros(108)
    n = lib().ss2o_compute(ctypes.byref(mode), width, height, _u8(a), _u8(b),
                           norms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           sums.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if n < 0:
        raise ValueError("minimum image size is 8x8 pixels")
    return n, norms, sums


def compute_many(mode: Mode, width: int, height: int, pairs, threads: int) -> np.ndarray:
    """norms[len(pairs), 108] with one pair per host thread at a time."""
    n = len(pairs)
    keep = [(np.ascontiguousarray(a, np.uint8), np.ascontiguousarray(b, np.uint8)) for a, b in pairs]
    P = ctypes.POINTER(ctypes.c_uint8)
    arr0 = (P * n)(*[_u8(a) for a, _ in keep])
    arr1 = (P * n)(*[_u8(b) for _, b in keep])
    out = np.zeros((n, 108))
    r = lib().ss2o_compute_many(ctypes.byref(mode), width, height, n, arr0, arr1,
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), threads)
    if r < 0:
        raise ValueError("oracle failed")
    return out


def score(levels: int, norms: np.ndarray) -> float:
    st = ctypes.c_int32()
    n = np.ascontiguousarray(norms, np.float64)
    v = lib().ss2o_score(levels, n.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(st))
    if st.value != 0:
        raise ValueError("invalid norms for the score mapping")
    return v


def plan_scales(width: int, height: int):
    sizes = (ctypes.c_uint32 * 12)()
    n = lib().ss2o_plan_scales(width, height, sizes)
    if n < 0:
        raise ValueError("minimum image size is 8x8 pixels")
    return [(sizes[2 * i], sizes[2 * i + 1]) for i in range(n)]


def cosine_taps() -> np.ndarray:
    """The recursive Gaussian's truncated-cosine kernel (FP64, 13-tap frame)."""
    out = np.zeros(13)
    lib().ss2o_cosine_taps(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def with_blur(mode: Mode, blur_kind: int) -> Mode:
    """A copy of `mode` with another blur kernel (0 sampled, 1 recursive)."""
    m = Mode.from_buffer_copy(mode)
    m.blur_kind = blur_kind
    return m


def gaussian_taps() -> np.ndarray:
    t = np.zeros(13, np.float32)
    lib().ss2o_gaussian_taps(t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return t


def channel_pruned(channel: int, scale: int, scales: int) -> bool:
    return bool(lib().ss2o_channel_pruned(channel, scale, scales))


def srgb_lut_f32() -> np.ndarray:
    t = np.zeros(256, np.float32)
    lib().ss2o_srgb_lut_f32(t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return t


def srgb_lut_f64() -> np.ndarray:
    t = np.zeros(256, np.float64)
    lib().ss2o_srgb_lut_f64(t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return t


def cbrt_dev(x: float) -> float:
    return float(lib().ss2o_cbrt_dev(x))


def cbrt_rn(x: float) -> float:
    """The cube root correctly rounded to FP32 (the FP32 ref form's root)."""
    return float(lib().ss2o_cbrt_rn(x))


def rcp_dev(x: float) -> float:
    return float(lib().ss2o_rcp_dev(x))


def xyb_f32(r: float, g: float, b: float, cbrt_kind: int = 1) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().ss2o_xyb_f32(r, g, b, cbrt_kind, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out


def xyb_f64(r: float, g: float, b: float) -> np.ndarray:
    out = np.zeros(3)
    lib().ss2o_xyb_f64(r, g, b, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def blur_f32(plane: np.ndarray, order: int = 0) -> np.ndarray:
    a = np.ascontiguousarray(plane, np.float32)
    out = np.zeros_like(a)
    h, w = a.shape
    lib().ss2o_blur_f32(a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), w, h, order)
    return out


def blur_f64(plane: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(plane, np.float64)
    out = np.zeros_like(a)
    h, w = a.shape
    lib().ss2o_blur_f64(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), w, h)
    return out


def xyb_planes(mode: Mode, width: int, height: int, image: np.ndarray):
    """Stored XYB planes under an FP32 mode: list over levels of (3, h_l, w_l)
    float32 arrays (zeros where a level is pruned)."""
    sizes = plan_scales(width, height)
    total = sum(3 * w * h for w, h in sizes)
    out = np.zeros(total, np.float32)
    a = np.ascontiguousarray(image, np.uint8)
    n = lib().ss2o_xyb_planes(ctypes.byref(mode), width, height, _u8(a),
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    assert n == len(sizes)
    res, at = [], 0
    for w, h in sizes:
        res.append(out[at:at + 3 * w * h].reshape(3, h, w))
        at += 3 * w * h
    return res


def pyramid(width: int, height: int, image: np.ndarray, f64: bool = True):
    """Linear RGB pyramid: list of (3, h_l, w_l) arrays."""
    levels = plan_scales(width, height)
    total = sum(3 * w * h for w, h in levels)
    a = np.ascontiguousarray(image, np.uint8)
    if f64:
        out = np.zeros(total)
        lib().ss2o_pyramid_f64(width, height, _u8(a), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    else:
        out = np.zeros(total, np.float32)
        lib().ss2o_pyramid_f32(width, height, _u8(a), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    res, off = [], 0
    for w, h in levels:
        res.append(out[off:off + 3 * w * h].reshape(3, h, w))
        off += 3 * w * h
    return res


def image_bytes(width: int, height: int) -> int:
    return int(lib().ss2o_image_bytes(width, height))


def round_plane(plane_format: int, v: float) -> float:
    """A stored plane value as `plane_format` keeps it (see ss2_oracle.h)."""
    return float(lib().ss2o_round_plane(plane_format, v))
