"""ctypes binding of libvqm_hip.so (include/vqm_ss2.h).

The library is built in-tree by ``python -m vqm_amd.build`` (or
``__graft_entry__.build()``). Loading fails loudly when it is missing: there is
no CPU fallback for the scoring path.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvqm_hip.so")
SYNTH_PATH = os.path.join(_HERE, "libvqm_synth.so")

VQM_OK = 0
VQM_INVALID_ARGUMENT = 1
VQM_LOGIC_ERROR = 2
VQM_RUNTIME_ERROR = 3
VQM_OVERFLOW = 4

PLANES_F32 = 0
PLANES_F16 = 1
PLANES_BF16 = 2
PLANES_U16 = 3  # unsigned 16-bit fixed point at every level (the mixed form)

u8, u32, u64, i32, f32, f64 = (ctypes.c_uint8, ctypes.c_uint32, ctypes.c_uint64,
                               ctypes.c_int32, ctypes.c_float, ctypes.c_double)
vp = ctypes.c_void_p


class ImageFormatC(ctypes.Structure):
    _fields_ = [("type", u8), ("family", u8), ("alpha", u8), ("range", u8),
                ("matrix", u8), ("transfer", u8), ("primaries", u8),
                ("chroma_location", u8), ("subsample_w", u8), ("subsample_h", u8),
                ("depth", u8), ("reserved", u8 * 5)]


class ImageDescC(ctypes.Structure):
    _fields_ = [("width", u32), ("height", u32), ("format", ImageFormatC), ("bytes", u64)]


class HostViewC(ctypes.Structure):
    _fields_ = [("width", u32), ("height", u32), ("format", ImageFormatC), ("planes", vp * 4),
                ("strides", u64 * 4)]


class DeviceImageC(ctypes.Structure):
    _fields_ = [("data", vp), ("width", u32), ("height", u32), ("format", ImageFormatC),
                ("bytes", u64)]


class ConfigC(ctypes.Structure):
    _fields_ = [("width", u32), ("height", u32), ("resident_images", u32),
                ("plane_format", u32), ("streaming_inputs", u32), ("strip_rows", u32),
                ("input_bytes", u64), ("graph", u32), ("chroma_filter", u32), ("blur", u32),
                ("map_form", u32), ("strip_cols", u32)]


class MemoryC(ctypes.Structure):
    _fields_ = [("workspace_bytes", u64), ("resident_bytes", u64), ("streaming_bytes", u64),
                ("input_pool_bytes", u64), ("pinned_host_bytes", u64),
                ("total_device_bytes", u64)]


class MemoryHeapC(ctypes.Structure):
    _fields_ = [("size_bytes", u64), ("device_local", u32), ("budget_available", u32),
                ("budget_bytes", u64), ("usage_bytes", u64)]


class BatchConfigC(ctypes.Structure):
    _fields_ = [("plane_format", u32), ("strip_rows", u32), ("max_pairs", u32),
                ("chunks", u32), ("hbm_budget", u64), ("autotune", u32), ("blur", u32),
                ("persistent", u32), ("map_form", u32), ("strip_cols", u32), ("graph", u32)]


class TuneResultC(ctypes.Structure):
    _fields_ = [("strip_rows", u32), ("chunks", u32), ("candidates", u32), ("cached", u32),
                ("ms_per_launch", f32), ("default_ms", f32), ("strip_cols", u32)]


class DevicePairC(ctypes.Structure):
    _fields_ = [("original", vp), ("distorted", vp), ("width", u32), ("height", u32)]


class ServerConfigC(ctypes.Structure):
    _fields_ = [("plane_format", u32), ("lanes", u32), ("max_batch", u32), ("strip_rows", u32),
                ("hbm_budget", u64), ("blur", u32), ("map_form", u32), ("strip_cols", u32),
                ("graph", u32)]


class RequestResultC(ctypes.Structure):
    _fields_ = [("status", i32), ("scales", u32), ("batch_size", u32), ("lane", u32),
                ("submit_ns", u64), ("launch_ns", u64), ("done_ns", u64), ("norms", f64 * 108)]


class ServerStatsC(ctypes.Structure):
    _fields_ = [("submitted", u64), ("completed", u64), ("failed", u64), ("batches", u64),
                ("batched_requests", u64), ("max_batch", u64), ("pool_bytes", u64),
                ("peak_pool_bytes", u64), ("max_lanes_busy", u32), ("queued", u32)]


class ReplayRequestC(ctypes.Structure):
    _fields_ = [("original", vp), ("distorted", vp), ("width", u32), ("height", u32),
                ("host", u32), ("reserved", u32), ("arrival_ns", u64)]


class CbrtCheckC(ctypes.Structure):
    _fields_ = [("mismatches", u64), ("exact_steps", u64), ("max_seed_error", ctypes.c_double),
                ("first_mismatch", u32), ("reserved", u32)]


class DivCheckC(ctypes.Structure):
    _fields_ = [("mismatches", u64), ("reciprocal_mismatches", u64), ("first_mismatch", u32),
                ("reserved", u32)]


# Every function include/vqm_ss2.h declares: name -> (restype, argtypes).
SIGNATURES = {
    "vqm_last_error": (ctypes.c_char_p, []),
    "vqm_ss2_config_init": (None, [ctypes.POINTER(ConfigC), u32, u32]),
    "vqm_image_format_srgb8": (None, [ctypes.POINTER(ImageFormatC)]),
    "vqm_required_bytes": (ctypes.c_int, [ctypes.POINTER(ImageFormatC), u32, u32,
                                          ctypes.POINTER(u64)]),
    "vqm_validate_image_format": (ctypes.c_int, [ctypes.POINTER(ImageFormatC), u32, u32]),
    "vqm_copy_image": (ctypes.c_int, [ctypes.POINTER(HostViewC), vp, ctypes.POINTER(u64)]),
    "vqm_plane_scale_offset": (ctypes.c_int, [ctypes.POINTER(ImageFormatC), ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(f32), ctypes.POINTER(f32)]),
    "vqm_ss2_plan_scales": (ctypes.c_int, [u32, u32, ctypes.POINTER(u32), ctypes.POINTER(u32)]),
    "vqm_ss2_gaussian_taps": (None, [ctypes.POINTER(f32)]),
    "vqm_ss2_blur_taps": (ctypes.c_int, [u32, ctypes.POINTER(f32)]),
    "vqm_ss2_channel_pruned": (ctypes.c_int, [u32, u32, u32]),
    "vqm_ss2_score": (ctypes.c_int, [u32, ctypes.POINTER(f64), ctypes.POINTER(f64)]),
    "vqm_ss2_srgb_table": (None, [ctypes.POINTER(f32)]),
    "vqm_ss2_memory_plan": (ctypes.c_int, [ctypes.POINTER(ConfigC), ctypes.POINTER(MemoryC)]),
    "vqm_ss2_algorithmic_bytes": (ctypes.c_int, [u32, u32, u32, ctypes.POINTER(u64)]),
    "vqm_ss2_create": (ctypes.c_int, [ctypes.POINTER(ConfigC), ctypes.POINTER(vp)]),
    "vqm_ss2_destroy": (None, [vp]),
    "vqm_ss2_upload": (ctypes.c_int, [vp, ctypes.POINTER(ImageDescC), vp, ctypes.POINTER(u64)]),
    "vqm_ss2_release": (ctypes.c_int, [vp, u64]),
    "vqm_ss2_submit_host": (ctypes.c_int, [vp, ctypes.POINTER(ImageDescC), vp,
                                           ctypes.POINTER(ImageDescC), vp, f32,
                                           ctypes.POINTER(u64)]),
    "vqm_ss2_submit_views": (ctypes.c_int, [vp, ctypes.POINTER(HostViewC), ctypes.POINTER(HostViewC), f32,
                                            ctypes.POINTER(u64)]),
    "vqm_ss2_submit_resident": (ctypes.c_int, [vp, u64, u64, f32, ctypes.POINTER(u64)]),
    "vqm_ss2_submit_resident_host": (ctypes.c_int, [vp, u64, ctypes.POINTER(ImageDescC), vp, f32,
                                                    ctypes.POINTER(u64)]),
    "vqm_ss2_submit_device": (ctypes.c_int, [vp, ctypes.POINTER(DeviceImageC),
                                             ctypes.POINTER(DeviceImageC), vp, f32,
                                             ctypes.POINTER(u64)]),
    "vqm_ss2_finish": (ctypes.c_int, [vp, u64, ctypes.POINTER(u32), ctypes.POINTER(f64)]),
    "vqm_ss2_memory_usage": (ctypes.c_int, [vp, ctypes.POINTER(MemoryC)]),
    "vqm_ss2_allocate_image": (ctypes.c_int, [vp, ctypes.POINTER(ImageFormatC),
                                              ctypes.POINTER(u64), ctypes.POINTER(u64)]),
    "vqm_ss2_map_image": (ctypes.c_int, [vp, u64, ctypes.POINTER(vp)]),
    "vqm_ss2_release_image": (ctypes.c_int, [vp, u64]),
    "vqm_ss2_submit_buffers": (ctypes.c_int, [vp, u64, u64, f32, ctypes.POINTER(u64)]),
    "vqm_ss2_stage_times": (ctypes.c_int, [vp, ctypes.POINTER(f32)]),
    "vqm_ss2_prepared_images": (ctypes.c_int, [vp, ctypes.POINTER(u32)]),
    "vqm_ss2_batch_create": (ctypes.c_int, [ctypes.POINTER(BatchConfigC), ctypes.POINTER(vp)]),
    "vqm_ss2_batch_destroy": (None, [vp]),
    "vqm_ss2_batch_submit": (ctypes.c_int, [vp, ctypes.POINTER(DevicePairC), u32,
                                            ctypes.POINTER(u64)]),
    "vqm_ss2_batch_finish": (ctypes.c_int, [vp, u64, ctypes.POINTER(u32), ctypes.POINTER(f64)]),
    "vqm_ss2_batch_stage_times": (ctypes.c_int, [vp, ctypes.POINTER(f32)]),
    "vqm_ss2_batch_pool_bytes": (ctypes.c_int, [vp, ctypes.POINTER(u64)]),
    "vqm_memory_heaps": (ctypes.c_int, [ctypes.POINTER(MemoryHeapC)]),
    "vqm_ss2_batch_set_shape": (ctypes.c_int, [vp, u32, u32, u32, u32]),
    "vqm_ss2_autotune": (ctypes.c_int, [u32, u32, u32, u32, u32, ctypes.POINTER(TuneResultC)]),
    "vqm_ss2_score_device_noalloc": (ctypes.c_int, [ctypes.POINTER(ConfigC), ctypes.POINTER(DeviceImageC),
                                                    ctypes.POINTER(DeviceImageC), vp, vp,
                                                    ctypes.POINTER(u32)]),
    "vqm_ss2_server_config_init": (None, [ctypes.POINTER(ServerConfigC)]),
    "vqm_ss2_server_create": (ctypes.c_int, [ctypes.POINTER(ServerConfigC), ctypes.POINTER(vp)]),
    "vqm_ss2_server_destroy": (None, [vp]),
    "vqm_ss2_server_submit_device": (ctypes.c_int, [vp, ctypes.POINTER(DevicePairC),
                                                    ctypes.POINTER(u64)]),
    "vqm_ss2_server_submit_host": (ctypes.c_int, [vp, ctypes.POINTER(ImageDescC), vp,
                                                  ctypes.POINTER(ImageDescC), vp,
                                                  ctypes.POINTER(u64)]),
    "vqm_ss2_server_wait": (ctypes.c_int, [vp, u64, ctypes.POINTER(RequestResultC)]),
    "vqm_ss2_server_get_stats": (ctypes.c_int, [vp, ctypes.POINTER(ServerStatsC)]),
    "vqm_now_ns": (u64, []),
    "vqm_ss2_server_replay": (ctypes.c_int, [vp, ctypes.POINTER(ReplayRequestC), u32,
                                             ctypes.POINTER(RequestResultC)]),
    "vqm_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "vqm_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p,
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(u64)]),
    "vqm_device_malloc": (ctypes.c_int, [u64, ctypes.POINTER(vp)]),
    "vqm_device_free": (ctypes.c_int, [vp]),
    "vqm_host_malloc": (ctypes.c_int, [u64, ctypes.POINTER(vp)]),
    "vqm_host_free": (ctypes.c_int, [vp]),
    "vqm_memcpy_htod": (ctypes.c_int, [vp, vp, u64]),
    "vqm_memcpy_dtoh": (ctypes.c_int, [vp, vp, u64]),
    "vqm_device_synchronize": (ctypes.c_int, []),
    "vqm_device_copy": (ctypes.c_int, [vp, vp, u64]),
    "vqm_ss2_check_cbrt": (ctypes.c_int, [u32, u32, ctypes.POINTER(CbrtCheckC)]),
    "vqm_ss2_check_div": (ctypes.c_int, [u32, u32, ctypes.POINTER(DivCheckC)]),
}

_lock = threading.Lock()
_lib = None
_synth = None


def lib() -> ctypes.CDLL:
    """The loaded libvqm_hip.so with typed signatures; raises if not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with `python -m vqm_amd.build` "
                    "(the scoring path has no CPU fallback)")
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
        return _lib


def synth() -> ctypes.CDLL:
    """The loaded libvqm_synth.so (synthetic pair generator)."""
    global _synth
    with _lock:
        if _synth is None:
            if not os.path.exists(SYNTH_PATH):
                raise RuntimeError(f"{SYNTH_PATH} is missing: build it with `python -m vqm_amd.build`")
            h = ctypes.CDLL(SYNTH_PATH)
            h.vqm_synth_image_bytes.restype = u64
            h.vqm_synth_image_bytes.argtypes = [u32, u32]
            h.vqm_synth_reference.restype = None
            h.vqm_synth_reference.argtypes = [u32, u32, u32, u32, f64, vp, u32]
            h.vqm_synth_distort.restype = None
            h.vqm_synth_distort.argtypes = [u32, u32, u32, u32, vp, i32, f64, vp, u32]
            h.vqm_synth_draw.restype = f64
            h.vqm_synth_draw.argtypes = [u32, u32, u32]
            h.vqm_synth_hash.restype = u64
            h.vqm_synth_hash.argtypes = [vp, u64]
            _synth = h
        return _synth
