"""ctypes mirror of include/meteoros_cloud.h.

The structures are byte-identical to the C header (which in turn keeps the
std140 layouts of the reference's CameraUBO, TimeUBO and SunAndSkyUBO,
camera.h:12-18, Scene.h:12-21, Sky.h:9-15); `tests/test_abi.py` checks every
size and offset.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmeteoros_cloud.so")

MTC_OK = 0
MTC_ERR_INVALID_ARG = 1
MTC_ERR_OUT_OF_MEMORY = 2
MTC_ERR_HIP = 3
MTC_ERR_UNSUPPORTED = 4
MTC_ERR_STATE = 5
MTC_GOD_RAYS_NOALLOC_MAX_ENTRIES = 2359296  # (width + 1) * (height + 1) limit of mtc_god_rays_noalloc
MTC_GOD_RAYS_NOALLOC_MAX_HEIGHT = 4320  # height limit of mtc_god_rays_noalloc
MTC_GOD_RAYS_NOALLOC_MAX_WIDTH = 16384  # width limit of mtc_god_rays_noalloc
MTC_GOD_RAYS_NOALLOC_SCRATCH_BYTES = 18874368  # its static scratch, bytes per device

MTC_STEPS_REFERENCE = 0
MTC_STEPS_FIXED = 1
MTC_COVERAGE_CONSTANT = 0
MTC_COVERAGE_WEATHER = 1
MTC_FMT_UNORM8 = 0
MTC_FMT_FP16 = 1
MTC_KERNEL_PLAIN = 0
MTC_KERNEL_FAST = 1
MTC_KERNEL_PRECISION = 2
MTC_KERNEL_LEARNED = 3
MTC_FLAG_COUNTERS = 1
MTC_FLAG_FULL_MARCH = 2
MTC_FLAG_TIMING = 4
MTC_FLAG_LDS_BRICKS = 8
MTC_DISPLAY_RGBA8 = 0
MTC_DISPLAY_RGBA32F = 1


class MtcCamera(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("proj", C.c_float * 16), ("eye", C.c_float * 4),
                ("tan_fov_by2", C.c_float * 2)]


class MtcTime(C.Structure):
    _fields_ = [("halton1", C.c_float * 4), ("halton2", C.c_float * 4), ("halton3", C.c_float * 4),
                ("halton4", C.c_float * 4), ("time", C.c_float * 2), ("frame_count_mod16", C.c_int32)]


class MtcSun(C.Structure):
    _fields_ = [("sun_location", C.c_float * 4), ("sun_direction", C.c_float * 4),
                ("light_color", C.c_float * 4), ("sun_intensity", C.c_float)]


class MtcCloudParams(C.Structure):
    _fields_ = [
        ("earth_radius", C.c_double), ("shell_inner_radius", C.c_double), ("shell_outer_radius", C.c_double),
        ("sky_sun_location", C.c_float * 4),
        ("step_mode", C.c_int32), ("steps_min", C.c_int32), ("steps_max", C.c_int32),
        ("light_samples", C.c_int32),
        ("base_density_factor", C.c_float), ("jitter_factor", C.c_float), ("jitter_dim", C.c_float),
        ("sample_divisor", C.c_float), ("erosion_scale", C.c_float), ("accum_scale", C.c_float),
        ("cone_erosion_scale", C.c_float), ("brightness", C.c_float),
        ("hg_eccentricity", C.c_float), ("hg_silver_intensity", C.c_float), ("hg_silver_spread", C.c_float),
        ("wind_direction", C.c_float * 4), ("wind_upward_bias", C.c_float), ("cloud_speed", C.c_float),
        ("cloud_top_offset", C.c_float), ("wind_skew", C.c_float),
        ("coverage_mode", C.c_int32), ("coverage_constant", C.c_float), ("weather_scale", C.c_float),
        ("background_min", C.c_float), ("cloud_fade_start", C.c_float), ("cloud_fade_end", C.c_float),
        ("motion_blur_samples", C.c_int32), ("reserved", C.c_int32 * 3),
    ]


class MtcCameraState(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("ref", C.c_float * 3), ("forward", C.c_float * 3),
                ("right", C.c_float * 3), ("up", C.c_float * 3), ("world_up", C.c_float * 3),
                ("width", C.c_int32), ("height", C.c_int32), ("fovy_deg", C.c_float), ("aspect", C.c_float),
                ("near_clip", C.c_float), ("far_clip", C.c_float)]


class MtcVolumeDesc(C.Structure):
    _fields_ = [("lf_dim", C.c_int32), ("hf_dim", C.c_int32), ("curl_dim", C.c_int32),
                ("weather_dim", C.c_int32), ("format", C.c_int32)]


class MtcNoiseDesc(C.Structure):
    _fields_ = [("seed_volumes", C.c_uint32), ("seed_weather", C.c_uint32)]


class MtcDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("volumes", MtcVolumeDesc),
                ("kernel", C.c_int32), ("flags", C.c_int32), ("device", C.c_int32),
                ("lanes_per_ray", C.c_int32)]


class MtcCounters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "rays_marched", "rays_ocean", "rays_sky", "primary_steps", "cloud_steps", "cone_samples",
        "cone_erosions", "early_exits", "mlp_queries", "mlp_fallbacks", "mlp_audits", "mlp_audit_sq_err",
        "lf_brick_stages", "lf_brick_samples")]


class MtcTimings(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("reproject_ms", C.c_double), ("march_ms", C.c_double)]


class MtcBatchTuning(C.Structure):
    _fields_ = [("lanes_per_ray", C.c_int32), ("tile_rays_x", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("order", C.c_int32)]


MTC_MLP_INPUTS = 9
MTC_MLP_PARAMS = 9025


class MtcMlpTrainDesc(C.Structure):
    _fields_ = [("steps", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint32), ("holdout", C.c_int64)]


class MtcMlpCalibration(C.Structure):
    _fields_ = [("lo", C.c_float), ("hi", C.c_float), ("rms", C.c_float * 16), ("count", C.c_uint32 * 16),
                ("mask", C.c_uint32), ("max_rms", C.c_float), ("fallback_density", C.c_float),
                ("rms_all", C.c_float), ("rms_mean_predictor", C.c_float)]


P = C.POINTER
# (name, restype, argtypes) of every exported entry point of meteoros_cloud.h
SIGNATURES = [
    ("mtc_last_error", C.c_char_p, []),
    ("mtc_version", C.c_char_p, []),
    ("mtc_device_count", C.c_int, [P(C.c_int32)]),
    ("mtc_camera_init", C.c_int, [P(MtcCameraState), P(C.c_float), P(C.c_float), C.c_int32, C.c_int32, C.c_float,
                                  C.c_float, C.c_float]),
    ("mtc_camera_rotate_about_up", C.c_int, [P(MtcCameraState), C.c_float]),
    ("mtc_camera_rotate_about_right", C.c_int, [P(MtcCameraState), C.c_float]),
    ("mtc_camera_translate", C.c_int, [P(MtcCameraState), C.c_float, C.c_float, C.c_float]),
    ("mtc_camera_ubo", C.c_int, [P(MtcCameraState), P(MtcCamera)]),
    ("mtc_halton", C.c_float, [C.c_int32, C.c_int32]),
    ("mtc_time_init", C.c_int, [P(MtcTime)]),
    ("mtc_time_advance", C.c_int, [P(MtcTime), C.c_float]),
    ("mtc_params_reference", C.c_int, [P(MtcCloudParams), P(MtcSun)]),
    ("mtc_params_preset", C.c_int, [C.c_int32, P(MtcCloudParams), P(MtcSun)]),
    ("mtc_sun_from_angles", C.c_int, [C.c_float, C.c_float, C.c_float, P(MtcSun), P(MtcCloudParams)]),
    ("mtc_wind_speed_units", C.c_float, [P(MtcCloudParams), C.c_float]),
    ("mtc_create", C.c_int, [P(MtcDesc), P(C.c_void_p)]),
    ("mtc_destroy", C.c_int, [C.c_void_p]),
    ("mtc_set_volumes", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mtc_generate_volumes", C.c_int, [C.c_void_p, P(MtcNoiseDesc)]),
    ("mtc_get_volumes", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mtc_frame", C.c_int, [C.c_void_p, P(MtcCamera), P(MtcCamera), P(MtcTime), P(MtcSun), P(MtcCloudParams),
                            C.c_void_p]),
    ("mtc_march_full", C.c_int, [C.c_void_p, P(MtcCamera), P(MtcTime), P(MtcSun), P(MtcCloudParams), C.c_void_p]),
    ("mtc_synchronize", C.c_int, [C.c_void_p]),
    ("mtc_read_output", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("mtc_output_device_ptrs", C.c_int, [C.c_void_p, P(C.c_void_p), P(C.c_void_p)]),
    ("mtc_set_history", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("mtc_reset_history", C.c_int, [C.c_void_p]),
    ("mtc_get_counters", C.c_int, [C.c_void_p, P(MtcCounters)]),
    ("mtc_reset_counters", C.c_int, [C.c_void_p]),
    ("mtc_get_timings", C.c_int, [C.c_void_p, P(MtcTimings)]),
    ("mtc_reset_timings", C.c_int, [C.c_void_p]),
    ("mtc_set_timing_stride", C.c_int, [C.c_void_p, C.c_int32]),
    ("mtc_reprojection_indices", C.c_int, [C.c_void_p, P(MtcCamera), P(MtcCamera), P(MtcTime), P(MtcCloudParams),
                                           C.c_void_p, C.c_size_t]),
    ("mtc_load_image_rgba8", C.c_int, [C.c_char_p, P(C.c_int32), P(C.c_int32), C.c_void_p, C.c_size_t]),
    ("mtc_load_tga_slices", C.c_int, [C.c_char_p, C.c_char_p, C.c_int32, C.c_void_p, C.c_size_t]),
    ("mtc_god_rays", C.c_int, [C.c_void_p, P(MtcCamera), P(MtcSun), C.c_void_p]),
    ("mtc_god_rays_noalloc", C.c_int, [C.c_int32, C.c_int32, P(MtcCamera), P(MtcSun), C.c_void_p, C.c_void_p,
                                       C.c_void_p]),
    ("mtc_tone_map", C.c_int, [C.c_void_p, P(MtcTime), C.c_int32, C.c_void_p]),
    ("mtc_read_display", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("mtc_txaa", C.c_int, [C.c_void_p, P(MtcCamera), P(MtcCamera), P(MtcTime), C.c_void_p]),
    ("mtc_read_txaa", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("mtc_set_txaa_history", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("mtc_batch_create", C.c_int, [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_void_p)]),
    ("mtc_batch_destroy", C.c_int, [C.c_void_p]),
    ("mtc_batch_set_tuning", C.c_int, [C.c_void_p, P(MtcBatchTuning)]),
    ("mtc_batch_get_tuning", C.c_int, [C.c_void_p, P(MtcBatchTuning)]),
    ("mtc_batch_render", C.c_int, [C.c_void_p, P(MtcCamera), P(MtcTime), P(MtcSun), P(MtcCloudParams), C.c_void_p]),
    ("mtc_batch_autotune", C.c_int, [C.c_void_p, P(MtcCamera), P(MtcTime), P(MtcSun), P(MtcCloudParams),
                                     P(MtcBatchTuning), P(C.c_double)]),
    ("mtc_batch_last_ms", C.c_int, [C.c_void_p, P(C.c_double)]),
    ("mtc_batch_read", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("mtc_mlp_create", C.c_int, [C.c_int32, C.c_int64, P(C.c_void_p)]),
    ("mtc_mlp_destroy", C.c_int, [C.c_void_p]),
    ("mtc_mlp_init", C.c_int, [C.c_void_p, C.c_uint32]),
    ("mtc_mlp_set_params", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("mtc_mlp_get_params", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("mtc_mlp_collect", C.c_int, [C.c_void_p, C.c_void_p, P(MtcCamera), P(MtcTime), P(MtcSun), P(MtcCloudParams),
                                  C.c_uint32, C.c_float, P(C.c_int64)]),
    ("mtc_mlp_sample_count", C.c_int, [C.c_void_p, P(C.c_int64)]),
    ("mtc_mlp_clear_samples", C.c_int, [C.c_void_p]),
    ("mtc_mlp_set_samples", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    ("mtc_mlp_read_samples", C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mtc_mlp_train", C.c_int, [C.c_void_p, P(MtcMlpTrainDesc), C.c_void_p]),
    ("mtc_mlp_gradient", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, P(C.c_float)]),
    ("mtc_mlp_predict", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    ("mtc_mlp_calibrate", C.c_int, [C.c_void_p, C.c_int64, C.c_float, C.c_float, P(MtcMlpCalibration)]),
    ("mtc_set_learned_light", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
]

_lib = None


class MtcError(RuntimeError):
    """A non-zero mtc_status from the C ABI, with the library's message."""

    def __init__(self, code: int, fn: str, msg: str):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code
        self.fn = fn


def lib() -> C.CDLL:
    """Load libmeteoros_cloud.so (built in-tree by __graft_entry__.build()).

    Raises loudly when the native library is missing: there is no fallback
    implementation of the stage in Python.
    """
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        handle = C.CDLL(LIB_PATH)
        missing = []
        for name, res, args in SIGNATURES:
            try:
                fn = getattr(handle, name)
            except AttributeError:
                # a library older than this binding: the entry point raises
                # AttributeError when called; the rest stay usable
                missing.append(name)
                continue
            fn.restype = res
            fn.argtypes = args
        handle.mtc_missing_symbols = missing
        _lib = handle
    return _lib


def check(code: int, fn: str) -> None:
    if code != MTC_OK:
        raise MtcError(code, fn, lib().mtc_last_error().decode())
