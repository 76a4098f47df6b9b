"""Host-side mirror of the reference's cloud-pass interface.

The reference drives its compute stage from C++ objects (camera.{h,cpp},
Scene.{h,cpp}, Sky.{h,cpp}, Renderer.{h,cpp}); the same roles exist here as
thin Python classes over the C ABI of libmeteoros_cloud.so:

  Camera      camera.h:20-68      lookAt/perspective uniforms, RotateAboutUp ...
  Time        Scene.h:12-21       Halton table, (dt, total), frameCount mod 16
  CloudStage  Renderer.cpp:653-722 + Sky.cpp:25-58
              volumes, frame() = reprojection then march, ping-pong history

Every call goes to the native library; nothing here computes the stage.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi as A
from ._abi import check, lib


class Camera:
    """camera.h:20-68: eye/ref/up basis, UpdateBuffer() -> CameraUBO."""

    def __init__(self, eye, ref, width: int, height: int, fovy_deg: float = 45.0,
                 near_clip: float = 0.1, far_clip: float = 1000.0):
        self.state = A.MtcCameraState()
        e = (C.c_float * 3)(*eye)
        r = (C.c_float * 3)(*ref)
        check(lib().mtc_camera_init(C.byref(self.state), e, r, int(width), int(height), float(fovy_deg),
                                    float(near_clip), float(far_clip)), "mtc_camera_init")

    def rotate_about_up(self, deg: float) -> None:
        check(lib().mtc_camera_rotate_about_up(C.byref(self.state), float(deg)), "mtc_camera_rotate_about_up")

    def rotate_about_right(self, deg: float) -> None:
        check(lib().mtc_camera_rotate_about_right(C.byref(self.state), float(deg)), "mtc_camera_rotate_about_right")

    def translate(self, look: float = 0.0, right: float = 0.0, up: float = 0.0) -> None:
        check(lib().mtc_camera_translate(C.byref(self.state), float(look), float(right), float(up)),
              "mtc_camera_translate")

    def ubo(self) -> A.MtcCamera:
        out = A.MtcCamera()
        check(lib().mtc_camera_ubo(C.byref(self.state), C.byref(out)), "mtc_camera_ubo")
        return out


class Time:
    """Scene.h:12-21 / Scene.cpp:65-119 with the wall clock replaced by dt."""

    def __init__(self):
        self.ubo = A.MtcTime()
        check(lib().mtc_time_init(C.byref(self.ubo)), "mtc_time_init")

    def advance(self, dt: float) -> None:
        check(lib().mtc_time_advance(C.byref(self.ubo), float(dt)), "mtc_time_advance")

    @property
    def frame(self) -> int:
        return int(self.ubo.frame_count_mod16)


def halton(index: int, base: int) -> float:
    """Scene::HaltonSequenceAt (Scene.cpp:125-138)."""
    return float(lib().mtc_halton(int(index), int(base)))


def reference_params():
    """The shipped #define values of cloudRayMarch.comp / reprojection.comp."""
    p, s = A.MtcCloudParams(), A.MtcSun()
    check(lib().mtc_params_reference(C.byref(p), C.byref(s)), "mtc_params_reference")
    return p, s


def preset_params(preset: int):
    """Workload presets 1..5 (shell R+1500..R+4000 m, weather coverage,
    sun at 30 deg, 10 m/s wind; 64/128/192/64/128 primary steps)."""
    p, s = A.MtcCloudParams(), A.MtcSun()
    check(lib().mtc_params_preset(int(preset), C.byref(p), C.byref(s)), "mtc_params_preset")
    return p, s


def sun_from_angles(elevation_deg: float, azimuth_deg: float, sun: A.MtcSun, params: A.MtcCloudParams | None = None,
                    distance_m: float = 1.0e8) -> None:
    """Place the lighting sun (and the background sun of `params`) at an
    elevation / azimuth (0 = -z) and distance."""
    check(lib().mtc_sun_from_angles(float(elevation_deg), float(azimuth_deg), float(distance_m), C.byref(sun),
                                    C.byref(params) if params is not None else None), "mtc_sun_from_angles")


@dataclass
class VolumeSpec:
    lf_dim: int = 128
    hf_dim: int = 32
    curl_dim: int = 128
    weather_dim: int = 512
    fmt: int = A.MTC_FMT_UNORM8

    def desc(self) -> A.MtcVolumeDesc:
        return A.MtcVolumeDesc(self.lf_dim, self.hf_dim, self.curl_dim, self.weather_dim, self.fmt)


def device_count() -> int:
    n = C.c_int32(0)
    check(lib().mtc_device_count(C.byref(n)), "mtc_device_count")
    return int(n.value)


class CloudStage:
    """The compute half of Renderer: resources + Frame() for one context."""

    def __init__(self, width: int, height: int, volumes: VolumeSpec | None = None,
                 kernel: int = A.MTC_KERNEL_PLAIN, flags: int = 0, device: int = 0, lanes_per_ray: int = 0):
        self.width, self.height = int(width), int(height)
        self.volumes = volumes or VolumeSpec()
        d = A.MtcDesc()
        d.width, d.height = self.width, self.height
        d.volumes = self.volumes.desc()
        d.kernel, d.flags, d.device = int(kernel), int(flags), int(device)
        d.lanes_per_ray = int(lanes_per_ray)
        self._ctx = C.c_void_p()
        check(lib().mtc_create(C.byref(d), C.byref(self._ctx)), "mtc_create")

    def close(self) -> None:
        if self._ctx:
            lib().mtc_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._ctx

    # -- resources (Sky::CreateCloudResources) --------------------------------
    def set_volumes(self, lf: np.ndarray, hf: np.ndarray, curl: np.ndarray, weather: np.ndarray) -> None:
        arrs = [np.ascontiguousarray(a) for a in (lf, hf, curl, weather)]
        check(lib().mtc_set_volumes(self._ctx, *[a.ctypes.data for a in arrs]), "mtc_set_volumes")

    def generate_volumes(self, seed_volumes: int = 1, seed_weather: int = 2) -> None:
        nd = A.MtcNoiseDesc(seed_volumes, seed_weather)
        check(lib().mtc_generate_volumes(self._ctx, C.byref(nd)), "mtc_generate_volumes")

    def get_volumes(self):
        v = self.volumes
        tb = 8 if v.fmt == A.MTC_FMT_FP16 else 4
        dt = np.uint16 if v.fmt == A.MTC_FMT_FP16 else np.uint8
        lf = np.empty(v.lf_dim ** 3 * tb // dt().itemsize, dtype=dt)
        hf = np.empty(v.hf_dim ** 3 * tb // dt().itemsize, dtype=dt)
        curl = np.empty(v.curl_dim ** 2 * 4, dtype=np.uint8)
        weather = np.empty(v.weather_dim ** 2 * 4, dtype=np.uint8)
        check(lib().mtc_get_volumes(self._ctx, lf.ctypes.data, hf.ctypes.data, curl.ctypes.data,
                                    weather.ctypes.data), "mtc_get_volumes")
        return lf, hf, curl, weather

    # -- frames (Renderer::Frame) ---------------------------------------------
    def frame(self, cam: A.MtcCamera, old: A.MtcCamera, time: A.MtcTime, sun: A.MtcSun,
              params: A.MtcCloudParams, stream: int | None = None) -> None:
        check(lib().mtc_frame(self._ctx, C.byref(cam), C.byref(old), C.byref(time), C.byref(sun),
                              C.byref(params), stream), "mtc_frame")

    def march_full(self, cam: A.MtcCamera, time: A.MtcTime, sun: A.MtcSun, params: A.MtcCloudParams,
                   stream: int | None = None) -> None:
        check(lib().mtc_march_full(self._ctx, C.byref(cam), C.byref(time), C.byref(sun), C.byref(params), stream),
              "mtc_march_full")

    def synchronize(self) -> None:
        check(lib().mtc_synchronize(self._ctx), "mtc_synchronize")

    def read_output(self):
        n = self.width * self.height * 4
        rgba = np.empty(n, dtype=np.float32)
        mask = np.empty(n, dtype=np.float32)
        check(lib().mtc_read_output(self._ctx, rgba.ctypes.data, mask.ctypes.data, n), "mtc_read_output")
        return rgba.reshape(self.height, self.width, 4), mask.reshape(self.height, self.width, 4)

    def set_history(self, rgba: np.ndarray) -> None:
        a = np.ascontiguousarray(rgba, dtype=np.float32).reshape(-1)
        check(lib().mtc_set_history(self._ctx, a.ctypes.data, a.size), "mtc_set_history")

    def reset_history(self) -> None:
        check(lib().mtc_reset_history(self._ctx), "mtc_reset_history")

    def counters(self) -> dict:
        c = A.MtcCounters()
        check(lib().mtc_get_counters(self._ctx, C.byref(c)), "mtc_get_counters")
        return {name: int(getattr(c, name)) for name, _ in A.MtcCounters._fields_}

    def reset_counters(self) -> None:
        check(lib().mtc_reset_counters(self._ctx), "mtc_reset_counters")

    def timings(self) -> dict:
        t = A.MtcTimings()
        check(lib().mtc_get_timings(self._ctx, C.byref(t)), "mtc_get_timings")
        return {"frames": int(t.frames), "reproject_ms": float(t.reproject_ms), "march_ms": float(t.march_ms)}

    def reset_timings(self) -> None:
        check(lib().mtc_reset_timings(self._ctx), "mtc_reset_timings")

    def set_timing_stride(self, stride: int) -> None:
        """With MTC_FLAG_TIMING: time one frame in `stride`."""
        check(lib().mtc_set_timing_stride(self._ctx, int(stride)), "mtc_set_timing_stride")

    # -- post-process passes (postProcess_GodRays.frag, postProcess_ToneMap.frag)
    def god_rays(self, cam: A.MtcCamera, sun: A.MtcSun, stream: int | None = None) -> None:
        check(lib().mtc_god_rays(self._ctx, C.byref(cam), C.byref(sun), stream), "mtc_god_rays")

    def tone_map(self, time: A.MtcTime, as_float: bool = False, stream: int | None = None) -> None:
        fmt = A.MTC_DISPLAY_RGBA32F if as_float else A.MTC_DISPLAY_RGBA8
        check(lib().mtc_tone_map(self._ctx, C.byref(time), fmt, stream), "mtc_tone_map")
        self._display_float = as_float

    def read_display(self) -> np.ndarray:
        as_float = getattr(self, "_display_float", False)
        out = np.empty((self.height, self.width, 4), dtype=np.float32 if as_float else np.uint8)
        check(lib().mtc_read_display(self._ctx, out.ctypes.data, out.nbytes), "mtc_read_display")
        return out

    def txaa(self, cam: A.MtcCamera, old: A.MtcCamera, time: A.MtcTime, stream: int | None = None) -> None:
        check(lib().mtc_txaa(self._ctx, C.byref(cam), C.byref(old), C.byref(time), stream), "mtc_txaa")

    def read_txaa(self) -> np.ndarray:
        out = np.empty((self.height, self.width, 4), dtype=np.uint8)
        check(lib().mtc_read_txaa(self._ctx, out.ctypes.data, out.nbytes), "mtc_read_txaa")
        return out

    def set_txaa_history(self, rgba8: np.ndarray) -> None:
        a = np.ascontiguousarray(rgba8, dtype=np.uint8)
        check(lib().mtc_set_txaa_history(self._ctx, a.ctypes.data, a.nbytes), "mtc_set_txaa_history")

    def reprojection_indices(self, cam, old, time, params) -> np.ndarray:
        n = self.width * self.height * int(params.motion_blur_samples)
        idx = np.empty(n, dtype=np.int32)
        check(lib().mtc_reprojection_indices(self._ctx, C.byref(cam), C.byref(old), C.byref(time), C.byref(params),
                                             idx.ctypes.data, n), "mtc_reprojection_indices")
        return idx.reshape(self.height, self.width, int(params.motion_blur_samples))


class Batch:
    """Many views of uneven size rendered by one persistent launch (every
    pixel marched); the launch is a hipGraph replayed per render."""

    def __init__(self, stage: CloudStage, sizes):
        self.stage = stage
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        n = len(self.sizes)
        ws = (C.c_int32 * n)(*[w for w, _ in self.sizes])
        hs = (C.c_int32 * n)(*[h for _, h in self.sizes])
        self._b = C.c_void_p()
        check(lib().mtc_batch_create(stage.handle, n, ws, hs, C.byref(self._b)), "mtc_batch_create")

    def close(self) -> None:
        if self._b:
            lib().mtc_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _arrays(cams, times, suns, params):
        n = len(cams)
        return ((A.MtcCamera * n)(*cams), (A.MtcTime * n)(*times), (A.MtcSun * n)(*suns),
                (A.MtcCloudParams * n)(*params))

    def render(self, cams, times, suns, params, stream=None) -> None:
        check(lib().mtc_batch_render(self._b, *self._arrays(cams, times, suns, params), stream), "mtc_batch_render")

    def autotune(self, cams, times, suns, params):
        best = A.MtcBatchTuning()
        ms = C.c_double()
        check(lib().mtc_batch_autotune(self._b, *self._arrays(cams, times, suns, params), C.byref(best), C.byref(ms)),
              "mtc_batch_autotune")
        return {n: int(getattr(best, n)) for n, _ in A.MtcBatchTuning._fields_}, float(ms.value)

    def set_tuning(self, lanes_per_ray=8, tile_rays_x=8, blocks_per_cu=4, order=0) -> None:
        t = A.MtcBatchTuning(lanes_per_ray, tile_rays_x, blocks_per_cu, order)
        check(lib().mtc_batch_set_tuning(self._b, C.byref(t)), "mtc_batch_set_tuning")

    def tuning(self) -> dict:
        t = A.MtcBatchTuning()
        check(lib().mtc_batch_get_tuning(self._b, C.byref(t)), "mtc_batch_get_tuning")
        return {n: int(getattr(t, n)) for n, _ in A.MtcBatchTuning._fields_}

    def last_ms(self) -> float:
        ms = C.c_double()
        check(lib().mtc_batch_last_ms(self._b, C.byref(ms)), "mtc_batch_last_ms")
        return float(ms.value)

    def read(self, view: int):
        w, h = self.sizes[view]
        rgba = np.empty(w * h * 4, dtype=np.float32)
        mask = np.empty(w * h * 4, dtype=np.float32)
        check(lib().mtc_batch_read(self._b, int(view), rgba.ctypes.data, mask.ctypes.data, rgba.size),
              "mtc_batch_read")
        return rgba.reshape(h, w, 4), mask.reshape(h, w, 4)


def load_image_rgba8(path: str) -> np.ndarray:
    """A TGA or PNG image as RGBA8 [h, w, 4], top row first (stbi_load(..., 4))."""
    w, h = C.c_int32(), C.c_int32()
    check(lib().mtc_load_image_rgba8(path.encode(), C.byref(w), C.byref(h), None, 0), "mtc_load_image_rgba8")
    out = np.empty((h.value, w.value, 4), dtype=np.uint8)
    check(lib().mtc_load_image_rgba8(path.encode(), None, None, out.ctypes.data, out.nbytes), "mtc_load_image_rgba8")
    return out


def load_tga_slices(folder: str, base: str, n: int) -> np.ndarray:
    """An n^3 RGBA8 volume from "<folder>/<base>(z+1).tga" slices, flat, x fastest
    (ImageLoadingUtility.cpp:75-139)."""
    out = np.empty(n * n * n * 4, dtype=np.uint8)
    check(lib().mtc_load_tga_slices(folder.encode(), base.encode(), int(n), out.ctypes.data, out.nbytes),
          "mtc_load_tga_slices")
    return out


def load_shipped_volumes(texture_dir: str):
    """The four resources of Sky::CreateCloudResources (Sky.cpp:25-58) from
    Meteoros' CloudTextures folder: (lf 128^3, hf 32^3, curl 128^2, weather
    512^2), RGBA8, ready for CloudStage.set_volumes with
    VolumeSpec(128, 32, 128, 512)."""
    lf = load_tga_slices(f"{texture_dir}/LowFrequency", "LowFrequency", 128)
    hf = load_tga_slices(f"{texture_dir}/HighFrequency", "HighFrequency", 32)
    curl = load_image_rgba8(f"{texture_dir}/curlNoise.png").reshape(-1)
    weather = load_image_rgba8(f"{texture_dir}/weatherMap.png").reshape(-1)
    return lf, hf, curl, weather


class LightNet:
    """The learned-light network (9 -> 64 -> 64 -> 64 -> 1, SiLU) with its
    device store of training samples.  No reference counterpart: it stands
    in for the cone march's optical depth (cloudRayMarch.comp:338-373).
    Training, inference and sample generation run on the GPU."""

    N_PARAMS = A.MTC_MLP_PARAMS
    N_INPUTS = A.MTC_MLP_INPUTS

    def __init__(self, capacity: int = 1 << 20, device: int = 0):
        self.capacity = int(capacity)
        self._m = C.c_void_p()
        check(lib().mtc_mlp_create(int(device), self.capacity, C.byref(self._m)), "mtc_mlp_create")

    def close(self) -> None:
        if self._m:
            lib().mtc_mlp_destroy(self._m)
            self._m = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._m

    def init(self, seed: int = 1) -> None:
        check(lib().mtc_mlp_init(self._m, int(seed)), "mtc_mlp_init")

    @property
    def params(self) -> np.ndarray:
        p = np.empty(self.N_PARAMS, dtype=np.float32)
        check(lib().mtc_mlp_get_params(self._m, p.ctypes.data, p.size), "mtc_mlp_get_params")
        return p

    @params.setter
    def params(self, p: np.ndarray) -> None:
        a = np.ascontiguousarray(p, dtype=np.float32).reshape(-1)
        check(lib().mtc_mlp_set_params(self._m, a.ctypes.data, a.size), "mtc_mlp_set_params")

    def collect(self, stage: CloudStage, cam, time, sun, params, seed: int, keep_fraction: float) -> int:
        n = C.c_int64()
        check(lib().mtc_mlp_collect(self._m, stage.handle, C.byref(cam), C.byref(time), C.byref(sun), C.byref(params),
                                    int(seed), float(keep_fraction), C.byref(n)), "mtc_mlp_collect")
        return int(n.value)

    def sample_count(self) -> int:
        n = C.c_int64()
        check(lib().mtc_mlp_sample_count(self._m, C.byref(n)), "mtc_mlp_sample_count")
        return int(n.value)

    def clear_samples(self) -> None:
        check(lib().mtc_mlp_clear_samples(self._m), "mtc_mlp_clear_samples")

    def set_samples(self, xs: np.ndarray, ys: np.ndarray) -> None:
        x = np.ascontiguousarray(xs, dtype=np.float32).reshape(-1, self.N_INPUTS)
        y = np.ascontiguousarray(ys, dtype=np.float32).reshape(-1)
        check(lib().mtc_mlp_set_samples(self._m, x.ctypes.data, y.ctypes.data, y.size), "mtc_mlp_set_samples")

    def read_samples(self, first: int = 0, n: int | None = None):
        if n is None:
            n = self.sample_count() - first
        xs = np.empty((n, self.N_INPUTS), dtype=np.float32)
        ys = np.empty(n, dtype=np.float32)
        keys = np.empty(n, dtype=np.uint64)
        check(lib().mtc_mlp_read_samples(self._m, int(first), int(n), xs.ctypes.data, ys.ctypes.data,
                                         keys.ctypes.data), "mtc_mlp_read_samples")
        return xs, ys, keys

    def train(self, steps: int, batch: int = 4096, lr: float = 2e-3, beta1: float = 0.9, beta2: float = 0.999,
              eps: float = 1e-8, seed: int = 1, holdout: int = 0) -> np.ndarray:
        d = A.MtcMlpTrainDesc(int(steps), int(batch), float(lr), float(beta1), float(beta2), float(eps), int(seed),
                              int(holdout))
        curve = np.empty(int(steps), dtype=np.float32)
        check(lib().mtc_mlp_train(self._m, C.byref(d), curve.ctypes.data), "mtc_mlp_train")
        return curve

    def gradient(self, xs: np.ndarray, ys: np.ndarray):
        x = np.ascontiguousarray(xs, dtype=np.float32).reshape(-1, self.N_INPUTS)
        y = np.ascontiguousarray(ys, dtype=np.float32).reshape(-1)
        g = np.empty(self.N_PARAMS, dtype=np.float32)
        loss = C.c_float()
        check(lib().mtc_mlp_gradient(self._m, x.ctypes.data, y.ctypes.data, y.size, g.ctypes.data, C.byref(loss)),
              "mtc_mlp_gradient")
        return g, float(loss.value)

    def predict(self, xs: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(xs, dtype=np.float32).reshape(-1, self.N_INPUTS)
        y = np.empty(x.shape[0], dtype=np.float32)
        check(lib().mtc_mlp_predict(self._m, x.ctypes.data, x.shape[0], y.ctypes.data), "mtc_mlp_predict")
        return y

    def calibrate(self, holdout: int, max_rms: float, fallback_density: float = 0.8) -> dict:
        c = A.MtcMlpCalibration()
        check(lib().mtc_mlp_calibrate(self._m, int(holdout), float(max_rms), float(fallback_density), C.byref(c)),
              "mtc_mlp_calibrate")
        return {"lo": c.lo, "hi": c.hi, "rms": list(c.rms), "count": list(c.count), "mask": int(c.mask),
                "max_rms": c.max_rms, "fallback_density": c.fallback_density, "rms_all": c.rms_all,
                "rms_mean_predictor": c.rms_mean_predictor}

    def attach(self, stage: CloudStage, audit_period: int = 0) -> None:
        """Use this network in a MTC_KERNEL_LEARNED stage's frames."""
        check(lib().mtc_set_learned_light(stage.handle, self._m, int(audit_period)), "mtc_set_learned_light")
