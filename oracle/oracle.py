"""Python binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker (or the timed CPU
baseline); the product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# METEOROS_ORACLE_LIB points the tests at another build of this oracle (the
# mutation tests build deliberately broken copies on the CPU and check that
# the pins reject them).
LIB_PATH = os.environ.get("METEOROS_ORACLE_LIB") or os.path.join(_HERE, "_build", "liboracle.so")

# The boundary structs are shared with the product's C header; import the
# ctypes mirrors from there (they are plain type definitions).
from meteoros_amd import _abi as A  # noqa: E402

_lib = None
P = C.POINTER


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing; run `make -C oracle`")
        h = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        h.orc_march_pixels.restype = C.c_int
        h.orc_march_pixels.argtypes = [vp, vp, vp, vp, P(A.MtcVolumeDesc), P(A.MtcCamera), P(A.MtcTime), P(A.MtcSun),
                                       P(A.MtcCloudParams), C.c_int, C.c_int, vp, C.c_int64, C.c_int, vp, vp, vp, vp,
                                       C.c_int]
        h.orc_light_samples.restype = C.c_int64
        h.orc_light_samples.argtypes = [vp, vp, vp, vp, P(A.MtcVolumeDesc), P(A.MtcCamera), P(A.MtcTime),
                                        P(A.MtcSun), P(A.MtcCloudParams), C.c_int, C.c_int, vp, C.c_int64, C.c_int,
                                        C.c_int64, vp, vp, vp, vp, vp]
        h.orc_march_trace.restype = C.c_int64
        h.orc_march_trace.argtypes = [vp, vp, vp, vp, P(A.MtcVolumeDesc), P(A.MtcCamera), P(A.MtcTime), P(A.MtcSun),
                                      P(A.MtcCloudParams), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64,
                                      vp, vp, vp]
        h.orc_flop_counts.restype = C.c_int
        h.orc_flop_counts.argtypes = [vp, vp, vp, vp, P(A.MtcVolumeDesc), P(A.MtcCamera), P(A.MtcCamera), P(A.MtcTime),
                                      P(A.MtcSun), P(A.MtcCloudParams), C.c_int, C.c_int, vp, C.c_int64, vp, vp]
        h.orc_god_rays.restype = C.c_int
        h.orc_god_rays.argtypes = [P(A.MtcCamera), P(A.MtcSun), C.c_int, C.c_int, vp, vp, vp, C.c_int, vp]
        h.orc_tone_map.restype = C.c_int
        h.orc_tone_map.argtypes = [C.c_uint32, C.c_int, C.c_int, vp, vp, vp, C.c_int]
        h.orc_txaa.restype = C.c_int
        h.orc_txaa.argtypes = [P(A.MtcCamera), P(A.MtcCamera), P(A.MtcTime), C.c_int, C.c_int, vp, vp, vp, vp, vp,
                               C.c_int]
        h.orc_wang_hash.restype = C.c_uint32
        h.orc_wang_hash.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        h.orc_uncharted2.restype = C.c_float
        h.orc_uncharted2.argtypes = [C.c_float]
        h.orc_select_pixels.restype = C.c_int64
        h.orc_select_pixels.argtypes = [C.c_int, C.c_int, C.c_int, vp]
        h.orc_frame.restype = C.c_int
        h.orc_frame.argtypes = [vp, vp, vp, vp, P(A.MtcVolumeDesc), P(A.MtcCamera), P(A.MtcCamera), P(A.MtcTime),
                                P(A.MtcSun), P(A.MtcCloudParams), C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp,
                                C.c_int]
        h.orc_reprojection_indices.restype = C.c_int
        h.orc_reprojection_indices.argtypes = [P(A.MtcCamera), P(A.MtcCamera), P(A.MtcTime), P(A.MtcCloudParams),
                                               C.c_int, C.c_int, C.c_int, vp, C.c_int]
        h.orc_encode_float_rgba.restype = None
        h.orc_encode_float_rgba.argtypes = [C.c_float, P(C.c_float)]
        h.orc_encode_float_rgba64.restype = None
        h.orc_encode_float_rgba64.argtypes = [C.c_double, P(C.c_double)]
        h.orc_hg.restype = C.c_double
        h.orc_hg.argtypes = [C.c_double, C.c_double, C.c_int]
        h.orc_remap.restype = C.c_double
        h.orc_remap.argtypes = [C.c_double] * 5 + [C.c_int]
        h.orc_smoothstep.restype = C.c_double
        h.orc_smoothstep.argtypes = [C.c_double] * 3
        h.orc_ray_sphere.restype = None
        h.orc_ray_sphere.argtypes = [P(C.c_double), P(C.c_double), P(C.c_double), C.c_double, C.c_int, P(C.c_double)]
        h.orc_sample3d.restype = None
        h.orc_sample3d.argtypes = [vp, C.c_int, C.c_int, P(C.c_double), C.c_int, P(C.c_double)]
        h.orc_sample2d.restype = None
        h.orc_sample2d.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int, P(C.c_double)]
        h.orc_atmosphere.restype = None
        h.orc_atmosphere.argtypes = [P(A.MtcSun), P(A.MtcCloudParams), P(C.c_double), P(C.c_double), C.c_int,
                                     P(C.c_double)]
        h.orc_noise_generate.restype = C.c_int
        h.orc_noise_generate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_int, vp, C.c_int]
        h.orc_to_half.restype = C.c_uint16
        h.orc_to_half.argtypes = [C.c_float]
        h.orc_light_energy.restype = C.c_double
        h.orc_light_energy.argtypes = [C.c_double] * 6 + [C.c_int]
        h.orc_air_mass.restype = C.c_double
        h.orc_air_mass.argtypes = [C.c_double]
        h.orc_rayleigh_phase.restype = C.c_double
        h.orc_rayleigh_phase.argtypes = [C.c_double]
        h.orc_sky_constants.restype = None
        h.orc_sky_constants.argtypes = [P(C.c_double)]
        _lib = h
    return _lib


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


COUNTER_NAMES = ("rays_marched", "rays_ocean", "rays_sky", "primary_steps", "cloud_steps", "cone_samples",
                 "cone_erosions", "early_exits")


class Volumes:
    """The four stage resources as numpy arrays (x fastest, 4 channels)."""

    def __init__(self, lf, hf, curl, weather, lf_dim, hf_dim, curl_dim, weather_dim, fmt=A.MTC_FMT_UNORM8):
        self.lf, self.hf = np.ascontiguousarray(lf), np.ascontiguousarray(hf)
        self.curl, self.weather = np.ascontiguousarray(curl), np.ascontiguousarray(weather)
        self.desc = A.MtcVolumeDesc(lf_dim, hf_dim, curl_dim, weather_dim, fmt)

    def ptrs(self):
        return (self.lf.ctypes.data, self.hf.ctypes.data, self.curl.ctypes.data, self.weather.ctypes.data)


def noise(kind: int, n: int, fmt: int = A.MTC_FMT_UNORM8, seed: int = 1, z0: int = 0, nz: int = 0,
          threads: int = 0) -> np.ndarray:
    """CPU generator: kind 0 = LF (n^3), 1 = HF (n^3), 2 = curl (n^2), 3 = weather (n^2)."""
    depth = (n if nz <= 0 else nz) if kind <= 1 else 1
    count = depth * n * n * 4
    out = np.empty(count, dtype=np.uint16 if fmt == A.MTC_FMT_FP16 else np.uint8)
    rc = lib().orc_noise_generate(kind, n, fmt, seed, z0, nz, out.ctypes.data, threads)
    if rc:
        raise ValueError("orc_noise_generate rejected its arguments")
    return out


def generate_volumes(lf_dim=128, hf_dim=32, curl_dim=128, weather_dim=512, fmt=A.MTC_FMT_UNORM8, seed_volumes=1,
                     seed_weather=2, threads=0) -> Volumes:
    return Volumes(noise(0, lf_dim, fmt, seed_volumes, threads=threads),
                   noise(1, hf_dim, fmt, seed_volumes, threads=threads),
                   noise(2, curl_dim, A.MTC_FMT_UNORM8, seed_volumes, threads=threads),
                   noise(3, weather_dim, A.MTC_FMT_UNORM8, seed_weather, threads=threads),
                   lf_dim, hf_dim, curl_dim, weather_dim, fmt)


def select_pixels(W: int, H: int, frame: int) -> np.ndarray:
    xy = np.empty(((W + 3) // 4) * ((H + 3) // 4) * 2, dtype=np.int32)
    n = lib().orc_select_pixels(W, H, frame, xy.ctypes.data)
    return xy[: 2 * n].reshape(n, 2)


def march_pixels(vol: Volumes, cam, time, sun, params, W, H, xy: np.ndarray, precision=32, threads=0):
    """Returns (rgba[n,4], mask[n,4], steps[n], counters dict), all float64/uint64."""
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1, 2)
    n = xy.shape[0]
    rgba = np.empty((n, 4), dtype=np.float64)
    mask = np.empty((n, 4), dtype=np.float64)
    steps = np.empty(n, dtype=np.uint64)
    cnt = np.zeros(8, dtype=np.uint64)
    rc = lib().orc_march_pixels(*vol.ptrs(), C.byref(vol.desc), C.byref(cam), C.byref(time), C.byref(sun),
                                C.byref(params), W, H, xy.ctypes.data, n, precision, rgba.ctypes.data,
                                mask.ctypes.data, steps.ctypes.data, cnt.ctypes.data, threads)
    if rc:
        raise ValueError("orc_march_pixels rejected its arguments")
    return rgba, mask, steps, dict(zip(COUNTER_NAMES, (int(c) for c in cnt)))


def light_samples(vol: Volumes, cam, time, sun, params, W, H, xy: np.ndarray, precision=64, extra=False):
    """Learned-light training samples of the listed pixels (every in-cloud
    step).  Returns (pixel index into xy [m], step index [m], inputs [m, 9],
    target [m]) with target = -(cone optical depth); with extra=True also
    [m, 8]: the six cone terms and the weather (coverage, type) at the
    cone's first sample."""
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1, 2)
    n = xy.shape[0]
    cap = max(1, n * 64)
    while True:
        pix = np.empty(cap, dtype=np.int32)
        step = np.empty(cap, dtype=np.int32)
        x = np.empty((cap, 9), dtype=np.float64)
        y = np.empty(cap, dtype=np.float64)
        ex = np.empty((cap, 8), dtype=np.float64)
        m = lib().orc_light_samples(*vol.ptrs(), C.byref(vol.desc), C.byref(cam), C.byref(time), C.byref(sun),
                                    C.byref(params), W, H, xy.ctypes.data, n, precision, cap, pix.ctypes.data,
                                    step.ctypes.data, x.ctypes.data, y.ctypes.data, ex.ctypes.data)
        if m < 0:
            raise ValueError("orc_light_samples rejected its arguments")
        if m <= cap:
            out = (pix[:m], step[:m], x[:m], y[:m])
            return out + (ex[:m],) if extra else out
        cap = int(m)


STEP_TRACE = np.dtype([("k", "f8"), ("in_cloud", "f8"), ("t", "f8"), ("pos", "f8", 3), ("sp", "f8", 3),
                       ("sk", "f8", 3), ("h", "f8"), ("lf", "f8", 4), ("base", "f8"), ("curl", "f8", 2),
                       ("hf", "f8", 3), ("ero", "f8", 3), ("hfd", "f8"), ("accum", "f8"), ("dal", "f8"),
                       ("energy", "f8"), ("transmittance", "f8"), ("cone_lf_r", "f8", 6), ("cone_pos", "f8", (6, 3)),
                       ("cone_e", "f8", 6)])
assert STEP_TRACE.itemsize == 61 * 8


def march_trace(vol: Volumes, cam, time, sun, params, W, H, px, py, precision=64):
    """Every executed primary step of pixel (px, py)'s march as a STEP_TRACE
    record array (position, sample point, height, wind-skewed point, the LF /
    curl / HF samples, density, accumulation, cone sample points and terms,
    light energy, transmittance), plus the pixel's colour and mask."""
    cap = 4096
    buf = np.zeros(cap * 61, dtype=np.float64)
    rgba = np.zeros(4, dtype=np.float64)
    mask = np.zeros(4, dtype=np.float64)
    n = lib().orc_march_trace(*vol.ptrs(), C.byref(vol.desc), C.byref(cam), C.byref(time), C.byref(sun),
                              C.byref(params), W, H, int(px), int(py), precision, cap, buf.ctypes.data,
                              rgba.ctypes.data, mask.ctypes.data)
    if n < 0 or n > cap:
        raise ValueError("orc_march_trace rejected its arguments")
    return buf[: n * 61].view(STEP_TRACE).copy(), rgba, mask


FLOP_UNITS = ("ray_marched", "ray_ocean", "ray_sky", "primary_step", "cloud_step", "cone_sample", "cone_erosion",
              "reproject_pixel")


def flop_counts(vol: Volumes, cam, old, time, sun, params, W, H, xy: np.ndarray):
    """Arithmetic of the GLSL as written, counted by running the march and
    the reprojection of the listed pixels through the oracle's templates with
    an operation-counting scalar (oracle/flop_count.hpp): returns
    {unit: (operations, units)} for the units of FLOP_UNITS."""
    xy = np.ascontiguousarray(xy, dtype=np.int32).reshape(-1, 2)
    ops = np.zeros(8, dtype=np.uint64)
    units = np.zeros(8, dtype=np.uint64)
    if lib().orc_flop_counts(*vol.ptrs(), C.byref(vol.desc), C.byref(cam), C.byref(old), C.byref(time), C.byref(sun),
                             C.byref(params), W, H, xy.ctypes.data, xy.shape[0], ops.ctypes.data, units.ctypes.data):
        raise ValueError("orc_flop_counts rejected its arguments")
    return {k: (int(o), int(u)) for k, o, u in zip(FLOP_UNITS, ops, units)}


def god_rays(cam, sun, rgba: np.ndarray, mask: np.ndarray, precision=32):
    """postProcess_GodRays.frag over an HxWx4 image and its god-ray mask.
    Returns (image, (sun_ss_x, sun_ss_y, blend))."""
    H, W = rgba.shape[:2]
    a = np.ascontiguousarray(rgba, dtype=np.float32)
    m = np.ascontiguousarray(mask, dtype=np.float32)
    out = np.empty_like(a)
    consts = np.zeros(3, dtype=np.float64)
    if lib().orc_god_rays(C.byref(cam), C.byref(sun), W, H, a.ctypes.data, m.ctypes.data, out.ctypes.data, precision,
                          consts.ctypes.data):
        raise ValueError("orc_god_rays rejected its arguments")
    return out, tuple(float(c) for c in consts)


def tone_map(time_s: int, rgba: np.ndarray, precision=32):
    """postProcess_ToneMap.frag: (rgb before quantisation [H,W,3], RGBA8 [H,W,4])."""
    H, W = rgba.shape[:2]
    a = np.ascontiguousarray(rgba, dtype=np.float32)
    rgb = np.empty((H, W, 3), dtype=np.float64)
    out8 = np.empty((H, W, 4), dtype=np.uint8)
    if lib().orc_tone_map(int(time_s), W, H, a.ctypes.data, rgb.ctypes.data, out8.ctypes.data, precision):
        raise ValueError("orc_tone_map rejected its arguments")
    return rgb, out8


def txaa(cam, old, time, cur8: np.ndarray, prev8: np.ndarray, precision=32):
    """postProcess_TXAA.frag over HxWx4 RGBA8 current and history images.
    Returns (blend before quantisation [H,W,4], RGBA8 [H,W,4], old uv [H,W,2])."""
    H, W = cur8.shape[:2]
    c = np.ascontiguousarray(cur8, dtype=np.uint8)
    p = np.ascontiguousarray(prev8, dtype=np.uint8)
    rgba = np.empty((H, W, 4), dtype=np.float64)
    out8 = np.empty((H, W, 4), dtype=np.uint8)
    uv = np.empty((H, W, 2), dtype=np.float64)
    if lib().orc_txaa(C.byref(cam), C.byref(old), C.byref(time), W, H, c.ctypes.data, p.ctypes.data,
                      rgba.ctypes.data, out8.ctypes.data, uv.ctypes.data, precision):
        raise ValueError("orc_txaa rejected its arguments")
    return rgba, out8, uv


def wang_hash(u: int, v: int, s: int) -> int:
    return int(lib().orc_wang_hash(u, v, s))


def uncharted2(x: float) -> float:
    return float(lib().orc_uncharted2(x))


def frame(vol: Volumes, cam, old, time, sun, params, W, H, prev=None, precision=32, full=False, threads=0,
          mask=None):
    """One CPU frame: reprojection of prev (if given) then the march.
    Returns (rgba[H,W,4] float32, mask[H,W,4] float32, counters)."""
    cur = np.zeros((H, W, 4), dtype=np.float32)
    m = np.zeros((H, W, 4), dtype=np.float32) if mask is None else np.ascontiguousarray(mask, dtype=np.float32).copy()
    cnt = np.zeros(8, dtype=np.uint64)
    prev_p = None
    if prev is not None:
        prev = np.ascontiguousarray(prev, dtype=np.float32)
        prev_p = prev.ctypes.data
    rc = lib().orc_frame(*vol.ptrs(), C.byref(vol.desc), C.byref(cam), C.byref(old) if old is not None else None,
                         C.byref(time), C.byref(sun), C.byref(params), W, H, precision, int(full), prev_p,
                         cur.ctypes.data, m.ctypes.data, cnt.ctypes.data, threads)
    if rc:
        raise ValueError("orc_frame rejected its arguments")
    return cur, m, dict(zip(COUNTER_NAMES, (int(c) for c in cnt)))


def reprojection_indices(cam, old, time, params, W, H, precision=32, threads=0) -> np.ndarray:
    N = int(params.motion_blur_samples)
    idx = np.empty(W * H * N, dtype=np.int32)
    rc = lib().orc_reprojection_indices(C.byref(cam), C.byref(old), C.byref(time), C.byref(params), W, H, precision,
                                        idx.ctypes.data, threads)
    if rc:
        raise ValueError("orc_reprojection_indices rejected its arguments")
    return idx.reshape(H, W, N)


def encode_float_rgba(v: float, precision=32):
    if precision == 64:
        out = (C.c_double * 4)()
        lib().orc_encode_float_rgba64(v, out)
    else:
        out = (C.c_float * 4)()
        lib().orc_encode_float_rgba(v, out)
    return np.array(list(out))


def hg(c, g, precision=32):
    return lib().orc_hg(c, g, precision)


def remap(v, a, b, c, d, kind=0):
    return lib().orc_remap(v, a, b, c, d, kind)


def smoothstep(e0, e1, x):
    return lib().orc_smoothstep(e0, e1, x)


def ray_sphere(o, d, c, R, precision=32):
    out = (C.c_double * 5)()
    lib().orc_ray_sphere(_d3(o), _d3(d), _d3(c), float(R), precision, out)
    return bool(out[0]), np.array(out[1:4]), float(out[4])


def sample3d(vol: np.ndarray, fmt: int, n: int, s, precision=32):
    out = (C.c_double * 4)()
    v = np.ascontiguousarray(vol)
    lib().orc_sample3d(v.ctypes.data, fmt, n, _d3(s), precision, out)
    return np.array(list(out))


def sample2d(m: np.ndarray, n: int, sx, sy, precision=32):
    out = (C.c_double * 4)()
    v = np.ascontiguousarray(m, dtype=np.uint8)
    lib().orc_sample2d(v.ctypes.data, n, float(sx), float(sy), precision, out)
    return np.array(list(out))


def atmosphere(sun, params, direction, origin=(0.0, 0.0, 0.0), precision=32):
    out = (C.c_double * 3)()
    lib().orc_atmosphere(C.byref(sun), C.byref(params), _d3(direction), _d3(origin), precision, out)
    return np.array(list(out))


def light_energy(h, dl, ds, phase, cosang, brightness=1.0, precision=64) -> float:
    return float(lib().orc_light_energy(float(h), float(dl), float(ds), float(phase), float(cosang), float(brightness),
                                        int(precision)))


def air_mass(zenith_rad) -> float:
    return float(lib().orc_air_mass(float(zenith_rad)))


def rayleigh_phase(c) -> float:
    return float(lib().orc_rayleigh_phase(float(c)))


def sky_constants():
    """(RAYLEIGH_TOTAL, MIE_CONST) as the oracle holds them."""
    out = (C.c_double * 6)()
    lib().orc_sky_constants(out)
    v = np.array(list(out))
    return v[:3], v[3:]


def to_half(f: float) -> int:
    return int(lib().orc_to_half(f))


def ray_directions(cam, time, xy, W, H):
    """castRay (cloudRayMarch.comp:194-226) in float64 numpy for a list of
    pixels: used to locate the sun disk in an image."""
    xy = np.asarray(xy).reshape(-1, 2)
    v = np.array(cam.view, dtype=np.float64).reshape(4, 4)  # v[c][r], column-major
    right = v[:3, 0] / np.linalg.norm(v[:3, 0])
    up = v[:3, 1] / np.linalg.norm(v[:3, 1])
    look = -v[:3, 2] / np.linalg.norm(v[:3, 2])
    idx = int(time.frame_count_mod16) // 2
    hs = [time.halton1, time.halton2, time.halton3, time.halton4]
    jx, jy = (hs[0][idx], hs[1][idx]) if idx < 4 else (hs[2][idx - 4], hs[3][idx - 4])
    ux = xy[:, 0] / W
    uy = 1.0 - xy[:, 1] / H
    nx = ux * 2 - 1 + jx / W
    ny = uy * 2 - 1 + jy / H
    d = (look[None] + (nx * cam.tan_fov_by2[0])[:, None] * right[None]
         + (ny * cam.tan_fov_by2[1])[:, None] * up[None])
    return d / np.linalg.norm(d, axis=1, keepdims=True)


SUN_DISK_RADIUS_DEG = 0.6


def outside_sun_disk(cam, time, params, xy, W, H, radius_deg=SUN_DISK_RADIUS_DEG):
    """True for pixels farther than radius_deg from the background sun.

    The sun-disk term of the Preetham background (cloudRayMarch.comp:460-461)
    is smoothstep(cos 0.533 deg, +2e-5, cos theta) * 15000 * sunE: over the
    disk's edge a change of one float ulp in the ray direction moves a
    pixel's value by ~1% of ~3e4, so those few pixels are not a measure of
    arithmetic accuracy and the image PSNR leaves them out."""
    d = ray_directions(cam, time, xy, W, H)
    s = np.array(params.sky_sun_location[:3], dtype=np.float64)
    s /= np.linalg.norm(s)
    return d @ s < np.cos(np.radians(radius_deg))


def error_vs_fp64(got_rgb, ref_rgb, keep):
    """Per-pixel error of the kept pixels against the FP64 reference: PSNR
    (as psnr_vs_fp64), and the largest and the 99.9th-percentile absolute
    RGB error, each as a fraction of the peak (max of the FP64 reference)."""
    g = np.asarray(got_rgb, dtype=np.float64)[keep]
    r = np.asarray(ref_rgb, dtype=np.float64)[keep]
    peak = float(r.max())
    err = np.abs(g - r).max(axis=1) / peak
    return {"psnr_db": psnr_vs_fp64(got_rgb, ref_rgb, keep), "max_err_frac_of_peak": float(err.max()),
            "p999_err_frac_of_peak": float(np.quantile(err, 0.999)), "rays": int(r.shape[0])}


def psnr_vs_fp64(got_rgb, ref_rgb, keep):
    """PSNR over RGB of the kept pixels, peak = max of the FP64 reference."""
    g = np.asarray(got_rgb, dtype=np.float64)[keep]
    r = np.asarray(ref_rgb, dtype=np.float64)[keep]
    mse = float(np.mean((g - r) ** 2))
    if mse == 0.0:
        return float("inf")
    return 10.0 * np.log10(float(r.max()) ** 2 / mse)
