"""Independent pins of the march oracle's composition (oracle/cloud_oracle.hpp),
on the CPU, against cloudRayMarch.comp.

The FP64 twin is driven with inputs for which the shader's arithmetic has a
closed form derived here from the GLSL text, not from the oracle's code:

* constant volumes, a vertical ray and the sun overhead: every texture()
  returns its texel, so the march is the per-step recurrence of :627-686
  (base density :499-540, erosion :542-563, the six cone samples of
  :652-669 with term 0 at the step's own point, accum += 0.5 e, T <-
  mix(T, E, 1 - accum), early exit) written out below in numpy;
* volumes that are linear ramps along one axis: a trilinear sample of a ramp
  gives back the coordinate it was taken at, so the recorded samples say
  where the LF, curl and HF volumes were read.  That pins the wind skew
  (:489-497), the curl displacement of the erosion point (:546-549) and the
  cone's sample points (unskewed and not divided by 8, :658-661);
* the step positions, relative position (:188-191) and relative height
  (:171-186) against the ray geometry (law of cosines about the planet
  centre).

Zero jitter (a Halton table of zeros) keeps every ray on its pixel's
direction.  Tolerances are relative 1e-9 (FP64; the closed forms round
differently from the oracle only in the last bits).  The runtime parameters
travel as float32 (mtc_cloud_params), so the closed forms use their float32
values (f32 below).
"""
import math

import numpy as np
import pytest

import meteoros_amd as M
from meteoros_amd import _abi as A
from oracle import oracle as O

W = H = 8
RE = 6371000.0
R_IN, R_OUT = RE + 1500.0, RE + 4000.0
THICK = R_OUT - R_IN
KV = np.array([[0.1, 0.25, -0.15], [0.2, 0.5, 0.2], [-0.2, 0.1, -0.1], [-0.05, 0.75, 0.05], [-0.1, 1.0, 0.0],
               [0.0, 3.0, 0.0]])  # the cone kernel, cloudRayMarch.comp:598-609


def up_camera(tan_half=0.5):
    """A camera at the origin looking straight up: right = +x, up = +z,
    look = +y (view rows (1,0,0), (0,0,1), (0,-1,0))."""
    cam = A.MtcCamera()
    rows = [(1.0, 0.0, 0.0), (0.0, 0.0, 1.0), (0.0, -1.0, 0.0)]
    for r in range(3):
        for c in range(3):
            cam.view[c * 4 + r] = rows[r][c]
    cam.view[15] = 1.0
    cam.tan_fov_by2[0] = cam.tan_fov_by2[1] = tan_half
    return cam


def pixel_dir(px, py, tan_half=0.5):
    """castRay (:194-226) of up_camera with zero jitter, in float64."""
    nx = (px / W) * 2.0 - 1.0
    ny = (1.0 - py / H) * 2.0 - 1.0
    d = np.array([0.0, 1.0, 0.0]) + nx * tan_half * np.array([1.0, 0.0, 0.0]) + ny * tan_half * np.array([0.0, 0.0, 1.0])
    return d / np.linalg.norm(d)


def still_time(total=0.0):
    t = A.MtcTime()  # Halton tables of zeros: no jitter
    t.time[1] = total
    t.frame_count_mod16 = 0
    return t


def scene(steps=64, coverage=0.6, speed=0.08, light_samples=6):
    p, sun = M.preset_params(1)
    p.earth_radius, p.shell_inner_radius, p.shell_outer_radius = RE, R_IN, R_OUT
    p.step_mode = A.MTC_STEPS_FIXED
    p.steps_min = p.steps_max = steps
    p.light_samples = light_samples
    p.coverage_mode = A.MTC_COVERAGE_CONSTANT
    p.coverage_constant = coverage
    p.base_density_factor, p.erosion_scale, p.accum_scale, p.cone_erosion_scale = 0.38, 1.4, 0.5, 1.5
    p.sample_divisor, p.brightness = 8.0, 5.0
    p.hg_eccentricity, p.hg_silver_intensity, p.hg_silver_spread = 0.6, 0.7, 0.1
    p.wind_direction[:] = (1.0, 0.0, 0.0, 0.0)
    p.wind_upward_bias, p.cloud_speed, p.cloud_top_offset, p.wind_skew = 0.1, speed, 1.0, 0.009
    p.cloud_fade_start, p.cloud_fade_end = 0.06, 0.2
    sun.sun_location[:] = (0.0, 1.0e8, 0.0, 1.0)  # overhead: the cone kernel points straight up
    return p, sun


def volumes(lf_code=(150, 200, 100, 50), hf_code=(180, 60, 120, 0), curl_code=(90, 170, 0, 255),
            lf=None, hf=None, curl=None):
    n_lf, n_hf, n_c, n_w = 128, 32, 128, 16
    lf = np.broadcast_to(np.array(lf_code, np.uint8), (n_lf, n_lf, n_lf, 4)) if lf is None else lf
    hf = np.broadcast_to(np.array(hf_code, np.uint8), (n_hf, n_hf, n_hf, 4)) if hf is None else hf
    curl = np.broadcast_to(np.array(curl_code, np.uint8), (n_c, n_c, 4)) if curl is None else curl
    weather = np.full((n_w, n_w, 4), 255, np.uint8)
    return O.Volumes(lf, hf, curl, weather, n_lf, n_hf, n_c, n_w)


def f32(x):
    return float(np.float32(x))


def hg_forward(g):
    """Henyey-Greenstein at cos = 1: (1 + g) / (4 pi (1 - g)^2)."""
    return (1.0 + g) / (4.0 * math.pi * (1.0 - g) ** 2)


def shell_t(R):
    """raySphereIntersection (:229-273) of the vertical ray from the origin:
    the hit is at height R - RE, and t = length(hit - rO / R) with the
    normalised origin rO / R = (0, RE / R, 0) (the shader's unit mix)."""
    return (R - RE) - RE / R


def closed_form_vertical(lf_code, hf_code, steps, coverage, n_taken=None):
    """The march of the vertical ray through constant volumes, from the
    GLSL text (cloudRayMarch.comp:499-563, :565-688, :331-388).
    Returns (colour, accum, steps, early_exit)."""
    r, g, b, a = (c / 255.0 for c in lf_code)
    fbm = min(max(0.625 * g + 0.25 * b + 0.125 * a, 0.0), 1.0)
    base_lf = min(max((r - (fbm - 0.9)) / (1.0 - (fbm - 0.9)), 0.0), 1.0)
    coverage = f32(coverage)
    s_density = (min(max(base_lf, coverage), 1.0) - coverage) / (1.0 - coverage) * coverage
    base = s_density * f32(0.38)
    hr, hg, hb = (c / 255.0 for c in hf_code[:3])
    f_hf = 0.625 * hr + 0.25 * hg + 0.125 * hb

    def erode(x, h):
        m = min(max(f_hf * (1.0 - min(2.0 * h, 1.0)) + (1.0 - f_hf) * min(2.0 * h, 1.0), 0.0), 1.0)
        return (x - 0.005 * m) / (1.0 - 0.005 * m)

    phase = max(hg_forward(f32(0.6)), f32(0.7) * hg_forward(0.99 - f32(0.1)))
    start, end = shell_t(R_IN), shell_t(R_OUT)
    step = (end - start) / steps
    t, accum, trans, col, k = start, 0.0, 1.0, 0.0, 0
    while t < end if n_taken is None else k < n_taken:
        h = abs(t - (R_IN - RE)) / THICK
        k += 1
        if base > 0.0:
            accum += 0.5 * erode(f32(1.4) * base, h)
            dal = 6.0 * erode(1.5 * s_density, h) if s_density > 0.0 else 0.0
            # toward the sun (cos = 1): attenuation = exp(-dl); depth
            # exponent clamps to 0.5 (h / 8 < 0.3)
            vert = min(max(0.1 + (1.5 * h - 0.07) / 0.27 * 0.9, 0.1), 1.0) ** 0.8
            energy = math.exp(-dal) * math.exp(-dal) * (0.05 + math.sqrt(base)) * vert * phase * 5.0
            trans = trans * accum + energy * (1.0 - accum)
            col += trans
        if accum >= 1.0:
            return col, 1.0, k, True
        t += step
    return col, accum, k, False


CENTRE = (W // 2, H // 2)


@pytest.mark.parametrize("lf_code,steps,exits", [((150, 200, 100, 50), 64, True), ((128, 200, 100, 50), 64, False),
                                                 ((170, 90, 230, 30), 48, True)])
def test_constant_volume_march_is_the_closed_form_recurrence(lf_code, steps, exits):
    p, sun = scene(steps=steps)
    vol = volumes(lf_code=lf_code)
    cam, tm = up_camera(), still_time(0.0)
    xy = np.array([CENTRE], np.int32)
    rgba, mask, nsteps, cnt = O.march_pixels(vol, cam, tm, sun, p, W, H, xy, precision=64)
    n = int(nsteps[0])
    assert n in (steps, steps + 1) or exits
    col, accum, k, early = closed_form_vertical(lf_code, (180, 60, 120, 0), steps, 0.6,
                                                n_taken=None if exits else n)
    assert early == exits and k == n
    assert cnt["primary_steps"] == n and cnt["cloud_steps"] == n
    assert cnt["cone_samples"] == 6 * n and cnt["cone_erosions"] == 6 * n
    assert cnt["early_exits"] == int(exits)
    if exits:
        # mix(background, colour, 1) is the colour alone; the mask encodes 0
        assert np.allclose(rgba[0, :3], col, rtol=1e-9, atol=0)
        assert np.all(mask[0] == 0.0)
    else:
        bg = O.atmosphere(sun, p, (0.0, 1.0, 0.0), precision=64) * max(p.background_min, 1.0)
        assert np.allclose(rgba[0, :3], bg * (1.0 - accum) + col * accum, rtol=1e-9, atol=0)
        # the mask encodes 25 * min(0.05, 1 - accum); decoding returns its
        # fractional part (EncodeFloatRGBA keeps fract(v) only)
        grey = 25.0 * min(0.05, 1.0 - accum)
        dec = mask[0] @ (1.0 / np.array([1.0, 255.0, 65025.0, 16581375.0]))
        assert abs(dec - (grey - math.floor(grey))) < 1e-9
    print(f"lf {lf_code}: {n} steps, accum {accum:.6f}, colour {col:.6f}, early exit {early}")


def test_constant_volume_with_no_cone_samples_keeps_full_light():
    """light_samples = 0: dl = 0, so every in-cloud step's energy is the
    closed form with exp(-dl) = 1."""
    p, sun = scene(steps=32, light_samples=0)
    vol = volumes(lf_code=(128, 200, 100, 50))
    rgba, _, nsteps, cnt = O.march_pixels(vol, up_camera(), still_time(), sun, p, W, H,
                                          np.array([CENTRE], np.int32), precision=64)
    assert cnt["cone_samples"] == 0
    tr, _, _ = O.march_trace(vol, up_camera(), still_time(), sun, p, W, H, *CENTRE)
    assert np.all(tr["dal"] == 0.0)
    phase = max(hg_forward(f32(0.6)), f32(0.7) * hg_forward(0.99 - f32(0.1)))
    vert = np.clip(0.1 + (1.5 * tr["h"] - 0.07) / 0.27 * 0.9, 0.1, 1.0) ** 0.8
    assert np.allclose(tr["energy"], (0.05 + np.sqrt(tr["base"])) * vert * phase * 5.0, rtol=1e-9, atol=0)


def _trace(vol, px, py, total=0.0, **kw):
    p, sun = scene(**kw)
    tr, rgba, mask = O.march_trace(vol, up_camera(), still_time(total), sun, p, W, H, px, py)
    assert len(tr) > 0
    return tr, p


@pytest.mark.parametrize("px,py", [CENTRE, (1, 2), (7, 0), (6, 5)])
def test_step_positions_relative_position_and_height(px, py):
    """pos_k = eye + t_k d (zero jitter); sample point = (pos - (c.x,
    R_in - R, c.z)) / thickness / 8; relative height = |cos(d, radial) *
    (t_k - t_s)| / thickness, with t_s the distance to the inner-shell hit
    and cos from the law of cosines about the planet centre."""
    d = pixel_dir(px, py)
    tr, _ = _trace(volumes(), px, py)
    t = tr["t"]
    assert np.allclose(tr["pos"], t[:, None] * d[None], rtol=1e-12, atol=1e-9)
    c = np.array([0.0, -RE, 0.0])
    sp = (tr["pos"] - np.array([0.0, R_IN - RE, 0.0])) / THICK / 8.0
    assert np.allclose(tr["sp"], sp, rtol=1e-12, atol=1e-15)
    cos_z = d[1]
    t_s = -RE * cos_z + math.sqrt(RE * RE * cos_z * cos_z + R_IN * R_IN - RE * RE)
    r_k = np.sqrt(RE * RE + t * t + 2.0 * RE * t * cos_z)
    cos_a = (t + RE * cos_z) / r_k
    h = np.abs(cos_a * (t - t_s)) / THICK
    assert np.allclose(tr["h"], h, rtol=1e-9, atol=1e-12)
    # the first step starts about a metre before the shell (t = length(hit - rO / R))
    assert 0.5 < t_s - t[0] < 1.5
    assert np.all(np.diff(t) > 0)


def _ramp3(axis, n=128, lo=64):
    """An n^3 RGBA8 volume whose r code is lo + i along `axis` (0 = x, 1 = y,
    2 = z) and whose g, b, a are 255."""
    v = np.full((n, n, n, 4), 255, np.uint8)
    i = np.arange(n, dtype=np.int32) + lo
    shape = [1, 1, 1]
    shape[2 - axis] = n  # array axes are (z, y, x)
    v[..., 0] = np.broadcast_to(i.reshape(shape), (n, n, n)).astype(np.uint8)
    return v


def _coord_from_ramp(r, n, lo=64):
    """Sample coordinate (mod 1) that a trilinear sample of _ramp3 came from,
    valid where the sample lies between texel centres 0 and n - 1 (not
    across the wrap seam, where the ramp jumps back)."""
    u = r * 255.0 - lo  # u = s * n - 0.5 between texels 0 and n - 1
    return (u + 0.5) / n


def _frac(x):
    return x - np.floor(x)


def _clear_of_seam(s, n):
    """Coordinates whose sample does not straddle the wrap seam (mod 1
    within half a texel of 0)."""
    u = _frac(s) * n
    return (u > 0.5 + 1e-6) & (u < n - 0.5 - 1e-6)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("total", [0.0, 3.0, 11.5])
def test_wind_skew_moves_the_lf_sample_point(axis, total):
    """skewSamplePointWithWind (:489-497): sk = sp + h * 0.009 * wind +
    (wind + (0, 0.1, 0)) * 0.08 * time, wind = (1, 0, 0).  A ramp along each
    axis returns the coordinate the LF was read at."""
    lf = _ramp3(axis)
    vol = volumes(lf=lf)
    tr, _ = _trace(vol, 6, 5, total=total, coverage=0.05)
    expect = (tr["sp"] + np.outer(tr["h"], [f32(0.009), 0.0, 0.0])
              + (np.array([1.0, f32(0.1), 0.0]) * f32(0.08)) * f32(total))
    assert np.allclose(tr["sk"], expect, rtol=1e-12, atol=1e-14)
    s = _coord_from_ramp(tr["lf"][:, 0], 128)
    ok = _clear_of_seam(expect[:, axis], 128)
    assert ok.sum() > 0.8 * len(tr)
    assert np.allclose(s[ok], _frac(expect[ok, axis]), rtol=0, atol=1e-9)


def test_cone_samples_read_the_lf_unskewed_and_undivided():
    """:652-669: cone sample i reads the LF at getRelativePositionInAtmosphere(
    pos + step * kernel_i * i), neither wind-skewed nor divided by 8, and
    sample 0 is the step's own point.  With the sun overhead the kernel is
    (0, k_i.y, 0) (the max-|component| frame of :598-624 degenerates to the
    sun axis), and a y ramp returns each sample's y."""
    vol = volumes(lf=_ramp3(1))
    tr, p = _trace(vol, *CENTRE, total=7.0, coverage=0.05)
    cloud = tr[tr["in_cloud"] == 1.0]
    assert len(cloud) > 10
    step = (shell_t(R_OUT) - shell_t(R_IN)) / p.steps_max
    checked = 0
    for st in cloud:
        pos_i = st["pos"][None] + step * np.arange(6)[:, None] * np.outer(KV[:, 1], [0.0, 1.0, 0.0])
        slp = (pos_i - np.array([0.0, R_IN - RE, 0.0])) / THICK
        assert np.allclose(st["cone_pos"], slp, rtol=1e-12, atol=1e-12)
        s = _coord_from_ramp(st["cone_lf_r"], 128)
        ok = _clear_of_seam(slp[:, 1], 128)
        checked += int(ok.sum())
        assert np.allclose(s[ok], _frac(slp[ok, 1]), rtol=0, atol=1e-9)
    assert checked > 4 * len(cloud)
    # sample 0 sits at the step's point: its sample is the primary one only
    # when neither skew nor the divisor moves it, which they do here
    assert not np.allclose(cloud["cone_pos"][:, 0], cloud["sk"])


def test_erosion_reads_curl_at_the_skewed_point_and_hf_at_the_curled_point():
    """erodeCloudWithHighFrequency (:542-563): curl = texture(curl, sk.xy);
    sk.xy += curl.xy * (1 - h) * 0.5; hf = texture(HF, sk).  An x ramp in the
    curl map and an x ramp in the HF volume return both sample points."""
    curl = np.zeros((128, 128, 4), np.uint8)
    curl[..., 0] = (np.arange(128) + 64).astype(np.uint8)[None, :]
    curl[..., 1] = 170
    curl[..., 3] = 255
    hf = np.zeros((32, 32, 32, 4), np.uint8)
    hf[..., 0] = (np.arange(32) + 64).astype(np.uint8)[None, None, :]
    hf[..., 1:3] = 128
    vol = volumes(lf_code=(130, 255, 255, 255), hf=hf, curl=curl)
    tr, _ = _trace(vol, 6, 5, total=5.0, coverage=0.05)
    cloud = tr[tr["in_cloud"] == 1.0]
    assert len(cloud) > 10
    s_c = _coord_from_ramp(cloud["curl"][:, 0], 128)
    ok_c = _clear_of_seam(cloud["sk"][:, 0], 128)
    assert ok_c.sum() > 0.8 * len(cloud)
    assert np.allclose(s_c[ok_c], _frac(cloud["sk"][ok_c, 0]), rtol=0, atol=1e-9)
    assert np.allclose(cloud["curl"][:, 1], 170.0 / 255.0, rtol=1e-12)
    k = (1.0 - cloud["h"]) * 0.5
    ero = cloud["sk"].copy()
    ero[:, 0] += cloud["curl"][:, 0] * k
    ero[:, 1] += cloud["curl"][:, 1] * k
    assert np.allclose(cloud["ero"], ero, rtol=1e-12, atol=1e-14)
    s_h = _coord_from_ramp(cloud["hf"][:, 0], 32)
    ok_h = _clear_of_seam(ero[:, 0], 32)
    assert ok_h.sum() > 0.8 * len(cloud)
    assert np.allclose(s_h[ok_h], _frac(ero[ok_h, 0]), rtol=0, atol=1e-9)


def test_cone_terms_and_transmittance_follow_the_step_values():
    """Per in-cloud step, from the recorded samples alone: the base density
    and each cone term are the GLSL formulas of :499-563 applied to them,
    dl is the sum of all six terms (term 0 included), and accum / T follow
    accum += 0.5 e, T <- mix(T, E, 1 - accum)."""
    vol = volumes(lf=_ramp3(0), hf_code=(40, 230, 90, 0))
    tr, p = _trace(vol, 6, 5, total=5.0, coverage=0.1)
    cloud = tr["in_cloud"] == 1.0
    assert cloud.sum() > 10
    g = b = a = 1.0
    fbm = min(0.625 * g + 0.25 * b + 0.125 * a, 1.0)

    def density(r, cov):
        base = np.clip((r - (fbm - 0.9)) / (1.0 - (fbm - 0.9)), 0.0, 1.0)
        return (np.clip(base, cov, 1.0) - cov) / (1.0 - cov) * cov

    assert np.allclose(tr["base"], density(tr["lf"][:, 0], f32(0.1)) * f32(0.38), rtol=1e-12, atol=1e-15)
    f_hf = (0.625 * 40 + 0.25 * 230 + 0.125 * 90) / 255.0
    h = tr["h"]
    m = np.clip(f_hf * (1 - np.clip(2 * h, 0, 1)) + (1 - f_hf) * np.clip(2 * h, 0, 1), 0, 1)
    cone_d = density(tr["cone_lf_r"], f32(0.1))
    e = (1.5 * cone_d - 0.005 * m[:, None]) / (1.0 - 0.005 * m[:, None])
    e = np.where(cone_d > 0.0, e, 0.0)
    assert np.allclose(tr["cone_e"][cloud], e[cloud], rtol=1e-9, atol=1e-15)
    assert np.allclose(tr["dal"][cloud], e[cloud].sum(1), rtol=1e-9, atol=1e-15)
    hfd = (f32(1.4) * tr["base"] - 0.005 * m) / (1.0 - 0.005 * m)
    acc = np.cumsum(np.where(cloud, 0.5 * hfd, 0.0))
    assert np.allclose(tr["accum"], np.minimum(acc, 1.0), rtol=1e-9, atol=1e-15)
    trans, t_prev = [], 1.0
    for k in range(len(tr)):
        if cloud[k]:
            t_prev = t_prev * acc[k] + tr["energy"][k] * (1.0 - acc[k])
        trans.append(t_prev)
    assert np.allclose(tr["transmittance"], trans, rtol=1e-9, atol=1e-15)


def _encode(v):
    """EncodeFloatRGBA (:105-112) in float64: fract(v * (1, 255, 65025,
    16581375)), each component minus the next one / 255 (the last kept)."""
    e = v * np.array([1.0, 255.0, 65025.0, 16581375.0])
    e = e - np.floor(e)
    return e - np.array([e[1], e[2], e[3], 0.0]) / 255.0


def test_mask_of_rays_below_0_05_is_scaled_by_five():
    """main (:772-779): greyScale = EncodeFloatRGBA(25 min(0.05, 1 - accum)),
    multiplied by max(5, dir.y) when dir.y < 0.05.  That branch is reached
    only with a cloud fade start below 0.05: here 0.01, and a camera
    looking along -z with a narrow field so that some rays climb at
    0.01 < dir.y < 0.05 and others above 0.05."""
    cam = A.MtcCamera()
    for i in range(4):
        cam.view[i * 5] = 1.0  # identity view: right +x, up +y, look -z
    cam.tan_fov_by2[0] = cam.tan_fov_by2[1] = 0.1
    p, sun = scene(steps=48, coverage=0.3)
    p.cloud_fade_start = 0.01
    vol = volumes(lf=_ramp3(0))
    seen = set()
    for py in range(H):
        tr, rgba, mask = O.march_trace(vol, cam, still_time(3.0), sun, p, W, H, 5, py)
        dy = (1.0 - py / H) * 2.0 - 1.0
        dy = dy * 0.1 / math.sqrt(1.0 + (dy * 0.1) ** 2 + ((5 / W) * 2.0 - 1.0) ** 2 * 0.01)
        if dy < 0.01 or abs(dy - 0.05) < 2e-3:
            continue
        assert len(tr) > 0
        accum = tr["accum"][-1] if tr["accum"][-1] < 1.0 else 1.0
        expect = _encode(25.0 * min(0.05, 1.0 - accum))
        if dy < 0.05:
            expect = expect * 5.0
        seen.add(dy < 0.05)
        assert np.allclose(mask, expect, rtol=1e-9, atol=1e-12), (py, dy, mask, expect)
    assert seen == {True, False}
