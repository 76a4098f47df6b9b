"""Pins of the CPU oracle (oracle/cloud_oracle.hpp) against what the
mathematics fixes: closed forms, brute force, invariants.  The reference has
no tests or golden vectors of its own (its shaders only run inside a Vulkan
driver, absent here), so each function of the oracle is pinned here in a way
that a dropped term, a wrong sign, a transposed index or a wrong constant
fails."""
import math

import numpy as np
import pytest

import meteoros_amd as M
from oracle import oracle as O
from tests.common import ground_camera, psnr, volumes


# --- EncodeFloatRGBA (cloudRayMarch.comp:105-112) ---------------------------
def test_encode_float_rgba_closed_form_and_decode():
    bitenc = np.array([1.0, 255.0, 65025.0, 16581375.0])
    for v in [0.0, 0.25, 0.5, 0.7, 1.0 / 3.0, 0.123456, 0.999]:
        enc = O.encode_float_rgba(v, 64)
        frac = np.modf(bitenc * v)[0]
        exp = frac - np.array([frac[1], frac[2], frac[3], 0.0]) / np.array([255.0, 255.0, 255.0, 1.0])
        assert np.allclose(enc, exp, atol=1e-12)
        # the god-ray pass decodes with dot(x, 1/bitEnc) (postProcess_GodRays.frag:39-43)
        assert abs(float(np.dot(enc, 1.0 / bitenc)) - v) < 1e-9
    # values above 1 wrap: 1.25 -> x = fract(1.25) - fract(318.75)/255
    assert abs(O.encode_float_rgba(1.25)[0] - (0.25 - 0.75 / 255.0)) < 1e-6


# --- remap family (:153-169) -------------------------------------------------
def test_remap_closed_forms():
    assert O.remap(0.5, 0.0, 1.0, 10.0, 20.0, 0) == 15.0
    assert O.remap(2.0, 0.0, 1.0, 0.0, 1.0, 0) == 2.0          # unclamped
    assert O.remap(2.0, 0.0, 1.0, 0.0, 1.0, 1) == 1.0          # clamped after
    assert O.remap(-1.0, 0.0, 1.0, 0.0, 1.0, 2) == 0.0         # clamped before and after
    # reversed output range (used with nmin > nmax at :342)
    assert abs(O.remap(0.85, 0.7, 1.0, 2.0, 0.5, 0) - (2.0 + 0.5 * -1.5)) < 1e-6
    # remapClamped clamps to [nmin, nmax] in that order: min(max(t, nmin), nmax)
    assert O.remap(5.0, 0.0, 1.0, 0.0, 2.0, 1) == 2.0


def test_smoothstep_closed_form():
    for x in np.linspace(-0.5, 1.5, 21):
        t = min(max(x, 0.0), 1.0)
        assert abs(O.smoothstep(0.0, 1.0, x) - t * t * (3 - 2 * t)) < 1e-6


# --- Henyey-Greenstein (:295-300) --------------------------------------------
@pytest.mark.parametrize("g", [0.0, 0.3, 0.6, 0.8, -0.4])
def test_hg_is_normalised_over_the_sphere(g):
    # integral over solid angle = 2*pi * int_{-1}^{1} HG(mu) dmu = 1
    mu, w = np.polynomial.legendre.leggauss(400)
    total = 2 * math.pi * sum(wi * O.hg(float(m), g, 64) for m, wi in zip(mu, w))
    assert abs(total - 1.0) < 1e-4


def test_hg_endpoint_closed_forms():
    for g in (0.6, 0.89):
        assert abs(O.hg(1.0, g, 64) - (1 + g) / (4 * math.pi * (1 - g) ** 2)) < 1e-9 * (1 + g) / (1 - g) ** 2
        assert abs(O.hg(-1.0, g, 64) - (1 - g) / (4 * math.pi * (1 + g) ** 2)) < 1e-9


# --- ray / sphere (:229-273) -------------------------------------------------
def test_ray_sphere_against_analytic_roots_and_the_t_quirk():
    R = 6371000.0 + 4000.0
    c = np.array([0.0, -6371000.0, 0.0])
    rng = np.random.default_rng(5)
    for _ in range(50):
        d = rng.normal(size=3)
        d[1] = abs(d[1]) + 0.05
        d /= np.linalg.norm(d)
        o = np.array([rng.uniform(-50, 50), rng.uniform(0, 100), rng.uniform(-50, 50)])
        valid, p, t = O.ray_sphere(o, d, c, R, 64)
        assert valid
        # textbook quadratic in world units: |o + s d - c| = R, far root (o is inside)
        oc = o - c
        b = np.dot(d, oc)
        s = -b + math.sqrt(b * b - (np.dot(oc, oc) - R * R))
        assert np.allclose(p, o + s * d, rtol=0, atol=1e-6)
        assert abs(np.linalg.norm(p - c) - R) < 1e-6
        # t = length(point_world - rO_normalised) (:269), not the distance from o
        rn = (o - c) / R
        assert abs(t - np.linalg.norm(p - rn)) < 1e-6
        assert abs(t - np.linalg.norm(p - o)) > 1e-3 or np.linalg.norm(o - rn) < 1e-3


def test_ray_sphere_miss_and_behind():
    c = np.zeros(3)
    valid, p, t = O.ray_sphere([0, 0, 10], [0, 1, 0], c, 1.0, 64)
    assert not valid and np.all(p == 0)
    valid, p, t = O.ray_sphere([0, 0, 10], [0, 0, 1], c, 1.0, 64)  # sphere behind the ray
    assert not valid


def test_fp32_shell_intersection_cancellation_is_real():
    """The FP32 quadratic cancels at C = dot(rO, rO) - 1: its hit point moves
    by metres against FP64, the reason the precision variant keeps FP64."""
    c = np.array([0.0, -6371000.0, 0.0])
    R = 6371000.0 + 1500.0
    errs = []
    for dy in (0.05, 0.1, 0.2, 0.3):
        d = np.array([0.0, dy, -math.sqrt(1 - dy * dy)])
        _, p32, _ = O.ray_sphere([0, 0, 0], d, c, R, 32)
        _, p64, _ = O.ray_sphere([0, 0, 0], d, c, R, 64)
        assert abs(np.linalg.norm(p64 - c) - R) < 1e-6
        errs.append(np.linalg.norm(p32 - p64))
    # FP32 loses tens of centimetres to metres near the horizon
    assert max(errs) > 0.3 and min(errs) > 0.01


# --- trilinear sampler vs the Vulkan formula, brute force --------------------
def _vulkan_trilinear(vol, n, s):
    """Vulkan spec 16.8: u = s*N - 0.5, weights (1-a)(1-b)(1-c) ..., REPEAT."""
    u = np.array(s) * n - 0.5
    i0 = np.floor(u).astype(int)
    a = u - np.floor(u)
    out = np.zeros(4)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (a[0] if dx else 1 - a[0]) * (a[1] if dy else 1 - a[1]) * (a[2] if dz else 1 - a[2])
                x, y, z = (i0[0] + dx) % n, (i0[1] + dy) % n, (i0[2] + dz) % n
                out += w * vol[z, y, x].astype(np.float64)
    return out / 255.0


def test_trilinear_matches_vulkan_formula_brute_force():
    rng = np.random.default_rng(11)
    n = 4
    vol = rng.integers(0, 256, size=(n, n, n, 4), dtype=np.uint8)
    flat = vol.reshape(-1)
    for _ in range(200):
        s = rng.uniform(-2.0, 3.0, size=3)
        exp = _vulkan_trilinear(vol, n, s)
        got64 = O.sample3d(flat, 0, n, s, 64)
        got32 = O.sample3d(flat, 0, n, s, 32)
        assert np.allclose(got64, exp, atol=1e-12)
        assert np.allclose(got32, exp, atol=2e-6)


def test_trilinear_texel_centres_are_exact_and_repeat_wraps():
    rng = np.random.default_rng(3)
    n = 8
    vol = rng.integers(0, 256, size=(n, n, n, 4), dtype=np.uint8)
    flat = vol.reshape(-1)
    k = np.float32(1.0 / 255.0)
    for (x, y, z) in [(0, 0, 0), (3, 5, 7), (7, 7, 7), (2, 0, 6)]:
        s = ((x + 0.5) / n, (y + 0.5) / n, (z + 0.5) / n)
        got = O.sample3d(flat, 0, n, s, 32)
        exp = vol[z, y, x].astype(np.float32) * k
        assert np.array_equal(got.astype(np.float32), exp)
        s2 = (s[0] + 1.0, s[1] - 2.0, s[2] + 3.0)  # REPEAT (exact shifts in binary)
        assert np.array_equal(O.sample3d(flat, 0, n, s2, 32), got)


def test_bilinear_2d_matches_formula():
    rng = np.random.default_rng(4)
    n = 8
    m = rng.integers(0, 256, size=(n, n, 4), dtype=np.uint8)
    for _ in range(100):
        sx, sy = rng.uniform(-1, 2, size=2)
        u = np.array([sx, sy]) * n - 0.5
        i0 = np.floor(u).astype(int)
        a = u - np.floor(u)
        exp = np.zeros(4)
        for dy in (0, 1):
            for dx in (0, 1):
                w = (a[0] if dx else 1 - a[0]) * (a[1] if dy else 1 - a[1])
                exp += w * m[(i0[1] + dy) % n, (i0[0] + dx) % n]
        assert np.allclose(O.sample2d(m.reshape(-1), n, sx, sy, 64), exp / 255.0, atol=1e-12)


# --- march invariants ---------------------------------------------------------
def _scene(W=64, H=48, pitch=20.0, preset=1):
    params, sun = M.preset_params(preset)
    cam = ground_camera(W, H, pitch)
    t = M.Time()
    t.advance(1.0 / 60.0)
    return params, sun, cam.ubo(), t.ubo


def test_empty_volume_gives_the_background_exactly():
    W, H = 48, 32
    params, sun, ubo, tm = _scene(W, H)
    v = volumes(16, 8, 16, 32)
    # r = 0 with fbm = 1 makes remapClamped(0, 0.1, 1, 0, 1) = 0 (:513), so
    # every sample has zero density (an all-zero volume does not: its base
    # is remap(0, -0.9, 1, 0, 1) = 0.47)
    lf = np.zeros_like(v.lf).reshape(-1, 4)
    lf[:, 1:] = 255
    empty = O.Volumes(lf.reshape(-1), np.zeros_like(v.hf), v.curl, v.weather, 16, 8, 16, 32)
    xy = np.array([[x, y] for y in range(H) for x in range(W)], dtype=np.int32)
    rgba, mask, steps, cnt = O.march_pixels(empty, ubo, tm, sun, params, W, H, xy, 32)
    assert cnt["cloud_steps"] == 0 and cnt["early_exits"] == 0
    # with zero density the march returns colour 0 and accum 0, so the pixel
    # is the background: mix(bg, 0, 0) == bg, and the mask encodes 25*0.05
    enc = O.encode_float_rgba(1.25, 32)
    marched = mask.sum(1) != 0
    assert marched.any()
    assert np.allclose(mask[marched], enc, atol=0)
    for i in np.nonzero(marched)[0][:40]:
        assert rgba[i, 3] == 1.0 and np.all(rgba[i, :3] > 0)


def test_ocean_colour_closed_form():
    W, H = 32, 32
    params, sun, ubo, tm = _scene(W, H, pitch=-60.0)
    v = volumes(16, 8, 16, 32)
    xy = np.array([[x, y] for y in range(H) for x in range(W)], dtype=np.int32)
    rgba, mask, steps, cnt = O.march_pixels(v, ubo, tm, sun, params, W, H, xy, 64)
    assert cnt["rays_ocean"] == W * H
    c1 = np.array([0.0, 0.16, 0.51]) * 0.4
    c2 = np.array([0.0, 0.73, 0.95]) * 0.5
    # every colour lies on the segment c1 -> c2 extended by a = -5.5*dir.y
    a = (rgba[:, 2] - c1[2]) / (c2[2] - c1[2])
    assert np.allclose(rgba[:, :3], c1[None] * (1 - a[:, None]) + c2[None] * a[:, None], atol=1e-12)
    # a = -5.5 * dir.y with dir.y in [-1, -sin(30 deg)] for this frustum
    assert np.all(a > 5.5 * np.sin(np.radians(30))) and np.all(a <= 5.5)
    assert np.all(mask == 0) and np.all(steps == 0)


def test_step_counts_and_accumulation_bounds():
    W, H = 96, 64
    params, sun, ubo, tm = _scene(W, H)
    v = volumes(32, 16, 32, 64)
    xy = O.select_pixels(W, H, tm.frame_count_mod16)
    rgba, mask, steps, cnt = O.march_pixels(v, ubo, tm, sun, params, W, H, xy, 32)
    n_steps = int(params.steps_max)
    marched = steps > 0
    # a ray that is not cut short by the early exit makes N or N+1 steps
    assert np.all(steps[marched] <= n_steps + 1)
    assert cnt["primary_steps"] == int(steps.sum())
    assert cnt["cone_samples"] == 6 * cnt["cloud_steps"]
    assert cnt["cone_erosions"] <= cnt["cone_samples"]
    # mask decodes to 25*min(0.05, 1-accum) in [0, 1.25]: accum in [0, 1]
    dec = mask @ (1.0 / np.array([1.0, 255.0, 65025.0, 16581375.0]))
    assert np.all(dec[marched] >= -1e-6) and np.all(dec[marched] <= 1.25 + 1e-6)


def test_fp32_and_fp64_twins_agree_on_a_frame():
    W, H = 96, 64
    params, sun, ubo, tm = _scene(W, H)
    v = volumes(32, 16, 32, 64)
    xy = np.array([[x, y] for y in range(0, H, 2) for x in range(0, W, 2)], dtype=np.int32)
    a, _, _, _ = O.march_pixels(v, ubo, tm, sun, params, W, H, xy, 32)
    b, _, _, _ = O.march_pixels(v, ubo, tm, sun, params, W, H, xy, 64)
    assert psnr(a[:, :3], b[:, :3], peak=np.percentile(b[:, :3], 99.9)) > 35.0


def test_reference_preset_marches_35_to_60_steps():
    W, H = 64, 64
    params, sun = M.reference_params()
    cam = M.Camera((0.0, 0.0, 2.0), (0.0, 0.6, 1.0), W, H, 45.0)
    t = M.Time()
    t.advance(1.0 / 60.0)
    v = volumes(32, 16, 32, 64)
    xy = np.array([[x, y] for y in range(0, H, 3) for x in range(0, W, 3)], dtype=np.int32)
    _, _, steps, cnt = O.march_pixels(v, cam.ubo(), t.ubo, sun, params, W, H, xy, 32)
    s = steps[steps > 0]
    assert s.size and s.max() <= 61 and cnt["early_exits"] + (s >= 35).sum() >= s.size


# --- 4x4 selection ------------------------------------------------------------
@pytest.mark.parametrize("W,H", [(16, 12), (19, 13), (480, 270)])
def test_4x4_selection_covers_every_pixel_once_in_16_frames(W, H):
    hits = np.zeros((H, W), dtype=np.int32)
    for f in range(16):
        xy = O.select_pixels(W, H, f)
        np.add.at(hits, (xy[:, 1], xy[:, 0]), 1)
        # the frame's phase: pixel = 4*gid + (f/4, f%4) (cloudRayMarch.comp:695-701)
        assert np.all(xy[:, 0] % 4 == f // 4) and np.all(xy[:, 1] % 4 == f % 4)
    assert np.all(hits == 1)


# --- reprojection -------------------------------------------------------------
def test_reprojection_identity_camera_reads_its_own_pixel():
    W, H = 64, 48
    params, sun, ubo, tm = _scene(W, H)
    for frame in range(16):
        tm.frame_count_mod16 = frame
        idx = O.reprojection_indices(ubo, ubo, tm, params, W, H, 32)
        own = (np.arange(H)[:, None] * W + np.arange(W)[None, :])
        # zero motion: every tap lands on round(x + jitter/2) = x (jitter/2 < 0.45)
        assert np.array_equal(idx, np.repeat(own[:, :, None], 10, axis=2)), frame


def test_reprojection_pure_yaw_shifts_horizontally():
    W, H = 160, 90
    params, sun = M.preset_params(2)
    cam = ground_camera(W, H, 0.0)
    old = cam.ubo()
    yaw = 2.0
    cam.rotate_about_up(yaw)
    cur = cam.ubo()
    t = M.Time()
    t.advance(1 / 60)
    idx = O.reprojection_indices(cur, old, t.ubo, params, W, H, 64)
    x0 = idx[:, :, 0] % W
    # tap 0 reads the old position of the same world point: shift ~ yaw/fov_x * W
    fov_x = 2 * math.degrees(math.atan(float(cur.tan_fov_by2[0])))
    shift = (x0 - np.arange(W)[None, :])[H // 2, 40:120]
    assert abs(np.median(shift) - (-yaw / fov_x * W)) < 1.5 or abs(np.median(shift) - (yaw / fov_x * W)) < 1.5
    assert abs(np.median(np.abs(shift)) - yaw / fov_x * W) < 1.5


def test_reprojection_averages_taps_of_the_previous_frame():
    W, H = 40, 30
    params, sun, ubo, tm = _scene(W, H)
    rng = np.random.default_rng(8)
    prev = rng.uniform(0, 2, size=(H, W, 4)).astype(np.float32)
    v = volumes(16, 8, 16, 32)
    cur, _, _ = O.frame(v, ubo, ubo, tm, sun, params, W, H, prev=prev, precision=32)
    sel = np.zeros((H, W), bool)
    xy = O.select_pixels(W, H, tm.frame_count_mod16)
    sel[xy[:, 1], xy[:, 0]] = True
    # identical cameras: the 10 taps all read the pixel itself
    assert np.allclose(cur[~sel], prev[~sel], rtol=1e-6)
