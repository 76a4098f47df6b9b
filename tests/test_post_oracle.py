"""Pins of the post-process oracle (oracle/post_oracle.hpp): god rays
(postProcess_GodRays.frag) and tone map + dither (postProcess_ToneMap.frag).

The reference has no tests for either pass; the pins are closed forms:
the folded sun position and blend factor for a camera whose basis is known,
the accumulated god-ray colour for a constant and for a linear decoded mask
(bilinear filtering reproduces linear functions, so the 100-sample walk sums
an arithmetic series), the Uncharted-2 curve's closed values and white
point, the dither's range and mean, and the Wang hash's invertibility (each
of its steps undone by an independently derived inverse).
"""
import math

import numpy as np
import pytest

import meteoros_amd as M
from oracle import oracle as O

ENC = np.array([1.0, 255.0, 65025.0, 16581375.0])


def _sun(light=(1.0, 0.9, 0.8)):
    _, sun = M.preset_params(2)
    for k in range(3):
        sun.light_color[k] = light[k]
    sun.light_color[3] = 1.0
    return sun


def _camera(W, H, pitch_deg, yaw_deg=0.0, fovy=45.0):
    p, y = math.radians(pitch_deg), math.radians(yaw_deg)
    d = (math.cos(p) * math.sin(y), math.sin(p), -math.cos(p) * math.cos(y))
    return M.Camera((0.0, 0.0, 0.0), d, W, H, fovy).ubo()


def _mask_from_decoded(d):
    """RGBA mask whose decode dot(x, 1/ENC) is d (all of it in the r channel)."""
    m = np.zeros(d.shape + (4,), np.float32)
    m[..., 0] = d
    return m


def test_sun_position_and_blend_closed_form():
    W, H = 64, 48
    for pitch in (70.0, 80.0, 85.0):
        cam = _camera(W, H, pitch)
        _, (sx, sy, blend) = O.god_rays(cam, _sun(), np.zeros((H, W, 4), np.float32),
                                        np.zeros((H, W, 4), np.float32), precision=64)
        # view-space sun (0, cos p, -sin p); clip.y = -cos p / tan(fovy/2)
        # (y-flipped projection), no perspective divide before the clamp
        t = math.tan(math.radians(22.5))
        assert abs(sx - 0.5) < 1e-6
        assert abs(sy - (1.0 - math.cos(math.radians(pitch)) / t) / 2.0) < 2e-5
        assert abs(blend - math.sin(math.radians(pitch))) < 2e-5


def test_sun_behind_camera_leaves_the_image_unchanged():
    W, H = 32, 24
    cam = _camera(W, H, -30.0)  # looking down: the sun at (0, 1, 0) is behind
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 3, (H, W, 4)).astype(np.float32)
    out, (_, _, blend) = O.god_rays(cam, _sun(), img, _mask_from_decoded(np.full((H, W), 0.7)))
    assert blend < 0
    assert np.array_equal(out, img)


@pytest.mark.parametrize("precision", [32, 64])
def test_god_rays_constant_mask_closed_form(precision):
    W, H = 80, 60
    cam = _camera(W, H, 80.0)
    light = (1.0, 0.9, 0.8)
    img = np.zeros((H, W, 4), np.float32)
    out, (_, _, blend) = O.god_rays(cam, _sun(light), img, _mask_from_decoded(np.full((H, W), 0.6)), precision)
    # every sample decodes to 0.6: accum = 100 * light * 0.6 * 0.001
    for k in range(3):
        assert np.allclose(out[..., k], 0.06 * light[k] * blend, rtol=1e-5)
    assert np.allclose(out[..., 3], blend, rtol=1e-6)


def test_god_rays_linear_mask_is_an_arithmetic_series():
    W, H = 96, 64
    cam = _camera(W, H, 80.0)
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    alpha, beta, gamma = 0.3, 0.004, -0.002
    d = alpha + beta * xs + gamma * ys  # decoded value of texel (x, y)
    out, (sx, sy, blend) = O.god_rays(cam, _sun((1.0, 1.0, 1.0)), np.zeros((H, W, 4), np.float32),
                                      _mask_from_decoded(d.astype(np.float32)), precision=64)
    # sample i at uv_i = uv - i (uv - sun) / 100, texel coordinate x = u W - 0.5
    px, py = np.meshgrid(np.arange(W), np.arange(H))
    u0, v0 = (px + 0.5) / W, (py + 0.5) / H
    i = np.arange(100)
    ui = u0[..., None] - i * (u0[..., None] - sx) / 100.0
    vi = v0[..., None] - i * (v0[..., None] - sy) / 100.0
    val = alpha + beta * (ui * W - 0.5) + gamma * (vi * H - 0.5)
    expect = 0.001 * val.sum(-1) * blend
    # samples within half a texel of the border blend in black border texels
    inside = ((ui * W - 0.5 >= 0) & (ui * W - 0.5 <= W - 1) & (vi * H - 0.5 >= 0) & (vi * H - 0.5 <= H - 1)).all(-1)
    assert inside.mean() > 0.8
    # the image is stored in FP32
    assert np.allclose(out[..., 0][inside], expect[inside], rtol=2e-6, atol=1e-9)
    # and the twins agree
    out32, _ = O.god_rays(cam, _sun((1.0, 1.0, 1.0)), np.zeros((H, W, 4), np.float32),
                          _mask_from_decoded(d.astype(np.float32)), precision=32)
    assert np.allclose(out32, out, rtol=1e-4, atol=1e-6)


def test_mask_decode_matches_the_march_encoding():
    # decode(EncodeFloatRGBA(v)) = fract(v): the god-ray pass reads what the
    # march wrote (the march's 25 * min(0.05, 1 - accum) reaches 1.25 in clear
    # sky, which the encoding wraps to 0.25)
    W, H = 16, 12
    cam = _camera(W, H, 80.0)
    for v in (0.0, 0.25, 0.5, 0.7, 0.999, 1.25):
        v_dec = v - math.floor(v)
        enc = np.tile(np.array(O.encode_float_rgba(v, 64), np.float32), (H, W, 1))
        out, (_, _, blend) = O.god_rays(cam, _sun((1.0, 1.0, 1.0)), np.zeros((H, W, 4), np.float32), enc, 64)
        assert np.allclose(out[..., 0], 0.1 * v_dec * blend, rtol=1e-6, atol=1e-9)


def test_uncharted2_closed_forms():
    # U2(x) = ((x(Ax+CB)+DE)/(x(Ax+B)+DF)) - E/F with A..F = .15 .5 .1 .2 .02 .3
    u2 = lambda x: (x * (0.15 * x + 0.05) + 0.004) / (x * (0.15 * x + 0.5) + 0.06) - 0.02 / 0.3  # noqa: E731
    for x in (0.0, 0.5, 2.5, 10.0, 100.0):
        assert abs(O.uncharted2(x) - u2(x)) < 2e-6
    assert abs(O.uncharted2(0.0)) < 1e-7
    assert abs(O.uncharted2(1e6) - (1.0 - 0.02 / 0.3)) < 1e-5
    xs = np.linspace(0, 200, 401)
    assert np.all(np.diff([O.uncharted2(float(x)) for x in xs]) > 0)


def test_tone_map_white_point_dither_and_black():
    W, H = 64, 64
    # 100 / exposure maps to exactly the white point: pow(1, 1/2.2) = 1
    img = np.full((H, W, 4), 40.0, np.float32)
    rgb, rgba8 = O.tone_map(7, img, precision=64)
    noise = rgb[..., 0] - 1.0
    assert np.all(noise >= -1e-6) and np.all(noise < 0.01 + 1e-6)
    assert abs(noise.mean() - 0.005) < 4e-4          # uniform dither in [0, 0.01)
    assert np.allclose(rgb[..., 0], rgb[..., 1]) and np.allclose(rgb[..., 1], rgb[..., 2])
    assert np.all(rgba8[..., :3] == 255) and np.all(rgba8[..., 3] == 255)
    # black: U2(0) is 0 up to rounding, so the output is the dither
    rgb0, _ = O.tone_map(7, np.zeros((H, W, 4), np.float32), precision=64)
    assert np.all(np.abs(rgb0[..., 0] - noise) < 1e-3)
    # a different second gives a different dither pattern, the same mean
    rgb2, _ = O.tone_map(8, img, precision=64)
    assert not np.allclose(rgb2, rgb)
    # FP32 twin
    rgb32, _ = O.tone_map(7, img, precision=32)
    assert np.allclose(rgb32, rgb, atol=2e-6)


def _inv_xorshift_right(y, s):
    x = y
    for _ in range(32 // s + 1):
        x = y ^ (x >> s)
    return x & 0xFFFFFFFF


def test_wang_hash_is_invertible_step_by_step():
    inv9 = pow(9, -1, 1 << 32)
    invk = pow(0x27D4EB2D, -1, 1 << 32)
    rng = np.random.default_rng(3)
    for _ in range(2000):
        u, v, s = (int(a) for a in rng.integers(0, 1 << 32, 3, dtype=np.uint64))
        h = O.wang_hash(u, v, s)
        x = _inv_xorshift_right(h, 15)
        x = (x * invk) & 0xFFFFFFFF
        x = _inv_xorshift_right(x, 4)
        x = (x * inv9) & 0xFFFFFFFF
        hi = x >> 16
        lo = (x & 0xFFFF) ^ 61 ^ hi
        seed = (hi << 16) | lo
        assert seed == ((u * 1664525 + v) + s) & 0xFFFFFFFF
    # uniform over 16 bins (chi-square, 15 dof, p ~ 1e-4 bound)
    hs = np.array([O.wang_hash(x, y, 5) for y in range(64) for x in range(64)], dtype=np.uint64)
    counts = np.bincount((hs >> np.uint64(28)).astype(np.int64), minlength=16)
    exp = len(hs) / 16
    assert ((counts - exp) ** 2 / exp).sum() < 45


# --- TXAA (postProcess_TXAA.frag) ---------------------------------------------
def _time_frame(frame):
    t = M.Time()
    t.ubo.frame_count_mod16 = frame
    return t.ubo


def _jitter(tm, frame, W, H):
    idx = frame // 2
    jx = tm.halton1[idx] if idx < 4 else tm.halton3[idx - 4]
    jy = tm.halton2[idx] if idx < 4 else tm.halton4[idx - 4]
    return jx / W, jy / H


@pytest.mark.parametrize("frame", [0, 5, 11])
def test_txaa_reprojection_with_a_still_camera_is_the_jitter(frame):
    # same camera: the hit point projects back along its own jittered ray,
    # so old_uv = uv + jitter / 2 (NDC = 2 uv - 1 + jitter)
    W, H = 48, 32
    cam = _camera(W, H, 25.0, 10.0)
    tm = _time_frame(frame)
    img = np.zeros((H, W, 4), np.uint8)
    _, _, uv = O.txaa(cam, cam, tm, img, img, precision=64)
    jx, jy = _jitter(tm, frame, W, H)
    px, py = np.meshgrid(np.arange(W), np.arange(H))
    assert np.allclose(uv[..., 0], (px + 0.5) / W + jx / 2, atol=1e-9)
    assert np.allclose(uv[..., 1], (py + 0.5) / H + jy / 2, atol=1e-9)


def test_txaa_constant_frame_is_a_fixed_point():
    # a constant current frame collapses the clip box onto it: the output is
    # the current colour whatever the history holds (interior pixels)
    W, H = 24, 16
    cam = _camera(W, H, 30.0)
    cur = np.zeros((H, W, 4), np.uint8)
    cur[...] = (90, 140, 200, 255)
    rng = np.random.default_rng(5)
    prev = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    _, out8, _ = O.txaa(cam, cam, _time_frame(3), cur, prev, precision=64)
    assert np.array_equal(out8[1:-1, 1:-1], cur[1:-1, 1:-1])


def test_txaa_feedback_weight_closed_form():
    # checkerboard current frame (box = [0, 1] per channel, so no clipping)
    # and a constant history P: out = mix(P, cur, 0.5 (1 - |l0 - l1| / max(l0, l1, 0.2))^2)
    W, H = 20, 14
    cam = _camera(W, H, 30.0)
    ys, xs = np.mgrid[0:H, 0:W]
    cur = np.zeros((H, W, 4), np.uint8)
    cur[(xs + ys) % 2 == 0] = (255, 255, 255, 255)
    cur[(xs + ys) % 2 == 1] = (0, 0, 0, 255)
    prev = np.zeros((H, W, 4), np.uint8)
    prev[...] = (100, 120, 140, 255)
    rgba, _, _ = O.txaa(cam, cam, _time_frame(0), cur, prev, precision=64)
    c = cur.astype(np.float64) / 255.0
    p = prev.astype(np.float64) / 255.0
    lum = lambda a: a[..., 0] * 0.2125 + a[..., 1] * 0.7154 + a[..., 2] * 0.0721  # noqa: E731
    l0, l1 = lum(c), lum(p)
    k = 0.5 * (1.0 - np.abs(l0 - l1) / np.maximum(l0, np.maximum(l1, 0.2))) ** 2
    exp = p * (1 - k[..., None]) + c * k[..., None]
    inner = (slice(2, -2), slice(2, -2))
    assert np.allclose(rgba[inner], exp[inner], atol=1e-6)


def test_txaa_twins_agree():
    W, H = 40, 30
    cam = _camera(W, H, 25.0)
    c2 = M.Camera((0.0, 0.0, 0.0), (0.05, 0.42, -0.9), W, H, 45.0)
    rng = np.random.default_rng(8)
    cur = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    prev = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    a, a8, _ = O.txaa(cam, c2.ubo(), _time_frame(7), cur, prev, precision=64)
    b, b8, _ = O.txaa(cam, c2.ubo(), _time_frame(7), cur, prev, precision=32)
    # the FP32 shell hit (radius 6.4e6 m) moves the reprojected uv by a
    # fraction of a texel; on a random history that is visible
    assert np.abs(a - b).max() < 5e-3
    assert np.abs(a8.astype(int) - b8.astype(int)).max() <= 2
