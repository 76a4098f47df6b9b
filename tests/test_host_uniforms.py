"""The C++ uniform producers (meteoros_amd/csrc/host_uniforms.cpp) against
closed forms: Scene::HaltonSequenceAt / UpdateTime (Scene.cpp:65-138) and
Camera::UpdateBuffer / RotateAboutUp (camera.cpp:31-96, glm lookAtRH and
perspectiveRH_ZO)."""
import math

import numpy as np
import pytest

import meteoros_amd as M


def _radical_inverse(i, b):
    r, f = 0.0, 1.0
    while i > 0:
        f /= b
        r += f * (i % b)
        i //= b
    return r


def test_halton_base3_closed_forms():
    expect = [1 / 3, 2 / 3, 1 / 9, 4 / 9, 7 / 9, 2 / 9, 5 / 9, 8 / 9, 1 / 27, 10 / 27, 19 / 27, 4 / 27, 13 / 27,
              22 / 27, 7 / 27, 16 / 27]
    for i, e in enumerate(expect, start=1):
        assert abs(M.halton(i, 3) - e) < 1e-7
        assert abs(M.halton(i, 3) - _radical_inverse(i, 3)) < 1e-7
    assert abs(M.halton(5, 2) - 0.625) < 1e-7


def test_time_table_and_frame_counter():
    t = M.Time()
    flat = list(t.ubo.halton1) + list(t.ubo.halton2) + list(t.ubo.halton3) + list(t.ubo.halton4)
    assert np.allclose(flat, [_radical_inverse(i, 3) for i in range(1, 17)], atol=1e-7)
    assert t.frame == 0  # InitializeTime sets 0 (Scene.cpp:116)
    frames = []
    for _ in range(40):
        t.advance(1.0 / 60.0)  # UpdateTime increments first (Scene.cpp:74-75)
        frames.append(t.frame)
    assert frames[:3] == [1, 2, 3] and frames[15] == 0 and frames[16] == 1
    assert abs(t.ubo.time[1] - 40 / 60.0) < 1e-5 and abs(t.ubo.time[0] - 1 / 60.0) < 1e-9


def _mat(ubo_field):
    return np.array(list(ubo_field), dtype=np.float64).reshape(4, 4).T  # row-major view of column-major storage


@pytest.mark.parametrize("eye,ref", [((0.0, 0.0, 2.0), (0.0, 0.0, 1.0)), ((0.0, 0.0, 0.0), (0.0, 0.34, -0.94)),
                                     ((10.0, 5.0, -3.0), (11.0, 5.5, -4.0))])
def test_view_matrix_is_lookat(eye, ref):
    W, H, fovy = 1284, 720, 45.0
    cam = M.Camera(eye, ref, W, H, fovy, 0.1, 1000.0)
    u = cam.ubo()
    V = _mat(u.view)
    R = V[:3, :3]
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-6)              # orthonormal basis
    e = np.array(eye)
    assert np.allclose(V @ np.append(e, 1.0), [0, 0, 0, 1], atol=1e-4)  # eye -> origin
    f = np.array(ref) - e
    f /= np.linalg.norm(f)
    assert np.allclose(-R[2], f, atol=1e-6)                         # camera looks down -z
    assert R[1, 1] > 0                                              # up keeps world-up's sign
    assert list(u.eye) == [eye[0], eye[1], eye[2], 1.0]


def test_projection_and_tan_fov():
    W, H, fovy, n, f = 1920, 1080, 45.0, 0.1, 1000.0
    u = M.Camera((0, 0, 0), (0, 0, -1), W, H, fovy, n, f).ubo()
    P = _mat(u.proj)
    t = math.tan(math.radians(fovy) / 2)
    a = W / H
    assert abs(P[0, 0] - 1 / (a * t)) < 1e-5
    assert abs(P[1, 1] + 1 / t) < 1e-5              # y flipped (camera.cpp:36)
    assert abs(P[2, 2] - f / (n - f)) < 1e-5
    assert P[3, 2] == -1.0
    assert abs(P[2, 3] + f * n / (f - n)) < 1e-5
    # tanFovBy2 uses PI = 3.14159 (camera.h:10, camera.cpp:40-41)
    ty = abs(math.tan(fovy * 0.5 * (3.14159 / 180.0)))
    assert abs(u.tan_fov_by2[1] - ty) < 1e-7
    assert abs(u.tan_fov_by2[0] - a * ty) < 1e-6


def test_rotate_about_up_yaws_by_the_given_angle():
    cam = M.Camera((0, 0, 0), (0, math.sin(math.radians(20)), -math.cos(math.radians(20))), 1920, 1080)
    f0 = np.array(cam.state.forward)
    up0 = np.array(cam.state.up)
    for _ in range(8):
        cam.rotate_about_up(0.5)
    f1 = np.array(cam.state.forward)
    # rotation about the camera's own up axis: the angle between forwards is 4 deg
    ang = math.degrees(math.acos(np.clip(np.dot(f0, f1), -1, 1)))
    assert abs(ang - 4.0) < 1e-3
    # the basis is re-derived from world up after each turn (RecomputeAttributes,
    # camera.cpp:68-77), so the pitch is kept up to a small drift
    assert abs(math.degrees(math.asin(f1[1])) - math.degrees(math.asin(f0[1]))) < 0.1
    assert np.dot(np.array(cam.state.up), up0) > 0.999
    assert abs(np.linalg.norm(np.array(cam.state.ref) - np.array(cam.state.eye)) - 1.0) < 1e-5


def test_translate_moves_eye_and_ref_together():
    cam = M.Camera((0, 0, 0), (0, 0, -1), 64, 64)
    cam.translate(look=2.0, right=1.0, up=0.5)
    assert np.allclose(cam.state.eye, [1.0, 0.5, -2.0], atol=1e-6)
    assert np.allclose(np.array(cam.state.ref) - np.array(cam.state.eye), [0, 0, -1], atol=1e-6)


def test_presets_and_wind_conversion():
    p, s = M.reference_params()
    assert p.shell_inner_radius == 6378500.0 and p.shell_outer_radius == 6391000.0
    assert p.steps_min == 35 and p.steps_max == 60 and p.light_samples == 6
    assert abs(s.sun_location[1] - 5751900.0) < 1 and abs(s.sun_location[2] + 5751900.0) < 1
    assert abs(s.sun_intensity - 0.78) < 1e-7
    for k, steps in zip(range(1, 6), (64, 128, 192, 64, 128)):
        q, sun = M.preset_params(k)
        assert q.steps_max == steps and q.shell_inner_radius == 6371000.0 + 1500.0
        # 10 m/s over one sample unit = 8 x 2500 m
        assert abs(q.cloud_speed - 10.0 / (2500.0 * 8.0)) < 1e-9
        d = np.array(sun.sun_location[:3], dtype=np.float64)
        assert abs(math.degrees(math.asin(d[1] / np.linalg.norm(d))) - 30.0) < 1e-4
