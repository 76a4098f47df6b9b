"""The C-ABI library loads, exports every entry point of
include/meteoros_cloud.h, keeps the std140 uniform layouts of the reference
(camera.h:12-18, Scene.h:12-21, Sky.h:9-15), and rejects bad input on the
host before any device work."""
import ctypes as C
import os
import re

import pytest

import meteoros_amd as M
from meteoros_amd import _abi as A

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "meteoros_cloud.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mtc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = M.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {s[0] for s in A.SIGNATURES}
    assert set(names) <= bound, set(names) - bound
    assert lib.mtc_missing_symbols == []


def test_header_constants_match_the_bindings():
    """Every integer #define of the header that the Python bindings restate
    carries the same value there."""
    with open(HEADER) as f:
        text = f.read()
    defines = dict(re.findall(r"^#define\s+(MTC_[A-Z0-9_]+)\s+(-?\d+)\s*$", text, flags=re.M))
    assert "MTC_GOD_RAYS_NOALLOC_SCRATCH_BYTES" in defines
    shared = [k for k in defines if hasattr(A, k)]
    assert len(shared) >= 4
    for k in shared:
        assert getattr(A, k) == int(defines[k]), k


def test_std140_layouts():
    assert C.sizeof(A.MtcCamera) == 152
    assert A.MtcCamera.proj.offset == 64 and A.MtcCamera.eye.offset == 128 and A.MtcCamera.tan_fov_by2.offset == 144
    assert C.sizeof(A.MtcTime) == 76
    assert A.MtcTime.halton2.offset == 16 and A.MtcTime.halton4.offset == 48
    assert A.MtcTime.time.offset == 64 and A.MtcTime.frame_count_mod16.offset == 72
    assert C.sizeof(A.MtcSun) == 52 and A.MtcSun.sun_intensity.offset == 48


def test_version_and_device_count_never_fail():
    assert b"gfx950" in M.lib().mtc_version()
    assert M.device_count() >= 0


def _desc(**kw):
    d = A.MtcDesc()
    d.width, d.height = kw.get("width", 64), kw.get("height", 32)
    d.volumes = A.MtcVolumeDesc(*kw.get("vol", (32, 16, 32, 64, A.MTC_FMT_UNORM8)))
    d.kernel = kw.get("kernel", A.MTC_KERNEL_PLAIN)
    d.flags = kw.get("flags", 0)
    d.lanes_per_ray = kw.get("lanes", 0)
    return d


@pytest.mark.parametrize("kw", [
    dict(width=0), dict(height=-3), dict(width=20000),
    dict(vol=(30, 16, 32, 64, 0)), dict(vol=(32, 16, 32, 64, 7)), dict(kernel=17), dict(flags=64),
    dict(vol=(2048, 16, 32, 64, 0)), dict(vol=(1024, 16, 32, 64, 0), kernel=A.MTC_KERNEL_FAST),
    dict(vol=(32, 1024, 32, 64, 1), kernel=A.MTC_KERNEL_PRECISION),
    dict(flags=8), dict(flags=8, kernel=A.MTC_KERNEL_LEARNED), dict(flags=8, kernel=A.MTC_KERNEL_FAST, lanes=4),
    dict(flags=8, kernel=A.MTC_KERNEL_PRECISION, vol=(4, 16, 32, 64, 0)),
])
def test_create_rejects_invalid_descriptions(kw):
    ctx = C.c_void_p()
    rc = M.lib().mtc_create(C.byref(_desc(**kw)), C.byref(ctx))
    assert rc == A.MTC_ERR_INVALID_ARG
    assert not ctx
    assert M.lib().mtc_last_error()


def test_create_without_device_is_an_error_not_a_crash():
    if M.device_count() > 0:
        pytest.skip("a device is present")
    ctx = C.c_void_p()
    rc = M.lib().mtc_create(C.byref(_desc()), C.byref(ctx))
    assert rc == A.MTC_ERR_HIP
    assert b"device" in M.lib().mtc_last_error()


def test_null_context_calls_are_rejected():
    lib = M.lib()
    assert lib.mtc_synchronize(None) == A.MTC_ERR_INVALID_ARG
    assert lib.mtc_set_timing_stride(None, 4) == A.MTC_ERR_INVALID_ARG
    assert lib.mtc_read_output(None, None, None, 0) == A.MTC_ERR_INVALID_ARG
    assert lib.mtc_frame(None, None, None, None, None, None, None) == A.MTC_ERR_INVALID_ARG
    assert lib.mtc_destroy(None) == A.MTC_OK


def test_camera_helpers_validate():
    lib = M.lib()
    st = A.MtcCameraState()
    e = (C.c_float * 3)(0, 0, 0)
    assert lib.mtc_camera_init(C.byref(st), e, e, 64, 64, 45.0, 0.1, 100.0) == A.MTC_ERR_INVALID_ARG
    r = (C.c_float * 3)(0, 0, -1)
    assert lib.mtc_camera_init(C.byref(st), e, r, 64, 64, 190.0, 0.1, 100.0) == A.MTC_ERR_INVALID_ARG
    assert lib.mtc_camera_init(C.byref(st), e, r, 64, 64, 45.0, 0.1, 100.0) == A.MTC_OK
    t = A.MtcTime()
    assert lib.mtc_time_init(C.byref(t)) == A.MTC_OK
    assert lib.mtc_time_advance(C.byref(t), float("nan")) == A.MTC_ERR_INVALID_ARG
    assert lib.mtc_time_advance(C.byref(t), -1.0) == A.MTC_ERR_INVALID_ARG
    p, s = A.MtcCloudParams(), A.MtcSun()
    assert lib.mtc_params_preset(0, C.byref(p), C.byref(s)) == A.MTC_ERR_INVALID_ARG
    assert lib.mtc_params_preset(6, C.byref(p), C.byref(s)) == A.MTC_ERR_INVALID_ARG
    for k in range(1, 6):
        assert lib.mtc_params_preset(k, C.byref(p), C.byref(s)) == A.MTC_OK


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(A, "_lib", None)
    monkeypatch.setattr(A, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(ImportError):
        A.lib()


def test_learned_light_structs_and_null_handling():
    assert C.sizeof(A.MtcCounters) == 14 * 8
    assert C.sizeof(A.MtcMlpTrainDesc) == 40 and A.MtcMlpTrainDesc.holdout.offset == 32
    assert C.sizeof(A.MtcMlpCalibration) == 156 and A.MtcMlpCalibration.mask.offset == 136
    lib = M.lib()
    assert lib.mtc_mlp_destroy(None) == A.MTC_OK
    assert lib.mtc_mlp_init(None, 1) == A.MTC_ERR_INVALID_ARG
    assert lib.mtc_mlp_predict(None, None, 0, None) == A.MTC_ERR_INVALID_ARG
    assert lib.mtc_mlp_train(None, None, None) == A.MTC_ERR_INVALID_ARG
    assert lib.mtc_set_learned_light(None, None, 0) == A.MTC_ERR_INVALID_ARG
    m = C.c_void_p()
    assert lib.mtc_mlp_create(0, 0, C.byref(m)) == A.MTC_ERR_INVALID_ARG and not m
    if M.device_count() == 0:
        assert lib.mtc_mlp_create(0, 1024, C.byref(m)) == A.MTC_ERR_HIP and not m


def test_god_rays_noalloc_rejects_bad_input_on_the_host():
    """mtc_god_rays_noalloc validates everything before touching a device:
    null pointers, non-positive sizes, sizes beyond the limits that keep
    its decoded mask and blur table inside the static scratch, and
    non-finite uniforms.  The pointers passed with
    bad sizes are never dereferenced."""
    lib = M.lib()
    cam, sun = A.MtcCamera(), A.MtcSun()
    for k in range(16):
        cam.view[k] = cam.proj[k] = 1.0 if k % 5 == 0 else 0.0
    dummy = C.c_void_p(256)
    f = lib.mtc_god_rays_noalloc
    assert f(64, 32, None, C.byref(sun), dummy, dummy, None) == A.MTC_ERR_INVALID_ARG
    assert f(64, 32, C.byref(cam), C.byref(sun), None, dummy, None) == A.MTC_ERR_INVALID_ARG
    assert f(0, 32, C.byref(cam), C.byref(sun), dummy, dummy, None) == A.MTC_ERR_INVALID_ARG
    assert f(64, -1, C.byref(cam), C.byref(sun), dummy, dummy, None) == A.MTC_ERR_INVALID_ARG
    limit = A.MTC_GOD_RAYS_NOALLOC_MAX_ENTRIES
    assert f(2047, limit // 2048, C.byref(cam), C.byref(sun), dummy, dummy, None) == A.MTC_ERR_INVALID_ARG
    assert b"MTC_GOD_RAYS_NOALLOC_MAX_ENTRIES" in lib.mtc_last_error()
    # few entries but more rows than the blur's row table holds
    tall = A.MTC_GOD_RAYS_NOALLOC_MAX_HEIGHT + 1
    assert 2 * (tall + 1) <= limit
    assert f(1, tall, C.byref(cam), C.byref(sun), dummy, dummy, None) == A.MTC_ERR_INVALID_ARG
    assert b"MTC_GOD_RAYS_NOALLOC_MAX_HEIGHT" in lib.mtc_last_error()
    wide = A.MTC_GOD_RAYS_NOALLOC_MAX_WIDTH + 1
    assert (wide + 1) * 2 <= limit
    assert f(wide, 1, C.byref(cam), C.byref(sun), dummy, dummy, None) == A.MTC_ERR_INVALID_ARG
    assert b"MTC_GOD_RAYS_NOALLOC_MAX_WIDTH" in lib.mtc_last_error()
    assert f(3840, 2160, C.byref(cam), C.byref(sun), dummy, dummy, None) == A.MTC_ERR_INVALID_ARG
    cam.eye[1] = float("nan")
    assert f(64, 32, C.byref(cam), C.byref(sun), dummy, dummy, None) == A.MTC_ERR_INVALID_ARG
