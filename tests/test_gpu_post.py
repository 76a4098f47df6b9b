"""The post-process passes on the GPU against the CPU oracle
(oracle/post_oracle.hpp): god rays over a marched frame's result and mask,
tone map + dither in RGBA32F (before quantisation) and RGBA8, and TXAA over
two frames of a moving camera (RGBA8, same one-code bar).

Bars: god rays within 1e-5 of the frame's peak (the kernel decodes the mask
before filtering, the shader after: the same linear map, FP32 rounding
apart); tone map within 4e-6 absolute before quantisation (pow ulps) and
RGBA8 bytes equal except where that rounding crosses a code boundary
(at most one code, at most 0.1% of values)."""
import numpy as np
import pytest

import meteoros_amd as M
from meteoros_amd import _abi as A
from oracle import oracle as O
from tests.common import ground_camera, volumes

pytestmark = pytest.mark.gpu

DIMS = (128, 32, 128, 512)


def _frame(W, H, pitch, yaw=0.0, frame=3):
    params, sun = M.preset_params(2)
    for k in range(3):
        sun.light_color[k] = (1.0, 0.95, 0.85)[k]
    cam = ground_camera(W, H, pitch, yaw).ubo()
    t = M.Time()
    t.ubo.frame_count_mod16 = frame
    t.ubo.time[1] = 12.7
    st = M.CloudStage(W, H, M.VolumeSpec(*DIMS), kernel=A.MTC_KERNEL_FAST)
    v = volumes(*DIMS)
    st.set_volumes(v.lf, v.hf, v.curl, v.weather)
    st.march_full(cam, t.ubo, sun, params)
    return st, cam, t.ubo, sun


@pytest.mark.parametrize("pitch,yaw", [(60.0, 0.0), (35.0, 20.0), (-20.0, 0.0)])
def test_god_rays_match_oracle(pitch, yaw):
    W, H = 480, 270
    st, cam, tm, sun = _frame(W, H, pitch, yaw)
    with st:
        rgba, mask = st.read_output()
        st.god_rays(cam, sun)
        got, _ = st.read_output()
    ref, (sx, sy, blend) = O.god_rays(cam, sun, rgba, mask, precision=32)
    peak = float(np.abs(ref).max())
    err = np.abs(got.astype(np.float64) - ref)
    print(f"pitch {pitch} yaw {yaw}: sun_ss ({sx:.3f}, {sy:.3f}) blend {blend:.3f}; max err {err.max():.2e} "
          f"(peak {peak:.2f}), god-ray energy added {float((ref - rgba)[..., :3].sum()):.1f}")
    if blend < 0:
        assert np.array_equal(got, rgba)
    else:
        assert err.max() <= 1e-5 * peak
        assert float((got - rgba)[..., :3].sum()) > 0.0


@pytest.mark.parametrize("as_float", [True, False])
def test_tone_map_matches_oracle(as_float):
    W, H = 480, 270
    st, cam, tm, sun = _frame(W, H, 20.0)
    with st:
        rgba, _ = st.read_output()
        st.tone_map(tm, as_float=as_float)
        got = st.read_display()
    rgb, rgba8 = O.tone_map(int(tm.time[1]), rgba, precision=32)
    if as_float:
        err = np.abs(got[..., :3].astype(np.float64) - rgb)
        print(f"tone map (float): max err {err.max():.2e}")
        assert err.max() <= 4e-6
        assert np.all(got[..., 3] == 1.0)
    else:
        d = np.abs(got.astype(np.int16) - rgba8.astype(np.int16))
        print(f"tone map (RGBA8): {int((d > 0).sum())} of {d.size} bytes differ, max {int(d.max())}")
        assert d.max() <= 1 and (d > 0).mean() <= 1e-3
        assert np.all(got[..., 3] == 255)


@pytest.mark.parametrize("W,H,pitch", [(65, 37, 85.0), (1, 1, 75.0), (130, 17, 60.0), (200, 300, 89.0),
                                       (2560, 1440, 70.0)])
def test_god_rays_odd_sizes_match_oracle(W, H, pitch):
    """The blur's workgroups cover 64 columns x 16 rows: partial waves (65,
    130, 200 columns), a 1x1 image, partial row groups (37, 17, 300 rows) and
    the sun inside the image (pitch 75-89 deg puts it at 19-48% of the
    height) must all match the oracle, and so must an image larger than
    1080p (2560 x 1440: 3 600 workgroups, more than the device holds at
    once)."""
    st, cam, tm, sun = _frame(W, H, pitch)
    with st:
        rgba, mask = st.read_output()
        st.god_rays(cam, sun)
        got, _ = st.read_output()
    ref, (sx, sy, blend) = O.god_rays(cam, sun, rgba, mask, precision=32)
    peak = float(np.abs(ref).max())
    err = np.abs(got.astype(np.float64) - ref)
    print(f"{W}x{H} pitch {pitch}: sun_ss ({sx:.3f}, {sy:.3f}) blend {blend:.3f}; max err {err.max():.2e} "
          f"(peak {peak:.2f})")
    assert blend > 0.0
    assert err.max() <= 1e-5 * peak


def test_god_rays_then_tone_map_at_1080p():
    W, H = 1920, 1080
    st, cam, tm, sun = _frame(W, H, 55.0)
    with st:
        rgba, mask = st.read_output()
        st.god_rays(cam, sun)
        st.tone_map(tm)
        out = st.read_display()
        st.tone_map(tm, as_float=True)  # a format change re-allocates the display image
        outf = st.read_display()
    ref, _ = O.god_rays(cam, sun, rgba, mask, precision=32)
    rgb, rgba8 = O.tone_map(int(tm.time[1]), ref, precision=32)
    d = np.abs(out.astype(np.int16) - rgba8.astype(np.int16))
    assert d.max() <= 1 and (d > 0).mean() <= 1e-3
    assert np.abs(outf[..., :3] - rgb).max() <= 1e-4


@pytest.mark.parametrize("yaw", [0.0, 0.5, 3.0])
def test_txaa_matches_oracle(yaw):
    """Two frames: tone map -> TXAA (black history), camera yaws, tone map ->
    TXAA against the first output; each TXAA output against the FP32 oracle."""
    W, H = 480, 270
    params, sun = M.preset_params(2)
    camera = ground_camera(W, H, 25.0)
    v = volumes(*DIMS)
    with M.CloudStage(W, H, M.VolumeSpec(*DIMS), kernel=A.MTC_KERNEL_FAST) as st:
        st.set_volumes(v.lf, v.hf, v.curl, v.weather)
        history = np.zeros((H, W, 4), np.uint8)
        old = camera.ubo()
        for k in range(2):
            camera.rotate_about_up(yaw)
            cur = camera.ubo()
            t = M.Time()
            t.ubo.frame_count_mod16 = 4 + k
            t.ubo.time[1] = 2.0 + k / 60.0
            st.march_full(cur, t.ubo, sun, params)
            st.tone_map(t.ubo)
            disp = st.read_display()
            st.txaa(cur, old, t.ubo)
            got = st.read_txaa()
            _, ref8, _ = O.txaa(cur, old, t.ubo, disp, history, precision=32)
            d = np.abs(got.astype(np.int16) - ref8.astype(np.int16))
            print(f"yaw {yaw} frame {k}: {int((d > 0).sum())} of {d.size} bytes differ, max {int(d.max())}")
            assert d.max() <= 1 and (d > 0).mean() <= 1e-3
            history = got
            old = cur


@pytest.mark.parametrize("W,H", [(65, 37), (130, 17), (3, 2), (200, 300)])
def test_tone_map_and_txaa_odd_sizes_match_oracle(W, H):
    """Both passes read neighbouring texels without bounds tests in waves
    (one pixel row of 64 columns) whose row, columns and neighbours lie
    inside the image, and with them elsewhere: 65 x 37 and 3 x 2 have only
    bounds-tested waves, 130 x 17 and 200 x 300 both kinds, with partial
    waves at the right edge."""
    params, sun = M.preset_params(2)
    camera = ground_camera(W, H, 25.0)
    v = volumes(*DIMS)
    with M.CloudStage(W, H, M.VolumeSpec(*DIMS), kernel=A.MTC_KERNEL_FAST) as st:
        st.set_volumes(v.lf, v.hf, v.curl, v.weather)
        history = np.zeros((H, W, 4), np.uint8)
        old = camera.ubo()
        for k in range(2):
            camera.rotate_about_up(0.5)
            cur = camera.ubo()
            t = M.Time()
            t.ubo.frame_count_mod16 = 4 + k
            t.ubo.time[1] = 2.0 + k / 60.0
            st.march_full(cur, t.ubo, sun, params)
            rgba, _ = st.read_output()
            st.tone_map(t.ubo)
            disp = st.read_display()
            _, disp_ref = O.tone_map(int(t.ubo.time[1]), rgba, precision=32)
            d = np.abs(disp.astype(np.int16) - disp_ref.astype(np.int16))
            print(f"{W}x{H} frame {k}: tone map {int((d > 0).sum())} of {d.size} bytes differ", end="")
            assert d.max() <= 1 and (d > 0).sum() <= max(1, d.size // 1000)
            st.txaa(cur, old, t.ubo)
            got = st.read_txaa()
            _, ref8, _ = O.txaa(cur, old, t.ubo, disp, history, precision=32)
            d = np.abs(got.astype(np.int16) - ref8.astype(np.int16))
            print(f"; TXAA {int((d > 0).sum())} differ, max {int(d.max())}")
            assert d.max() <= 1 and (d > 0).sum() <= max(1, d.size // 1000)
            history = got
            old = cur


def test_post_error_paths():
    with M.CloudStage(32, 16, M.VolumeSpec(32, 16, 32, 64)) as st:
        with pytest.raises(M.MtcError):
            st.read_display()
        t = M.Time().ubo
        t.time[1] = float("nan")
        with pytest.raises(M.MtcError):
            st.tone_map(t)
        assert M.lib().mtc_tone_map(st.handle, M.Time().ubo, 7, None) == A.MTC_ERR_INVALID_ARG
        cam = ground_camera(32, 16).ubo()
        with pytest.raises(M.MtcError):
            st.txaa(cam, cam, M.Time().ubo)  # no RGBA8 tone-mapped image yet
        with pytest.raises(M.MtcError):
            st.read_txaa()
