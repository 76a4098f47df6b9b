"""Parity of the HIP kernels (through the C ABI) with the CPU oracle.

Bars: noise volumes, 4x4 selection, step counts and reprojection tap
indices bit-exact; colours and god-ray mask of the plain FP32 kernel within
1e-3 relative per value (floor 1e-3 of the frame's peak).
"""
import numpy as np
import pytest

import meteoros_amd as M
from meteoros_amd import _abi as A
from oracle import oracle as O
from tests.common import ground_camera, volumes, within_rel

pytestmark = pytest.mark.gpu

SMALL = (32, 16, 32, 64)
FULL = (128, 32, 128, 512)


def _stage(W, H, dims=FULL, fmt=A.MTC_FMT_UNORM8, flags=0, kernel=A.MTC_KERNEL_PLAIN, vol=None):
    st = M.CloudStage(W, H, M.VolumeSpec(*dims, fmt), kernel=kernel, flags=flags)
    v = vol if vol is not None else volumes(*dims, fmt=fmt)
    st.set_volumes(v.lf, v.hf, v.curl, v.weather)
    return st, v


def _time(frame=1, total=1.0 / 60.0):
    t = M.Time()
    t.ubo.frame_count_mod16 = frame
    t.ubo.time[0] = 1.0 / 60.0
    t.ubo.time[1] = total
    return t


@pytest.mark.parametrize("dims,fmt", [(SMALL, 0), (FULL, 0), ((64, 32, 64, 128), 1)])
def test_noise_generation_is_bit_exact(dims, fmt):
    with M.CloudStage(16, 16, M.VolumeSpec(*dims, fmt)) as st:
        st.generate_volumes(1, 2)
        lf, hf, curl, weather = st.get_volumes()
    ref = O.generate_volumes(*dims, fmt=fmt, seed_volumes=1, seed_weather=2)
    assert np.array_equal(lf, ref.lf)
    assert np.array_equal(hf, ref.hf)
    assert np.array_equal(curl, ref.curl)
    assert np.array_equal(weather, ref.weather)


def _compare_selected(rgba, mask, W, H, vol, cam, tm, sun, params):
    xy = O.select_pixels(W, H, tm.frame_count_mod16)
    ref, refm, steps, cnt = O.march_pixels(vol, cam, tm, sun, params, W, H, xy, 32)
    got = rgba[xy[:, 1], xy[:, 0]]
    gotm = mask[xy[:, 1], xy[:, 0]]
    bad = within_rel(got, ref)
    badm = within_rel(gotm, refm, floor_frac=1.0)
    assert not bad.any(), f"{int(bad.sum())} colour values off; worst rel {np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-6))}"
    assert not badm.any(), f"{int(badm.sum())} mask values off"
    return xy, cnt


@pytest.mark.parametrize("W,H,preset,pitch", [(480, 270, 1, 20.0), (97, 61, 1, 10.0), (256, 144, 2, 30.0)])
def test_plain_march_matches_oracle(W, H, preset, pitch):
    params, sun = M.preset_params(preset)
    cam = ground_camera(W, H, pitch).ubo()
    for frame in (1, 6, 15):
        tm = _time(frame, frame / 60.0).ubo
        st, v = _stage(W, H, flags=A.MTC_FLAG_COUNTERS)
        with st:
            st.frame(cam, cam, tm, sun, params)
            rgba, mask = st.read_output()
            c = st.counters()
        xy, cnt = _compare_selected(rgba, mask, W, H, v, cam, tm, sun, params)
        for k in O.COUNTER_NAMES:
            assert c[k] == cnt[k], (k, c[k], cnt[k])
        assert c["rays_marched"] + c["rays_ocean"] + c["rays_sky"] == xy.shape[0]


def test_plain_march_c2_frame_1080p():
    W, H = 1920, 1080
    params, sun = M.preset_params(2)
    cam = ground_camera(W, H, 20.0).ubo()
    tm = _time(1).ubo
    st, v = _stage(W, H, flags=A.MTC_FLAG_COUNTERS)
    with st:
        st.frame(cam, cam, tm, sun, params)
        rgba, mask = st.read_output()
        c = st.counters()
    xy, cnt = _compare_selected(rgba, mask, W, H, v, cam, tm, sun, params)
    assert c["primary_steps"] == cnt["primary_steps"]
    assert xy.shape[0] == 129600


def test_reference_preset_matches_oracle():
    W, H = 160, 120
    params, sun = M.reference_params()
    cam = M.Camera((0.0, 0.0, 2.0), (0.0, 0.5, 1.0), W, H, 45.0).ubo()
    tm = _time(3, 0.5).ubo
    st, v = _stage(W, H)
    with st:
        st.frame(cam, cam, tm, sun, params)
        rgba, mask = st.read_output()
    _compare_selected(rgba, mask, W, H, v, cam, tm, sun, params)


def test_full_march_every_pixel():
    W, H = 120, 68
    params, sun = M.preset_params(1)
    cam = ground_camera(W, H, 15.0).ubo()
    tm = _time(4).ubo
    st, v = _stage(W, H, dims=SMALL, vol=volumes(*SMALL))
    with st:
        st.march_full(cam, tm, sun, params)
        rgba, mask = st.read_output()
    ref, refm, _ = O.frame(v, cam, None, tm, sun, params, W, H, precision=32, full=True)
    assert not within_rel(rgba, ref).any()
    assert not within_rel(mask, refm, floor_frac=1.0).any()


def test_ocean_and_empty_volume_edge_cases():
    W, H = 64, 40
    params, sun = M.preset_params(1)
    v = volumes(*SMALL)
    lf = np.zeros_like(v.lf).reshape(-1, 4)
    lf[:, 1:] = 255  # zero density everywhere
    empty = O.Volumes(lf.reshape(-1), np.zeros_like(v.hf), v.curl, v.weather, *SMALL)
    for pitch, vol in ((-70.0, v), (25.0, empty)):
        cam = ground_camera(W, H, pitch).ubo()
        tm = _time(0).ubo
        st, _ = _stage(W, H, dims=SMALL, vol=vol)
        with st:
            st.march_full(cam, tm, sun, params)
            rgba, mask = st.read_output()
        ref, refm, cnt = O.frame(vol, cam, None, tm, sun, params, W, H, precision=32, full=True)
        assert not within_rel(rgba, ref).any()
        assert np.array_equal(mask, refm)


@pytest.mark.parametrize("W,H", [(480, 270), (1920, 1080)])
def test_reprojection_indices_bit_exact(W, H):
    params, sun = M.preset_params(2)
    camera = ground_camera(W, H, 20.0)
    old = camera.ubo()
    camera.rotate_about_up(0.5)
    camera.rotate_about_right(0.3)
    cur = camera.ubo()
    tm = _time(7).ubo
    with M.CloudStage(W, H, M.VolumeSpec(*SMALL)) as st:
        got = st.reprojection_indices(cur, old, tm, params)
    ref = O.reprojection_indices(cur, old, tm, params, W, H, 32)
    assert np.array_equal(got, ref), f"{int((got != ref).sum())} tap indices differ"


def test_frame_sequence_with_reprojection_matches_oracle():
    """Eight frames, yaw 0.5 deg/frame, 4x4 reprojection (the frame loop of
    main.cpp:172-186): time -> frame -> old camera <- camera."""
    W, H = 240, 136
    params, sun = M.preset_params(2)
    camera = ground_camera(W, H, 20.0)
    time = M.Time()
    st, v = _stage(W, H, dims=SMALL, vol=volumes(*SMALL))
    old = camera.ubo()
    prev = np.zeros((H, W, 4), np.float32)
    mask_ref = np.zeros((H, W, 4), np.float32)
    with st:
        for k in range(8):
            camera.rotate_about_up(0.5)
            time.advance(1.0 / 60.0)
            cur = camera.ubo()
            st.frame(cur, old, time.ubo, sun, params)
            rgba, mask = st.read_output()
            ref, mask_ref, _ = O.frame(v, cur, old, time.ubo, sun, params, W, H, prev=prev, precision=32,
                                       mask=mask_ref)
            bad = within_rel(rgba, ref)
            assert not bad.any(), (k, int(bad.sum()))
            assert not within_rel(mask, mask_ref, floor_frac=1.0).any(), k
            prev = ref
            old = cur


def test_history_and_error_paths():
    W, H = 32, 24
    params, sun = M.preset_params(1)
    cam = ground_camera(W, H).ubo()
    tm = _time(2).ubo
    with M.CloudStage(W, H, M.VolumeSpec(*SMALL)) as st:
        with pytest.raises(M.MtcError) as e:
            st.frame(cam, cam, tm, sun, params)
        assert e.value.code == A.MTC_ERR_STATE
        st.generate_volumes(1, 2)
        bad_t = _time(2).ubo
        bad_t.frame_count_mod16 = 16
        with pytest.raises(M.MtcError):
            st.frame(cam, cam, bad_t, sun, params)
        nan_cam = ground_camera(W, H).ubo()
        nan_cam.view[3] = float("nan")
        with pytest.raises(M.MtcError):
            st.frame(nan_cam, cam, tm, sun, params)
        hist = np.full((H, W, 4), 0.25, np.float32)
        st.set_history(hist)
        st.frame(cam, cam, tm, sun, params)
        rgba, _ = st.read_output()
        sel = np.zeros((H, W), bool)
        xy = O.select_pixels(W, H, 2)
        sel[xy[:, 1], xy[:, 0]] = True
        assert np.allclose(rgba[~sel], 0.25, rtol=1e-6)
        st.reset_history()
        st.frame(cam, cam, tm, sun, params)
        rgba, _ = st.read_output()
        assert np.all(rgba[~sel] == 0.0)


@pytest.mark.parametrize("yaw,pitch", [(0.5, 0.0), (0.5, 0.4), (12.0, 6.0)])
def test_reprojection_of_the_fast_path_is_bit_identical(yaw, pitch):
    """Small and large camera motion: every pixel the march does not
    overwrite equals the plain context's and the FP32 oracle's bit for bit."""
    W, H = 640, 360
    params, sun = M.preset_params(2)
    camera = ground_camera(W, H, 20.0)
    old = camera.ubo()
    camera.rotate_about_up(yaw)
    camera.rotate_about_right(pitch)
    cur = camera.ubo()
    tm = _time(6).ubo
    rng = np.random.default_rng(21)
    hist = rng.uniform(0.0, 4.0, size=(H, W, 4)).astype(np.float32)
    outs = []
    for kernel in (A.MTC_KERNEL_PLAIN, A.MTC_KERNEL_FAST):
        st, _ = _stage(W, H, dims=SMALL, vol=volumes(*SMALL), kernel=kernel)
        with st:
            st.set_history(hist)
            st.frame(cur, old, tm, sun, params)
            outs.append(st.read_output()[0])
    sel = np.zeros((H, W), bool)
    xy = O.select_pixels(W, H, 6)
    sel[xy[:, 1], xy[:, 0]] = True
    assert np.array_equal(outs[0][~sel], outs[1][~sel])
    ref, _, _ = O.frame(volumes(*SMALL), cur, old, tm, sun, params, W, H, prev=hist, precision=32)
    assert np.array_equal(outs[1][~sel], ref[~sel])


@pytest.mark.parametrize("kernel", [A.MTC_KERNEL_FAST, A.MTC_KERNEL_PRECISION, A.MTC_KERNEL_LEARNED])
def test_reprojection_strips_of_a_ragged_frame_are_bit_identical(kernel):
    """The optimised launches carry the reprojection as 64 x 4 pixel strips
    after the march tiles; on a frame whose width and height are not
    multiples of the strip (203 x 75) every pixel the march does not
    overwrite equals the FP32 oracle's bit for bit."""
    W, H = 203, 75
    params, sun = M.preset_params(2)
    camera = ground_camera(W, H, 20.0)
    old = camera.ubo()
    camera.rotate_about_up(0.5)
    camera.rotate_about_right(0.4)
    cur = camera.ubo()
    tm = _time(9).ubo
    hist = np.random.default_rng(22).uniform(0.0, 4.0, size=(H, W, 4)).astype(np.float32)
    with M.LightNet(capacity=1 << 10) as net:
        net.init(3)
        st, _ = _stage(W, H, dims=SMALL, vol=volumes(*SMALL), kernel=kernel)
        with st:
            if kernel == A.MTC_KERNEL_LEARNED:
                net.attach(st)
            st.set_history(hist)
            st.frame(cur, old, tm, sun, params)
            out = st.read_output()[0]
    sel = np.zeros((H, W), bool)
    xy = O.select_pixels(W, H, 9)
    sel[xy[:, 1], xy[:, 0]] = True
    ref, _, _ = O.frame(volumes(*SMALL), cur, old, tm, sun, params, W, H, prev=hist, precision=32)
    assert np.array_equal(out[~sel].view(np.uint32), ref[~sel].view(np.uint32))


@pytest.mark.parametrize("kernel", [A.MTC_KERNEL_PLAIN, A.MTC_KERNEL_PRECISION])
def test_timing_events_sample_frames_without_changing_results(kernel):
    """MTC_FLAG_TIMING attaches events to one frame in `stride`; the frames
    it times and the ones it does not give the same pixels as an untimed
    context."""
    W, H = 160, 96
    params, sun = M.preset_params(1)
    cam = ground_camera(W, H, 25.0)
    seq = []
    old = cam.ubo()
    for f in range(8):
        cam.rotate_about_up(0.5)
        t = _time(f)
        seq.append((cam.ubo(), old, t.ubo))
        old = seq[-1][0]
    outs = []
    for flags in (0, A.MTC_FLAG_TIMING):
        st, _ = _stage(W, H, dims=SMALL, vol=volumes(*SMALL), kernel=kernel, flags=flags)
        with st:
            if flags:
                st.set_timing_stride(3)
            for cur, prev, tm in seq:
                st.frame(cur, prev, tm, sun, params)
            outs.append(st.read_output()[0])
            if flags:
                tt = st.timings()
                assert tt["frames"] == 3  # frames 0, 3, 6
                assert tt["march_ms"] > 0.0
                if kernel == A.MTC_KERNEL_PLAIN:
                    assert tt["reproject_ms"] > 0.0
                else:  # the reprojection runs inside the march launch
                    assert tt["reproject_ms"] == 0.0
                with pytest.raises(M.MtcError):
                    st.set_timing_stride(0)
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


@pytest.mark.parametrize("kernel", [A.MTC_KERNEL_PLAIN, A.MTC_KERNEL_FAST, A.MTC_KERNEL_PRECISION])
def test_low_fade_start_scales_the_horizon_mask(kernel):
    """With cloud_fade_start = 0.01 rays with 0.01 <= dir.y < 0.05 are
    marched, and their god-ray mask is multiplied by max(5, dir.y)
    (cloudRayMarch.comp:775-779, unreachable with the shipped 0.06).  The
    plain kernel matches the FP32 oracle to 1e-3; the optimised kernels'
    decoded masks match the FP64 oracle's on average to 5 x 2e-3."""
    W, H = 160, 90
    params, sun = M.preset_params(1)
    params.cloud_fade_start = 0.01
    cam = ground_camera(W, H, 2.0).ubo()
    tm = _time(2).ubo
    st, v = _stage(W, H, kernel=kernel)
    with st:
        st.march_full(cam, tm, sun, params)
        rgba, mask = st.read_output()
    xy = np.stack(np.meshgrid(np.arange(W), np.arange(H)), -1).reshape(-1, 2).astype(np.int32)
    dy = O.ray_directions(cam, tm, xy, W, H)[:, 1].reshape(H, W)
    low = (dy >= 0.011) & (dy < 0.049)
    assert low.sum() > 500
    if kernel == A.MTC_KERNEL_PLAIN:
        ref, refm, _ = O.frame(v, cam, None, tm, sun, params, W, H, precision=32, full=True)
        assert not within_rel(rgba, ref).any()
        assert not within_rel(mask, refm, floor_frac=1.0).any()
        assert np.abs(refm[low]).max() > 1.0  # a scaled encoding: components above 1
    else:
        r64, m64, _, _ = O.march_pixels(v, cam, tm, sun, params, W, H, xy, 64)
        dec = lambda m: m @ (1.0 / np.array([1.0, 255.0, 65025.0, 16581375.0]))  # noqa: E731
        got = dec(mask.reshape(-1, 4).astype(np.float64))
        ref = dec(m64)
        assert np.abs(got - ref)[low.reshape(-1)].mean() < 1e-2
        assert np.abs(m64[low.reshape(-1)]).max() > 1.0
