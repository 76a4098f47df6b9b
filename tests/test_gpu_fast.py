"""Accuracy of the optimised march kernels against the FP64 CPU oracle.

The fast kernel (footprint-packed FP16 texels, FP32 filtering, hoisted
erosion, exp2/log2 pow) and the precision kernel (FP16 volumes, packed-FP16
filtering, FP64 shell quadratic) are judged by PSNR against the FP64 oracle
on the same rays: PSNR = 10 log10(peak^2 / MSE) over RGB of the marched
pixels outside the sun disk (oracle.outside_sun_disk: the disk's edge is a
smoothstep over 2e-5 in cos theta, ill-conditioned in the ray direction),
peak = max of the FP64 reference over those pixels.  Bar: 45 dB.
"""
import numpy as np
import pytest

import meteoros_amd as M
from meteoros_amd import _abi as A
from oracle import oracle as O
from tests.common import ground_camera, psnr, volumes, within_rel

pytestmark = pytest.mark.gpu

PSNR_DB = 45.0


def _render(kernel, W, H, dims, fmt, cam, tm, sun, params, flags=0, vol=None):
    st = M.CloudStage(W, H, M.VolumeSpec(*dims, fmt), kernel=kernel, flags=flags)
    if vol is None:
        st.generate_volumes(1, 2)
        lf, hf, curl, weather = st.get_volumes()
        vol = O.Volumes(lf, hf, curl, weather, *dims, fmt)
    else:
        st.set_volumes(vol.lf, vol.hf, vol.curl, vol.weather)
    st.frame(cam, cam, tm, sun, params)
    rgba, mask = st.read_output()
    cnt = st.counters() if flags & A.MTC_FLAG_COUNTERS else None
    st.close()
    return rgba, mask, cnt, vol


def _time(frame):
    t = M.Time()
    t.ubo.frame_count_mod16 = frame
    t.ubo.time[1] = frame / 60.0
    return t.ubo


@pytest.mark.parametrize("kernel", [A.MTC_KERNEL_FAST, A.MTC_KERNEL_PRECISION])
@pytest.mark.parametrize("pitch,frame", [(20.0, 1), (8.0, 9)])
def test_fast_kernels_psnr_vs_fp64_at_1080p(kernel, pitch, frame):
    W, H = 1920, 1080
    params, sun = M.preset_params(2)
    cam = ground_camera(W, H, pitch).ubo()
    tm = _time(frame)
    rgba, mask, cnt, vol = _render(kernel, W, H, (128, 32, 128, 512), A.MTC_FMT_UNORM8, cam, tm, sun, params,
                                   flags=A.MTC_FLAG_COUNTERS)
    xy = O.select_pixels(W, H, frame)[::3]
    ref, refm, _, ocnt = O.march_pixels(vol, cam, tm, sun, params, W, H, xy, 64)
    got = rgba[xy[:, 1], xy[:, 0], :3]
    keep = O.outside_sun_disk(cam, tm, params, xy, W, H)
    p = O.psnr_vs_fp64(got, ref[:, :3], keep)
    print(f"kernel {kernel} pitch {pitch}: PSNR {p:.2f} dB vs FP64 ({int((~keep).sum())} sun-disk rays left out), "
          f"unmasked {psnr(got, ref[:, :3]):.2f} dB")
    assert p >= PSNR_DB
    # the god-ray mask encodes 25*min(0.05, 1-accum): compare decoded values
    dec = lambda m: m @ (1.0 / np.array([1.0, 255.0, 65025.0, 16581375.0]))  # noqa: E731
    assert np.abs(dec(mask[xy[:, 1], xy[:, 0]].astype(np.float64)) - dec(refm)).mean() < 2e-3
    # the same rays march the same number of steps up to one per ray: whether
    # t_N (~ the outer shell) still satisfies t < end_t is a rounding coin flip
    # that the FP32 and FP64 oracles also decide differently
    full_xy = O.select_pixels(W, H, frame)
    _, _, steps32, c32 = O.march_pixels(vol, cam, tm, sun, params, W, H, full_xy, 32)
    assert cnt["rays_marched"] == c32["rays_marched"]
    assert abs(cnt["primary_steps"] - c32["primary_steps"]) <= c32["rays_marched"]
    assert abs(cnt["primary_steps"] - c32["primary_steps"]) <= 0.01 * c32["primary_steps"]


@pytest.mark.parametrize("kernel", [A.MTC_KERNEL_FAST, A.MTC_KERNEL_PRECISION])
def test_fast_kernels_at_480x270_with_64_steps(kernel):
    W, H = 480, 270
    params, sun = M.preset_params(1)
    cam = ground_camera(W, H, 20.0).ubo()
    tm = _time(5)
    vol = volumes(128, 32, 128, 512)
    rgba, _, _, _ = _render(kernel, W, H, (128, 32, 128, 512), 0, cam, tm, sun, params, vol=vol)
    xy = O.select_pixels(W, H, 5)
    ref, _, _, _ = O.march_pixels(vol, cam, tm, sun, params, W, H, xy, 64)
    got = rgba[xy[:, 1], xy[:, 0], :3]
    p = O.psnr_vs_fp64(got, ref[:, :3], O.outside_sun_disk(cam, tm, params, xy, W, H))
    print(f"kernel {kernel} 480x270: PSNR {p:.2f} dB")
    assert p >= PSNR_DB
    # ocean and sky-only pixels take no march: they agree to float rounding
    r32, _, steps, _ = O.march_pixels(vol, cam, tm, sun, params, W, H, xy, 32)
    nomarch = steps == 0
    assert nomarch.any()
    assert not within_rel(got[nomarch], r32[nomarch, :3], rel=1e-4).any()


@pytest.mark.parametrize("kernel", [A.MTC_KERNEL_FAST, A.MTC_KERNEL_PRECISION])
def test_reference_preset(kernel):
    """The shipped constants (mtc_params_reference: shell R+7.5..R+20 km,
    35-60 steps, constant coverage 0.6): the preset closest to the bar,
    because the FP32 shell quadratic of the reference cancels there."""
    W, H = 320, 180
    params, sun = M.reference_params()
    cam = M.Camera((0.0, 0.0, 2.0), (0.0, 0.5, 1.0), W, H, 45.0).ubo()
    tm = _time(3)
    vol = volumes(128, 32, 128, 512)
    rgba, _, _, _ = _render(kernel, W, H, (128, 32, 128, 512), 0, cam, tm, sun, params, vol=vol)
    xy = O.select_pixels(W, H, 3)
    ref, _, _, _ = O.march_pixels(vol, cam, tm, sun, params, W, H, xy, 64)
    got = rgba[xy[:, 1], xy[:, 0], :3]
    p = O.psnr_vs_fp64(got, ref[:, :3], O.outside_sun_disk(cam, tm, params, xy, W, H))
    print(f"reference preset, kernel {kernel}: PSNR {p:.2f} dB")
    assert p >= PSNR_DB


def test_precision_variant_4k_fp16_volumes_over_poses():
    """3840x2160, 192 steps, LF 256^3 / HF 64^3 FP16 volumes, poses spread
    over yaw and elevation 0-60 deg; a strided subset of each pose's marched
    rays is checked against the FP64 oracle."""
    W, H = 3840, 2160
    dims = (256, 64, 128, 512)
    params, sun = M.preset_params(3)
    st = M.CloudStage(W, H, M.VolumeSpec(*dims, A.MTC_FMT_FP16), kernel=A.MTC_KERNEL_PRECISION)
    st.generate_volumes(1, 2)
    vol = O.Volumes(*st.get_volumes(), *dims, A.MTC_FMT_FP16)
    worst = 1e9
    rng = np.random.default_rng(3)
    for k, (yaw, elev) in enumerate([(y, e) for y in (0.0, 90.0, 180.0, 270.0) for e in (0.0, 20.0, 40.0, 60.0)]):
        cam = ground_camera(W, H, elev, yaw).ubo()
        tm = _time(k % 16)
        st.frame(cam, cam, tm, sun, params)
        rgba, _ = st.read_output()
        xy = O.select_pixels(W, H, k % 16)
        xy = xy[rng.choice(xy.shape[0], 3000, replace=False)]
        ref, _, _, _ = O.march_pixels(vol, cam, tm, sun, params, W, H, xy, 64)
        got = rgba[xy[:, 1], xy[:, 0], :3]
        p = O.psnr_vs_fp64(got, ref[:, :3], O.outside_sun_disk(cam, tm, params, xy, W, H))
        print(f"pose yaw {yaw} elev {elev}: PSNR {p:.2f} dB")
        worst = min(worst, p)
    st.close()
    assert worst >= PSNR_DB


@pytest.mark.parametrize("kernel", [A.MTC_KERNEL_FAST, A.MTC_KERNEL_PRECISION])
def test_results_do_not_depend_on_lanes_per_ray(kernel):
    """Each lane group replays its steps' accumulation in step order, so a
    ray's colour, mask and the counters are the same bits whether 1, 2, 4 or
    8 lanes share it (the grouping only changes which lane computes which
    step and how the cone work is dealt out)."""
    W, H = 480, 270
    params, sun = M.preset_params(2)
    cam = ground_camera(W, H, 15.0, 30.0).ubo()
    tm = _time(7)
    vol = None
    outs = []
    for lanes in (1, 2, 4, 8):
        st = M.CloudStage(W, H, M.VolumeSpec(128, 32, 128, 512), kernel=kernel, flags=A.MTC_FLAG_COUNTERS,
                          lanes_per_ray=lanes)
        if vol is None:
            st.generate_volumes(1, 2)
            vol = st.get_volumes()
        else:
            st.set_volumes(*vol)
        st.frame(cam, cam, tm, sun, params)
        rgba, mask = st.read_output()
        outs.append((rgba, mask, st.counters()))
        st.close()
    bits = lambda a: np.ascontiguousarray(a).view(np.uint32)  # noqa: E731
    for rgba, mask, cnt in outs[1:]:
        assert np.array_equal(bits(rgba), bits(outs[0][0]))
        assert np.array_equal(bits(mask), bits(outs[0][1]))
        assert cnt == outs[0][2]


@pytest.mark.parametrize("W,H,full,lanes", [(416, 172, False, 8), (416, 172, False, 4), (300, 77, True, 8),
                                             (1920, 1080, False, 8), (36, 20, True, 2)])
def test_strip_order_marches_every_ray_once(W, H, full, lanes):
    """The fast launch deals its tiles to the XCDs in column strips (a
    different tile order per image width, with rows past the last whole
    block of strips in raster order).  Every ray of the frame must be marched
    exactly once: the counters' marched + ocean + sky rays equal the frame's
    rays (a tile dealt twice or never would move the sum by a tile's rays),
    and every one of them holds a written colour (the march writes alpha 1
    on every ray it visits; the context's images start as zeros)."""
    params, sun = M.preset_params(2)
    cam = ground_camera(W, H, 10.0, 25.0).ubo()
    tm = _time(5)
    st = M.CloudStage(W, H, M.VolumeSpec(128, 32, 128, 512), kernel=A.MTC_KERNEL_PRECISION,
                      flags=A.MTC_FLAG_COUNTERS, lanes_per_ray=lanes)
    st.generate_volumes(1, 2)
    if full:
        st.march_full(cam, tm, sun, params)
        n_rays = W * H
    else:
        st.frame(cam, cam, tm, sun, params)
        n_rays = len(O.select_pixels(W, H, 5))
    cnt = st.counters()
    rgba = st.read_output()[0]
    st.close()
    assert cnt["rays_marched"] + cnt["rays_ocean"] + cnt["rays_sky"] == n_rays
    if full:
        assert np.all(rgba[..., 3] == 1.0)
    else:
        xy = O.select_pixels(W, H, 5)
        assert np.all(rgba[xy[:, 1], xy[:, 0], 3] == 1.0)


@pytest.mark.parametrize("kernel", [A.MTC_KERNEL_FAST, A.MTC_KERNEL_PRECISION])
def test_single_and_paired_cone_rounds_agree(kernel):
    """Launches of few tiles per compute unit run the cone rounds two at a
    time (5 waves per SIMD), longer ones one at a time (6 waves per SIMD).
    A 1920x1080 4x4 frame (16 200 tiles) takes the paired kernel and the
    full march of the same view (259 200 tiles) the single-round one; a
    pixel's result depends only on its ray and the frame's uniforms, so the
    frame's marched pixels must be the same bits in both."""
    W, H = 1920, 1080
    params, sun = M.preset_params(2)
    cam = ground_camera(W, H, 15.0, 20.0).ubo()
    tm = _time(3)
    spec = M.VolumeSpec(128, 32, 128, 512)
    with M.CloudStage(W, H, spec, kernel=kernel) as a, M.CloudStage(W, H, spec, kernel=kernel) as b:
        a.generate_volumes(1, 2)
        b.generate_volumes(1, 2)
        a.frame(cam, cam, tm, sun, params)
        b.march_full(cam, tm, sun, params)
        fa, ma = a.read_output()
        fb, mb = b.read_output()
    xy = O.select_pixels(W, H, 3)
    bits = lambda x: np.ascontiguousarray(x[xy[:, 1], xy[:, 0]]).view(np.uint32)  # noqa: E731
    assert np.array_equal(bits(fa), bits(fb))
    assert np.array_equal(bits(ma), bits(mb))


@pytest.mark.parametrize("kernel", [A.MTC_KERNEL_FAST, A.MTC_KERNEL_PRECISION, A.MTC_KERNEL_LEARNED])
def test_march_history_changes_no_pixel(kernel):
    """Whatever a context keeps from march to march (the learned march deals
    its tiles longest first by the previous march's measured durations; the
    others keep raster order), each frame of a moving sequence, with a full
    march in the middle, must give the same bits as the same frame run as
    the first march of a fresh context from the same history."""
    W, H = 480, 270
    params, sun = M.preset_params(2)
    cam = ground_camera(W, H, 12.0, 40.0)
    ops, old = [], cam.ubo()
    for f in range(7):
        cam.rotate_about_up(0.5)
        ops.append(("full" if f == 3 else "frame", cam.ubo(), old, _time(f)))
        old = ops[-1][1]
    spec = M.VolumeSpec(128, 32, 128, 512)
    with M.LightNet(capacity=1 << 10) as net:
        net.init(5)  # an untrained network answers as deterministically as a trained one
        a = M.CloudStage(W, H, spec, kernel=kernel)
        if kernel == A.MTC_KERNEL_LEARNED:
            net.attach(a)
        a.generate_volumes(1, 2)
        vol = a.get_volumes()
        history, results = [], []
        for op, cur, prev, tm in ops:
            history.append(a.read_output()[0])
            if op == "full":
                a.march_full(cur, tm, sun, params)
            else:
                a.frame(cur, prev, tm, sun, params)
            results.append(a.read_output())
        a.close()
        bits = lambda x: np.ascontiguousarray(x).view(np.uint32)  # noqa: E731
        for k, (op, cur, prev, tm) in enumerate(ops):
            if k == 0:
                continue  # the first march of both contexts
            with M.CloudStage(W, H, spec, kernel=kernel) as b:
                if kernel == A.MTC_KERNEL_LEARNED:
                    net.attach(b)
                b.set_volumes(*vol)
                b.set_history(history[k])
                if op == "full":
                    b.march_full(cur, tm, sun, params)
                else:
                    b.frame(cur, prev, tm, sun, params)
                rgba, mask = b.read_output()
            assert np.array_equal(bits(rgba), bits(results[k][0])), f"{op} {k}: colour differs"
            if op == "full":
                assert np.array_equal(bits(mask), bits(results[k][1])), f"{op} {k}: mask differs"
            else:
                xy = O.select_pixels(W, H, tm.frame_count_mod16)
                m0, m1 = mask[xy[:, 1], xy[:, 0]], results[k][1][xy[:, 1], xy[:, 0]]
                assert np.array_equal(bits(m0), bits(m1)), f"{op} {k}: mask differs"
