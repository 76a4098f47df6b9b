"""The LDS-staged LF bricks (MTC_FLAG_LDS_BRICKS) against the default
footprint-table path of the same kernel.

With the flag, the primary steps' LF samples of each wave read an 8^3-texel
box staged in LDS from a 4^3-bricked copy of the LF, where their 2x2x2
footprint fits the box, and the footprint table otherwise.  Both copies
hold the same half2 (r, fbm) texels and feed the same filter, so every
colour, mask and counter must be the same bits as without the flag; the
counters also show how many samples came from LDS.  Cases: the 1080p
frame sequence with reprojection (weather coverage, both variants), the
reference preset (constant coverage, 35-60 steps) marched in full, an 8^3
LF (every box wraps around the volume) and the 4K workload's 256^3 FP16
volumes."""
import numpy as np
import pytest

import meteoros_amd as M
from meteoros_amd import _abi as A
from tests.common import ground_camera

pytestmark = pytest.mark.gpu

STD = ("rays_marched", "rays_ocean", "rays_sky", "primary_steps", "cloud_steps", "cone_samples", "cone_erosions",
       "early_exits")


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _run(kernel, flags, W, H, spec, frames, params, sun, vols=None, full=False):
    st = M.CloudStage(W, H, spec, kernel=kernel, flags=flags | A.MTC_FLAG_COUNTERS)
    try:
        if vols is None:
            st.generate_volumes(1, 2)
        else:
            st.set_volumes(*vols)
        for cur, old, tm in frames:
            if full:
                st.march_full(cur, tm, sun, params)
            else:
                st.frame(cur, old, tm, sun, params)
        rgba, mask = st.read_output()
        return rgba, mask, st.counters(), (st.get_volumes() if vols is None else vols)
    finally:
        st.close()


def _sequence(W, H, pitch, yaw, n, start=0):
    cam = ground_camera(W, H, pitch, yaw)
    t = M.Time()
    for _ in range(start):
        t.advance(1.0 / 60.0)
    old, out = cam.ubo(), []
    for _ in range(n):
        cam.rotate_about_up(0.5)
        t.advance(1.0 / 60.0)
        tm = A.MtcTime.from_buffer_copy(t.ubo)
        out.append((cam.ubo(), old, tm))
        old = out[-1][0]
    return out


def _compare(kernel, W, H, spec, frames, params, sun, full=False, label=""):
    print(f"bricks vs table: {label} kernel {kernel} {W}x{H}", flush=True)
    a_rgba, a_mask, a_cnt, vols = _run(kernel, 0, W, H, spec, frames, params, sun, full=full)
    b_rgba, b_mask, b_cnt, _ = _run(kernel, A.MTC_FLAG_LDS_BRICKS, W, H, spec, frames, params, sun, vols, full)
    served = b_cnt["lf_brick_samples"] / max(b_cnt["primary_steps"], 1)
    print(f"  {b_cnt['lf_brick_samples']} primary LF samples from LDS ({served:.1%} of the primary steps), "
          f"{b_cnt['lf_brick_stages']} box stagings; colour values differing: "
          f"{int((_bits(a_rgba) != _bits(b_rgba)).sum())}", flush=True)
    assert np.array_equal(_bits(a_rgba), _bits(b_rgba))
    assert np.array_equal(_bits(a_mask), _bits(b_mask))
    assert {k: a_cnt[k] for k in STD} == {k: b_cnt[k] for k in STD}
    assert a_cnt["lf_brick_samples"] == 0 and a_cnt["lf_brick_stages"] == 0
    assert b_cnt["lf_brick_samples"] > 0 and b_cnt["lf_brick_stages"] > 0
    return b_cnt


@pytest.mark.parametrize("kernel", [A.MTC_KERNEL_FAST, A.MTC_KERNEL_PRECISION])
def test_bricks_bit_identical_on_the_1080p_sequence(kernel):
    params, sun = M.preset_params(2)
    frames = _sequence(1920, 1080, 20.0, 0.0, 4)
    _compare(kernel, 1920, 1080, M.VolumeSpec(128, 32, 128, 512), frames, params, sun, label="1080p sequence")


@pytest.mark.parametrize("kernel", [A.MTC_KERNEL_FAST, A.MTC_KERNEL_PRECISION])
def test_bricks_bit_identical_on_the_reference_preset(kernel):
    params, sun = M.reference_params()
    frames = _sequence(320, 180, 25.0, 30.0, 1, start=5)
    _compare(kernel, 320, 180, M.VolumeSpec(128, 32, 128, 512), frames, params, sun, full=True,
             label="reference preset, full march")


def test_bricks_bit_identical_when_every_box_wraps():
    params, sun = M.preset_params(2)
    frames = _sequence(480, 270, 15.0, 60.0, 2)
    _compare(A.MTC_KERNEL_PRECISION, 480, 270, M.VolumeSpec(8, 8, 16, 32), frames, params, sun, full=True,
             label="8^3 LF, full march")


def test_bricks_bit_identical_on_fp16_256_volumes():
    params, sun = M.preset_params(3)
    frames = _sequence(960, 540, 40.0, 90.0, 2)
    _compare(A.MTC_KERNEL_PRECISION, 960, 540, M.VolumeSpec(256, 64, 128, 512, A.MTC_FMT_FP16), frames, params,
             sun, label="256^3 FP16 LF")
