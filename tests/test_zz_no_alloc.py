"""The allocation-free god-ray entry point (mtc_god_rays_noalloc) against
the context's own pass (mtc_god_rays), through the C ABI.

Two contexts march the same frame (the march is deterministic, so their
result images and masks are equal, which is checked first); one runs
mtc_god_rays on its own images, the other mtc_god_rays_noalloc on the
device pointers of its images.  The results must be bit-identical: both
run the same decode and blur kernels on the same inputs, one with the
context's footprint map, the other with the library's static scratch.
Two calls in a row, the second smaller, so the second finds the first
call's contents in the shared scratch."""
import ctypes as C

import numpy as np
import pytest

import meteoros_amd as M
from meteoros_amd import _abi as A
from tests.common import ground_camera, volumes

pytestmark = pytest.mark.gpu

DIMS = (64, 16, 64, 128)


def _marched_stage(W, H, cam, tm, sun, params):
    st = M.CloudStage(W, H, M.VolumeSpec(*DIMS), kernel=A.MTC_KERNEL_FAST)
    v = volumes(*DIMS)
    st.set_volumes(v.lf, v.hf, v.curl, v.weather)
    st.march_full(cam, tm, sun, params)
    return st


@pytest.mark.parametrize("pitch", [35.0])
def test_noalloc_god_rays_equal_the_context_pass(pitch):
    lib = M.lib()
    params, sun = M.preset_params(2)
    for k in range(3):
        sun.light_color[k] = (1.0, 0.95, 0.85)[k]
    t = M.Time()
    t.ubo.frame_count_mod16 = 5
    t.ubo.time[1] = 3.5
    for W, H in [(400, 224), (256, 144)]:
        print(f"noalloc god rays: {W}x{H}, pitch {pitch}", flush=True)
        assert (W + 1) * (H + 1) <= A.MTC_GOD_RAYS_NOALLOC_MAX_ENTRIES
        cam = ground_camera(W, H, pitch, 10.0).ubo()
        ref_st = _marched_stage(W, H, cam, t.ubo, sun, params)
        new_st = _marched_stage(W, H, cam, t.ubo, sun, params)
        with ref_st, new_st:
            rgba_a, mask_a = ref_st.read_output()
            rgba_b, mask_b = new_st.read_output()
            assert np.array_equal(rgba_a, rgba_b) and np.array_equal(mask_a, mask_b)
            ref_st.god_rays(cam, sun)
            ref, _ = ref_st.read_output()
            img, msk = C.c_void_p(), C.c_void_p()
            assert lib.mtc_output_device_ptrs(new_st.handle, C.byref(img), C.byref(msk)) == A.MTC_OK
            assert img.value and msk.value
            new_st.synchronize()
            rc = lib.mtc_god_rays_noalloc(W, H, C.byref(cam), C.byref(sun), msk, img, None)
            assert rc == A.MTC_OK, lib.mtc_last_error()
            new_st.synchronize()
            got, _ = new_st.read_output()
        added = float((ref - rgba_a)[..., :3].sum())
        print(f"  energy added {added:.2f}; {int((got != ref).sum())} values differ", flush=True)
        assert added > 0.0
        assert np.array_equal(got, ref)
