"""Batched views: one persistent launch (hipGraph) over many views of uneven
size, autotuned.  Each ray's result does not depend on lane grouping, tile
shape, item order or the cone compaction, so every batched view must equal a
single-view full march of the same kernel bit for bit; accuracy against the
FP64 oracle is checked on a sample of rays."""
import numpy as np
import pytest

import meteoros_amd as M
from meteoros_amd import _abi as A
from oracle import oracle as O
from tests.common import ground_camera

pytestmark = pytest.mark.gpu

DIMS = (128, 32, 128, 512)


def _views(rng, sizes):
    cams, times, suns, params = [], [], [], []
    for k, (w, h) in enumerate(sizes):
        p, s = M.preset_params(4)
        steps = int(rng.integers(64, 161))
        p.steps_min = p.steps_max = steps
        M.sun_from_angles(float(rng.uniform(10, 60)), float(rng.uniform(0, 360)), s, p)
        cam = ground_camera(w, h, float(rng.uniform(5, 50)), float(rng.uniform(0, 360))).ubo()
        t = M.Time()
        t.ubo.frame_count_mod16 = k % 16
        t.ubo.time[1] = 0.25 * k
        cams.append(cam)
        times.append(t.ubo)
        suns.append(s)
        params.append(p)
    return cams, times, suns, params


def test_batch_views_equal_single_view_marches_and_graph_replays_agree():
    rng = np.random.default_rng(3)
    sizes = [(96, 64), (64, 64), (200, 40), (33, 17), (128, 72)]
    cams, times, suns, params = _views(rng, sizes)
    with M.CloudStage(16, 16, M.VolumeSpec(*DIMS), kernel=A.MTC_KERNEL_FAST) as st:
        st.generate_volumes(1, 2)
        with M.Batch(st, sizes) as b:
            b.set_tuning(8, 8, 2, 0)
            b.render(cams, times, suns, params)
            first = [b.read(v) for v in range(len(sizes))]
            # replays of the captured graph, with other work in between
            for tune in ((8, 16, 4, 1), (4, 8, 2, 0), (8, 8, 2, 0)):
                b.set_tuning(*tune)
                b.render(cams, times, suns, params)
                st.synchronize()
                for v in range(len(sizes)):
                    got = b.read(v)
                    assert np.array_equal(got[0], first[v][0]) and np.array_equal(got[1], first[v][1]), (tune, v)
        for v, (w, h) in enumerate(sizes):
            with M.CloudStage(w, h, M.VolumeSpec(*DIMS), kernel=A.MTC_KERNEL_FAST) as one:
                one.generate_volumes(1, 2)
                one.march_full(cams[v], times[v], suns[v], params[v])
                rgba, mask = one.read_output()
            assert np.array_equal(rgba, first[v][0]), v
            assert np.array_equal(mask, first[v][1]), v


def test_batch_autotune_and_accuracy_vs_fp64():
    rng = np.random.default_rng(5)
    sizes = [(256, 256)] * 2 + [(128, 128)] * 4 + [(256, 144)] * 4 + [(480, 270)]
    cams, times, suns, params = _views(rng, sizes)
    with M.CloudStage(16, 16, M.VolumeSpec(*DIMS), kernel=A.MTC_KERNEL_FAST) as st:
        st.generate_volumes(1, 2)
        vol = O.Volumes(*st.get_volumes(), *DIMS)
        with M.Batch(st, sizes) as b:
            best, ms = b.autotune(cams, times, suns, params)
            assert ms > 0 and best == b.tuning()
            print(f"autotuned {best}: {ms:.3f} ms for {sum(w * h for w, h in sizes)} px")
            b.render(cams, times, suns, params)
            for v in (0, 3, 7, 10):
                w, h = sizes[v]
                rgba, _ = b.read(v)
                xy = np.stack(np.meshgrid(np.arange(0, w, 3), np.arange(0, h, 3)), -1).reshape(-1, 2).astype(np.int32)
                ref, _, _, _ = O.march_pixels(vol, cams[v], times[v], suns[v], params[v], w, h, xy, 64)
                got = rgba[xy[:, 1], xy[:, 0], :3]
                keep = O.outside_sun_disk(cams[v], times[v], params[v], xy, w, h)
                p = O.psnr_vs_fp64(got, ref[:, :3], keep)
                print(f"view {v} {w}x{h}: PSNR {p:.2f} dB")
                assert p >= 45.0


def test_batch_rejects_bad_input():
    with M.CloudStage(16, 16, M.VolumeSpec(32, 16, 32, 64)) as plain:
        with pytest.raises(M.MtcError) as e:
            M.Batch(plain, [(8, 8)])
        assert e.value.code == A.MTC_ERR_UNSUPPORTED
    with M.CloudStage(16, 16, M.VolumeSpec(32, 16, 32, 64), kernel=A.MTC_KERNEL_FAST) as st:
        with pytest.raises(M.MtcError):
            M.Batch(st, [(0, 8)])
        with M.Batch(st, [(8, 8)]) as b:
            with pytest.raises(M.MtcError):
                b.set_tuning(8, 5, 2, 0)
            p, s = M.preset_params(4)
            cam = ground_camera(8, 8).ubo()
            t = M.Time()
            with pytest.raises(M.MtcError) as e:
                b.render([cam], [t.ubo], [s], [p])  # no volumes yet
            assert e.value.code == A.MTC_ERR_STATE
