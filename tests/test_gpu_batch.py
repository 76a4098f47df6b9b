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


def _stride(w, h, target=2000):
    return max(1, int(np.sqrt(w * h / target)))


@pytest.mark.parametrize("kernel", [A.MTC_KERNEL_FAST, A.MTC_KERNEL_PRECISION])
def test_96_view_batch_every_view_vs_fp64(kernel):
    """The 96-view workload exactly as the bench renders it (bench.BATCH_SIZES
    and bench.batch_views(): 6 cube faces 512^2 at 90 deg, 40 probes 128^2,
    40 views 256x144, 10 views 1920x1080; seeded poses and suns, 64-160
    steps), autotuned: every view reaches 45 dB against the FP64 oracle on a
    strided subset of its pixels (about 2000 rays), and the first view of
    each size class equals its single-view mtc_march_full bit for bit."""
    import bench

    sizes = bench.BATCH_SIZES
    cams, times, suns, params = bench.batch_views()
    with M.CloudStage(16, 16, M.VolumeSpec(*DIMS), kernel=kernel) as st:
        st.generate_volumes(1, 2)
        vols = st.get_volumes()
        vol = O.Volumes(*vols, *DIMS)
        with M.Batch(st, sizes) as b:
            best, ms = b.autotune(cams, times, suns, params)
            print(f"kernel {kernel}: autotuned {best}, {ms:.3f} ms per 96-view render")
            b.render(cams, times, suns, params)
            st.synchronize()
            views = [b.read(v) for v in range(len(sizes))]
    bits = lambda x: np.ascontiguousarray(x).view(np.uint32)  # noqa: E731
    firsts = {}
    for v, s in enumerate(sizes):
        firsts.setdefault(s, v)
    for (w, h), v in firsts.items():
        with M.CloudStage(w, h, M.VolumeSpec(*DIMS), kernel=kernel) as one:
            one.set_volumes(*vols)
            one.march_full(cams[v], times[v], suns[v], params[v])
            rgba, mask = one.read_output()
        assert np.array_equal(bits(rgba), bits(views[v][0])), f"view {v} ({w}x{h}) colour differs from its single-view march"
        assert np.array_equal(bits(mask), bits(views[v][1])), f"view {v} ({w}x{h}) mask differs from its single-view march"
    per_class, below = {}, []
    for v, (w, h) in enumerate(sizes):
        s = _stride(w, h)
        xy = np.stack(np.meshgrid(np.arange(v % s, w, s), np.arange(v % s, h, s)), -1).reshape(-1, 2).astype(np.int32)
        ref, _, _, _ = O.march_pixels(vol, cams[v], times[v], suns[v], params[v], w, h, xy, 64)
        got = views[v][0][xy[:, 1], xy[:, 0], :3]
        keep = O.outside_sun_disk(cams[v], times[v], params[v], xy, w, h)
        e = O.error_vs_fp64(got, ref[:, :3], keep)
        per_class.setdefault(f"{w}x{h}", []).append(e)
        if not e["psnr_db"] >= 45.0:
            below.append((v, w, h, e))
    for cls, es in per_class.items():
        ps = np.array([e["psnr_db"] for e in es])
        print(f"kernel {kernel} {cls}: {len(es)} views, PSNR min {ps.min():.2f} median {np.median(ps):.2f} dB, "
              f"max err {max(e['max_err_frac_of_peak'] for e in es):.2e} of peak")
    assert not below, f"views below 45 dB: {below}"
