"""The roofline's FLOP weights come from the oracle's instrumented operation
count, not from hand-weighted constants.

oracle.flop_counts() runs the march and the reprojection of a sample of the
bench workload's rays through the oracle's templates with an
operation-counting scalar (oracle/flop_count.hpp: add, sub, mul, div, sqrt
and transcendentals 1, fma 2) and attributes each operation to its unit of
work (per-ray set-up by ray kind, primary step, in-cloud step, cone sample,
cone erosion, reprojected pixel; cloudRayMarch.comp:565-826,
reprojection.comp:193-244).  bench.FLOPS must equal the counted averages.
"""
import os
import sys

import numpy as np

import meteoros_amd as M
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _counts(stride, offset):
    sys.path.insert(0, ROOT)
    import bench

    seq = bench.make_sequence(3)
    cur, old, tm = seq[2]
    params, sun = M.preset_params(bench.PRESET)
    vol = O.generate_volumes(128, 32, 128, 512, threads=8)
    xy = O.select_pixels(bench.W, bench.H, tm.frame_count_mod16)[offset::stride]
    return bench, O.flop_counts(vol, cur, old, tm, sun, params, bench.W, bench.H, xy)


def test_bench_flop_weights_are_the_counted_ones():
    bench, c = _counts(7, 0)
    for unit, (ops, n) in c.items():
        assert n > 0, unit
        print(f"{unit}: {ops / n:.3f} operations per unit over {n} units (bench {bench.FLOPS[unit]})")
        assert bench.FLOPS[unit] == round(ops / n), unit
    assert set(bench.FLOPS) == set(c)


def test_step_costs_do_not_depend_on_the_rays_sampled():
    """Per-step units cost the same on another sample of rays (the only
    data-dependent difference is the loop increment a step that exits
    early does not execute)."""
    _, a = _counts(7, 0)
    _, b = _counts(11, 3)
    for unit in ("primary_step", "cloud_step", "cone_sample", "cone_erosion", "reproject_pixel"):
        ma, mb = a[unit][0] / a[unit][1], b[unit][0] / b[unit][1]
        assert abs(ma - mb) < 0.01, (unit, ma, mb)
        assert abs(ma - round(ma)) < 0.01, (unit, ma)


def test_march_flops_combines_counters_and_weights():
    sys.path.insert(0, ROOT)
    import bench

    one = {"rays_marched": 1, "rays_ocean": 2, "rays_sky": 3, "primary_steps": 4, "cloud_steps": 5,
           "cone_samples": 6, "cone_erosions": 7}
    F = bench.FLOPS
    assert bench.march_flops(one) == (F["ray_marched"] + 2 * F["ray_ocean"] + 3 * F["ray_sky"]
                                      + 4 * F["primary_step"] + 5 * F["cloud_step"] + 6 * F["cone_sample"]
                                      + 7 * F["cone_erosion"])
