"""The oracle's pins must reject plausible mistakes (CPU only).

Each case plants one mistake in a copy of the oracle's source (a swapped
weight, a sign flip, a dropped term, a wrong sample point), builds that copy
on the CPU, and runs the pins that are meant to catch it against the broken
build (METEOROS_ORACLE_LIB points oracle.oracle at it).  The pins must fail;
the same pins must pass against an unmodified copy built the same way.
Nothing here touches the product library's kernels or a GPU.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

MARCH = "tests/test_march_pins.py"
CLOSED = MARCH + "::test_constant_volume_march_is_the_closed_form_recurrence"
TERMS = MARCH + "::test_cone_terms_and_transmittance_follow_the_step_values"
WIND = MARCH + "::test_wind_skew_moves_the_lf_sample_point"
CONE = MARCH + "::test_cone_samples_read_the_lf_unskewed_and_undivided"
ERO = MARCH + "::test_erosion_reads_curl_at_the_skewed_point_and_hf_at_the_curled_point"
GEOM = MARCH + "::test_step_positions_relative_position_and_height"
MASK = MARCH + "::test_mask_of_rays_below_0_05_is_scaled_by_five"
NOISE = "tests/test_noise_pins.py"

# (name, file, text, replacement, pins that must fail)
MUTATIONS = [
    ("lf fbm weights .625/.25 swapped", "cloud_oracle.hpp",
     "T fbm = (n.g * T(0.625)) + (n.b * T(0.25))", "T fbm = (n.g * T(0.25)) + (n.b * T(0.625))", [CLOSED]),
    ("hf fbm weights .625/.25 swapped", "cloud_oracle.hpp",
     "T fbm = (hn.r * T(0.625)) + (hn.g * T(0.25))", "T fbm = (hn.r * T(0.25)) + (hn.g * T(0.625))",
     [CLOSED, TERMS]),
    ("coverage factor dropped", "cloud_oracle.hpp", "  bc = bc * cov;\n", "\n", [CLOSED, TERMS]),
    ("cone reads the LF at the skewed point", "cloud_oracle.hpp",
     "T cb = sample_low(f, vol, slp, slp, h, &cls_i);", "T cb = sample_low(f, vol, sk, slp, h, &cls_i);", [CONE]),
    ("cone point divided by 8", "cloud_oracle.hpp",
     "v3<T> slp = rel_pos(f, lp, c);", "v3<T> slp = rel_pos(f, lp, c) / f.sample_divisor;", [CONE]),
    ("cone sum without term 0", "cloud_oracle.hpp",
     "for (int i = 0; i < f.light_samples; ++i) {\n        Region<T>::set(FR_CONE);",
     "for (int i = 1; i < f.light_samples; ++i) {\n        Region<T>::set(FR_CONE);", [CLOSED]),
    ("cone erosion at the primary scale", "cloud_oracle.hpp",
     "T e = erode(f, vol, f.cone_erosion_scale * cb, sk, h);", "T e = erode(f, vol, f.erosion_scale * cb, sk, h);",
     [CLOSED, TERMS]),
    ("wind skew sign flipped", "cloud_oracle.hpp",
     "p = p + ((h * f.wind_dir) * f.top_offset) * f.wind_skew;",
     "p = p - ((h * f.wind_dir) * f.top_offset) * f.wind_skew;", [WIND]),
    ("wind drift sign flipped", "cloud_oracle.hpp",
     "p = p + ((f.wind_dir + V(T(0), f.wind_bias, T(0))) * f.cloud_speed) * f.total_time;",
     "p = p - ((f.wind_dir + V(T(0), f.wind_bias, T(0))) * f.cloud_speed) * f.total_time;", [WIND]),
    ("wind upward bias on z", "cloud_oracle.hpp",
     "V(T(0), f.wind_bias, T(0))", "V(T(0), T(0), f.wind_bias)", [WIND]),
    ("accumulation scale dropped", "cloud_oracle.hpp",
     "accum += hfd * f.accum_scale;", "accum += hfd;", [CLOSED, TERMS]),
    ("transmittance mix weight reversed", "cloud_oracle.hpp",
     "transmittance = mixT(transmittance, E, T(1) - accum);", "transmittance = mixT(transmittance, E, accum);",
     [CLOSED, TERMS]),
    ("no early exit", "cloud_oracle.hpp", "if (accum >= T(1)) {", "if (accum >= T(2)) {", [CLOSED]),
    ("relative height without abs", "cloud_oracle.hpp",
     "T num = m_fabs(cosT * (lr - ls));", "T num = cosT * (lr - ls);", [GEOM]),
    ("relative position without the inner offset", "cloud_oracle.hpp",
     "return (pos - V(c.x, f.inner_off, c.z)) / f.thickness;", "return (pos - V(c.x, T(0), c.z)) / f.thickness;",
     [GEOM]),
    ("curl displacement sign flipped", "cloud_oracle.hpp",
     "p.x = p.x + (cn.r * k) * T(0.5);", "p.x = p.x - (cn.r * k) * T(0.5);", [ERO]),
    ("curl displacement not height-weighted", "cloud_oracle.hpp", "T k = T(1) - h;", "T k = T(1);", [ERO]),
    ("erosion height blend from h, not 2h", "cloud_oracle.hpp",
     "clampT(h * T(2), T(0), T(1))", "clampT(h, T(0), T(1))", [CLOSED, TERMS]),
    ("horizon mask multiplier dropped", "cloud_oracle.hpp", "if (ray.d.y < T(0.05)) {", "if (ray.d.y < T(-1)) {",
     [MASK]),
    ("curl map sign flipped (not a curl)", "noise_oracle.hpp",
     "float cx = dpdy * 0.05f, cy = -dpdx * 0.05f;", "float cx = dpdy * 0.05f, cy = dpdx * 0.05f;",
     [NOISE + "::test_curl_map_is_divergence_free"]),
    ("worley cells not wrapped in x", "noise_oracle.hpp",
     "        uint32_t h = hash3((uint32_t)wrapi(ix + dx, cells),", "        uint32_t h = hash3((uint32_t)(ix + dx),",
     [NOISE + "::test_volumes_tile_across_the_wrap_seam"]),
    ("LF worley channels over-stretched", "noise_oracle.hpp",
     "out[1] = clamp01((w4 - 0.426f) * 0.77f + 0.687f);", "out[1] = clamp01((w4 - 0.426f) * 0.97f + 0.687f);",
     [NOISE + "::test_channel_statistics_match_the_shipped_volumes"]),
    ("weather type four-valued", "noise_oracle.hpp",
     "float type = (float)(th % 3u) * 0.5f;", "float type = (float)(th % 4u) * (1.0f / 3.0f);",
     [NOISE + "::test_weather_type_is_three_valued_and_blue_is_empty"]),
]

CXX = ["g++", "-O1", "-std=c++17", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
       "-shared"]


def _build(work, name, mutation):
    """Copy the oracle's sources into work/<name>/oracle (with the header
    beside it as ../include), plant the mutation, build liboracle.so."""
    d = os.path.join(work, name)
    src = os.path.join(d, "oracle")
    os.makedirs(src)
    shutil.copytree(os.path.join(ROOT, "include"), os.path.join(d, "include"))
    for f in ("oracle_capi.cpp", "cloud_oracle.hpp", "flop_count.hpp", "noise_oracle.hpp", "post_oracle.hpp"):
        shutil.copy(os.path.join(ORACLE, f), src)
    if mutation is not None:
        fname, text, repl = mutation
        path = os.path.join(src, fname)
        body = open(path).read()
        assert body.count(text) == 1, f"{name}: the text to mutate is not in {fname} exactly once"
        open(path, "w").write(body.replace(text, repl))
    lib = os.path.join(d, "liboracle.so")
    subprocess.run(CXX + ["-o", lib, os.path.join(src, "oracle_capi.cpp")], check=True, capture_output=True)
    return lib


def _pins(lib, nodes):
    env = dict(os.environ, METEOROS_ORACLE_LIB=lib)
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                           *nodes], cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)


def test_every_planted_mistake_fails_its_pins(tmp_path):
    work = str(tmp_path)
    jobs = [("control", None)] + [(f"m{i:02d}", (m[1], m[2], m[3])) for i, m in enumerate(MUTATIONS)]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        libs = dict(zip([j[0] for j in jobs], ex.map(lambda j: _build(work, *j), jobs)))
    all_nodes = sorted({n for m in MUTATIONS for n in m[4]})
    control = _pins(libs["control"], all_nodes)
    assert control.returncode == 0, "unmodified oracle copy fails its pins:\n" + control.stdout[-3000:]

    def run(i):
        name, _, _, _, nodes = MUTATIONS[i]
        return name, _pins(libs[f"m{i:02d}"], nodes)

    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        results = list(ex.map(run, range(len(MUTATIONS))))
    survived = []
    for name, r in results:
        # exit status 1: tests ran and some failed (anything else is a
        # harness error, not a caught mistake)
        print(f"{name}: {'caught' if r.returncode == 1 else 'NOT caught (rc %d)' % r.returncode}")
        if r.returncode != 1:
            survived.append((name, r.returncode, r.stdout[-1500:]))
    assert not survived, survived
