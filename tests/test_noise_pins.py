"""Pins of the noise generator (oracle/noise_oracle.hpp), on the CPU.

The reference has no generator: it ships pre-authored volumes
(src/CloudScapes/textures/CloudTextures/, loaded by Sky.cpp:28-57 through
ImageLoadingUtility.cpp:75-139).  The generator is this build's own
specification, so it is pinned by what the shipped data and the mathematics
fix, not by its own formulas:

* channel statistics of the generated LF and HF volumes against the
  shipped volumes' (tests/golden/texture_stats.json, computed from the
  reference's data files by tests/golden/make_texture_stats.py);
* tileability: the REPEAT sampler (Texture3D.cpp:92-134) reads across the
  wrap seam, so the difference across it must look like any interior
  neighbour difference, on every axis;
* the weather map's cloud type takes only the values 0, 1/2, 1 and its
  blue channel is empty (the shader reads r and g only, :515-525);
* the curl map is the curl of a potential, so its discrete divergence
  vanishes up to 8-bit quantisation while its vorticity does not;
* determinism and seed dependence.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
STATS = json.load(open(os.path.join(HERE, "golden", "texture_stats.json")))


def _vol(kind, n, seed=1):
    a = O.noise(kind, n, seed=seed, threads=8)
    return a.reshape(n, n, n, 4) if kind <= 1 else a.reshape(n, n, 4)


@pytest.fixture(scope="module")
def lf():
    return _vol(0, 128)


@pytest.fixture(scope="module")
def hf():
    return _vol(1, 32)


def _moments(v, channels):
    a = v.reshape(-1, 4)[:, channels].astype(np.float64) / 255.0
    return a.mean(0), a.std(0)


@pytest.mark.parametrize("name,channels", [("lf", [0, 1, 2, 3]), ("hf", [0, 1, 2])])
def test_channel_statistics_match_the_shipped_volumes(name, channels, lf, hf):
    v = lf if name == "lf" else hf
    mean, std = _moments(v, channels)
    ref_mean = np.array(STATS[name]["mean"])[channels]
    ref_std = np.array(STATS[name]["std"])[channels]
    print(name, "mean", mean.round(4), "shipped", ref_mean.round(4), "std", std.round(4), "shipped", ref_std.round(4))
    assert np.all(np.abs(mean - ref_mean) <= 0.01)
    assert np.all(np.abs(std - ref_std) <= 0.01)


def test_hf_alpha_is_empty_like_the_shipped_volume(hf):
    assert STATS["hf"]["max"][3] == 0.0
    assert not hf[..., 3].any()


def _seams(v, axis):
    """Mean |difference| across the wrap seam and across each interior pair
    of neighbouring planes, per channel."""
    a = np.moveaxis(v.astype(np.float64) / 255.0, axis, 0)
    inner = np.abs(np.diff(a, axis=0)).reshape(a.shape[0] - 1, -1, 4).mean(1)
    wrap = np.abs(a[0] - a[-1]).reshape(-1, 4).mean(0)
    return wrap, inner


@pytest.mark.parametrize("name,channels", [("lf", [0, 1, 2, 3]), ("hf", [0, 1, 2])])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_volumes_tile_across_the_wrap_seam(name, channels, axis, lf, hf):
    """Array axes are (z, y, x).  The seam is one more pair of neighbouring
    planes: its mean |difference| lies within the range the interior pairs
    span (10% slack), channel by channel, and within 10% of the interior
    mean for the 128^3 volume (one seam plane of 32^2 texels is too small a
    sample for that).  A volume that does not tile differs across the seam
    like two unrelated planes (the shipped curl map: 6x its interior)."""
    v = lf if name == "lf" else hf
    wrap, inner = _seams(v, axis)
    wrap, inner = wrap[channels], inner[:, channels]
    print(name, "axis", axis, "wrap", wrap.round(4), "interior", inner.mean(0).round(4), "range",
          inner.min(0).round(4), inner.max(0).round(4))
    assert np.all(wrap <= 1.1 * inner.max(0)) and np.all(wrap >= 0.9 * inner.min(0))
    if name == "lf":
        assert np.all(np.abs(wrap - inner.mean(0)) <= 0.1 * inner.mean(0))


def test_weather_type_is_three_valued_and_blue_is_empty():
    w = _vol(3, 512, seed=2)
    assert set(np.unique(w[..., 1]).tolist()) == {0, 128, 255}
    assert not w[..., 2].any()
    assert np.all(w[..., 3] == 255)
    # coverage spans the unit range and is smooth: a smoothstep blend of a
    # 16^2 lattice changes by at most 1.5 x 255 / 32 codes per texel at 512^2
    r = w[..., 0].astype(np.int32)
    assert r.min() < 30 and r.max() > 225
    slope = 1.5 * 255.0 / 32.0 + 1.0
    assert np.abs(np.diff(r, axis=1)).max() <= slope and np.abs(np.diff(r, axis=0)).max() <= slope
    # ... and tiles: the wrap difference obeys the same bound
    assert np.abs(r[:, 0] - r[:, -1]).max() <= slope and np.abs(r[0] - r[-1]).max() <= slope


def test_curl_map_is_divergence_free():
    """(r, g) encode c = 2 * code / 255 - 1 of the curl of a periodic
    potential: the central-difference divergence of c cancels term by term
    (d/dx d/dy psi - d/dy d/dx psi), so only the 8-bit quantisation remains
    (each code is within half a step, 1/255 of c); the vorticity, minus the
    potential's Laplacian, does not vanish."""
    m = _vol(2, 128)
    c = m[..., :2].astype(np.float64) * (2.0 / 255.0) - 1.0
    cx, cy = c[..., 0], c[..., 1]  # array axes (y, x)
    unclipped = (m[..., 0] > 0) & (m[..., 0] < 255) & (m[..., 1] > 0) & (m[..., 1] < 255)
    assert unclipped.mean() > 0.99
    dx = lambda f: 0.5 * (np.roll(f, -1, axis=1) - np.roll(f, 1, axis=1))  # noqa: E731
    dy = lambda f: 0.5 * (np.roll(f, -1, axis=0) - np.roll(f, 1, axis=0))  # noqa: E731
    div = dx(cx) + dy(cy)
    vort = dx(cy) - dy(cx)
    quant = 2.0 * (1.0 / 255.0)  # |d/dx| + |d/dy| of half-step errors
    print("rms divergence", np.sqrt(np.mean(div ** 2)), "rms vorticity", np.sqrt(np.mean(vort ** 2)))
    assert np.abs(div[unclipped]).max() <= quant + 1e-9
    assert np.sqrt(np.mean(vort ** 2)) > 5.0 * np.sqrt(np.mean(div ** 2))
    # the curl map spans a useful range of both components
    assert cx.std() > 0.05 and cy.std() > 0.05


def test_generator_is_deterministic_and_seeded():
    a = O.noise(0, 32, seed=1)
    b = O.noise(0, 32, seed=1)
    c = O.noise(0, 32, seed=7)
    assert np.array_equal(a, b)
    assert (a != c).mean() > 0.5
    # a slab generated on its own equals the same slab of the whole volume
    whole = O.noise(0, 32, seed=1).reshape(32, 32, 32, 4)
    slab = O.noise(0, 32, seed=1, z0=8, nz=4).reshape(4, 32, 32, 4)
    assert np.array_equal(slab, whole[8:12])
