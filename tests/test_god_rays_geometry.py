"""The geometry the LDS-staged god-ray blur relies on, checked on the CPU by
replaying the kernel's FP32 sample walks (numpy float32: one rounding per
operation, as the kernel's `-ffp-contract=off` build computes them).

The blur (meteoros_amd/csrc/post.hip, god_rays_kernel) stages, per
workgroup of 64 columns x 16 rows and per sample, an 18-row x 72-column
region of the decoded mask map: rows from the group's lowest texel row,
columns from the first lane's column aligned down to 4.  Every tap must
read inside that region and the region inside the padded map, for every
image size and sun position:

  * the 64 lanes' texel columns never decrease across lanes (the region
    starts at lane 0's column);
  * column offset + 1 <= 71 and row offset + 1 <= 17 in the region;
  * the staged rows and columns lie inside the map's (H + 19) x pitch.

The walks follow postProcess_GodRays.frag:96-130: uv starts at the pixel
centre, delta_uv = ((uv - sun) / 100) * 1.0 and uv -= delta_uv after each
of the 100 samples; the sampler reads at uv * dim - 0.5.
"""
import numpy as np
import pytest

F = np.float32
SAMPLES, GROUP_ROWS, STAGE_ROWS, STAGE_COLS = 100, 16, 18, 72


def map_pitch(W):
    return (W + 76 + 3) // 4 * 4


def walk(n, dim, sun):
    """Texel index floor(uv * dim - 0.5), clamped to [-1, dim - 1], of the
    100 samples of positions 0..n-1 along one axis: shape (100, n)."""
    p = np.minimum(np.arange(n), dim - 1).astype(F)
    u = (p + F(0.5)) / F(dim)
    du = ((u - F(sun)) / F(SAMPLES)) * F(1.0)
    out = np.empty((SAMPLES, n), np.int64)
    for i in range(SAMPLES):
        x = u * F(dim) - F(0.5)
        out[i] = np.clip(np.floor(x).astype(np.int64), -1, dim - 1)
        u = (u - du).astype(F)
    return out


def check(W, H, su, sv):
    # columns: every 64-lane block of the image (lanes right of it walk the
    # last column's line)
    cols = walk(((W + 63) // 64) * 64, W, su).reshape(SAMPLES, -1, 64)
    assert np.all(np.diff(cols, axis=2) >= 0), "lane columns decrease"
    cs = (cols[:, :, :1] + 4) & ~3  # stored column of the region's start
    rel_c = cols + 4 - cs
    assert rel_c.min() >= 0 and rel_c.max() + 1 <= STAGE_COLS - 1
    assert (cs + STAGE_COLS - 1).max() < map_pitch(W)
    # rows: every 16-row group (rows past the image repeat the last row)
    groups = (H + GROUP_ROWS - 1) // GROUP_ROWS
    rows = walk(groups * GROUP_ROWS, H, sv).reshape(SAMPLES, groups, GROUP_ROWS)
    lo = rows.min(axis=2, keepdims=True)
    rel_r = rows - lo
    assert rel_r.max() + 1 <= STAGE_ROWS - 1
    assert (lo + 1 + STAGE_ROWS - 1).max() < H + 19  # stored rows of the padded map
    return int(rel_r.max()), int(rel_c.max())


SUNS = [(0.0, 0.0), (0.5, 0.0), (1.0, 0.0), (0.0, 0.5), (0.5, 0.5), (1.0, 0.5), (0.0, 1.0), (0.5, 1.0),
        (1.0, 1.0), (0.37, 0.81)]


@pytest.mark.parametrize("W,H", [(1, 1), (2, 3), (63, 5), (65, 37), (130, 17), (480, 270), (1920, 1080),
                                 (3840, 2160)])
def test_staged_region_holds_every_tap(W, H):
    worst = [check(W, H, su, sv) for su, sv in SUNS]
    # the bounds are reached, not just respected, at the larger sizes
    if W >= 480 and H >= 270:
        assert max(r for r, _ in worst) >= 15 and max(c for _, c in worst) >= 66


def test_a_wider_walk_would_leave_the_region():
    """The check can fail: an 18-row group would need 19 staged rows."""
    rows = walk(18, 1080, 0.0)
    assert (rows - rows.min(axis=1, keepdims=True)).max() + 1 > STAGE_ROWS - 1
