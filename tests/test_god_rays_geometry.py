"""The geometry the LDS-staged god-ray blur relies on, checked on the CPU by
replaying the kernel's FP32 sample walks (numpy float32: one rounding per
operation, as the kernel's `-ffp-contract=off` build computes them).

The blur (meteoros_amd/csrc/post.hip, god_rays_kernel) stages, per
workgroup of 64 columns x 16 rows and per sample, a region of the decoded
mask map: rows from the group's lowest texel row, columns from the first
lane's column aligned down to 4.  The region is 18 rows x 72 columns for
samples 0-24 and shrinks in three steps (14 x 56, 10 x 40, 7 x 24 for
samples 25-49, 50-74, 75-99), since sample i of adjacent pixels lies
1 - i / 100 texels apart.  Every tap must read inside its sample's region
and the region inside the padded map, for every image size (up to the
16384 the library accepts) and sun position (the host clamps it to
[0, 1]^2):

  * no lane's texel column lies more than one column below lane 0's (the
    region starts one column before lane 0's; the FP32 walk does put a lane
    one column below lane 0's at widths of 8192 and more);
  * column offset + 1 <= columns - 1 and row offset + 1 <= rows - 1 in the
    region of the sample's phase;
  * the staged rows and columns lie inside the map's (H + 19) x pitch.

The walks follow postProcess_GodRays.frag:96-130: uv starts at the pixel
centre, delta_uv = ((uv - sun) / 100) * 1.0 and uv -= delta_uv after each
of the 100 samples; the sampler reads at uv * dim - 0.5.
"""
import numpy as np
import pytest

F = np.float32
SAMPLES, GROUP_ROWS, STAGE_ROWS, STAGE_COLS = 100, 16, 18, 72
# kGodRayPhaseRows / kGodRayPhaseBlocks (meteoros_amd/csrc/runtime_internal.h)
PHASE_ROWS, PHASE_COLS = (18, 14, 10, 7), (72, 56, 40, 24)
PHASE = 25


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


def offsets(W, H, su, sv):
    """Per sample: the largest row and column offset of a tap in its
    workgroup's region, and the region's first stored column and row."""
    # columns: every 64-lane block of the image (lanes right of it walk the
    # last column's line)
    cols = walk(((W + 63) // 64) * 64, W, su).reshape(SAMPLES, -1, 64)
    assert (cols - cols[:, :, :1]).min() >= -1, "a lane two columns below lane 0"
    cs = (cols[:, :, :1] + 3) & ~3  # stored column of the region's start
    rel_c = cols + 4 - cs
    assert rel_c.min() >= 0
    # rows: every 16-row group (rows past the image repeat the last row)
    groups = (H + GROUP_ROWS - 1) // GROUP_ROWS
    rows = walk(groups * GROUP_ROWS, H, sv).reshape(SAMPLES, groups, GROUP_ROWS)
    lo = rows.min(axis=2, keepdims=True)
    rel_r = rows - lo
    return rel_r.max(axis=(1, 2)), rel_c.max(axis=(1, 2)), cs.max(), lo.max()


def check(W, H, su, sv, phase_rows=PHASE_ROWS, phase_cols=PHASE_COLS):
    rel_r, rel_c, cs_max, lo_max = offsets(W, H, su, sv)
    for p in range(4):
        s = slice(PHASE * p, PHASE * (p + 1))
        assert rel_c[s].max() + 1 <= phase_cols[p] - 1, (p, int(rel_c[s].max()))
        assert rel_r[s].max() + 1 <= phase_rows[p] - 1, (p, int(rel_r[s].max()))
    # the first phase's region is the largest and every region starts at
    # the same column and row: it bounds what any phase reads
    assert cs_max + STAGE_COLS - 1 < map_pitch(W)
    assert lo_max + 1 + STAGE_ROWS - 1 < H + 19  # stored rows of the padded map
    return int(rel_r.max()), int(rel_c.max())


SUNS = [(0.0, 0.0), (0.5, 0.0), (1.0, 0.0), (0.0, 0.5), (0.5, 0.5), (1.0, 0.5), (0.0, 1.0), (0.5, 1.0),
        (1.0, 1.0), (0.37, 0.81)]


@pytest.mark.parametrize("W,H", [(1, 1), (2, 3), (63, 5), (65, 37), (130, 17), (480, 270), (1920, 1080),
                                 (3840, 2160), (16384, 9)])
def test_staged_region_holds_every_tap(W, H):
    worst = [check(W, H, su, sv) for su, sv in SUNS]
    # the bounds are reached, not just respected, at the larger sizes
    if W >= 480 and H >= 270:
        assert max(r for r, _ in worst) >= 15 and max(c for _, c in worst) >= 66


def test_largest_images():
    """16384 pixels along each axis, where the walk's FP32 rounding is
    largest (a texel is 2^-14 of uv), with the sun at the corners and
    edges of the screen."""
    for su, sv in [(0.0, 0.0), (1.0, 1.0), (0.5, 0.0), (1.0, 0.5), (0.37, 0.81)]:
        check(16384, 16384, su, sv)


def test_a_wider_walk_would_leave_the_region():
    """The check can fail: an 18-row group would need 19 staged rows."""
    rows = walk(18, 1080, 0.0)
    assert (rows - rows.min(axis=1, keepdims=True)).max() + 1 > STAGE_ROWS - 1


def test_lanes_can_fall_one_column_behind_lane_0():
    """Why the region starts one column early: at 8192 columns the FP32
    walk puts a lane one column below lane 0's (sample 99, sun at 0.9)."""
    cols = walk(8192, 8192, 0.9).reshape(SAMPLES, -1, 64)
    assert (cols - cols[:, :, :1]).min() == -1


@pytest.mark.parametrize("p", [1, 2, 3])
def test_a_smaller_phase_region_would_miss_taps(p):
    """Each later phase's shape is close to the smallest that holds its
    taps: one 4-column block less, or one row less (two in the last phase,
    which keeps a row to spare), fails at 1080p."""
    for shrink_rows in (True, False):
        rows, cols = list(PHASE_ROWS), list(PHASE_COLS)
        if shrink_rows:
            rows[p] -= 2 if p == 3 else 1
        else:
            cols[p] -= 4
        with pytest.raises(AssertionError):
            for su, sv in SUNS:
                check(1920, 1080, su, sv, rows, cols)
