"""Shared scene set-up of the tests (camera poses, presets, tolerances)."""
import numpy as np

import meteoros_amd as M
from oracle import oracle as O


def ground_camera(W, H, pitch_deg=20.0, yaw_deg=0.0, eye=(0.0, 0.0, 0.0)):
    """Eye at the origin (altitude 0), looking down -z pitched up."""
    p, y = np.radians(pitch_deg), np.radians(yaw_deg)
    d = (np.cos(p) * np.sin(y), np.sin(p), -np.cos(p) * np.cos(y))
    ref = tuple(float(eye[i] + d[i]) for i in range(3))
    return M.Camera(eye, ref, W, H, 45.0)


def within_rel(got, ref, rel=1e-3, floor_frac=1e-3):
    """|g - o| <= rel * max(|o|, floor_frac * peak) elementwise; returns the
    boolean array of failures."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    peak = float(np.abs(ref).max()) if ref.size else 0.0
    tol = rel * np.maximum(np.abs(ref), floor_frac * peak)
    return np.abs(got - ref) > tol


def psnr(test, ref, peak=None):
    test = np.asarray(test, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if peak is None:
        peak = float(ref.max())
    mse = float(np.mean((test - ref) ** 2))
    if mse == 0.0:
        return float("inf")
    return 10.0 * np.log10(peak * peak / mse)


_VOL_CACHE = {}


def volumes(lf=128, hf=32, curl=128, weather=512, fmt=0, seed_volumes=1, seed_weather=2):
    key = (lf, hf, curl, weather, fmt, seed_volumes, seed_weather)
    if key not in _VOL_CACHE:
        _VOL_CACHE[key] = O.generate_volumes(lf, hf, curl, weather, fmt, seed_volumes, seed_weather)
    return _VOL_CACHE[key]
