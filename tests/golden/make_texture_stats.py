"""Write tests/golden/texture_stats.json: channel statistics of the noise
volumes and maps that ship with the reference
(src/CloudScapes/textures/CloudTextures/, loaded by Sky.cpp:28-57 through
ImageLoadingUtility.cpp:75-139: slice z <- "Name(z+1).tga", RGBA8, x fastest).

Only data files are read (TGA/PNG decoded with PIL); no reference code is
imported or run.  The JSON holds numbers only and is what the noise
generator's statistical tests compare against.
"""
import json
import os
import sys

import numpy as np
from PIL import Image

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/CloudScapes/textures/CloudTextures"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "texture_stats.json")


def load_volume(folder, name, n):
    vol = np.zeros((n, n, n, 4), dtype=np.uint8)
    for z in range(n):
        img = Image.open(os.path.join(SRC, folder, f"{name}({z + 1}).tga")).convert("RGBA")
        vol[z] = np.asarray(img)
    return vol


def seam_stats(v, axis):
    """Mean |difference| across the wrap seam vs between interior neighbours."""
    a = v.astype(np.float64) / 255.0
    n = a.shape[axis]
    first = np.take(a, 0, axis=axis)
    last = np.take(a, n - 1, axis=axis)
    interior = np.abs(np.diff(a, axis=axis)).mean(axis=tuple(i for i in range(a.ndim - 1)))
    wrap = np.abs(first - last).mean(axis=tuple(i for i in range(first.ndim - 1)))
    return {"wrap": wrap.tolist(), "interior": interior.tolist()}


def stats(arr):
    a = arr.reshape(-1, 4).astype(np.float64) / 255.0
    return {"mean": a.mean(0).tolist(), "std": a.std(0).tolist(), "min": a.min(0).tolist(), "max": a.max(0).tolist()}


def main():
    lf = load_volume("LowFrequency", "LowFrequency", 128)   # [z, y, x, c]
    hf = load_volume("HighFrequency", "HighFrequency", 32)
    curl = np.asarray(Image.open(os.path.join(SRC, "curlNoise.png")).convert("RGBA"))
    weather = np.asarray(Image.open(os.path.join(SRC, "weatherMap.png")).convert("RGBA"))
    out = {
        "source": "src/CloudScapes/textures/CloudTextures (reference data files)",
        "lf": {"dim": 128, **stats(lf), "seam_x": seam_stats(lf, 2), "seam_y": seam_stats(lf, 1),
               "seam_z": seam_stats(lf, 0)},
        "hf": {"dim": 32, **stats(hf), "seam_x": seam_stats(hf, 2), "seam_y": seam_stats(hf, 1),
               "seam_z": seam_stats(hf, 0)},
        "curl": {"dim": curl.shape[0], **stats(curl), "seam_x": seam_stats(curl, 1), "seam_y": seam_stats(curl, 0)},
        "weather": {"dim": weather.shape[0], **stats(weather)},
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
