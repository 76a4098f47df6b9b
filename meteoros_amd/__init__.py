"""meteoros_amd -- MI355X-native cloud compute stage of Meteoros.

The stage (noise volumes -> ray march -> 4x4 reprojection, HDR RGBA32F plus
the god-ray mask) lives in the C-ABI library libmeteoros_cloud.so, built
from meteoros_amd/csrc for gfx950.  This package only loads it and mirrors
the reference's host objects (Camera, Scene time, Sky/Renderer) over it.
"""
from ._abi import LIB_PATH, MtcError, lib  # noqa: F401
from .stage import (Batch, Camera, CloudStage, LightNet, Time, VolumeSpec, device_count, halton,  # noqa: F401
                    load_image_rgba8, load_shipped_volumes, load_tga_slices, preset_params, reference_params,
                    sun_from_angles)

__all__ = ["Batch", "Camera", "CloudStage", "LightNet", "Time", "VolumeSpec", "device_count", "halton", "preset_params",
           "reference_params", "sun_from_angles", "load_image_rgba8", "load_tga_slices", "load_shipped_volumes", "lib",
           "LIB_PATH", "MtcError"]
