"""MI355X-native smart interpolation (drop-in for biomedisa.interpolation).

    from biomedisa_amd import smart_interpolation
    results = smart_interpolation(img, labels, nbrw=10, sorw=4000, allaxis=False, smooth=0)

The compute runs in libbmx_hip.so (hand-written HIP kernels for gfx950, C ABI
in include/bmx.h); this package holds the host side.
"""
from .interpolation import (smart_interpolation, smart_interpolation_device, run_pipeline, DeviceArray,  # noqa: F401
                            DeviceVolume)
from .batch import smart_interpolation_batch  # noqa: F401
from . import helpers  # noqa: F401

__all__ = ["smart_interpolation", "smart_interpolation_device", "smart_interpolation_batch", "run_pipeline", "DeviceArray",
           "DeviceVolume", "helpers"]
