"""libfusion_amd — MI355X (gfx950) dense-reconstruction hot path of libfusion.

The native library libfusion_amd.so holds the HIP kernels, the C-ABI
(include/libfusion_amd.h) and the C++ drop-in classes (include/fusion_amd.hpp).
This package is the Python-side binding (capi) plus the seeded synthetic RGB-D
generator used by the tests and the benchmark (synth).
"""
from . import capi  # noqa: F401

__all__ = ["capi", "synth", "configs"]
