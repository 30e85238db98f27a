"""MI355X-native NUFFT (gfx950 HIP) behind the ``@nufft_3d`` operator API.

``from nufft_3d_amd import nufft_3d`` gives the operator class; the HIP
library is loaded on first use and its absence is an error.
"""
from .nufft_3d import nufft_3d  # noqa: F401
