"""kgeo_amd: MI355X-native analytic Kerr raytracer (FP64 HIP kernels for gfx950).

Drop-in for kgeo's ``raytrace_ana`` hot path (kerr_raytracing_ana.py:46-122)
and its elliptic layer; see include/kgeo.h for the C-ABI.
"""

__version__ = "0.1.0"

from ._lib import KgeoError  # noqa: F401
