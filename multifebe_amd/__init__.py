"""MI355X-native BEM influence-matrix assembly and dense solve (MultiFEBE hot path).

The compute path is libmfbem.so (hand-written HIP for gfx950, C ABI in include/mfbem.h); the
Python modules here mirror the reference's interface on top of it.
"""
from .mesh import Mesh, cube, sphere, cylinder, static_cube_bcs, pressure_hemisphere_bcs, elastic_constants  # noqa: F401

__all__ = ["Mesh", "cube", "sphere", "cylinder", "static_cube_bcs", "pressure_hemisphere_bcs",
           "elastic_constants"]
