"""MI355X-native BVH build and ray traversal (gfx950 HIP kernels behind a C ABI).

The compute path is libsequoia_amd.so (include/sequoia.h); this package is a
thin ctypes mirror of that ABI for Python callers, tests and benchmarks.
"""
from ._lib import (  # noqa: F401
    BUILDERS, OUT_ALL, OUT_COUNTS, OUT_HIT_ID, OUT_HIT_T, OUT_RGB, Batch, HEATMAP_BUNDLE, HEATMAP_PACKET, HEATMAP_RAY, LAYOUT_FIXED, LAYOUT_REFERENCE, MISS, NODE_DTYPE,
    RAY_DTYPE, TRACE_CLOSEST, TRACE_EXHAUSTIVE, Bvh, Camera, Context, Scene, SqError, as_tris, gen_clusters,
    gen_icosphere, gen_secondary_rays, gen_soup, lib, pinhole_camera, reference_camera,
)
