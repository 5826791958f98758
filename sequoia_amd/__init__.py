"""MI355X-native BVH build and ray traversal (gfx950 HIP kernels behind a C ABI).

The compute path is libsequoia_amd.so (include/sequoia.h); this package is a
thin ctypes mirror of that ABI for Python callers, tests and benchmarks.
"""
from ._lib import (  # noqa: F401
    BUILDERS, BVHK_NO_COLLAPSE, ERR_BUDGET, ERR_CANCELLED, ERR_INVALID, ERR_TIMEOUT, HEATMAP_BUNDLE, HEATMAP_PACKET, HEATMAP_RAY,
    KERNEL_AUTO, KERNEL_PACKETS, KERNEL_RAYS, LAYOUT_FIXED, LAYOUT_REFERENCE, MISS, NODE2_DTYPE, NODE_DTYPE,
    OUT_ALL, OUT_COUNTS, OUT_HIT_ID, OUT_HIT_T, OUT_RGB, PLUCKER_DTYPE, RAY_DTYPE, REQ_DONE, REQ_FAILED,
    REQ_QUEUED, REQ_RUNNING, TRACE_CLOSEST, TRACE_EXHAUSTIVE, Batch, Bvh, Bvh2, BvhK, Camera, Context, Scene, Server,
    SqError, as_tris, encode_bmp, gen_clusters, gen_icosphere, gen_secondary_rays, gen_soup, lib, load_obj,
    nodek_dtype, pinhole_camera, reference_camera, write_bmp, write_obj,
)
