"""ctypes binding of libsequoia_amd.so (the C ABI declared in include/sequoia.h).

The library is the product: if it is missing or fails to load, every entry
point raises. There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsequoia_amd.so")

SQ_OK = 0
ERRORS = {-1: "INVALID", -2: "HIP", -3: "NOMEM", -4: "BUDGET", -5: "UNSUPPORTED", -6: "OVERFLOW", -7: "TIMEOUT",
          -8: "CANCELLED"}
ERR_INVALID, ERR_BUDGET, ERR_TIMEOUT, ERR_CANCELLED = -1, -4, -7, -8
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3

BUILDERS = {
    "recursive": 0, "horizontal": 1, "grid": 2, "grid_sse": 3, "grid_avx": 4,
    "morton": 5, "morton2": 6, "morton63": 7,
}
LAYOUT_REFERENCE, LAYOUT_FIXED = 0, 1
TRACE_EXHAUSTIVE, TRACE_CLOSEST = 0, 1
HEATMAP_RAY, HEATMAP_PACKET, HEATMAP_BUNDLE = 0, 1, 2
KERNEL_AUTO, KERNEL_RAYS, KERNEL_PACKETS = 0, 1, 2
MISS = 0xFFFFFFFF
BVHK_NO_COLLAPSE = 1

NODE_DTYPE = np.dtype([("min", "<f4", 3), ("count", "<u2"), ("flags", "<u2"), ("max", "<f4", 3), ("child", "<u4")])
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
NODE2_DTYPE = np.dtype([("c", "<f4", 3), ("w0", "<u4"), ("w1", "<u4"), ("q", "u1", 12)])
PLUCKER_DTYPE = np.dtype([("nu", "<f4"), ("nv", "<f4"), ("np", "<f4"), ("pu", "<f4"), ("pv", "<f4"), ("ci", "<i4"),
                          ("e0u", "<f4"), ("e0v", "<f4"), ("e1u", "<f4"), ("e1v", "<f4"), ("pad0", "<i4"),
                          ("pad1", "<i4")])
assert NODE_DTYPE.itemsize == 32 and RAY_DTYPE.itemsize == 32


class SqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sequoia error %d (%s): %s" % (code, ERRORS.get(code, "?"), msg))
        self.code = code


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("origin", Vec3), ("tan_half_fov", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32),
                ("spp", C.c_uint32), ("jitter", C.c_uint32), ("seed", C.c_uint32), ("tmin", C.c_float),
                ("tmax", C.c_float)]


class BuildParams(C.Structure):
    _fields_ = [("morton_bits_per_axis", C.c_uint32), ("radix_digit_bits", C.c_uint32), ("num_bins", C.c_uint32),
                ("grid_dim", C.c_uint32), ("layout", C.c_uint32), ("no_refit", C.c_uint32),
                ("subtree_threshold", C.c_uint32), ("reserved", C.c_uint32)]


class TraceParams(C.Structure):
    _fields_ = [("semantics", C.c_uint32), ("heatmap", C.c_uint32), ("width", C.c_uint32),
                ("block_size", C.c_uint32), ("stack_slots", C.c_uint32), ("count_work", C.c_uint32),
                ("kernel", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("tri_test", C.c_uint32),
                ("lists_per_step", C.c_uint32)]


class TraceOut(C.Structure):
    _fields_ = [("hit_id", C.c_void_p), ("hit_t", C.c_void_p), ("counts", C.c_void_p), ("rgb", C.c_void_p)]


class TuneResult(C.Structure):
    _fields_ = [("radix_digit_bits", C.c_uint32), ("subtree_threshold", C.c_uint32), ("stack_slots", C.c_uint32),
                ("block_size", C.c_uint32), ("build_ms", C.c_float), ("trace_ms", C.c_float),
                ("default_build_ms", C.c_float), ("default_trace_ms", C.c_float), ("candidates_tried", C.c_uint32),
                ("kernel", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("lists_per_step", C.c_uint32),
                ("num_bins", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("max_wave_requests", C.c_uint32), ("max_wave_rays", C.c_uint32),
                ("max_wave_tris", C.c_uint32), ("semantics", C.c_uint32), ("inflight_waves", C.c_uint32),
                ("reserved", C.c_uint32 * 3)]


class Request(C.Structure):
    _fields_ = [("key", C.c_uint64), ("tris", C.c_void_p), ("ntris", C.c_uint32), ("nrays", C.c_uint32),
                ("rays", C.c_void_p), ("camera", Camera), ("out", TraceOut), ("timeout_ms", C.c_float),
                ("flags", C.c_uint32)]


class RequestStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("latency_ms", C.c_float), ("queue_ms", C.c_float),
                ("wave", C.c_uint64), ("fault", C.c_uint32), ("reserved", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64), ("waves", C.c_uint64),
                ("evictions", C.c_uint64), ("resident_scenes", C.c_uint64), ("resident_bytes", C.c_uint64),
                ("inflight_bytes", C.c_uint64), ("pool_reserved_bytes", C.c_uint64), ("rays_traced", C.c_uint64),
                ("tris_built", C.c_uint64), ("p50_ms", C.c_float), ("p99_ms", C.c_float), ("mean_ms", C.c_float),
                ("max_ms", C.c_float), ("span_s", C.c_float), ("mean_wave_requests", C.c_float),
                ("stage_in_ms", C.c_float), ("stage_out_ms", C.c_float), ("wave_ms", C.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Timings(C.Structure):
    _fields_ = [("bounds_ms", C.c_float), ("morton_ms", C.c_float), ("sort_ms", C.c_float), ("emit_ms", C.c_float),
                ("refit_ms", C.c_float), ("sah_ms", C.c_float), ("grid_ms", C.c_float),
                ("total_build_ms", C.c_float), ("trace_ms", C.c_float), ("node_tests", C.c_uint64),
                ("tri_tests", C.c_uint64), ("sort_passes", C.c_uint32), ("tree_depth", C.c_uint32),
                ("top_levels", C.c_uint32), ("small_tasks", C.c_uint32), ("top_retries", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

_P = C.c_void_p
_U32 = C.c_uint32
_SIGS = {
    "sq_last_error": (C.c_char_p, []),
    "sq_version": (C.c_char_p, []),
    "sq_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "sq_ctx_destroy": (None, [_P]),
    "sq_ctx_sync": (C.c_int, [_P]),
    "sq_ctx_stream": (_P, [_P]),
    "sq_get_timings": (C.c_int, [_P, C.POINTER(Timings)]),
    "sq_ctx_set_phase_timing": (C.c_int, [_P, C.c_int]),
    "sq_device_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "sq_device_free": (None, [_P, _P]),
    "sq_copy_to_host": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "sq_copy_to_device": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "sq_device_copy": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "sq_scene_upload": (C.c_int, [_P, _P, _U32, C.POINTER(_P)]),
    "sq_scene_wrap_device": (C.c_int, [_P, _P, _U32, C.POINTER(_P)]),
    "sq_scene_free": (None, [_P]),
    "sq_scene_device_ptr": (_P, [_P]),
    "sq_scene_count": (_U32, [_P]),
    "sq_build": (C.c_int, [_P, _P, C.c_int, C.POINTER(BuildParams), C.POINTER(_P)]),
    "sq_bvh_from_host": (C.c_int, [_P, _P, _P, _U32, _U32, C.POINTER(_P)]),
    "sq_bvh_info": (C.c_int, [_P, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U32)]),
    "sq_bvh_download": (C.c_int, [_P, _P, _U32]),
    "sq_bvh_device_nodes": (_P, [_P]),
    "sq_bvh_free": (None, [_P]),
    "sq_morton_codes": (C.c_int, [_P, _P, _U32, _P]),
    "sq_radix_sort_pairs": (C.c_int, [_P, _P, _P, _U32, _U32, _U32]),
    "sq_refit": (C.c_int, [_P, _P]),
    "sq_trace_camera": (C.c_int, [_P, _P, C.POINTER(Camera), C.POINTER(TraceParams), C.POINTER(TraceOut), C.c_int]),
    "sq_trace_rays": (C.c_int, [_P, _P, _P, _U32, C.POINTER(TraceParams), C.POINTER(TraceOut), C.c_int]),
    "sq_trace_rays_noalloc": (C.c_int, [_P, _P, _P, _U32, C.POINTER(TraceParams), C.POINTER(TraceOut)]),
    "sq_batch_create": (C.c_int, [_P, _P, _P, _U32, C.c_int, C.POINTER(_P)]),
    "sq_batch_free": (None, [_P]),
    "sq_batch_count": (_U32, [_P]),
    "sq_batch_build": (C.c_int, [_P, _P, C.c_int, C.POINTER(BuildParams)]),
    "sq_batch_bvh": (C.c_int, [_P, _U32, C.POINTER(_P)]),
    "sq_batch_set_cameras": (C.c_int, [_P, _P, _P, C.POINTER(TraceParams), _U32]),
    "sq_batch_trace": (C.c_int, [_P, _P]),
    "sq_batch_run": (C.c_int, [_P, _P, C.c_int, C.POINTER(BuildParams)]),
    "sq_batch_outputs": (C.c_int, [_P, _U32, C.POINTER(TraceOut)]),
    "sq_batch_download": (C.c_int, [_P, _P, _U32, C.POINTER(TraceOut)]),
    "sq_batch_times": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "sq_batch_autotune": (C.c_int, [_P, _P, C.c_int, _U32, C.POINTER(TuneResult)]),
    "sq_server_create": (C.c_int, [C.c_int, C.POINTER(ServerConfig), C.POINTER(_P)]),
    "sq_server_destroy": (None, [_P]),
    "sq_server_submit": (C.c_int, [_P, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "sq_server_wait": (C.c_int, [_P, C.c_uint64, C.c_float, C.POINTER(RequestStatus)]),
    "sq_server_cancel": (C.c_int, [_P, C.c_uint64]),
    "sq_server_drain": (C.c_int, [_P]),
    "sq_server_evict": (C.c_int, [_P, C.c_uint64]),
    "sq_server_get_stats": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "sq_server_reset_stats": (C.c_int, [_P]),
    "sq_gen_icosphere_count": (_U32, [_U32]),
    "sq_gen_icosphere": (C.c_int, [_U32, _U32, C.c_float, _P]),
    "sq_gen_soup": (C.c_int, [_U32, _U32, C.c_float, _P]),
    "sq_gen_clusters": (C.c_int, [_U32, _U32, _U32, _P]),
    "sq_gen_secondary_rays": (C.c_int, [_P, _U32, _U32, _U32, C.c_float, _P]),
    "sq_reference_camera": (C.c_int, [_P, _U32, _U32, _U32, C.POINTER(Camera)]),
    "sq_pinhole_camera": (C.c_int, [_P, _U32, _U32, _U32, C.c_float, C.c_float, _U32, _U32, _U32,
                                    C.POINTER(Camera)]),
    "sq_bvh_plucker": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "sq_bvh2_create": (C.c_int, [_P, _P, _U32, C.POINTER(_P)]),
    "sq_bvh2_info": (C.c_int, [_P, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U32), _P]),
    "sq_bvh2_download": (C.c_int, [_P, _P, _U32, _P, _P, _U32]),
    "sq_bvh2_free": (None, [_P]),
    "sq_trace2_camera": (C.c_int, [_P, _P, C.POINTER(Camera), C.POINTER(TraceParams), C.POINTER(TraceOut), C.c_int]),
    "sq_trace2_rays": (C.c_int, [_P, _P, _P, _U32, C.POINTER(TraceParams), C.POINTER(TraceOut), C.c_int]),
    "sq_bvhk_create": (C.c_int, [_P, _P, _U32, _U32, C.POINTER(_P)]),
    "sq_bvhk_info": (C.c_int, [_P, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U32), _P]),
    "sq_bvhk_download": (C.c_int, [_P, _P, _U32]),
    "sq_bvhk_free": (None, [_P]),
    "sq_tracek_camera": (C.c_int, [_P, _P, C.POINTER(Camera), C.POINTER(TraceParams), C.POINTER(TraceOut), C.c_int]),
    "sq_tracek_rays": (C.c_int, [_P, _P, _P, _U32, C.POINTER(TraceParams), C.POINTER(TraceOut), C.c_int]),
    "sq_load_obj": (C.c_int, [C.c_char_p, C.POINTER(_P), C.POINTER(_U32), C.POINTER(_U32)]),
    "sq_host_free": (None, [_P]),
    "sq_bmp_size": (C.c_size_t, [_U32, _U32]),
    "sq_encode_bmp": (C.c_int, [_P, _U32, _U32, _P, C.c_size_t]),
    "sq_write_bmp": (C.c_int, [C.c_char_p, _P, _U32, _U32]),
}


def lib():
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libsequoia_amd.so is not built: run python -m sequoia_amd.build")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def declared_symbols():
    return list(_SIGS)


def _check(rc):
    if rc != SQ_OK:
        raise SqError(rc, lib().sq_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def as_tris(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    assert a.ndim == 2 and a.shape[1] == 9, a.shape
    return a


# ---- generators (host) ----
def gen_icosphere(target, seed=1, amplitude=0.1):
    n = lib().sq_gen_icosphere_count(target)
    out = np.empty((n, 9), np.float32)
    r = lib().sq_gen_icosphere(target, seed, amplitude, _ptr(out))
    if r < 0:
        _check(r)
    return out


def gen_soup(n, seed=2, sigma=0.01):
    out = np.empty((n, 9), np.float32)
    r = lib().sq_gen_soup(n, seed, sigma, _ptr(out))
    if r < 0:
        _check(r)
    return out


def gen_clusters(n, clusters=64, seed=3):
    out = np.empty((n, 9), np.float32)
    r = lib().sq_gen_clusters(n, clusters, seed, _ptr(out))
    if r < 0:
        _check(r)
    return out


def gen_secondary_rays(tris, nrays, seed=3, tmin=1e-4):
    tris = as_tris(tris)
    out = np.empty(nrays, RAY_DTYPE)
    r = lib().sq_gen_secondary_rays(_ptr(tris), len(tris), nrays, seed, tmin, _ptr(out))
    if r < 0:
        _check(r)
    return out


def reference_camera(tris, w, h):
    tris = as_tris(tris)
    cam = Camera()
    _check(lib().sq_reference_camera(_ptr(tris), len(tris), w, h, C.byref(cam)))
    return cam


def pinhole_camera(tris, w, h, fov_deg=60.0, distance=3.0, spp=1, jitter=0, seed=0):
    tris = as_tris(tris)
    cam = Camera()
    _check(lib().sq_pinhole_camera(_ptr(tris), len(tris), w, h, fov_deg, distance, spp, jitter, seed, C.byref(cam)))
    return cam


# ---- scene files (host) ----
def load_obj(path):
    """Triangles of a Wavefront OBJ, read as the reference's loader reads it;
    returns (tris (n, 9) float32, vertex count)."""
    p, n, nv = _P(), _U32(), _U32()
    _check(lib().sq_load_obj(os.fsencode(path), C.byref(p), C.byref(n), C.byref(nv)))
    try:
        out = np.empty((n.value, 9), np.float32)
        if n.value:
            C.memmove(out.ctypes.data, p, out.nbytes)
    finally:
        lib().sq_host_free(p)
    return out, nv.value


def write_obj(path, tris):
    """Writes triangles as OBJ text: three 'v' records per triangle with 9
    significant digits (which read back as the same float32 values), then one
    'f' record per triangle."""
    v = as_tris(tris).reshape(-1, 3)
    with open(path, "w", newline="\n") as f:
        f.write("# %d triangles\n" % (len(v) // 3))
        f.writelines("v %.9g %.9g %.9g\n" % (float(a), float(b), float(c)) for a, b, c in v)
        f.writelines("f %d %d %d\n" % (3 * i + 1, 3 * i + 2, 3 * i + 3) for i in range(len(v) // 3))


def encode_bmp(rgb):
    """The reference's heatmap BMP bytes for an (h, w, 3) uint8 image."""
    rgb = np.ascontiguousarray(rgb, np.uint8)
    assert rgb.ndim == 3 and rgb.shape[2] == 3, rgb.shape
    h, w = rgb.shape[:2]
    out = np.empty(lib().sq_bmp_size(w, h), np.uint8)
    _check(lib().sq_encode_bmp(_ptr(rgb), w, h, _ptr(out), out.nbytes))
    return out.tobytes()


def write_bmp(path, rgb):
    rgb = np.ascontiguousarray(rgb, np.uint8)
    assert rgb.ndim == 3 and rgb.shape[2] == 3, rgb.shape
    h, w = rgb.shape[:2]
    _check(lib().sq_write_bmp(os.fsencode(path), _ptr(rgb), w, h))


# ---- device objects ----
class Scene:
    def __init__(self, ctx, handle, n, keep=None):
        self.ctx, self.h, self.n, self._keep = ctx, handle, n, keep

    def free(self):
        if self.h:
            lib().sq_scene_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Bvh:
    def __init__(self, ctx, handle, scene):
        self.ctx, self.h, self.scene = ctx, handle, scene

    def info(self):
        r, cap, d = _U32(), _U32(), _U32()
        _check(lib().sq_bvh_info(self.h, C.byref(r), C.byref(cap), C.byref(d)))
        return r.value, cap.value, d.value

    @property
    def root(self):
        return self.info()[0]

    def download(self):
        _, cap, _ = self.info()
        out = np.empty(cap, NODE_DTYPE)
        _check(lib().sq_bvh_download(self.h, _ptr(out), cap))
        return out

    def free(self):
        if self.h:
            lib().sq_bvh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Node2Root(C.Structure):
    _fields_ = [("min", C.c_float * 3), ("max", C.c_float * 3), ("kind", C.c_uint32), ("child", C.c_uint32),
                ("count", C.c_uint32), ("reserved", C.c_uint32)]


def nodek_dtype(k):
    """Node<K> (aabb.h:26-32): min x/y/z[K], flags[K], max x/y/z[K], child[K]."""
    return np.dtype([("min_x", "<f4", k), ("min_y", "<f4", k), ("min_z", "<f4", k), ("flags", "<u4", k),
                     ("max_x", "<f4", k), ("max_y", "<f4", k), ("max_z", "<f4", k), ("child", "<u4", k)])


class BvhK:
    """Wide (Node<K>, K = 4 or 8) tree converted from a Node1 tree; its leaves
    index the Node1 tree's scene."""

    def __init__(self, ctx, handle, scene):
        self.ctx, self.h, self.scene = ctx, handle, scene

    def info(self):
        k, n, lv, r = _U32(), _U32(), _U32(), Node2Root()
        _check(lib().sq_bvhk_info(self.h, C.byref(k), C.byref(n), C.byref(lv), C.byref(r)))
        return {"k": k.value, "nodes": n.value, "levels": lv.value, "root": r}

    def download(self):
        """(nodes, root record)."""
        i = self.info()
        nodes = np.empty(i["nodes"], nodek_dtype(i["k"]))
        _check(lib().sq_bvhk_download(self.h, _ptr(nodes), len(nodes)))
        return nodes, i["root"]

    def free(self):
        if self.h:
            lib().sq_bvhk_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _trace_fns(bvh):
    """(camera, rays) entry points for a Node1, Node2 or Node<K> tree."""
    if isinstance(bvh, Bvh2):
        return lib().sq_trace2_camera, lib().sq_trace2_rays
    if isinstance(bvh, BvhK):
        return lib().sq_tracek_camera, lib().sq_tracek_rays
    return lib().sq_trace_camera, lib().sq_trace_rays


class Bvh2:
    """Quantised two-child (Node2) tree converted from a Node1 tree."""

    def __init__(self, ctx, handle):
        self.ctx, self.h = ctx, handle

    def info(self):
        n, t, lv, r = _U32(), _U32(), _U32(), Node2Root()
        _check(lib().sq_bvh2_info(self.h, C.byref(n), C.byref(t), C.byref(lv), C.byref(r)))
        return {"nodes": n.value, "tris": t.value, "levels": lv.value, "root": r}

    def download(self):
        """(nodes, root record, reordered triangles, original indices)."""
        i = self.info()
        nodes = np.empty(i["nodes"], NODE2_DTYPE)
        tris = np.empty((i["tris"], 9), np.float32)
        perm = np.empty(i["tris"], np.uint32)
        _check(lib().sq_bvh2_download(self.h, _ptr(nodes), len(nodes), _ptr(tris), _ptr(perm), len(perm)))
        return nodes, i["root"], tris, perm

    def free(self):
        if self.h:
            lib().sq_bvh2_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


OUT_HIT_ID, OUT_HIT_T, OUT_COUNTS, OUT_RGB = 1, 2, 4, 8
SQ_REQ_DEVICE_OUT = 1
OUT_ALL = 15


class BatchBvh(Bvh):
    """Tree of one scene of a batch (storage owned by the batch)."""

    def free(self):
        self.h = None


class Batch:
    """Many scenes built and traced together (sq_batch_*)."""

    def __init__(self, ctx, scenes):
        self.ctx = ctx
        arrs = [as_tris(t) for t in scenes]
        self.counts = np.array([len(a) for a in arrs], np.uint32)
        self._tris = np.ascontiguousarray(np.concatenate(arrs))
        h = C.c_void_p()
        _check(lib().sq_batch_create(ctx.h, _ptr(self._tris), _ptr(self.counts), len(arrs), 0, C.byref(h)))
        self.h = h
        self.cams = None

    def __len__(self):
        return len(self.counts)

    def free(self):
        if self.h:
            lib().sq_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @staticmethod
    def _bp(kind, digit_bits, subtree_threshold, layout, morton_bits=0, num_bins=0):
        return BuildParams(morton_bits, digit_bits, num_bins, 0, layout, 0, subtree_threshold, 0)

    def build(self, kind="morton63", digit_bits=0, subtree_threshold=0, layout=LAYOUT_REFERENCE, num_bins=0):
        p = self._bp(kind, digit_bits, subtree_threshold, layout, num_bins=num_bins)
        _check(lib().sq_batch_build(self.ctx.h, self.h, BUILDERS[kind], C.byref(p)))

    def bvh(self, i):
        h = C.c_void_p()
        _check(lib().sq_batch_bvh(self.h, i, C.byref(h)))
        return BatchBvh(self.ctx, h, None)

    def set_cameras(self, cams, semantics=TRACE_EXHAUSTIVE, outputs=OUT_ALL, stack_slots=0, block_size=0,
                    count_work=False, kernel=0, lists_per_step=0, blocks_per_cu=0):
        arr = (Camera * len(cams))(*cams)
        p = TraceParams(semantics, HEATMAP_RAY, 0, block_size, stack_slots, 1 if count_work else 0,
                        kernel, blocks_per_cu, 0, lists_per_step)
        _check(lib().sq_batch_set_cameras(self.ctx.h, self.h, arr, C.byref(p), outputs))
        self.cams = list(cams)
        self._cam_arr = arr

    def trace(self):
        _check(lib().sq_batch_trace(self.ctx.h, self.h))

    def run(self, kind="morton63", digit_bits=0, subtree_threshold=0, layout=LAYOUT_REFERENCE, num_bins=0):
        p = self._bp(kind, digit_bits, subtree_threshold, layout, num_bins=num_bins)
        _check(lib().sq_batch_run(self.ctx.h, self.h, BUILDERS[kind], C.byref(p)))

    def times(self):
        b, t, r = C.c_float(), C.c_float(), C.c_float()
        _check(lib().sq_batch_times(self.h, C.byref(b), C.byref(t), C.byref(r)))
        return {"build_ms": b.value, "trace_ms": t.value, "run_ms": r.value}

    def outputs(self, i):
        o = TraceOut()
        _check(lib().sq_batch_outputs(self.h, i, C.byref(o)))
        return o

    def download(self, i, hits=True, counts=True, rgb=True):
        cam = self.cams[i]
        npix = cam.width * cam.height
        nray = npix * cam.spp
        hid = np.empty(nray, np.uint32) if hits else None
        ht = np.empty(nray, np.float32) if hits else None
        cnt = np.empty(npix, np.uint32) if counts else None
        im = np.empty((cam.height, cam.width, 3), np.uint8) if rgb else None
        out = TraceOut(_ptr(hid), _ptr(ht), _ptr(cnt), _ptr(im))
        _check(lib().sq_batch_download(self.ctx.h, self.h, i, C.byref(out)))
        return {"hit_id": hid, "hit_t": ht, "counts": cnt, "rgb": im}

    def autotune(self, kind="morton63", repeats=2):
        r = TuneResult()
        _check(lib().sq_batch_autotune(self.ctx.h, self.h, BUILDERS[kind], repeats, C.byref(r)))
        return r.as_dict()


class Server:
    """Continuous-batching request loop (sq_server_*). Inputs and output
    arrays of a request are kept alive here until wait() returns for it."""

    def __init__(self, device=0, budget_bytes=0, max_wave_requests=0, max_wave_rays=0, max_wave_tris=0,
                 semantics=TRACE_EXHAUSTIVE, inflight_waves=0):
        cfg = ServerConfig(budget_bytes, max_wave_requests, max_wave_rays, max_wave_tris, semantics, inflight_waves,
                           (C.c_uint32 * 3)())
        h = C.c_void_p()
        _check(lib().sq_server_create(device, C.byref(cfg), C.byref(h)))
        self.h = h
        self._live = {}

    def close(self):
        if self.h:
            lib().sq_server_destroy(self.h)
            self.h = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, key, tris=None, camera=None, rays=None, hits=True, counts=True, rgb=True, timeout_ms=0.0,
               device_out=None):
        """Build `tris` under `key` (or trace the resident scene `key` when
        tris is None) and trace `camera` or the ray list `rays`. device_out:
        a dict of device pointers (hit_id, hit_t, counts, rgb; missing or None
        = not written) that the trace writes in place; wait() then returns no
        host arrays."""
        keep = []
        r = Request()
        r.key = key
        if tris is not None:
            tris = as_tris(tris)
            keep.append(tris)
            r.tris, r.ntris = _ptr(tris), len(tris)
        if rays is not None:
            rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
            keep.append(rays)
            r.rays, r.nrays = _ptr(rays), len(rays)
            nray, npix = len(rays), 0
        else:
            r.camera = camera
            npix = camera.width * camera.height
            nray = npix * camera.spp
        if device_out is not None:
            def dp(k):
                v = device_out.get(k)
                return v.value if isinstance(v, C.c_void_p) else v
            out = {"hit_id": None, "hit_t": None, "counts": None, "rgb": None}
            r.out = TraceOut(dp("hit_id"), dp("hit_t"), dp("counts"), dp("rgb") if rays is None else None)
            r.flags = SQ_REQ_DEVICE_OUT
        else:
            out = {"hit_id": np.empty(nray, np.uint32) if hits else None,
                   "hit_t": np.empty(nray, np.float32) if hits else None,
                   "counts": np.empty(nray if rays is not None else npix, np.uint32) if counts else None,
                   "rgb": np.empty((camera.height, camera.width, 3), np.uint8) if (rgb and rays is None) else None}
            r.out = TraceOut(_ptr(out["hit_id"]), _ptr(out["hit_t"]), _ptr(out["counts"]), _ptr(out["rgb"]))
        r.timeout_ms = timeout_ms
        rid = C.c_uint64()
        _check(lib().sq_server_submit(self.h, C.byref(r), C.byref(rid)))
        self._live[rid.value] = (keep, out)
        return rid.value

    def wait(self, rid, timeout_ms=-1.0):
        """Returns (status dict, outputs); raises SqError(TIMEOUT) if the wait times out."""
        st = RequestStatus()
        rc = lib().sq_server_wait(self.h, rid, timeout_ms, C.byref(st))
        if rc != SQ_OK:
            _check(rc)
        _, out = self._live.pop(rid)
        return {"state": st.state, "status": st.status, "latency_ms": st.latency_ms, "queue_ms": st.queue_ms,
                "wave": st.wave, "fault": st.fault}, out

    def cancel(self, rid):
        _check(lib().sq_server_cancel(self.h, rid))

    def drain(self):
        _check(lib().sq_server_drain(self.h))

    def evict(self, key):
        _check(lib().sq_server_evict(self.h, key))

    def stats(self):
        s = ServerStats()
        _check(lib().sq_server_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        _check(lib().sq_server_reset_stats(self.h))


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().sq_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().sq_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        _check(lib().sq_ctx_sync(self.h))

    def timings(self):
        t = Timings()
        _check(lib().sq_get_timings(self.h, C.byref(t)))
        return t.as_dict()

    def set_phase_timing(self, enable=True):
        """Time each build phase (bounds_ms .. grid_ms) from now on; off by default."""
        _check(lib().sq_ctx_set_phase_timing(self.h, 1 if enable else 0))

    def alloc(self, nbytes):
        p = C.c_void_p()
        _check(lib().sq_device_alloc(self.h, nbytes, C.byref(p)))
        return p

    def free(self, p):
        lib().sq_device_free(self.h, p)

    def copy(self, dst, src, nbytes):
        _check(lib().sq_device_copy(self.h, dst, src, nbytes))

    def to_host(self, dev, nbytes, dtype):
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
        _check(lib().sq_copy_to_host(self.h, _ptr(out), dev, nbytes))
        return out

    def to_device(self, dev, host):
        """Copies a contiguous host array into device memory at `dev`."""
        host = np.ascontiguousarray(host)
        _check(lib().sq_copy_to_device(self.h, dev, _ptr(host), host.nbytes))

    def trace_camera_device(self, bvh, cam, hit_id=None, hit_t=None, counts=None, rgb=None,
                            semantics=TRACE_EXHAUSTIVE, stack_slots=0, count_work=False, kernel=0):
        """Trace with outputs already in device memory (no host copies)."""
        out = TraceOut(hit_id, hit_t, counts, rgb)
        p = self._params(semantics, HEATMAP_RAY, 0, stack_slots, count_work, kernel)
        f = _trace_fns(bvh)[0]
        _check(f(self.h, bvh.h, C.byref(cam), C.byref(p), C.byref(out), 1))

    def trace_rays_device(self, bvh, rays, n, hit_id=None, hit_t=None, counts=None, semantics=TRACE_EXHAUSTIVE,
                          count_work=False, lists_per_step=0, noalloc=False):
        """Traces n rays already in device memory into device outputs; noalloc
        selects the allocation-free entry point (fixed per-device scratch)."""
        out = TraceOut(hit_id, hit_t, counts, None)
        p = self._params(semantics, HEATMAP_RAY, 0, 0, count_work, 0, 0, lists_per_step)
        if noalloc:
            _check(lib().sq_trace_rays_noalloc(self.h, bvh.h, rays, n, C.byref(p), C.byref(out)))
        else:
            _check(_trace_fns(bvh)[1](self.h, bvh.h, rays, n, C.byref(p), C.byref(out), 1))

    def upload(self, tris):
        tris = as_tris(tris)
        h = C.c_void_p()
        _check(lib().sq_scene_upload(self.h, _ptr(tris), len(tris), C.byref(h)))
        return Scene(self, h, len(tris))

    def build(self, scene, kind="morton", morton_bits=0, digit_bits=0, num_bins=0, grid_dim=0,
              layout=LAYOUT_REFERENCE, refit=True, subtree_threshold=0):
        p = BuildParams(morton_bits, digit_bits, num_bins, grid_dim, layout, 0 if refit else 1, subtree_threshold, 0)
        h = C.c_void_p()
        _check(lib().sq_build(self.h, scene.h, BUILDERS[kind] if isinstance(kind, str) else kind, C.byref(p),
                              C.byref(h)))
        return Bvh(self, h, scene)

    def bvh_from_host(self, scene, nodes, root):
        nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
        h = C.c_void_p()
        _check(lib().sq_bvh_from_host(self.h, scene.h, _ptr(nodes), len(nodes), root, C.byref(h)))
        return Bvh(self, h, scene)

    def morton_codes(self, scene, bits_per_axis):
        out = np.empty(scene.n, np.uint64)
        _check(lib().sq_morton_codes(self.h, scene.h, bits_per_axis, _ptr(out)))
        return out

    def radix_sort(self, keys, vals, key_bits, digit_bits=8):
        k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
        v = np.ascontiguousarray(vals, dtype=np.uint32).copy()
        _check(lib().sq_radix_sort_pairs(self.h, _ptr(k), _ptr(v), len(k), key_bits, digit_bits))
        return k, v

    def refit(self, bvh):
        _check(lib().sq_refit(self.h, bvh.h))

    @staticmethod
    def _params(semantics, heatmap, width, stack_slots, count_work, kernel=0, tri_test=0, lists_per_step=0):
        return TraceParams(semantics, heatmap, width, 0, stack_slots, 1 if count_work else 0, kernel, 0, tri_test,
                           lists_per_step)

    def trace_camera(self, bvh, cam, semantics=TRACE_EXHAUSTIVE, heatmap=HEATMAP_RAY, width=0, want_hits=True,
                     want_counts=True, want_rgb=True, stack_slots=0, count_work=False, kernel=0, tri_test=0,
                     lists_per_step=0):
        npix = cam.width * cam.height
        nray = npix * cam.spp
        res = {}
        hid = np.empty(nray, np.uint32) if want_hits else None
        ht = np.empty(nray, np.float32) if want_hits else None
        cnt = np.empty(npix, np.uint32) if want_counts else None
        rgb = np.empty((cam.height, cam.width, 3), np.uint8) if want_rgb else None
        out = TraceOut(_ptr(hid), _ptr(ht), _ptr(cnt), _ptr(rgb))
        p = self._params(semantics, heatmap, width, stack_slots, count_work, kernel, tri_test, lists_per_step)
        f = _trace_fns(bvh)[0]
        _check(f(self.h, bvh.h, C.byref(cam), C.byref(p), C.byref(out), 0))
        res.update(hit_id=hid, hit_t=ht, counts=cnt, rgb=rgb)
        return res

    def trace_rays(self, bvh, rays, semantics=TRACE_EXHAUSTIVE, want_counts=True, stack_slots=0, count_work=False,
                   tri_test=0, lists_per_step=0):
        """Ray list through a Node1 (Bvh), Node2 (Bvh2) or Node<K> (BvhK) tree."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        n = len(rays)
        hid = np.empty(n, np.uint32)
        ht = np.empty(n, np.float32)
        cnt = np.empty(n, np.uint32) if want_counts else None
        out = TraceOut(_ptr(hid), _ptr(ht), _ptr(cnt), None)
        p = self._params(semantics, HEATMAP_RAY, 0, stack_slots, count_work, 0, tri_test, lists_per_step)
        f = _trace_fns(bvh)[1]
        _check(f(self.h, bvh.h, _ptr(rays), n, C.byref(p), C.byref(out), 0))
        return {"hit_id": hid, "hit_t": ht, "counts": cnt}

    def plucker(self, bvh):
        """Precomputes the tree's triangles for tri_test=1; returns them (host copy)."""
        p = C.c_void_p()
        _check(lib().sq_bvh_plucker(self.h, bvh.h, C.byref(p)))
        n = lib().sq_scene_count(bvh.scene.h)
        return self.to_host(p.value, n * PLUCKER_DTYPE.itemsize, np.uint8).view(PLUCKER_DTYPE)

    def nodek(self, bvh, k=4, collapse=True):
        """Converts a Node1 tree into a Node<k> tree (k = 4 or 8) on the device."""
        h = C.c_void_p()
        _check(lib().sq_bvhk_create(self.h, bvh.h, k, 0 if collapse else BVHK_NO_COLLAPSE, C.byref(h)))
        return BvhK(self, h, bvh.scene)

    def node2(self, bvh, max_leaf=0):
        """Converts a Node1 tree into a Node2 tree on the device."""
        h = C.c_void_p()
        _check(lib().sq_bvh2_create(self.h, bvh.h, max_leaf, C.byref(h)))
        return Bvh2(self, h)
