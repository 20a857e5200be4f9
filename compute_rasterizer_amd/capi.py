"""ctypes binding of libpcr.so (the C ABI declared in include/pcr.h).

The structures below mirror include/pcr.h field for field. The library is
loaded from this package directory; if it is missing the import fails loudly
(there is no CPU fallback: the product path is the HIP library).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpcr.so")

PCR_OK = 0
PCR_E_INVALID = -1
PCR_E_NO_DEVICE = -2
PCR_E_HIP = -3
PCR_E_OOM = -4
PCR_E_BUDGET = -5
PCR_E_NOT_READY = -6
PCR_E_CANCELLED = -7

POINTS_PER_BATCH = 10240
STATIC_MAX_BATCHES = 262144
FB_CLEAR = 0xFFFFFFFFFFFFFFFF
DEPTH_CLEAR = 0xFFFFFFFF
BG_LAS = 0x00443322
BG_BLACK = 0xFF000000

MODE_PLAIN16, MODE_LAS, MODE_LAS2, MODE_HQS, MODE_NODES, MODE_HQS16, MODE_NODES_HQS = range(7)
LOD_SPACING, LOD_SIZE80 = 0, 1
SCENE_CUBE, SCENE_TERRAIN, SCENE_CITY = 0, 1, 2
FB_KEY64, FB_HQS = 0, 1


class Batch(C.Structure):
    _fields_ = [("state", C.c_int32), ("min", C.c_float * 3), ("max", C.c_float * 3),
                ("num_points", C.c_uint32), ("first_point", C.c_uint32), ("level", C.c_uint32),
                ("parent", C.c_int32), ("spacing", C.c_float), ("pad", C.c_uint32 * 4)]


assert C.sizeof(Batch) == 64

# numpy view of pcr_batch arrays
BATCH_DTYPE = np.dtype([("state", "<i4"), ("min", "<f4", 3), ("max", "<f4", 3), ("num_points", "<u4"),
                        ("first_point", "<u4"), ("level", "<u4"), ("parent", "<i4"), ("spacing", "<f4"),
                        ("pad", "<u4", 4)])
assert BATCH_DTYPE.itemsize == 64


class View(C.Structure):
    _fields_ = [("transform", C.c_float * 16), ("transform_frustum", C.c_float * 16),
                ("view_world", C.c_float * 16), ("proj", C.c_float * 16),
                ("width", C.c_int32), ("height", C.c_int32), ("pad", C.c_int32 * 2)]


class RenderParams(C.Structure):
    _fields_ = [("mode", C.c_int32), ("frustum_culling", C.c_int32), ("lod_rule", C.c_int32),
                ("lod_px", C.c_float), ("background", C.c_uint32), ("block_size", C.c_int32),
                ("clear_before", C.c_int32), ("time_stages", C.c_int32), ("tile_px", C.c_int32),
                ("batch_begin", C.c_uint32), ("batch_end", C.c_uint32), ("resolve_window", C.c_int32),
                ("edl_strength", C.c_float), ("occlusion_skip", C.c_int32), ("debug_counters", C.c_int32)]


class LasInfo(C.Structure):
    _fields_ = [("num_points", C.c_uint64), ("offset_to_points", C.c_uint32), ("record_length", C.c_uint16),
                ("point_format", C.c_uint8), ("version_major", C.c_uint8), ("version_minor", C.c_uint8),
                ("rgb_offset", C.c_uint8), ("compressed", C.c_uint8), ("pad", C.c_uint8 * 5),
                ("scale", C.c_double * 3),
                ("offset", C.c_double * 3), ("min", C.c_double * 3), ("max", C.c_double * 3)]


class LazInfo(C.Structure):
    _fields_ = [("num_points", C.c_uint64), ("record_length", C.c_uint32), ("num_chunks", C.c_uint32),
                ("chunk_size", C.c_uint32), ("compressor", C.c_uint16), ("num_items", C.c_uint16),
                ("item_type", C.c_uint16 * 8), ("item_size", C.c_uint16 * 8), ("item_version", C.c_uint16 * 8),
                ("point_format", C.c_uint8), ("has_rgb", C.c_uint8), ("rgb_offset", C.c_uint8),
                ("pad", C.c_uint8 * 5)]


LAZ_ALL_FIELDS = 0
LAZ_XYZ_RGB = 1


class PotreeInfo(C.Structure):
    _fields_ = [("scale", C.c_double * 3), ("offset", C.c_double * 3), ("min", C.c_double * 3),
                ("max", C.c_double * 3), ("spacing", C.c_double), ("num_points", C.c_uint64),
                ("first_chunk_size", C.c_uint64), ("bytes_per_point", C.c_uint32), ("rgb_offset", C.c_uint32)]


class LasFile(C.Structure):
    _fields_ = [("bytes", C.c_void_p), ("size", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("batches_total", C.c_uint64), ("batches_culled", C.c_uint64),
                ("batches_per_level", C.c_uint64 * 5), ("points_drawn", C.c_uint64),
                ("bytes_algorithmic", C.c_uint64), ("ms_cull", C.c_float), ("ms_raster", C.c_float),
                ("ms_resolve", C.c_float), ("ms_clear", C.c_float), ("batches_occluded", C.c_uint64),
                ("points_occluded", C.c_uint64), ("bytes_occluded", C.c_uint64), ("atomics_issued", C.c_uint64),
                ("pixels_visible", C.c_uint64)]


class GenCfg(C.Structure):
    _fields_ = [("scene", C.c_int32), ("num_buildings", C.c_int32), ("seed", C.c_uint64),
                ("num_points", C.c_uint64), ("extent", C.c_float), ("amplitude", C.c_float),
                ("jitter", C.c_float), ("terrain_frac", C.c_float)]


class BuildOpts(C.Structure):
    _fields_ = [("batch_size", C.c_uint32), ("morton_sort", C.c_int32), ("keep_points16", C.c_int32)]


class CloudInfo(C.Structure):
    _fields_ = [("num_points", C.c_uint64), ("num_batches", C.c_uint32), ("batch_size", C.c_uint32),
                ("has_points16", C.c_int32), ("has_batches", C.c_int32), ("num_levels", C.c_int32),
                ("aabb_min", C.c_float * 3), ("aabb_max", C.c_float * 3), ("device_bytes", C.c_uint64)]


class OctreeOpts(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("root_spacing_div", C.c_float), ("max_node_points", C.c_uint32),
                ("max_depth", C.c_int32)]


class Tune(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("points_per_thread", C.c_int32), ("persistent_groups", C.c_int32),
                ("tile_px", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("cloud_id", C.c_int32), ("mode", C.c_int32), ("lod_px", C.c_float), ("pad", C.c_int32),
                ("view", View)]


class MultiViewStats(C.Structure):
    _fields_ = [("batches_any_view", C.c_uint64), ("bytes_batched", C.c_uint64), ("bytes_per_view_sum", C.c_uint64),
                ("points_per_view_sum", C.c_uint64), ("points_any_view", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("refused", C.c_uint64), ("launches", C.c_uint64),
                ("points_rendered", C.c_uint64), ("fb_bytes_in_use", C.c_uint64), ("fb_bytes_peak", C.c_uint64),
                ("points_drawn", C.c_uint64), ("bytes_algorithmic", C.c_uint64)]


P = C.c_void_p
PP = C.POINTER(C.c_void_p)
i32, u32, u64, f32, f64 = C.c_int32, C.c_uint32, C.c_uint64, C.c_float, C.c_double
dp = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol declared in include/pcr.h
SIGNATURES = {
    "pcr_last_error": (C.c_char_p, []),
    "pcr_version": (C.c_int, []),
    "pcr_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "pcr_create": (C.c_int, [C.c_int, PP]),
    "pcr_destroy": (C.c_int, [P]),
    "pcr_synchronize": (C.c_int, [P]),
    "pcr_stream": (P, [P]),
    "pcr_view_look_at": (C.c_int, [dp, dp, dp, f64, f64, f64, dp, C.c_int, C.c_int, C.POINTER(View)]),
    "pcr_view_from_matrices": (C.c_int, [dp, dp, dp, C.c_int, C.c_int, C.POINTER(View)]),
    "pcr_encode_batches_host": (C.c_int, [P, u64, u32, P, P, P, P]),
    "pcr_cloud_upload_points16": (C.c_int, [P, P, P, u64, PP]),
    "pcr_cloud_upload_batched": (C.c_int, [P, P, u32, P, P, P, P, u64, PP]),
    "pcr_generate_points": (C.c_int, [P, C.POINTER(GenCfg), P, P]),
    "pcr_scene_ground_height": (C.c_int, [C.POINTER(GenCfg), f32, f32, C.POINTER(C.c_float)]),
    "pcr_cloud_build_device": (C.c_int, [P, P, P, u64, C.POINTER(BuildOpts), PP]),
    "pcr_cloud_generate": (C.c_int, [P, C.POINTER(GenCfg), C.POINTER(BuildOpts), PP]),
    "pcr_cloud_info_get": (C.c_int, [P, C.POINTER(CloudInfo)]),
    "pcr_cloud_download": (C.c_int, [P, P, P, P, P, P, P, P]),
    "pcr_cloud_destroy": (C.c_int, [P]),
    "pcr_cloud_build_octree_host": (C.c_int, [P, P, P, u64, C.POINTER(OctreeOpts), PP]),
    "pcr_cloud_build_octree_device": (C.c_int, [P, P, P, u64, C.POINTER(OctreeOpts), PP]),
    "pcr_lod_selection": (C.c_int, [P, P, P, u32]),
    "pcr_octree_generate": (C.c_int, [P, C.POINTER(GenCfg), C.POINTER(OctreeOpts), PP]),
    "pcr_octree_layout_host": (C.c_int, [P, P, u64, C.POINTER(OctreeOpts), P, C.POINTER(u32), P, P, P, P,
                                         C.POINTER(u64)]),
    "pcr_fb_create": (C.c_int, [P, C.c_int, C.c_int, C.c_int, PP]),
    "pcr_fb_destroy": (C.c_int, [P]),
    "pcr_fb_clear": (C.c_int, [P, P]),
    "pcr_fb_readback": (C.c_int, [P, P, P]),
    "pcr_fb_readback_hqs": (C.c_int, [P, P, P, P]),
    "pcr_fb_readback_rgba": (C.c_int, [P, P, P]),
    "pcr_fb_device_keys": (P, [P]),
    "pcr_fb_device_rgba": (P, [P]),
    "pcr_render": (C.c_int, [P, P, C.POINTER(View), C.POINTER(RenderParams), P]),
    "pcr_resolve": (C.c_int, [P, P, C.POINTER(RenderParams), P]),
    "pcr_render_static": (C.c_int, [P, u32, P, P, P, u64, C.POINTER(View), C.POINTER(RenderParams), P, P]),
    "pcr_stats_get": (C.c_int, [P, C.POINTER(Stats)]),
    "pcr_las_parse_header": (C.c_int, [P, C.c_uint64, P]),
    "pcr_laz_info_get": (C.c_int, [P, C.c_uint64, P]),
    "pcr_laz_chunk_table": (C.c_int, [P, C.c_uint64, P, P, C.c_uint32, P]),
    "pcr_laz_decode": (C.c_int, [P, C.c_uint64, C.c_uint64, C.c_uint64, P, C.c_uint32, C.c_int]),
    "pcr_cloud_from_las": (C.c_int, [P, P, C.c_int, P, C.POINTER(C.c_void_p)]),
    "pcr_cloud_from_potree": (C.c_int, [P, P, P, C.c_uint64, P, C.c_uint64, C.POINTER(C.c_void_p)]),
    "pcr_frame_times": (C.c_int, [P, C.c_int, P, C.POINTER(C.c_int)]),
    "pcr_selftest_divide": (C.c_int, [P, P, P, P, C.c_uint64, P]),
    "pcr_multiview_create": (C.c_int, [P, P, C.POINTER(View), C.c_int, C.POINTER(C.c_void_p),
                                       C.POINTER(RenderParams), C.POINTER(Tune), C.c_int, PP]),
    "pcr_multiview_set_views": (C.c_int, [P, C.POINTER(View), C.c_int]),
    "pcr_multiview_launch": (C.c_int, [P]),
    "pcr_multiview_destroy": (C.c_int, [P]),
    "pcr_multiview_stats_get": (C.c_int, [P, C.POINTER(MultiViewStats)]),
    "pcr_autotune": (C.c_int, [P, P, C.POINTER(View), C.c_int, C.POINTER(C.c_void_p), C.POINTER(RenderParams),
                               C.c_int, C.POINTER(Tune), C.POINTER(C.c_float)]),
    "pcr_server_create": (C.c_int, [P, u64, C.c_int, C.c_int, PP]),
    "pcr_server_add_cloud": (C.c_int, [P, P, C.POINTER(C.c_int)]),
    "pcr_server_submit": (C.c_int, [P, C.POINTER(Request), C.POINTER(u64)]),
    "pcr_server_poll": (C.c_int, [P, u64, P, dp]),
    "pcr_server_wait": (C.c_int, [P, u64, P, dp]),
    "pcr_server_wait_pinned": (C.c_int, [P, C.c_uint64, C.POINTER(C.c_void_p), C.POINTER(C.c_double)]),
    "pcr_server_release": (C.c_int, [P, C.c_uint64]),
    "pcr_server_cancel": (C.c_int, [P, u64]),
    "pcr_server_stats_get": (C.c_int, [P, C.POINTER(ServerStats)]),
    "pcr_server_destroy": (C.c_int, [P]),
}


class PcrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pcr error {code}: {msg}")
        self.code = code


_lib = None


def load() -> C.CDLL:
    """Load libpcr.so from the package directory (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                          f"make -C compute_rasterizer_amd/csrc")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> int:
    if rc != PCR_OK:
        raise PcrError(rc, load().pcr_last_error().decode())
    return rc


def ptr(a: np.ndarray | None):
    """Raw pointer of a contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    return C.c_void_p(a.ctypes.data)


def dvec(v) -> C.Array:
    return (C.c_double * 3)(*[float(x) for x in v])
