"""ctypes wrapper of the CPU oracle liborc.so (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to check the GPU's results against. It restates the reference's per-frame
shader logic in plain C (pcr_oracle.c cites the reference file:line of each
function).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from compute_rasterizer_amd.capi import BATCH_DTYPE, FB_CLEAR, GenCfg, LasInfo, PotreeInfo, View

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liborc.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        P = C.c_void_p
        f = _lib
        f.orc_mix64.restype = C.c_uint64; f.orc_mix64.argtypes = [C.c_uint64]
        f.orc_rand.restype = C.c_uint64; f.orc_rand.argtypes = [C.c_uint64, C.c_uint64]
        f.orc_u01.restype = C.c_float; f.orc_u01.argtypes = [C.c_uint64]
        f.orc_generate.argtypes = [C.POINTER(GenCfg), C.c_uint64, C.c_uint64, P, P]
        f.orc_terrain_height.restype = C.c_float
        f.orc_terrain_height.argtypes = [C.c_uint64, C.c_float, C.c_float, C.c_float, C.c_float]
        f.orc_morton3.restype = C.c_uint64; f.orc_morton3.argtypes = [C.c_uint32] * 3
        f.orc_morton_order.argtypes = [P, C.c_uint64, P, P]
        f.orc_encode_batches.argtypes = [P, C.c_uint64, C.c_uint32, P, P, P, P]
        f.orc_quantize30.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_uint32)]
        f.orc_frustum_visible.restype = C.c_int; f.orc_frustum_visible.argtypes = [P, P, P]
        f.orc_projected_px.restype = C.c_float
        f.orc_projected_px.argtypes = [P, P, C.c_float, P, P, C.c_int, C.c_int]
        f.orc_precision_level.restype = C.c_int
        f.orc_precision_level.argtypes = [P, P, C.POINTER(View), C.c_float]
        f.orc_decode.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, P, P, P]
        f.orc_decode_n.argtypes = [C.c_int, C.c_uint64, P, P, P, P, P, P]
        f.orc_project.restype = C.c_int
        f.orc_project.argtypes = [P, P, C.c_int, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        f.orc_clear64.argtypes = [P, C.c_uint64]
        f.orc_render_points16.argtypes = [P, P, C.c_uint64, C.POINTER(View), P]
        f.orc_render_batched.argtypes = [P, C.c_uint32, P, P, P, C.POINTER(View), C.c_int, P, P]
        f.orc_resolve_index.argtypes = [P, C.c_uint64, P, C.c_uint64, C.c_uint32, P]
        f.orc_resolve_color.argtypes = [P, C.c_uint64, C.c_uint32, P]
        f.orc_edl_exp.argtypes = [C.c_float]
        f.orc_edl_exp.restype = C.c_float
        f.orc_resolve_edl.argtypes = [P, C.c_int, C.c_int, P, C.c_uint64, C.c_uint32, C.c_float, P]
        f.orc_hqs_batched.argtypes = [P, C.c_uint32, P, P, P, P, C.POINTER(View), C.c_int, P, P]
        f.orc_hqs_points16.argtypes = [P, P, C.c_uint64, C.POINTER(View), P, P]
        f.orc_hqs_resolve.argtypes = [P, C.c_uint64, C.c_uint32, P]
        f.orc_lod_select.argtypes = [P, C.c_uint32, C.POINTER(View), C.c_int, C.c_int, C.c_float, P]
        f.orc_render_nodes.argtypes = [P, C.c_uint32, P, P, P, C.POINTER(View), C.c_int, C.c_int, C.c_float, P, P]
        f.orc_render_batched_mt.argtypes = [P, C.c_uint32, P, P, P, C.POINTER(View), C.c_int, P, C.c_int]
        f.orc_render_nodes_mt.argtypes = [P, C.c_uint32, P, P, P, C.POINTER(View), C.c_int, C.c_int, C.c_float, P,
                                          C.c_int]
        f.orc_las_parse_header.restype = C.c_int
        f.orc_las_parse_header.argtypes = [P, C.c_uint64, C.POINTER(LasInfo)]
        f.orc_las_encode.argtypes = [P, C.c_uint64, C.POINTER(LasInfo), C.c_uint32, C.c_uint64, C.c_uint32,
                                     P, P, P, P, P]
        f.orc_hqs_nodes.argtypes = [P, C.c_uint32, P, P, P, P, C.POINTER(View), C.c_int, C.c_int, C.c_float, P, P]
        f.orc_hqs_resolve_window.argtypes = [P, P, C.c_int, C.c_int, C.c_int, P]
        f.orc_potree_nodes.restype = C.c_int64
        f.orc_potree_nodes.argtypes = [C.POINTER(PotreeInfo), P, C.c_uint64, P, P, C.c_uint64]
        f.orc_potree_encode.argtypes = [C.POINTER(PotreeInfo), P, P, P, C.c_uint64, P, P, P, P]
        D3 = C.POINTER(C.c_double)
        f.orc_look_at_view.argtypes = [D3, D3, D3, C.c_double, C.c_double, C.c_double, D3, C.c_int, C.c_int,
                                       C.POINTER(View)]
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def f32(a):
    return np.ascontiguousarray(a, np.float32)


def generate(cfg: GenCfg, first=0, count=None):
    count = cfg.num_points - first if count is None else count
    xyz = np.zeros((count, 3), np.float32)
    rgba = np.zeros(count, np.uint32)
    load().orc_generate(C.byref(cfg), first, count, _p(xyz), _p(rgba))
    return xyz, rgba


def morton_order(xyz):
    xyz = f32(xyz).reshape(-1, 3)
    n = xyz.shape[0]
    perm = np.zeros(n, np.uint32)
    keys = np.zeros(n, np.uint64)
    load().orc_morton_order(_p(xyz), n, _p(perm), _p(keys))
    return perm, keys


def encode_batches(xyz, batch_size=10240):
    xyz = f32(xyz).reshape(-1, 3)
    n = xyz.shape[0]
    nb = (n + batch_size - 1) // batch_size
    b = np.zeros(nb, BATCH_DTYPE)
    lo, med, hi = (np.zeros(n, np.uint32) for _ in range(3))
    load().orc_encode_batches(_p(xyz), n, batch_size, _p(b), _p(lo), _p(med), _p(hi))
    return b, lo, med, hi


def quantize30(x, mn, size):
    q = C.c_uint32()
    load().orc_quantize30(float(x), float(mn), float(size), C.byref(q))
    return q.value


def frustum_visible(tf, mn, mx):
    return bool(load().orc_frustum_visible(_p(f32(tf)), _p(f32(mn)), _p(f32(mx))))


def projected_px(mn, mx, length, vw, proj, w, h):
    return load().orc_projected_px(_p(f32(mn)), _p(f32(mx)), length, _p(f32(vw)), _p(f32(proj)), w, h)


def precision_level(mn, mx, view: View, first=100.0):
    return load().orc_precision_level(_p(f32(mn)), _p(f32(mx)), C.byref(view), first)


def decode(level, lo, med, hi, mn, mx):
    out = np.zeros(3, np.float32)
    load().orc_decode(level, lo, med, hi, _p(f32(mn)), _p(f32(mx)), _p(out))
    return out


def decode_n(level, lo, med, hi, mn, mx):
    """orc_decode over arrays of plane words sharing one box: (n, 3) float32."""
    lo = np.ascontiguousarray(lo, np.uint32)
    med = None if med is None else np.ascontiguousarray(med, np.uint32)
    hi = None if hi is None else np.ascontiguousarray(hi, np.uint32)
    out = np.zeros((lo.size, 3), np.float32)
    load().orc_decode_n(level, lo.size, _p(lo), _p(med), _p(hi), _p(f32(mn)), _p(f32(mx)), _p(out))
    return out


def project(t, p, w, h):
    pix, d = C.c_uint32(), C.c_uint32()
    ok = load().orc_project(_p(f32(t)), _p(f32(p)), w, h, C.byref(pix), C.byref(d))
    return (pix.value, d.value) if ok else None


def new_fb(view: View):
    return np.full(view.width * view.height, FB_CLEAR, np.uint64)


def render_points16(xyz, rgba, view: View, fb=None):
    xyz = f32(xyz).reshape(-1, 3)
    rgba = np.ascontiguousarray(rgba, np.uint32)
    fb = new_fb(view) if fb is None else fb
    load().orc_render_points16(_p(xyz), _p(rgba), xyz.shape[0], C.byref(view), _p(fb))
    return fb


def render_batched(b, lo, med, hi, view: View, culling=1, fb=None):
    fb = new_fb(view) if fb is None else fb
    levels = np.zeros(len(b), np.uint8)
    load().orc_render_batched(_p(b), len(b), _p(lo), _p(med), _p(hi), C.byref(view), culling, _p(fb), _p(levels))
    return fb, levels


def render_batched_mt(b, lo, med, hi, view: View, culling=1, nthreads=8, fb=None):
    fb = new_fb(view) if fb is None else fb
    load().orc_render_batched_mt(_p(b), len(b), _p(lo), _p(med), _p(hi), C.byref(view), culling, _p(fb), nthreads)
    return fb


def resolve_index(fb, rgba, bg):
    rgba = np.ascontiguousarray(rgba, np.uint32)
    fb = np.ascontiguousarray(fb, np.uint64)
    out = np.zeros(fb.shape[0], np.uint32)
    load().orc_resolve_index(_p(fb), fb.shape[0], _p(rgba), rgba.shape[0], bg, _p(out))
    return out


def edl_exp(x) -> float:
    """The oracle's e^x (x <= 0) of the eye-dome lighting shade."""
    return load().orc_edl_exp(float(x))


def resolve_edl(fb, width, height, rgba, bg, strength):
    """Index-payload resolve with eye-dome lighting (compute_loop_las/resolve.cs:148-191)."""
    rgba = np.ascontiguousarray(rgba, np.uint32)
    out = np.zeros(width * height, np.uint32)
    load().orc_resolve_edl(_p(np.ascontiguousarray(fb, np.uint64)), width, height, _p(rgba), rgba.shape[0], bg,
                           strength, _p(out))
    return out


def resolve_color(fb, bg):
    out = np.zeros(fb.shape[0], np.uint32)
    load().orc_resolve_color(_p(fb), fb.shape[0], bg, _p(out))
    return out


def hqs_batched(b, lo, med, hi, rgba, view: View, culling=1):
    npix = view.width * view.height
    depth = np.full(npix, 0xFFFFFFFF, np.uint32)
    acc = np.zeros(4 * npix, np.uint32)
    load().orc_hqs_batched(_p(b), len(b), _p(lo), _p(med), _p(hi), _p(rgba), C.byref(view), culling, _p(depth),
                           _p(acc))
    return depth, acc


def hqs_points16(xyz, rgba, view: View):
    xyz = f32(xyz).reshape(-1, 3)
    npix = view.width * view.height
    depth = np.full(npix, 0xFFFFFFFF, np.uint32)
    acc = np.zeros(4 * npix, np.uint32)
    load().orc_hqs_points16(_p(xyz), _p(np.ascontiguousarray(rgba, np.uint32)), xyz.shape[0], C.byref(view),
                            _p(depth), _p(acc))
    return depth, acc


def hqs_resolve(acc, bg):
    npix = acc.shape[0] // 4
    out = np.zeros(npix, np.uint32)
    load().orc_hqs_resolve(_p(acc), npix, bg, _p(out))
    return out


def lod_select(nodes, view: View, culling=1, rule=0, lod_px=1.0):
    flags = np.zeros(len(nodes), np.uint8)
    load().orc_lod_select(_p(nodes), len(nodes), C.byref(view), culling, rule, lod_px, _p(flags))
    return flags


def render_nodes(nodes, lo, med, hi, view: View, culling=1, rule=0, lod_px=1.0):
    fb = new_fb(view)
    flags = np.zeros(len(nodes), np.uint8)
    load().orc_render_nodes(_p(nodes), len(nodes), _p(lo), _p(med), _p(hi), C.byref(view), culling, rule, lod_px,
                            _p(fb), _p(flags))
    return fb, flags


def render_nodes_mt(nodes, lo, med, hi, view: View, culling=1, rule=0, lod_px=1.0, nthreads=8):
    fb = new_fb(view)
    load().orc_render_nodes_mt(_p(nodes), len(nodes), _p(lo), _p(med), _p(hi), C.byref(view), culling, rule,
                               lod_px, _p(fb), nthreads)
    return fb


def look_at(eye, target, up=(0, 1, 0), fovy=60.0, near=0.1, far=1000.0, origin=(0, 0, 0), width=1280,
            height=720) -> View:
    v = View()
    d = lambda x: (C.c_double * 3)(*map(float, x))  # noqa: E731
    load().orc_look_at_view(d(eye), d(target), d(up), fovy, near, far, d(origin), width, height, C.byref(v))
    return v


def las_parse_header(data: bytes):
    """pcr_las_info of a LAS file's bytes, or None when the header is invalid."""
    buf = np.frombuffer(data, np.uint8)
    info = LasInfo()
    rc = load().orc_las_parse_header(_p(buf), buf.size, C.byref(info))
    return info if rc == 0 else None


def las_encode(data: bytes, info: LasInfo, batch_size=10240, sparse_offset=0, file_index=0, laz=False):
    """(batches, lo, med, hi, rgba) of one LAS file, as LasLoaderSparse::loadLas builds them.

    laz=True: the records came from a LAZ file, which LasLoaderSparse::loadLaz
    (LasLoaderSparse.cpp:214-330) encodes the same way except for colour: it
    takes R, G, B from the LASzip point (:283-285), which holds zeros for point
    formats without colour, where loadLas reads the record's first six bytes."""
    buf = np.frombuffer(data, np.uint8)
    n = int(info.num_points)
    rec = np.ascontiguousarray(buf[info.offset_to_points:info.offset_to_points + n * info.record_length])
    nb = (n + batch_size - 1) // batch_size
    b = np.zeros(nb, BATCH_DTYPE)
    lo, med, hi, rgba = (np.zeros(n, np.uint32) for _ in range(4))
    load().orc_las_encode(_p(rec), n, C.byref(info), batch_size, sparse_offset, file_index, _p(b), _p(lo), _p(med),
                          _p(hi), _p(rgba))
    if laz and info.point_format not in (2, 3, 5, 7, 8, 10):
        rgba[:] = 0
    return b, lo, med, hi, rgba


def potree_nodes(info: PotreeInfo, hierarchy: bytes, max_nodes=1 << 22):
    """(batches, boxes) of a Potree 2.0 hierarchy in octree.bin order, or None when it is malformed."""
    h = np.frombuffer(hierarchy, np.uint8)
    b = np.zeros(max_nodes, BATCH_DTYPE)
    boxes = np.zeros((max_nodes, 6), np.float64)
    n = load().orc_potree_nodes(C.byref(info), _p(h), h.size, _p(b), _p(boxes), max_nodes)
    if n < 0:
        return None
    return b[:n].copy(), boxes[:n].copy()


def potree_encode(info: PotreeInfo, octree: bytes, b, boxes):
    """(lo, med, hi, rgba) of a Potree 2.0 octree.bin."""
    o = np.frombuffer(octree, np.uint8)
    n = int(info.num_points)
    lo, med, hi, rgba = (np.zeros(n, np.uint32) for _ in range(4))
    load().orc_potree_encode(C.byref(info), _p(o), _p(b), _p(np.ascontiguousarray(boxes)), len(b), _p(lo), _p(med),
                             _p(hi), _p(rgba))
    return lo, med, hi, rgba


def hqs_nodes(nodes, lo, med, hi, rgba, view: View, culling=1, rule=0, lod_px=1.0):
    """(depth, acc) of the hqs passes over the LOD node cut."""
    npix = view.width * view.height
    depth = np.full(npix, 0xFFFFFFFF, np.uint32)
    acc = np.zeros(4 * npix, np.uint32)
    load().orc_hqs_nodes(_p(nodes), len(nodes), _p(lo), _p(med), _p(hi), _p(rgba), C.byref(view), culling, rule,
                         lod_px, _p(depth), _p(acc))
    return depth, acc


def hqs_resolve_window(depth, acc, width, height, window=0):
    out = np.zeros(width * height, np.uint32)
    load().orc_hqs_resolve_window(_p(np.ascontiguousarray(depth, np.uint32)), _p(np.ascontiguousarray(acc, np.uint32)),
                                  width, height, window, _p(out))
    return out
