"""GPU: pcr_render_static, the allocation-free render, through the C ABI.

Its inputs (batch headers, planes) and its output key buffer are plain device
buffers the test allocates itself with the HIP runtime's C API (the entry
point's callers own them); its scratch is the library's fixed device array.
Every frame must equal pcr_render's frame of the same layout, view and
parameters bit for bit, for two calls in a row (the second with a smaller
cloud and image, so the scratch still holds the first call's list and levels),
in each mode it takes; one frame is also checked against the oracle.
"""
import ctypes as C

import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from compute_rasterizer_amd import capi
from oracle import orc

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


class DeviceArrays:
    """Device copies of host arrays, freed on close."""

    def __init__(self):
        self.hip = _hip()
        self.ptrs = []

    def put(self, a: np.ndarray) -> C.c_void_p:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 16)) == 0
        self.ptrs.append(p)
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, dtype, n) -> np.ndarray:
        out = np.zeros(n, dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _pad4(a):
    n = (a.size + 3) // 4 * 4
    out = np.zeros(n, np.uint32)
    out[:a.size] = a
    return out


def _terrain(n, seed):
    xyz, rgba = orc.generate(pcr.gen_cfg(pcr.SCENE_TERRAIN, n, seed=seed))
    perm, _ = orc.morton_order(xyz)
    b, lo, med, hi = orc.encode_batches(xyz[perm])
    return b, lo, med, hi, rgba[perm]


def _octree(n, seed):
    xyz, rgba = orc.generate(pcr.gen_cfg(pcr.SCENE_TERRAIN, n, seed=seed))
    return pcr.octree_layout_host(xyz, rgba, seed=seed, max_node_points=4000)


def _view(w, h, eye=(1000.0, -300.0, 500.0)):
    return pcr.look_at(eye, (1000.0, 1000.0, 40.0), up=(0, 0, 1), near=1.0, far=5000.0, width=w, height=h)


def static_frame(layout, view, prm, stream=None):
    b, lo, med, hi = layout[:4]
    dev = DeviceArrays()
    try:
        planes = [_pad4(x) for x in (lo, med, hi)]
        npix = view.width * view.height
        ptrs = [dev.put(np.ascontiguousarray(b, capi.BATCH_DTYPE))] + [dev.put(p) for p in planes]
        keys = dev.put(np.zeros(npix, np.uint64))      # garbage-free but not the clear value: the call clears
        capi.check(capi.load().pcr_render_static(ptrs[0], len(b), ptrs[1], ptrs[2], ptrs[3], planes[0].size,
                                                 C.byref(view), C.byref(prm), keys, stream))
        return dev.get(keys, np.uint64, npix)
    finally:
        dev.close()


def context_frame(ctx, layout, view, prm):
    b, lo, med, hi, rgba = layout
    cloud = ctx.upload_batched(b, lo, med, hi, rgba)
    fb = ctx.framebuffer(view.width, view.height)
    try:
        ctx.render(cloud, view, prm, fb)
        return fb.keys()
    finally:
        fb.close()
        cloud.close()


@pytest.mark.parametrize("mode,tile", [(pcr.MODE_LAS2, 2048), (pcr.MODE_LAS2, 0), (pcr.MODE_LAS, 0)])
def test_static_render_equals_render_twice(ctx, mode, tile):
    big, small = _terrain(400_000, 3), _terrain(90_000, 4)
    for k, (layout, view) in enumerate(((big, _view(640, 360)), (small, _view(320, 200, (900.0, -150.0, 300.0))))):
        prm = pcr.params(mode, tile_px=tile)
        print(f"render_static mode={mode} tile={tile} call={k} batches={len(layout[0])} "
              f"{view.width}x{view.height}", flush=True)
        got = static_frame(layout, view, prm, stream=capi.load().pcr_stream(ctx.h) if k else None)
        want = context_frame(ctx, layout, view, prm)
        assert (want != pcr.FB_CLEAR).sum() > 1000
        assert np.array_equal(got, want), (k, int((got != want).sum()))


def test_static_render_equals_oracle(ctx):
    layout = _terrain(200_000, 5)
    view = _view(480, 270)
    got = static_frame(layout, view, pcr.params(pcr.MODE_LAS2))
    ref, _ = orc.render_batched(*layout[:4], view, 1)
    assert (ref != pcr.FB_CLEAR).sum() > 1000
    assert np.array_equal(got, ref)


def test_static_render_nodes_twice(ctx):
    for k, (n, w, h) in enumerate(((300_000, 512, 384), (60_000, 256, 160))):
        layout = _octree(n, 6 + k)
        view = _view(w, h)
        for rule in (pcr.LOD_SPACING, pcr.LOD_SIZE80):
            prm = pcr.params(pcr.MODE_NODES, lod_rule=rule, lod_px=1.0)
            print(f"render_static nodes rule={rule} call={k} nodes={len(layout[0])}", flush=True)
            got = static_frame(layout, view, prm)
            want = context_frame(ctx, layout, view, prm)
            assert (want != pcr.FB_CLEAR).sum() > 500
            assert np.array_equal(got, want), (k, rule, int((got != want).sum()))


def test_static_render_refuses_beyond_its_scratch(ctx):
    layout = _terrain(20_000, 7)
    b, lo, med, hi = layout[:4]
    dev = DeviceArrays()
    try:
        p = [dev.put(np.ascontiguousarray(b, capi.BATCH_DTYPE))] + [dev.put(_pad4(x)) for x in (lo, med, hi)]
        keys = dev.put(np.zeros(64 * 64, np.uint64))
        view = _view(64, 64)
        lib = capi.load()
        rc = lib.pcr_render_static(p[0], capi.STATIC_MAX_BATCHES + 1, p[1], p[2], p[3], _pad4(lo).size,
                                   C.byref(view), C.byref(pcr.params(pcr.MODE_LAS2)), keys, None)
        assert rc == capi.PCR_E_INVALID and b"limit" in lib.pcr_last_error()
        rc = lib.pcr_render_static(p[0], len(b), p[1], p[2], p[3], _pad4(lo).size, C.byref(view),
                                   C.byref(pcr.params(pcr.MODE_HQS)), keys, None)
        assert rc == capi.PCR_E_INVALID
        rc = lib.pcr_render_static(p[0], len(b), p[1], p[2], p[3], _pad4(lo).size, C.byref(view),
                                   C.byref(pcr.params(pcr.MODE_LAS2, time_stages=1)), keys, None)
        assert rc == capi.PCR_E_INVALID
    finally:
        dev.close()
