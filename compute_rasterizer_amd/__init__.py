"""compute_rasterizer_amd — MI355X-native point-cloud software rasterizer.

Python face of libpcr.so (include/pcr.h). The per-frame call pair
`Context.render` + `Context.resolve` stands in for the reference's
Method::render (include/Method.h:10-24, compute_loop_las.h:117-304).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import (BATCH_DTYPE, BG_BLACK, BG_LAS, FB_CLEAR, FB_HQS, FB_KEY64, LOD_SIZE80, LOD_SPACING,
                   MODE_HQS, MODE_HQS16, MODE_LAS, MODE_LAS2, MODE_NODES, MODE_NODES_HQS, MODE_PLAIN16,
                   POINTS_PER_BATCH,
                   SCENE_CITY, SCENE_CUBE, SCENE_TERRAIN, BuildOpts, GenCfg, OctreeOpts, PcrError,
                   RenderParams, Stats, Tune, View, check, ptr)

__all__ = ["Context", "Cloud", "Framebuffer", "look_at", "params", "PcrError", "View", "RenderParams",
           "MODE_PLAIN16", "MODE_LAS", "MODE_LAS2", "MODE_HQS", "MODE_NODES", "MODE_HQS16", "MODE_NODES_HQS",
           "LOD_SPACING", "LOD_SIZE80", "SCENE_CUBE", "SCENE_TERRAIN", "SCENE_CITY", "FB_KEY64", "FB_HQS",
           "BG_LAS", "BG_BLACK", "FB_CLEAR", "POINTS_PER_BATCH", "BATCH_DTYPE"]


def look_at(eye, target, up=(0.0, 1.0, 0.0), fovy=60.0, near=0.1, far=1000.0, origin=(0.0, 0.0, 0.0),
            width=1280, height=720) -> View:
    """Camera -> pcr_view (double-precision glm lookAt/perspective, cast to float)."""
    v = View()
    check(capi.load().pcr_view_look_at(capi.dvec(eye), capi.dvec(target), capi.dvec(up), fovy, near, far,
                                       capi.dvec(origin), width, height, C.byref(v)))
    return v


def params(mode=MODE_LAS2, culling=1, background=None, lod_rule=LOD_SPACING, lod_px=1.0, block_size=0,
           clear_before=1, time_stages=0, tile_px=None, batch_begin=0, batch_end=0, resolve_window=0,
           edl_strength=0.0, occlusion_skip=True, debug_counters=False) -> RenderParams:
    """Render parameters; LAS2, NODES, HQS and NODES_HQS default to the LDS-tiled
    kernels (2048-px tile)."""
    if background is None:
        background = BG_BLACK if mode in (MODE_PLAIN16,) else BG_LAS
    if tile_px is None:
        tile_px = 2048 if mode in (MODE_LAS2, MODE_NODES, MODE_HQS, MODE_NODES_HQS) else 0
    return RenderParams(mode, culling, lod_rule, lod_px, background, block_size, clear_before, time_stages,
                        tile_px, batch_begin, batch_end, resolve_window, edl_strength,
                        0 if occlusion_skip else -1, int(bool(debug_counters)))


def potree_info(metadata_json: str) -> capi.PotreeInfo:
    """pcr_potree_info from a Potree 2.0 metadata.json (PotreeData.h:144-182)."""
    import json
    m = json.loads(metadata_json)
    if m.get("encoding", "DEFAULT") != "DEFAULT":
        raise ValueError(f"Potree encoding {m.get('encoding')!r} is not read (only DEFAULT)")
    info = capi.PotreeInfo()
    for a in range(3):
        info.scale[a] = m["scale"][a]
        info.offset[a] = m["offset"][a]
        info.min[a] = m["boundingBox"]["min"][a]
        info.max[a] = m["boundingBox"]["max"][a]
    info.spacing = m["spacing"]
    info.num_points = m["points"]
    info.first_chunk_size = m["hierarchy"]["firstChunkSize"]
    bpp, rgb = 0, 0
    for att in m["attributes"]:
        if att["name"] == "rgb":
            rgb = bpp
        bpp += att["size"]
    info.bytes_per_point, info.rgb_offset = bpp, rgb
    return info


def laz_info(data) -> capi.LazInfo:
    """Header, LASzip VLR and chunk table of LAZ file bytes (host only)."""
    b = np.frombuffer(data, np.uint8)
    out = capi.LazInfo()
    check(capi.load().pcr_laz_info_get(b.ctypes.data, b.size, C.byref(out)))
    return out


def laz_chunk_table(data):
    """(chunk_start, chunk_first) of a LAZ file: chunk c holds records
    [chunk_first[c], chunk_first[c+1]) in bytes [chunk_start[c], chunk_start[c+1])."""
    b = np.frombuffer(data, np.uint8)
    n = C.c_uint32()
    check(capi.load().pcr_laz_chunk_table(b.ctypes.data, b.size, None, None, 0, C.byref(n)))
    starts, firsts = np.zeros(n.value + 1, np.uint64), np.zeros(n.value + 1, np.uint64)
    if n.value:
        check(capi.load().pcr_laz_chunk_table(b.ctypes.data, b.size, ptr(starts), ptr(firsts), n.value + 1,
                                              C.byref(n)))
    return starts, firsts


def laz_decode(data, first=0, count=None, fields=capi.LAZ_ALL_FIELDS, threads=0) -> np.ndarray:
    """Records [first, first + count) of LAZ file bytes as a (count, record_length)
    uint8 array: the uncompressed LAS records (host only)."""
    b = np.frombuffer(data, np.uint8)
    info = laz_info(data)
    n = int(info.num_points) - first if count is None else count
    out = np.empty((max(n, 0), info.record_length), np.uint8)
    check(capi.load().pcr_laz_decode(b.ctypes.data, b.size, first, n, out.ctypes.data if n else None, fields,
                                     threads))
    return out


def gen_cfg(scene, n, seed=1, extent=2000.0, amplitude=150.0, jitter=0.5, num_buildings=256,
            terrain_frac=0.2) -> GenCfg:
    return GenCfg(scene, num_buildings, seed, n, extent, amplitude, jitter, terrain_frac)


def ground_height(cfg: GenCfg, x: float, y: float) -> float:
    """Ground height of a TERRAIN or CITY scene at (x, y) (pcr_scene_ground_height)."""
    h = C.c_float()
    check(capi.load().pcr_scene_ground_height(C.byref(cfg), x, y, C.byref(h)))
    return h.value


def encode_batches_host(xyz: np.ndarray, batch_size: int = POINTS_PER_BATCH):
    """Host encoder (LasLoaderSparse.cpp:37-212): returns (batches, lo, med, hi)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    n = xyz.shape[0]
    nb = (n + batch_size - 1) // batch_size
    b = np.zeros(nb, BATCH_DTYPE)
    lo, med, hi = (np.zeros(n, np.uint32) for _ in range(3))
    check(capi.load().pcr_encode_batches_host(ptr(xyz), n, batch_size, ptr(b), ptr(lo), ptr(med), ptr(hi)))
    return b, lo, med, hi


def octree_layout_host(xyz, rgba, seed=1, root_spacing_div=128.0, max_node_points=20000, max_depth=12):
    """Potree-style node layout on the host: returns (nodes, lo, med, hi, rgba)."""
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    rgba = np.ascontiguousarray(rgba, np.uint32)
    o = OctreeOpts(seed, root_spacing_div, max_node_points, max_depth)
    nn, ns = C.c_uint32(0), C.c_uint64(0)
    lib = capi.load()
    check(lib.pcr_octree_layout_host(ptr(xyz), ptr(rgba), xyz.shape[0], C.byref(o), None, C.byref(nn), None, None,
                                     None, None, C.byref(ns)))
    nodes = np.zeros(nn.value, BATCH_DTYPE)
    lo, med, hi, col = (np.zeros(ns.value, np.uint32) for _ in range(4))
    check(lib.pcr_octree_layout_host(ptr(xyz), ptr(rgba), xyz.shape[0], C.byref(o), ptr(nodes), C.byref(nn),
                                     ptr(lo), ptr(med), ptr(hi), ptr(col), C.byref(ns)))
    return nodes, lo, med, hi, col


class Framebuffer:
    def __init__(self, ctx: "Context", width: int, height: int, kind: int = FB_KEY64):
        self.ctx, self.width, self.height, self.kind = ctx, width, height, kind
        h = C.c_void_p()
        check(capi.load().pcr_fb_create(ctx.h, width, height, kind, C.byref(h)))
        self.h = h

    @property
    def npix(self):
        return self.width * self.height

    def clear(self):
        check(capi.load().pcr_fb_clear(self.ctx.h, self.h))

    def keys(self) -> np.ndarray:
        out = np.empty(self.npix, np.uint64)
        check(capi.load().pcr_fb_readback(self.ctx.h, self.h, ptr(out)))
        return out

    def hqs(self):
        d = np.empty(self.npix, np.uint32)
        a = np.empty(4 * self.npix, np.uint32)
        check(capi.load().pcr_fb_readback_hqs(self.ctx.h, self.h, ptr(d), ptr(a)))
        return d, a

    def rgba(self) -> np.ndarray:
        out = np.empty(self.npix, np.uint32)
        check(capi.load().pcr_fb_readback_rgba(self.ctx.h, self.h, ptr(out)))
        return out

    def close(self):
        if self.h:
            capi.load().pcr_fb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Cloud:
    def __init__(self, ctx: "Context", handle: C.c_void_p):
        self.ctx, self.h = ctx, handle

    def info(self) -> capi.CloudInfo:
        i = capi.CloudInfo()
        check(capi.load().pcr_cloud_info_get(self.h, C.byref(i)))
        return i

    def download(self, batched=True, points16=False):
        i = self.info()
        out = {}
        if batched and i.has_batches:
            out["batches"] = np.zeros(i.num_batches, BATCH_DTYPE)
            for k in ("lo", "med", "hi", "rgba"):
                out[k] = np.zeros(i.num_points, np.uint32)
        if points16 and i.has_points16:
            out["xyz16"] = np.zeros((i.num_points, 3), np.float32)
            out["rgba16"] = np.zeros(i.num_points, np.uint32)
        check(capi.load().pcr_cloud_download(self.h, ptr(out.get("batches")), ptr(out.get("lo")),
                                             ptr(out.get("med")), ptr(out.get("hi")), ptr(out.get("rgba")),
                                             ptr(out.get("xyz16")), ptr(out.get("rgba16"))))
        return out

    def close(self):
        if self.h:
            capi.load().pcr_cloud_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MultiView:
    """One persistent launch rendering many views of a cloud (pcr_multiview)."""

    def __init__(self, ctx: "Context", cloud: Cloud, views, fbs, prm: RenderParams, tune: Tune | None = None,
                 use_graph=True):
        self.ctx, self.cloud, self.fbs = ctx, cloud, list(fbs)
        n = len(views)
        self._views = (View * n)(*views)
        self._fbs = (C.c_void_p * n)(*[f.h.value for f in self.fbs])
        h = C.c_void_p()
        check(capi.load().pcr_multiview_create(ctx.h, cloud.h, self._views, n, self._fbs, C.byref(prm),
                                               C.byref(tune) if tune is not None else None, int(use_graph),
                                               C.byref(h)))
        self.h = h

    def set_views(self, views):
        n = len(views)
        self._views = (View * n)(*views)
        check(capi.load().pcr_multiview_set_views(self.h, self._views, n))

    def launch(self):
        check(capi.load().pcr_multiview_launch(self.h))

    def stats(self) -> capi.MultiViewStats:
        """Batches and algorithmic bytes of the last launch (synchronises)."""
        st = capi.MultiViewStats()
        check(capi.load().pcr_multiview_stats_get(self.h, C.byref(st)))
        return st

    def close(self):
        if self.h:
            capi.load().pcr_multiview_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching render server (pcr_server)."""

    def __init__(self, ctx: "Context", budget_bytes: int, max_batch=16, num_streams=4):
        h = C.c_void_p()
        check(capi.load().pcr_server_create(ctx.h, budget_bytes, max_batch, num_streams, C.byref(h)))
        self.ctx, self.h, self.clouds = ctx, h, []
        self._npix = {}   # ticket -> pixels of its image, recorded at submit

    def _pixels(self, ticket: int, npix):
        """The ticket's image size; a caller's npix must agree with it (the
        library copies exactly width * height pixels)."""
        n = self._npix.get(ticket)
        if n is None:
            raise PcrError(capi.PCR_E_INVALID, f"unknown ticket {ticket}")
        if npix is not None and npix != n:
            raise ValueError(f"ticket {ticket} has {n} pixels, not {npix}")
        return n

    def add_cloud(self, cloud: Cloud) -> int:
        cid = C.c_int()
        check(capi.load().pcr_server_add_cloud(self.h, cloud.h, C.byref(cid)))
        self.clouds.append(cloud)
        return cid.value

    def submit(self, cloud_id: int, view: View, mode=MODE_LAS2, lod_px=1.0) -> int:
        req = capi.Request(cloud_id, mode, lod_px, 0, view)
        t = C.c_uint64()
        check(capi.load().pcr_server_submit(self.h, C.byref(req), C.byref(t)))
        self._npix[t.value] = view.width * view.height
        return t.value

    def poll(self, ticket: int, npix: int | None = None):
        """(image, latency_ms) or None while the request is in flight."""
        out = np.empty(self._pixels(ticket, npix), np.uint32)
        lat = C.c_double()
        rc = capi.load().pcr_server_poll(self.h, ticket, ptr(out), C.byref(lat))
        if rc == capi.PCR_E_NOT_READY:
            return None
        self._npix.pop(ticket, None)
        check(rc)
        return out, lat.value

    def wait(self, ticket: int, npix: int | None = None):
        out = np.empty(self._pixels(ticket, npix), np.uint32)
        lat = C.c_double()
        rc = capi.load().pcr_server_wait(self.h, ticket, ptr(out), C.byref(lat))
        self._npix.pop(ticket, None)
        check(rc)
        return out, lat.value

    def wait_pinned(self, ticket: int, npix: int | None = None):
        """(image, latency_ms) without a copy: the image is a view of the
        server's pinned result memory, valid until release(ticket)."""
        n = self._pixels(ticket, npix)
        p = C.c_void_p()
        lat = C.c_double()
        check(capi.load().pcr_server_wait_pinned(self.h, ticket, C.byref(p), C.byref(lat)))
        img = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), shape=(n,))
        return img, lat.value

    def release(self, ticket: int):
        check(capi.load().pcr_server_release(self.h, ticket))
        self._npix.pop(ticket, None)

    def cancel(self, ticket: int):
        check(capi.load().pcr_server_cancel(self.h, ticket))
        self._npix.pop(ticket, None)

    def stats(self) -> capi.ServerStats:
        s = capi.ServerStats()
        check(capi.load().pcr_server_stats_get(self.h, C.byref(s)))
        return s

    def close(self):
        if self.h:
            capi.load().pcr_server_destroy(self.h)
            self.h = None


def autotune(ctx: "Context", cloud: Cloud, views, fbs, prm: RenderParams, iters=5):
    """Sweep block size x points per thread for a multi-view frame: (Tune, ms)."""
    n = len(views)
    vs = (View * n)(*views)
    fs = (C.c_void_p * n)(*[f.h.value for f in fbs])
    best, ms = Tune(), C.c_float()
    check(capi.load().pcr_autotune(ctx.h, cloud.h, vs, n, fs, C.byref(prm), iters, C.byref(best), C.byref(ms)))
    return best, ms.value


class Context:
    """One HIP device + stream (pcr_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(capi.load().pcr_create(device, C.byref(h)))
        self.h = h

    # --- clouds
    def upload_points16(self, xyz, rgba) -> Cloud:
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        rgba = np.ascontiguousarray(rgba, np.uint32)
        h = C.c_void_p()
        check(capi.load().pcr_cloud_upload_points16(self.h, ptr(xyz), ptr(rgba), xyz.shape[0], C.byref(h)))
        return Cloud(self, h)

    def upload_batched(self, batches, lo, med, hi, rgba) -> Cloud:
        batches = np.ascontiguousarray(batches, BATCH_DTYPE)
        arrs = [np.ascontiguousarray(a, np.uint32) for a in (lo, med, hi, rgba)]
        h = C.c_void_p()
        check(capi.load().pcr_cloud_upload_batched(self.h, ptr(batches), len(batches), *[ptr(a) for a in arrs],
                                                   arrs[0].shape[0], C.byref(h)))
        return Cloud(self, h)

    def generate(self, cfg: GenCfg, batch_size=POINTS_PER_BATCH, morton_sort=True, keep_points16=False) -> Cloud:
        h = C.c_void_p()
        opts = BuildOpts(batch_size, int(morton_sort), int(keep_points16))
        check(capi.load().pcr_cloud_generate(self.h, C.byref(cfg), C.byref(opts), C.byref(h)))
        return Cloud(self, h)

    def octree_host(self, xyz, rgba, seed=1, root_spacing_div=128.0, max_node_points=20000, max_depth=12) -> Cloud:
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        rgba = np.ascontiguousarray(rgba, np.uint32)
        h = C.c_void_p()
        o = OctreeOpts(seed, root_spacing_div, max_node_points, max_depth)
        check(capi.load().pcr_cloud_build_octree_host(self.h, ptr(xyz), ptr(rgba), xyz.shape[0], C.byref(o),
                                                      C.byref(h)))
        return Cloud(self, h)

    def octree_generate(self, cfg: GenCfg, seed=1, root_spacing_div=128.0, max_node_points=20000,
                        max_depth=12) -> Cloud:
        """Generate a scene on the device and build its LOD octree there."""
        h = C.c_void_p()
        o = OctreeOpts(seed, root_spacing_div, max_node_points, max_depth)
        check(capi.load().pcr_octree_generate(self.h, C.byref(cfg), C.byref(o), C.byref(h)))
        return Cloud(self, h)

    def from_potree(self, metadata_json: str, hierarchy: bytes, octree: bytes) -> Cloud:
        """LOD cloud from a Potree 2.0 dataset (metadata.json text, hierarchy.bin and octree.bin bytes)."""
        info = potree_info(metadata_json)
        hb = np.frombuffer(hierarchy, np.uint8)
        ob = np.frombuffer(octree, np.uint8)
        h = C.c_void_p()
        check(capi.load().pcr_cloud_from_potree(self.h, C.byref(info), ptr(hb), hb.size, ptr(ob), ob.size,
                                                C.byref(h)))
        return Cloud(self, h)

    def from_las(self, files):
        """One cloud from LAS file bytes: (Cloud, first_batch) with file k's batches in
        [first_batch[k], first_batch[k+1])."""
        bufs = [np.frombuffer(d, np.uint8) for d in files]
        arr = (capi.LasFile * len(bufs))(*[capi.LasFile(b.ctypes.data, b.size) for b in bufs])
        first = np.zeros(len(bufs) + 1, np.uint32)
        h = C.c_void_p()
        check(capi.load().pcr_cloud_from_las(self.h, arr, len(bufs), ptr(first), C.byref(h)))
        return Cloud(self, h), first

    def framebuffer(self, width, height, kind=FB_KEY64) -> Framebuffer:
        return Framebuffer(self, width, height, kind)

    # --- frame
    def render(self, cloud: Cloud, view: View, prm: RenderParams, fb: Framebuffer):
        check(capi.load().pcr_render(self.h, cloud.h, C.byref(view), C.byref(prm), fb.h))

    def resolve(self, cloud: Cloud, prm: RenderParams, fb: Framebuffer):
        check(capi.load().pcr_resolve(self.h, cloud.h, C.byref(prm), fb.h))

    def stats(self) -> Stats:
        s = Stats()
        check(capi.load().pcr_stats_get(self.h, C.byref(s)))
        return s

    def frame_times(self, n: int) -> np.ndarray:
        """[frames, 5] ms of clear, cull, raster, resolve, whole frame for the last n timed frames."""
        out = np.zeros((n, 5), np.float32)
        got = C.c_int()
        check(capi.load().pcr_frame_times(self.h, n, ptr(out), C.byref(got)))
        return out[:got.value]

    def lod_selection(self, cloud: Cloud) -> np.ndarray:
        n = cloud.info().num_batches
        out = np.zeros(n, np.uint8)
        check(capi.load().pcr_lod_selection(self.h, cloud.h, ptr(out), n))
        return out

    def synchronize(self):
        check(capi.load().pcr_synchronize(self.h))

    def close(self):
        if self.h:
            capi.load().pcr_destroy(self.h)
            self.h = None
