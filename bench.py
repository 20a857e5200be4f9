#!/usr/bin/env python3
"""bench.py — 1e9-point adaptive-precision frames at 1920x1080 on one MI355X.

Workload: a city-scale cloud of 1e9 points (256 box buildings sampled on
their faces over a 2 km fBm terrain), generated on the device, Morton-sorted
into 10,240-point batches with 30-bit coordinates in three 10-bit planes.
One step = one frame of a 64-view seeded flythrough 100 m above the ground:
clear, cull + precision select, rasterise (64-bit atomicMin), resolve to
RGBA8.

Prints ONE JSON line: metric, value, ms_per_step, p50/p99 frame time, the
raster kernel's roofline, the frame counters (points drawn, skipped as
occluded, atomics issued, pixels visible), the CPU-oracle baseline on the
host cores, and the other legs: the 250 m flythrough, a frame with the whole
city in view, batched multi-view, stereo hqs, hqs on the 1M-point cube,
serving with continuous batching, ingest.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import struct
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import compute_rasterizer_amd as pcr  # noqa: E402
from compute_rasterizer_amd import capi  # noqa: E402

# HBM bytes per launch of the default raster kernel on the default workload,
# from the rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in profiles/
# over the same 64 views (a counter pass cannot run inside the timed loop)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "raster_tiled_city1e9_1080p_traffic.json")

HBM_PEAK_GBS = 8000.0
FRAME_BYTES_PER_PIXEL = 24   # clear 8, resolve: key 8 + colour gather 4 + RGBA8 write 4
HQS_FRAME_BYTES_PER_PIXEL = 40   # depth clear 4 + sums clear 16, resolve: sums 16 + write 4


def measured_traffic(kernel: str, config: dict):
    """Counter-measured HBM bytes per launch and the algorithmic bytes of the
    same profiled launches, or None when the profile is of another kernel or
    workload."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    want = d.get("config", {})
    if d.get("kernel") != kernel or any(config.get(k) != v for k, v in want.items()):
        return None
    return d


def city_cfg(n_points):
    return pcr.gen_cfg(pcr.SCENE_CITY, n_points, seed=42, num_buildings=256, terrain_frac=0.2)


def flythrough(n, width, height, cfg, seed=42, agl=100.0, altitude=None):
    """Seeded circular flythrough over the city, looking ahead and down: the
    eye agl m above the ground (altitude=None) or at a fixed altitude above
    the datum, +-10 m."""
    rng = np.random.default_rng(seed)
    views = []
    c = cfg.extent / 2
    for k in range(n):
        a = 2 * math.pi * k / n + rng.uniform(-0.02, 0.02)
        r = 650.0 + rng.uniform(-40, 40)
        ex, ey = c + r * math.cos(a), c + r * math.sin(a)
        dz = rng.uniform(-10, 10)
        ahead = a + math.pi / 2 + rng.uniform(-0.3, 0.3)
        tx, ty = ex + 400 * math.cos(ahead), ey + 400 * math.sin(ahead)
        if altitude is None:
            eye = (ex, ey, pcr.ground_height(cfg, ex, ey) + agl + dz)
            tgt = (tx, ty, pcr.ground_height(cfg, tx, ty) + 20.0)
        else:
            eye, tgt = (ex, ey, altitude + dz), (tx, ty, 60.0)
        views.append(pcr.look_at(eye, tgt, up=(0, 0, 1), fovy=60.0, near=1.0, far=5000.0, width=width,
                                 height=height))
    return views


def overview_view(width, height, cfg):
    """The whole city in view: looking straight down from above its centre,
    the 2 km square (plus 5%) spanning the image height."""
    c = cfg.extent / 2
    d = 1.05 * c / math.tan(math.radians(30.0))
    top = cfg.amplitude + 80.0          # highest ground plus the tallest building
    return pcr.look_at((c, c, top + d), (c, c, 0.0), up=(0, 1, 0), fovy=60.0, near=1.0, far=5000.0, width=width,
                       height=height)


def frame_counters(ctx, cloud, views, prm):
    """Per-view counters of one frame each (pcr_stats with debug counters)."""
    out = []
    fb = ctx.framebuffer(views[0].width, views[0].height)
    for v in views:
        ctx.render(cloud, v, prm, fb)
        ctx.resolve(cloud, prm, fb)
        st = ctx.stats()
        drawn = st.batches_total - st.batches_culled
        planes = st.bytes_algorithmic - 64 * st.batches_total
        out.append({"drawn_batches": drawn, "points_drawn": st.points_drawn, "points_occluded": st.points_occluded,
                    # raster, algorithmic: headers of the frustum-visible batches, a level byte each, and the
                    # planes their precision levels need; read: less the planes of batches the occlusion test
                    # skipped
                    "raster_bytes": 64 * drawn + drawn + planes,
                    "raster_bytes_read": 64 * drawn + drawn + planes - st.bytes_occluded,
                    "frame_bytes": st.bytes_algorithmic + FRAME_BYTES_PER_PIXEL * v.width * v.height,
                    "atomics": st.atomics_issued, "pixels_visible": st.pixels_visible,
                    "levels": list(st.batches_per_level)})
    fb.close()
    return out


def timed_frames(ctx, cloud, views, prm_timed, fb, frames, warmup):
    """Wall time and per-frame raster times (HIP events around the raster
    stage) of `frames` frames cycling through the views."""
    for k in range(max(warmup, 1)):
        ctx.render(cloud, views[k % len(views)], prm_timed, fb)
        ctx.resolve(cloud, prm_timed, fb)
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(frames):
        ctx.render(cloud, views[k % len(views)], prm_timed, fb)
        ctx.resolve(cloud, prm_timed, fb)
    ctx.synchronize()
    return time.perf_counter() - t0, ctx.frame_times(frames)


def second_flythrough_leg(ctx, cloud, cfg, n_points, width, height, frames, tile):
    """The flythrough at a fixed 250 m above the datum (the previous default
    workload): frame time, raster time and counters."""
    views = flythrough(64, width, height, cfg, altitude=250.0)
    prm = pcr.params(pcr.MODE_LAS2, tile_px=tile, debug_counters=True)
    cnt = frame_counters(ctx, cloud, views, prm)
    fb = ctx.framebuffer(width, height)
    wall, ft = timed_frames(ctx, cloud, views, pcr.params(pcr.MODE_LAS2, tile_px=tile, time_stages=2), fb, frames, 8)
    fb.close()
    idx = [k % len(views) for k in range(frames)]
    raster = ft[:, 2].astype(np.float64)
    return {"altitude_m": 250.0, "frames": frames, "ms_per_frame": round(wall * 1e3 / frames, 4),
            "Gpoints_s": round(n_points * frames / wall / 1e9, 1),
            "raster_ms_mean": round(float(raster.mean()), 4),
            "points_drawn_mean": int(np.mean([cnt[i]["points_drawn"] for i in idx])),
            "points_occluded_mean": int(np.mean([cnt[i]["points_occluded"] for i in idx])),
            "raster_GBs": round(sum(cnt[i]["raster_bytes"] for i in idx) / (raster.sum() / 1e3) / 1e9, 1)}


def overview_leg(ctx, cloud, cfg, n_points, width, height, frames, tile):
    """One frame with the whole city in view (every batch drawn): the shape of
    the reference's 1e9-points-per-frame figure."""
    v = overview_view(width, height, cfg)
    prm = pcr.params(pcr.MODE_LAS2, tile_px=tile, debug_counters=True)
    cnt = frame_counters(ctx, cloud, [v], prm)[0]
    fb = ctx.framebuffer(width, height)
    wall, ft = timed_frames(ctx, cloud, [v], pcr.params(pcr.MODE_LAS2, tile_px=tile, time_stages=2), fb, frames, 4)
    fb.close()
    raster = ft[:, 2].astype(np.float64)
    return {"frames": frames, "ms_per_frame": round(wall * 1e3 / frames, 4),
            "Gpoints_s": round(n_points * frames / wall / 1e9, 1), "raster_ms_mean": round(float(raster.mean()), 4),
            "batches_drawn": int(cnt["drawn_batches"]), "points_drawn": int(cnt["points_drawn"]),
            "points_occluded": int(cnt["points_occluded"]), "pixels_visible": int(cnt["pixels_visible"]),
            "atomics_issued": int(cnt["atomics"]), "levels": cnt["levels"],
            "roofline_raster": {"achieved": round(cnt["raster_bytes"] / (raster.mean() / 1e3) / 1e9, 1),
                                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(cnt["raster_bytes"] / (raster.mean() / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                                "bytes_per_launch": int(cnt["raster_bytes"])}}


MV_SIZES = [(256, 256), (640, 480), (1280, 720), (1920, 1080), (2560, 1440)]


def multiview_views(n, seed=7, extent=2000.0):
    """Seeded cameras over the city at uneven resolutions."""
    rng = np.random.default_rng(seed)
    views = []
    for _ in range(n):
        w, h = MV_SIZES[int(rng.integers(0, len(MV_SIZES)))]
        x, y = rng.uniform(0.15 * extent, 0.85 * extent, 2)
        a = rng.uniform(0, 2 * math.pi)
        eye = (x, y, rng.uniform(150.0, 400.0))
        tgt = (x + 500 * math.cos(a), y + 500 * math.sin(a), rng.uniform(0.0, 80.0))
        views.append(pcr.look_at(eye, tgt, up=(0, 0, 1), fovy=60.0, near=1.0, far=5000.0, width=w, height=h))
    return views


def multiview_leg(ctx, cloud, n_points, nviews, frames, tile_px):
    """Plain per-view frames vs one autotuned persistent launch for all views,
    with both forms' algorithmic bytes (the batched form reads each batch's
    planes once, at the most precise level any view needs)."""
    views = multiview_views(nviews)
    fbs = [ctx.framebuffer(v.width, v.height) for v in views]
    prm = pcr.params(pcr.MODE_LAS2, tile_px=tile_px)

    def plain_frame():
        for v, fb in zip(views, fbs):
            ctx.render(cloud, v, prm, fb)
            ctx.resolve(cloud, prm, fb)

    plain_frame()
    ctx.synchronize()
    ref = [(fbs[k].keys(), fbs[k].rgba()) for k in range(nviews)]
    t0 = time.perf_counter()
    for _ in range(frames):
        plain_frame()
    ctx.synchronize()
    plain_s = (time.perf_counter() - t0) / frames

    t = time.perf_counter()
    best, best_ms = pcr.autotune(ctx, cloud, views, fbs, prm, iters=3)
    tune_s = time.perf_counter() - t
    mv = pcr.MultiView(ctx, cloud, views, fbs, prm, tune=best)
    mv.launch()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(frames):
        mv.launch()
    ctx.synchronize()
    mv_s = (time.perf_counter() - t0) / frames
    st = mv.stats()
    exact = all(np.array_equal(fbs[k].keys(), ref[k][0]) and np.array_equal(fbs[k].rgba(), ref[k][1])
                for k in range(nviews))
    mv.close()
    for fb in fbs:
        fb.close()
    pix = sum(v.width * v.height for v in views)
    batched_bytes = st.bytes_batched + FRAME_BYTES_PER_PIXEL * pix
    per_view_bytes = st.bytes_per_view_sum + FRAME_BYTES_PER_PIXEL * pix

    def roof(b, s):
        return {"bound": "hbm", "achieved": round(b / s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(b / s / 1e9 / HBM_PEAK_GBS, 4), "bytes_per_frame": int(b)}

    return {"views": nviews, "megapixels": round(pix / 1e6, 2), "frames": frames,
            "plain_ms_per_frame": round(plain_s * 1e3, 3), "batched_ms_per_frame": round(mv_s * 1e3, 3),
            "plain_Gpoints_s": round(n_points * nviews / plain_s / 1e9, 1),
            "batched_Gpoints_s": round(n_points * nviews / mv_s / 1e9, 1),
            "batched_drawn_Gpoints_s": round(st.points_per_view_sum / mv_s / 1e9, 1),
            "speedup": round(plain_s / mv_s, 3),
            "batches_any_view": int(st.batches_any_view),
            "roofline_batched": roof(batched_bytes, mv_s),
            "roofline_plain_per_view": roof(per_view_bytes, plain_s),
            "tune": {"block_size": best.block_size, "points_per_thread": best.points_per_thread,
                     "tile_px": best.tile_px, "median_ms": round(best_ms, 3), "sweep_s": round(tune_s, 2)},
            "bit_exact_vs_plain": bool(exact)}


def hqs_bytes(st, width, height):
    """Algorithmic bytes of one hqs frame from its counters: both passes read
    the frustum-visible batches' headers and planes, pass 2 also their
    colours; the framebuffer adds 40 B per pixel."""
    drawn = st.batches_total - st.batches_culled
    planes = st.bytes_algorithmic - 64 * st.batches_total
    return 2 * (64 * drawn + planes) + 4 * st.points_drawn + HQS_FRAME_BYTES_PER_PIXEL * width * height


def stereo_hqs_leg(ctx, cloud, n_points, frames, width, height):
    """hqs for two eyes (the VR methods): per-eye renders vs one shared-decode
    multi-view launch that reads each batch once for both eyes."""
    pairs = stereo_pairs(frames, width, height)
    prm = pcr.params(pcr.MODE_HQS, tile_px=2048)
    plain_fbs = [ctx.framebuffer(width, height, pcr.FB_HQS) for _ in range(2)]
    mv_fbs = [ctx.framebuffer(width, height, pcr.FB_HQS) for _ in range(2)]
    eye_bytes, occluded = [], []
    for pair in pairs:                      # counters of every eye (untimed)
        for v, fb in zip(pair, plain_fbs):
            ctx.render(cloud, v, prm, fb)
            st = ctx.stats()
            eye_bytes.append(hqs_bytes(st, width, height))
            occluded.append(st.points_occluded / max(1, 2 * st.points_drawn))

    def plain_frame(pair):
        for v, fb in zip(pair, plain_fbs):
            ctx.render(cloud, v, prm, fb)
            ctx.resolve(cloud, prm, fb)

    plain_frame(pairs[0])
    ctx.synchronize()
    t0 = time.perf_counter()
    for pair in pairs:
        plain_frame(pair)
    ctx.synchronize()
    plain_s = (time.perf_counter() - t0) / frames
    best, best_ms = pcr.autotune(ctx, cloud, pairs[0], mv_fbs, prm, iters=3)
    mv = pcr.MultiView(ctx, cloud, pairs[0], mv_fbs, prm, tune=best)
    mv_bytes = 0
    for pair in pairs:                      # shared-form bytes of every frame (untimed)
        mv.set_views(pair)
        mv.launch()
        st = mv.stats()
        mv_bytes += 2 * st.bytes_batched + 4 * st.points_any_view + 2 * HQS_FRAME_BYTES_PER_PIXEL * width * height
    mv_bytes /= len(pairs)
    t0 = time.perf_counter()
    for pair in pairs:
        mv.set_views(pair)
        mv.launch()
    ctx.synchronize()
    mv_s = (time.perf_counter() - t0) / frames
    exact = True
    for p, m in zip(plain_fbs, mv_fbs):   # both hold the last pair
        pd, pa = p.hqs()
        md, ma = m.hqs()
        exact = exact and bool(np.array_equal(pd, md) and np.array_equal(pa, ma) and
                               np.array_equal(p.rgba(), m.rgba()))
    mv.close()
    for fb in plain_fbs + mv_fbs:
        fb.close()
    per_frame = float(np.mean(eye_bytes)) * 2
    return {"eyes": 2, "width": width, "height": height, "frames": frames,
            "per_eye_ms_per_frame": round(plain_s * 1e3, 3), "shared_decode_ms_per_frame": round(mv_s * 1e3, 3),
            "shared_decode_Gpoints_s": round(n_points * 2 / mv_s / 1e9, 1), "speedup": round(plain_s / mv_s, 3),
            "occluded_fraction_of_pass_points": round(float(np.mean(occluded)), 3),
            "roofline_per_eye": {"bound": "hbm", "achieved": round(per_frame / plain_s / 1e9, 1),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(per_frame / plain_s / 1e9 / HBM_PEAK_GBS, 4),
                                 "bytes_per_frame": int(per_frame)},
            "roofline_shared": {"bound": "hbm", "achieved": round(mv_bytes / mv_s / 1e9, 1), "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": round(mv_bytes / mv_s / 1e9 / HBM_PEAK_GBS, 4),
                                "bytes_per_frame": int(mv_bytes)},
            "tune": {"block_size": best.block_size, "points_per_thread": best.points_per_thread,
                     "tile_px": best.tile_px, "median_ms": round(best_ms, 3)},
            "bit_exact_vs_per_eye": exact}


def hqs_cube_leg(ctx, frames, width=1920, height=1080):
    """hqs of the 1M-point cube (16 B/pt, unsorted) at 1920x1080: both passes
    read 16 B per point (its parity with the oracle is a GPU test)."""
    cfg = pcr.gen_cfg(pcr.SCENE_CUBE, 1_000_000, seed=1)
    cloud = ctx.generate(cfg, morton_sort=False, keep_points16=True)
    v = pcr.look_at((0.5, 0.5, 2.5), (0.5, 0.5, 0.5), up=(0, 1, 0), fovy=60.0, near=0.1, far=1000.0, width=width,
                    height=height)
    fb = ctx.framebuffer(width, height, pcr.FB_HQS)
    prm_t = pcr.params(pcr.MODE_HQS16, time_stages=2)
    wall, ft = timed_frames(ctx, cloud, [v], prm_t, fb, frames, 8)
    fb.close()
    cloud.close()
    raster = ft[:, 2].astype(np.float64)
    nbytes = 32 * cfg.num_points + HQS_FRAME_BYTES_PER_PIXEL * width * height
    return {"points": cfg.num_points, "width": width, "height": height, "frames": frames,
            "ms_per_frame": round(wall * 1e3 / frames, 4), "passes_ms_mean": round(float(raster.mean()), 4),
            "Gpoints_s": round(cfg.num_points * frames / wall / 1e9, 2),
            "roofline": {"bound": "hbm", "achieved": round(nbytes / (wall / frames) / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(nbytes / (wall / frames) / 1e9 / HBM_PEAK_GBS, 4),
                         "bytes_per_frame": int(nbytes), "kernels": "k_hqs_points16<0>, <1>, resolve"}}


def stereo_pairs(n, width, height, ipd=0.065, seed=3, extent=2000.0, altitude=250.0):
    """Two-eye views along a circular flythrough: each eye sits ipd/2 to the
    side of the head position, both looking at the same target offset."""
    rng = np.random.default_rng(seed)
    c = extent / 2
    pairs = []
    for k in range(n):
        a = 2 * math.pi * k / n + rng.uniform(-0.02, 0.02)
        head = np.array([c + 650.0 * math.cos(a), c + 650.0 * math.sin(a), altitude + rng.uniform(-10, 10)])
        ahead = a + math.pi / 2 + rng.uniform(-0.3, 0.3)
        fwd = np.array([400 * math.cos(ahead), 400 * math.sin(ahead), 60.0 - head[2]])
        right = np.array([fwd[1], -fwd[0], 0.0])
        right *= (ipd / 2) / np.linalg.norm(right)
        pairs.append([pcr.look_at(tuple(head + s * right), tuple(head + s * right + fwd), up=(0, 0, 1), fovy=90.0,
                                  near=0.1, far=5000.0, width=width, height=height) for s in (-1.0, 1.0)])
    return pairs


def layout_sweep(ctx, n_points, batch_sizes, nviews, frames):
    """Autotuned multi-view frame time per batch size (each a separate layout of
    the same 1e9-point scene; the batch AABBs, hence the decoded points, differ)."""
    views = multiview_views(nviews)
    fbs = [ctx.framebuffer(v.width, v.height) for v in views]
    prm = pcr.params(pcr.MODE_LAS2)
    res = []
    for bs in batch_sizes:
        cloud = ctx.generate(pcr.gen_cfg(pcr.SCENE_CITY, n_points, seed=42), batch_size=bs)
        best, best_ms = pcr.autotune(ctx, cloud, views, fbs, prm, iters=3)
        mv = pcr.MultiView(ctx, cloud, views, fbs, prm, tune=best)
        mv.launch()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(frames):
            mv.launch()
        ctx.synchronize()
        ms = (time.perf_counter() - t0) / frames * 1e3
        res.append({"batch_size": bs, "batched_ms_per_frame": round(ms, 3),
                    "batched_Gpoints_s": round(n_points * nviews / ms / 1e6, 1),
                    "tune": {"block_size": best.block_size, "points_per_thread": best.points_per_thread,
                             "tile_px": best.tile_px}})
        mv.close()
        cloud.close()
    for fb in fbs:
        fb.close()
    return res


def poisson_phase(srv, ids, n_clouds, rate, seconds, seed):
    """Poisson(rate) requests for `seconds`, each for a random cloud, view and
    resolution; latencies from submit to image ready (the server's clock). The
    client takes each image in place from the server's pinned result memory
    and releases it, so a slow client copy cannot hold the memory that
    admission needs."""
    import threading
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * seconds * 1.5) + 16))
    arrivals = arrivals[arrivals < seconds]
    n_req = len(arrivals)
    which = rng.integers(0, n_clouds, n_req)
    views = multiview_views(n_req, seed=seed + 1)
    tickets = [None] * n_req
    ready = threading.Semaphore(0)
    errors = []
    s0 = srv.stats()
    start = time.perf_counter()

    def producer():
        try:
            for i in range(n_req):
                dt = start + arrivals[i] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                tickets[i] = srv.submit(ids[which[i]], views[i], pcr.MODE_NODES, lod_px=1.0)
                ready.release()
        except Exception as e:  # surface the error in the main thread instead of blocking it
            errors.append(e)
            ready.release()

    th = threading.Thread(target=producer)
    th.start()
    lat = []
    for i in range(n_req):
        if not ready.acquire(timeout=60):
            raise RuntimeError(f"serving: request {i} was not submitted within 60 s")
        if errors:
            raise errors[0]
        _, ms = srv.wait_pinned(tickets[i])   # the image in place, in the server's pinned memory
        srv.release(tickets[i])
        lat.append(ms)
    wall = time.perf_counter() - start
    th.join()
    s1 = srv.stats()
    lat = np.array(lat)
    launches = s1.launches - s0.launches
    return {"rate_per_s": rate, "requests": n_req, "seconds": round(wall, 2),
            "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
            "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
            "latency_ms_max": round(float(lat.max()), 3), "launches": int(launches),
            "mean_requests_per_launch": round(n_req / max(1, launches), 2),
            "Gpoints_s": round((s1.points_rendered - s0.points_rendered) / wall / 1e9, 2),
            "drawn_Gpoints_s": round((s1.points_drawn - s0.points_drawn) / wall / 1e9, 3),
            "algorithmic_GBs": round((s1.bytes_algorithmic - s0.bytes_algorithmic) / wall / 1e9, 2)}


def serving_leg(ctx, seconds, rates, budget_gb, sizes, seed=11):
    """Render requests against device-resident LOD clouds, served with
    continuous batching under a framebuffer budget: Poisson arrivals at each
    rate (p50/p99 latency), then a burst against the plain per-view path."""
    clouds, build_s = [], []
    for k, n in enumerate(sizes):
        t = time.perf_counter()
        cfg = pcr.gen_cfg(pcr.SCENE_CITY, n, seed=200 + k, num_buildings=256, terrain_frac=0.2)
        clouds.append(ctx.octree_generate(cfg, seed=k + 1))
        build_s.append(round(time.perf_counter() - t, 2))
    srv = pcr.Server(ctx, budget_bytes=int(budget_gb * (1 << 30)), max_batch=16, num_streams=4)
    ids = [srv.add_cloud(c) for c in clouds]
    phases = [poisson_phase(srv, ids, len(clouds), rate, seconds if k == 0 else min(seconds, 2.0), seed + 10 * k)
              for k, rate in enumerate(rates)]

    # Burst: requests all submitted at once, served with continuous batching,
    # against rendering them one at a time through the plain per-view path.
    # Both hand every image to the host; a sample of images is compared bit
    # for bit.
    rng = np.random.default_rng(seed)
    nb = 512
    bviews = multiview_views(nb, seed=seed + 2)
    bwhich = rng.integers(0, len(clouds), nb)
    prm = pcr.params(pcr.MODE_NODES, lod_px=1.0)
    fbs = {}
    drawn = 0
    for i in range(nb):                     # counters of the plain renders (untimed)
        v = bviews[i]
        fb = fbs.get((v.width, v.height)) or fbs.setdefault((v.width, v.height), ctx.framebuffer(v.width, v.height))
        ctx.render(clouds[bwhich[i]], v, prm, fb)
        drawn += ctx.stats().points_drawn
    t0 = time.perf_counter()
    for i in range(nb):
        v = bviews[i]
        fb = fbs[(v.width, v.height)]
        ctx.render(clouds[bwhich[i]], v, prm, fb)
        ctx.resolve(clouds[bwhich[i]], prm, fb)
        fb.rgba()
    ctx.synchronize()
    plain_s = time.perf_counter() - t0
    s0 = srv.stats()
    t0 = time.perf_counter()
    bt = [srv.submit(ids[bwhich[i]], bviews[i], pcr.MODE_NODES, lod_px=1.0) for i in range(nb)]
    imgs = {}
    for i, t in enumerate(bt):
        # zero copy: the image is read in the server's pinned result memory
        img, _ = srv.wait_pinned(t)
        if i % 64 == 0:
            imgs[i] = img.copy()
        srv.release(t)
    served_s = time.perf_counter() - t0
    s1 = srv.stats()
    exact = True
    for i, img in imgs.items():
        v = bviews[i]
        fb = fbs[(v.width, v.height)]
        ctx.render(clouds[bwhich[i]], v, prm, fb)
        ctx.resolve(clouds[bwhich[i]], prm, fb)
        exact = exact and bool(np.array_equal(fb.rgba(), img))
    for fb in fbs.values():
        fb.close()
    pts = float(sum(int(clouds[k].info().num_points) for k in bwhich))
    pix = float(sum(v.width * v.height for v in bviews))
    served_bytes = (s1.bytes_algorithmic - s0.bytes_algorithmic) + FRAME_BYTES_PER_PIXEL * pix
    burst = {"requests": nb, "plain_s": round(plain_s, 3), "served_s": round(served_s, 3),
             "plain_Gpoints_s": round(pts / plain_s / 1e9, 1), "served_Gpoints_s": round(pts / served_s / 1e9, 1),
             "drawn_points": int(drawn), "served_drawn_Gpoints_s": round(drawn / served_s / 1e9, 2),
             "speedup": round(plain_s / served_s, 3), "launches": int(s1.launches - s0.launches),
             "mean_requests_per_launch": round(nb / max(1, s1.launches - s0.launches), 2),
             "roofline": {"bound": "hbm", "achieved": round(served_bytes / served_s / 1e9, 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(served_bytes / served_s / 1e9 / HBM_PEAK_GBS, 4),
                          "bytes": int(served_bytes),
                          "note": "whole served job: node headers and planes per launch plus 24 B/px, over wall time"},
             "bit_exact_vs_plain_sample": exact}
    st = srv.stats()
    srv.close()
    for c in clouds:
        c.close()
    return {"clouds": list(sizes), "octree_build_s": build_s, "fb_budget_GB": budget_gb,
            "fb_peak_GB": round(st.fb_bytes_peak / (1 << 30), 3), "lod_px": 1.0, "mode": "nodes",
            "poisson": phases, "burst": burst}


def laz_ingest(ctx, n_points):
    """LAZ ingest: two LAZ files of about n_points / 2 points each (point
    formats 3 and 7, made by repeating the chunks of two LAZ fixtures that the
    reference's LASzip wrote, tests/golden/laz/), decoded on the host threads
    and encoded on the device; against the same records as uncompressed LAS."""
    from compute_rasterizer_amd.datasets import laz_tile
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden")
    store = np.load(os.path.join(golden, "laz_records.npz"))
    lazs, lases = [], []
    for name, key in (("f3_extra3_variable", "f3_extra3_variable"), ("f7_v4", "f7_v3")):
        with open(os.path.join(golden, "laz", name + ".laz"), "rb") as f:
            seed = f.read()
        rec = store[key + "_records"]
        copies = max(1, int(n_points // 2 // rec.shape[0]))
        lazs.append(laz_tile(seed, copies))
        hdr = bytearray(store[key + "_las_header"].tobytes())
        n = copies * rec.shape[0]
        struct.pack_into("<I", hdr, 107, n)
        if len(hdr) >= 255:
            struct.pack_into("<Q", hdr, 247, n)
        lases.append(bytes(hdr) + np.tile(rec, (copies, 1)).tobytes())
    npts = sum(int(pcr.laz_info(d).num_points) for d in lazs)
    res = {"files": "formats 3 (pointwise, extra bytes) and 7 (layered v4)", "points": npts,
           "laz_bytes": sum(len(d) for d in lazs), "las_bytes": sum(len(d) for d in lases),
           "host_threads": min(16, os.cpu_count() or 1)}
    for fields, key in ((capi.LAZ_ALL_FIELDS, "decode_all_fields"), (capi.LAZ_XYZ_RGB, "decode_xyz_rgb")):
        t = time.perf_counter()
        for d in lazs:
            pcr.laz_decode(d, fields=fields)
        dt = time.perf_counter() - t
        res[key + "_Mpoints_s"] = round(npts / dt / 1e6, 1)
    t = time.perf_counter()
    a, _ = ctx.from_las(lazs)
    dt = time.perf_counter() - t
    res["laz_to_cloud_s"] = round(dt, 3)
    res["laz_to_cloud_Mpoints_s"] = round(npts / dt / 1e6, 1)
    t = time.perf_counter()
    b, _ = ctx.from_las(lases)
    dt = time.perf_counter() - t
    res["las_to_cloud_Mpoints_s"] = round(npts / dt / 1e6, 1)
    la, lb = a.download(), b.download()
    res["bit_exact_vs_las"] = bool(all(np.array_equal(la[k], lb[k]) for k in ("lo", "med", "hi", "rgba")) and
                                   np.array_equal(la["batches"].view(np.uint8), lb["batches"].view(np.uint8)))
    a.close()
    b.close()
    return res


def ingest_leg(ctx, n_las, seed=5):
    """LAS and Potree 2.0 ingest (host bytes -> device cloud, including the
    host-to-device copy of the records), and LOD frames of the Potree cloud
    with and without hqs."""
    from compute_rasterizer_amd.datasets import make_las, make_potree
    out = {}
    specs = [(2, (1, 2)), (3, (1, 3)), (7, (1, 4)), (8, (1, 4))]
    files = [make_las(n_las // 4 + k, point_format=f, version=v, seed=seed + k, extent=(2000.0, 2000.0, 200.0),
                      offset=(500000.0 + 500.0 * k, 4000000.0, 100.0))[0] for k, (f, v) in enumerate(specs)]
    warm, _ = ctx.from_las([make_las(1000, seed=seed)[0]])
    warm.close()
    t = time.perf_counter()
    cloud, first = ctx.from_las(files)
    las_s = time.perf_counter() - t
    n = int(cloud.info().num_points)
    out["las"] = {"files": len(files), "points": n, "bytes": sum(len(f) for f in files), "seconds": round(las_s, 3),
                  "Mpoints_s": round(n / las_s / 1e6, 1), "batches": int(first[-1])}
    cloud.close()
    del files
    out["laz"] = laz_ingest(ctx, n_las // 2)
    meta, hier, octree, _ = make_potree(seed=seed, max_depth=5, points_per_node=(2000, 8000))
    t = time.perf_counter()
    pc = ctx.from_potree(meta, hier, octree)
    pot_s = time.perf_counter() - t
    info = pc.info()
    out["potree"] = {"nodes": int(info.num_batches), "points": int(info.num_points), "bytes": len(octree) + len(hier),
                     "seconds": round(pot_s, 3), "Mpoints_s": round(info.num_points / pot_s / 1e6, 1)}
    pinfo = pcr.potree_info(meta)
    o = tuple(pinfo.min)
    v = pcr.look_at((o[0] + 512, o[1] - 600, o[2] + 700), (o[0] + 512, o[1] + 512, o[2] + 200), up=(0, 0, 1),
                    near=1.0, far=10000.0, origin=o, width=1920, height=1080)
    for name, mode, kind in (("nodes_ms", pcr.MODE_NODES, pcr.FB_KEY64), ("nodes_hqs_ms", pcr.MODE_NODES_HQS, pcr.FB_HQS)):
        fb = ctx.framebuffer(1920, 1080, kind)
        prm = pcr.params(mode, lod_px=1.0, time_stages=1, resolve_window=1, tile_px=2048)
        for _ in range(20):
            ctx.render(pc, v, prm, fb)
            ctx.resolve(pc, prm, fb)
        ft = ctx.frame_times(16)
        out["potree"][name] = round(float(np.median(ft[:, 4])), 4)
        fb.close()
    pc.close()
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(ctx, cloud, views, gpu_fb, prm, frames, n_points):
    """The oracle's renderer (multithreaded C) on the host cores: a warm-up
    frame, then the median of `frames` frames with all threads of this GPU's
    share of the host, each compared bit for bit with the GPU's framebuffer,
    and the median of `frames` frames on one thread."""
    from oracle import orc
    avail = len(os.sched_getaffinity(0))
    cores = min(avail, 16)   # a one-GPU job gets a 16-core share of the host, whatever the node has
    lay = cloud.download()
    b, lo, med, hi = lay["batches"], lay["lo"], lay["med"], lay["hi"]

    def run(v, threads):
        t0 = time.perf_counter()
        ref = orc.render_batched_mt(b, lo, med, hi, v, 1, threads)
        return time.perf_counter() - t0, ref

    run(views[0], cores)   # warm-up: pages the 16 GB layout in
    times, exact = [], True
    for j in range(frames):
        v = views[1 + j]
        dt, ref = run(v, cores)
        times.append(dt)
        ctx.render(cloud, v, prm, gpu_fb)
        exact = exact and bool(np.array_equal(gpu_fb.keys(), ref))
    t1 = [run(views[1 + j], 1)[0] for j in range(frames)]
    del lay, b, lo, med, hi
    cpu_s, one_s = float(np.median(times)), float(np.median(t1))
    return {"value": round(n_points / cpu_s / 1e9, 4), "unit": "Gpoints/s", "cores": cores, "kind": "port",
            "cores_available": avail, "cpu_model": cpu_model(),
            "value_1thread": round(n_points / one_s / 1e9, 4),
            "ms_per_frame": round(cpu_s * 1e3, 1), "ms_per_frame_1thread": round(one_s * 1e3, 1),
            "sample": f"1 warm-up then {frames} full 1920x1080 frames of the 1e9-point cloud on {cores} threads "
                      f"(median {cpu_s:.3f} s/frame) and {frames} on 1 thread (median {one_s:.2f} s/frame), "
                      f"oracle/pcr_oracle.c fetch-min renderer; framebuffers bit-exact to the GPU: {exact}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=128)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--points", type=int, default=1_000_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--views", type=int, default=64)
    ap.add_argument("--agl", type=float, default=100.0, help="flythrough height above the ground (m)")
    ap.add_argument("--mode", default="las2", choices=["las", "las2"])
    ap.add_argument("--cpu-frames", type=int, default=5, help="0 skips the CPU baseline")
    ap.add_argument("--tile-px", type=int, default=2048, help="LDS tile size of the tiled raster (0 = direct)")
    ap.add_argument("--block", type=int, default=0, help="raster workgroup size (0 = default)")
    ap.add_argument("--extra-frames", type=int, default=32, help="frames of the 250 m and overview legs (0 skips)")
    ap.add_argument("--mv-views", type=int, default=48, help="views of the batched multi-view leg (0 skips it)")
    ap.add_argument("--mv-frames", type=int, default=8)
    ap.add_argument("--stereo-frames", type=int, default=16, help="frames of the two-eye hqs leg (0 skips it)")
    ap.add_argument("--cube-hqs-frames", type=int, default=64, help="frames of the 1M-point cube hqs leg (0 skips)")
    ap.add_argument("--mv-batch-sizes", default="5120,20480",
                    help="extra layouts for the multi-view autotune sweep ('' skips it)")
    ap.add_argument("--serve-seconds", type=float, default=4.0, help="first serving phase duration (0 skips it)")
    ap.add_argument("--serve-rates", default="200,1000,2000,3000", help="Poisson arrival rates of the serving phases")
    ap.add_argument("--serve-budget-gb", type=float, default=16.0)
    ap.add_argument("--serve-clouds", default="10e6,100e6,500e6,1e9")
    ap.add_argument("--ingest-points", type=float, default=20e6, help="LAS ingest leg size (0 skips the leg)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 4 device copies of 1 GiB (reference for the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(4):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 4 << 28, "copies": 4}))
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: the renderer runs on one GPU; --gpus {args.gpus} (WORLD_SIZE {world}) is not "
              f"supported (sharding batches across GPUs with a min reduce-scatter of the framebuffer is "
              f"designed, not built).", file=sys.stderr)
        sys.exit(2)

    ctx = pcr.Context(0)
    t = time.perf_counter()
    cfg = city_cfg(args.points)
    cloud = ctx.generate(cfg)
    gen_s = time.perf_counter() - t
    info = cloud.info()
    print(f"[bench] generated {info.num_points} points in {info.num_batches} batches ({gen_s:.1f} s)",
          file=sys.stderr, flush=True)

    W, H = args.width, args.height
    views = flythrough(args.views, W, H, cfg, agl=args.agl)
    mode = pcr.MODE_LAS2 if args.mode == "las2" else pcr.MODE_LAS
    tile = args.tile_px if mode == pcr.MODE_LAS2 else 0
    prm = pcr.params(mode, time_stages=0, tile_px=tile, block_size=args.block)
    # timed frames record HIP events around the raster stage only (two a
    # frame; each event costs the stream a few µs); the clear, cull and
    # resolve stage times come from a separate pass with all six events
    prm_timed = pcr.params(mode, time_stages=2, tile_px=tile, block_size=args.block)
    prm_stages = pcr.params(mode, time_stages=1, tile_px=tile, block_size=args.block)
    cnt = frame_counters(ctx, cloud, views, pcr.params(mode, tile_px=tile, block_size=args.block,
                                                       debug_counters=True))
    hist = np.sum([c["levels"] for c in cnt], axis=0)
    print(f"[bench] per-view counters done; level histogram {hist.tolist()}", file=sys.stderr, flush=True)
    fb = ctx.framebuffer(W, H)
    K = args.steps
    wall, ft = timed_frames(ctx, cloud, views, prm_timed, fb, K, args.warmup)
    print(f"[bench] timed {K} frames in {wall:.3f} s", file=sys.stderr, flush=True)
    for k in range(32):   # stage pass (not timed): events around every stage
        ctx.render(cloud, views[k % len(views)], prm_stages, fb)
        ctx.resolve(cloud, prm_stages, fb)
    st_ft = ctx.frame_times(32)
    # the resolve again with eye-dome lighting (reference strength 0.0005),
    # on the last frame's framebuffer, outside the timed frames
    prm_edl = pcr.params(mode, edl_strength=0.0005)
    ctx.resolve(cloud, prm_edl, fb)
    ctx.synchronize()
    t1 = time.perf_counter()
    for _ in range(32):
        ctx.resolve(cloud, prm_edl, fb)
    ctx.synchronize()
    edl_ms = (time.perf_counter() - t1) / 32 * 1e3

    idx = [k % len(views) for k in range(K)]
    raster_ms = ft[:, 2].astype(np.float64)
    frame_ms = ft[:, 4].astype(np.float64)   # raster end to raster end
    frame_ms = frame_ms[frame_ms > 0]
    mean = lambda key: float(np.mean([cnt[i][key] for i in idx]))  # noqa: E731
    achieved = sum(cnt[i]["raster_bytes"] for i in idx) / (raster_ms.sum() / 1e3) / 1e9
    frame_gbs = mean("frame_bytes") / (frame_ms.mean() / 1e3) / 1e9
    value = info.num_points * K / wall / 1e9
    kernel = ("k_raster_tiled<%d,%d>" % (args.block if args.block in (128, 256) else 256, args.tile_px)
              if args.tile_px and mode == pcr.MODE_LAS2 else
              "k_raster_batched<%d,%d>" % (args.block or (256 if mode == pcr.MODE_LAS2 else 128),
                                          4 if mode == pcr.MODE_LAS2 else 1))
    config = {"workload": f"city-1e9-adaptive-1080p-flythrough-{args.agl:g}m-agl", "points": info.num_points,
              "batches": info.num_batches, "batch_size": 10240, "width": W, "height": H, "views": len(views),
              "mode": args.mode, "tile_px": args.tile_px}
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel,
            "bytes_per_launch_mean": int(mean("raster_bytes")),
            "bytes_read_per_launch_mean": int(mean("raster_bytes_read")),
            # framebuffer atomicMin rate: atomics the kernel issued (its own count, per frame) over the
            # timed raster time of the same views
            "atomics_G_per_s": round(sum(cnt[i]["atomics"] for i in idx) / (raster_ms.sum() / 1e3) / 1e9, 2)}
    prof = measured_traffic(kernel, config)
    if prof:
        roof["traffic"] = int(prof["traffic_bytes_per_launch"])
        roof["traffic_over_algorithmic"] = round(prof["traffic_bytes_per_launch"] /
                                                 prof["algorithmic_bytes_per_launch"], 3)
        # issue-side bound beside the HBM fraction, from the same profiled launches
        for k in ("valu_busy_frac", "tcc_hit_frac"):
            if k in prof:
                roof[k] = prof[k]
        roof["counters_source"] = "profiles/raster_tiled_city1e9_1080p_traffic.json (rocprofv3, the 64 views)"
    out = {
        "metric": "Gpoints/s (1e9-pt adaptive-precision 1920x1080 frames)",
        "value": round(value, 3),
        "unit": "Gpoints/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
        "frame_ms_p50": round(float(np.percentile(frame_ms, 50)), 4),
        "frame_ms_p99": round(float(np.percentile(frame_ms, 99)), 4),
        "stage_ms_mean": {"clear": round(float(st_ft[:, 0].mean()), 4), "cull": round(float(st_ft[:, 1].mean()), 4),
                          "raster": round(float(raster_ms.mean()), 4),
                          "resolve": round(float(st_ft[:, 3].mean()), 4)},
        "stage_ms_source": "raster: the timed frames; clear, cull, resolve: a 32-frame pass with all stage events",
        "points_in_drawn_batches_mean": int(mean("points_drawn")),
        "drawn_Gpoints_s": round(mean("points_drawn") * K / wall / 1e9, 1),
        "points_occluded_mean": int(mean("points_occluded")),
        "atomics_issued_mean": int(mean("atomics")),
        "pixels_visible_mean": int(mean("pixels_visible")),
        "level_histogram_per_frame": (hist / len(views)).round(1).tolist(),
        "frame_algorithmic_GBs": round(frame_gbs, 1),
        "roofline": roof,
        "resolve_edl_ms": round(edl_ms, 4),
        "generate_s": round(gen_s, 2),
    }
    if args.extra_frames > 0:
        out["flythrough_250m"] = second_flythrough_leg(ctx, cloud, cfg, info.num_points, W, H, args.extra_frames,
                                                       tile)
        print(f"[bench] 250 m flythrough {out['flythrough_250m']}", file=sys.stderr, flush=True)
        out["overview"] = overview_leg(ctx, cloud, cfg, info.num_points, W, H, args.extra_frames, tile)
        print(f"[bench] overview {out['overview']}", file=sys.stderr, flush=True)
    if args.mv_views > 0:
        out["multiview"] = multiview_leg(ctx, cloud, info.num_points, args.mv_views, args.mv_frames, args.tile_px)
        print(f"[bench] multiview {out['multiview']}", file=sys.stderr, flush=True)
    if args.stereo_frames > 0:
        out["stereo_hqs"] = stereo_hqs_leg(ctx, cloud, info.num_points, args.stereo_frames, W, H)
        print(f"[bench] stereo hqs {out['stereo_hqs']}", file=sys.stderr, flush=True)
    if args.cpu_frames > 0:
        out["cpu_baseline"] = cpu_baseline(ctx, cloud, views, fb, prm, args.cpu_frames, info.num_points)
        print(f"[bench] cpu baseline {out['cpu_baseline']}", file=sys.stderr, flush=True)
    fb.close()
    cloud.close()
    if args.cube_hqs_frames > 0:
        out["hqs_cube"] = hqs_cube_leg(ctx, args.cube_hqs_frames)
        print(f"[bench] hqs cube {out['hqs_cube']}", file=sys.stderr, flush=True)
    if args.mv_views > 0 and args.mv_batch_sizes:
        sizes = [int(x) for x in args.mv_batch_sizes.split(",")]
        out["multiview"]["layout_sweep"] = layout_sweep(ctx, args.points, sizes, args.mv_views, args.mv_frames)
        print(f"[bench] layout sweep {out['multiview']['layout_sweep']}", file=sys.stderr, flush=True)
    if args.serve_seconds > 0:
        sizes = [int(float(x)) for x in args.serve_clouds.split(",")]
        rates = [float(x) for x in args.serve_rates.split(",")]
        out["serving"] = serving_leg(ctx, args.serve_seconds, rates, args.serve_budget_gb, sizes)
        print(f"[bench] serving {out['serving']}", file=sys.stderr, flush=True)
    if args.ingest_points > 0:
        out["ingest"] = ingest_leg(ctx, int(args.ingest_points))
        print(f"[bench] ingest {out['ingest']}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
