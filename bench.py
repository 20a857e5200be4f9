#!/usr/bin/env python3
"""bench.py — 1e9-point adaptive-precision frames at 1920x1080 on one MI355X.

Workload: a city-scale cloud of 1e9 points (256 box buildings sampled on
their faces over a 2 km fBm terrain), generated on the device, Morton-sorted
into 10,240-point batches with 30-bit coordinates in three 10-bit planes.
One step = one frame of a 64-view seeded flythrough at 250 m: clear, cull +
precision select, rasterise (64-bit atomicMin), resolve to RGBA8.

Prints ONE JSON line (metric, value, ms_per_step, p50/p99 frame time,
roofline of the raster kernel, CPU-oracle baseline on the host cores).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import compute_rasterizer_amd as pcr  # noqa: E402

# HBM bytes per launch of the default raster kernel on the default workload,
# from the rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in profiles/
# (a counter pass cannot run inside the timed loop)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "raster_tiled_city1e9_1080p_traffic.json")


def measured_traffic(kernel: str, config: dict):
    """Counter-measured HBM bytes per launch for this kernel and workload, or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    want = d.get("config", {})
    if d.get("kernel") != kernel or any(config.get(k) != v for k, v in want.items()):
        return None
    return int(d["traffic_bytes_per_launch"])

HBM_PEAK_GBS = 8000.0


def flythrough(n, width, height, seed=42, extent=2000.0, altitude=250.0):
    """Seeded circular flythrough above the city, looking ahead and down."""
    rng = np.random.default_rng(seed)
    views = []
    c = extent / 2
    for k in range(n):
        a = 2 * math.pi * k / n + rng.uniform(-0.02, 0.02)
        r = 650.0 + rng.uniform(-40, 40)
        eye = (c + r * math.cos(a), c + r * math.sin(a), altitude + rng.uniform(-10, 10))
        ahead = a + math.pi / 2 + rng.uniform(-0.3, 0.3)
        tgt = (eye[0] + 400 * math.cos(ahead), eye[1] + 400 * math.sin(ahead), 60.0)
        views.append(pcr.look_at(eye, tgt, up=(0, 0, 1), fovy=60.0, near=1.0, far=5000.0, width=width,
                                 height=height))
    return views


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
    """Plain per-view frames vs one autotuned persistent launch for all views."""
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
    exact = all(np.array_equal(fbs[k].keys(), ref[k][0]) and np.array_equal(fbs[k].rgba(), ref[k][1])
                for k in range(nviews))
    mv.close()
    for fb in fbs:
        fb.close()
    pix = sum(v.width * v.height for v in views)
    return {"views": nviews, "megapixels": round(pix / 1e6, 2), "frames": frames,
            "plain_ms_per_frame": round(plain_s * 1e3, 3), "batched_ms_per_frame": round(mv_s * 1e3, 3),
            "plain_Gpoints_s": round(n_points * nviews / plain_s / 1e9, 1),
            "batched_Gpoints_s": round(n_points * nviews / mv_s / 1e9, 1),
            "speedup": round(plain_s / mv_s, 3),
            "tune": {"block_size": best.block_size, "points_per_thread": best.points_per_thread,
                     "tile_px": best.tile_px, "median_ms": round(best_ms, 3), "sweep_s": round(tune_s, 2)},
            "bit_exact_vs_plain": bool(exact)}


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


def stereo_hqs_leg(ctx, cloud, n_points, frames, width, height):
    """hqs for two eyes (the VR methods): per-eye renders vs one shared-decode
    multi-view launch that reads each batch once for both eyes."""
    pairs = stereo_pairs(frames, width, height)
    prm = pcr.params(pcr.MODE_HQS, tile_px=2048)
    plain_fbs = [ctx.framebuffer(width, height, pcr.FB_HQS) for _ in range(2)]
    mv_fbs = [ctx.framebuffer(width, height, pcr.FB_HQS) for _ in range(2)]

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
    mv.launch()
    ctx.synchronize()
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
    return {"eyes": 2, "width": width, "height": height, "frames": frames,
            "per_eye_ms_per_frame": round(plain_s * 1e3, 3), "shared_decode_ms_per_frame": round(mv_s * 1e3, 3),
            "shared_decode_Gpoints_s": round(n_points * 2 / mv_s / 1e9, 1), "speedup": round(plain_s / mv_s, 3),
            "tune": {"block_size": best.block_size, "points_per_thread": best.points_per_thread,
                     "tile_px": best.tile_px, "median_ms": round(best_ms, 3)},
            "bit_exact_vs_per_eye": exact}


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


def serving_leg(ctx, seconds, rate, budget_gb, sizes, seed=11):
    """Poisson(rate) render requests against device-resident LOD clouds, served
    with continuous batching under a framebuffer budget; p50/p99 latency."""
    import threading
    clouds, build_s = [], []
    for k, n in enumerate(sizes):
        t = time.perf_counter()
        cfg = pcr.gen_cfg(pcr.SCENE_CITY, n, seed=200 + k, num_buildings=256, terrain_frac=0.2)
        clouds.append(ctx.octree_generate(cfg, seed=k + 1))
        build_s.append(round(time.perf_counter() - t, 2))
    srv = pcr.Server(ctx, budget_bytes=int(budget_gb * (1 << 30)), max_batch=16, num_streams=4)
    ids = [srv.add_cloud(c) for c in clouds]
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * seconds * 1.5) + 16))
    arrivals = arrivals[arrivals < seconds]
    n_req = len(arrivals)
    which = rng.integers(0, len(clouds), n_req)
    views = multiview_views(n_req, seed=seed + 1)
    tickets = [None] * n_req
    ready = threading.Semaphore(0)
    start = time.perf_counter()

    errors = []

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
        _, ms = srv.wait(tickets[i], views[i].width * views[i].height)
        lat.append(ms)
    wall = time.perf_counter() - start
    th.join()
    st = srv.stats()

    # Saturated phase: the same kind of requests all submitted at once, served
    # with continuous batching, against rendering them one at a time through
    # the plain per-view path. Both hand every image to the host; a sample of
    # images is compared bit for bit.
    nb = min(n_req, 512)
    bviews = multiview_views(nb, seed=seed + 2)
    bwhich = rng.integers(0, len(clouds), nb)
    prm = pcr.params(pcr.MODE_NODES, lod_px=1.0)
    fbs = {}
    t0 = time.perf_counter()
    for i in range(nb):
        v = bviews[i]
        fb = fbs.get((v.width, v.height)) or fbs.setdefault((v.width, v.height), ctx.framebuffer(v.width, v.height))
        ctx.render(clouds[bwhich[i]], v, prm, fb)
        ctx.resolve(clouds[bwhich[i]], prm, fb)
        fb.rgba()
    ctx.synchronize()
    plain_s = time.perf_counter() - t0
    launches0 = srv.stats().launches
    t0 = time.perf_counter()
    bt = [srv.submit(ids[bwhich[i]], bviews[i], pcr.MODE_NODES, lod_px=1.0) for i in range(nb)]
    imgs = {}
    for i, t in enumerate(bt):
        # zero copy: the image is read in the server's pinned result memory
        img, _ = srv.wait_pinned(t, bviews[i].width * bviews[i].height)
        if i % 64 == 0:
            imgs[i] = img.copy()
        srv.release(t)
    served_s = time.perf_counter() - t0
    burst_launches = srv.stats().launches - launches0
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
    burst = {"requests": nb, "plain_s": round(plain_s, 3), "served_s": round(served_s, 3),
             "plain_Gpoints_s": round(pts / plain_s / 1e9, 1), "served_Gpoints_s": round(pts / served_s / 1e9, 1),
             "speedup": round(plain_s / served_s, 3), "launches": int(burst_launches),
             "mean_requests_per_launch": round(nb / max(1, burst_launches), 2),
             "bit_exact_vs_plain_sample": exact}
    srv.close()
    for c in clouds:
        c.close()
    lat = np.array(lat)
    return {"burst": burst, "clouds": list(sizes), "octree_build_s": build_s, "requests": n_req, "rate_per_s": rate,
            "seconds": round(wall, 2), "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
            "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
            "latency_ms_max": round(float(lat.max()), 3),
            "Gpoints_s": round(st.points_rendered / wall / 1e9, 2), "launches": int(st.launches),
            "mean_requests_per_launch": round(n_req / max(1, st.launches), 2),
            "fb_budget_GB": budget_gb, "fb_peak_GB": round(st.fb_bytes_peak / (1 << 30), 3),
            "lod_px": 1.0, "mode": "nodes"}


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


def cpu_baseline(ctx, cloud, views, gpu_fb, prm, frames, n_points):
    """Oracle (multithreaded C) on the host cores for `frames` full frames;
    the GPU framebuffers of the same views are compared bit for bit."""
    from oracle import orc
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, 16)
    lay = cloud.download()
    times, exact = [], True
    for j in range(frames):
        v = views[j]
        t0 = time.perf_counter()
        ref = orc.render_batched_mt(lay["batches"], lay["lo"], lay["med"], lay["hi"], v, 1, cores)
        times.append(time.perf_counter() - t0)
        ctx.render(cloud, v, prm, gpu_fb)
        exact = exact and bool(np.array_equal(gpu_fb.keys(), ref))
    del lay
    cpu_s = float(np.median(times))
    return {"value": round(n_points / cpu_s / 1e9, 4), "unit": "Gpoints/s", "cores": cores, "kind": "port",
            "sample": f"{frames} full 1920x1080 frames of the 1e9-point cloud (median {cpu_s:.2f} s/frame), "
                      f"oracle/pcr_oracle.c multithreaded fetch-min, framebuffers bit-exact to the GPU: {exact}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=128)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--points", type=int, default=1_000_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--views", type=int, default=64)
    ap.add_argument("--mode", default="las2", choices=["las", "las2"])
    ap.add_argument("--cpu-frames", type=int, default=2, help="0 skips the CPU baseline")
    ap.add_argument("--tile-px", type=int, default=2048, help="LDS tile size of the tiled raster (0 = direct)")
    ap.add_argument("--block", type=int, default=0, help="raster workgroup size (0 = default)")
    ap.add_argument("--mv-views", type=int, default=48, help="views of the batched multi-view leg (0 skips it)")
    ap.add_argument("--mv-frames", type=int, default=8)
    ap.add_argument("--stereo-frames", type=int, default=16, help="frames of the two-eye hqs leg (0 skips it)")
    ap.add_argument("--mv-batch-sizes", default="5120,20480",
                    help="extra layouts for the multi-view autotune sweep ('' skips it)")
    ap.add_argument("--serve-seconds", type=float, default=8.0, help="serving leg duration (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
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
        print(f"bench.py: this build renders on one GPU; --gpus {args.gpus} (WORLD_SIZE {world}) is not "
              f"supported (sharding batches across GPUs with a min reduce-scatter of the framebuffer is "
              f"designed, not built).", file=sys.stderr)
        sys.exit(2)

    ctx = pcr.Context(0)
    t = time.perf_counter()
    cfg = pcr.gen_cfg(pcr.SCENE_CITY, args.points, seed=42, num_buildings=256, terrain_frac=0.2)
    cloud = ctx.generate(cfg)
    gen_s = time.perf_counter() - t
    info = cloud.info()
    print(f"[bench] generated {info.num_points} points in {info.num_batches} batches ({gen_s:.1f} s)",
          file=sys.stderr, flush=True)

    W, H = args.width, args.height
    views = flythrough(args.views, W, H)
    mode = pcr.MODE_LAS2 if args.mode == "las2" else pcr.MODE_LAS
    tile = args.tile_px if mode == pcr.MODE_LAS2 else 0
    prm = pcr.params(mode, time_stages=0, tile_px=tile, block_size=args.block)
    # timed frames record HIP events around the raster stage only (two a
    # frame; each event costs the stream a few µs); the clear, cull and
    # resolve stage times come from a separate pass with all six events
    prm_timed = pcr.params(mode, time_stages=2, tile_px=tile, block_size=args.block)
    prm_stages = pcr.params(mode, time_stages=1, tile_px=tile, block_size=args.block)
    fb = ctx.framebuffer(W, H)

    # per-view algorithmic bytes of the raster kernel: headers of drawn batches,
    # the planes their precision level needs, one level byte per batch
    raster_bytes, frame_bytes, drawn_pts, hist = [], [], [], np.zeros(5, np.int64)
    for v in views:
        ctx.render(cloud, v, prm, fb)
        ctx.resolve(cloud, prm, fb)
        st = ctx.stats()
        drawn = st.batches_total - st.batches_culled
        plane = st.bytes_algorithmic - 64 * st.batches_total
        raster_bytes.append(plane + 64 * drawn + st.batches_total)
        # whole frame: select reads all headers; clear 8, resolve 8 + 4 gather + 4 write per pixel
        frame_bytes.append(st.bytes_algorithmic + 24 * W * H)
        drawn_pts.append(st.points_drawn)
        hist += np.array(list(st.batches_per_level))
    print(f"[bench] per-view stats done; level histogram {hist.tolist()}", file=sys.stderr, flush=True)

    for k in range(max(args.warmup, 1)):
        ctx.render(cloud, views[k % len(views)], prm_timed, fb)
        ctx.resolve(cloud, prm_timed, fb)
    ctx.synchronize()

    K = args.steps
    t0 = time.perf_counter()
    for k in range(K):
        ctx.render(cloud, views[k % len(views)], prm_timed, fb)
        ctx.resolve(cloud, prm_timed, fb)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ft = ctx.frame_times(K)
    print(f"[bench] timed {K} frames in {wall:.3f} s", file=sys.stderr, flush=True)
    # stage pass (not timed): 32 frames with events around every stage
    for k in range(32):
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
    achieved = sum(raster_bytes[i] for i in idx) / (raster_ms.sum() / 1e3) / 1e9
    frame_gbs = float(np.mean([frame_bytes[i] for i in idx])) / (frame_ms.mean() / 1e3) / 1e9
    value = info.num_points * K / wall / 1e9

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
        "config": {"workload": "city-1e9-adaptive-1080p-flythrough", "points": info.num_points,
                   "batches": info.num_batches, "batch_size": 10240, "width": W, "height": H,
                   "views": len(views), "mode": args.mode, "tile_px": args.tile_px},
        "frame_ms_p50": round(float(np.percentile(frame_ms, 50)), 4),
        "frame_ms_p99": round(float(np.percentile(frame_ms, 99)), 4),
        "stage_ms_mean": {"clear": round(float(st_ft[:, 0].mean()), 4), "cull": round(float(st_ft[:, 1].mean()), 4),
                          "raster": round(float(raster_ms.mean()), 4),
                          "resolve": round(float(st_ft[:, 3].mean()), 4)},
        "stage_ms_source": "raster: the timed frames; clear, cull, resolve: a 32-frame pass with all stage events",
        "points_in_drawn_batches_mean": int(np.mean([drawn_pts[i] for i in idx])),
        "level_histogram_per_frame": (hist / len(views)).round(1).tolist(),
        "frame_algorithmic_GBs": round(frame_gbs, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": ("k_raster_tiled<%d,%d>" % (args.block if args.block in (128, 256) else 256, args.tile_px)
                                if args.tile_px and mode == pcr.MODE_LAS2 else
                                "k_raster_batched<%d,%d>" % (args.block or (256 if mode == pcr.MODE_LAS2 else 128),
                                                             4 if mode == pcr.MODE_LAS2 else 1)),
                     "bytes_per_launch_mean": int(np.mean([raster_bytes[i] for i in idx]))},
        "resolve_edl_ms": round(edl_ms, 4),
        "generate_s": round(gen_s, 2),
    }
    out["roofline"]["traffic"] = measured_traffic(out["roofline"]["kernel"], out["config"])
    if args.mv_views > 0:
        out["multiview"] = multiview_leg(ctx, cloud, info.num_points, args.mv_views, args.mv_frames, args.tile_px)
        print(f"[bench] multiview {out['multiview']}", file=sys.stderr, flush=True)
    if args.stereo_frames > 0:
        out["stereo_hqs"] = stereo_hqs_leg(ctx, cloud, info.num_points, args.stereo_frames, W, H)
        print(f"[bench] stereo hqs {out['stereo_hqs']}", file=sys.stderr, flush=True)
    if args.cpu_frames > 0:
        out["cpu_baseline"] = cpu_baseline(ctx, cloud, views, fb, prm, args.cpu_frames, info.num_points)
    if args.mv_views > 0 and args.mv_batch_sizes:
        fb.close()
        cloud.close()
        sizes = [int(x) for x in args.mv_batch_sizes.split(",")]
        out["multiview"]["layout_sweep"] = layout_sweep(ctx, args.points, sizes, args.mv_views, args.mv_frames)
        print(f"[bench] layout sweep {out['multiview']['layout_sweep']}", file=sys.stderr, flush=True)
    if args.serve_seconds > 0:
        fb.close()
        cloud.close()
        sizes = [int(float(x)) for x in args.serve_clouds.split(",")]
        out["serving"] = serving_leg(ctx, args.serve_seconds, args.serve_rate, args.serve_budget_gb, sizes)
        print(f"[bench] serving {out['serving']}", file=sys.stderr, flush=True)
    if args.ingest_points > 0:
        out["ingest"] = ingest_leg(ctx, int(args.ingest_points))
        print(f"[bench] ingest {out['ingest']}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
