"""GPU: batched multi-view launch and the continuous-batching server.

Each view of a multi-view launch, and each served request, must equal the
single-view render of the same view bit for bit (and that render equals the
oracle, checked here for one view per test and exhaustively in
test_gpu_parity.py).
"""
import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from compute_rasterizer_amd import capi
from oracle import orc

pytestmark = pytest.mark.gpu

SIZES = [(256, 256), (640, 480), (1280, 720), (1920, 1080), (2560, 1440)]


def views_for(sizes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for w, h in sizes:
        x, y = rng.uniform(300, 1700, 2)
        ang = rng.uniform(0, 2 * np.pi)
        eye = (x, y, rng.uniform(120, 400))
        tgt = (x + 300 * np.cos(ang), y + 300 * np.sin(ang), 40.0)
        out.append(pcr.look_at(eye, tgt, up=(0, 0, 1), near=1.0, far=5000.0, width=w, height=h))
    return out


@pytest.fixture(scope="module")
def city(ctx):
    cloud = ctx.generate(pcr.gen_cfg(pcr.SCENE_CITY, 2_000_000, seed=17, num_buildings=32))
    yield cloud
    cloud.close()


@pytest.fixture(scope="module")
def nodes(ctx):
    xyz, rgba = orc.generate(pcr.gen_cfg(pcr.SCENE_TERRAIN, 300_000, seed=19))
    cloud = ctx.octree_host(xyz, rgba, seed=4)
    yield cloud
    cloud.close()


def single_view(ctx, cloud, view, prm):
    fb = ctx.framebuffer(view.width, view.height)
    ctx.render(cloud, view, prm, fb)
    ctx.resolve(cloud, prm, fb)
    out = fb.keys(), fb.rgba()
    fb.close()
    return out


TUNES = {"tiled128": (128, 40, 0, 2048), "tiled256_4096": (256, 40, 0, 4096), "untiled256": (256, 40, 0, 0),
         "untiled64_160": (64, 160, 0, 0)}


@pytest.mark.parametrize("use_graph", [1, 0])
@pytest.mark.parametrize("tune", sorted(TUNES))
def test_multiview_equals_single_views(ctx, city, use_graph, tune):
    prm = pcr.params(pcr.MODE_LAS2)
    views = views_for(SIZES * 2, seed=1)
    fbs = [ctx.framebuffer(v.width, v.height) for v in views]
    mv = pcr.MultiView(ctx, city, views, fbs, prm, tune=pcr.Tune(*TUNES[tune]), use_graph=bool(use_graph))
    for rnd in range(3):
        if rnd:
            views = views_for(SIZES * 2, seed=1 + rnd)
            mv.set_views(views)
        mv.launch()
        ctx.synchronize()
        # eager work between replays (a single-view render on the same stream)
        for v, fb in zip(views, fbs):
            keys, img = single_view(ctx, city, v, prm)
            assert np.array_equal(fb.keys(), keys)
            assert np.array_equal(fb.rgba(), img)
    lay = city.download()
    ref, _ = orc.render_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], views[1], 1)
    assert np.array_equal(fbs[1].keys(), ref)
    mv.close()


def eye_pairs(seed, n_pairs, ipd=0.065, sizes=((1280, 720), (1280, 720))):
    """n_pairs stereo pairs (eyes ipd apart) of the city from about 250 m, the
    two eyes at sizes[0] and sizes[1]."""
    rng = np.random.default_rng(seed)
    views = []
    for _ in range(n_pairs):
        x, y = rng.uniform(600, 1400, 2)
        eye = np.array([x, y - 700.0, rng.uniform(200, 300)])
        tgt = np.array([x, y + 300.0, 40.0])
        for off, (w, h) in zip((0.0, ipd), sizes):
            r = np.array([off, 0.0, 0.0])
            views.append(pcr.look_at(tuple(eye + r), tuple(tgt + r), up=(0, 0, 1), near=1.0, far=5000.0,
                                     width=w, height=h))
    return views


@pytest.mark.parametrize("occl", [True, False])
@pytest.mark.parametrize("n_pairs,tune,sizes", [(1, (256, 40, 0, 2048), ((1280, 720), (1280, 720))),
                                               (1, (128, 40, 0, 1024), ((1280, 720), (1280, 720))),
                                               (2, (256, 40, 0, 4096), ((1280, 720), (1280, 720))),
                                               (1, (256, 40, 0, 2048), ((1280, 720), (1920, 1080)))])
def test_view_pairs_equal_oracle(ctx, city, n_pairs, tune, sizes, occl):
    """Two adjacent views that draw a batch at one level take it in one pass
    over the points (the two halves of the LDS tile); stereo pairs take that
    path for most batches, also when the two views differ in size. Every
    view's keys equal the oracle's."""
    prm = pcr.params(pcr.MODE_LAS2, occlusion_skip=occl)
    views = eye_pairs(11 + n_pairs, n_pairs, sizes=sizes)
    fbs = [ctx.framebuffer(v.width, v.height) for v in views]
    mv = pcr.MultiView(ctx, city, views, fbs, prm, tune=pcr.Tune(*tune))
    mv.launch()
    ctx.synchronize()
    lay = city.download()
    for k, v in enumerate(views):
        ref, _ = orc.render_batched(lay["batches"], lay["lo"], lay["med"], lay["hi"], v, 1)
        keys = fbs[k].keys()
        assert (ref != pcr.FB_CLEAR).sum() > 1000
        assert np.array_equal(keys, ref), (k, int((keys != ref).sum()))
        kk, img = single_view(ctx, city, v, prm)
        assert np.array_equal(fbs[k].rgba(), img), k
    mv.close()
    for fb in fbs:
        fb.close()


def test_multiview_lod_nodes(ctx, nodes):
    prm = pcr.params(pcr.MODE_NODES, lod_px=1.0)
    views = views_for(SIZES, seed=7)
    fbs = [ctx.framebuffer(v.width, v.height) for v in views]
    mv = pcr.MultiView(ctx, nodes, views, fbs, prm)
    mv.launch()
    ctx.synchronize()
    for v, fb in zip(views, fbs):
        keys, img = single_view(ctx, nodes, v, prm)
        assert np.array_equal(fb.keys(), keys)
        assert np.array_equal(fb.rgba(), img)
    mv.close()


def test_autotune_picks_a_valid_config(ctx, city):
    prm = pcr.params(pcr.MODE_LAS2)
    views = views_for(SIZES, seed=3)
    fbs = [ctx.framebuffer(v.width, v.height) for v in views]
    best, ms = pcr.autotune(ctx, city, views, fbs, prm, iters=3)
    assert best.block_size in (64, 128, 256, 512) and best.points_per_thread in (40, 80, 160)
    assert best.tile_px in (0, 1024, 2048, 4096)
    assert ms > 0
    mv = pcr.MultiView(ctx, city, views, fbs, prm, tune=best)
    mv.launch()
    ctx.synchronize()
    keys, _ = single_view(ctx, city, views[2], prm)
    assert np.array_equal(fbs[2].keys(), keys)
    mv.close()


def test_server_same_result_alone_and_in_mix(ctx, city, nodes):
    big = 12 * 1920 * 1080
    srv = pcr.Server(ctx, budget_bytes=3 * big, max_batch=4, num_streams=3)
    c0, c1 = srv.add_cloud(city), srv.add_cloud(nodes)
    rng = np.random.default_rng(5)
    reqs = []
    for k, v in enumerate(views_for(SIZES[:4] * 4, seed=11)):
        cid = int(rng.integers(0, 2))
        mode = pcr.MODE_LAS2 if cid == c0 else pcr.MODE_NODES
        reqs.append((cid, mode, v))
    alone = []
    for cid, mode, v in reqs:
        t = srv.submit(cid, v, mode)
        img, lat = srv.wait(t, v.width * v.height)
        alone.append(img)
        assert lat > 0
    tickets = [srv.submit(cid, v, mode) for cid, mode, v in reqs]
    mixed = [srv.wait(t, v.width * v.height)[0] for t, (_, _, v) in zip(tickets, reqs)]
    for a, b in zip(alone, mixed):
        assert np.array_equal(a, b)
    # both equal the plain single-view render and the oracle
    cid, mode, v = reqs[0]
    _, img = single_view(ctx, city if cid == c0 else nodes, v, pcr.params(mode, lod_px=1.0))
    assert np.array_equal(alone[0], img)
    lays = {c0: city.download(), c1: nodes.download()}
    for (cid, mode, v), img in zip(reqs, mixed):
        L = lays[cid]
        if cid == c0:
            ref, _ = orc.render_batched(L["batches"], L["lo"], L["med"], L["hi"], v, 1)
        else:
            ref, _ = orc.render_nodes(L["batches"], L["lo"], L["med"], L["hi"], v, 1, pcr.LOD_SPACING, 1.0)
        assert np.array_equal(img, orc.resolve_index(ref, L["rgba"], pcr.BG_LAS))
    st = srv.stats()
    assert st.completed == 2 * len(reqs)
    assert st.fb_bytes_peak <= 3 * big and st.fb_bytes_in_use == 0
    srv.close()


def test_server_image_size_is_the_tickets(ctx, city):
    srv = pcr.Server(ctx, budget_bytes=12 * 640 * 480 * 2, max_batch=2, num_streams=1)
    cid = srv.add_cloud(city)
    v = views_for([(640, 480)], seed=3)[0]
    t = srv.submit(cid, v)
    with pytest.raises(ValueError):
        srv.wait(t, 100)                 # the image has 640 x 480 pixels: a shorter buffer is refused
    img, _ = srv.wait(t)
    assert img.shape == (640 * 480,)
    with pytest.raises(pcr.PcrError):
        srv.wait(t)                      # taken
    srv.close()


def test_server_refuses_and_cancels(ctx, city):
    srv = pcr.Server(ctx, budget_bytes=12 * 640 * 480 + 4096, max_batch=2, num_streams=1)
    cid = srv.add_cloud(city)
    with pytest.raises(pcr.PcrError) as e:
        srv.submit(cid, views_for([(1920, 1080)], seed=1)[0])
    assert e.value.code == capi.PCR_E_BUDGET
    vs = views_for([(640, 480)] * 6, seed=2)
    ts = [srv.submit(cid, v) for v in vs]
    srv.cancel(ts[-1])
    for t, v in zip(ts[:-1], vs[:-1]):
        img, _ = srv.wait(t, v.width * v.height)
        assert img.shape == (640 * 480,)
    st = srv.stats()
    assert st.refused == 1 and st.completed == 5 and st.fb_bytes_in_use == 0
    srv.close()


def test_server_pinned_results(ctx, city):
    """Zero-copy results: the pinned image equals the copied one, stays
    counted against the budget until released, and cannot be taken twice."""
    srv = pcr.Server(ctx, budget_bytes=12 * 1280 * 720 * 2, max_batch=2, num_streams=2)
    cid = srv.add_cloud(city)
    vs = views_for([(640, 480), (1280, 720)], seed=21)
    copied = []
    for v in vs:
        t = srv.submit(cid, v)
        copied.append(srv.wait(t, v.width * v.height)[0])
    ts = [srv.submit(cid, v) for v in vs]
    held = []
    for t, v in zip(ts, vs):
        img, lat = srv.wait_pinned(t, v.width * v.height)
        assert lat > 0
        held.append(img)
    assert srv.stats().fb_bytes_in_use > 0
    for img, ref in zip(held, copied):
        assert np.array_equal(img, ref)
    with pytest.raises(pcr.PcrError) as e:
        srv.wait(ts[0], vs[0].width * vs[0].height)
    assert "already taken" in str(e.value)
    srv.release(ts[0])
    srv.cancel(ts[1])   # cancelling a held result releases it too
    with pytest.raises(pcr.PcrError):
        srv.release(ts[0])
    assert srv.stats().fb_bytes_in_use == 0
    srv.close()


def test_server_concurrent_clients(ctx, city, nodes):
    """Four host threads submit and collect at once (the server is the one
    thread-safe object of the library); every image equals the one the same
    request gets alone, and the budget ends empty."""
    import threading
    srv = pcr.Server(ctx, budget_bytes=12 * 1280 * 720 * 6, max_batch=8, num_streams=3)
    c0, c1 = srv.add_cloud(city), srv.add_cloud(nodes)
    per_thread = []
    for t in range(4):
        vs = views_for(SIZES[:3] * 4, seed=40 + t)
        per_thread.append([((c0, pcr.MODE_LAS2) if k % 2 == 0 else (c1, pcr.MODE_NODES)) + (v,)
                           for k, v in enumerate(vs)])
    alone = {}
    for t, reqs in enumerate(per_thread):
        for k, (cid, mode, v) in enumerate(reqs):
            alone[(t, k)] = srv.wait(srv.submit(cid, v, mode), v.width * v.height)[0]
    got, errors = {}, []

    def client(t):
        try:
            reqs = per_thread[t]
            tickets = [srv.submit(cid, v, mode) for cid, mode, v in reqs]
            for k, (tk, (_, _, v)) in enumerate(zip(tickets, reqs)):
                if k % 3 == 2:   # collect some results zero-copy
                    img, _ = srv.wait_pinned(tk, v.width * v.height)
                    got[(t, k)] = img.copy()
                    srv.release(tk)
                else:
                    got[(t, k)] = srv.wait(tk, v.width * v.height)[0]
        except Exception as e:   # reported below, in the main thread
            errors.append(e)

    threads = [threading.Thread(target=client, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads)
    assert not errors, errors
    assert len(got) == len(alone)
    for key, img in alone.items():
        assert np.array_equal(got[key], img), key
    st = srv.stats()
    assert st.completed == 2 * len(alone) and st.fb_bytes_in_use == 0
    srv.close()
