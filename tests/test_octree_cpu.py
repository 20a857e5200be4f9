"""LOD octree layout (pcr_octree_layout_host) checked on the CPU against the
invariants of a Potree-style subsampled octree, and the oracle's LOD cut on
it."""
import numpy as np
import pytest

import compute_rasterizer_amd as pcr
from oracle import orc


def cloud(n=60_000, seed=3):
    cfg = pcr.gen_cfg(pcr.SCENE_TERRAIN, n, seed=seed, extent=400.0, amplitude=40.0)
    xyz, _ = orc.generate(cfg)
    ids = np.arange(n, dtype=np.uint32)          # unique "colours" identify every point
    return xyz, ids


@pytest.fixture(scope="module")
def tree():
    xyz, ids = cloud()
    return xyz, ids, pcr.octree_layout_host(xyz, ids, seed=9, max_node_points=2000)


def test_every_point_stored_once_and_decodes_close(tree):
    xyz, ids, (nodes, lo, med, hi, col) = tree
    seen = []
    for k, nd in enumerate(nodes):
        f, n = int(nd["first_point"]), int(nd["num_points"])
        assert f % 4 == 0 and 0 < n <= 2000
        seen.append(col[f:f + n])
        size = (nd["max"] - nd["min"]).astype(np.float64)
        for i in range(f, f + n, 97):
            p = orc.decode(0, int(lo[i]), int(med[i]), int(hi[i]), nd["min"], nd["max"])
            q = xyz[col[i]]
            assert np.all(np.abs(p - q) <= size * 2.0 ** -22 + np.abs(q) * 2.0 ** -22 + 1e-5)
    seen = np.concatenate(seen)
    assert len(seen) == len(ids) and np.array_equal(np.sort(seen), ids)


def test_hierarchy_parents_levels_spacing(tree):
    xyz, ids, (nodes, *_rest) = tree
    assert nodes[0]["parent"] == -1 and nodes[0]["level"] == 0
    s0 = float(nodes[0]["spacing"])
    diag = float(np.linalg.norm(xyz.max(0).astype(np.float64) - xyz.min(0)))
    assert abs(s0 - diag / 128) / s0 < 1e-6
    for k, nd in enumerate(nodes):
        lv = int(nd["level"])
        assert nd["spacing"] == np.float32(diag / 128 / 2 ** lv)
        if k and nd["parent"] == -1:
            assert lv == 0                     # extra chunks of the root
        if nd["parent"] >= 0:
            assert nd["parent"] < k
            assert nodes[nd["parent"]]["level"] == lv - 1
    assert nodes["level"].max() >= 2


def _nodes_of(nodes):
    """Header indices grouped into nodes (a node's chunks are consecutive and
    share level, parent and box); returns (groups, index of each header's group)."""
    groups, of = [], np.zeros(len(nodes), np.int64)
    for k, nd in enumerate(nodes):
        if groups:
            h = nodes[groups[-1][0]]
            if (h["level"] == nd["level"] and h["parent"] == nd["parent"] and np.array_equal(h["min"], nd["min"])
                    and np.array_equal(h["max"], nd["max"])):
                groups[-1].append(k)
                of[k] = len(groups) - 1
                continue
        groups.append([k])
        of[k] = len(groups) - 1
    return groups, of


def test_inner_nodes_hold_maximal_poisson_disk_samples(tree):
    """A node with children holds a seeded Poisson-disk sample of the points of
    its region: no two of its points closer than r = cube / (128 * 2^level)
    (exactly, in the 21-bit Morton coordinates), and every point stored deeper
    in its subtree within r of one of them (maximal)."""
    from scipy.spatial import cKDTree
    xyz, ids, (nodes, lo, med, hi, col) = tree
    mn = xyz.min(0)
    side = float(max(np.float32(xyz.max(0)[a] - mn[a]) for a in range(3)))
    scale = 2097152.0 / side
    q = np.minimum(((xyz.astype(np.float64) - mn.astype(np.float64)) * scale).astype(np.int64), 2097151)
    groups, of = _nodes_of(nodes)
    children = [[] for _ in groups]
    for k, nd in enumerate(nodes):
        if nd["parent"] >= 0 and k == groups[of[k]][0]:
            children[of[nd["parent"]]].append(of[k])

    def points(g):
        return np.concatenate([col[int(nodes[k]["first_point"]):int(nodes[k]["first_point"]) +
                                   int(nodes[k]["num_points"])] for k in groups[g]])

    def subtree(g):
        out, stack = [], list(children[g])
        while stack:
            c = stack.pop()
            out.append(points(c))
            stack += children[c]
        return np.concatenate(out)

    checked = 0
    for g in range(len(groups)):
        if not children[g]:
            continue
        lv = int(nodes[groups[g][0]]["level"])
        r = 1 << (21 - 7 - lv)
        s = q[points(g)]
        t = cKDTree(s)
        for i, j in t.query_pairs(r + 1.0):
            d2 = int(((s[i] - s[j]) ** 2).sum())
            assert d2 >= r * r, f"node {g}: two samples {d2 ** 0.5:.1f} apart, spacing {r}"
        deeper = q[subtree(g)]
        dist, nearest = t.query(deeper)
        d2 = ((deeper - s[nearest]) ** 2).sum(1)
        assert np.all(d2 < r * r), f"node {g}: a point of its subtree is {np.sqrt(d2.max()):.1f} from every sample"
        checked += 1
    assert checked >= 3


def test_layout_is_deterministic_and_seed_dependent():
    xyz, ids = cloud(20_000, seed=5)
    a = pcr.octree_layout_host(xyz, ids, seed=1, max_node_points=1000)
    b = pcr.octree_layout_host(xyz, ids, seed=1, max_node_points=1000)
    c = pcr.octree_layout_host(xyz, ids, seed=2, max_node_points=1000)
    assert all(np.array_equal(x, y) for x, y in zip(a[1:], b[1:]))
    assert np.array_equal(a[0].view(np.uint8), b[0].view(np.uint8))
    assert not np.array_equal(a[4], c[4])


def test_oracle_lod_cut_on_octree(tree):
    xyz, ids, (nodes, lo, med, hi, col) = tree
    far = orc.look_at((200, 200, 20000), (200, 200, 0), (0, 1, 0), near=1, far=1e6, width=640, height=480)
    sel = orc.lod_select(nodes, far, 1, pcr.LOD_SPACING, 1.0)
    assert sel[nodes["level"] == 0].all() and sel[nodes["level"] > 0].sum() == 0
    near = orc.look_at((200, -50, 120), (200, 200, 0), (0, 0, 1), near=1, far=1e5, width=1920, height=1080)
    sel = orc.lod_select(nodes, near, 1, pcr.LOD_SPACING, 1.0)
    assert (sel & (nodes["level"] >= 2)).any()
    for k, nd in enumerate(nodes):
        if sel[k] and nd["parent"] >= 0:
            assert sel[nd["parent"]]
    fb, flags = orc.render_nodes(nodes, lo, med, hi, near, 1, pcr.LOD_SPACING, 1.0)
    assert np.array_equal(flags, sel) and (fb != pcr.FB_CLEAR).sum() > 1000


def test_rejects_non_power_of_two_spacing_div():
    xyz, ids = cloud(1000)
    with pytest.raises(pcr.PcrError):
        pcr.octree_layout_host(xyz, ids, root_spacing_div=100.0)


def test_nodes_mt_equals_single_thread():
    """The oracle's multithreaded LOD renderer (CAS fetch-min over the drawn
    nodes) gives the single-thread frame for both LOD rules."""
    xyz, rgba = orc.generate(pcr.gen_cfg(pcr.SCENE_TERRAIN, 200_000, seed=5))
    nodes, lo, med, hi, rgb = pcr.octree_layout_host(xyz, rgba, seed=3, max_node_points=4000)
    v = pcr.look_at((1000, -300, 400), (1000, 1000, 30), up=(0, 0, 1), near=1, far=5000, width=320, height=200)
    for rule in (pcr.LOD_SPACING, pcr.LOD_SIZE80):
        ref, _ = orc.render_nodes(nodes, lo, med, hi, v, 1, rule, 1.0)
        assert (ref != pcr.FB_CLEAR).sum() > 1000
        for threads in (1, 4, 7):
            assert np.array_equal(orc.render_nodes_mt(nodes, lo, med, hi, v, 1, rule, 1.0, threads), ref)
