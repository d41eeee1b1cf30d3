"""Flat-triangle accelerators (reference uploadDirectAcceleration,
src/scene/scene.cu:241-294, with the policies nexus_bvh2, nexus_bvh8,
median_bvh and spatial_split_bvh; traversal device_traversal.cuh:880-1033).

Every accelerator must give the nearest hit of the triangles, so the oracle
walking each exported tree is checked against an independent FP64 brute
force, and the trees are checked structurally: the binary trees hold every
triangle in one leaf (spatial splits: in one or more leaves whose bounds lie
inside the triangle's), the median tree has depth ceil(log2 n), and the
scene build reports the wide tree's stack need."""
import math

import numpy as np
import pytest

import test_cpu_loaders as loaders
import wpt

ACCELS = [wpt.FLAT_SAH_BVH2, wpt.FLAT_MEDIAN_BVH2, wpt.FLAT_SPATIAL_SPLIT_BVH2, wpt.FLAT_WIDE8]


def flat_scene(accel, n=3000, seed=4, slivers=200):
    """Random small triangles in a 6^3 box plus long thin slivers that cross
    much of it (they give spatial splits something to cut)."""
    rng = np.random.default_rng(seed)
    b = wpt.SceneBuilder()
    for k in range(3):
        b.material(wpt.make_material(wpt.DIFFUSE, (0.3 + 0.2 * k, 0.5, 0.5)))
    for i in range(n):
        c = rng.uniform(-3, 3, 3)
        b.flat_triangles.append(wpt.make_triangle(c + rng.normal(0, 0.15, 3), c + rng.normal(0, 0.15, 3),
                                                  c + rng.normal(0, 0.15, 3), i % 3))
    for i in range(slivers):
        a = rng.uniform(-3, 3, 3)
        d = rng.normal(size=3)
        d *= 5.0 / np.linalg.norm(d)
        w = rng.normal(size=3) * 0.05
        b.flat_triangles.append(wpt.make_triangle(a, a + d, a + d + w, i % 3))
    b.flat_accelerator = accel
    return b


def random_rays(n, seed):
    rng = np.random.default_rng(seed)
    rays = np.zeros((6, n), np.float32)
    rays[0:3] = rng.uniform(-4.5, 4.5, (3, n))
    d = rng.normal(size=(3, n))
    rays[3:6] = d / np.linalg.norm(d, axis=0)
    return rays


def leaves_of(exported):
    """Exported tree (reference convention): (primitive, bounds) per leaf."""
    nodes, count = exported
    return [(int(nodes[i].right), list(nodes[i].bmin), list(nodes[i].bmax)) for i in range(count)
            if nodes[i].left == 0xFFFFFFFF]


@pytest.mark.parametrize("accel", ACCELS)
def test_tree_structure(accel):
    b = flat_scene(accel)
    scene = wpt.Scene(b.desc(), device=False, owner=b)
    info = scene.info()
    n = len(b.flat_triangles)
    assert info.flat_triangle_count == n and info.flat_accelerator == accel
    nodes = scene.export_bvh2(1)
    leaves = leaves_of(nodes)
    prims = [p for p, _, _ in leaves]
    tri = np.array([[list(t.v0), list(t.v1), list(t.v2)] for t in b.flat_triangles], np.float64)
    if accel == wpt.FLAT_SPATIAL_SPLIT_BVH2:
        # Every triangle is referenced; some more than once; each reference's
        # box lies within the triangle's bounds (up to the widening).
        assert set(prims) == set(range(n))
        assert n < len(prims) <= 2 * n and info.flat_leaf_slots == len(prims)
        lo, hi = tri.min(axis=1), tri.max(axis=1)
        for p, bmin, bmax in leaves:
            tol = 1e-6 * np.maximum(1.0, np.abs(lo[p]))
            assert np.all(np.asarray(bmin) >= lo[p] - tol) and np.all(np.asarray(bmax) <= hi[p] + 1e-6 * np.maximum(1.0, np.abs(hi[p])))
    else:
        assert sorted(prims) == list(range(n)) and info.flat_leaf_slots == n
    if accel == wpt.FLAT_MEDIAN_BVH2:
        # Halving at every level: the deepest leaf is ceil(log2 n) levels down.
        assert info.max_stack_depth == math.ceil(math.log2(n)) + 1
    if accel == wpt.FLAT_WIDE8:
        assert info.flat_wide_node_count > 0
        # Every wide node holds 2 to 8 children: between (n-1)/7 and n-1 nodes.
        assert (n - 1) / 7 <= info.flat_wide_node_count <= n - 1
    else:
        assert info.flat_wide_node_count == 0
    assert info.max_stack_depth <= 64


@pytest.mark.parametrize("accel", ACCELS)
def test_oracle_on_each_tree_matches_fp64_brute_force(oracle_mod, accel):
    b = flat_scene(accel, n=1500, slivers=100)
    scene = wpt.Scene(b.desc(), device=False, owner=b)
    oracle = oracle_mod.Oracle(scene)
    rays = random_rays(3000, 7)
    hits = oracle.trace(rays)
    t, mat, _ = oracle.brute_fp64(rays)
    hit = hits["kind"] != wpt.HIT_NONE
    assert np.array_equal(hit, mat != wpt.INVALID)
    assert hit.sum() > 500
    assert np.all(np.abs(hits["t"][hit] - t[hit]) < 3e-4 * np.maximum(1.0, t[hit]))


def test_every_tree_gives_the_same_hits(oracle_mod):
    rays = random_rays(4000, 9)
    results = []
    for accel in ACCELS:
        b = flat_scene(accel, n=2000, slivers=150)
        scene = wpt.Scene(b.desc(), device=False, owner=b)
        results.append(oracle_mod.Oracle(scene).trace(rays).tobytes())
    assert all(r == results[0] for r in results[1:])


@pytest.mark.parametrize("policy,accel", [("nexus_bvh2", wpt.FLAT_SAH_BVH2), ("nexus_bvh8", wpt.FLAT_WIDE8),
                                          ("median_bvh", wpt.FLAT_MEDIAN_BVH2),
                                          ("spatial_split_bvh", wpt.FLAT_SPATIAL_SPLIT_BVH2),
                                          ("tlas_blas_bvh2", wpt.FLAT_SAH_BVH2)])
def test_manifest_policy_selects_the_accelerator(tmp_path, policy, accel):
    host = wpt.HostScene(loaders.deer_manifest(tmp_path, policy))
    d = host.desc()
    assert d.flat_accelerator == accel
    assert (d.flat_triangle_count == 0) == (policy == "tlas_blas_bvh2")
    scene = wpt.Scene(d, device=False, owner=host)
    assert scene.info().flat_accelerator == (accel if d.flat_triangle_count else wpt.FLAT_SAH_BVH2)


def test_unknown_accelerator_is_rejected():
    b = flat_scene(wpt.FLAT_SAH_BVH2, n=10, slivers=0)
    d = b.desc()
    d.flat_accelerator = 4
    with pytest.raises(wpt.WptError):
        wpt.Scene(d, device=False, owner=b)
