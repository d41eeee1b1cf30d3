"""The device BVH builder (PLOC over Morton order, csrc/kernels/bvh_build.hip):
the exported tree is a valid BVH2 over every triangle once (leaf bounds
equal the triangles' bounds, every box contains its children), repeated
builds are identical, and traversal and rendering through it match the
oracle bit for bit (the oracle traverses the exported tree) and the
host-built scene's closest hits (topology cannot change the nearest hit).
The reference builds its BLASes on the GPU with NexusBVH
(src/scene/scene.cu:123-239), which is not available here."""
import ctypes as C

import numpy as np
import pytest

import wpt

pytestmark = pytest.mark.gpu


def with_builder(desc, builder, radius=0):
    d = wpt.SceneDesc.from_buffer_copy(desc)
    d.bvh_builder, d.bvh_search_radius = builder, radius
    return d


NODE = np.dtype([("bmin", "<f4", 3), ("bmax", "<f4", 3), ("left", "<u4"), ("right", "<u4")])


def check_tree(nodes, n_nodes, tris):
    # Reference node convention: a leaf has left == ~0 and right = its
    # triangle; the root is the last node.
    n = len(tris)
    assert n_nodes == 2 * n - 1
    arr = np.frombuffer(C.string_at(nodes, n_nodes * NODE.itemsize), NODE)
    bmin, bmax = arr["bmin"], arr["bmax"]
    left, right = arr["left"].astype(np.int64), arr["right"].astype(np.int64)
    leaf = left == 0xFFFFFFFF
    assert leaf.sum() == n
    prim = right[leaf]
    assert np.array_equal(np.sort(prim), np.arange(n))
    assert np.array_equal(bmin[leaf], tris[prim].min(axis=1)) and np.array_equal(bmax[leaf], tris[prim].max(axis=1))
    inner = np.nonzero(~leaf)[0]
    seen = np.zeros(n_nodes, np.int64)
    for c in (left[inner], right[inner]):
        assert np.all(bmin[inner] <= bmin[c]) and np.all(bmax[inner] >= bmax[c])
        np.add.at(seen, c, 1)
    root = n_nodes - 1
    assert seen[root] == 0 and np.all(np.delete(seen, root) == 1)  # a tree: one parent each


def mesh_triangles(desc, mesh=0):
    m = desc.meshes[mesh]
    size = C.sizeof(wpt.Triangle)
    base = C.cast(desc.mesh_triangles, C.c_void_p).value + m.first_triangle * size
    raw = np.frombuffer(C.string_at(base, m.triangle_count * size), np.uint8).reshape(m.triangle_count, size)
    return np.ascontiguousarray(raw[:, :36]).view(np.float32).reshape(m.triangle_count, 3, 3)


def random_mesh_scene(n, seed):
    rng = np.random.default_rng(seed)
    b = wpt.SceneBuilder()
    mat = b.material(wpt.make_material(wpt.DIFFUSE, (0.6, 0.6, 0.6)))
    centers = rng.uniform(-1, 1, (n, 3))
    tris = [wpt.make_triangle(c, c + rng.normal(0, 0.05, 3), c + rng.normal(0, 0.05, 3), mat) for c in centers]
    if n >= 4:  # a few exact duplicates: equal Morton codes and equal boxes
        tris[1] = tris[0]
        tris[3] = tris[2]
    mesh = b.mesh(tris)
    b.instance(mesh)
    b.instance(mesh, wpt.scaled_transform((0.5, 0.2, -0.3), 0.7))
    return b


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (3, 2), (17, 3), (1000, 4), (4099, 5)])
@pytest.mark.parametrize("radius", [1, 16, 64])
def test_device_tree_is_valid_and_matches_oracle(torch_cuda, oracle_mod, n, seed, radius):
    import torch
    b = random_mesh_scene(n, seed)
    desc = b.desc()
    scene = wpt.Scene(with_builder(desc, wpt.BVH_DEVICE_PLOC, radius), device=True, owner=b)
    info = scene.info()
    assert info.bvh_builder == wpt.BVH_DEVICE_PLOC and info.blas_count == 1
    nodes, count = scene.export_bvh2(2)
    check_tree(nodes, count, mesh_triangles(desc))
    rng = np.random.default_rng(seed)
    m = 2048
    o = rng.uniform(-1.5, 1.5, (m, 3))
    d = rng.normal(size=(m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.ascontiguousarray(np.concatenate([o, d], 1).T.astype(np.float32))
    got = wpt.hits_to_numpy(scene.trace_closest(torch.from_numpy(rays).cuda()))
    want = oracle_mod.Oracle(scene).trace(rays)
    for f in ("kind", "instance", "primitive", "t", "u", "v"):
        assert np.array_equal(got[f], want[f]), f
    # Exact ties (the duplicated triangles) resolve by a fixed order, so the
    # two trees give the same hits, identities included.
    host = wpt.Scene(with_builder(desc, wpt.BVH_HOST_SAH), device=True, owner=b)
    ref = wpt.hits_to_numpy(host.trace_closest(torch.from_numpy(rays).cuda()))
    for f in ("kind", "instance", "primitive", "t", "u", "v"):
        assert np.array_equal(got[f], ref[f]), f


def test_sphere_field_blas_on_device(torch_cuda, oracle_mod):
    host = wpt.HostScene("sphere-field")
    desc = host.desc()
    dev = wpt.Scene(with_builder(desc, wpt.BVH_DEVICE_PLOC), device=True, owner=host)
    again = wpt.Scene(with_builder(desc, wpt.BVH_DEVICE_PLOC), device=True, owner=host)
    sah = wpt.Scene(with_builder(desc, wpt.BVH_HOST_SAH), device=True, owner=host)
    di, si = dev.info(), sah.info()
    print(f"1,003,520-triangle BLAS: device build {di.blas_build_milliseconds:.2f} ms, depth {di.max_blas_depth}; "
          f"host SAH {si.blas_build_milliseconds:.1f} ms, depth {si.max_blas_depth}")
    nodes, count = dev.export_bvh2(2)
    nodes2, count2 = again.export_bvh2(2)
    assert count == count2 == 2 * 1_003_520 - 1
    assert C.string_at(nodes, count * C.sizeof(wpt.Bvh2Node)) == C.string_at(nodes2, count * C.sizeof(wpt.Bvh2Node))
    check_tree(nodes, count, mesh_triangles(desc))
    cam = host.camera(64, 36)
    cfg = wpt.render_config(64, 36, spp=2, max_depth=8, rr_start=3, mode=wpt.MODE_REGENERATE)
    img, _ = wpt.Renderer(dev, cam, cfg).render_frame()
    want, _ = oracle_mod.Oracle(dev).render(cam, cfg)
    assert np.array_equal(img, want)
    # Nearest-hit distances do not depend on the tree.
    import torch
    rng = np.random.default_rng(11)
    m = 8192
    o = np.tile(np.array([[13.0, 2.0, 3.0]]), (m, 1)) + rng.normal(0, 0.2, (m, 3))
    d = np.array([-13.0, -1.2, -3.0]) + rng.normal(0, 1.5, (m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = torch.from_numpy(np.ascontiguousarray(np.concatenate([o, d], 1).T.astype(np.float32))).cuda()
    a = wpt.hits_to_numpy(dev.trace_closest(rays))
    b = wpt.hits_to_numpy(sah.trace_closest(rays))
    assert (a["kind"] != wpt.HIT_NONE).mean() > 0.5
    for f in ("kind", "instance", "primitive", "t", "u", "v"):
        assert np.array_equal(a[f], b[f]), f
    ref, _ = wpt.Renderer(sah, cam, cfg).render_frame()
    assert np.array_equal(img, ref)


def test_exact_ties_are_tree_independent(torch_cuda, oracle_mod):
    # Coincident triangles in two coincident instances: both trees and the
    # oracle pick instance 0 and the first copy.
    import torch
    import test_cpu_oracle as cpu
    b = cpu.tie_scene()
    rays = torch.from_numpy(cpu.tie_rays()).cuda()
    want = oracle_mod.Oracle(wpt.Scene(b.desc(), device=False, owner=b)).trace(cpu.tie_rays())
    for builder in (wpt.BVH_HOST_SAH, wpt.BVH_DEVICE_PLOC):
        scene = wpt.Scene(with_builder(b.desc(), builder), device=True, owner=b)
        got = wpt.hits_to_numpy(scene.trace_closest(rays))
        for f in ("kind", "instance", "primitive", "t", "u", "v"):
            assert np.array_equal(got[f], want[f]), (builder, f)
        mesh = got["kind"] == wpt.HIT_INSTANCE
        assert mesh.sum() > 100 and np.all(got["instance"][mesh] == 0) and np.all(got["primitive"][mesh] < 40)


def test_device_builder_needs_a_device_scene():
    b = random_mesh_scene(8, 0)
    with pytest.raises(wpt.WptError):
        wpt.Scene(with_builder(b.desc(), wpt.BVH_DEVICE_PLOC), device=False, owner=b)
    with pytest.raises(wpt.WptError):
        wpt.Scene(with_builder(b.desc(), wpt.BVH_DEVICE_PLOC, 65), device=True, owner=b)
