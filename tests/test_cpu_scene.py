"""CPU tests of the scene build behind wpt_scene_create: the AccelerationMode
partition invariants (tests/scene_acceleration_tests.cu:137-179), the light
table and Walker alias (src/scene/scene.cu:82-121, scene_host.cpp:60-106)
against the oracle's independent construction, BVH structure invariants and
description validation."""
import ctypes as C

import numpy as np
import pytest

import scenes
import wpt


def host_scene(builder):
    return wpt.Scene(builder.desc(), device=False, owner=builder)


@pytest.mark.parametrize("case", [c[0] for c in scenes.reference_cases()])
def test_partition_invariants(oracle_mod, case):
    builder = dict(scenes.reference_cases())[case]
    scene = host_scene(builder)
    spheres, instances = builder.spheres, builder.instances
    sphere_uses = np.zeros(len(spheres), int)
    instance_uses = np.zeros(len(instances), int)
    for i in scene.list(0):
        sphere_uses[i] += 1
        assert spheres[i].acceleration == wpt.ACCEL_DIRECT
    for i in scene.list(1):
        instance_uses[i] += 1
        assert instances[i].acceleration == wpt.ACCEL_DIRECT
    for leaf in scene.list(2):
        if leaf & 0x80000000:
            idx = leaf & 0x7FFFFFFF
            sphere_uses[idx] += 1
            assert spheres[idx].acceleration == wpt.ACCEL_BVH
        else:
            instance_uses[leaf] += 1
            assert instances[leaf].acceleration == wpt.ACCEL_BVH
    assert np.all(sphere_uses == 1)
    for i, inst in enumerate(instances):
        empty = builder.meshes[inst.mesh].triangle_count == 0
        assert instance_uses[i] == (0 if empty and inst.acceleration == wpt.ACCEL_BVH else 1)
    for m in range(len(builder.meshes)):
        needs = any(inst.mesh == m and inst.acceleration == wpt.ACCEL_BVH for inst in instances) and \
            builder.meshes[m].triangle_count > 0
        _, n = scene.export_bvh2(2 + m)
        assert (n > 0) == needs
    # The oracle derives the same lists on its own.
    oracle = oracle_mod.Oracle(scene)
    for which in (0, 1, 2):
        assert np.array_equal(scene.list(which), oracle.list(which))


def check_bvh(nodes, n, prim_bounds):
    """Reference node convention: root = n-1, leaf iff left == ~0, one
    primitive per leaf, each primitive exactly once, parents contain children."""
    if n == 0:
        return
    seen = np.zeros(len(prim_bounds), int)
    stack = [n - 1]
    visited = 0
    while stack:
        i = stack.pop()
        visited += 1
        node = nodes[i]
        lo, hi = np.array(node.bmin), np.array(node.bmax)
        if node.left == wpt.INVALID:
            p = node.right
            seen[p] += 1
            assert np.array_equal(lo, prim_bounds[p][0]) and np.array_equal(hi, prim_bounds[p][1])
            continue
        for c in (node.left, node.right):
            assert np.all(np.array(nodes[c].bmin) >= lo) and np.all(np.array(nodes[c].bmax) <= hi)
            stack.append(c)
    assert visited == n
    assert np.all(seen == 1)


def triangle_bounds(tris):
    out = []
    for t in tris:
        v = np.array([t.v0[:], t.v1[:], t.v2[:]], np.float32)
        out.append((v.min(axis=0), v.max(axis=0)))
    return out


@pytest.mark.parametrize("preset,freq", [("demo", 0), ("prism", 0), ("sphere-field", 20)])
def test_bvh_invariants(preset, freq):
    host = wpt.HostScene(preset, icosphere_frequency=freq)
    scene = wpt.Scene(host.desc(), device=False, owner=host)
    d = host.desc()
    for m in range(d.mesh_count):
        mesh = d.meshes[m]
        tris = [d.mesh_triangles[mesh.first_triangle + i] for i in range(mesh.triangle_count)]
        nodes, n = scene.export_bvh2(2 + m)
        if n:
            assert n == 2 * len(tris) - 1
            check_bvh(nodes, n, triangle_bounds(tris))
    nodes, n = scene.export_bvh2(1)
    if n:
        check_bvh(nodes, n, triangle_bounds([d.flat_triangles[i] for i in range(d.flat_triangle_count)]))
    nodes, n = scene.export_bvh2(0)
    assert n == 2 * scene.info().tlas_leaf_count - 1 or n == 0


def test_bvh_deterministic_across_builds():
    a = wpt.HostScene("sphere-field", icosphere_frequency=40)
    b = wpt.HostScene("sphere-field", icosphere_frequency=40)
    na, ca = wpt.Scene(a.desc(), device=False, owner=a).export_bvh2(2)
    nb, cb = wpt.Scene(b.desc(), device=False, owner=b).export_bvh2(2)
    assert ca == cb and C.string_at(na, ca * 32) == C.string_at(nb, cb * 32)


@pytest.mark.parametrize("preset", ["cornell-box", "prism"])
def test_light_table_matches_oracle(oracle_mod, preset):
    host = wpt.HostScene(preset)
    scene = wpt.Scene(host.desc(), device=False, owner=host)
    lights, alias = scene.lights()
    olights, oalias = oracle_mod.Oracle(scene).lights()
    assert lights.shape[0] == 2
    assert lights.tobytes() == olights.tobytes()
    assert alias.tobytes() == oalias.tobytes()
    pdf = lights[:, 14].astype(np.float64)
    assert abs(pdf.sum() - 1.0) < 1e-6
    # Walker table reproduces the selection pdf: P(i) = (p_i + sum over j
    # aliasing to i of (1 - p_j)) / n.
    n = len(pdf)
    prob = alias[:, 0].astype(np.float64)
    target = alias[:, 1].view(np.uint32)
    selected = prob.copy()
    for j in range(n):
        selected[target[j]] += 1.0 - prob[j]
    assert np.allclose(selected / n, pdf, atol=1e-6)


def test_alias_table_many_lights(oracle_mod):
    b = wpt.SceneBuilder()
    rng = np.random.default_rng(3)
    for i in range(64):
        e = float(np.exp(rng.normal()))
        b.material(wpt.make_material(wpt.EMISSIVE, (1, 1, 1), emission=(e, e * 0.5, e * 0.25)))
        x = float(i)
        b.static_triangles.append(wpt.make_triangle((x, 0, 0), (x + 1, 0, 0), (x, float(rng.uniform(0.1, 2)), 0), i))
    scene = host_scene(b)
    lights, alias = scene.lights()
    olights, oalias = oracle_mod.Oracle(scene).lights()
    assert lights.tobytes() == olights.tobytes() and alias.tobytes() == oalias.tobytes()
    n = len(lights)
    prob = alias[:, 0].astype(np.float64)
    target = alias[:, 1].view(np.uint32)
    selected = prob.copy()
    for j in range(n):
        selected[target[j]] += 1.0 - prob[j]
    assert np.allclose(selected / n, lights[:, 14], atol=1e-5)


def test_instance_light_offsets():
    b = wpt.SceneBuilder()
    b.material(wpt.make_material(wpt.DIFFUSE))
    b.material(wpt.make_material(wpt.EMISSIVE, emission=(4, 4, 4)))
    m = b.mesh([wpt.make_triangle((0, 0, 0), (1, 0, 0), (0, 1, 0), 1), wpt.make_triangle((0, 0, 1), (1, 0, 1), (0, 1, 1), 0),
                wpt.make_triangle((0, 0, 2), (1, 0, 2), (0, 1, 2), 1)])
    b.instance(m, wpt.scaled_transform((0, 0, 0), 1))
    b.instance(m, wpt.scaled_transform((5, 0, 0), 2), material_override=0)
    b.static_triangles.append(wpt.make_triangle((0, 5, 0), (1, 5, 0), (0, 6, 0), 1))
    scene = host_scene(b)
    offsets, static_idx, mesh_idx = scene.light_indices()
    assert offsets.tolist() == [0, 2]
    assert mesh_idx.tolist() == [0, wpt.INVALID, 1]
    assert static_idx.tolist() == [4]
    lights, _ = scene.lights()
    assert len(lights) == 5


@pytest.mark.parametrize("mutate,message", [
    (lambda b: b.spheres.append(wpt.Sphere(wpt.vec3((0, 0, 0)), 1, 99, 0)), "sphere material"),
    (lambda b: b.instances.append(wpt.Instance(7, wpt.INVALID, 0, 0, wpt.scaled_transform())), "instance mesh"),
    (lambda b: b.static_triangles.append(wpt.make_triangle((0, 0, 0), (1, 0, 0), (0, 1, 0), 50)), "static triangle"),
    (lambda b: b.spheres.append(wpt.Sphere(wpt.vec3((0, 0, 0)), 1, 0, 5)), "acceleration mode"),
])
def test_invalid_descriptions_rejected(mutate, message):
    b = scenes.base_scene()
    mutate(b)
    with pytest.raises(wpt.WptError) as err:
        wpt.Scene(b.desc(), device=False, owner=b)
    assert message in str(err.value)
