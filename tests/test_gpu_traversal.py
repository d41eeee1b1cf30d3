"""GPU parity of closest-hit and shadow traversal (C ABI wpt_trace_closest /
wpt_trace_occluded) against the oracle, on the reference acceleration-test
scenes and ray set (tests/scene_acceleration_tests.cu:181-354) and the
scaled-instance cases (tests/material_transport_tests.cu:13-52).

Bar: hit records (t, u, v, kind, instance, primitive, material, light index)
and occlusion bits are bit-identical to the oracle's FP32 restatement."""
import numpy as np
import pytest

import scenes
import wpt

pytestmark = pytest.mark.gpu


def gpu_trace(torch, scene, rays):
    d_rays = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    hits = wpt.hits_to_numpy(scene.trace_closest(d_rays))
    return hits


def gpu_occluded(torch, scene, rays, tmax):
    d_rays = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    d_t = torch.from_numpy(np.ascontiguousarray(tmax, dtype=np.float32)).cuda()
    return scene.trace_occluded(d_rays, d_t).cpu().numpy()


def assert_hits_equal(name, got, want):
    g = got.view(np.uint32).reshape(len(got), -1)
    w = want.view(np.uint32).reshape(len(want), -1)
    bad = np.where((g != w).any(axis=1))[0]
    assert bad.size == 0, f"{name}: {bad.size} hit records differ, first ray {bad[0]}: gpu {got[bad[0]]} oracle {want[bad[0]]}"


@pytest.mark.parametrize("case", [c[0] for c in scenes.reference_cases()])
def test_reference_cases_bit_exact(torch_cuda, oracle_mod, case):
    builder = dict(scenes.reference_cases())[case]
    desc = builder.desc()
    scene = wpt.Scene(desc, device=True, owner=builder)
    oracle = oracle_mod.Oracle(scene)
    rays = oracle_mod.reference_rays(12345, 2048)
    want = oracle.trace(rays)
    got = gpu_trace(torch_cuda, scene, rays)
    assert_hits_equal(case, got, want)
    hit = want["kind"] != wpt.HIT_NONE
    limits = np.where(hit, want["t"], 1000.0).astype(np.float32)
    before = gpu_occluded(torch_cuda, scene, rays, limits - 0.02)
    after = gpu_occluded(torch_cuda, scene, rays, limits + 0.02)
    assert np.array_equal(before, oracle.occluded(rays, limits - 0.02))
    assert np.array_equal(after, oracle.occluded(rays, limits + 0.02))
    # The reference test's own semantics: nothing blocks before the closest
    # hit, and the closest hit blocks just after it.
    assert not before.any()
    assert np.array_equal(after.astype(bool), hit)


@pytest.mark.parametrize("scale", [0.01, 1.0, 100.0, 10000.0])
def test_scaled_instances(torch_cuda, oracle_mod, scale):
    builder = scenes.scaled_instance_scene(scale)
    scene = wpt.Scene(builder.desc(), device=True, owner=builder)
    oracle = oracle_mod.Oracle(scene)
    rays = np.array([[0, 0], [0, 0], [2, -2], [0, 0], [0, 0], [-1, 1]], dtype=np.float32)
    got = gpu_trace(torch_cuda, scene, rays)
    assert_hits_equal(f"scale {scale}", got, oracle.trace(rays))
    assert np.all(got["kind"] == wpt.HIT_INSTANCE)
    assert np.all(np.abs(got["t"] - 2.0) <= 1e-5)
    assert gpu_occluded(torch_cuda, scene, rays, np.full(2, 2.5, np.float32)).all()
    assert not gpu_occluded(torch_cuda, scene, rays, np.full(2, 1.5, np.float32)).any()


@pytest.mark.parametrize("preset,freq", [("cornell-box", 0), ("demo", 0), ("prism", 0), ("sphere-field", 32)])
def test_preset_scenes_random_rays(torch_cuda, oracle_mod, preset, freq):
    host = wpt.HostScene(preset, icosphere_frequency=freq)
    scene = wpt.Scene(host.desc(), device=True, owner=host)
    oracle = oracle_mod.Oracle(scene)
    rng = np.random.default_rng(7)
    n = 1 << 15
    origins = rng.uniform(-3, 3, size=(3, n)).astype(np.float32)
    if preset == "sphere-field":
        origins *= 3.0
        origins[1] = np.abs(origins[1]) + 0.05
    dirs = rng.normal(size=(3, n)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=0, keepdims=True).astype(np.float32)
    rays = np.concatenate([origins, dirs.astype(np.float32)]).astype(np.float32)
    want = oracle.trace(rays)
    got = gpu_trace(torch_cuda, scene, rays)
    assert_hits_equal(preset, got, want)
    tmax = np.where(want["kind"] != 0, want["t"] * 0.75, 50.0).astype(np.float32)
    assert np.array_equal(gpu_occluded(torch_cuda, scene, rays, tmax), oracle.occluded(rays, tmax))


def test_empty_ray_batch(torch_cuda):
    builder = scenes.base_scene()
    scene = wpt.Scene(builder.desc(), device=True, owner=builder)
    rays = np.zeros((6, 0), dtype=np.float32)
    d = torch_cuda.zeros((6, 0), dtype=torch_cuda.float32, device="cuda")
    assert scene.trace_closest(d).shape[0] == 0
