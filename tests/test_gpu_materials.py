"""GPU parity of textured and glTF materials: a room covering every texture
slot (tiled and transformed albedo, alpha masks tested in traversal, normal
maps on both texture-coordinate sets with vertex tangents, emissive,
metallic-roughness, specular-glossiness and thickness textures, glTF
specular and clearcoat lobes with delta reflections, a textured instanced
mesh) rendered by all three schedulers, the debug views and the batch
consumer, against the oracle's FP32 restatement bit for bit; and the
reference's glTF smooth-coat energy check on the device."""
import numpy as np
import pytest

import scenes
import wpt

pytestmark = pytest.mark.gpu

MODES = [wpt.MODE_GRAPH, wpt.MODE_PERSISTENT, wpt.MODE_REGENERATE]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("lit", [False, True])
def test_textured_room_matches_oracle(torch_cuda, oracle_mod, mode, lit):
    b = scenes.textured_room(with_env=lit, with_lights=lit)
    scene = wpt.Scene(b.desc(), device=True, owner=b)
    cam = scenes.room_camera()
    cfg = wpt.render_config(80, 80, spp=3, max_depth=8, rr_start=3, mode=mode, clamp=10.0)
    r = wpt.Renderer(scene, cam, cfg)
    img, _ = r.render_frame()
    want, ost = oracle_mod.Oracle(scene).render(cam, cfg)
    assert np.array_equal(img, want), f"max |diff| {np.abs(img - want).max()}"
    st = r.stats()
    assert st.closest_queries == ost.closest_queries and st.shadow_queries == ost.shadow_queries


@pytest.mark.parametrize("debug", ["intersection_debug", "shading_normal_debug"])
def test_textured_room_debug_views(torch_cuda, oracle_mod, debug):
    b = scenes.textured_room()
    scene = wpt.Scene(b.desc(), device=True, owner=b)
    cam = scenes.room_camera()
    cfg = wpt.render_config(64, 64, spp=1, max_depth=8, rr_start=3)
    setattr(cfg, debug, 1)
    img, _ = wpt.Renderer(scene, cam, cfg).render_frame()
    want, _ = oracle_mod.Oracle(scene).render(cam, cfg)
    assert np.array_equal(img, want)


@pytest.mark.parametrize("scheduler", [wpt.BATCH_STAGED, wpt.BATCH_REGENERATE], ids=["staged", "regenerate"])
def test_textured_batch_equals_render_task(torch_cuda, scheduler):
    rooms = [scenes.textured_room(), scenes.textured_room(with_env=False, with_lights=False)]
    dev = [wpt.Scene(b.desc(), device=True, owner=b) for b in rooms]
    w, h = 96, 64
    cam = scenes.room_camera(w / h)
    tasks = [wpt.batch_task(i, cam, w, h, x, y, 32, 32, 1, 2, max_depth=8, rr_start=3, clamp=10.0)
             for i in range(2) for (x, y) in [(0, 0), (32, 32), (64, 0)]]
    outs = wpt.Batch(dev, scheduler=scheduler).render(tasks)
    for t, out in zip(tasks, outs):
        cfg = wpt.render_config(w, h, spp=3, max_depth=8, rr_start=3, tile=(32, 32), clamp=10.0)
        want = wpt.Renderer(dev[t.scene], cam, cfg).render_task(
            wpt.Task(0, 0, t.sample_start, t.sample_count, t.x, t.y, t.tile_width, t.tile_height))
        assert np.array_equal(out, want)


def test_trace_queries_see_texture_alpha(torch_cuda, oracle_mod):
    # The C-ABI traversal queries apply the textured alpha mask like the
    # oracle: random rays through the room, hit identities bit-exact.
    import torch
    b = scenes.textured_room()
    scene = wpt.Scene(b.desc(), device=True, owner=b)
    rng = np.random.default_rng(7)
    n = 4096
    o = rng.uniform((-2.5, 0.1, -2.5), (2.5, 2.8, 2.5), (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.ascontiguousarray(np.concatenate([o, d], axis=1).T.astype(np.float32))
    got = wpt.hits_to_numpy(scene.trace_closest(torch.from_numpy(rays).cuda()))
    want = oracle_mod.Oracle(scene).trace(rays)
    for f in ("kind", "instance", "primitive", "material", "t", "u", "v"):
        assert np.array_equal(got[f], want[f]), f


def test_gltf_smooth_coat_energy_on_device(torch_cuda):
    albedo = np.array([0.2, 0.4, 0.6])
    b = scenes.env_floor(wpt.make_material(wpt.DIFFUSE, albedo, ior=1.5), scenes.constant_map(4.0), intensity=2.5)
    b.material_exts[0] = wpt.make_material_ext(gltf_pbr_workflow=True, specular_weight=0.0, clearcoat=1.0,
                                               clearcoat_roughness=0.0)
    scene = wpt.Scene(b.desc(), device=True, owner=b)
    cfg = wpt.render_config(128, 128, spp=16, max_depth=12, rr_start=64, mode=wpt.MODE_REGENERATE, clamp=0.0)
    img, n = wpt.Renderer(scene, scenes.look_down(fov=0.02), cfg).render_frame()
    mean = img.reshape(-1, 3).astype(np.float64).mean(axis=0) / n
    assert np.linalg.norm(mean - 10 * (0.04 + 0.96 * albedo)) < 0.07, mean


def test_differentiable_renderer_rejects_textured_scenes(torch_cuda):
    b = scenes.textured_room()
    scene = wpt.Scene(b.desc(), device=True, owner=b)
    with pytest.raises(wpt.WptError):
        wpt.DiffRenderer(scene, scenes.room_camera(), 32, 32, 1, 4, (8, 8), textured_material=0)
