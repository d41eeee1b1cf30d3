"""The shading-heavy terrain workload on the GPU: textured terrain, rocks
with normal maps and glTF metallic-roughness, alpha-masked foliage,
dielectric crystals with Beer attenuation and dispersion, 512 emissive
triangles, an HDR sky with a sun and a spectral caustic photon map, all in
one scene.  At reduced size every scheduler matches the oracle bit for bit;
at full size (2,006,680 triangles, 16 x 4096^2 textures, 2048 x 1024 sky,
4M photons) the schedulers agree with each other bit for bit."""
import json

import numpy as np
import pytest

import wpt

pytestmark = pytest.mark.gpu

MODES = [wpt.MODE_GRAPH, wpt.MODE_PERSISTENT, wpt.MODE_REGENERATE]


def small_manifest(tmp_path, seed=3):
    p = tmp_path / "terrain.json"
    p.write_text(json.dumps({"name": "terrain-small", "preset": "terrain", "procedural": {
        "seed": seed, "icosphere_frequency": 8, "terrain_resolution": 24, "texture_size": 32}}))
    return str(p)


def config(manifest, w, h, spp, photons, extra=()):
    status, rc, _, fp = wpt.parse_command_line(
        ["--scene-file", manifest, "--caustics", "--caustic-photons", str(photons), "--resolution", f"{w}x{h}",
         "--spp", str(spp), "--max-depth", "10", "--rr-depth", "3", "--rng", "pcg32"] + list(extra))
    assert status == wpt.OK and rc.caustic_photons == photons and rc.caustic_spectral == 1
    return rc, fp


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("builder", [wpt.BVH_HOST_SAH, wpt.BVH_DEVICE_PLOC])
def test_small_terrain_matches_oracle(torch_cuda, oracle_mod, tmp_path, mode, builder):
    manifest = small_manifest(tmp_path)
    host = wpt.HostScene(manifest)
    desc = wpt.SceneDesc.from_buffer_copy(host.desc())
    desc.bvh_builder = builder
    scene = wpt.Scene(desc, device=True, owner=host)
    rc, _ = config(manifest, 64, 40, 2, 40000)
    rc.mode = mode
    cam = host.camera(64, 40)
    r = wpt.Renderer(scene, cam, rc)
    img, _ = r.render_frame()
    want, ost = oracle_mod.Oracle(scene).render(cam, rc)
    st = r.stats()
    assert st.caustic_photons_stored == ost.caustic_photons_stored > 0
    assert np.array_equal(img, want), f"max |diff| {np.abs(img - want).max()}"
    assert st.closest_queries == ost.closest_queries and st.shadow_queries == ost.shadow_queries


def test_small_terrain_batch_equals_render_task(torch_cuda, tmp_path):
    manifest = small_manifest(tmp_path, seed=5)
    host = wpt.HostScene(manifest)
    scene = wpt.Scene(host.desc(), device=True, owner=host)
    w, h = 64, 48
    cam = host.camera(w, h)
    tasks = [wpt.batch_task(0, cam, w, h, x, y, 32, 24, 1, 3, max_depth=10, rr_start=3, clamp=10.0)
             for (x, y) in [(0, 0), (32, 24), (32, 0)]]
    outs = wpt.Batch([scene]).render(tasks)
    for t, out in zip(tasks, outs):
        cfg = wpt.render_config(w, h, spp=4, max_depth=10, rr_start=3, tile=(32, 24), clamp=10.0)
        want = wpt.Renderer(scene, cam, cfg).render_task(
            wpt.Task(0, 0, t.sample_start, t.sample_count, t.x, t.y, t.tile_width, t.tile_height))
        assert np.array_equal(out, want)


def test_full_terrain_schedulers_agree(torch_cuda):
    host = wpt.HostScene("terrain")
    desc = host.desc()
    assert desc.mesh_triangle_count == 2_006_680 and desc.texture_count == 16
    assert desc.textures[0].width == 4096 and desc.environment_width == 2048
    scene = wpt.Scene(desc, device=True, owner=host)
    assert scene.info().light_count == 512
    status, rc, _, _ = wpt.parse_command_line(["--scene", "terrain", "--caustics", "--caustic-photons", "4194304",
                                               "--resolution", "160x90", "--spp", "2", "--max-depth", "10"])
    cam = host.camera(160, 90)
    images, stored = [], []
    for mode in (wpt.MODE_GRAPH, wpt.MODE_REGENERATE):
        rc.mode = mode
        r = wpt.Renderer(scene, cam, rc)
        img, n = r.render_frame()
        st = r.stats()
        print(f"mode {mode}: {st.caustic_photons_stored} of 4194304 photons stored, photon map "
              f"{st.caustic_build_milliseconds:.1f} ms, frame {st.trace_milliseconds:.1f} ms")
        images.append(img)
        stored.append(st.caustic_photons_stored)
        r.close()
    assert np.all(np.isfinite(images[0])) and images[0].mean() > 0
    assert stored[0] == stored[1] > 1000
    assert np.array_equal(images[0], images[1])


@pytest.mark.parametrize("scheduler", [wpt.BATCH_STAGED, wpt.BATCH_REGENERATE], ids=["staged", "regenerate"])
def test_mixed_batch_with_caustic_terrain(torch_cuda, oracle_mod, tmp_path, scheduler):
    # A batch holding a small caustic terrain beside two other scenes: the
    # terrain's photon map is built for the batch from the same caustic
    # settings a renderer takes, and every task equals wpt_render_task on a
    # renderer with those settings bit for bit; small terrain tasks equal the
    # oracle's caustic render.
    manifest = small_manifest(tmp_path, seed=7)
    terrain_host = wpt.HostScene(manifest)
    others = [wpt.HostScene("cornell-box"), wpt.HostScene("sphere-field", icosphere_frequency=8)]
    hosts = [others[0], terrain_host, others[1]]
    scenes = [wpt.Scene(h.desc(), device=True, owner=h) for h in hosts]
    w, h = 96, 64
    rc, _ = config(manifest, w, h, 8, 30000)
    batch = wpt.Batch(scenes, block_size=256, scheduler=scheduler)
    stored, ms = batch.set_caustics(1, rc)
    assert stored > 0
    rng = np.random.default_rng(11)
    tasks = []
    for i in range(18):
        s = i % 3
        tw, th = int(rng.choice([16, 32, 48])), int(rng.choice([16, 24, 32]))
        x, y = int(rng.integers(0, w - tw + 1)), int(rng.integers(0, h - th + 1))
        tasks.append(wpt.batch_task(s, hosts[s].camera(w, h), w, h, x, y, tw, th, int(rng.integers(0, 3)),
                                    int(rng.integers(1, 5)), max_depth=10, rr_start=3, rng=wpt.RNG_PCG32,
                                    clamp=rc.indirect_clamp))
    outs = batch.render(tasks)
    st = batch.stats()
    closest = shadow = 0
    for t, out in zip(tasks, outs):
        cfg = wpt.RenderConfig.from_buffer_copy(rc) if t.scene == 1 else wpt.render_config(
            w, h, spp=8, max_depth=10, rr_start=3, rng=wpt.RNG_PCG32, clamp=rc.indirect_clamp)
        cfg.tiled, cfg.tile_width, cfg.tile_height = 1, t.tile_width, t.tile_height
        r = wpt.Renderer(scenes[t.scene], t.camera, cfg)
        if t.scene == 1:
            assert r.stats().caustic_photons_stored == stored
        want = r.render_task(wpt.Task(0, 0, t.sample_start, t.sample_count, t.x, t.y, t.tile_width, t.tile_height))
        q = r.stats()
        closest += q.closest_queries
        shadow += q.shadow_queries
        r.close()
        assert np.array_equal(out, want), f"task on scene {t.scene}: max |diff| {np.abs(out - want).max()}"
    assert (st.closest_queries, st.shadow_queries) == (closest, shadow)
    # Terrain tasks against the oracle with the same photon map settings.
    oracle = oracle_mod.Oracle(scenes[1])
    for t, out in zip(tasks, outs):
        if t.scene != 1:
            continue
        fb = np.zeros((h, w, 3), np.float32)
        want, ost = oracle.render(t.camera, rc, tile=(t.x, t.y, t.tile_width, t.tile_height),
                                  sample_start=t.sample_start, sample_count=t.sample_count, framebuffer=fb)
        assert ost.caustic_photons_stored == stored
        tile = want[t.y:t.y + t.tile_height, t.x:t.x + t.tile_width]
        assert np.array_equal(out, tile)
    batch.close()
