"""GPU parity of environment-map and punctual-light rendering: images of
scenes lit by an equirectangular map and by point, spot and directional
lights (with and without triangle lights, every material kind, rotated and
scaled maps) from all three schedulers and the batch consumer, against the
oracle's FP32 restatement bit for bit; and the reference's HDR-sun energy
checks (tests/environment_shading_tests.cu) on the device."""
import math

import numpy as np
import pytest

import scenes
import test_cpu_environment as cpu_env
import wpt

pytestmark = pytest.mark.gpu

MODES = [wpt.MODE_GRAPH, wpt.MODE_PERSISTENT, wpt.MODE_REGENERATE]


def random_map(width=32, height=16, seed=3):
    rng = np.random.default_rng(seed)
    env = (rng.random((height, width, 3)) ** 3 * 4).astype(np.float32)
    env[height // 4, width // 3] = (400, 380, 300)  # a small sun
    return env


class PresetWithMap:
    """A host preset's scene description with an environment map (None: no
    map) and punctual lights attached."""

    def __init__(self, preset, env, intensity, rotation_degrees, freq=0, lights=()):
        self.host = wpt.HostScene(preset, icosphere_frequency=freq)
        d = wpt.SceneDesc.from_buffer_copy(self.host.desc())
        if env is not None:
            self.env = np.ascontiguousarray(env, np.float32)
            d.environment_rgb = self.env.ctypes.data_as(wpt.C.POINTER(wpt.C.c_float))
            d.environment_height, d.environment_width = self.env.shape[0], self.env.shape[1]
            d.environment_intensity = intensity
            d.environment_rotation_degrees = rotation_degrees
        if lights:
            self.lights = (wpt.PunctualLight * len(lights))(*lights)
            d.punctual_lights = wpt.C.cast(self.lights, wpt.C.POINTER(wpt.PunctualLight))
            d.punctual_light_count = len(lights)
        self._desc = d

    def desc(self):
        return self._desc

    def camera(self, w, h):
        return self.host.camera(w, h)


def preset_cases():
    return [("cornell-box", 0, 1.0, 30.0), ("rtweek", 8, 0.7, -75.0), ("prism", 8, 2.0, 200.0)]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("preset,freq,intensity,rotation", preset_cases())
def test_preset_with_map_matches_oracle(torch_cuda, oracle_mod, mode, preset, freq, intensity, rotation):
    src = PresetWithMap(preset, random_map(), intensity, rotation, freq)
    scene = wpt.Scene(src.desc(), device=True, owner=src)
    cam = src.camera(96, 64)
    cfg = wpt.render_config(96, 64, spp=3, max_depth=10, rr_start=3, mode=mode, clamp=10.0)
    r = wpt.Renderer(scene, cam, cfg)
    img, n = r.render_frame()
    want, ost = oracle_mod.Oracle(scene).render(cam, cfg)
    assert n == 3
    assert np.array_equal(img, want), f"max |diff| {np.abs(img - want).max()}"
    st = r.stats()
    assert st.closest_queries == ost.closest_queries
    assert st.shadow_queries == ost.shadow_queries


def mixed_lights(center=(0.5, 0.9, 0.5)):
    c = np.asarray(center, np.float32)
    return [wpt.make_punctual_light(wpt.LIGHT_POINT, c + (0.2, 0, 0.1), intensity=(0.6, 0.5, 0.4)),
            wpt.make_punctual_light(wpt.LIGHT_DIRECTIONAL, direction=(0.3, -1, 0.2), intensity=(0.8, 0.8, 0.7)),
            wpt.make_punctual_light(wpt.LIGHT_SPOT, c + (-0.2, 0, 0), direction=(0, -1, 0), intensity=(1.5, 1, 1),
                                    range_=3.0, inner_cone_cos=0.95, outer_cone_cos=0.8),
            wpt.make_punctual_light(wpt.LIGHT_DIRECTIONAL, direction=(-0.5, -1, -0.1), intensity=(0.2, 0.3, 0.5))]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("with_map", [False, True])
def test_punctual_lights_match_oracle(torch_cuda, oracle_mod, mode, with_map):
    src = PresetWithMap("cornell-box", random_map() if with_map else None, 0.5, 45.0, lights=mixed_lights())
    scene = wpt.Scene(src.desc(), device=True, owner=src)
    cam = src.camera(96, 64)
    cfg = wpt.render_config(96, 64, spp=3, max_depth=8, rr_start=3, mode=mode, clamp=10.0)
    r = wpt.Renderer(scene, cam, cfg)
    img, _ = r.render_frame()
    want, ost = oracle_mod.Oracle(scene).render(cam, cfg)
    assert np.array_equal(img, want), f"max |diff| {np.abs(img - want).max()}"
    assert r.stats().shadow_queries == ost.shadow_queries


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("directional", [7, 10])
def test_many_directional_lights_match_oracle(torch_cuda, oracle_mod, mode, directional):
    # Every directional light is a radiance event of a diffuse vertex: with
    # the area light, the positional light, the map and two escape slots, 7
    # lights fill the 12 events a vertex may store; 10 exceed it, and such
    # scenes trace their shadow rays inline.  Both must give the oracle's
    # image and counts.
    rng = np.random.default_rng(directional)
    lights = [wpt.make_punctual_light(wpt.LIGHT_DIRECTIONAL, direction=(rng.uniform(-1, 1), -1.0, rng.uniform(-1, 1)),
                                      intensity=tuple(rng.uniform(0.05, 0.2, 3))) for _ in range(directional)]
    lights.append(wpt.make_punctual_light(wpt.LIGHT_POINT, (0.4, 0.8, 0.6), intensity=(0.4, 0.3, 0.2)))
    src = PresetWithMap("cornell-box", random_map(), 0.5, 20.0, lights=lights)
    scene = wpt.Scene(src.desc(), device=True, owner=src)
    cam = src.camera(64, 48)
    cfg = wpt.render_config(64, 48, spp=2, max_depth=6, rr_start=3, mode=mode, clamp=10.0)
    r = wpt.Renderer(scene, cam, cfg)
    img, _ = r.render_frame()
    want, ost = oracle_mod.Oracle(scene).render(cam, cfg)
    assert np.array_equal(img, want), f"max |diff| {np.abs(img - want).max()}"
    st = r.stats()
    assert st.closest_queries == ost.closest_queries and st.shadow_queries == ost.shadow_queries


@pytest.mark.parametrize("mode", MODES)
def test_tiled_render_task_with_map(torch_cuda, oracle_mod, mode):
    src = PresetWithMap("rtweek", random_map(48, 24, 9), 1.3, 10.0, 8)
    scene = wpt.Scene(src.desc(), device=True, owner=src)
    cam = src.camera(80, 48)
    cfg = wpt.render_config(80, 48, spp=4, max_depth=8, rr_start=3, mode=mode, tile=(32, 16), clamp=0.0)
    full, _ = wpt.Renderer(scene, cam, cfg).render_frame()
    want, _ = oracle_mod.Oracle(scene).render(cam, cfg)
    assert np.array_equal(full, want)
    r = wpt.Renderer(scene, cam, cfg)
    task = wpt.Task(0, 0, 1, 2, 32, 16, 32, 16)
    tile = r.render_task(task)
    part, _ = oracle_mod.Oracle(scene).render(cam, cfg, tile=(32, 16, 32, 16), sample_start=1, sample_count=2)
    assert np.array_equal(tile, part[16:32, 32:64])


@pytest.mark.parametrize("case", ["diffuse", "glass", "mirror", "slab"])
def test_small_map_scenes_match_oracle(torch_cuda, oracle_mod, case):
    if case == "slab":
        b = scenes.env_slab(wpt.make_material(wpt.DIFFUSE, (0.2, 0.4, 0.6)), scenes.sun_map())
    elif case == "glass":
        b = scenes.env_floor(wpt.make_material(wpt.DIELECTRIC, (0.2, 0.4, 0.6), ior=1.5), scenes.constant_map(4.0),
                             intensity=2.5)
    elif case == "mirror":
        b = scenes.env_floor(wpt.make_material(wpt.METAL, (0.5, 0.5, 0.5), roughness=0.0),
                             scenes.constant_map(50000.0))
    else:
        b = scenes.env_floor(wpt.make_material(wpt.DIFFUSE, (0.2, 0.4, 0.6)), scenes.constant_map(4.0), intensity=2.5)
    scene = wpt.Scene(b.desc(), device=True, owner=b)
    cam = scenes.look_down(fov=0.02)
    for mode in MODES:
        cfg = wpt.render_config(32, 32, spp=4, max_depth=12, rr_start=64, mode=mode, clamp=10.0)
        img, _ = wpt.Renderer(scene, cam, cfg).render_frame()
        want, _ = oracle_mod.Oracle(scene).render(cam, cfg)
        assert np.array_equal(img, want), (case, mode)


@pytest.mark.parametrize("scheduler", [wpt.BATCH_STAGED, wpt.BATCH_REGENERATE], ids=["staged", "regenerate"])
def test_batch_with_map_scenes_equals_render_task(torch_cuda, scheduler):
    srcs = [PresetWithMap("cornell-box", random_map(), 1.0, 30.0),
            PresetWithMap("rtweek", random_map(64, 32, 4), 0.5, 120.0, 8)]
    dev = [wpt.Scene(s.desc(), device=True, owner=s) for s in srcs]
    w, h = 96, 64
    tasks = []
    for i, s in enumerate(srcs):
        for (x, y) in [(0, 0), (32, 16), (64, 48)]:
            tasks.append(wpt.batch_task(i, s.camera(w, h), w, h, x, y, 32, 16, 1, 3, max_depth=8, rr_start=3,
                                        clamp=10.0))
    outs = wpt.Batch(dev, scheduler=scheduler).render(tasks)
    for t, out in zip(tasks, outs):
        cfg = wpt.render_config(w, h, spp=4, max_depth=8, rr_start=3, tile=(32, 16), clamp=10.0)
        r = wpt.Renderer(dev[t.scene], srcs[t.scene].camera(w, h), cfg)
        want = r.render_task(wpt.Task(0, 0, t.sample_start, t.sample_count, t.x, t.y, t.tile_width, t.tile_height))
        assert np.array_equal(out, want)


@pytest.mark.parametrize("case", ["direct", "metal", "slab"])
def test_hdr_sun_energy_on_device(torch_cuda, case):
    # environment_shading_tests.cu:180-222 on the regenerating scheduler at
    # 128 x 128 x 16 paths: within 3% of the independent integral.
    sun = scenes.sun_map()
    if case == "slab":
        b = scenes.env_slab(wpt.make_material(wpt.DIFFUSE, (0.2, 0.4, 0.6)), sun)
        depth, want = 12, cpu_env.integrated_sun(170, 64, 512, 256, 50000, slab=True) * 0.2
    elif case == "metal":
        b = scenes.env_floor(wpt.make_material(wpt.METAL, (0.8, 0.8, 0.8), roughness=0.5), sun)
        depth, want = 2, cpu_env.integrated_sun(170, 64, 512, 256, 50000, metal=True)
    else:
        b = scenes.env_floor(wpt.make_material(wpt.DIFFUSE, (0.2, 0.4, 0.6)), sun)
        depth, want = 2, cpu_env.integrated_sun(170, 64, 512, 256, 50000) * 0.2
    scene = wpt.Scene(b.desc(), device=True, owner=b)
    cfg = wpt.render_config(128, 128, spp=16, max_depth=depth, rr_start=64, mode=wpt.MODE_REGENERATE, clamp=0.0)
    img, n = wpt.Renderer(scene, scenes.look_down(fov=0.02), cfg).render_frame()
    mean = img[..., 0].astype(np.float64).mean() / n
    assert abs(mean - want) < 0.03 * want, (case, mean, want)
    assert math.isfinite(mean)
