"""The oracle-pinning cases on the device, through the C ABI: camera mapping,
indirect clamp, thin glass pane, tilted emitter and Russian-roulette energy,
rendered by all three schedulers; every image equals the oracle's bit for
bit and passes the same closed-form checks (test_cpu_pins.py)."""
import numpy as np
import pytest

import test_cpu_pins as pins
import wpt

pytestmark = pytest.mark.gpu

MODES = [wpt.MODE_GRAPH, wpt.MODE_PERSISTENT, wpt.MODE_REGENERATE]


def render_both(oracle_mod, builder, cam, cfg):
    scene = wpt.Scene(builder.desc(), device=True, owner=builder)
    r = wpt.Renderer(scene, cam, cfg)
    img, n = r.render_frame()
    want, ost = oracle_mod.Oracle(scene).render(cam, cfg)
    st = r.stats()
    r.close()
    assert n == cfg.samples_per_pixel
    assert np.array_equal(img, want), f"max |diff| {np.abs(img - want).max()}"
    assert st.closest_queries == ost.closest_queries and st.shadow_queries == ost.shadow_queries
    return img / cfg.samples_per_pixel


@pytest.mark.parametrize("mode", MODES)
def test_camera_mapping_on_device(torch_cuda, oracle_mod, mode):
    b, cam = pins.camera_scene()
    spp = 16
    cfg = wpt.render_config(pins.W, pins.H, spp=spp, max_depth=1, rr_start=64, clamp=0.0, mode=mode)
    pins.check_camera_image(render_both(oracle_mod, b, cam, cfg), spp)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("mirror_albedo,clamp", [(1.0, 10.0), (0.001, 10.0), (1.0, 0.0)])
def test_clamp_known_answers_on_device(torch_cuda, oracle_mod, mode, mirror_albedo, clamp):
    b, cam = pins.clamp_scene(mirror_albedo)
    cfg = wpt.render_config(24, 24, spp=4, max_depth=2, rr_start=64, clamp=clamp, mode=mode)
    img = render_both(oracle_mod, b, cam, cfg)
    pins.check_clamp_image(img, pins.clamp_expectations(mirror_albedo, clamp))


@pytest.mark.parametrize("mode", MODES)
def test_thin_pane_on_device(torch_cuda, oracle_mod, mode):
    b, cam = pins.thin_pane_scene()
    spp = 64
    cfg = wpt.render_config(32, 32, spp=spp, max_depth=64, rr_start=64, clamp=0.0, mode=mode)
    pins.check_thin_pane(render_both(oracle_mod, b, cam, cfg), 32 * 32 * spp)


def test_thin_pane_exit_trace_on_device(torch_cuda, oracle_mod):
    # Entry at z = 0, origin pushed off the entry point: the next closest hit
    # is the pane's exit face, from behind, within 1.5 mm.
    torch = torch_cuda
    b, _ = pins.thin_pane_scene()
    scene = wpt.Scene(b.desc(), device=True, owner=b)
    origin = oracle_mod.offset_ray_origin((0, 0, 0), (0, 0, 1), (0, 0, -1), 1.0)
    rays = np.array([[0, origin[0]], [0, origin[1]], [1, origin[2]], [0, 0], [0, 0], [-1, -1]], np.float32)
    hits = wpt.hits_to_numpy(scene.trace_closest(torch.from_numpy(rays).cuda()))
    want = oracle_mod.Oracle(scene).trace(rays)
    assert hits.tobytes() == want.tobytes()
    assert hits["t"][0] == 1.0 and hits["primitive"][1] in (2, 3) and 0 < hits["t"][1] <= pins.PANE


@pytest.mark.parametrize("mode", MODES)
def test_tilted_emitter_on_device(torch_cuda, oracle_mod, mode):
    images = []
    for ignore in (False, True):
        b, cam = pins.tilted_emitter_scene(ignore)
        cfg = wpt.render_config(24, 24, spp=4, max_depth=1, rr_start=64, clamp=0.0, mode=mode)
        images.append(render_both(oracle_mod, b, cam, cfg))
    assert np.array_equal(images[0], images[1]) and images[0].mean() > 0.05


@pytest.mark.parametrize("mode", MODES)
def test_roulette_energy_on_device(torch_cuda, oracle_mod, mode):
    b, cam = pins.rr_floor_scene()
    cfg = wpt.render_config(32, 32, spp=32, max_depth=64, rr_start=1, clamp=0.0, mode=mode)
    pins.check_rr_energy(render_both(oracle_mod, b, cam, cfg))
