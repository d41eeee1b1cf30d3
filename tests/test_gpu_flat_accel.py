"""The flat-triangle accelerators on the device: closest-hit records and
occlusion bits of every tree (SAH, median, spatial-split, eight-wide) equal
the oracle's bit for bit, and images of a lit scene over each tree equal the
oracle's under all three schedulers."""
import numpy as np
import pytest

import test_cpu_flat_accel as flat
import test_gpu_traversal as trav
import wpt

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("accel", flat.ACCELS)
def test_traversal_bit_exact(torch_cuda, oracle_mod, accel):
    b = flat.flat_scene(accel)
    scene = wpt.Scene(b.desc(), device=True, owner=b)
    oracle = oracle_mod.Oracle(scene)
    rays = flat.random_rays(20000, 3)
    want = oracle.trace(rays)
    got = trav.gpu_trace(torch_cuda, scene, rays)
    trav.assert_hits_equal(f"accelerator {accel}", got, want)
    hit = want["kind"] != wpt.HIT_NONE
    limits = np.where(hit, want["t"], 1000.0).astype(np.float32)
    for tmax in (limits * 0.5, limits + 0.01):
        assert np.array_equal(trav.gpu_occluded(torch_cuda, scene, rays, tmax), oracle.occluded(rays, tmax))


def lit_flat_scene(accel):
    b = flat.flat_scene(accel, n=1200, slivers=80)
    light = b.material(wpt.make_material(wpt.EMISSIVE, emission=(6.0, 5.0, 4.0)))
    glass = b.material(wpt.make_material(wpt.DIELECTRIC, (1, 1, 1), ior=1.5))
    b.flat_triangles.append(wpt.make_triangle((-2, 4.5, -2), (2, 4.5, -2), (2, 4.5, 2), light))
    b.flat_triangles.append(wpt.make_triangle((-2, 4.5, -2), (2, 4.5, 2), (-2, 4.5, 2), light))
    b.flat_triangles.append(wpt.make_triangle((-1, -1, 3.5), (1, -1, 3.5), (0, 1.5, 3.5), glass))
    return b


@pytest.mark.parametrize("mode", [wpt.MODE_GRAPH, wpt.MODE_PERSISTENT, wpt.MODE_REGENERATE])
@pytest.mark.parametrize("accel", flat.ACCELS)
def test_images_match_oracle(torch_cuda, oracle_mod, accel, mode):
    import scenes
    b = lit_flat_scene(accel)
    scene = wpt.Scene(b.desc(), device=True, owner=b)
    cam = scenes.camera((0.2, 0.3, 7.5), (0, 0, -1), fov=0.9)
    cfg = wpt.render_config(48, 40, spp=2, max_depth=6, rr_start=3, mode=mode)
    r = wpt.Renderer(scene, cam, cfg)
    img, _ = r.render_frame()
    st = r.stats()
    r.close()
    want, ost = oracle_mod.Oracle(scene).render(cam, cfg)
    assert np.array_equal(img, want), f"max |diff| {np.abs(img - want).max()}"
    assert (st.closest_queries, st.shadow_queries) == (ost.closest_queries, ost.shadow_queries)
    assert img.mean() > 0
