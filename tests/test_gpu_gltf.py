"""GPU parity of glTF-imported scenes (src/scene/gltf_loader.cpp:307-503,
scene_host.cpp:242-385): a sponza-preset scene built from a glTF file
(PNG textures with transforms, alpha masking, normal maps,
metallic-roughness, specular-glossiness, clearcoat, transmission with
volume attenuation, punctual lights, the top light) and a crystal room
around the same file with caustics render identically to the oracle on
every scheduler."""
import json

import numpy as np
import pytest

import wpt
from gltf_fixtures import gltf_document, write_glb

pytestmark = pytest.mark.gpu

MODES = [wpt.MODE_GRAPH, wpt.MODE_PERSISTENT, wpt.MODE_REGENERATE]


def manifest(tmp_path, body):
    doc, binary, _ = gltf_document()
    body["object"]["file"] = write_glb(tmp_path / "scene.glb", doc, binary)
    p = tmp_path / "scene.json"
    p.write_text(json.dumps(body))
    return str(p)


@pytest.mark.parametrize("mode", MODES)
def test_gltf_scene_matches_oracle(torch_cuda, oracle_mod, tmp_path, mode):
    m = manifest(tmp_path, {"preset": "sponza", "object": {"source": "glb"}, "gltf_scale": 0.8,
                            "lighting": {"add_sponza_top_light": True}})
    host = wpt.HostScene(m)
    desc = host.desc()
    assert desc.texture_count == 3 and desc.punctual_light_count == 3
    scene = wpt.Scene(desc, device=True, owner=host)
    status, rc, _, _ = wpt.parse_command_line(["--scene-file", m, "--resolution", "64x48", "--spp", "2",
                                               "--max-depth", "6", "--rr-depth", "3", "--rng", "pcg32"])
    assert status == wpt.OK
    rc.mode = mode
    cam = host.camera(64, 48)
    r = wpt.Renderer(scene, cam, rc)
    img, _ = r.render_frame()
    want, ost = oracle_mod.Oracle(scene).render(cam, rc)
    assert np.isfinite(img).all() and img.mean() > 0
    assert np.array_equal(img, want), f"max |diff| {np.abs(img - want).max()}"
    assert r.stats().closest_queries == ost.closest_queries


@pytest.mark.parametrize("mode", MODES)
def test_gltf_crystal_room_caustics_match_oracle(torch_cuda, oracle_mod, tmp_path, mode):
    m = manifest(tmp_path, {"preset": "crystal", "object": {"source": "glb", "scale": 0.3, "translation": [0, -0.4, -0.5],
                                                             "dielectric": True},
                            "room": {"neutral": True}, "lighting": {"profile": "crystal"}})
    host = wpt.HostScene(m)
    scene = wpt.Scene(host.desc(), device=True, owner=host)
    status, rc, _, _ = wpt.parse_command_line(["--scene-file", m, "--caustics", "--caustic-photons", "20000",
                                               "--resolution", "48x32", "--spp", "2", "--max-depth", "8",
                                               "--rr-depth", "3", "--rng", "pcg32"])
    assert status == wpt.OK
    rc.mode = mode
    cam = host.camera(48, 32)
    r = wpt.Renderer(scene, cam, rc)
    img, _ = r.render_frame()
    want, ost = oracle_mod.Oracle(scene).render(cam, rc)
    assert r.stats().caustic_photons_stored == ost.caustic_photons_stored
    assert np.array_equal(img, want), f"max |diff| {np.abs(img - want).max()}"
