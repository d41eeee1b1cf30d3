"""Textured and glTF materials on the CPU: the oracle's texture lookups,
material channels, normal maps and glTF lobes against the reference's
known answers (tests/gltf_support_tests.cpp:40-131,
tests/gltf_shading_tests.cu:40-87) and closed forms, and whole oracle images
with textured and glTF materials against closed forms
(tests/environment_shading_tests.cu:140-170).  The product's host build of
the same scenes is checked for validation errors."""
import math

import numpy as np
import pytest

import scenes
import wpt

PI = 3.14159265358979323846


def near(a, b, tol=1e-5):
    return abs(float(a) - float(b)) < tol


def one_texel_scene(exts, texel=(128, 64, 32, 128), materials=None):
    b = wpt.SceneBuilder()
    b.texture(np.array(texel, np.uint8).reshape(1, 1, 4))
    for m, e in zip(materials, exts):
        b.material(m, e)
    return b


def oracle_of(oracle_mod, b):
    scene = wpt.Scene(b.desc(), device=False, owner=b)
    return scene, oracle_mod.Oracle(scene)


def srgb_channel(v):
    v = np.float32(v)
    return v / np.float32(12.92) if v <= np.float32(0.04045) else ((v + np.float32(0.055)) / np.float32(1.055)) ** 2.4


# ---- texture channels (gltf_shading_tests.cu:40-76)

def test_texture_channels_decode_like_the_reference(oracle_mod):
    diffuse = wpt.make_material(wpt.DIFFUSE, (0.5, 0.5, 0.5))
    glossy = wpt.make_material(wpt.SPECULAR_GLOSSINESS, (0.8, 0.7, 0.6), specular=(0.5, 0.4, 0.3), glossiness=0.8,
                               opacity=0.6)
    glossy_ext = wpt.make_material_ext(specular_glossiness_workflow=True, albedo_texture=0,
                                       specular_glossiness_texture=0)
    pbr = wpt.make_material(wpt.DIFFUSE, (0.5, 0.5, 0.5), ior=1.5)
    pbr_ext = wpt.make_material_ext(gltf_pbr_workflow=True, specular_weight=0.6, specular_color=(2, 1, 0.5),
                                    specular_weight_texture=0, specular_color_texture=0, clearcoat=0.8,
                                    clearcoat_roughness=0.4, clearcoat_texture=0, clearcoat_roughness_texture=0)
    b = one_texel_scene([None, glossy_ext, pbr_ext], materials=[diffuse, glossy, pbr])
    _, oracle = oracle_of(oracle_mod, b)
    srgb = srgb_channel(128 / 255)
    g = oracle.material_terms(1)
    assert near(g["specular_f0"][0], 0.5 * srgb, 1e-6)          # specular rgb decoded as sRGB
    assert near(g["roughness"], 1 - 0.8 * 128 / 255, 1e-6)      # glossiness alpha stays linear
    assert near(g["opacity"], 0.6 * 128 / 255, 1e-6)            # opacity alpha stays linear
    assert near(g["albedo"][0], 0.8 * srgb, 1e-6)               # diffuse factor x linear texture
    p = oracle.material_terms(2)
    assert near(p["dielectric_f0"][0], 0.04 * 2 * srgb * 0.6 * 128 / 255, 1e-6)  # specular color and strength
    assert np.allclose([p["f90"], p["coat"], p["coat_roughness"]],
                       [0.6 * 128 / 255, 0.8 * 128 / 255, 0.4 * 64 / 255], atol=1e-6)  # channels stay linear
    assert np.allclose(p["coat_normal"], (0, 0, 1), atol=1e-6)  # no coat normal map: the coat keeps the normal


def alpha_probe(oracle_mod, alpha, binding=None):
    """Opacity 1 x texture alpha: the raw filtered alpha channel."""
    h, w = alpha.shape
    t = np.zeros((h, w, 4), np.uint8)
    t[..., 3] = alpha
    b = wpt.SceneBuilder()
    b.texture(t)
    b.material(wpt.make_material(wpt.DIFFUSE), wpt.make_material_ext(albedo_texture=binding or
                                                                      wpt.texture_binding(0)))
    _, oracle = oracle_of(oracle_mod, b)
    return lambda u, v: oracle.material_terms(0, (u, v))["opacity"]


def test_bilinear_texture_lookup_closed_forms(oracle_mod):
    alpha = np.array([[10, 200, 40, 90], [250, 0, 120, 60]], np.uint8)
    look = alpha_probe(oracle_mod, alpha)
    h, w = alpha.shape
    for y in range(h):
        for x in range(w):
            # Texel centers; the image's row 0 is read at v = 1.
            u, v = (x + 0.5) / w, 1 - (y + 0.5) / h
            assert near(look(u, v), alpha[y, x] / 255, 1e-6), (x, y)
            assert near(look(u + 1, v - 2), alpha[y, x] / 255, 1e-5)  # wrapping in both directions
    # Halfway between horizontal neighbours, and across the u seam.
    assert near(look(1.0 / w, 1 - 0.5 / h), (alpha[0, 0] + alpha[0, 1]) / 510, 1e-6)
    assert near(look(0.0, 1 - 0.5 / h), (alpha[0, 0] + alpha[0, w - 1]) / 510, 1e-6)
    # The four-texel average at a shared corner.
    assert near(look(2.0 / w, 0.5), alpha[:, 1:3].astype(np.float64).mean() / 255, 1e-6)
    assert look(float("nan"), 0.25) == look(0.0, 0.25)  # non-finite coordinates read 0


def test_texture_transform_scale_rotation_offset(oracle_mod):
    alpha = (np.arange(64).reshape(8, 8) * 4).astype(np.uint8)
    plain = alpha_probe(oracle_mod, alpha)
    scaled = alpha_probe(oracle_mod, alpha, wpt.texture_binding(0, scale=(2.0, 2.0), transform=True))
    rotated = alpha_probe(oracle_mod, alpha, wpt.texture_binding(0, rotation=math.pi / 2, offset=(0.25, 0.0),
                                                                 transform=True))
    for (u, v) in ((0.1, 0.2), (0.3, 0.45), (0.05, 0.4)):
        assert near(scaled(u, v), plain(2 * u, 2 * v), 1e-5)
        # rotation pi/2: (u, v) -> (-v, u), then the offset.
        assert near(rotated(u, v), plain(-v + 0.25, u), 2e-3)


def test_normal_map_frames(oracle_mod):
    def scene(texel, use=True, min_cos=0.0, scale=1.0):
        b = wpt.SceneBuilder()
        b.texture(np.array(texel, np.uint8).reshape(1, 1, 4))
        b.material(wpt.make_material(wpt.DIFFUSE), wpt.make_material_ext(normal_texture=0, normal_scale=scale))
        b.use_normal_maps = use
        b.normal_map_minimum_cosine = min_cos
        return oracle_of(oracle_mod, b)[1].material_terms(0)["normal"]
    flat = scene((128, 128, 255, 255))
    assert np.allclose(flat, (0, 0, 1), atol=0.01)
    tilt = (200, 90, 220, 255)
    x, y, z = (np.array(tilt[:3]) / 255) * 2 - 1
    want = np.array([x * 0.7, y * 0.7, z])
    assert np.allclose(scene(tilt, scale=0.7), want / np.linalg.norm(want), atol=1e-5)
    assert np.allclose(scene(tilt, use=False), (0, 0, 1))            # switched off
    clamped = scene((255, 128, 140, 255), min_cos=0.8)                # grazing texel
    assert near(clamped[2], 0.8, 1e-5) and near(np.linalg.norm(clamped), 1.0, 1e-5)


# ---- glTF lobes (gltf_support_tests.cpp:76-131)

def test_gltf_layered_sampler_matches_independent_quadrature(oracle_mod):
    n = np.array([0, 0, 1.0])
    view = np.array([0.3, 0, 1.0]) / np.linalg.norm([0.3, 0, 1.0])
    s = oracle_mod.gltf_surface((0.5, 0.3, 0.1), n, n, 0.3, 0.55, (0.04, 0.04, 0.04), 1.0, 0.8, 0.4)

    def kernel(nv, nl, nh, rough):  # independent double-precision GGX, no oracle calls
        a2 = rough ** 4
        d = a2 / (PI * (1 - nh * nh + a2 * nh * nh) ** 2)
        g = lambda z: 2 * z / (z + math.sqrt(a2 + (1 - a2) * z * z))
        return d * g(nv) * g(nl) / (4 * nv * nl)

    rng = np.random.default_rng(139)
    nv = float(view @ n)
    coat = 0.8 * (0.04 + 0.96 * (1 - nv) ** 5)
    count = 40000
    reference = estimate = 0.0
    for _ in range(count):
        z, phi = rng.random(), 2 * PI * rng.random()
        r = math.sqrt(1 - z * z)
        light = np.array([r * math.cos(phi), r * math.sin(phi), z])
        half = (view + light) / np.linalg.norm(view + light)
        vh, nh = float(view @ half), half[2]
        w5 = (1 - vh) ** 5
        fd, fm = 0.04 + 0.96 * w5, 0.5 + 0.5 * w5
        base = (1 - 0.3) * (1 - fd) * 0.5 / PI + ((1 - 0.3) * fd + 0.3 * fm) * kernel(nv, z, nh, 0.55)
        reference += ((1 - coat) * base + coat * kernel(nv, z, nh, 0.4)) * z * 2 * PI
        choice, u, v = rng.random(3)
        _, _, weights = oracle_mod.gltf_eval(s, view, (0, 0, 1), 1 / PI)
        if choice < weights[2]:
            light = oracle_mod.gltf_sample_reflection(n, view, 0.4, u, v)
        elif choice < weights[2] + weights[1]:
            light = oracle_mod.gltf_sample_reflection(n, view, 0.55, u, v)
        else:
            rr = math.sqrt(u)
            light = np.array([rr * math.cos(2 * PI * v), rr * math.sin(2 * PI * v), math.sqrt(1 - u)])
        value, pdf, _ = oracle_mod.gltf_eval(s, view, light, max(0.0, float(light[2])) / PI)
        if pdf > 0:
            estimate += value[0] / pdf
    assert abs(reference / count - estimate / count) < 0.015
    assert 0 < estimate / count < 1


def test_gltf_lambert_recovery_and_delta_lobes(oracle_mod):
    n = (0, 0, 1.0)
    s = oracle_mod.gltf_surface((0.5, 0.3, 0.1), n, n, 0.0, 0.55, (0, 0, 0), 0.0, 0.0, 0.4)
    value, pdf, _ = oracle_mod.gltf_eval(s, n, n, 1 / PI)
    assert near(value[0], 0.5 / PI) and near(pdf, 1 / PI)   # specular zero recovers Lambert exactly
    s = oracle_mod.gltf_surface((0.5, 0.3, 0.1), n, n, 0.25, 0.0, (0, 0, 0), 0.0, 1.0, 0.0)
    value, pdf, w = oracle_mod.gltf_eval(s, n, n, 1 / PI)
    assert np.allclose(w, (0.96 * 0.75, 0.96 * 0.25, 0.04), atol=1e-5)       # delta selection masses
    assert near(value[0], 0.96 * 0.75 * 0.5 / PI) and near(pdf, 0.96 * 0.75 / PI)  # deltas out of NEE/pdf
    s = oracle_mod.gltf_surface((0.5, 0.3, 0.1), n, n, 1.0, 0.0, (0, 0, 0), 0.0, 0.0, 0.0)
    value, pdf, _ = oracle_mod.gltf_eval(s, n, n, 1 / PI)
    assert not value.any() and pdf == 0                                          # perfect metal stays a delta


def test_gltf_reflection_frame_correction(oracle_mod):
    n = np.array([0, 0, 1.0], np.float32)
    smooth = np.array([-0.57, 0, 0.82]) / np.linalg.norm([-0.57, 0, 0.82])
    oblique = np.array([0.7846, 0, 0.62]) / np.linalg.norm([0.7846, 0, 0.62])
    assert n @ (2 * (smooth @ oblique) * smooth - oblique) < -0.4
    c = oracle_mod.gltf_reflection_normal(smooth, n, oblique).astype(np.float64)
    reflected = n @ (2 * (c @ oblique) * c - oblique)
    assert 0.009999 <= reflected < 0.01001          # meets the horizon margin with minimal rotation
    assert np.array_equal(oracle_mod.gltf_reflection_normal(n, n, oblique), n)  # valid frame unchanged


# ---- whole images

def render(oracle_mod, b, cam, size=16, spp=8, depth=4, rr=64, clamp=0.0):
    scene = wpt.Scene(b.desc(), device=False, owner=b)
    cfg = wpt.render_config(size, size, spp=spp, max_depth=depth, rr_start=rr, clamp=clamp)
    fb, _ = oracle_mod.Oracle(scene).render(cam, cfg)
    return fb / spp


def test_textured_emitter_is_exact(oracle_mod):
    # An emissive quad under a uniform texel c: Le x (c / 255)^2.2 exactly.
    b = wpt.SceneBuilder()
    b.black_background = True
    b.texture(np.full((4, 4, 4), 150, np.uint8))
    m = b.material(wpt.make_material(wpt.EMISSIVE, emission=(2.0, 3.0, 4.0)),
                   wpt.make_material_ext(emissive_texture=0))
    scenes.textured_quad(b, (-5, 0, -5), (-5, 0, 5), (5, 0, 5), (5, 0, -5), m)
    img = render(oracle_mod, b, scenes.look_down(0.5, fov=0.3), depth=1)
    want = np.array([2.0, 3.0, 4.0]) * (150 / 255) ** 2.2
    assert np.allclose(img.reshape(-1, 3), want, rtol=2e-6)


@pytest.mark.parametrize("alpha,visible", [(0, False), (200, True)])
def test_alpha_masked_texture_passes_rays(oracle_mod, alpha, visible):
    # A masked quad over an emitter: invisible below the cutoff, opaque above.
    b = wpt.SceneBuilder()
    b.black_background = True
    t = np.zeros((2, 2, 4), np.uint8)
    t[..., 3] = alpha
    b.texture(t)
    leaf = b.material(wpt.make_material(wpt.DIFFUSE, (0, 0, 0), alpha_masked=True, alpha_cutoff=0.5),
                      wpt.make_material_ext(albedo_texture=0))
    e = b.material(wpt.make_material(wpt.EMISSIVE, emission=(1.5, 1.5, 1.5)))
    scenes.quad(b, (-5, 0, -5), (-5, 0, 5), (5, 0, 5), (5, 0, -5), e)
    scenes.textured_quad(b, (-5, 0.2, -5), (-5, 0.2, 5), (5, 0.2, 5), (5, 0.2, -5), leaf)
    img = render(oracle_mod, b, scenes.look_down(0.5, fov=0.3), depth=2)
    assert np.allclose(img, 0.0 if visible else 1.5)


def test_gltf_smooth_coat_under_constant_map(oracle_mod):
    # environment_shading_tests.cu:160-170: specular weight 0 and a smooth
    # coat of 1 over the base colour reflect 0.04 + 0.96 x base under a
    # constant map (4 x 2.5 = 10); NEE, the delta coat split and the miss
    # must not double count.
    albedo = np.array([0.2, 0.4, 0.6])
    ext = wpt.make_material_ext(gltf_pbr_workflow=True, specular_weight=0.0, clearcoat=1.0, clearcoat_roughness=0.0)
    b = scenes.env_floor(wpt.make_material(wpt.DIFFUSE, albedo, ior=1.5), scenes.constant_map(4.0), intensity=2.5)
    b.material_exts[0] = ext
    img = render(oracle_mod, b, scenes.look_down(fov=0.02), size=64, spp=16, depth=12)
    mean = img.reshape(-1, 3).mean(axis=0)
    assert np.linalg.norm(mean - 10 * (0.04 + 0.96 * albedo)) < 0.07, mean
    # A perfect metal under the coat: every sample is exact.
    ext.gltf_pbr_workflow = 1
    b2 = scenes.env_floor(wpt.make_material(wpt.DIFFUSE, albedo, ior=1.5, metallic=1.0), scenes.constant_map(4.0),
                          intensity=2.5)
    b2.material_exts[0] = wpt.make_material_ext(gltf_pbr_workflow=True, specular_weight=0.0, clearcoat=1.0,
                                                clearcoat_roughness=0.0)
    img = render(oracle_mod, b2, scenes.look_down(fov=0.02), size=16, spp=4, depth=12)
    assert np.abs(img.reshape(-1, 3) - 10 * (0.04 + 0.96 * albedo)).max() < 1e-4


def test_gltf_lambert_floor_under_uniform_emitter(oracle_mod):
    # Specular weight 0, no coat, metallic 0: the glTF floor is Lambertian
    # and reflects its albedo under the uniform emitting plane.
    rho = 0.6
    b = scenes.floor_under_emitter(wpt.make_material(wpt.DIFFUSE, (rho, rho, rho), ior=1.5))
    b.material_exts = [wpt.make_material_ext(gltf_pbr_workflow=True, specular_weight=0.0), None]
    img = render(oracle_mod, b, scenes.look_down(0.5, fov=0.3), size=32, spp=32, depth=4)
    flat = img.reshape(-1, 3).astype(np.float64)
    sem = flat.std(axis=0).max() / math.sqrt(flat.shape[0])
    assert np.all(np.abs(flat.mean(axis=0) - rho) < max(0.002, 5 * sem))


def test_textured_room_builds_and_renders_finite(oracle_mod):
    b = scenes.textured_room()
    scene = wpt.Scene(b.desc(), device=False, owner=b)
    assert scene.info().sphere_count == 5
    img = render(oracle_mod, b, scenes.room_camera(), size=24, spp=2, depth=6, rr=3, clamp=10.0)
    assert np.isfinite(img).all() and img.mean() > 0


def test_invalid_texture_inputs_are_rejected():
    b = wpt.SceneBuilder()
    b.material(wpt.make_material(), wpt.make_material_ext(albedo_texture=wpt.texture_binding(0, texcoord=2)))
    b.texture(np.zeros((2, 2, 4), np.uint8))
    with pytest.raises(wpt.WptError):
        wpt.Scene(b.desc(), device=False, owner=b)
    d = b.desc()
    d.texture_count = 1
    d.textures = None
    with pytest.raises(wpt.WptError):
        wpt.Scene(d, device=False, owner=b)
