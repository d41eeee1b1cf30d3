"""Punctual lights on the CPU: the oracle's samplePunctualLight against the
reference's known answers (tests/gltf_support_tests.cpp:19-39), whole oracle
images against the reference's point/directional checks
(tests/gltf_shading_tests.cu:55-87) in image form, and the host build's
validation and directional-first ordering."""
import math

import numpy as np
import pytest

import scenes
import wpt


def near(a, b, tol=1e-5):
    return abs(a - b) < tol


def test_point_spot_directional_known_answers(oracle_mod):
    light = wpt.make_punctual_light(wpt.LIGHT_POINT, (0, 0, 2), intensity=(8, 4, 2))
    _, inc, dist = oracle_mod.punctual_sample(light, (0, 0, 0))
    assert near(inc[0], 2) and dist == 2.0                       # inverse square
    assert near(oracle_mod.punctual_sample(light, (0, 0, -2))[1][0], 0.5)
    light.range = 2.0
    assert oracle_mod.punctual_sample(light, (0, 0, 0))[1][0] == 0  # range window ends at the range
    light.range = 0.0
    light.type = wpt.LIGHT_SPOT
    light.inner_cone_cos, light.outer_cone_cos = math.cos(0.2), math.cos(0.4)
    assert near(oracle_mod.punctual_sample(light, (0, 0, 0))[1][0], 2)   # on the axis
    assert oracle_mod.punctual_sample(light, (3, 0, 0))[1][0] == 0       # outside the cone
    angle = 0.3
    cone = (math.cos(angle) - light.outer_cone_cos) / (light.inner_cone_cos - light.outer_cone_cos)
    point = (2 * math.sin(angle), 0, 2 - 2 * math.cos(angle))
    assert near(oracle_mod.punctual_sample(light, point)[1][0], 2 * cone * cone)
    light.type = wpt.LIGHT_DIRECTIONAL
    d, inc, dist = oracle_mod.punctual_sample(light, (100, 0, 0))
    assert near(inc[0], 8) and d[2] == 1 and dist == np.float32(1e30)


def test_range_window_is_smooth(oracle_mod):
    light = wpt.make_punctual_light(wpt.LIGHT_POINT, (0, 0, 0), intensity=(1, 1, 1), range_=4.0)
    for r in (0.5, 1.0, 2.0, 3.0, 3.9):
        _, inc, _ = oracle_mod.punctual_sample(light, (r, 0, 0))
        want = (1 / r ** 2) * max(0.0, 1 - (r / 4) ** 4)
        assert abs(inc[0] - want) < 1e-5 * max(1, want)


def floor_with_lights(lights, blockers=(), albedo=0.5):
    b = wpt.SceneBuilder()
    b.black_background = True
    f = b.material(wpt.make_material(wpt.DIFFUSE, (albedo, albedo, albedo)))
    h = 50.0
    scenes.quad(b, (-h, 0, -h), (-h, 0, h), (h, 0, h), (h, 0, -h), f)
    for center, radius in blockers:
        b.sphere(center, radius, f)
    for light in lights:
        b.punctual_light(light)
    return b


def render(oracle_mod, builder, size=8, spp=2, max_depth=1, precision=32):
    scene = wpt.Scene(builder.desc(), device=False, owner=builder)
    oracle = oracle_mod.Oracle(scene)
    cfg = wpt.render_config(size, size, spp=spp, max_depth=max_depth, rr_start=64, clamp=0.0)
    fb, stats = oracle.render(scenes.look_down(0.5, fov=0.002), cfg, precision=precision)
    return fb / spp, stats


@pytest.mark.parametrize("precision", [32, 64])
def test_point_light_irradiance_closed_form(oracle_mod, precision):
    # A Lambertian floor (albedo 0.5) 2 below a point light of intensity 8:
    # albedo / pi x 8 / 2^2 = 1 / pi at the foot of the light.
    light = wpt.make_punctual_light(wpt.LIGHT_POINT, (0, 2, 0), intensity=(8, 8, 8))
    img, stats = render(oracle_mod, floor_with_lights([light]), precision=precision)
    assert np.abs(img - 1 / math.pi).max() < 2e-5
    assert stats.shadow_queries == 8 * 8 * 2


def test_point_light_shadow_segment(oracle_mod):
    light = wpt.make_punctual_light(wpt.LIGHT_POINT, (0, 2, 0), intensity=(8, 8, 8))
    img, _ = render(oracle_mod, floor_with_lights([light], blockers=[((0, 1, 0), 0.2)]))
    assert not img.any()                     # a blocker between: occluded
    img, _ = render(oracle_mod, floor_with_lights([light], blockers=[((0, 3, 0), 0.2)]))
    assert np.abs(img - 1 / math.pi).max() < 2e-5  # a blocker beyond the light is not


def test_directional_plus_positional_sum(oracle_mod):
    point = wpt.make_punctual_light(wpt.LIGHT_POINT, (0, 2, 0), intensity=(8, 8, 8))
    sun = wpt.make_punctual_light(wpt.LIGHT_DIRECTIONAL, direction=(0, -1, 0), intensity=(2, 2, 2))
    img, stats = render(oracle_mod, floor_with_lights([point, sun]))  # given positional first
    assert np.abs(img - 2 / math.pi).max() < 4e-5
    assert stats.shadow_queries == 2 * 8 * 8 * 2


def test_positional_lights_are_sampled_uniformly(oracle_mod):
    # Two equal point lights at equal distance: one is picked per vertex and
    # weighted by the count, so every sample is exactly the two-light sum.
    a = wpt.make_punctual_light(wpt.LIGHT_POINT, (0.5, 2, 0), intensity=(8, 8, 8))
    b = wpt.make_punctual_light(wpt.LIGHT_POINT, (-0.5, 2, 0), intensity=(8, 8, 8))
    img, _ = render(oracle_mod, floor_with_lights([a, b]))
    d2 = 0.25 + 4
    want = 2 * 0.5 / math.pi * 8 / d2 * (2 / math.sqrt(d2))
    assert np.abs(img - want).max() < 1e-4


def test_spot_cone_in_image(oracle_mod):
    # A spot aimed straight down with its outer cone inside the view: the
    # center pixel is lit as a point light, a floor point outside is dark.
    spot = wpt.make_punctual_light(wpt.LIGHT_SPOT, (0, 2, 0), direction=(0, -1, 0), intensity=(8, 8, 8),
                                   inner_cone_cos=math.cos(0.1), outer_cone_cos=math.cos(0.2))
    img, _ = render(oracle_mod, floor_with_lights([spot]))
    assert np.abs(img - 1 / math.pi).max() < 2e-5
    off = wpt.make_punctual_light(wpt.LIGHT_SPOT, (3, 2, 0), direction=(0, -1, 0), intensity=(8, 8, 8),
                                  inner_cone_cos=math.cos(0.1), outer_cone_cos=math.cos(0.2))
    img, _ = render(oracle_mod, floor_with_lights([off]))
    assert not img.any()


def test_product_orders_directional_first_and_validates():
    lights = [wpt.make_punctual_light(wpt.LIGHT_POINT, (0, 1, 0)),
              wpt.make_punctual_light(wpt.LIGHT_DIRECTIONAL, direction=(0, -1, 0)),
              wpt.make_punctual_light(wpt.LIGHT_SPOT, (1, 1, 0))]
    b = floor_with_lights(lights)
    wpt.Scene(b.desc(), device=False, owner=b)  # accepted
    bad = floor_with_lights([wpt.make_punctual_light(wpt.LIGHT_POINT, (0, float("nan"), 0))])
    with pytest.raises(wpt.WptError):
        wpt.Scene(bad.desc(), device=False, owner=bad)
    kind = floor_with_lights([wpt.make_punctual_light(7, (0, 1, 0))])
    with pytest.raises(wpt.WptError):
        wpt.Scene(kind.desc(), device=False, owner=kind)
