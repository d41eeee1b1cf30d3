"""Closed-form and known-answer checks that pin the oracle's camera mapping,
ray-origin offset, indirect clamp and Russian-roulette weighting, plus the
reference's thin-glass and tilted-emitter cases as whole renders.

Each check is chosen so that a plausible slip fails it: a flipped or
transposed image axis, an aspect ratio inverted, pixel coordinates divided
by (w - 1), a jitter outside [0, 1), an offset without the nextafter step
or with the wrong magnitude, a clamp on the max instead of the RGB mean or
before the throughput, a missing 1/p on roulette survivors.

Reference: include/scene/camera.cuh:17-27, src/renderer/renderer.cu:25-41
(camera and primary jitter), include/core/ray.cuh:8-21 (offsetRayOrigin),
include/renderer/path_state.cuh:26-43 (addPathRadiance),
include/renderer/device_path_shading.cuh:144-164 (roulette),
tests/material_transport_tests.cu:144-158 (thin pane), :180-211 (tilted
emitter), :213-233 (clamp known answers)."""
import math

import numpy as np
import pytest

import scenes
import wpt


def f32(x):
    return np.float32(x)


# ---------------------------------------------------------------- offset

def test_offset_ray_origin_known_answers(oracle_mod):
    # magnitude = max(1, |t|, |p|_inf); the push is 32 x 2^-23 x magnitude
    # along the normal toward the outgoing direction, then one more ulp per
    # pushed axis, away from the surface.
    push = 32.0 * 2.0 ** -23  # 3.814697265625e-06, exact in FP32
    cases = [
        # (position, normal, direction, t, expected)
        ((0, 0, 0), (0, 0, 1), (0, 0, 1), 0.0,
         (0.0, 0.0, float(np.nextafter(f32(push), f32(np.inf))))),
        ((0, 0, 0), (0, 0, 1), (0.3, 0, -1), 0.0,                     # leaving below the surface
         (0.0, 0.0, float(np.nextafter(f32(-push), f32(-np.inf))))),
        ((100, 0, 0), (0, 1, 0), (0, -1, 0), 0.0,                     # |p| = 100 scales the push
         (100.0, float(np.nextafter(f32(-100 * push), f32(-np.inf))), 0.0)),
        ((0.25, 0.5, 0.125), (1, 0, 0), (1, 1, 1), 250.0,            # the hit distance scales it
         (float(np.nextafter(f32(0.25) + f32(250 * push), f32(np.inf))), 0.5, 0.125)),
    ]
    for p, n, d, t, want in cases:
        got = oracle_mod.offset_ray_origin(p, n, d, t)
        assert np.array_equal(got, np.asarray(want, np.float32)), (p, n, d, t, got, want)
    # A normal with every component set pushes (and steps) every axis.
    n = np.array([0.6, 0.0, 0.8], np.float32)
    got = oracle_mod.offset_ray_origin((1, 2, 3), n, (0, 0, 1), 0.0)
    step = f32(3 * push)
    want = np.array([np.nextafter(f32(1) + n[0] * step, f32(np.inf)), 2.0,
                     np.nextafter(f32(3) + n[2] * step, f32(np.inf))], np.float32)
    assert np.array_equal(got, want)
    # The offset point lies strictly on the outgoing side, within 2 pushes.
    rng = np.random.default_rng(5)
    for _ in range(200):
        p = rng.uniform(-50, 50, 3).astype(np.float32)
        nn = rng.normal(size=3)
        nn = (nn / np.linalg.norm(nn)).astype(np.float32)
        d = rng.normal(size=3).astype(np.float32)
        t = float(rng.uniform(0, 80))
        o = oracle_mod.offset_ray_origin(p, nn, d, t).astype(np.float64)
        side = 1.0 if float(np.dot(d, nn)) >= 0 else -1.0
        h = side * float(np.dot(o - p, nn))
        mag = max(1.0, t, float(np.abs(p).max()))
        assert 0.0 < h <= 2.0 * push * mag, (h, push * mag)


# ---------------------------------------------------------------- camera

W, H = 48, 32
ASPECT = W / H
FOV = math.pi / 2  # tan(fov / 2) = 1
# Emitting rectangles in continuous pixel coordinates (column range, row
# range; row 0 at the top), and their radiance.  A touches no frame edge;
# B and C run past the bottom-right and top-left corners.
RECTS = [((5.5, 30.25), (2.75, 20.5), (1.0, 0.5, 0.25)),
         ((40.5, 60.0), (26.5, 50.0), (0.25, 2.0, 0.5)),
         ((-10.0, 2.5), (-10.0, 1.5), (0.5, 0.25, 3.0))]


def camera_frame():
    return scenes.camera((0.5, -0.25, 1.0), (0.3, -0.2, -1.0), up_hint=(0, 1, 0), fov=FOV, aspect=ASPECT)


def camera_scene():
    """Emitters in the plane one unit along `forward`: the camera ray through
    image position (x, y) meets it at position + f + r*x + u*y, with
    x = (2 cx / W - 1) * aspect * tan(fov/2), y = (1 - 2 cy / H) * tan(fov/2)
    for continuous pixel coordinates (cx, cy) -- pixel (0, 0) at the top left."""
    cam = camera_frame()
    p = np.array(list(cam.position), np.float64)
    f = np.array(list(cam.forward), np.float64)
    r = np.array(list(cam.right), np.float64)
    u = np.array(list(cam.up), np.float64)
    t = math.tan(FOV / 2)
    b = wpt.SceneBuilder()
    b.black_background = True

    def at(cx, cy):
        x = (2 * cx / W - 1) * ASPECT * t
        y = (1 - 2 * cy / H) * t
        return tuple(p + f + r * x + u * y)

    for (c0, c1), (r0, r1), le in RECTS:
        m = b.material(wpt.make_material(wpt.EMISSIVE, emission=le))
        # Corners ordered so that the face normal points back at the camera.
        scenes.quad(b, at(c0, r1), at(c1, r1), at(c1, r0), at(c0, r0), m)
    return b, cam


def coverage(lo, hi, n):
    """Overlap of each unit cell [i, i+1) with [lo, hi]."""
    i = np.arange(n, dtype=np.float64)
    return np.clip(np.minimum(i + 1, hi) - np.maximum(i, lo), 0.0, 1.0)


def check_camera_image(img, spp):
    """img: (H, W, 3) mean radiance per pixel."""
    full = np.zeros((H, W), bool)
    empty = np.ones((H, W), bool)
    lit_pixels = half_pixels = 0
    fracs, covs = [], []
    for (c0, c1), (r0, r1), le in RECTS:
        cov = coverage(r0, r1, H)[:, None] * coverage(c0, c1, W)[None, :]
        le = np.asarray(le, np.float32)
        whole = cov == 1.0
        assert np.allclose(img[whole], le, rtol=0, atol=2e-6 * le.max()), "fully covered pixel is not exactly Le"
        lit_pixels += whole.sum()
        # Edge pixels half covered are strictly between 0 and Le (all 16
        # samples on one side has probability 2^-15); over all partly
        # covered pixels, the lit fraction averages to the coverage, the
        # jitter's expectation.
        half = cov == 0.5
        frac = img[half][:, 0] / le[0]
        assert half.sum() > 0 and np.all((frac > 0) & (frac < 1)), "half-covered pixel is all or nothing"
        half_pixels += half.sum()
        part = (cov > 0) & (cov < 1)
        fracs.append(img[part][:, 0] / le[0])
        covs.append(cov[part])
        full |= whole
        empty &= cov == 0.0
    assert np.array_equal(img[empty], np.zeros_like(img[empty])), "uncovered pixel received light"
    assert half_pixels > 40
    frac, cov = np.concatenate(fracs), np.concatenate(covs)
    sem = math.sqrt(0.25 / (frac.size * spp))
    assert abs(frac.mean() - cov.mean()) < 5 * sem, (frac.mean(), cov.mean(), sem)
    assert lit_pixels > 300 and full[0, 0] and full[H - 1, W - 1], "corner pixels must see B and C"


def test_camera_mapping_closed_form(oracle_mod):
    b, cam = camera_scene()
    scene = wpt.Scene(b.desc(), device=False, owner=b)
    spp = 16
    cfg = wpt.render_config(W, H, spp=spp, max_depth=1, rr_start=64, clamp=0.0)
    fb, _ = oracle_mod.Oracle(scene).render(cam, cfg)
    check_camera_image(fb / spp, spp)


def test_camera_scene_geometry(oracle_mod):
    # The emitters sit where the closed form says: the centre direction
    # (`forward`, image position (24, 16)) meets A, the corner directions
    # normalize(f -/+ r*aspect*tan(fov/2) +/- u*tan(fov/2)) meet C (top left)
    # and B (bottom right).  Materials are 0 = A, 1 = B, 2 = C.
    b, cam = camera_scene()
    scene = wpt.Scene(b.desc(), device=False, owner=b)
    oracle = oracle_mod.Oracle(scene)
    p = np.array(list(cam.position), np.float64)
    f = np.array(list(cam.forward), np.float64)
    r = np.array(list(cam.right), np.float64)
    u = np.array(list(cam.up), np.float64)
    t = math.tan(FOV / 2)
    dirs = [f, f - r * ASPECT * t + u * t, f + r * ASPECT * t - u * t, f + r * ASPECT * t * (2 * 35 / W - 1)]
    rays = np.zeros((6, len(dirs)), np.float32)
    for k, d in enumerate(dirs):
        rays[0:3, k] = p
        rays[3:6, k] = d / np.linalg.norm(d)
    hits = oracle.trace(rays)
    assert hits["material"].tolist() == [0, 2, 1, wpt.INVALID]


# ---------------------------------------------------------------- clamp

CLAMP_INTENSITY = (1000.0, 500.0, 250.0)


def clamp_scene(mirror_albedo):
    """A white Lambertian square (z = 0, normal +z) under a directional light
    shining straight down, seen only in a mirror wall (x = 2, normal -x): the
    light reaches the eye through one specular bounce, so the square's light
    sample is an indirect contribution (path depth 1)."""
    b = wpt.SceneBuilder()
    b.black_background = True
    white = b.material(wpt.make_material(wpt.DIFFUSE, (1, 1, 1)))
    mirror = b.material(wpt.make_material(wpt.METAL, (mirror_albedo,) * 3, roughness=0.0))
    scenes.quad(b, (-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0), white)
    scenes.quad(b, (2, -1, 0.05), (2, -1, 3), (2, 1, 3), (2, 1, 0.05), mirror)
    b.punctual_light(wpt.make_punctual_light(wpt.LIGHT_DIRECTIONAL, direction=(0, 0, -1), intensity=CLAMP_INTENSITY))
    cam = scenes.camera((-3, 0, 4), (5, 0, -3.2), up_hint=(0, 0, 1), fov=0.2)
    return b, cam


def clamp_render_oracle(oracle_mod, mirror_albedo, clamp, size=24, spp=4):
    b, cam = clamp_scene(mirror_albedo)
    scene = wpt.Scene(b.desc(), device=False, owner=b)
    cfg = wpt.render_config(size, size, spp=spp, max_depth=2, rr_start=64, clamp=clamp)
    fb, _ = oracle_mod.Oracle(scene).render(cam, cfg)
    return fb / spp


def clamp_expectations(mirror_albedo, clamp):
    raw = np.asarray(CLAMP_INTENSITY) / math.pi * mirror_albedo
    if clamp > 0 and raw.mean() > clamp:
        return raw * (clamp / raw.mean())
    return raw


def check_clamp_image(img, want, min_pixels=100):
    px = img.reshape(-1, 3).astype(np.float64)
    on = np.all(np.abs(px - want) <= 1e-4 * np.abs(want).max() + 1e-6, axis=1)
    off = np.all(px == 0, axis=1)
    assert on.sum() >= min_pixels, on.sum()
    # Every other pixel mixes the two (samples on and off the square).
    mixed = ~(on | off)
    frac = px[mixed] / want
    assert np.allclose(frac, frac[:, :1], rtol=1e-4, atol=1e-6) and np.all((frac > 0) & (frac < 1))


def test_clamp_known_answers(oracle_mod):
    # (1000, 500, 250)/pi has RGB mean 185.7 > 10: scaled to mean 10 with the
    # hue kept, (120/7, 60/7, 30/7) (material_transport_tests.cu:223-224).
    want = np.array([120 / 7, 60 / 7, 30 / 7])
    assert np.allclose(clamp_expectations(1.0, 10.0), want)
    check_clamp_image(clamp_render_oracle(oracle_mod, 1.0, 10.0), want)
    # The clamp acts after the throughput: a 0.001 path is raw and linear.
    check_clamp_image(clamp_render_oracle(oracle_mod, 0.001, 10.0), clamp_expectations(0.001, 10.0))
    # Clamp 0 keeps raw transport.
    check_clamp_image(clamp_render_oracle(oracle_mod, 1.0, 0.0), clamp_expectations(1.0, 0.0))


# ---------------------------------------------------------------- thin pane

PANE = 0.0015


def thin_pane_scene():
    """A 1.5 mm glass pane (ior 1.5; outward normals +z at z = 0 and -z at
    z = -0.0015) over an emitter of radiance 1 (z = -1, normal +z), seen
    head-on through a narrow camera; black background."""
    b = wpt.SceneBuilder()
    b.black_background = True
    glass = b.material(wpt.make_material(wpt.DIELECTRIC, (1, 1, 1), ior=1.5))
    light = b.material(wpt.make_material(wpt.EMISSIVE, emission=(1, 1, 1)))
    s = 2.0
    scenes.quad(b, (-s, -s, 0), (s, -s, 0), (s, s, 0), (-s, s, 0), glass)                 # normal +z
    scenes.quad(b, (-s, -s, -PANE), (-s, s, -PANE), (s, s, -PANE), (s, -s, -PANE), glass)  # normal -z
    scenes.quad(b, (-s, -s, -1), (s, -s, -1), (s, s, -1), (-s, s, -1), light)             # normal +z
    cam = scenes.camera((0, 0, 1), (0, 0, -1), up_hint=(0, 1, 0), fov=0.05)
    return b, cam


def thin_pane_expectation():
    # Normal incidence: R = ((1 - 1.5)/(1 + 1.5))^2 = 0.04 at each face; the
    # eta^2 radiance scalings of entry and exit cancel; summing the internal
    # reflections, the emitter is reached with probability (1-R)^2/(1-R^2).
    r = 0.04
    return (1 - r) / (1 + r)


def check_thin_pane(img, n_samples):
    mean = float(img.mean())
    se = math.sqrt(0.08 / n_samples)
    want = thin_pane_expectation()
    # Skipping the exit face (an offset or cutoff wider than the pane) gives
    # 1 - R = 0.96 instead: about 35 standard errors away.
    assert abs(mean - want) < 5 * se, (mean, want, se)


def test_thin_pane_exit_is_found(oracle_mod):
    b, cam = thin_pane_scene()
    scene = wpt.Scene(b.desc(), device=False, owner=b)
    oracle = oracle_mod.Oracle(scene)
    # The reference's unit case: enter at z = 0, push off the entry point,
    # and the exit face is the next hit, from behind, within the pane.
    entry = oracle.trace(np.array([[0], [0], [1], [0], [0], [-1]], np.float32))
    assert entry["t"][0] == 1.0
    origin = oracle_mod.offset_ray_origin((0, 0, 0), (0, 0, 1), (0, 0, -1), 1.0)
    exit_ = oracle.trace(np.array([[origin[0]], [origin[1]], [origin[2]], [0], [0], [-1]], np.float32))
    assert exit_["primitive"][0] in (2, 3) and 0 < exit_["t"][0] <= PANE
    spp = 64
    cfg = wpt.render_config(32, 32, spp=spp, max_depth=64, rr_start=64, clamp=0.0)
    fb, _ = oracle.render(cam, cfg)
    check_thin_pane(fb / spp, 32 * 32 * spp)


# ---------------------------------------------------------------- tilted emitter

def tilted_emitter_scene(ignore_occlusion):
    """The reference's tilted emitter (-10,-10,1), (0,10,3), (10,-10,2) over a
    white Lambertian floor: shadow rays from the floor to points on the
    emitter must not be stopped by the emitter itself, so rendering with and
    without occlusion tests gives the same image."""
    b = wpt.SceneBuilder()
    b.black_background = True
    b.ignore_direct_light_occlusion = ignore_occlusion
    white = b.material(wpt.make_material(wpt.DIFFUSE, (1, 1, 1)))
    light = b.material(wpt.make_material(wpt.EMISSIVE, emission=(1, 1, 1)))
    s = 20.0
    scenes.quad(b, (-s, -s, 0), (s, -s, 0), (s, s, 0), (-s, s, 0), white)
    b.static_triangles.append(wpt.make_triangle((-10, -10, 1), (0, 10, 3), (10, -10, 2), light))
    cam = scenes.camera((0.3, -0.2, 0.6), (0.1, 0.2, -1), up_hint=(0, 1, 0), fov=1.2)
    return b, cam


def test_tilted_emitter_is_not_self_occluded(oracle_mod):
    images = []
    for ignore in (False, True):
        b, cam = tilted_emitter_scene(ignore)
        scene = wpt.Scene(b.desc(), device=False, owner=b)
        cfg = wpt.render_config(24, 24, spp=4, max_depth=1, rr_start=64, clamp=0.0)
        fb, _ = oracle_mod.Oracle(scene).render(cam, cfg)
        images.append(fb)
    assert np.array_equal(images[0], images[1])
    assert images[0].mean() > 0.05


# ---------------------------------------------------------------- roulette

def rr_floor_scene(rho=0.6):
    return scenes.floor_under_emitter(wpt.make_material(wpt.DIFFUSE, (rho, rho, rho))), \
        scenes.camera((0, 0.5, 0), (0, -1, 0), up_hint=(0, 0, 1), fov=0.3)


def check_rr_energy(img, rho=0.6):
    # Roulette from depth 1: the continuation off the floor survives with
    # p = clamp(max throughput, .05, .95) = rho and carries throughput / p.
    # Light sampling of the 1000 x 1000 emitter is almost never preferred by
    # MIS, so nearly all of albedo x Le arrives through that continuation: a
    # missing 1/p gives about rho^2 = 0.36 instead of 0.6.
    mean = img.mean(axis=(0, 1))
    sem = img.reshape(-1, 3).std(axis=0) / math.sqrt(img.shape[0] * img.shape[1])
    assert np.all(np.abs(mean - rho) < max(0.003, 5 * sem.max())), (mean, sem)


def test_roulette_keeps_energy(oracle_mod):
    b, cam = rr_floor_scene()
    scene = wpt.Scene(b.desc(), device=False, owner=b)
    spp = 32
    cfg = wpt.render_config(32, 32, spp=spp, max_depth=64, rr_start=1, clamp=0.0)
    fb, st = oracle_mod.Oracle(scene).render(cam, cfg)
    check_rr_energy(fb / spp)
