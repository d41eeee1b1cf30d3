"""Environment-map tests on the CPU: the oracle's importance table, lookups
and sampling against the checks of the reference's tests/environment_tests.cpp
and against an independent numpy restatement; the product's host-built
tables against the oracle's, bit for bit; and whole oracle images with a map
against closed forms and independent integrals, restating
tests/environment_shading_tests.cu in image form."""
import math

import numpy as np
import pytest

import scenes
import wpt

PI_F = np.float32(3.14159265359)


def env_scene(env, intensity=1.0, rotation_degrees=0.0):
    b = wpt.SceneBuilder()
    b.material(wpt.make_material(wpt.DIFFUSE, (0.5, 0.5, 0.5)))
    b.environment(env, intensity, rotation_degrees)
    return wpt.Scene(b.desc(), device=False, owner=b)


def oracle_for(oracle_mod, env, intensity=1.0, rotation_degrees=0.0):
    scene = env_scene(env, intensity, rotation_degrees)
    return scene, oracle_mod.Oracle(scene)


def numpy_distribution(env):
    """The published construction, restated in numpy (libm sin): tent
    [1/8 3/4 1/8] luminance with u wrap and v clamp, times the texel solid
    angle; 0.99 of it plus 0.01 uniform solid angle; normalized."""
    h, w, _ = env.shape
    lum = env.astype(np.float64) @ np.array([0.2126, 0.7152, 0.0722])
    tent = np.array([0.125, 0.75, 0.125])
    padded = np.vstack([lum[:1], lum, lum[-1:]])
    padded = np.hstack([padded[:, -1:], padded, padded[:, :1]])
    avg = np.zeros_like(lum)
    for dy in range(3):
        for dx in range(3):
            avg += tent[dy] * tent[dx] * padded[dy:dy + h, dx:dx + w]
    rows = np.arange(h)
    omega = (4 * np.pi / w) * np.sin(np.pi * (rows + 0.5) / h) * np.sin(np.pi / (2 * h))
    weights = avg * omega[:, None]
    uniform = np.broadcast_to(omega[:, None] / (4 * np.pi), (h, w))
    total = weights.sum()
    p = 0.99 * weights / total + 0.01 * uniform if total > 0 else uniform
    return (p / p.sum()).reshape(-1)


def alias_masses(prob, alias_p, alias_i):
    """Mass each index receives from a Walker table: its own column's kept
    share plus the rejected shares of the columns that alias to it."""
    n = prob.size
    mass = alias_p.astype(np.float64).copy()
    np.add.at(mass, alias_i, 1.0 - alias_p.astype(np.float64))
    return mass / n


# ---- environment_tests.cpp:testDistribution

def test_constant_map_normalization_and_uniform_pdf(oracle_mod):
    scene, oracle = oracle_for(oracle_mod, scenes.constant_map(8.0))
    prob, _, _ = oracle.env_tables()
    solid = sum(16 * oracle_mod.Oracle.env_solid_angle(y, 16, 8) for y in range(8))
    assert abs(solid - 4 * math.pi) < 1e-5
    assert abs(prob.astype(np.float64).sum() - 1.0) < 1e-6
    dirs = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1)], np.float32)
    _, pdf, _, _ = oracle.env_eval(dirs)
    assert np.all(np.abs(pdf - 1 / (4 * PI_F)) < 1e-5)


def test_panorama_conventions(oracle_mod):
    _, oracle = oracle_for(oracle_mod, scenes.constant_map(1.0))
    uv, _, _, _ = oracle.env_eval(np.array([(1, 0, 0), (0, 0, 1), (0, 1, 0), (0, -1, 0)], np.float32))
    assert abs(uv[0, 0] - 0.5) < 1e-5       # +X at the center column
    assert abs(uv[1, 0] - 0.75) < 1e-5      # +Z at three quarters
    assert uv[2, 1] == 0.0 and uv[3, 1] == 1.0  # +Y on the top row
    _, rotated = oracle_for(oracle_mod, scenes.constant_map(1.0), rotation_degrees=90.0)
    uv, _, _, _ = rotated.env_eval(np.array([(0, 0, -1)], np.float32))
    assert abs(uv[0, 0] - 0.5) < 1e-5       # yaw turns the map about +Y


def test_black_map_falls_back_to_uniform(oracle_mod):
    _, oracle = oracle_for(oracle_mod, scenes.constant_map(0.0))
    _, pdf, _, _ = oracle.env_eval(np.array([(0, 1, 0)], np.float32))
    assert abs(pdf[0] - 1 / (4 * PI_F)) < 1e-5


def draws(n, seed):
    r = np.random.default_rng(seed).random((n, 4), dtype=np.float32)
    return r[:, 0].astype(np.float64), r[:, 1:]


def test_constant_map_samples_agree_with_pdf_and_integrate_irradiance(oracle_mod):
    _, oracle = oracle_for(oracle_mod, scenes.constant_map(8.0))
    cand, rest = draws(200_000, 41)
    dirs, pdf = oracle.env_sample(cand, rest)
    _, pdf_eval, _, _ = oracle.env_eval(dirs)
    assert np.all(np.abs(pdf - pdf_eval) < 1e-5)
    irradiance = np.mean(np.maximum(0, dirs[:, 1].astype(np.float64)) / pdf)
    assert abs(irradiance - math.pi) < 0.025
    assert np.allclose(np.linalg.norm(dirs.astype(np.float64), axis=1), 1.0, atol=1e-6)


def test_bright_cell_frequency_shoulders_and_support(oracle_mod):
    env = scenes.constant_map(0.0)
    env[0, 0] = 10000.0
    _, oracle = oracle_for(oracle_mod, env, rotation_degrees=math.degrees(0.37))
    prob, _, _ = oracle.env_tables()
    cand, rest = draws(20_000, 7)
    dirs, pdf = oracle.env_sample(cand, rest)
    _, pdf_eval, _, texel = oracle.env_eval(dirs)
    assert np.all(np.abs(pdf - pdf_eval) < 1e-3)     # rotated, near the pole
    assert abs(np.mean(texel == 0) - prob[0]) < 0.015
    w = 16
    assert prob[-1] > 0                               # black texels keep support
    # The bilinear footprint's shoulders: longitude wrap, next column, next row.
    assert prob[w - 1] > 0.01 and prob[1] > 0.01 and prob[w] > 0.01


@pytest.mark.parametrize("bad", ["negative", "infinite", "empty", "nan"])
def test_invalid_maps_are_rejected(bad):
    env = scenes.constant_map(1.0)
    if bad == "negative":
        env[0, 0, 0] = -1.0
    elif bad == "infinite":
        env[0, 0, 0] = np.inf
    elif bad == "nan":
        env[3, 2, 1] = np.nan
    else:
        env = np.zeros((0, 4, 3), np.float32)
    with pytest.raises(wpt.WptError):
        env_scene(env)


def test_invalid_intensity_and_rotation_are_rejected():
    with pytest.raises(wpt.WptError):
        env_scene(scenes.constant_map(1.0), intensity=-1.0)
    with pytest.raises(wpt.WptError):
        env_scene(scenes.constant_map(1.0), rotation_degrees=float("inf"))


# ---- the tables: oracle vs numpy restatement vs product

def random_maps():
    rng = np.random.default_rng(5)
    yield "constant", scenes.constant_map(4.0)
    yield "sun", scenes.sun_map(width=64, height=32, x=21, y=8)
    yield "random", (rng.random((8, 16, 3)) ** 4 * 100).astype(np.float32)
    yield "one column", rng.random((5, 1, 3)).astype(np.float32)
    yield "one row", rng.random((1, 7, 3)).astype(np.float32)
    yield "tall", rng.random((33, 3, 3)).astype(np.float32)


@pytest.mark.parametrize("name,env", list(random_maps()))
def test_oracle_table_matches_numpy_restatement(oracle_mod, name, env):
    _, oracle = oracle_for(oracle_mod, env)
    prob, alias_p, alias_i = oracle.env_tables()
    want = numpy_distribution(env)
    assert np.allclose(prob, want, rtol=2e-6, atol=1e-12), name
    assert np.all(alias_i < prob.size)
    assert np.all((alias_p > 0) & (alias_p <= 1))
    assert np.allclose(alias_masses(prob, alias_p, alias_i), prob, rtol=1e-5, atol=1e-9), name


@pytest.mark.parametrize("name,env", list(random_maps()))
def test_product_tables_equal_oracle(oracle_mod, name, env):
    scene, oracle = oracle_for(oracle_mod, env, 2.0, 33.0)
    got = scene.environment_tables()
    want = oracle.env_tables()
    for g, w in zip(got, want):
        assert np.array_equal(g, w), name


# ---- radiance lookups (environment_shading_tests.cu:260-275)

def test_texel_center_lookup_keeps_hdr_values_and_rotation(oracle_mod):
    env = np.zeros((2, 4, 3), np.float32)
    env[0, 0] = (16, 8, 4)
    theta, phi = PI_F * np.float32(0.25), -PI_F * np.float32(0.75)
    d = np.array([[np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)]], np.float32)
    _, oracle = oracle_for(oracle_mod, env, intensity=2.0)
    _, _, rgb, _ = oracle.env_eval(d)
    assert abs(rgb[0, 0] - 32) < 0.01 and abs(rgb[0, 1] - 16) < 0.01
    _, turned = oracle_for(oracle_mod, env, intensity=2.0, rotation_degrees=90.0)
    c, s = np.cos(np.pi / 2), np.sin(np.pi / 2)
    r = np.array([[c * d[0, 0] + s * d[0, 2], d[0, 1], -s * d[0, 0] + c * d[0, 2]]], np.float32)
    _, _, rgb2, _ = turned.env_eval(r)
    assert np.abs(rgb2 - rgb).max() < 0.01
    _, dark = oracle_for(oracle_mod, env, intensity=0.0)
    _, _, rgb3, _ = dark.env_eval(d)
    assert not rgb3.any()


def test_bilinear_lookup_between_texel_centers(oracle_mod):
    # Halfway between two texel centers of one row: the mean of the two.
    env = np.zeros((4, 8, 3), np.float32)
    env[1, 2] = (8, 4, 2)
    env[1, 3] = (0, 4, 6)
    u = np.float64(3.0) / 8        # between column centers 2.5/8 and 3.5/8
    v = np.float64(1.5) / 4        # row 1 center
    phi = (u - 0.5) * 2 * np.pi
    theta = v * np.pi
    d = np.array([[np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)]], np.float32)
    _, oracle = oracle_for(oracle_mod, env)
    _, _, rgb, _ = oracle.env_eval(d)
    assert np.allclose(rgb[0], (4, 4, 4), atol=2e-3)


# ---- whole images (environment_shading_tests.cu in image form)

def render(oracle_mod, builder, cam, size, spp, max_depth, rr_start=64, clamp=0.0, precision=32):
    scene = wpt.Scene(builder.desc(), device=False, owner=builder)
    oracle = oracle_mod.Oracle(scene)
    cfg = wpt.render_config(size, size, spp=spp, max_depth=max_depth, rr_start=rr_start, clamp=clamp)
    fb, _ = oracle.render(cam, cfg, precision=precision)
    return fb / spp


def mean_and_sem(img):
    flat = img.reshape(-1, 3).astype(np.float64)
    return flat.mean(axis=0), flat.std(axis=0) / math.sqrt(flat.shape[0])


@pytest.mark.parametrize("precision", [32, 64])
def test_constant_map_diffuse_energy(oracle_mod, precision):
    # A Lambertian floor under a constant map (4 x 2.5 = 10) reflects
    # albedo x 10: environment NEE, the mixture continuation and the miss's
    # MIS weight must add to it; a missing or doubled strategy fails.
    albedo = (0.2, 0.4, 0.6)
    b = scenes.env_floor(wpt.make_material(wpt.DIFFUSE, albedo), scenes.constant_map(4.0), intensity=2.5)
    img = render(oracle_mod, b, scenes.look_down(), 64, 16, 4, precision=precision)
    mean, sem = mean_and_sem(img)
    want = np.array(albedo) * 10
    assert np.all(np.abs(mean - want) < np.maximum(0.01 * want, 5 * sem)), (mean, sem)


@pytest.mark.parametrize("precision", [32, 64])
def test_glass_terminal_split_is_exact(oracle_mod, precision):
    # A glass floor under a constant map: both Fresnel branches escape and
    # are integrated at the interface, so every pixel is albedo x 10 x
    # (R + (1 - R) eta^2) whatever the Fresnel coin, and the path ends.
    albedo = (0.2, 0.4, 0.6)
    b = scenes.env_floor(wpt.make_material(wpt.DIELECTRIC, albedo, ior=1.5), scenes.constant_map(4.0), intensity=2.5)
    img = render(oracle_mod, b, scenes.look_down(fov=0.02), 16, 4, 12, precision=precision)
    want = np.array(albedo) * 10 * (0.04 + 0.96 * 4 / 9)
    assert np.abs(img.reshape(-1, 3) - want).max() < 2e-4 * want.max()
    capped = render(oracle_mod, b, scenes.look_down(fov=0.02), 16, 4, 1, precision=precision)
    assert not capped.any()  # the split respects the depth cap


@pytest.mark.parametrize("kind", [wpt.DIFFUSE, wpt.DIELECTRIC])
def test_enclosed_floor_sees_no_environment(oracle_mod, kind):
    # A floor inside a black sphere: its environment NEE and its glass
    # escape tests are occluded, and every continuation ends on the black
    # sphere (whose outward normals let it see the map, with zero albedo).
    b = scenes.env_floor(wpt.make_material(kind, (1, 1, 1), ior=1.5), scenes.constant_map(4.0), half=2.0)
    b.material(wpt.make_material(wpt.DIFFUSE, (0, 0, 0)))
    b.sphere((0, 0, 0), 10.0, 1)
    img = render(oracle_mod, b, scenes.look_down(fov=0.6), 16, 4, 8)
    assert not img.any()


def test_first_bounce_escape_is_not_clamped(oracle_mod):
    # A mirror floor: the reflected camera ray reaches the map at depth 1,
    # which counts as direct light and bypasses the indirect clamp.
    b = scenes.env_floor(wpt.make_material(wpt.METAL, (0.5, 0.5, 0.5), roughness=0.0), scenes.constant_map(50000.0))
    img = render(oracle_mod, b, scenes.look_down(), 8, 2, 4, clamp=10.0)
    assert np.abs(img.reshape(-1, 3) - 25000.0).max() < 0.1


def integrated_sun(x, y, width, height, radiance, slab=False, metal=False, sub=64):
    """Independent integral of the sun texel's bilinear footprint (3 x 3
    cells, sub x sub points each) times cos, the BSDF and, under a glass
    slab, the slab's transmission with exits at depths 2..10."""
    i = (np.arange(sub) + 0.5) / sub
    total = 0.0
    for row in range(y - 1, y + 2):
        j = np.arange(sub)
        v = row + (j + 0.5) / sub
        omega_row = (2 * np.pi / (width * sub)) * (np.cos(np.pi * (row + j / sub) / height) -
                                                   np.cos(np.pi * (row + (j + 1) / sub) / height))
        cosine = np.cos(np.pi * v / height)
        trans = np.ones_like(cosine)
        if slab:
            cos_t = np.sqrt(1 - (1 - cosine ** 2) / 2.25)
            rs = (cosine - 1.5 * cos_t) / (cosine + 1.5 * cos_t)
            rp = (1.5 * cosine - cos_t) / (1.5 * cosine + cos_t)
            f = (rs * rs + rp * rp) / 2
            trans = (1 - f) / (1 + f) * (1 - f ** 10)
        bsdf = np.full_like(cosine, 1 / np.pi)
        if metal:
            nh = np.sqrt((1 + cosine) / 2)
            a2 = 0.0625
            den = 1 - nh * nh + a2 * nh * nh
            d = a2 / (np.pi * den * den)
            g = 2 * cosine / (cosine + np.sqrt(a2 + (1 - a2) * cosine * cosine))
            fr = 0.8 + 0.2 * (1 - nh) ** 5
            bsdf = fr * d * g / (4 * cosine)
        for col in range(x - 1, x + 2):
            u = col + i
            tent = np.maximum(0, 1 - np.abs(u[None, :] - 0.5 - x)) * np.maximum(0, 1 - np.abs(v[:, None] - 0.5 - y))
            total += np.sum(radiance * tent * np.maximum(0, cosine)[:, None] * trans[:, None] *
                            omega_row[:, None] * bsdf[:, None])
    return total


@pytest.mark.parametrize("case", ["direct", "metal", "slab"])
def test_hdr_sun_energy_matches_independent_integral(oracle_mod, case):
    sun = scenes.sun_map()
    if case == "slab":
        b = scenes.env_slab(wpt.make_material(wpt.DIFFUSE, (0.2, 0.4, 0.6)), sun)
        max_depth, want = 12, integrated_sun(170, 64, 512, 256, 50000, slab=True) * 0.2
    elif case == "metal":
        b = scenes.env_floor(wpt.make_material(wpt.METAL, (0.8, 0.8, 0.8), roughness=0.5), sun)
        max_depth, want = 2, integrated_sun(170, 64, 512, 256, 50000, metal=True)
    else:
        b = scenes.env_floor(wpt.make_material(wpt.DIFFUSE, (0.2, 0.4, 0.6)), sun)
        max_depth, want = 2, integrated_sun(170, 64, 512, 256, 50000) * 0.2
    img = render(oracle_mod, b, scenes.look_down(fov=0.02), 64, 16, max_depth)
    mean, sem = mean_and_sem(img)
    assert abs(mean[0] - want) < 0.03 * want, (case, mean[0], want, sem[0])
    if case == "slab":
        assert sem[0] < 0.01 * want  # the traced continuation finds the sun
