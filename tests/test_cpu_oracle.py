"""CPU tests that pin the oracle: RNG golden vectors and seed uniqueness,
elementary functions against libm, BSDF/Fresnel known answers from the
reference's material tests, and the oracle's traversal against an
independent FP64 brute force on the reference's acceleration-test scenes."""
import json
import os

import numpy as np
import pytest

import scenes
import wpt

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_rng_streams_match_golden(oracle_mod):
    # tests/golden/make_rng_golden.py: published PCG32 / xorshift32 /
    # SplitMix64 definitions with the reference's keys and float mapping.
    data = json.load(open(os.path.join(GOLDEN, "rng_streams.json")))
    for case in data["cases"]:
        seeded, floats = oracle_mod.rng_stream(case["pixel"], case["sample"], case["strategy"], len(case["floats"]))
        assert seeded == case["seeded_state"], case
        assert np.array_equal(floats, np.array(case["floats"], dtype=np.float32)), case


def test_rng_values_in_unit_interval(oracle_mod):
    for strategy in (0, 1):
        _, f = oracle_mod.rng_stream(77, 3, strategy, 4096)
        assert f.min() >= 0.0 and f.max() < 1.0
        assert abs(f.mean() - 0.5) < 0.02


def test_pcg_seed_uniqueness():
    # tests/rng_tests.cu:14-63: no seed collision over 64x64 pixels x 256 samples.
    pixels = np.arange(64 * 64, dtype=np.uint64)
    samples = np.arange(256, dtype=np.uint64)
    keys = (samples[:, None] << np.uint64(32)) | pixels[None, :]
    with np.errstate(over="ignore"):
        z = keys.reshape(-1) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    assert np.unique(z).size == z.size


def ulps(got, want):
    want32 = want.astype(np.float32)
    return np.abs(got.astype(np.float64) - want) / np.spacing(np.abs(want32)).astype(np.float64)


@pytest.mark.parametrize("op,fn,lo,hi", [(0, np.sin, 0.0, 6.2831855), (1, np.cos, 0.0, 6.2831855),
                                         (0, np.sin, -100.0, 100.0), (1, np.cos, -100.0, 100.0),
                                         (2, np.tan, -1.5, 1.5), (3, np.exp, -87.0, 88.0), (4, np.log, 1e-30, 1e30)])
def test_elementary_functions_correctly_rounded(oracle_mod, op, fn, lo, hi):
    x = np.random.default_rng(op * 7 + int(hi)).uniform(lo, hi, 400_000).astype(np.float32)
    if op == 4:
        x = np.exp(np.random.default_rng(9).uniform(np.log(lo), np.log(hi), 400_000)).astype(np.float32)
    got = oracle_mod.elementary(op, x)
    assert ulps(got, fn(x.astype(np.float64))).max() <= 1.0


def test_pow_correctly_rounded(oracle_mod):
    rng = np.random.default_rng(5)
    x = rng.uniform(0.0, 1.0, 200_000).astype(np.float32)
    for y in (5.0, 2.2, 2.4):
        got = oracle_mod.elementary(5, x, np.full_like(x, y))
        want = np.power(x.astype(np.float64), np.float64(np.float32(y)))
        nz = want > 1e-30
        assert ulps(got[nz], want[nz]).max() <= 1.0
    assert oracle_mod.elementary(5, np.array([0.0, 1.0], np.float32), np.array([5.0, 2.2], np.float32)).tolist() == [0.0, 1.0]


def test_fresnel_known_answers(oracle_mod):
    # tests/material_transport_tests.cu:118-128
    f = oracle_mod.dielectric_reflectance
    assert abs(f(1.0, np.float32(1 / 1.5)) - 0.04) <= 1e-6
    assert abs(f(0.767, 1.5) - 0.23811942) <= 2e-6
    assert f(0.7, 1.5) == 1.0
    assert f(0.0, 1.0) == 0.0
    cosine, eta = np.float32(0.8), np.float32(1 / 1.5)
    transmitted = np.sqrt(np.float32(1) - eta * eta * (np.float32(1) - cosine * cosine))
    assert abs(f(cosine, eta) - f(transmitted, np.float32(1) / eta)) <= 1e-6


def test_ggx_normal_incidence_closed_form(oracle_mod):
    # tests/material_transport_tests.cu:100-108: at normal incidence
    # f = F0 / (4 pi r^4), G = 1; the sampler mixes 95% GGX and 5% cosine.
    albedo = np.array([0.9, 0.8, 0.7], np.float32)
    r = np.float32(0.47)
    n = np.array([0, 0, 1], np.float32)
    value, pdf = oracle_mod.eval_specgloss(albedo, np.zeros(3), r, n, n, n)
    lobe = 1.0 / (4 * np.pi * float(r) ** 4)
    assert np.allclose(value, albedo * lobe, atol=1e-5)
    assert abs(pdf - (0.95 * lobe + 0.05 / np.pi)) <= 1e-5


@pytest.mark.parametrize("case", [c[0] for c in scenes.reference_cases()])
def test_oracle_traversal_matches_fp64_brute_force(oracle_mod, case):
    # tests/scene_acceleration_tests.cu:181-251 tolerances: material exact,
    # |dt| < 3e-4 max(1, t), |dn| < 3e-3, sphere attributes zeroed.
    builder = dict(scenes.reference_cases())[case]
    scene = wpt.Scene(builder.desc(), device=False, owner=builder)
    oracle = oracle_mod.Oracle(scene)
    rays = oracle_mod.reference_rays(12345, 2048)
    hits = oracle.trace(rays)
    t, mat, normal = oracle.brute_fp64(rays)
    expected_hit = mat != wpt.INVALID
    assert np.array_equal(hits["kind"] != wpt.HIT_NONE, expected_hit)
    h = expected_hit
    assert np.array_equal(hits["material"][h], mat[h])
    assert np.all(np.abs(hits["t"][h] - t[h]) < 3e-4 * np.maximum(1.0, t[h]))
    sphere = hits["kind"] == wpt.HIT_SPHERE
    assert np.all(hits["u"][sphere] == 0) and np.all(hits["v"][sphere] == 0)
    assert np.all(hits["light_index"][sphere] == wpt.INVALID)
    # Shadow semantics: nothing blocks before the closest hit, it blocks after.
    limits = np.where(h, t, 1000.0).astype(np.float32)
    assert not oracle.occluded(rays, limits - 0.02).any()
    assert np.array_equal(oracle.occluded(rays, limits + 0.02).astype(bool), h)


def test_brute_force_normals_are_unit_geometric(oracle_mod):
    # The brute force reports the outward geometric normal (not face-forwarded,
    # as in tests/scene_acceleration_tests.cu:216-232): unit length, and rays
    # that start inside the mesh-free room see both orientations.
    builder = dict(scenes.reference_cases())["mixed geometry and shared mesh"]
    scene = wpt.Scene(builder.desc(), device=False, owner=builder)
    oracle = oracle_mod.Oracle(scene)
    rays = oracle_mod.reference_rays(12345, 2048)
    t, mat, normal = oracle.brute_fp64(rays)
    h = mat != wpt.INVALID
    assert h.sum() > 100
    assert np.allclose(np.linalg.norm(normal[h].astype(np.float64), axis=1), 1.0, atol=1e-5)
    facing = np.einsum("ij,ij->i", normal[h].astype(np.float64), rays[3:6].T[h].astype(np.float64))
    assert (facing < 0).any() and (facing > 0).any()


def render_oracle(oracle_mod, builder, cam, size, spp, max_depth, rr_start, clamp, precision=32):
    scene = wpt.Scene(builder.desc(), device=False, owner=builder)
    oracle = oracle_mod.Oracle(scene)
    cfg = wpt.render_config(size, size, spp=spp, max_depth=max_depth, rr_start=rr_start, clamp=clamp)
    fb, _ = oracle.render(cam, cfg, precision=precision)
    return fb / spp


@pytest.mark.parametrize("precision", [32, 64])
def test_emissive_enclosure_is_exactly_le(oracle_mod, precision):
    # Camera rays end on an emitter at depth 0 with MIS weight 1
    # (device_path_shading.cuh:59-76): radiance = throughput 1 x Le.
    le = (1.5, 0.75, 0.25)
    img = render_oracle(oracle_mod, scenes.emissive_enclosure(le), scenes.camera((0.3, -0.2, 0.1), (1, 0.4, 0.2)),
                        16, 4, 8, 3, 10.0, precision)
    assert np.array_equal(img.reshape(-1, 3), np.tile(np.array(le, np.float32), (16 * 16, 1)))


@pytest.mark.parametrize("precision", [32, 64])
def test_mirror_floor_reflects_emitter_exactly(oracle_mod, precision):
    # Metal with roughness 0 (device_path_shading.cuh:115-131): throughput x
    # albedo, specular bounce, the emitter is then hit with MIS weight 1.
    albedo = (0.75, 0.5, 0.25)
    b = scenes.floor_under_emitter(wpt.make_material(wpt.METAL, albedo, roughness=0.0), le=(2.0, 2.0, 2.0))
    img = render_oracle(oracle_mod, b, scenes.camera((0, 0.5, 0), (0, -1, 0), up_hint=(0, 0, 1)), 16, 4, 8, 3, 10.0,
                        precision)
    want = np.float32(2.0) * np.array(albedo, np.float32)
    assert np.array_equal(img.reshape(-1, 3), np.tile(want, (16 * 16, 1)))


@pytest.mark.parametrize("kind,roughness", [(wpt.DIFFUSE, 0.0), (wpt.METAL, 0.3)])
def test_floor_under_uniform_emitter_energy(oracle_mod, kind, roughness):
    # A surface under a uniform emitting hemisphere reflects albedo x Le for
    # any energy-conserving BRDF whose directional albedo is the albedo;
    # for the Lambertian floor that is exact, and light sampling, BSDF
    # sampling and their MIS weights must add up to it.  A dropped MIS weight
    # doubles the result, a dropped cosine or 1/pi changes it by pi.
    rho = 0.6
    b = scenes.floor_under_emitter(wpt.make_material(kind, (rho, rho, rho), roughness=roughness))
    img = render_oracle(oracle_mod, b, scenes.camera((0, 0.5, 0), (0, -1, 0), up_hint=(0, 0, 1), fov=0.3), 32, 32, 4,
                        8, 0.0)
    mean = img.mean(axis=(0, 1))
    sem = img.reshape(-1, 3).std(axis=0) / np.sqrt(32 * 32)
    if kind == wpt.DIFFUSE:
        assert np.all(np.abs(mean - rho) < max(0.002, 5 * sem.max())), (mean, sem)
    else:
        # GGX reflectance at roughness 0.3 loses energy to masking, never gains.
        assert np.all(mean < rho + 5 * sem.max()) and np.all(mean > 0.5 * rho), (mean, sem)


# ---- differentiable estimator (no reference counterpart; pinned by calculus)

def diff_setup(oracle_mod, size=24, spp=2, texture=(12, 12), seed=1):
    host = wpt.HostScene("diff-cornell")
    scene = wpt.Scene(host.desc(), device=False, owner=host)
    oracle = oracle_mod.Oracle(scene)
    cfg = oracle.diff_config(size, size, spp, 6, texture, 5)
    t, r, e = oracle.diff_layout(cfg)
    rng = np.random.default_rng(seed)
    params = np.concatenate([rng.uniform(0.2, 0.8, t), rng.uniform(0.1, 0.9, r),
                             rng.uniform(0.5, 1.5, e)]).astype(np.float32)
    return host, scene, oracle, cfg, host.camera(size, size), params, (t, r, e)


def test_diff_layout_of_the_parameter_study_scene(oracle_mod):
    *_, (t, r, e) = diff_setup(oracle_mod, texture=(512, 512))
    assert (t, r, e) == (512 * 512 * 3, 64, 8)


def test_diff_reverse_adjoint_equals_forward_mode(oracle_mod):
    # The FP64 reverse sweep (hand-derived GGX derivative, no division) and
    # dual-number forward mode (derivative of the value expressions
    # themselves) agree along random directions.
    host, scene, oracle, cfg, cam, params, _ = diff_setup(oracle_mod)
    img = oracle.diff_render(cam, cfg, params, precision=64)
    g_img = (2.0 * (img - 0.9 * img) / img.size).astype(np.float32)
    grad = oracle.diff_adjoint64(cam, cfg, params, g_img)
    rng = np.random.default_rng(5)
    for _ in range(3):
        v = rng.normal(size=params.size).astype(np.float32)
        dual = float((g_img.astype(np.float64) * oracle.diff_directional64(cam, cfg, params, v)).sum())
        assert abs(grad @ v.astype(np.float64) - dual) <= 1e-9 * abs(dual)


def test_diff_forward_mode_matches_central_differences(oracle_mod):
    host, scene, oracle, cfg, cam, params, (t, r, e) = diff_setup(oracle_mod)
    rng = np.random.default_rng(9)
    for section in (slice(0, t), slice(t, t + r), slice(t + r, t + r + e)):
        v = np.zeros(params.size, np.float32)
        v[section] = rng.normal(size=v[section].size)
        eps = 1e-3
        plus = oracle.diff_render(cam, cfg, params + eps * v, precision=64).astype(np.float64)
        minus = oracle.diff_render(cam, cfg, params - eps * v, precision=64).astype(np.float64)
        fd = (plus - minus) / (2 * eps)
        dual = oracle.diff_directional64(cam, cfg, params, v)
        assert np.abs(fd - dual).sum() <= 2e-3 * np.abs(dual).sum()


def test_diff_image_is_linear_in_emission(oracle_mod):
    # Every path's radiance is a sum of terms that each carry exactly one
    # emission scale, so the image is homogeneous of degree 1 in the scales
    # and its derivative along the scales themselves is the image (Euler).
    host, scene, oracle, cfg, cam, params, (t, r, e) = diff_setup(oracle_mod)
    img = oracle.diff_render(cam, cfg, params, precision=64).astype(np.float64)
    v = np.zeros(params.size, np.float32)
    v[t + r:] = params[t + r:]
    dual = oracle.diff_directional64(cam, cfg, params, v)
    assert np.allclose(dual, img, rtol=1e-6, atol=1e-7)


def tie_scene():
    """Exact distance ties: a mesh holding each triangle twice (identical
    vertices, so identical t, u, v), instanced twice at the same place."""
    b = wpt.SceneBuilder()
    m = [b.material(wpt.make_material(wpt.DIFFUSE, (0.1 * (i + 1), 0.5, 0.5))) for i in range(2)]
    rng = np.random.default_rng(3)
    tris = []
    for k in range(40):
        c = rng.uniform(-1, 1, 3) * np.array([1, 0, 1])
        tris.append(wpt.make_triangle(c, c + (0.3, 0, 0), c + (0, 0, 0.3), m[1]))
    first = [wpt.make_triangle(list(t.v0), list(t.v1), list(t.v2), m[0]) for t in tris]
    mesh = b.mesh(first + tris)  # prim i and 40 + i coincide
    b.instance(mesh)
    b.instance(mesh)
    return b


def tie_rays(n=4096, seed=5):
    rng = np.random.default_rng(seed)
    o = np.concatenate([rng.uniform(-1.2, 1.2, (n, 1)), np.full((n, 1), 2.0), rng.uniform(-1.2, 1.2, (n, 1))], 1)
    d = np.tile([0.0, -1.0, 0.0], (n, 1))
    d[n // 2:, 0] = rng.uniform(-0.05, 0.05, n // 2)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.ascontiguousarray(np.concatenate([o, d], 1).T.astype(np.float32))


def test_exact_ties_take_the_lowest_instance_and_primitive(oracle_mod):
    # Whatever the tree visits first, an exact tie goes to the lower
    # instance, then the lower primitive: instance 0, the first copy.
    b = tie_scene()
    scene = wpt.Scene(b.desc(), device=False, owner=b)
    hits = oracle_mod.Oracle(scene).trace(tie_rays())
    mesh_hits = hits["kind"] == wpt.HIT_INSTANCE
    assert mesh_hits.sum() > 100
    assert np.all(hits["instance"][mesh_hits] == 0) and np.all(hits["primitive"][mesh_hits] < 40)
    assert np.all(hits["material"][mesh_hits] == 0)
