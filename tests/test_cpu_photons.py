"""Caustic photon map on the CPU: the oracle's gather
(src/renderer/photon_mapping.cu:96-163) against closed forms on synthetic
photon sets, and its emission and tracing (photon_mapping.cu:4-74) against
the invariants of a single light over a mirror: every stored photon carries
the light's power share times the mirror albedo, lies on the diffuse
receiver at the mirror image of its emission ray, and arrives from above;
spectral photons carry one channel at three times the share."""
import math

import numpy as np
import pytest

import scenes
import wpt

PI_F = np.float32(3.14159265359)


def config(radius=0.06, resolution=64, box=None, **kw):
    grid = None
    if box is not None:
        grid = (resolution, box[0], box[1])
    elif resolution != 64:
        grid = (resolution, (-2.5, -1.0, -2.5), (2.5, 3.0, 2.5))
    return wpt.render_config(8, 8, caustic_radius=radius, caustic_grid=grid, **kw)


def uniform_photons(spacing, half, flux, y=0.0, direction=(0, 1, 0), jitter=0.0, seed=0):
    rng = np.random.default_rng(seed)
    xs = np.arange(-half, half + 1e-9, spacing)
    X, Z = np.meshgrid(xs, xs)
    n = X.size
    pos = np.stack([X.ravel(), np.full(n, y), Z.ravel()], 1)
    if jitter:
        pos[:, [0, 2]] += rng.uniform(-jitter, jitter, (n, 2))
    d = np.tile(np.asarray(direction, np.float64) / np.linalg.norm(direction), (n, 1))
    f = np.tile(np.asarray(flux, np.float64), (n, 1))
    return np.concatenate([pos, d, f], 1).astype(np.float32)


def test_gather_of_uniform_density_is_the_irradiance(oracle_mod):
    # Photons of flux phi on a square lattice of spacing s at normal
    # incidence carry irradiance phi / s^2; the cone-filtered estimate over
    # the 64 nearest (normalized by 3 / (pi r^2)) recovers it to the
    # lattice's discretization error.
    s, phi = 0.01, 2e-4
    ph = uniform_photons(s, 0.5, (phi, 2 * phi, 3 * phi))
    e = oracle_mod.gather_photons(config(radius=0.01), ph, (0.0013, 0.0, -0.0021), (0, 1, 0))
    want = np.array([1, 2, 3]) * phi / s ** 2
    assert np.allclose(e, want, rtol=0.05), (e, want)


def test_gather_ignores_photons_from_behind_and_beyond_the_search(oracle_mod):
    cfg = config(radius=0.02)  # search radius max(0.08, 0.1) = 0.1
    front = uniform_photons(0.01, 0.3, (1e-4, 1e-4, 1e-4))
    back = uniform_photons(0.01, 0.3, (5e-3, 5e-3, 5e-3), direction=(0, -1, 0))
    e_front = oracle_mod.gather_photons(cfg, front, (0, 0, 0), (0, 1, 0))
    e_both = oracle_mod.gather_photons(cfg, np.concatenate([front, back]), (0, 0, 0), (0, 1, 0))
    assert np.array_equal(e_front, e_both)
    far = front.copy()
    far[:, 0] += 5.0
    assert not oracle_mod.gather_photons(cfg, far, (0, 0, 0), (0, 1, 0)).any()


def test_gather_with_few_photons_uses_the_search_radius(oracle_mod):
    # Below 16 photons the gather radius is the search radius
    # max(4 r, 0.1): a single photon at distance d gives
    # phi cos (1 - d / R) 3 / (pi R^2).
    cfg = config(radius=0.05)
    R = 0.2
    d, phi = 0.05, 1.0
    ph = np.array([[d, 0, 0, 0, 1, 0, phi, phi, phi]], np.float32)
    e = oracle_mod.gather_photons(cfg, ph, (0, 0, 0), (0, 1, 0))
    assert np.allclose(e, phi * (1 - d / R) * 3 / (math.pi * R * R), rtol=1e-5)
    tilted = ph.copy()
    tilted[0, 3:6] = (0.6, 0.8, 0.0)
    assert np.allclose(oracle_mod.gather_photons(cfg, tilted, (0, 0, 0), (0, 1, 0)), e * 0.8, rtol=1e-5)


def test_gather_keeps_the_64_nearest(oracle_mod):
    # 200 photons on a line toward +x: only the 64 nearest count, the
    # gather radius is the 64th distance (above the floor), and the result
    # does not depend on which cells the others fall in.
    cfg = config(radius=0.001)
    xs = 0.0005 * np.arange(1, 201)
    ph = np.zeros((200, 9), np.float32)
    ph[:, 0] = xs
    ph[:, 4] = 1.0
    ph[:, 6:] = 1e-3
    e = oracle_mod.gather_photons(cfg, ph, (0, 0, 0), (0, 1, 0))
    near = xs[:64].astype(np.float32)
    g = near.max()
    want = np.sum(1e-3 * np.maximum(0, 1 - near / g)) * 3 / (math.pi * g * g)
    assert np.allclose(e, want, rtol=1e-4), (e, want)
    shuffled = ph[np.random.default_rng(1).permutation(200)]
    assert np.allclose(oracle_mod.gather_photons(cfg, shuffled, (0, 0, 0), (0, 1, 0)), e, rtol=1e-6)


def mirror_scene(albedo=0.8, emission=(3.0, 2.0, 1.0)):
    """A small light facing down at y = 1 over a large mirror at y = 0 and
    a diffuse ceiling at y = 2 (facing down); the light's back faces up."""
    b = wpt.SceneBuilder()
    b.black_background = True
    mirror = b.material(wpt.make_material(wpt.METAL, (albedo, albedo, albedo), roughness=0.0))
    ceiling = b.material(wpt.make_material(wpt.DIFFUSE, (0.5, 0.5, 0.5)))
    lamp = b.material(wpt.make_material(wpt.EMISSIVE, (0, 0, 0), emission=emission))
    h = 2.4
    scenes.quad(b, (-h, 0, -h), (-h, 0, h), (h, 0, h), (h, 0, -h), mirror)       # normal +y
    scenes.quad(b, (-h, 2, -h), (h, 2, -h), (h, 2, h), (-h, 2, h), ceiling)      # normal -y
    a = 0.1
    scenes.quad(b, (-a, 1, -a), (a, 1, -a), (a, 1, a), (-a, 1, a), lamp)          # normal -y
    return b


@pytest.mark.parametrize("spectral", [False, True])
def test_photons_of_a_light_over_a_mirror(oracle_mod, spectral):
    b = mirror_scene()
    scene = wpt.Scene(b.desc(), device=False, owner=b)
    oracle = oracle_mod.Oracle(scene)
    n = 4000
    cfg = config(caustic_photons=n, caustic_spectral=spectral)
    ph = oracle.photon_map(cfg)
    area = np.float32(0.2 * 0.2 / 2)  # each light triangle
    share = np.array([3.0, 2.0, 1.0], np.float32) * (area * PI_F / np.float32(n)) * np.float32(0.8)
    # Photons that leave the 4.8-wide mirror and ceiling are lost: about
    # half of the cosine lobe reaches the ceiling over the mirror.
    assert 0.3 * n < len(ph) <= n
    assert np.allclose(ph[:, 1], 2.0, atol=1e-5)             # on the ceiling
    assert np.all(ph[:, 4] < 0)                               # the way back points down
    # A photon that the mirror sends back into the lamp continues with the
    # lamp's zero albedo and lands with no flux; only a few do.
    lit = ph[:, 6:].any(axis=1)
    assert (~lit).sum() < 0.02 * len(ph)
    full = ph[lit]
    if spectral:
        nonzero = (full[:, 6:] != 0).sum(axis=1)
        assert np.all(nonzero == 1)
        ch = np.argmax(full[:, 6:], axis=1)
        assert np.allclose(full[np.arange(len(full)), 6 + ch], 3 * share[ch], rtol=1e-6)
        assert min(np.bincount(ch, minlength=3)) > 0.25 * len(full)
    else:
        assert np.allclose(full[:, 6:], share, rtol=1e-6)
    ph = full
    # Mirror image: back from the ceiling along the stored direction to the
    # mirror at y = 0, reflected, up to y = 1: the point lies on the light.
    p, d = ph[:, :3].astype(np.float64), ph[:, 3:6].astype(np.float64)
    m = p + d * (p[:, 1] / -d[:, 1])[:, None]
    r = d * np.array([1, -1, 1])
    light = m + r * (1.0 / r[:, 1])[:, None]
    assert np.all(np.abs(light[:, [0, 2]]) < 0.1 + 2e-3)


def test_caustic_gather_adds_light_only_where_photons_land(oracle_mod):
    # Rendering with the gather: diffuse vertices add albedo x irradiance,
    # so the image brightens; without stored photons nothing changes.
    host = wpt.HostScene("cornell-box")
    scene = wpt.Scene(host.desc(), device=False, owner=host)
    oracle = oracle_mod.Oracle(scene)
    cam = host.camera(32, 32)
    base = wpt.render_config(32, 32, spp=2, max_depth=6, rr_start=3)
    plain, _ = oracle.render(cam, base)
    caustic_cfg = wpt.render_config(32, 32, spp=2, max_depth=6, rr_start=3, caustic_photons=20000)
    lit, st = oracle.render(cam, caustic_cfg)
    assert st.caustic_photons_stored > 0
    assert np.all(lit >= plain - 1e-6) and lit.sum() > plain.sum()
    traced_only = wpt.render_config(32, 32, spp=2, max_depth=6, rr_start=3, caustic_photons=20000,
                                    caustic_gather=False)
    same, _ = oracle.render(cam, traced_only)
    assert np.array_equal(same, plain)
