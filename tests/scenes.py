"""Scene cases of the reference's acceleration test
(tests/scene_acceleration_tests.cu:253-354), built through the C ABI's scene
description.  The reference's alpha-masked material uses a 1x1 texture with
alpha 0; the untextured equivalent is opacity 0 below the 0.5 cutoff."""
import copy

import wpt


def triangle_at(z, material):
    t = wpt.make_triangle((-2, -1, z), (2, -1, z), (0, 2, z), material,
                          uvs=((0, 0), (1, 0), (0.5, 1)), tangent=(1, 0, 0), bitangent=(0, 1, 0))
    return t


def base_scene():
    s = wpt.SceneBuilder()
    for _ in range(4):
        s.material(wpt.make_material(wpt.DIFFUSE, (0.5, 0.5, 0.5)))
    s.materials[3].alpha_masked = 1
    s.materials[3].opacity = 0.0
    return s


def clone(s):
    c = wpt.SceneBuilder()
    c.spheres = list(copy.deepcopy(s.spheres))
    c.static_triangles = list(copy.deepcopy(s.static_triangles))
    c.flat_triangles = list(copy.deepcopy(s.flat_triangles))
    c.mesh_triangles = list(copy.deepcopy(s.mesh_triangles))
    c.meshes = list(copy.deepcopy(s.meshes))
    c.instances = list(copy.deepcopy(s.instances))
    c.materials = list(copy.deepcopy(s.materials))
    c.black_background = s.black_background
    return c


def reference_cases():
    cases = []
    cases.append(("empty scene", base_scene()))
    spheres = base_scene()
    spheres.sphere((0, 0, 3), 1, 0)
    spheres.sphere((-3, 0, 3), 0.5, 1, wpt.ACCEL_DIRECT)
    spheres.sphere((0, -21, 0), 20, 2, wpt.ACCEL_DIRECT)
    cases.append(("sphere partitions", clone(spheres)))
    all_direct = clone(spheres)
    for sp in all_direct.spheres:
        sp.acceleration = wpt.ACCEL_DIRECT
    cases.append(("all direct spheres", all_direct))
    all_bvh = clone(spheres)
    for sp in all_bvh.spheres:
        sp.acceleration = wpt.ACCEL_BVH
    cases.append(("all accelerated spheres", all_bvh))
    spheres.sphere((0, 0, 3), 1, 1, wpt.ACCEL_DIRECT)
    cases.append(("coincident spheres across partitions", clone(spheres)))

    field = base_scene()
    field.sphere((0, -21, 0), 20, 2, wpt.ACCEL_DIRECT)
    for x in range(12):
        for z in range(12):
            for layer in range(2):
                field.sphere((-3.3 + 0.6 * x, -0.3 + 0.6 * layer, 1.0 + 0.6 * z), 0.15, (x + z + layer) % 3)
    cases.append(("large sphere hierarchy", field))

    mixed = clone(spheres)
    mesh = mixed.mesh([triangle_at(1, 3), triangle_at(5, 2)])
    mixed.instance(mesh, wpt.scaled_transform((0, 0, 0), 1))
    mixed.instance(mesh, wpt.scaled_transform((4, 0, 1), 0.5), 1, wpt.ACCEL_DIRECT)
    mixed.static_triangles.append(triangle_at(0.5, 3))
    cases.append(("mixed geometry and shared mesh", clone(mixed)))
    mixed.flat_triangles.append(triangle_at(7, 2))
    cases.append(("mesh TLAS with additional flat triangles", clone(mixed)))

    direct_mesh = base_scene()
    m = direct_mesh.mesh([triangle_at(1, 3), triangle_at(5, 2)])
    direct_mesh.instance(m, wpt.scaled_transform((0, 0, 0), 1), 1, wpt.ACCEL_DIRECT)
    cases.append(("direct mesh without BLAS or TLAS", clone(direct_mesh)))
    direct_mesh.instances[0].acceleration = wpt.ACCEL_BVH
    cases.append(("accelerated mesh without spheres", clone(direct_mesh)))

    empty_mesh = base_scene()
    m = empty_mesh.mesh([])
    empty_mesh.instance(m, wpt.scaled_transform((0, 0, 0), 1))
    cases.append(("empty accelerated mesh", clone(empty_mesh)))
    empty_mesh.instances[0].acceleration = wpt.ACCEL_DIRECT
    cases.append(("empty direct mesh", clone(empty_mesh)))

    reset = base_scene()
    reset.sphere((0, 0, 3), 1, 0)
    m = reset.mesh([triangle_at(5, 2)])
    reset.instance(m, wpt.scaled_transform((0, 0, 0), 1), wpt.INVALID, wpt.ACCEL_DIRECT)
    cases.append(("sphere replaces direct mesh attributes", reset))

    flat = clone(spheres)
    flat.flat_triangles = [triangle_at(1, 3), triangle_at(5, 2)]
    cases.append(("sphere TLAS with flat triangle BVH", flat))
    return cases


def scaled_instance_scene(scale):
    """material_transport_tests.cu:13-47: a 2e-4 triangle under an anisotropic
    scaled instance must be hit at t = 2 from both sides at every scale."""
    s = wpt.SceneBuilder()
    s.material(wpt.make_material(wpt.DIFFUSE, (0.5, 0.5, 0.5)))
    small = wpt.make_triangle((-.0001, -.0001, 0), (.0001, -.0001, 0), (0, .0001, 0), 0,
                              tangent=(1, 0, 0), bitangent=(0, 1, 0))
    m = s.mesh([small])
    s.instance(m, wpt.scaled_transform(axis_scales=(scale, scale * 0.5, scale * 2)), wpt.INVALID, wpt.ACCEL_DIRECT)
    s.instance(m, wpt.scaled_transform(axis_scales=(scale, scale * 0.5, scale * 2)), wpt.INVALID, wpt.ACCEL_BVH)
    return s


def quad(builder, a, b, c, d, material):
    """Two triangles a-b-c, a-c-d; the face normal is (b - a) x (c - a)."""
    builder.static_triangles.append(wpt.make_triangle(a, b, c, material))
    builder.static_triangles.append(wpt.make_triangle(a, c, d, material))


def emissive_enclosure(le=(1.5, 0.75, 0.25)):
    """A closed box of emissive walls whose normals face inward, around the
    camera at the origin: every camera ray's radiance is exactly Le."""
    b = wpt.SceneBuilder()
    b.black_background = True
    m = b.material(wpt.make_material(wpt.EMISSIVE, emission=le))
    s = 2.0
    p = [(x, y, z) for x in (-s, s) for y in (-s, s) for z in (-s, s)]
    # Corner index = 4*ix + 2*iy + iz.
    quad(b, p[0], p[4], p[6], p[2], m)   # z = -s, normal +z
    quad(b, p[1], p[3], p[7], p[5], m)   # z = +s, normal -z
    quad(b, p[0], p[1], p[5], p[4], m)   # y = -s, normal +y
    quad(b, p[2], p[6], p[7], p[3], m)   # y = +s, normal -y
    quad(b, p[0], p[2], p[3], p[1], m)   # x = -s, normal +x
    quad(b, p[4], p[5], p[7], p[6], m)   # x = +s, normal -x
    return b


def floor_under_emitter(floor_material, le=(1.0, 1.0, 1.0), half=500.0):
    """A floor at y = 0 under a parallel emitting plane at y = 1 (normal -y),
    both 1000 x 1000, black background.  Seen from the middle, the emitter
    fills the upper hemisphere to within 4e-6 of its solid-angle weight."""
    b = wpt.SceneBuilder()
    b.black_background = True
    f = b.material(floor_material)
    e = b.material(wpt.make_material(wpt.EMISSIVE, emission=le))
    h = half
    quad(b, (-h, 0, -h), (-h, 0, h), (h, 0, h), (h, 0, -h), f)  # normal +y
    quad(b, (-h, 1, -h), (h, 1, -h), (h, 1, h), (-h, 1, h), e)  # normal -y
    return b


def camera(position, forward, up_hint=(0, 1, 0), fov=0.6, aspect=1.0):
    """Pinhole camera; fov is the vertical field of view in radians."""
    import numpy as np
    f = np.asarray(forward, np.float64)
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up_hint, np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    return wpt.Camera(wpt.vec3(position), wpt.vec3(f), wpt.vec3(r), wpt.vec3(u), fov, aspect)


# ---- environment-map scenes (tests/environment_shading_tests.cu)

def constant_map(value, width=16, height=8):
    import numpy as np
    return np.full((height, width, 3), value, np.float32)


def sun_map(x=170, y=64, width=512, height=256, radiance=50000.0):
    """Black panorama with one bright texel, the reference's HDR sun."""
    import numpy as np
    m = np.zeros((height, width, 3), np.float32)
    m[y, x] = radiance
    return m


def env_floor(material, env, intensity=1.0, rotation_degrees=0.0, half=500.0):
    """A floor at y = 0 (normal +y) lit only by the environment."""
    b = wpt.SceneBuilder()
    f = b.material(material)
    h = half
    quad(b, (-h, 0, -h), (-h, 0, h), (h, 0, h), (h, 0, -h), f)
    b.environment(env, intensity, rotation_degrees)
    return b


def env_slab(floor_material, env, half=0.1):
    """A small floor square at y = 0 under a glass slab (y = 1 to 1.01,
    outward normals, 2000 x 2000), lit by the environment through the slab."""
    b = wpt.SceneBuilder()
    f = b.material(floor_material)
    g = b.material(wpt.make_material(wpt.DIELECTRIC, (1, 1, 1), ior=1.5))
    h = half
    quad(b, (-h, 0, -h), (-h, 0, h), (h, 0, h), (h, 0, -h), f)
    s = 1000.0
    quad(b, (-s, 1.0, -s), (s, 1.0, -s), (s, 1.0, s), (-s, 1.0, s), g)      # normal -y
    quad(b, (-s, 1.01, -s), (-s, 1.01, s), (s, 1.01, s), (s, 1.01, -s), g)  # normal +y
    b.environment(env, 1.0, 0.0)
    return b


def look_down(height=0.5, fov=0.05):
    return camera((0, height, 0), (0, -1, 0), up_hint=(0, 0, 1), fov=fov)


# ---- textured and glTF materials (device_traversal.cuh:325-525, gltf_pbr.cuh)

def checker_texture(size=16, cells=4, alpha=True, seed=0):
    import numpy as np
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:size, 0:size]
    on = ((x * cells // size) + (y * cells // size)) % 2
    t = np.zeros((size, size, 4), np.uint8)
    t[..., 0] = np.where(on, 230, 40)
    t[..., 1] = np.where(on, 180, 90)
    t[..., 2] = rng.integers(0, 256, (size, size))
    t[..., 3] = np.where((x // 4 + y // 4) % 3 == 0, 30, 255) if alpha else 255
    return t


def bump_normal_map(size=16):
    import numpy as np
    y, x = np.mgrid[0:size, 0:size]
    nx = 0.4 * np.sin(2 * np.pi * x / size)
    ny = 0.3 * np.cos(2 * np.pi * y / size)
    n = np.stack([nx, ny, np.ones_like(nx)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    t = np.zeros((size, size, 4), np.uint8)
    t[..., :3] = np.clip((n * 0.5 + 0.5) * 255 + 0.5, 0, 255).astype(np.uint8)
    t[..., 3] = 255
    return t


def textured_quad(builder, a, b, c, d, material, uv_scale=1.0, tangent=(1, 0, 0), bitangent=(0, 0, 1), uvs1=False,
                  vertex_tangents=False, normal=None):
    """Quad a-b-c-d with texture coordinates (0,0),(1,0),(1,1),(0,1) x scale."""
    s = uv_scale
    uv = ((0, 0), (s, 0), (s, s), (0, s))
    n = None if normal is None else (normal, normal, normal)
    tris = []
    for (i, j, k) in ((0, 1, 2), (0, 2, 3)):
        p = (a, b, c, d)
        kw = dict(normals=n, uvs=(uv[i], uv[j], uv[k]), tangent=tangent, bitangent=bitangent)
        if uvs1:
            kw["uvs1"] = ((uv[i][1], uv[i][0]), (uv[j][1], uv[j][0]), (uv[k][1], uv[k][0]))
            kw["tangent1"], kw["bitangent1"] = bitangent, tangent
        if vertex_tangents:
            kw["vertex_tangents"] = (tangent, tangent, tangent)
            kw["tangent_signs"] = (1.0, 1.0, -1.0) if (i, j, k) == (0, 2, 3) else (1.0, 1.0, 1.0)
        tris.append(wpt.make_triangle(p[i], p[j], p[k], material, **kw))
    builder.static_triangles.extend(tris)
    return tris


def textured_room(with_env=True, with_lights=True):
    """A floor, walls and objects covering every texture slot and the glTF
    lobes: tiled albedo with alpha, masked foliage, normal maps (first and
    second texture sets, vertex tangents), emissive, metallic-roughness and
    specular-glossiness textures, thickness under glass, clearcoat and
    specular extensions, delta glTF lobes, and a textured instanced mesh."""
    import numpy as np
    b = wpt.SceneBuilder()
    checker = b.texture(checker_texture(16, 4, alpha=True, seed=1))
    opaque = b.texture(checker_texture(8, 2, alpha=False, seed=2))
    bumps = b.texture(bump_normal_map(16))
    rough = b.texture(checker_texture(4, 2, alpha=False, seed=3))
    T = wpt.texture_binding
    floor = b.material(wpt.make_material(wpt.DIFFUSE, (0.8, 0.8, 0.8)),
                       wpt.make_material_ext(albedo_texture=T(opaque, scale=(3.0, 2.0), rotation=0.3, transform=True),
                                             normal_texture=T(bumps), normal_scale=0.8))
    leaves = b.material(wpt.make_material(wpt.DIFFUSE, (0.6, 0.9, 0.5), alpha_masked=True, alpha_cutoff=0.5),
                        wpt.make_material_ext(albedo_texture=checker))
    lamp = b.material(wpt.make_material(wpt.EMISSIVE, emission=(6.0, 5.0, 4.0)),
                      wpt.make_material_ext(emissive_texture=opaque))
    metal = b.material(wpt.make_material(wpt.DIFFUSE, (0.9, 0.7, 0.5), roughness=0.6, metallic=1.0),
                       wpt.make_material_ext(metallic_roughness_texture=rough))
    specgloss = b.material(wpt.make_material(wpt.SPECULAR_GLOSSINESS, (0.5, 0.6, 0.7), specular=(0.9, 0.8, 0.7),
                                             glossiness=0.7),
                           wpt.make_material_ext(specular_glossiness_workflow=True, albedo_texture=opaque,
                                                 specular_glossiness_texture=checker))
    glass = b.material(wpt.make_material(wpt.DIELECTRIC, (0.95, 0.9, 0.85), ior=1.45,
                                         attenuation_color=(0.7, 0.8, 0.9), attenuation_distance=0.8,
                                         volume_density=1.5),
                       wpt.make_material_ext(thickness_texture=rough, albedo_texture=opaque))
    coat = b.material(wpt.make_material(wpt.DIFFUSE, (0.7, 0.1, 0.1), roughness=0.45, ior=1.5),
                      wpt.make_material_ext(gltf_pbr_workflow=True, clearcoat=0.8, clearcoat_roughness=0.2,
                                            specular_weight=0.7, specular_color=(1.0, 0.9, 0.8),
                                            clearcoat_normal_texture=bumps, albedo_texture=opaque,
                                            specular_color_texture=checker, specular_weight_texture=rough))
    chrome = b.material(wpt.make_material(wpt.DIFFUSE, (0.9, 0.9, 0.95), roughness=0.0, metallic=1.0, ior=1.5),
                        wpt.make_material_ext(gltf_pbr_workflow=True, clearcoat=1.0, clearcoat_roughness=0.0))
    paint = b.material(wpt.make_material(wpt.DIFFUSE, (0.2, 0.5, 0.3), roughness=0.7, metallic=0.3, ior=1.5),
                       wpt.make_material_ext(gltf_pbr_workflow=True, metallic_roughness_texture=rough,
                                             normal_texture=T(bumps, texcoord=1), normal_scale=1.2,
                                             clearcoat_texture=checker, clearcoat=0.5, clearcoat_roughness=0.0,
                                             clearcoat_roughness_texture=rough))
    plain = b.material(wpt.make_material(wpt.DIFFUSE, (0.5, 0.5, 0.5)))
    h = 3.0
    textured_quad(b, (-h, 0, -h), (h, 0, -h), (h, 0, h), (-h, 0, h), floor, uv_scale=2.0, vertex_tangents=True)
    textured_quad(b, (-h, 0, h), (h, 0, h), (h, 3, h), (-h, 3, h), coat, tangent=(1, 0, 0), bitangent=(0, 1, 0),
                  uvs1=True)
    textured_quad(b, (-h, 0, -h), (-h, 0, h), (-h, 3, h), (-h, 3, -h), paint, tangent=(0, 0, 1), bitangent=(0, 1, 0),
                  uvs1=True, vertex_tangents=True)
    textured_quad(b, (-1.0, 0.2, 0.5), (1.0, 0.2, 0.5), (1.0, 1.6, 0.5), (-1.0, 1.6, 0.5), leaves, uv_scale=1.5)
    textured_quad(b, (-0.6, 2.9, -0.6), (0.6, 2.9, -0.6), (0.6, 2.9, 0.6), (-0.6, 2.9, 0.6), lamp)
    b.sphere((1.4, 0.6, -0.8), 0.6, metal)
    b.sphere((-1.3, 0.5, -1.0), 0.5, glass)
    b.sphere((0.3, 0.35, -1.6), 0.35, specgloss)
    b.sphere((2.2, 0.4, 1.2), 0.4, chrome)
    # A textured mesh, instanced twice with a rotation and a scale.
    tris = []
    for (p, q, r) in (((0, 0, 0), (1, 0, 0), (0, 1, 0)), ((1, 0, 0), (1, 1, 0), (0, 1, 0))):
        tris.append(wpt.make_triangle(p, q, r, paint, uvs=((p[0], p[1]), (q[0], q[1]), (r[0], r[1])),
                                      tangent=(1, 0, 0), bitangent=(0, 1, 0), uvs1=((p[1], p[0]), (q[1], q[0]), (r[1], r[0])),
                                      tangent1=(0, 1, 0), bitangent1=(1, 0, 0),
                                      vertex_tangents=((1, 0.1, 0), (1, 0, 0.1), (0.9, 0.1, 0)),
                                      tangent_signs=(1.0, -1.0, 1.0)))
    m = b.mesh(tris)
    c, s_ = np.cos(0.6), np.sin(0.6)
    t1 = wpt.scaled_transform((0.5, 0.01, 1.8), 0.8)
    rot = wpt.Transform()
    for i in range(3):
        for j in range(3):
            rot.local_to_world[i][j] = [[c, 0, -s_], [0, 1, 0], [s_, 0, c]][i][j]
            rot.world_to_local[i][j] = [[c, 0, s_], [0, 1, 0], [-s_, 0, c]][i][j]
    rot.translation = wpt.vec3((-2.0, 0.3, 1.0))
    b.instance(m, t1)
    b.instance(m, rot, chrome)
    b.instance(m, rot, wpt.INVALID, wpt.ACCEL_DIRECT)
    if with_lights:
        b.punctual_light(wpt.make_punctual_light(wpt.LIGHT_POINT, (0.5, 2.0, 1.0), intensity=(2, 2, 2)))
        b.punctual_light(wpt.make_punctual_light(wpt.LIGHT_DIRECTIONAL, direction=(0.2, -1, 0.3),
                                                 intensity=(0.5, 0.5, 0.5)))
    if with_env:
        env = np.zeros((16, 32, 3), np.float32) + 0.3
        env[3, 10] = 60.0
        b.environment(env, 1.0, 20.0)
    else:
        b.black_background = True
    b.sphere((-2.4, 0.25, -2.4), 0.25, plain)
    return b


def room_camera(aspect=1.0):
    return camera((0.2, 1.6, 2.7), (-0.05, -0.3, -1), fov=1.1, aspect=aspect)
