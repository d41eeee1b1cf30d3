"""File loaders of the host layer on the CPU: Radiance HDR decoding
(the reference's environment_tests.cpp:78-85 known answer, run-length
scanlines, rejections), environment maps attached through manifests, OBJ/MTL
parsing (fans, negative and partial indices, materials, generated normals)
and the OBJ presets (crystal room, deer over a ground sphere) assembled as
src/scene/scene_host.cpp:242-418 does."""
import ctypes as C
import json
import os
import struct

import numpy as np
import pytest

import wpt

REFERENCE_CRYSTALS = "/root/reference/assets/crystals_blue.obj"


def write_hdr_flat(path, width, height, texels):
    with open(path, "wb") as f:
        f.write(f"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y {height} +X {width}\n".encode())
        f.write(bytes(np.asarray(texels, np.uint8).reshape(-1)))


def rgbe_decode(t):
    t = np.asarray(t, np.float64)
    scale = np.where(t[..., 3] > 0, np.ldexp(1.0, (t[..., 3] - 136).astype(int)), 0.0)
    return (t[..., :3] * scale[..., None]).astype(np.float32)


def test_hdr_reference_known_answer(tmp_path):
    # environment_tests.cpp:78-85: a 4x2 map whose first texel decodes to
    # (16, 8, 4): values above 1 survive.
    texels = [[128, 64, 32, 133], [128, 128, 128, 129], [128, 128, 128, 128], [128, 128, 128, 130],
              [128, 128, 128, 129], [128, 128, 128, 129], [128, 128, 128, 129], [128, 128, 128, 129]]
    write_hdr_flat(tmp_path / "t.hdr", 4, 2, texels)
    img = wpt.load_hdr(tmp_path / "t.hdr")
    assert img.shape == (2, 4, 3)
    assert np.array_equal(img[0, 0], [16, 8, 4]) and np.array_equal(img[0, 1], [1, 1, 1])
    assert np.array_equal(img[0, 2], [0.5, 0.5, 0.5]) and np.array_equal(img[0, 3], [2, 2, 2])
    assert np.array_equal(img.reshape(-1, 3), rgbe_decode(texels))


def encode_rle_scanline(row):
    # New-style scanline: 2, 2, width; then each channel as runs (count >
    # 128: a repeated byte) or literals (count <= 128).
    w = row.shape[0]
    out = bytearray([2, 2, w >> 8, w & 255])
    for k in range(4):
        ch = row[:, k].tolist()
        i = 0
        while i < w:
            j = i
            while j < w and j - i < 127 and ch[j] == ch[i]:
                j += 1
            if j - i >= 3:
                out += bytes([128 + (j - i), ch[i]])
                i = j
            else:
                lit = ch[i:min(w, i + 100)]
                out += bytes([len(lit)]) + bytes(lit)
                i += len(lit)
    return bytes(out)


def test_hdr_run_length_scanlines(tmp_path):
    rng = np.random.default_rng(0)
    w, h = 37, 5
    tex = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    tex[:, 5:20] = tex[:, 5:6]  # runs
    tex[1, :, 3] = 0            # zero exponent: black
    with open(tmp_path / "r.hdr", "wb") as f:
        f.write(f"#?RGBE\nGAMMA=1\nFORMAT=32-bit_rle_rgbe\n\n-Y {h} +X {w}\n".encode())
        for y in range(h):
            f.write(encode_rle_scanline(tex[y]))
    img = wpt.load_hdr(tmp_path / "r.hdr")
    assert np.array_equal(img, rgbe_decode(tex))
    assert not img[1].any()


@pytest.mark.parametrize("content", [
    b"P3\n1 1\n255\n255 255 255\n",                              # an LDR image
    b"#?RADIANCE\n\n-Y 1 +X 1\n\x80\x80\x80\x80",               # no FORMAT line
    b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n+Y 1 +X 1\n\x80\x80\x80\x80",  # other orientation
    b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 2\n\x80\x80\x80\x80",  # truncated
])
def test_hdr_rejections(tmp_path, content):
    (tmp_path / "bad.hdr").write_bytes(content)
    with pytest.raises(wpt.WptError):
        wpt.load_hdr(tmp_path / "bad.hdr")
    with pytest.raises(wpt.WptError):
        wpt.load_hdr(tmp_path / "missing.hdr")


def test_manifest_environment_is_loaded(tmp_path):
    texels = np.full((4, 8, 4), 128, np.uint8)
    texels[..., 3] = 136
    write_hdr_flat(tmp_path / "sky.hdr", 8, 4, texels)
    (tmp_path / "s.json").write_text(json.dumps({"preset": "demo", "environment": {
        "file": "sky.hdr", "intensity": 2.5, "rotation_degrees": 90}}))
    host = wpt.HostScene(str(tmp_path / "s.json"))
    d = host.desc()
    assert (d.environment_width, d.environment_height) == (8, 4)
    assert d.environment_intensity == 2.5 and d.environment_rotation_degrees == 90
    assert d.environment_rgb[0] == 128.0
    scene = wpt.Scene(d, device=False, owner=host)
    probabilities = scene.environment_tables()[0]
    assert len(probabilities) == 32 and abs(float(np.sum(probabilities, dtype=np.float64)) - 1) < 1e-5
    (tmp_path / "m.json").write_text(json.dumps({"preset": "demo", "environment": {"file": "none.hdr"}}))
    with pytest.raises(wpt.WptError):
        wpt.HostScene(str(tmp_path / "m.json"))


OBJ = """# a quad, a pentagon and a lamp in three objects
mtllib parts.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 2 0 0
v 3 0 1
v 3 1 2
v 0 0 5
vt 0 0
vt 1 0
vt 1 1
vn 0 0 1
o quad
usemtl red
f 1/1 2/2 3/3 4/1
o pentagon
usemtl unknown
f -8//1 -7//1 -4//1 -3//1 -2//1
g lamp
usemtl glow
f 8 2 3
"""

MTL = """newmtl red
Kd 0.8 0.1 0.1
Ni 1.33
newmtl glow
Ke 4 3 2
Kd 0.2 0.2 0.2
map_Kd -s 1 1 1 lamp.jpg
"""


def write_obj(tmp_path, obj=OBJ, mtl=MTL):
    (tmp_path / "parts.obj").write_text(obj)
    (tmp_path / "parts.mtl").write_text(mtl)
    (tmp_path / "lamp.jpg").write_bytes(b"\xff\xd8\xff\xe0" + bytes(60))  # JPEG: not decoded, Kd kept
    return str(tmp_path / "parts.obj")


def deer_manifest(tmp_path, acceleration="tlas_blas_bvh2"):
    obj = write_obj(tmp_path)
    p = tmp_path / "deer.json"
    p.write_text(json.dumps({"preset": "deer", "acceleration": acceleration, "object": {
        "source": "obj", "file": os.path.basename(obj), "scale": 2.0, "translation": [0.5, -0.5, 0.0]}}))
    return str(p)


def triangles(desc, first, count):
    return [desc.mesh_triangles[first + i] for i in range(count)]


def test_obj_preset_over_ground_sphere(tmp_path):
    # assembleObjScene: ground sphere (material 0), OBJ materials from 1
    # (default first), one instance of all OBJ triangles.
    host = wpt.HostScene(deer_manifest(tmp_path))
    d = host.desc()
    assert d.sphere_count == 1 and d.spheres[0].radius == 100.0 and d.spheres[0].acceleration == wpt.ACCEL_DIRECT
    assert (d.mesh_count, d.instance_count, d.mesh_triangle_count) == (1, 1, 2 + 3 + 1)
    mats = [d.materials[i] for i in range(d.material_count)]
    assert d.material_count == 1 + 1 + 2
    assert np.allclose(list(mats[1].albedo), [0.65, 0.3, 0.18])          # OBJ default material
    assert mats[2].type == wpt.DIFFUSE and np.allclose(list(mats[2].albedo), [0.8, 0.1, 0.1])
    assert abs(mats[2].ior - 1.33) < 1e-6
    assert mats[3].type == wpt.EMISSIVE and np.allclose(list(mats[3].emission), [4, 3, 2])
    tris = triangles(d, 0, d.mesh_triangle_count)
    assert [t.material for t in tris] == [2, 2, 1, 1, 1, 3]
    # The quad is a fan from its first corner, with texture coordinates.
    assert list(tris[1].v0) == [0, 0, 0] and list(tris[1].v2) == [0, 1, 0]
    assert tris[0].flags & wpt.TRI_HAS_TEXCOORDS and not tris[2].flags & wpt.TRI_HAS_TEXCOORDS
    # File normals where given, else normalized sums of face normals.
    assert list(tris[2].n0) == [0, 0, 1]
    # Vertex 8 lies on the lamp face alone: cross((1,0,-5), (1,1,-5)) = (5,0,1).
    assert np.allclose(list(tris[5].n0), np.array([5, 0, 1]) / np.sqrt(26), atol=1e-7)
    assert all(t.flags & wpt.TRI_HAS_NORMALS for t in tris)
    t = d.instances[0].transform
    assert t.local_to_world[0][0] == 2.0 and list(t.translation) == [0.5, -0.5, 0.0]
    flat_host = wpt.HostScene(deer_manifest(tmp_path, "median_bvh"))
    flat = flat_host.desc()
    assert flat.mesh_count == 0 and flat.flat_triangle_count == 6
    assert list(flat.flat_triangles[0].v1) == [2.5, -0.5, 0.0]


def test_obj_png_textures(tmp_path):
    # obj_loader.cpp:69-104: map_Kd images decoded to RGBA8, one texture per
    # distinct file shared by the materials naming it; the JPEG keeps Kd.
    from gltf_fixtures import checker_png
    png, px = checker_png(4, 9)
    (tmp_path / "wood.png").write_bytes(png)
    mtl = MTL + "newmtl red2\nKd 0.5 0.5 0.5\nmap_Kd wood.png\nnewmtl red3\nmap_Kd wood.png\n"
    mtl = mtl.replace("newmtl red\nKd 0.8 0.1 0.1\n", "newmtl red\nKd 0.8 0.1 0.1\nmap_Kd wood.png\n")
    write_obj(tmp_path, mtl=mtl)
    p = tmp_path / "deer.json"
    p.write_text(json.dumps({"preset": "deer", "object": {"source": "obj", "file": "parts.obj"}}))
    host = wpt.HostScene(str(p))
    d = host.desc()
    assert d.texture_count == 1 and (d.textures[0].width, d.textures[0].height) == (4, 4)
    got = np.frombuffer(C.string_at(d.textures[0].rgba, 64), np.uint8).reshape(4, 4, 4)
    assert np.array_equal(got, px)
    # Materials: ground, OBJ default, red, glow, red2, red3.
    assert [d.material_ext[i].albedo_texture.texture for i in range(6)] == [wpt.INVALID, wpt.INVALID, 0,
                                                                             wpt.INVALID, 0, 0]
    (tmp_path / "wood.png").unlink()
    with pytest.raises(wpt.WptError, match="texture"):
        wpt.HostScene(str(p))


@pytest.mark.parametrize("bad", ["f 1 2 9\n", "f 1 0 2\n", "f 1/x 2 3\n"])
def test_obj_rejections(tmp_path, bad):
    path = write_obj(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 1 0\n" + bad)
    (tmp_path / "d.json").write_text(json.dumps({"preset": "deer", "object": {"file": os.path.basename(path)}}))
    with pytest.raises(wpt.WptError):
        wpt.HostScene(str(tmp_path / "d.json"))


def crystal_manifest(tmp_path, obj, dielectric=True):
    p = tmp_path / "crystal.json"
    p.write_text(json.dumps({"name": "crystal", "preset": "crystal", "object": {
        "source": "obj", "file": obj, "scale": 0.35, "translation": [0.0, -0.18, 0.0], "dielectric": dielectric},
        "room": {"neutral": True}, "lighting": {"profile": "crystal"}}))
    return str(p)


def test_crystal_room_from_obj(tmp_path):
    # scene_host.cpp:242-355: room materials 0-4, object materials from 5
    # turned into glass (ior 1.52, dispersion 0.02); the dielectric switch
    # picks the crystal light (emission 32).
    host = wpt.HostScene(crystal_manifest(tmp_path, write_obj(tmp_path)))
    d = host.desc()
    assert d.material_count == 5 + 3 and d.mesh_count == 3 and d.instance_count == 3
    for m in range(5, 8):
        mat = d.materials[m]
        assert mat.type == wpt.DIELECTRIC and abs(mat.ior - 1.52) < 1e-6 and list(mat.albedo) == [1, 1, 1]
        assert abs(d.material_ext[m].dispersion - 0.02) < 1e-7
    assert list(d.materials[4].emission) == [32, 32, 32]
    assert np.allclose(list(d.materials[1].albedo), [0.68, 0.68, 0.68])  # neutral room
    assert all(d.mesh_triangles[i].material >= 5 for i in range(d.mesh_triangle_count))
    scene = wpt.Scene(d, device=False, owner=host)
    assert scene.info().instance_count == 3


@pytest.mark.skipif(not os.path.exists(REFERENCE_CRYSTALS), reason="reference assets are not present")
def test_reference_crystal_asset(tmp_path, oracle_mod):
    host = wpt.HostScene(crystal_manifest(tmp_path, REFERENCE_CRYSTALS))
    d = host.desc()
    assert d.mesh_triangle_count == 80
    scene = wpt.Scene(d, device=False, owner=host)
    cam = host.camera(48, 32)
    status, rc, _, _ = wpt.parse_command_line(["--scene-file", crystal_manifest(tmp_path, REFERENCE_CRYSTALS),
                                               "--caustics", "--caustic-photons", "20000", "--resolution", "48x32",
                                               "--spp", "1"])
    img, st = oracle_mod.Oracle(scene).render(cam, rc)
    assert st.caustic_photons_stored > 0 and np.isfinite(img).all() and img.mean() > 0
