"""Writers of small PNG and glTF files for the loader tests: a PNG encoder
(every colour type and bit depth, the five filters, Adam7, chosen zlib
levels) and a glTF scene exercising the reference loader's paths
(src/scene/gltf_loader.cpp:307-503)."""
import base64
import json
import struct
import zlib

import numpy as np

CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}
ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def pack_row(values, depth):
    values = [int(v) for v in values]
    if depth == 16:
        return b"".join(struct.pack(">H", v) for v in values)
    if depth == 8:
        return bytes(values)
    out, acc, bits = bytearray(), 0, 0
    for v in values:
        acc = (acc << depth) | v
        bits += depth
        if bits == 8:
            out.append(acc)
            acc, bits = 0, 0
    if bits:
        out.append(acc << (8 - bits))
    return bytes(out)


def paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def filter_row(kind, row, prev, bpp):
    out = bytearray(len(row))
    for i, x in enumerate(row):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i] if prev is not None else 0
        c = prev[i - bpp] if (prev is not None and i >= bpp) else 0
        pred = [0, a, b, (a + b) >> 1, paeth(a, b, c)][kind]
        out[i] = (x - pred) & 255
    return bytes([kind]) + bytes(out)


def encode_subimage(samples, depth, bpp, row_offset):
    raw, prev = b"", None
    for y in range(samples.shape[0]):
        row = pack_row(samples[y].reshape(-1), depth)
        raw += filter_row((y + row_offset) % 5, row, prev, bpp)
        prev = row
    return raw


def encode_png(samples, color, depth, palette=None, trns=None, interlace=False, level=6, idat_parts=2):
    """samples: (h, w, channels) integers below 2**depth."""
    samples = np.asarray(samples)
    h, w = samples.shape[:2]
    bpp = max(1, CHANNELS[color] * depth // 8)
    if interlace:
        raw = b""
        for n, (x0, y0, dx, dy) in enumerate(ADAM7):
            sub = samples[y0::dy, x0::dx]
            if sub.shape[0] and sub.shape[1]:
                raw += encode_subimage(sub, depth, bpp, n)
    else:
        raw = encode_subimage(samples, depth, bpp, 0)
    data = zlib.compress(raw, level)
    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color, 0, 0, 1 if interlace else 0))
    out += chunk(b"tEXt", b"Comment\x00ancillary chunks are skipped")
    if palette is not None:
        out += chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        out += chunk(b"tRNS", trns)
    step = max(1, (len(data) + idat_parts - 1) // idat_parts)
    for i in range(0, len(data), step):
        out += chunk(b"IDAT", data[i:i + step])
    return out + chunk(b"IEND", b"")


def data_uri(payload: bytes, mime="application/octet-stream") -> str:
    return f"data:{mime};base64," + base64.b64encode(payload).decode()


def checker_png(size=8, seed=0):
    rng = np.random.default_rng(seed)
    px = rng.integers(0, 256, (size, size, 4))
    return encode_png(px, 6, 8), px.astype(np.uint8)


class BufferBuilder:
    """Accumulates little-endian arrays into one buffer with 4-byte aligned
    views, returning accessor indices."""

    def __init__(self):
        self.data = bytearray()
        self.views, self.accessors = [], []

    def view(self, payload: bytes, stride=None):
        while len(self.data) % 4:
            self.data.append(0)
        v = {"buffer": 0, "byteOffset": len(self.data), "byteLength": len(payload)}
        if stride:
            v["byteStride"] = stride
        self.data += payload
        self.views.append(v)
        return len(self.views) - 1

    def accessor(self, array, ctype, kind, stride=None):
        array = np.ascontiguousarray(array)
        count = array.shape[0]
        if stride:  # interleave with padding to exercise byteStride
            elem = array.reshape(count, -1).view(np.uint8).reshape(count, -1)
            padded = np.zeros((count, stride), np.uint8)
            padded[:, :elem.shape[1]] = elem
            view = self.view(padded.tobytes(), stride)
        else:
            view = self.view(array.tobytes())
        self.accessors.append({"bufferView": view, "componentType": ctype, "count": count, "type": kind})
        return len(self.accessors) - 1


QUAD_P = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
QUAD_N = np.array([[0, 1, 0]] * 4, np.float32)
QUAD_UV = np.array([[0, 0], [2, 0], [2, 2], [0, 2]], np.float32)
QUAD_T = np.array([[1, 0, 0, 1]] * 4, np.float32)
TRI_P = np.array([[0, 0, 0], [0.5, 1.0, 0.1], [-0.4, 0.8, 0.3]], np.float32)
TRI_UV = np.array([[0.1, 0.2], [0.9, 0.3], [0.4, 0.8]], np.float32)
TRI_UV1 = np.array([[0.5, 0.5], [0.2, 0.9], [0.7, 0.1]], np.float32)


def quat(axis, degrees):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    h = np.radians(degrees) / 2
    return [*(axis * np.sin(h)), np.cos(h)]


def gltf_document(images_dir=None, external_image=None):
    """(json dict, binary buffer, expected image pixels) of the test scene.
    With images_dir the third image is an external file written there."""
    b = BufferBuilder()
    qp = b.accessor(QUAD_P, 5126, "VEC3", stride=16)
    qn = b.accessor(QUAD_N, 5126, "VEC3")
    quv = b.accessor(QUAD_UV, 5126, "VEC2")
    qt = b.accessor(QUAD_T, 5126, "VEC4")
    qi = b.accessor(np.array([0, 1, 2, 0, 2, 3], np.uint8), 5121, "SCALAR")
    tp = b.accessor(TRI_P, 5126, "VEC3")
    tuv = b.accessor(TRI_UV, 5126, "VEC2")
    tuv1 = b.accessor(TRI_UV1, 5126, "VEC2")
    ti16 = b.accessor(np.array([0, 1, 2], np.uint16), 5123, "SCALAR")
    ti32 = b.accessor(np.array([2, 1, 0, 0, 1], np.uint32), 5125, "SCALAR")  # trailing partial triangle
    png0, px0 = checker_png(8, 1)
    png1, px1 = checker_png(4, 2)
    png2, px2 = checker_png(2, 3)
    img1_view = b.view(png1)
    images = [{"uri": data_uri(png0, "image/png"), "name": "inline"}, {"bufferView": img1_view, "mimeType": "image/png"}]
    if images_dir is not None:
        (images_dir / "third image.png").write_bytes(png2)
        images.append({"uri": "third%20image.png"})
    else:
        images.append({"uri": data_uri(png2, "image/png")})
    materials = [
        {"name": "pbr", "pbrMetallicRoughness": {"baseColorFactor": [0.8, 0.6, 0.4, 0.9], "metallicFactor": 0.25,
                                                   "roughnessFactor": 0.6,
                                                   "baseColorTexture": {"index": 0, "extensions": {
                                                       "KHR_texture_transform": {"offset": [0.1, 0.2], "scale": [2, 3],
                                                                                 "rotation": 0.5, "texCoord": 1}}},
                                                   "metallicRoughnessTexture": {"index": 1}},
         "normalTexture": {"index": 2, "scale": 0.7}, "alphaMode": "MASK", "alphaCutoff": 0.3},
        {"name": "lamp", "emissiveFactor": [4.0, 3.0, 2.0], "emissiveTexture": {"index": 1, "texCoord": 1}},
        {"name": "glass", "extensions": {"KHR_materials_transmission": {"transmissionFactor": 1.0},
                                         "KHR_materials_ior": {"ior": 1.33},
                                         "KHR_materials_volume": {"attenuationColor": [0.9, 0.5, 0.2],
                                                                  "attenuationDistance": 2.5, "thicknessFactor": 0.4,
                                                                  "thicknessTexture": {"index": 0}}}},
        {"name": "specgloss", "extensions": {"KHR_materials_pbrSpecularGlossiness": {
            "diffuseFactor": [0.5, 0.4, 0.3, 0.8], "specularFactor": [0.2, 0.3, 0.4], "glossinessFactor": 0.6,
            "diffuseTexture": {"index": 2}, "specularGlossinessTexture": {"index": 0, "texCoord": 1}}}},
        {"name": "coated", "pbrMetallicRoughness": {"baseColorFactor": [0.2, 0.3, 0.9, 1.0]},
         "extensions": {"KHR_materials_clearcoat": {"clearcoatFactor": 0.8, "clearcoatRoughnessFactor": 0.1,
                                                    "clearcoatTexture": {"index": 1},
                                                    "clearcoatNormalTexture": {"index": 2, "scale": 0.5}},
                        "KHR_materials_specular": {"specularFactor": 0.7, "specularColorFactor": [1.0, 0.9, 0.8],
                                                   "specularTexture": {"index": 0}}}},
    ]
    meshes = [
        {"primitives": [{"attributes": {"POSITION": qp, "NORMAL": qn, "TEXCOORD_0": quv, "TANGENT": qt},
                         "indices": qi, "material": 0}]},
        {"primitives": [
            {"attributes": {"POSITION": tp, "TEXCOORD_0": tuv, "TEXCOORD_1": tuv1}, "indices": ti16, "material": 3},
            {"attributes": {"POSITION": tp}, "indices": ti32},
            {"attributes": {"POSITION": tp}, "indices": ti16, "mode": 1},   # lines: skipped
            {"attributes": {"POSITION": tp}, "material": 1},                # not indexed: skipped
        ]},
    ]
    nodes = [
        {"name": "root", "translation": [1.0, 2.0, 3.0], "rotation": quat([0, 1, 0], 30), "scale": [2.0, 2.0, 2.0],
         "mesh": 0, "children": [1, 2]},
        {"name": "child", "matrix": [0, 0, -1, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0.5, -0.5, 0.25, 1], "mesh": 1},
        {"name": "spot", "translation": [0, 4, 0], "rotation": quat([1, 0, 0], -90),
         "extensions": {"KHR_lights_punctual": {"light": 1}}},
        {"name": "sun", "rotation": quat([1, 0, 0], -45), "extensions": {"KHR_lights_punctual": {"light": 2}}},
        {"name": "bulb", "translation": [-1, 1, 0], "extensions": {"KHR_lights_punctual": {"light": 0}}},
        {"name": "unused", "mesh": 0},
    ]
    doc = {
        "asset": {"version": "2.0"},
        "extensionsUsed": ["KHR_lights_punctual", "KHR_texture_transform"],
        "extensions": {"KHR_lights_punctual": {"lights": [
            {"type": "point", "color": [1.0, 0.5, 0.25], "intensity": 8.0, "range": 5.0},
            {"type": "spot", "intensity": 20.0, "spot": {"innerConeAngle": 0.2, "outerConeAngle": 0.6}},
            {"type": "directional", "color": [1.0, 0.9, 0.8], "intensity": 3.0}]}},
        "scene": 1,
        "scenes": [{"nodes": [5]}, {"nodes": [0, 3, 4]}],
        "nodes": nodes,
        "meshes": meshes,
        "materials": materials,
        "textures": [{"source": 0}, {"source": 1}, {"source": 2}],
        "images": images,
        "bufferViews": b.views,
        "accessors": b.accessors,
        "buffers": [{"byteLength": len(b.data)}],
    }
    return doc, bytes(b.data), [px0, px1, px2]


def write_gltf(path, doc, binary, external_buffer=False):
    doc = json.loads(json.dumps(doc))
    if external_buffer:
        (path.parent / "scene.bin").write_bytes(binary)
        doc["buffers"][0]["uri"] = "scene.bin"
    else:
        doc["buffers"][0]["uri"] = data_uri(binary)
    path.write_text(json.dumps(doc))
    return str(path)


def write_glb(path, doc, binary):
    text = json.dumps(doc).encode()
    text += b" " * ((4 - len(text) % 4) % 4)
    binary = binary + b"\x00" * ((4 - len(binary) % 4) % 4)
    body = struct.pack("<II", len(text), 0x4E4F534A) + text + struct.pack("<II", len(binary), 0x004E4942) + binary
    path.write_bytes(struct.pack("<4sII", b"glTF", 2, 12 + len(body)) + body)
    return str(path)


def quat_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * y * y - 2 * z * z, 2 * x * y - 2 * w * z, 2 * x * z + 2 * w * y],
                     [2 * x * y + 2 * w * z, 1 - 2 * x * x - 2 * z * z, 2 * y * z - 2 * w * x],
                     [2 * x * z - 2 * w * y, 2 * y * z + 2 * w * x, 1 - 2 * x * x - 2 * y * y]])


def trs(t=(0, 0, 0), q=(0, 0, 0, 1), s=(1, 1, 1)):
    m = np.eye(4)
    m[:3, :3] = quat_matrix(q) @ np.diag(s)
    m[:3, 3] = t
    return m
