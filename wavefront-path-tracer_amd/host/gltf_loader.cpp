// glTF 2.0 (.gltf with external or base64 buffers, and binary .glb) to the
// scene description's arrays, following the reference's loadGltfScene
// (src/scene/gltf_loader.cpp:307-503, over tinygltf with deferred image
// decoding and stb_image for the pixels).
#include "loaders.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>
#include <sstream>

#include "image_decode.hpp"
#include "json.hpp"

namespace wpt {
namespace {

bool fail(std::string* error, const std::string& message) {
    if (error) *error = message;
    return false;
}

// Errors inside the nested readers travel as exceptions to loadGltfScene.
struct GltfError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

bool readFile(const std::string& path, std::vector<uint8_t>& bytes) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    bytes.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    return true;
}

std::string directoryOf(const std::string& path) {
    size_t slash = path.find_last_of('/');
    return slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
}

std::string percentDecode(const std::string& uri) {
    std::string out;
    for (size_t i = 0; i < uri.size(); ++i) {
        if (uri[i] == '%' && i + 2 < uri.size() && std::isxdigit(static_cast<unsigned char>(uri[i + 1])) &&
            std::isxdigit(static_cast<unsigned char>(uri[i + 2]))) {
            out.push_back(static_cast<char>(std::stoi(uri.substr(i + 1, 2), nullptr, 16)));
            i += 2;
        } else {
            out.push_back(uri[i]);
        }
    }
    return out;
}

bool decodeBase64(const std::string& text, size_t from, std::vector<uint8_t>& out) {
    auto value = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+') return 62;
        if (c == '/') return 63;
        return -1;
    };
    out.clear();
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = from; i < text.size(); ++i) {
        char c = text[i];
        if (c == '=') break;
        int v = value(c);
        if (v < 0) return false;
        acc = (acc << 6) | static_cast<uint32_t>(v);
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back(static_cast<uint8_t>((acc >> bits) & 0xff));
        }
    }
    return true;
}

// Bytes behind a buffer or image URI: a base64 data URI or a file beside
// the glTF file.
void loadUri(const std::string& uri, const std::string& directory, std::vector<uint8_t>& out) {
    if (uri.compare(0, 5, "data:") == 0) {
        size_t comma = uri.find(',');
        if (comma == std::string::npos || uri.rfind(";base64", comma) == std::string::npos ||
            !decodeBase64(uri, comma + 1, out))
            throw GltfError("glTF: unsupported or malformed data URI");
        return;
    }
    std::string path = directory + percentDecode(uri);
    if (!readFile(path, out)) throw GltfError("glTF: could not read " + path);
}

const Json& member(const Json& object, const std::string& key) {
    static const Json null;
    return object.contains(key) ? object.at(key) : null;
}

double numberOr(const Json& object, const std::string& key, double fallback) {
    const Json& v = member(object, key);
    return v.isNumber() ? v.number() : fallback;
}

int intOr(const Json& object, const std::string& key, int fallback) {
    const Json& v = member(object, key);
    if (!v.isNumber()) return fallback;
    double x = v.number();
    return (x >= -2147483648.0 && x <= 2147483647.0) ? static_cast<int>(x) : -1;  // out of range: invalid index
}

// A byte count, offset or element count: absent -> 0; negative, fractional
// or beyond 2^48 is an error.
size_t sizeField(const Json& object, const std::string& key) {
    const Json& v = member(object, key);
    if (!v.isNumber()) return 0;
    double x = v.number();
    if (!(x >= 0.0 && x <= 281474976710656.0) || x != static_cast<double>(static_cast<uint64_t>(x)))
        throw GltfError("glTF: bad " + key);
    return static_cast<size_t>(x);
}

std::vector<double> numbers(const Json& object, const std::string& key) {
    std::vector<double> out;
    const Json& v = member(object, key);
    if (v.isArray())
        for (const Json& item : v.items()) out.push_back(item.isNumber() ? item.number() : 0.0);
    return out;
}

struct Document {
    Json root;
    std::vector<std::vector<uint8_t>> buffers;
    const std::vector<Json>& items(const char* key) const {
        static const std::vector<Json> none;
        const Json& v = member(root, key);
        return v.isArray() ? v.items() : none;
    }
};

// Accessor view: base pointer, stride and count, checked against its
// buffer view and buffer (tinygltf Accessor::ByteStride semantics).
struct AccessorView {
    const uint8_t* base = nullptr;
    size_t stride = 0, count = 0;
    int componentType = 0, components = 0;

    const uint8_t* at(size_t index) const {
        if (index >= count) throw GltfError("glTF: accessor index out of range");
        return base + stride * index;
    }
};

int componentCount(const std::string& type) {
    if (type == "SCALAR") return 1;
    if (type == "VEC2") return 2;
    if (type == "VEC3") return 3;
    if (type == "VEC4") return 4;
    if (type == "MAT2") return 4;
    if (type == "MAT3") return 9;
    if (type == "MAT4") return 16;
    return 0;
}

int componentSize(int componentType) {
    switch (componentType) {
    case 5120: case 5121: return 1;
    case 5122: case 5123: return 2;
    case 5125: case 5126: return 4;
    default: return 0;
    }
}

AccessorView accessor(const Document& doc, int index) {
    const std::vector<Json>& accessors = doc.items("accessors");
    const std::vector<Json>& views = doc.items("bufferViews");
    if (index < 0 || static_cast<size_t>(index) >= accessors.size()) throw GltfError("glTF: bad accessor index");
    const Json& a = accessors[index];
    int viewIndex = intOr(a, "bufferView", -1);
    if (viewIndex < 0 || static_cast<size_t>(viewIndex) >= views.size())
        throw GltfError("glTF: accessor without a valid bufferView");
    const Json& view = views[viewIndex];
    int buffer = intOr(view, "buffer", -1);
    if (buffer < 0 || static_cast<size_t>(buffer) >= doc.buffers.size()) throw GltfError("glTF: bad buffer index");
    AccessorView out;
    out.componentType = intOr(a, "componentType", 0);
    out.components = componentCount(a.value("type", std::string()));
    int size = componentSize(out.componentType);
    if (out.components == 0 || size == 0) throw GltfError("glTF: bad accessor type");
    size_t elementBytes = static_cast<size_t>(out.components) * size;
    size_t byteStride = sizeField(view, "byteStride");
    out.stride = byteStride == 0 ? elementBytes : byteStride;
    if (out.stride % size != 0) throw GltfError("glTF: bad byteStride");
    out.count = sizeField(a, "count");
    size_t viewOffset = sizeField(view, "byteOffset");
    size_t viewLength = sizeField(view, "byteLength");
    size_t offset = sizeField(a, "byteOffset");
    const std::vector<uint8_t>& data = doc.buffers[buffer];
    if (viewOffset + viewLength > data.size()) throw GltfError("glTF: bufferView beyond its buffer");
    if (out.count > viewLength || out.stride > viewLength || offset > viewLength)  // bounds each term before the product below
        throw GltfError("glTF: accessor beyond its bufferView");
    if (out.count > 0 && offset + out.stride * (out.count - 1) + elementBytes > viewLength)
        throw GltfError("glTF: accessor beyond its bufferView");
    out.base = data.data() + viewOffset + offset;
    return out;
}

AccessorView floatAccessor(const Document& doc, int index, const char* what, int components) {
    AccessorView v = accessor(doc, index);
    if (v.componentType != 5126 || v.components != components)
        throw GltfError(std::string("glTF: ") + what + " must be FLOAT VEC" + std::to_string(components));
    return v;
}

void readFloats(const AccessorView& v, size_t index, float* out, int n) {
    std::memcpy(out, v.at(index), sizeof(float) * n);
}

uint32_t readIndex(const AccessorView& v, size_t index) {
    const uint8_t* p = v.at(index);
    if (v.componentType == 5121) return *p;
    if (v.componentType == 5123) {
        uint16_t x;
        std::memcpy(&x, p, 2);
        return x;
    }
    uint32_t x;
    std::memcpy(&x, p, 4);
    return x;
}

// Column-major 4x4 products and node matrices in double precision
// (gltf_loader.cpp:123-153).
using Matrix = std::array<double, 16>;

Matrix multiply(const Matrix& a, const Matrix& b) {
    Matrix r{};
    for (int column = 0; column < 4; ++column)
        for (int row = 0; row < 4; ++row)
            for (int i = 0; i < 4; ++i) r[column * 4 + row] += a[i * 4 + row] * b[column * 4 + i];
    return r;
}

Matrix nodeMatrix(const Json& node) {
    std::vector<double> m = numbers(node, "matrix");
    if (m.size() == 16) {
        Matrix out;
        std::copy(m.begin(), m.end(), out.begin());
        return out;
    }
    std::vector<double> r = numbers(node, "rotation"), s = numbers(node, "scale"), t = numbers(node, "translation");
    double x = r.size() == 4 ? r[0] : 0.0, y = r.size() == 4 ? r[1] : 0.0, z = r.size() == 4 ? r[2] : 0.0;
    double w = r.size() == 4 ? r[3] : 1.0;
    double sx = s.size() == 3 ? s[0] : 1.0, sy = s.size() == 3 ? s[1] : 1.0, sz = s.size() == 3 ? s[2] : 1.0;
    double tx = t.size() == 3 ? t[0] : 0.0, ty = t.size() == 3 ? t[1] : 0.0, tz = t.size() == 3 ? t[2] : 0.0;
    return {sx * (1.0 - 2.0 * y * y - 2.0 * z * z), sx * (2.0 * x * y + 2.0 * w * z), sx * (2.0 * x * z - 2.0 * w * y), 0.0,
        sy * (2.0 * x * y - 2.0 * w * z), sy * (1.0 - 2.0 * x * x - 2.0 * z * z), sy * (2.0 * y * z + 2.0 * w * x), 0.0,
        sz * (2.0 * x * z + 2.0 * w * y), sz * (2.0 * y * z - 2.0 * w * x), sz * (1.0 - 2.0 * x * x - 2.0 * y * y), 0.0,
        tx, ty, tz, 1.0};
}

void cross3(const float a[3], const float b[3], float out[3]) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

// The instance transform of a world matrix: the basis in FP32 and its
// inverse by cofactors over the determinant (makeTransform,
// gltf_loader.cpp:155-165).
wpt_transform makeTransform(const Matrix& m) {
    wpt_transform t{};
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k) t.local_to_world[c][k] = static_cast<float>(m[c * 4 + k]);
    const float* x = t.local_to_world[0];
    const float* y = t.local_to_world[1];
    const float* z = t.local_to_world[2];
    float yz[3], zx[3], xy[3];
    cross3(y, z, yz);
    cross3(z, x, zx);
    cross3(x, y, xy);
    float determinant = x[0] * yz[0] + x[1] * yz[1] + x[2] * yz[2];
    float inverse = 1.0f / determinant;
    float rows[3][3];
    for (int k = 0; k < 3; ++k) {
        rows[0][k] = yz[k] * inverse;
        rows[1][k] = zx[k] * inverse;
        rows[2][k] = xy[k] * inverse;
    }
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) t.world_to_local[c][r] = rows[r][c];
    for (int k = 0; k < 3; ++k) t.translation[k] = static_cast<float>(m[12 + k]);
    return t;
}

void normalize3(float v[3]) {
    float inv = 1.0f / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int k = 0; k < 3; ++k) v[k] *= inv;
}

// Face tangent frame of one texture-coordinate set
// (calculateTangentSpace, gltf_loader.cpp:199-225).
void tangentSpace(wpt_triangle& t, bool secondSet) {
    if (!(t.flags & (secondSet ? WPT_TRIANGLE_HAS_TEXCOORDS1 : WPT_TRIANGLE_HAS_TEXCOORDS))) return;
    const float* uv0 = secondSet ? t.uv1_0 : t.uv0;
    const float* uv1 = secondSet ? t.uv1_1 : t.uv1;
    const float* uv2 = secondSet ? t.uv1_2 : t.uv2;
    float e1[3], e2[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = t.v1[k] - t.v0[k];
        e2[k] = t.v2[k] - t.v0[k];
    }
    float determinant = (uv1[0] - uv0[0]) * (uv2[1] - uv0[1]) - (uv1[1] - uv0[1]) * (uv2[0] - uv0[0]);
    if (std::fabs(determinant) < 0.000001f) return;
    float inverse = 1.0f / determinant;
    float* tangent = secondSet ? t.tangent1 : t.tangent;
    float* bitangent = secondSet ? t.bitangent1 : t.bitangent;
    for (int k = 0; k < 3; ++k) {
        tangent[k] = (e1[k] * (uv2[1] - uv0[1]) - e2[k] * (uv1[1] - uv0[1])) * inverse;
        bitangent[k] = (e2[k] * (uv1[0] - uv0[0]) - e1[k] * (uv2[0] - uv0[0])) * inverse;
    }
    normalize3(tangent);
    normalize3(bitangent);
}

wpt_texture_binding noTexture() {
    wpt_texture_binding b{};
    b.texture = WPT_INVALID_INDEX;
    b.scale[0] = b.scale[1] = 1.0f;
    return b;
}

void readVec2(const Json& value, float out[2]) {
    if (!value.isArray() || value.items().size() != 2) return;
    for (int k = 0; k < 2; ++k) out[k] = static_cast<float>(value.items()[k].isNumber() ? value.items()[k].number() : 0.0);
}

// KHR_texture_transform (gltf_loader.cpp:52-61).
void applyTextureTransform(wpt_texture_binding& b, const Json& transform) {
    if (!transform.isObject()) return;
    if (transform.contains("offset")) readVec2(transform.at("offset"), b.offset);
    if (transform.contains("scale")) readVec2(transform.at("scale"), b.scale);
    if (transform.contains("rotation")) b.rotation = static_cast<float>(numberOr(transform, "rotation", 0.0));
    if (transform.contains("texCoord")) b.texcoord = static_cast<uint32_t>(intOr(transform, "texCoord", 0));
    b.has_transform = 1;
}

// A textureInfo object: texture index -> image (the texture's source),
// texCoord and an optional transform (gltf_loader.cpp:63-121).
wpt_texture_binding textureBinding(const Document& doc, const Json& info) {
    wpt_texture_binding b = noTexture();
    if (!info.isObject() || !info.contains("index")) return b;
    const std::vector<Json>& textures = doc.items("textures");
    int index = intOr(info, "index", -1);
    if (index < 0 || static_cast<size_t>(index) >= textures.size()) return b;
    int source = intOr(textures[index], "source", -1);
    if (source < 0) return b;
    b.texture = static_cast<uint32_t>(source);
    b.texcoord = static_cast<uint32_t>(intOr(info, "texCoord", 0));
    const Json& extensions = member(info, "extensions");
    if (extensions.contains("KHR_texture_transform")) applyTextureTransform(b, extensions.at("KHR_texture_transform"));
    return b;
}

const Json& extension(const Json& material, const char* name) {
    return member(member(material, "extensions"), name);
}

float extensionNumber(const Json& material, const char* name, const char* property, float fallback) {
    const Json& ext = extension(material, name);
    if (!ext.contains(property)) return fallback;
    const Json& v = ext.at(property);
    return static_cast<float>(v.isNumber() ? v.number() : 0.0);
}

void extensionColor(const Json& material, const char* name, const char* property, float out[3]) {
    const Json& ext = extension(material, name);
    if (!ext.contains(property)) return;
    const Json& v = ext.at(property);
    if (!v.isArray() || v.items().size() != 3) return;
    for (int k = 0; k < 3; ++k) out[k] = static_cast<float>(v.items()[k].isNumber() ? v.items()[k].number() : 0.0);
}

wpt_texture_binding extensionTexture(const Document& doc, const Json& material, const char* name, const char* property) {
    return textureBinding(doc, member(extension(material, name), property));
}

bool unit(float x) { return std::isfinite(x) && x >= 0.0f && x <= 1.0f; }

// One glTF material (gltf_loader.cpp:354-439).
void convertMaterial(const Document& doc, const Json& source, wpt_material& m, wpt_material_ext& e) {
    m = wpt_material{};
    e = wpt_material_ext{};
    e.albedo_texture = e.metallic_roughness_texture = e.emissive_texture = e.normal_texture = noTexture();
    e.thickness_texture = e.specular_glossiness_texture = e.specular_weight_texture = noTexture();
    e.specular_color_texture = e.clearcoat_texture = e.clearcoat_roughness_texture = noTexture();
    e.clearcoat_normal_texture = noTexture();
    const Json& pbr = member(source, "pbrMetallicRoughness");
    std::vector<double> factor = numbers(pbr, "baseColorFactor");
    if (factor.size() != 4) factor = {1.0, 1.0, 1.0, 1.0};
    std::vector<double> emissive = numbers(source, "emissiveFactor");
    if (emissive.size() != 3) emissive = {0.0, 0.0, 0.0};
    for (int k = 0; k < 3; ++k) {
        m.albedo[k] = static_cast<float>(factor[k]);
        m.emission[k] = static_cast<float>(emissive[k]);
    }
    float transmission = extensionNumber(source, "KHR_materials_transmission", "transmissionFactor", 0.0f);
    float emission2 = m.emission[0] * m.emission[0] + m.emission[1] * m.emission[1] + m.emission[2] * m.emission[2];
    m.type = emission2 > 0.0f ? WPT_MATERIAL_EMISSIVE : (transmission > 0.0f ? WPT_MATERIAL_DIELECTRIC : WPT_MATERIAL_DIFFUSE);
    m.roughness = static_cast<float>(numberOr(pbr, "roughnessFactor", 1.0));
    m.metallic = static_cast<float>(numberOr(pbr, "metallicFactor", 1.0));
    const Json& normalInfo = member(source, "normalTexture");
    e.normal_scale = static_cast<float>(numberOr(normalInfo, "scale", 1.0));
    m.alpha_cutoff = static_cast<float>(numberOr(source, "alphaCutoff", 0.5));
    m.alpha_masked = source.value("alphaMode", std::string("OPAQUE")) == "MASK" ? 1u : 0u;
    m.ior = extensionNumber(source, "KHR_materials_ior", "ior", 1.5f);
    m.attenuation_color[0] = m.attenuation_color[1] = m.attenuation_color[2] = 1.0f;
    extensionColor(source, "KHR_materials_volume", "attenuationColor", m.attenuation_color);
    m.attenuation_distance = extensionNumber(source, "KHR_materials_volume", "attenuationDistance", 0.0f);
    m.volume_density = extensionNumber(source, "KHR_materials_volume", "thicknessFactor", 1.0f);
    e.thickness_texture = extensionTexture(doc, source, "KHR_materials_volume", "thicknessTexture");
    e.albedo_texture = textureBinding(doc, member(pbr, "baseColorTexture"));
    e.metallic_roughness_texture = textureBinding(doc, member(pbr, "metallicRoughnessTexture"));
    e.emissive_texture = textureBinding(doc, member(source, "emissiveTexture"));
    e.normal_texture = textureBinding(doc, normalInfo);
    m.opacity = static_cast<float>(factor[3]);
    m.specular[0] = m.specular[1] = m.specular[2] = 1.0f;
    m.glossiness = 1.0f;
    e.gltf_pbr_workflow = m.type == WPT_MATERIAL_DIFFUSE ? 1u : 0u;
    e.specular_weight = extensionNumber(source, "KHR_materials_specular", "specularFactor", 1.0f);
    e.specular_color[0] = e.specular_color[1] = e.specular_color[2] = 1.0f;
    extensionColor(source, "KHR_materials_specular", "specularColorFactor", e.specular_color);
    e.specular_weight_texture = extensionTexture(doc, source, "KHR_materials_specular", "specularTexture");
    e.specular_color_texture = extensionTexture(doc, source, "KHR_materials_specular", "specularColorTexture");
    e.clearcoat = extensionNumber(source, "KHR_materials_clearcoat", "clearcoatFactor", 0.0f);
    e.clearcoat_roughness = extensionNumber(source, "KHR_materials_clearcoat", "clearcoatRoughnessFactor", 0.0f);
    e.clearcoat_texture = extensionTexture(doc, source, "KHR_materials_clearcoat", "clearcoatTexture");
    e.clearcoat_roughness_texture = extensionTexture(doc, source, "KHR_materials_clearcoat", "clearcoatRoughnessTexture");
    e.clearcoat_normal_texture = extensionTexture(doc, source, "KHR_materials_clearcoat", "clearcoatNormalTexture");
    e.clearcoat_normal_scale = 1.0f;
    const Json& coatNormal = member(extension(source, "KHR_materials_clearcoat"), "clearcoatNormalTexture");
    if (coatNormal.contains("scale")) e.clearcoat_normal_scale = static_cast<float>(numberOr(coatNormal, "scale", 0.0));
    auto nonnegative = [](float x) { return std::isfinite(x) && x >= 0.0f; };
    if (!unit(e.specular_weight) || !unit(e.clearcoat) || !unit(e.clearcoat_roughness) ||
        !nonnegative(e.specular_color[0]) || !nonnegative(e.specular_color[1]) || !nonnegative(e.specular_color[2]) ||
        !std::isfinite(e.clearcoat_normal_scale))
        throw GltfError("Invalid glTF specular/clearcoat factors");
    const Json& specGloss = extension(source, "KHR_materials_pbrSpecularGlossiness");
    if (specGloss.isObject()) {
        e.gltf_pbr_workflow = 0;
        e.specular_glossiness_workflow = 1;
        m.albedo[0] = m.albedo[1] = m.albedo[2] = 1.0f;
        m.opacity = 1.0f;
        if (specGloss.contains("diffuseFactor")) {
            const Json& diffuse = specGloss.at("diffuseFactor");
            if (!diffuse.isArray() || diffuse.items().size() != 4)
                throw GltfError("Invalid glTF specular-glossiness diffuseFactor");
            std::vector<double> d = numbers(specGloss, "diffuseFactor");
            for (int k = 0; k < 3; ++k) m.albedo[k] = static_cast<float>(d[k]);
            m.opacity = static_cast<float>(d[3]);
        }
        extensionColor(source, "KHR_materials_pbrSpecularGlossiness", "specularFactor", m.specular);
        m.glossiness = extensionNumber(source, "KHR_materials_pbrSpecularGlossiness", "glossinessFactor", 1.0f);
        e.albedo_texture = extensionTexture(doc, source, "KHR_materials_pbrSpecularGlossiness", "diffuseTexture");
        e.specular_glossiness_texture =
            extensionTexture(doc, source, "KHR_materials_pbrSpecularGlossiness", "specularGlossinessTexture");
        if (!unit(m.albedo[0]) || !unit(m.albedo[1]) || !unit(m.albedo[2]) || !unit(m.opacity) ||
            !unit(m.specular[0]) || !unit(m.specular[1]) || !unit(m.specular[2]) || !unit(m.glossiness))
            throw GltfError("Invalid glTF specular-glossiness factors");
        if (m.type == WPT_MATERIAL_DIFFUSE) m.type = WPT_MATERIAL_SPECULAR_GLOSSINESS;
    }
    std::string name = source.value("name", std::string());
    if (!e.gltf_pbr_workflow && e.clearcoat > 0.0f)
        std::cerr << "glTF clearcoat on non-opaque material is not implemented: " << name << '\n';
    if (m.type == WPT_MATERIAL_DIELECTRIC &&
        (e.specular_weight != 1.0f || e.specular_color[0] != 1.0f || e.specular_color[1] != 1.0f ||
         e.specular_color[2] != 1.0f || e.specular_weight_texture.texture != WPT_INVALID_INDEX ||
         e.specular_color_texture.texture != WPT_INVALID_INDEX))
        std::cerr << "glTF specular overrides on transmission are not implemented: " << name << '\n';
}

// KHR_lights_punctual light of a node at its world matrix
// (gltf_loader.cpp:230-266).
wpt_punctual_light makeLight(const Json& source, const Matrix& world) {
    wpt_punctual_light light{};
    std::string type = source.value("type", std::string());
    if (type == "directional") light.type = WPT_LIGHT_DIRECTIONAL;
    else if (type == "spot") light.type = WPT_LIGHT_SPOT;
    else if (type == "point") light.type = WPT_LIGHT_POINT;
    else throw GltfError("Unknown glTF punctual light type");
    for (int k = 0; k < 3; ++k) {
        light.position[k] = static_cast<float>(world[12 + k]);
        light.direction[k] = static_cast<float>(-world[8 + k]);
    }
    float length2 = light.direction[0] * light.direction[0] + light.direction[1] * light.direction[1] +
        light.direction[2] * light.direction[2];
    if (!(length2 > 0.0f)) throw GltfError("Degenerate glTF light transform");
    normalize3(light.direction);
    float color[3] = {1.0f, 1.0f, 1.0f};
    std::vector<double> c = numbers(source, "color");
    if (source.contains("color")) {
        if (c.size() < 3) throw GltfError("Invalid glTF light color");
        for (int k = 0; k < 3; ++k) color[k] = static_cast<float>(c[k]);
    }
    double intensity = numberOr(source, "intensity", 1.0), range = numberOr(source, "range", 0.0);
    const Json& spot = member(source, "spot");
    double inner = numberOr(spot, "innerConeAngle", 0.0), outer = numberOr(spot, "outerConeAngle", 0.7853981634);
    if (!std::isfinite(intensity) || intensity < 0.0 || !std::isfinite(range) || range < 0.0)
        throw GltfError("Invalid glTF light intensity or range");
    for (float v : color)
        if (!std::isfinite(v) || v < 0.0f) throw GltfError("Invalid glTF light color");
    if (light.type == WPT_LIGHT_SPOT && (!std::isfinite(inner) || !std::isfinite(outer) || inner < 0.0 ||
                                         outer <= inner || outer > 1.5707963267948966))
        throw GltfError("Invalid glTF spotlight cone angles");
    for (int k = 0; k < 3; ++k) light.intensity[k] = color[k] * static_cast<float>(intensity);
    light.range = static_cast<float>(range);
    light.inner_cone_cos = cosf(static_cast<float>(inner));
    light.outer_cone_cos = cosf(static_cast<float>(outer));
    return light;
}

void addNode(const Document& doc, GltfScene& scene, const std::vector<std::vector<uint32_t>>& meshAssets,
    const std::vector<Json>& nodes, const std::vector<Json>& lights, int index, const Matrix& parent, int depth) {
    if (index < 0 || static_cast<size_t>(index) >= nodes.size()) throw GltfError("glTF: bad node index");
    if (depth > 256) throw GltfError("glTF: node hierarchy too deep or cyclic");
    const Json& node = nodes[index];
    Matrix world = multiply(parent, nodeMatrix(node));
    const Json& lightRef = member(member(node, "extensions"), "KHR_lights_punctual");
    int light = intOr(lightRef, "light", -1);
    if (light >= 0) {
        if (static_cast<size_t>(light) >= lights.size())
            throw GltfError("glTF node references an invalid punctual light");
        scene.punctualLights.push_back(makeLight(lights[light], world));
    }
    int mesh = intOr(node, "mesh", -1);
    if (mesh >= 0) {
        if (static_cast<size_t>(mesh) >= meshAssets.size()) throw GltfError("glTF: bad mesh index");
        for (uint32_t asset : meshAssets[mesh]) {
            wpt_instance inst{};
            inst.mesh = asset;
            inst.material_override = WPT_INVALID_INDEX;
            inst.acceleration = WPT_ACCEL_BVH;
            inst.transform = makeTransform(world);
            scene.instances.push_back(inst);
        }
    }
    const Json& children = member(node, "children");
    if (children.isArray())
        for (const Json& child : children.items())
            addNode(doc, scene, meshAssets, nodes, lights, child.isNumber() ? static_cast<int>(child.number()) : -1,
                world, depth + 1);
}

void parseDocument(const std::string& filename, Document& doc) {
    std::vector<uint8_t> bytes;
    if (!readFile(filename, bytes)) throw GltfError("could not open glTF file " + filename);
    std::vector<uint8_t> binChunk;
    bool binary = false;
    std::string text;
    auto le32 = [&](size_t at) {
        return uint32_t(bytes[at]) | (uint32_t(bytes[at + 1]) << 8) | (uint32_t(bytes[at + 2]) << 16) |
            (uint32_t(bytes[at + 3]) << 24);
    };
    if (bytes.size() >= 12 && std::memcmp(bytes.data(), "glTF", 4) == 0) {
        // GLB: 12-byte header, a JSON chunk, then an optional BIN chunk.
        binary = true;
        if (le32(4) != 2) throw GltfError("glTF: unsupported GLB version");
        size_t total = std::min<size_t>(le32(8), bytes.size());
        size_t at = 12;
        bool sawJson = false;
        while (at + 8 <= total) {
            uint32_t length = le32(at), type = le32(at + 4);
            if (at + 8 + static_cast<size_t>(length) > total) throw GltfError("glTF: truncated GLB chunk");
            const uint8_t* body = bytes.data() + at + 8;
            if (type == 0x4E4F534Au && !sawJson) {
                text.assign(reinterpret_cast<const char*>(body), length);
                sawJson = true;
            } else if (type == 0x004E4942u && binChunk.empty()) {
                binChunk.assign(body, body + length);
            }
            at += 8 + static_cast<size_t>(length);
        }
        if (!sawJson) throw GltfError("glTF: GLB without a JSON chunk");
    } else {
        text.assign(bytes.begin(), bytes.end());
    }
    try {
        doc.root = Json::parse(text);
    } catch (const std::exception& e) {
        throw GltfError(std::string("glTF: ") + e.what());
    }
    if (!doc.root.isObject()) throw GltfError("glTF: the document is not an object");
    std::string directory = directoryOf(filename);
    const std::vector<Json>& buffers = doc.items("buffers");
    for (size_t i = 0; i < buffers.size(); ++i) {
        std::vector<uint8_t> data;
        if (buffers[i].contains("uri")) {
            loadUri(buffers[i].at("uri").str(), directory, data);
        } else if (binary && i == 0) {
            data = binChunk;
        } else {
            throw GltfError("glTF: buffer without data");
        }
        size_t byteLength = sizeField(buffers[i], "byteLength");
        if (data.size() < byteLength) throw GltfError("glTF: buffer shorter than its byteLength");
        doc.buffers.push_back(std::move(data));
    }
}

}  // namespace

bool loadGltfScene(const std::string& filename, GltfScene& out, std::string* error) {
    out = GltfScene{};
    try {
        Document doc;
        parseDocument(filename, doc);
        const Json& required = member(doc.root, "extensionsRequired");
        if (required.isArray()) {
            for (const Json& item : required.items()) {
                std::string r = item.isString() ? item.str() : std::string();
                if (r == "KHR_materials_transmission" || r == "KHR_materials_volume")
                    std::cerr << "glTF partial required extension " << r
                              << ": smooth solid dielectric only; thin/rough/partial transmission is not implemented\n";
                else if (r != "KHR_materials_specular" && r != "KHR_materials_clearcoat" && r != "KHR_materials_ior" &&
                         r != "KHR_materials_pbrSpecularGlossiness" && r != "KHR_texture_transform" &&
                         r != "KHR_lights_punctual")
                    std::cerr << "glTF unsupported required extension: " << r << '\n';
            }
        }

        // Images, in file order, decoded to RGBA8.
        std::string directory = directoryOf(filename);
        const std::vector<Json>& views = doc.items("bufferViews");
        for (const Json& image : doc.items("images")) {
            std::vector<uint8_t> encoded;
            std::string uri = image.value("uri", std::string());
            int view = intOr(image, "bufferView", -1);
            if (!uri.empty()) {
                loadUri(uri, directory, encoded);
            } else if (view >= 0 && static_cast<size_t>(view) < views.size()) {
                int buffer = intOr(views[view], "buffer", -1);
                size_t offset = sizeField(views[view], "byteOffset");
                size_t length = sizeField(views[view], "byteLength");
                if (buffer < 0 || static_cast<size_t>(buffer) >= doc.buffers.size() ||
                    offset + length > doc.buffers[buffer].size())
                    throw GltfError("glTF: bad image bufferView");
                encoded.assign(doc.buffers[buffer].begin() + offset, doc.buffers[buffer].begin() + offset + length);
            }
            std::vector<uint8_t> rgba;
            uint32_t w = 0, h = 0;
            std::string why;
            if (encoded.empty() || !decodeImage(encoded.data(), encoded.size(), rgba, w, h, &why))
                throw GltfError("Failed to load glTF image " + image.value("name", uri) + (why.empty() ? "" : ": " + why));
            out.textureData.push_back(std::move(rgba));
            out.textureSizes.push_back({w, h});
        }

        // Material 0 is the default for primitives without one.
        const std::vector<Json>& materials = doc.items("materials");
        out.materials.resize(materials.size() + 1);
        out.materialExt.resize(materials.size() + 1);
        wpt_material& fallback = out.materials[0];
        fallback = wpt_material{};
        fallback.type = WPT_MATERIAL_DIFFUSE;
        fallback.albedo[0] = fallback.albedo[1] = fallback.albedo[2] = 0.7f;
        fallback.ior = 1.0f;
        fallback.opacity = 1.0f;
        fallback.alpha_cutoff = 0.5f;
        fallback.attenuation_color[0] = fallback.attenuation_color[1] = fallback.attenuation_color[2] = 1.0f;
        fallback.volume_density = 1.0f;
        fallback.specular[0] = fallback.specular[1] = fallback.specular[2] = 1.0f;
        fallback.glossiness = 1.0f;
        wpt_material scratch;
        convertMaterial(doc, Json(), scratch, out.materialExt[0]);  // plain bindings and unit weights
        out.materialExt[0].gltf_pbr_workflow = 0;
        for (size_t i = 0; i < materials.size(); ++i) {
            convertMaterial(doc, materials[i], out.materials[i + 1], out.materialExt[i + 1]);
            out.specularGlossinessMaterials += out.materialExt[i + 1].specular_glossiness_workflow;
        }

        // One mesh asset per indexed triangle primitive.
        const std::vector<Json>& meshes = doc.items("meshes");
        std::vector<std::vector<uint32_t>> meshAssets(meshes.size());
        for (size_t meshIndex = 0; meshIndex < meshes.size(); ++meshIndex) {
            const Json& primitives = member(meshes[meshIndex], "primitives");
            if (!primitives.isArray()) continue;
            for (const Json& primitive : primitives.items()) {
                const Json& attributes = member(primitive, "attributes");
                int position = intOr(attributes, "POSITION", -1);
                int indicesIndex = intOr(primitive, "indices", -1);
                if (intOr(primitive, "mode", 4) != 4 || position < 0 || indicesIndex < 0) continue;
                AccessorView positions = floatAccessor(doc, position, "POSITION", 3);
                AccessorView indices = accessor(doc, indicesIndex);
                if ((indices.componentType != 5121 && indices.componentType != 5123 && indices.componentType != 5125) ||
                    indices.components != 1)
                    throw GltfError("glTF: indices must be unsigned integers");
                int n = intOr(attributes, "NORMAL", -1), tg = intOr(attributes, "TANGENT", -1);
                int uv = intOr(attributes, "TEXCOORD_0", -1), uv1 = intOr(attributes, "TEXCOORD_1", -1);
                AccessorView normals, tangents, texcoords, texcoords1;
                if (n >= 0) normals = floatAccessor(doc, n, "NORMAL", 3);
                if (tg >= 0) tangents = floatAccessor(doc, tg, "TANGENT", 4);
                if (uv >= 0) texcoords = floatAccessor(doc, uv, "TEXCOORD_0", 2);
                if (uv1 >= 0) texcoords1 = floatAccessor(doc, uv1, "TEXCOORD_1", 2);
                int material = intOr(primitive, "material", -1);
                if (material >= static_cast<int>(materials.size())) throw GltfError("glTF: bad material index");
                std::vector<wpt_triangle> triangles;
                triangles.reserve(indices.count / 3);
                for (size_t i = 0; i + 2 < indices.count; i += 3) {
                    uint32_t id[3] = {readIndex(indices, i), readIndex(indices, i + 1), readIndex(indices, i + 2)};
                    wpt_triangle t{};
                    readFloats(positions, id[0], t.v0, 3);
                    readFloats(positions, id[1], t.v1, 3);
                    readFloats(positions, id[2], t.v2, 3);
                    t.material = material >= 0 ? static_cast<uint32_t>(material + 1) : 0u;
                    if (n >= 0) {
                        t.flags |= WPT_TRIANGLE_HAS_NORMALS;
                        readFloats(normals, id[0], t.n0, 3);
                        readFloats(normals, id[1], t.n1, 3);
                        readFloats(normals, id[2], t.n2, 3);
                    }
                    if (tg >= 0) {
                        t.flags |= WPT_TRIANGLE_HAS_VERTEX_TANGENTS;
                        float* dst[3] = {t.vertex_tangent0, t.vertex_tangent1, t.vertex_tangent2};
                        for (int k = 0; k < 3; ++k) {
                            float v[4];
                            readFloats(tangents, id[k], v, 4);
                            std::memcpy(dst[k], v, sizeof(float) * 3);
                            t.vertex_tangent_sign[k] = v[3];
                        }
                    }
                    if (uv >= 0) {
                        t.flags |= WPT_TRIANGLE_HAS_TEXCOORDS;
                        float* dst[3] = {t.uv0, t.uv1, t.uv2};
                        for (int k = 0; k < 3; ++k) {
                            readFloats(texcoords, id[k], dst[k], 2);
                            dst[k][1] = 1.0f - dst[k][1];
                        }
                    }
                    if (uv1 >= 0) {
                        t.flags |= WPT_TRIANGLE_HAS_TEXCOORDS1;
                        float* dst[3] = {t.uv1_0, t.uv1_1, t.uv1_2};
                        for (int k = 0; k < 3; ++k) {
                            readFloats(texcoords1, id[k], dst[k], 2);
                            dst[k][1] = 1.0f - dst[k][1];
                        }
                    }
                    tangentSpace(t, false);
                    tangentSpace(t, true);
                    triangles.push_back(t);
                }
                if (!triangles.empty()) {
                    meshAssets[meshIndex].push_back(static_cast<uint32_t>(out.meshes.size()));
                    out.meshes.push_back(std::move(triangles));
                }
            }
        }

        // Instances and punctual lights from the default scene's node trees.
        const std::vector<Json>& nodes = doc.items("nodes");
        std::vector<Json> lights;
        const Json& lightsExt = member(member(doc.root, "extensions"), "KHR_lights_punctual");
        if (member(lightsExt, "lights").isArray()) lights = lightsExt.at("lights").items();
        const std::vector<Json>& scenes = doc.items("scenes");
        int sceneIndex = intOr(doc.root, "scene", -1);
        if (sceneIndex < 0) sceneIndex = 0;
        const Matrix identity{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        if (static_cast<size_t>(sceneIndex) < scenes.size()) {
            const Json& roots = member(scenes[sceneIndex], "nodes");
            if (roots.isArray())
                for (const Json& root : roots.items())
                    addNode(doc, out, meshAssets, nodes, lights, root.isNumber() ? static_cast<int>(root.number()) : -1,
                        identity, 0);
        }
    } catch (const std::exception& e) {
        out = GltfScene{};
        return fail(error, e.what());
    }
    std::cout << "Imported glTF: " << out.meshes.size() << " mesh assets, " << out.instances.size() << " instances, "
              << out.specularGlossinessMaterials << " specular-glossiness materials, " << out.punctualLights.size()
              << " punctual lights\n";
    return true;
}

}  // namespace wpt
