#include "loaders.hpp"

#include "image_decode.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>
#include <sstream>
#include <unordered_map>

namespace wpt {

namespace {

bool fail(std::string* error, const std::string& message) {
    if (error) *error = message;
    return false;
}

std::string directoryOf(const std::string& path) {
    size_t s = path.find_last_of("/\\");
    return s == std::string::npos ? std::string() : path.substr(0, s + 1);
}

wpt_material objMaterial(const float kd[3], const float ke[3], float ni) {
    wpt_material m{};
    bool emissive = ke[0] != 0.0f || ke[1] != 0.0f || ke[2] != 0.0f;
    m.type = emissive ? WPT_MATERIAL_EMISSIVE : WPT_MATERIAL_DIFFUSE;
    for (int k = 0; k < 3; ++k) {
        m.albedo[k] = kd[k];
        m.emission[k] = ke[k];
    }
    m.roughness = 0.0f;
    m.ior = ni > 0.0f ? ni : 1.0f;
    m.opacity = 1.0f;
    m.alpha_cutoff = 0.5f;
    m.attenuation_color[0] = m.attenuation_color[1] = m.attenuation_color[2] = 1.0f;
    m.volume_density = 1.0f;
    m.specular[0] = m.specular[1] = m.specular[2] = 1.0f;
    m.glossiness = 1.0f;
    return m;
}

struct MtlEntry {
    float kd[3] = {0.0f, 0.0f, 0.0f};  // tinyobjloader's defaults
    float ke[3] = {0.0f, 0.0f, 0.0f};
    float ni = 1.0f;
    std::string mapKd;
};

bool loadMtl(const std::string& filename, std::vector<std::pair<std::string, MtlEntry>>& out, std::string* error) {
    std::ifstream in(filename);
    if (!in) return fail(error, "could not open material library " + filename);
    std::string line;
    MtlEntry* current = nullptr;
    while (std::getline(in, line)) {
        std::istringstream s(line);
        std::string key;
        if (!(s >> key) || key[0] == '#') continue;
        if (key == "newmtl") {
            std::string name;
            s >> name;
            out.emplace_back(name, MtlEntry{});
            current = &out.back().second;
        } else if (!current) {
            continue;
        } else if (key == "Kd") {
            s >> current->kd[0] >> current->kd[1] >> current->kd[2];
        } else if (key == "Ke") {
            s >> current->ke[0] >> current->ke[1] >> current->ke[2];
        } else if (key == "Ni") {
            s >> current->ni;
        } else if (key == "map_Kd") {
            std::string token, last;
            while (s >> token) last = token;  // options precede the file name
            current->mapKd = last;
        }
    }
    return true;
}

struct Corner {
    int v, t, n;  // 0-based; -1: absent
};

// One "v", "v/t", "v//n" or "v/t/n" reference; negative indices count back
// from the latest element.
bool parseCorner(const std::string& token, size_t nv, size_t nt, size_t nn, Corner& c) {
    int idx[3] = {0, 0, 0};
    bool present[3] = {false, false, false};
    size_t pos = 0;
    for (int k = 0; k < 3 && pos <= token.size(); ++k) {
        size_t slash = token.find('/', pos);
        std::string part = token.substr(pos, slash == std::string::npos ? std::string::npos : slash - pos);
        if (!part.empty()) {
            char* end = nullptr;
            long value = std::strtol(part.c_str(), &end, 10);
            if (*end != '\0' || value == 0) return false;
            idx[k] = static_cast<int>(value);
            present[k] = true;
        }
        if (slash == std::string::npos) break;
        pos = slash + 1;
    }
    if (!present[0]) return false;
    auto resolve = [](int i, size_t count) -> int { return i > 0 ? i - 1 : static_cast<int>(count) + i; };
    c.v = resolve(idx[0], nv);
    c.t = present[1] ? resolve(idx[1], nt) : -1;
    c.n = present[2] ? resolve(idx[2], nn) : -1;
    return c.v >= 0 && static_cast<size_t>(c.v) < nv && (c.t < 0 || static_cast<size_t>(c.t) < nt) &&
        (c.n < 0 || static_cast<size_t>(c.n) < nn);
}

struct ObjFace {
    Corner c[3];
    int material;
};

}  // namespace

bool loadObjScene(const std::string& filename, ObjScene& out, std::string* error) {
    std::ifstream in(filename);
    if (!in) return fail(error, "could not open OBJ file " + filename);
    std::vector<float> positions, texcoords, normals;
    std::vector<std::vector<ObjFace>> shapes(1);
    std::vector<std::pair<std::string, MtlEntry>> mtl;
    std::unordered_map<std::string, int> materialIndex;
    int currentMaterial = -1;
    std::string line;
    size_t lineNumber = 0;
    while (std::getline(in, line)) {
        ++lineNumber;
        std::istringstream s(line);
        std::string key;
        if (!(s >> key) || key[0] == '#') continue;
        if (key == "v") {
            float x = 0, y = 0, z = 0;
            s >> x >> y >> z;
            positions.insert(positions.end(), {x, y, z});
        } else if (key == "vt") {
            float u = 0, v = 0;
            s >> u >> v;
            texcoords.insert(texcoords.end(), {u, v});
        } else if (key == "vn") {
            float x = 0, y = 0, z = 0;
            s >> x >> y >> z;
            normals.insert(normals.end(), {x, y, z});
        } else if (key == "f") {
            std::vector<Corner> poly;
            std::string token;
            while (s >> token) {
                Corner c;
                if (!parseCorner(token, positions.size() / 3, texcoords.size() / 2, normals.size() / 3, c))
                    return fail(error, filename + ":" + std::to_string(lineNumber) + ": bad face index " + token);
                poly.push_back(c);
            }
            for (size_t k = 1; k + 1 < poly.size(); ++k)
                shapes.back().push_back(ObjFace{{poly[0], poly[k], poly[k + 1]}, currentMaterial});
        } else if (key == "o" || key == "g") {
            if (!shapes.back().empty()) shapes.emplace_back();
        } else if (key == "usemtl") {
            std::string name;
            s >> name;
            auto it = materialIndex.find(name);
            currentMaterial = it == materialIndex.end() ? -1 : it->second;
        } else if (key == "mtllib") {
            std::string name;
            while (s >> name) {
                size_t before = mtl.size();
                std::string mtlError;
                if (!loadMtl(directoryOf(filename) + name, mtl, &mtlError)) continue;  // missing library: defaults
                for (size_t i = before; i < mtl.size(); ++i)
                    materialIndex.emplace(mtl[i].first, static_cast<int>(i));
            }
        }
    }

    out = ObjScene{};
    wpt_material fallback{};
    const float kd0[3] = {0.65f, 0.3f, 0.18f}, zero[3] = {0.0f, 0.0f, 0.0f};
    fallback = objMaterial(kd0, zero, 1.0f);
    out.materials.push_back(fallback);
    out.albedoMaps.emplace_back();
    for (const auto& entry : mtl) {
        out.materials.push_back(objMaterial(entry.second.kd, entry.second.ke, entry.second.ni));
        out.albedoMaps.push_back(entry.second.mapKd);
    }
    // Albedo textures, one per distinct file name (obj_loader.cpp:69-104).
    out.albedoTexture.assign(out.materials.size(), WPT_INVALID_INDEX);
    std::unordered_map<std::string, uint32_t> textureIndex;
    for (size_t i = 0; i < out.albedoMaps.size(); ++i) {
        const std::string& name = out.albedoMaps[i];
        if (name.empty()) continue;
        auto known = textureIndex.find(name);
        if (known != textureIndex.end()) {
            out.albedoTexture[i] = known->second;
            continue;
        }
        std::string path = directoryOf(filename) + name;
        std::ifstream image(path, std::ios::binary);
        if (!image) return fail(error, "failed to load texture " + path);
        std::vector<uint8_t> encoded((std::istreambuf_iterator<char>(image)), std::istreambuf_iterator<char>());
        std::vector<uint8_t> rgba;
        uint32_t w = 0, h = 0;
        std::string why;
        if (!decodeImage(encoded.data(), encoded.size(), rgba, w, h, &why)) {
            std::cerr << "Texture " << path << " not decoded (" << why << "); the material keeps its Kd\n";
            continue;
        }
        uint32_t index = static_cast<uint32_t>(out.textureData.size());
        out.textureData.push_back(std::move(rgba));
        out.textureSizes.push_back({w, h});
        textureIndex.emplace(name, index);
        out.albedoTexture[i] = index;
    }

    // Vertex normals where the file gives none: the normalized sum of the
    // unnormalized normals of every face around the vertex.
    std::vector<float> generated(positions.size(), 0.0f);
    auto pos = [&](int v, int k) { return positions[3 * static_cast<size_t>(v) + k]; };
    for (const auto& shape : shapes)
        for (const ObjFace& f : shape) {
            float e1[3], e2[3];
            for (int k = 0; k < 3; ++k) {
                e1[k] = pos(f.c[1].v, k) - pos(f.c[0].v, k);
                e2[k] = pos(f.c[2].v, k) - pos(f.c[0].v, k);
            }
            float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            for (const Corner& c : f.c)
                for (int k = 0; k < 3; ++k) generated[3 * static_cast<size_t>(c.v) + k] += n[k];
        }
    for (size_t v = 0; v < generated.size(); v += 3) {
        float l2 = generated[v] * generated[v] + generated[v + 1] * generated[v + 1] + generated[v + 2] * generated[v + 2];
        if (l2 > 0.0f) {
            float inv = 1.0f / std::sqrt(l2);
            for (int k = 0; k < 3; ++k) generated[v + k] *= inv;
        }
    }

    for (const auto& shape : shapes) {
        if (shape.empty()) continue;
        std::vector<wpt_triangle> mesh;
        mesh.reserve(shape.size());
        for (const ObjFace& f : shape) {
            wpt_triangle t{};
            float* verts[3] = {t.v0, t.v1, t.v2};
            float* norms[3] = {t.n0, t.n1, t.n2};
            float* uvs[3] = {t.uv0, t.uv1, t.uv2};
            bool fileNormals = f.c[0].n >= 0 && f.c[1].n >= 0 && f.c[2].n >= 0;
            bool texcoordsGiven = f.c[0].t >= 0 && f.c[1].t >= 0 && f.c[2].t >= 0;
            for (int i = 0; i < 3; ++i) {
                for (int k = 0; k < 3; ++k) {
                    verts[i][k] = pos(f.c[i].v, k);
                    norms[i][k] = fileNormals ? normals[3 * static_cast<size_t>(f.c[i].n) + k]
                                              : generated[3 * static_cast<size_t>(f.c[i].v) + k];
                }
                if (texcoordsGiven) {
                    uvs[i][0] = texcoords[2 * static_cast<size_t>(f.c[i].t)];
                    uvs[i][1] = texcoords[2 * static_cast<size_t>(f.c[i].t) + 1];
                }
            }
            t.flags = WPT_TRIANGLE_HAS_NORMALS | (texcoordsGiven ? WPT_TRIANGLE_HAS_TEXCOORDS : 0u);
            t.material = f.material >= 0 ? static_cast<uint32_t>(f.material + 1) : 0u;
            mesh.push_back(t);
        }
        out.meshes.push_back(std::move(mesh));
    }
    if (out.meshes.empty()) return fail(error, "OBJ contains no triangles: " + filename);
    return true;
}

bool loadRadianceHdr(const std::string& filename, std::vector<float>& rgb, uint32_t& width, uint32_t& height,
    std::string* error) {
    std::ifstream in(filename, std::ios::binary);
    if (!in) return fail(error, "Environment must be a readable Radiance .hdr file: " + filename);
    std::string line;
    if (!std::getline(in, line) || (line != "#?RADIANCE" && line != "#?RGBE"))
        return fail(error, "Environment must be a readable Radiance .hdr file: " + filename);
    bool format = false;
    while (std::getline(in, line) && !line.empty())
        if (line == "FORMAT=32-bit_rle_rgbe") format = true;
    if (!format) return fail(error, "unsupported HDR format in " + filename);
    if (!std::getline(in, line)) return fail(error, "HDR resolution line missing in " + filename);
    int h = 0, w = 0;
    char tail = 0;
    if (std::sscanf(line.c_str(), "-Y %d +X %d%c", &h, &w, &tail) != 2 || h <= 0 || w <= 0 || h > (1 << 24) ||
        w > (1 << 24))
        return fail(error, "unsupported HDR orientation or size in " + filename);
    width = static_cast<uint32_t>(w);
    height = static_cast<uint32_t>(h);
    rgb.assign(3ull * width * height, 0.0f);
    auto convert = [&](const uint8_t* texel, size_t pixel) {
        if (texel[3] == 0) return;
        float f = std::ldexp(1.0f, static_cast<int>(texel[3]) - (128 + 8));
        for (int k = 0; k < 3; ++k) rgb[3 * pixel + k] = static_cast<float>(texel[k]) * f;
    };
    auto readFlat = [&](size_t first) {
        uint8_t texel[4];
        for (size_t p = first; p < static_cast<size_t>(width) * height; ++p) {
            if (!in.read(reinterpret_cast<char*>(texel), 4)) return false;
            convert(texel, p);
        }
        return true;
    };
    if (width < 8 || width >= 32768) {
        if (!readFlat(0)) return fail(error, "HDR data truncated in " + filename);
        return true;
    }
    std::vector<uint8_t> scan(4ull * width);
    for (uint32_t y = 0; y < height; ++y) {
        uint8_t head[4];
        if (!in.read(reinterpret_cast<char*>(head), 4)) return fail(error, "HDR data truncated in " + filename);
        if (head[0] != 2 || head[1] != 2 || (head[2] & 0x80)) {
            // Not run-length encoded: this texel starts the flat remainder.
            convert(head, static_cast<size_t>(y) * width);
            if (!readFlat(static_cast<size_t>(y) * width + 1)) return fail(error, "HDR data truncated in " + filename);
            return true;
        }
        if ((static_cast<uint32_t>(head[2]) << 8 | head[3]) != width)
            return fail(error, "HDR scanline width mismatch in " + filename);
        for (int k = 0; k < 4; ++k) {
            uint32_t x = 0;
            while (x < width) {
                int count = in.get();
                if (count == EOF) return fail(error, "HDR data truncated in " + filename);
                if (count > 128) {
                    int value = in.get();
                    count -= 128;
                    if (value == EOF || x + count > width) return fail(error, "bad HDR run in " + filename);
                    for (int i = 0; i < count; ++i) scan[4ull * x++ + k] = static_cast<uint8_t>(value);
                } else {
                    if (count == 0 || x + count > width) return fail(error, "bad HDR run in " + filename);
                    for (int i = 0; i < count; ++i) {
                        int value = in.get();
                        if (value == EOF) return fail(error, "HDR data truncated in " + filename);
                        scan[4ull * x++ + k] = static_cast<uint8_t>(value);
                    }
                }
            }
        }
        for (uint32_t x = 0; x < width; ++x) convert(&scan[4ull * x], static_cast<size_t>(y) * width + x);
    }
    return true;
}

}  // namespace wpt
