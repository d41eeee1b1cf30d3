// The shading-heavy terrain preset: everything the textured, caustic path
// exercises at once (see terrainScene in scene_factory.hpp).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "scene_factory.hpp"

namespace wpt {

namespace {

constexpr float kPi = 3.14159265358979f;

uint32_t hashU32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

// Lattice value in [0, 1) of a 2-D grid with the given period (tileable).
float lattice(int32_t x, int32_t y, uint32_t period, uint32_t seed) {
    uint32_t px = static_cast<uint32_t>(x) % period, py = static_cast<uint32_t>(y) % period;
    return static_cast<float>(hashU32(px * 0x8da6b343u ^ py * 0xd8163841u ^ seed * 0x9e3779b9u) >> 8) *
        (1.0f / 16777216.0f);
}

// Smooth value noise in [0, 1), tileable with period `period` lattice cells.
float tileNoise(float x, float y, uint32_t period, uint32_t seed) {
    float fx = std::floor(x), fy = std::floor(y);
    int32_t ix = static_cast<int32_t>(fx), iy = static_cast<int32_t>(fy);
    float tx = x - fx, ty = y - fy;
    float sx = tx * tx * (3.0f - 2.0f * tx), sy = ty * ty * (3.0f - 2.0f * ty);
    float a = lattice(ix, iy, period, seed), b = lattice(ix + 1, iy, period, seed);
    float c = lattice(ix, iy + 1, period, seed), d = lattice(ix + 1, iy + 1, period, seed);
    return (a + (b - a) * sx) + ((c + (d - c) * sx) - (a + (b - a) * sx)) * sy;
}

// Fractal sum over octaves at (u, v) in [0, 1)^2, tileable, in [0, 1).
float tileFbm(float u, float v, uint32_t frequency, int octaves, uint32_t seed) {
    float sum = 0.0f, weight = 0.5f, norm = 0.0f;
    for (int o = 0; o < octaves; ++o) {
        uint32_t f = frequency << o;
        sum += weight * tileNoise(u * static_cast<float>(f), v * static_cast<float>(f), f, seed + 977u * o);
        norm += weight;
        weight *= 0.5f;
    }
    return sum / norm;
}

// Untiled fractal noise for the terrain height, in [0, 1).
float fieldFbm(float x, float y, uint32_t seed) {
    float sum = 0.0f, weight = 0.5f, norm = 0.0f, f = 1.0f;
    for (int o = 0; o < 5; ++o) {
        sum += weight * tileNoise(x * f + 1000.0f, y * f + 1000.0f, 0x7fffffffu, seed + 31u * o);
        norm += weight;
        weight *= 0.5f;
        f *= 2.03f;
    }
    return sum / norm;
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9e3779b97f4a7c15ull + 0x632be59bd9b4e019ull) {}
    uint64_t next() {
        s += 0x9e3779b97f4a7c15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    float uniform() { return static_cast<float>(next() >> 40) * (1.0f / 16777216.0f); }
    float range(float a, float b) { return a + (b - a) * uniform(); }
    float normal() {
        float u1 = std::max(uniform(), 1e-7f), u2 = uniform();
        return std::sqrt(-2.0f * std::log(u1)) * std::cos(2.0f * kPi * u2);
    }
};

template <class F>
void parallelRows(uint32_t rows, F&& f) {
    unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            for (uint32_t y = t; y < rows; y += threads) f(y);
        });
    for (std::thread& th : pool) th.join();
}

uint8_t toByte(float v) { return static_cast<uint8_t>(std::min(255.0f, std::max(0.0f, v * 255.0f + 0.5f))); }

float smoothstep(float a, float b, float x) {
    float t = std::min(1.0f, std::max(0.0f, (x - a) / (b - a)));
    return t * t * (3.0f - 2.0f * t);
}

enum TextureKind { kAlbedo, kNormal, kLeaf, kMetalRough };

struct TextureRecipe {
    TextureKind kind;
    float lo[3], hi[3];  // palette for albedo and leaves
    uint32_t frequency;
};

// The 16 textures: 0 terrain, 1-4 rock, 5 bark, 6-9 normal maps, 10-11
// leaves (alpha masks), 12-13 metallic-roughness, 14-15 spare albedo.
const TextureRecipe kTextures[16] = {
    {kAlbedo, {0.30f, 0.24f, 0.16f}, {0.26f, 0.45f, 0.16f}, 8},
    {kAlbedo, {0.30f, 0.29f, 0.27f}, {0.62f, 0.60f, 0.56f}, 6},
    {kAlbedo, {0.35f, 0.27f, 0.22f}, {0.66f, 0.55f, 0.46f}, 5},
    {kAlbedo, {0.20f, 0.22f, 0.25f}, {0.55f, 0.58f, 0.62f}, 7},
    {kAlbedo, {0.40f, 0.36f, 0.28f}, {0.72f, 0.68f, 0.58f}, 9},
    {kAlbedo, {0.22f, 0.14f, 0.08f}, {0.45f, 0.32f, 0.20f}, 12},
    {kNormal, {0, 0, 0}, {0, 0, 0}, 8},
    {kNormal, {0, 0, 0}, {0, 0, 0}, 6},
    {kNormal, {0, 0, 0}, {0, 0, 0}, 10},
    {kNormal, {0, 0, 0}, {0, 0, 0}, 5},
    {kLeaf, {0.10f, 0.28f, 0.06f}, {0.35f, 0.55f, 0.12f}, 6},
    {kLeaf, {0.45f, 0.25f, 0.05f}, {0.70f, 0.50f, 0.10f}, 5},
    {kMetalRough, {0, 0, 0}, {0, 0, 0}, 7},
    {kMetalRough, {0, 0, 0}, {0, 0, 0}, 9},
    {kAlbedo, {0.55f, 0.50f, 0.40f}, {0.85f, 0.80f, 0.70f}, 4},
    {kAlbedo, {0.15f, 0.20f, 0.30f}, {0.40f, 0.50f, 0.70f}, 11},
};

std::vector<uint8_t> makeTexture(const TextureRecipe& r, uint32_t size, uint32_t seed) {
    std::vector<uint8_t> t(static_cast<size_t>(size) * size * 4);
    float inv = 1.0f / static_cast<float>(size);
    if (r.kind == kNormal) {
        // Tangent-space normal of the tileable height field (central
        // differences with wrap-around).
        std::vector<float> height(static_cast<size_t>(size) * size);
        parallelRows(size, [&](uint32_t y) {
            for (uint32_t x = 0; x < size; ++x)
                height[static_cast<size_t>(y) * size + x] = tileFbm((x + 0.5f) * inv, (y + 0.5f) * inv, r.frequency, 4, seed);
        });
        parallelRows(size, [&](uint32_t y) {
            for (uint32_t x = 0; x < size; ++x) {
                auto h = [&](uint32_t xx, uint32_t yy) { return height[static_cast<size_t>(yy % size) * size + xx % size]; };
                float hx = h(x + 1, y) - h(x + size - 1, y), hy = h(x, y + 1) - h(x, y + size - 1);
                float k = 0.02f * static_cast<float>(size);
                float nx = -hx * k, ny = -hy * k, nz = 1.0f;
                float l = 1.0f / std::sqrt(nx * nx + ny * ny + nz * nz);
                uint8_t* p = &t[(static_cast<size_t>(y) * size + x) * 4];
                p[0] = toByte(0.5f + 0.5f * nx * l);
                p[1] = toByte(0.5f + 0.5f * ny * l);
                p[2] = toByte(0.5f + 0.5f * nz * l);
                p[3] = 255;
            }
        });
        return t;
    }
    parallelRows(size, [&](uint32_t y) {
        float v = (static_cast<float>(y) + 0.5f) * inv;
        for (uint32_t x = 0; x < size; ++x) {
            float u = (static_cast<float>(x) + 0.5f) * inv;
            uint8_t* p = &t[(static_cast<size_t>(y) * size + x) * 4];
            float n = tileFbm(u, v, r.frequency, 4, seed);
            switch (r.kind) {
            case kAlbedo: {
                float detail = tileFbm(u, v, r.frequency * 8, 2, seed + 7u);
                float m = smoothstep(0.35f, 0.65f, n);
                for (int k = 0; k < 3; ++k)
                    p[k] = toByte((r.lo[k] + (r.hi[k] - r.lo[k]) * m) * (0.75f + 0.5f * detail));
                p[3] = 255;
                break;
            }
            case kNormal:
                break;  // made above
            case kLeaf: {
                // Leaf blobs: cells of the noise above a threshold are opaque.
                float cell = tileFbm(u, v, r.frequency * 2, 3, seed + 3u);
                for (int k = 0; k < 3; ++k) p[k] = toByte(r.lo[k] + (r.hi[k] - r.lo[k]) * n);
                p[3] = cell > 0.48f ? 255 : 0;
                break;
            }
            case kMetalRough: {
                float metal = tileFbm(u, v, r.frequency * 2, 2, seed + 5u);
                p[0] = 255;
                p[1] = toByte(0.2f + 0.7f * n);                 // roughness
                p[2] = toByte(smoothstep(0.45f, 0.55f, metal)); // metallic
                p[3] = 255;
                break;
            }
            }
        }
    });
    return t;
}

wpt_texture_binding binding(uint32_t texture, float tiling) {
    wpt_texture_binding b{};
    b.texture = texture;
    b.texcoord = 0;
    b.scale[0] = b.scale[1] = tiling;
    b.has_transform = tiling != 1.0f ? 1u : 0u;
    return b;
}

void set3(float d[3], float x, float y, float z) {
    d[0] = x;
    d[1] = y;
    d[2] = z;
}

// Per-triangle tangent frame from positions and texture coordinates.
void tangentFrame(wpt_triangle& t) {
    float e1[3], e2[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = t.v1[k] - t.v0[k];
        e2[k] = t.v2[k] - t.v0[k];
    }
    float du1 = t.uv1[0] - t.uv0[0], dv1 = t.uv1[1] - t.uv0[1];
    float du2 = t.uv2[0] - t.uv0[0], dv2 = t.uv2[1] - t.uv0[1];
    float det = du1 * dv2 - du2 * dv1;
    if (std::fabs(det) < 1e-12f) {
        set3(t.tangent, 1, 0, 0);
        set3(t.bitangent, 0, 0, 1);
        return;
    }
    float r = 1.0f / det;
    for (int k = 0; k < 3; ++k) {
        t.tangent[k] = (e1[k] * dv2 - e2[k] * dv1) * r;
        t.bitangent[k] = (e2[k] * du1 - e1[k] * du2) * r;
    }
}

struct TerrainShape {
    float half, amplitude;
    uint32_t seed;
    // Flat garden within r < 3, rising hills beyond.
    float height(float x, float z) const {
        float r = std::sqrt(x * x + z * z);
        float falloff = smoothstep(3.0f, 10.0f, r);
        return amplitude * (fieldFbm(x * 0.06f, z * 0.06f, seed) - 0.35f) * falloff;
    }
};

wpt_transform placed(float x, float y, float z, float scale, float yaw) {
    wpt_transform t = makeScaledTransform(x, y, z, scale);
    float c = std::cos(yaw), s = std::sin(yaw);
    // Rotation about +Y applied before the uniform scale.
    float m[3][3] = {{c * scale, 0, s * scale}, {0, scale, 0}, {-s * scale, 0, c * scale}};
    float inv[3][3] = {{c / scale, 0, -s / scale}, {0, 1 / scale, 0}, {s / scale, 0, c / scale}};
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
            t.local_to_world[r][k] = m[r][k];
            t.world_to_local[r][k] = inv[r][k];
        }
    return t;
}

}  // namespace

void terrainScene(HostScene& s, const SceneConfig& c) {
    const uint32_t seed = c.seed;
    const uint32_t texSize = c.textureSize;
    Rng rng(seed ^ 0x5eedu);

    // ---- textures
    s.textureData.resize(16);
    for (uint32_t i = 0; i < 16; ++i) s.textureData[i] = makeTexture(kTextures[i], texSize, seed * 131u + i);
    s.textures.clear();
    for (uint32_t i = 0; i < 16; ++i) s.textures.push_back(wpt_texture{texSize, texSize, s.textureData[i].data()});

    // ---- materials (0 terrain, 1-4 rocks, 5-6 foliage, 7-8 crystals, 9.. lamps)
    auto add = [&](const wpt_material& m, const wpt_material_ext& e) {
        s.materials.push_back(m);
        s.materialExt.push_back(e);
        return static_cast<uint32_t>(s.materials.size() - 1);
    };
    wpt_material_ext e = plainMaterialExt();
    e.albedo_texture = binding(0, 24.0f);
    e.normal_texture = binding(6, 24.0f);
    uint32_t terrainMat = add(makeMaterial(WPT_MATERIAL_DIFFUSE, 1, 1, 1), e);
    uint32_t rockMat[4];
    for (int i = 0; i < 4; ++i) {
        e = plainMaterialExt();
        e.albedo_texture = binding(1 + i, 2.0f);
        if (i < 3) e.normal_texture = binding(7 + i, 2.0f);
        wpt_material m = makeMaterial(WPT_MATERIAL_DIFFUSE, 1, 1, 1);
        if (i == 2) {  // metallic-roughness glTF rock
            e.gltf_pbr_workflow = 1;
            e.metallic_roughness_texture = binding(12, 2.0f);
            m.metallic = 1.0f;
            m.roughness = 1.0f;
        }
        if (i == 3) {
            m = makeMaterial(WPT_MATERIAL_METAL, 1, 1, 1);
            m.roughness = 0.35f;
        }
        rockMat[i] = add(m, e);
    }
    uint32_t leafMat[2];
    for (int i = 0; i < 2; ++i) {
        e = plainMaterialExt();
        e.albedo_texture = binding(10 + i, 1.0f);
        if (i == 1) e.normal_texture = binding(7, 1.0f);
        wpt_material m = makeMaterial(WPT_MATERIAL_DIFFUSE, 1, 1, 1);
        m.alpha_masked = 1;
        m.alpha_cutoff = 0.5f;
        leafMat[i] = add(m, e);
    }
    uint32_t crystalMat[2];
    const float tint[2][3] = {{0.55f, 0.8f, 1.0f}, {1.0f, 0.6f, 0.8f}};
    for (int i = 0; i < 2; ++i) {
        e = plainMaterialExt();
        e.dispersion = 0.02f;
        wpt_material m = makeMaterial(WPT_MATERIAL_DIELECTRIC, 1, 1, 1);
        m.ior = 1.5f;
        set3(m.attenuation_color, tint[i][0], tint[i][1], tint[i][2]);
        m.attenuation_distance = 0.35f + 0.15f * i;
        m.volume_density = 1.0f;
        crystalMat[i] = add(m, e);
    }

    // ---- terrain mesh: r x r cells over [-30, 30]^2
    const uint32_t r = c.terrainResolution;
    const float half = 30.0f;
    TerrainShape shape{half, 5.0f, seed};
    auto vertex = [&](uint32_t i, uint32_t j, float p[3], float n[3], float uv[2]) {
        float x = -half + 2.0f * half * static_cast<float>(i) / static_cast<float>(r);
        float z = -half + 2.0f * half * static_cast<float>(j) / static_cast<float>(r);
        float h = 2.0f * half / static_cast<float>(r);
        set3(p, x, shape.height(x, z), z);
        float dx = shape.height(x + h, z) - shape.height(x - h, z);
        float dz = shape.height(x, z + h) - shape.height(x, z - h);
        float nx = -dx, ny = 2.0f * h, nz = -dz, l = 1.0f / std::sqrt(nx * nx + ny * ny + nz * nz);
        set3(n, nx * l, ny * l, nz * l);
        uv[0] = static_cast<float>(i) / static_cast<float>(r);
        uv[1] = static_cast<float>(j) / static_cast<float>(r);
    };
    std::vector<wpt_triangle> terrain(2ull * r * r);
    parallelRows(r, [&](uint32_t j) {
        for (uint32_t i = 0; i < r; ++i) {
            float p[4][3], n[4][3], uv[4][2];
            vertex(i, j, p[0], n[0], uv[0]);
            vertex(i + 1, j, p[1], n[1], uv[1]);
            vertex(i + 1, j + 1, p[2], n[2], uv[2]);
            vertex(i, j + 1, p[3], n[3], uv[3]);
            const int corner[2][3] = {{0, 3, 2}, {0, 2, 1}};  // counter-clockwise seen from above
            for (int k = 0; k < 2; ++k) {
                wpt_triangle t = makeTriangle(p[corner[k][0]], p[corner[k][1]], p[corner[k][2]], terrainMat);
                std::memcpy(t.n0, n[corner[k][0]], sizeof t.n0);
                std::memcpy(t.n1, n[corner[k][1]], sizeof t.n1);
                std::memcpy(t.n2, n[corner[k][2]], sizeof t.n2);
                std::memcpy(t.uv0, uv[corner[k][0]], sizeof t.uv0);
                std::memcpy(t.uv1, uv[corner[k][1]], sizeof t.uv1);
                std::memcpy(t.uv2, uv[corner[k][2]], sizeof t.uv2);
                t.flags = WPT_TRIANGLE_HAS_NORMALS | WPT_TRIANGLE_HAS_TEXCOORDS;
                tangentFrame(t);
                terrain[2ull * (static_cast<size_t>(j) * r + i) + k] = t;
            }
        }
    });
    auto addMesh = [&](std::vector<wpt_triangle>&& tris) {
        uint32_t first = static_cast<uint32_t>(s.meshTriangles.size());
        s.meshTriangles.insert(s.meshTriangles.end(), tris.begin(), tris.end());
        s.meshes.push_back(wpt_mesh{first, static_cast<uint32_t>(tris.size())});
        return static_cast<uint32_t>(s.meshes.size() - 1);
    };
    auto addInstance = [&](uint32_t mesh, const wpt_transform& t, uint32_t override) {
        wpt_instance inst{};
        inst.mesh = mesh;
        inst.material_override = override;
        inst.acceleration = WPT_ACCEL_BVH;
        inst.transform = t;
        s.instances.push_back(inst);
    };
    uint32_t terrainMesh = addMesh(std::move(terrain));
    addInstance(terrainMesh, makeScaledTransform(0, 0, 0, 1), WPT_INVALID_INDEX);

    // ---- rocks: four displaced icospheres with spherical texture coordinates
    uint32_t rockFrequency = std::max(2u, c.icosphereFrequency / 2);
    uint32_t rockMesh[4];
    for (int v = 0; v < 4; ++v) {
        std::vector<wpt_triangle> tris = displacedIcosphere(rockFrequency, seed * 7u + static_cast<uint32_t>(v), 0.25f);
        for (wpt_triangle& t : tris) {
            float* verts[3] = {t.v0, t.v1, t.v2};
            float* uvs[3] = {t.uv0, t.uv1, t.uv2};
            for (int k = 0; k < 3; ++k) {
                float x = verts[k][0], y = verts[k][1], z = verts[k][2];
                float l = std::sqrt(x * x + y * y + z * z);
                uvs[k][0] = 0.5f + std::atan2(z, x) / (2.0f * kPi);
                uvs[k][1] = 0.5f + std::asin(std::max(-1.0f, std::min(1.0f, y / l))) / kPi;
            }
            // Keep a triangle on one side of the longitude seam.
            float umax = std::max(t.uv0[0], std::max(t.uv1[0], t.uv2[0]));
            for (float* uv : uvs)
                if (umax - uv[0] > 0.5f) uv[0] += 1.0f;
            t.material = rockMat[v];
            t.flags |= WPT_TRIANGLE_HAS_TEXCOORDS;
            tangentFrame(t);
        }
        rockMesh[v] = addMesh(std::move(tris));
    }

    // ---- foliage: a bush of 48 crossed leaf cards
    std::vector<wpt_triangle> bush;
    for (int k = 0; k < 48; ++k) {
        float cx = rng.range(-0.45f, 0.45f), cy = rng.range(0.15f, 1.1f), cz = rng.range(-0.45f, 0.45f);
        float yaw = rng.range(0.0f, 2.0f * kPi), pitch = rng.range(-0.6f, 0.6f), h = 0.22f;
        float ax = std::cos(yaw), az = std::sin(yaw);
        float ux = -az * std::sin(pitch), uy = std::cos(pitch), uz = ax * std::sin(pitch);
        float p[4][3];
        const float sgn[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
        for (int q = 0; q < 4; ++q)
            set3(p[q], cx + h * (sgn[q][0] * ax + sgn[q][1] * ux), cy + h * sgn[q][1] * uy,
                cz + h * (sgn[q][0] * az + sgn[q][1] * uz));
        const float uvq[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        const int tri[2][3] = {{0, 1, 2}, {0, 2, 3}};
        for (auto& tr : tri) {
            wpt_triangle t = makeTriangle(p[tr[0]], p[tr[1]], p[tr[2]], leafMat[0]);
            std::memcpy(t.uv0, uvq[tr[0]], sizeof t.uv0);
            std::memcpy(t.uv1, uvq[tr[1]], sizeof t.uv1);
            std::memcpy(t.uv2, uvq[tr[2]], sizeof t.uv2);
            t.flags = WPT_TRIANGLE_HAS_TEXCOORDS;
            tangentFrame(t);
            bush.push_back(t);
        }
    }
    uint32_t bushMesh = addMesh(std::move(bush));

    // ---- crystal: hexagonal prism with pointed ends (24 triangles)
    std::vector<wpt_triangle> crystal;
    {
        const float radius = 0.12f, top = 0.5f, tip = 0.68f, bottom = 0.0f, base = -0.1f;
        for (int k = 0; k < 6; ++k) {
            float a0 = 2.0f * kPi * k / 6.0f, a1 = 2.0f * kPi * (k + 1) / 6.0f;
            float p0[3] = {radius * std::cos(a0), bottom, radius * std::sin(a0)};
            float p1[3] = {radius * std::cos(a1), bottom, radius * std::sin(a1)};
            float q0[3] = {p0[0], top, p0[2]}, q1[3] = {p1[0], top, p1[2]};
            float apex[3] = {0, tip, 0}, nadir[3] = {0, base, 0};
            crystal.push_back(makeTriangle(p0, q1, p1, crystalMat[0]));
            crystal.push_back(makeTriangle(p0, q0, q1, crystalMat[0]));
            crystal.push_back(makeTriangle(q0, apex, q1, crystalMat[0]));
            crystal.push_back(makeTriangle(p0, p1, nadir, crystalMat[0]));
        }
    }
    uint32_t crystalMesh = addMesh(std::move(crystal));

    // ---- props: 84 rocks, 48 bushes (30 % of the props, alpha-masked),
    // 28 crystals in the flat garden around the origin
    for (int k = 0; k < 84; ++k) {
        float rr = rng.range(3.2f, 24.0f), a = rng.range(0.0f, 2.0f * kPi);
        float x = rr * std::cos(a), z = rr * std::sin(a), sc = rng.range(0.3f, 1.4f);
        addInstance(rockMesh[k % 4], placed(x, shape.height(x, z) + 0.2f * sc, z, sc, rng.range(0.0f, 2 * kPi)),
            WPT_INVALID_INDEX);
    }
    for (int k = 0; k < 48; ++k) {
        float rr = rng.range(2.6f, 20.0f), a = rng.range(0.0f, 2.0f * kPi);
        float x = rr * std::cos(a), z = rr * std::sin(a), sc = rng.range(0.6f, 1.4f);
        addInstance(bushMesh, placed(x, shape.height(x, z), z, sc, rng.range(0.0f, 2 * kPi)), leafMat[k % 2]);
    }
    for (int k = 0; k < 28; ++k) {
        float rr = std::sqrt(rng.uniform()) * 2.2f, a = rng.range(0.0f, 2.0f * kPi);
        float x = rr * std::cos(a), z = rr * std::sin(a), sc = rng.range(0.6f, 1.6f);
        addInstance(crystalMesh, placed(x, 0.0f, z, sc, rng.range(0.0f, 2 * kPi)), crystalMat[k % 2]);
    }

    // ---- lamps: 256 downward quads, each with its own log-normal radiance
    std::vector<wpt_triangle> lamps;
    for (int k = 0; k < 256; ++k) {
        float rr = std::sqrt(rng.uniform()) * 14.0f, a = rng.range(0.0f, 2.0f * kPi);
        float x = rr * std::cos(a), z = rr * std::sin(a), y = shape.height(x, z) + rng.range(2.5f, 4.0f);
        float scale = std::exp(rng.normal()) * 6.0f;
        float warm = rng.uniform();
        uint32_t m = add(makeMaterial(WPT_MATERIAL_EMISSIVE, 0, 0, 0, scale, scale * (0.75f + 0.2f * warm),
                             scale * (0.45f + 0.4f * warm)),
            plainMaterialExt());
        const float h = 0.12f;
        float p0[3] = {x - h, y, z - h}, p1[3] = {x + h, y, z - h}, p2[3] = {x + h, y, z + h}, p3[3] = {x - h, y, z + h};
        lamps.push_back(makeTriangle(p0, p1, p2, m));  // normal -y
        lamps.push_back(makeTriangle(p0, p2, p3, m));
    }
    uint32_t lampMesh = addMesh(std::move(lamps));
    addInstance(lampMesh, makeScaledTransform(0, 0, 0, 1), WPT_INVALID_INDEX);

    // ---- sky: log-normal cloud modulation over a gradient, plus the sun
    const uint32_t ew = texSize >= 1024 ? 2048u : 256u, eh = ew / 2;
    s.environmentWidth = ew;
    s.environmentHeight = eh;
    s.environment.assign(3ull * ew * eh, 0.0f);
    const float sunElevation = 0.61f, sunAzimuth = 1.0f;
    const float sun[3] = {std::cos(sunElevation) * std::cos(sunAzimuth), std::sin(sunElevation),
        std::cos(sunElevation) * std::sin(sunAzimuth)};
    parallelRows(eh, [&](uint32_t y) {
        float elevation = 0.5f * kPi - (static_cast<float>(y) + 0.5f) / static_cast<float>(eh) * kPi;
        for (uint32_t x = 0; x < ew; ++x) {
            float azimuth = 2.0f * kPi * (static_cast<float>(x) + 0.5f) / static_cast<float>(ew);
            float d[3] = {std::cos(elevation) * std::cos(azimuth), std::sin(elevation),
                std::cos(elevation) * std::sin(azimuth)};
            float rgb[3];
            if (elevation < 0.0f) {
                set3(rgb, 0.12f, 0.10f, 0.08f);
            } else {
                float t = std::sqrt(std::sin(elevation));
                float cloud = tileFbm(static_cast<float>(x) / ew, static_cast<float>(y) / eh, 8, 4, seed + 99u);
                float g = std::exp(1.2f * (cloud - 0.5f));  // log-normal modulation
                rgb[0] = (0.75f + (0.25f - 0.75f) * t) * g;
                rgb[1] = (0.82f + (0.45f - 0.82f) * t) * g;
                rgb[2] = (0.95f + (0.95f - 0.95f) * t) * g;
            }
            float cosine = d[0] * sun[0] + d[1] * sun[1] + d[2] * sun[2];
            float angle = std::acos(std::max(-1.0f, std::min(1.0f, cosine)));
            float glow = 30.0f * std::exp(-angle * angle / (2.0f * 0.06f * 0.06f));
            float disk = angle < 0.012f ? 20000.0f : 0.0f;
            float* out = &s.environment[3ull * (static_cast<size_t>(y) * ew + x)];
            out[0] = rgb[0] + glow + disk;
            out[1] = rgb[1] + 0.9f * glow + 0.95f * disk;
            out[2] = rgb[2] + 0.75f * glow + 0.85f * disk;
        }
    });
}

}  // namespace wpt
