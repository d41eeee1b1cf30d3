#include "scene_factory.hpp"

#include "loaders.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace wpt {

namespace {

void set3(float d[3], float x, float y, float z) {
    d[0] = x;
    d[1] = y;
    d[2] = z;
}

// Reference nextSceneRandom (src/scene/scene_host.cpp:55-58).
float nextSceneRandom(uint32_t& state) {
    state = state * 1664525u + 1013904223u;
    return static_cast<float>(state >> 8) * (1.0f / 16777216.0f);
}

void addQuad(std::vector<wpt_triangle>& out, const float a[3], const float b[3], const float c[3], const float d[3],
    uint32_t material) {
    out.push_back(makeTriangle(a, b, c, material));
    out.push_back(makeTriangle(a, c, d, material));
}

wpt_sphere sphere(float x, float y, float z, float r, uint32_t material, uint32_t accel = WPT_ACCEL_BVH) {
    wpt_sphere s{};
    set3(s.center, x, y, z);
    s.radius = r;
    s.material = material;
    s.acceleration = accel;
    return s;
}

// ---- value noise (seeded lattice hash, smoothstep trilinear interpolation)
uint32_t hash3(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    uint32_t h = static_cast<uint32_t>(x) * 0x8da6b343u ^ static_cast<uint32_t>(y) * 0xd8163841u ^
        static_cast<uint32_t>(z) * 0xcb1ab31fu ^ seed * 0x9e3779b9u;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

double lattice(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    return static_cast<double>(hash3(x, y, z, seed) >> 8) / 8388608.0 - 1.0;  // [-1, 1)
}

double valueNoise(double x, double y, double z, uint32_t seed) {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int32_t ix = static_cast<int32_t>(fx), iy = static_cast<int32_t>(fy), iz = static_cast<int32_t>(fz);
    double tx = x - fx, ty = y - fy, tz = z - fz;
    auto smooth = [](double t) { return t * t * (3.0 - 2.0 * t); };
    double sx = smooth(tx), sy = smooth(ty), sz = smooth(tz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
    auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
    double x00 = lerp(c[0][0][0], c[1][0][0], sx), x10 = lerp(c[0][1][0], c[1][1][0], sx);
    double x01 = lerp(c[0][0][1], c[1][0][1], sx), x11 = lerp(c[0][1][1], c[1][1][1], sx);
    return lerp(lerp(x00, x10, sy), lerp(x01, x11, sy), sz);
}

double fbm(double x, double y, double z, uint32_t seed) {
    return 0.6 * valueNoise(3.0 * x, 3.0 * y, 3.0 * z, seed) + 0.3 * valueNoise(7.0 * x, 7.0 * y, 7.0 * z, seed + 1) +
        0.1 * valueNoise(17.0 * x, 17.0 * y, 17.0 * z, seed + 2);
}

// Cornell room of the reference (scene_host.cpp:317-347): walls and the
// ceiling light as static triangles, materials 0..4 first.
void referenceCornellRoom(HostScene& s, const SceneConfig& c, bool isCrystal, bool isPrism) {
    const float hw = 2.5f, floorY = -1.0f, ceilY = 3.0f, backZ = -2.5f, frontZ = 2.5f;
    auto q = [&](std::array<float, 3> a, std::array<float, 3> b, std::array<float, 3> cc, std::array<float, 3> d,
                 uint32_t m) { addQuad(s.staticTriangles, a.data(), b.data(), cc.data(), d.data(), m); };
    q({-hw, floorY, frontZ}, {hw, floorY, frontZ}, {hw, floorY, backZ}, {-hw, floorY, backZ}, 0);
    q({-hw, ceilY, backZ}, {hw, ceilY, backZ}, {hw, ceilY, frontZ}, {-hw, ceilY, frontZ}, 0);
    q({-hw, floorY, backZ}, {hw, floorY, backZ}, {hw, ceilY, backZ}, {-hw, ceilY, backZ}, 0);
    q({-hw, floorY, frontZ}, {-hw, floorY, backZ}, {-hw, ceilY, backZ}, {-hw, ceilY, frontZ}, 1);
    q({hw, floorY, backZ}, {hw, floorY, frontZ}, {hw, ceilY, frontZ}, {hw, ceilY, backZ}, 2);
    std::array<float, 3> corner = isCrystal ? std::array<float, 3>{-0.5f, 2.35f, -1.6f}
        : (isPrism ? std::array<float, 3>{-0.9f, 2.35f, -1.6f} : std::array<float, 3>{-0.75f, 2.95f, -0.75f});
    std::array<float, 3> eu = isCrystal ? std::array<float, 3>{1.0f, 0.0f, 0.0f}
        : (isPrism ? std::array<float, 3>{1.8f, 0.0f, 0.0f} : std::array<float, 3>{1.5f, 0.0f, 0.0f});
    std::array<float, 3> ev = isCrystal ? std::array<float, 3>{0.0f, 0.25f, 0.42f}
        : (isPrism ? std::array<float, 3>{0.0f, 0.45f, 0.75f} : std::array<float, 3>{0.0f, 0.0f, 1.5f});
    auto add = [](std::array<float, 3> a, std::array<float, 3> b) {
        return std::array<float, 3>{a[0] + b[0], a[1] + b[1], a[2] + b[2]};
    };
    q(corner, add(corner, eu), add(add(corner, eu), ev), add(corner, ev), 4);
    float e = isCrystal ? 32.0f : (isPrism ? 10.0f : 16.0f);
    std::vector<wpt_material> room = {
        makeMaterial(WPT_MATERIAL_DIFFUSE, 0.73f, 0.73f, 0.73f),
        c.neutralRoom ? makeMaterial(WPT_MATERIAL_DIFFUSE, 0.68f, 0.68f, 0.68f) : makeMaterial(WPT_MATERIAL_DIFFUSE, 0.65f, 0.05f, 0.05f),
        c.neutralRoom ? makeMaterial(WPT_MATERIAL_DIFFUSE, 0.58f, 0.58f, 0.58f) : makeMaterial(WPT_MATERIAL_DIFFUSE, 0.12f, 0.45f, 0.15f),
        makeMaterial(WPT_MATERIAL_DIFFUSE, 0.72f, 0.72f, 0.72f),
        makeMaterial(WPT_MATERIAL_EMISSIVE, 1.0f, 1.0f, 1.0f, e, e, e)};
    s.materials.insert(s.materials.begin(), room.begin(), room.end());
    s.blackBackground = true;
}

// RTIOW sphere field (scene_host.cpp:188-212 layout) and three instances of a
// displaced icosphere in place of the external Lucy mesh.
void sphereFieldScene(HostScene& s, const SceneConfig& c, bool baselineMix) {
    s.materials.push_back(makeMaterial(WPT_MATERIAL_DIFFUSE, 0.5f, 0.5f, 0.5f));  // ground
    s.spheres.push_back(sphere(0.0f, -1000.0f, 0.0f, 1000.0f, 0, WPT_ACCEL_DIRECT));
    uint32_t state = baselineMix ? (0x4d595df4u ^ (c.seed * 0x9e3779b9u)) : 0x4d595df4u;
    if (!baselineMix) {
        s.materials.push_back(makeMaterial(WPT_MATERIAL_DIFFUSE, 0.72f, 0.20f, 0.12f));
        s.materials.push_back(makeMaterial(WPT_MATERIAL_DIFFUSE, 0.16f, 0.35f, 0.72f));
        s.materials.push_back(makeMaterial(WPT_MATERIAL_DIFFUSE, 0.22f, 0.62f, 0.28f));
        s.materials.push_back(makeMaterial(WPT_MATERIAL_METAL, 0.82f, 0.84f, 0.88f));
        s.materials.back().roughness = 0.08f;
        s.materials.push_back(makeMaterial(WPT_MATERIAL_METAL, 0.88f, 0.60f, 0.20f));
        s.materials.back().roughness = 0.05f;
        s.materials.push_back(makeMaterial(WPT_MATERIAL_DIELECTRIC, 1.0f, 1.0f, 1.0f));
        s.materials.back().ior = 1.5f;
    }
    for (int x = -11; x < 11; ++x) {
        for (int z = -11; z < 11; ++z) {
            float choose = nextSceneRandom(state);
            float cx = static_cast<float>(x) + nextSceneRandom(state);
            float cz = static_cast<float>(z) + nextSceneRandom(state);
            uint32_t material;
            if (!baselineMix) {
                if (choose < 0.70f) {
                    float color = nextSceneRandom(state);
                    material = color < 0.34f ? 1u : (color < 0.67f ? 2u : 3u);
                } else if (choose < 0.85f) {
                    material = nextSceneRandom(state) < 0.5f ? 4u : 5u;
                } else {
                    material = 6u;
                }
            } else {
                material = static_cast<uint32_t>(s.materials.size());
                if (choose < 0.80f) {
                    float r = nextSceneRandom(state), g = nextSceneRandom(state), b = nextSceneRandom(state);
                    s.materials.push_back(makeMaterial(WPT_MATERIAL_DIFFUSE, r, g, b));
                } else if (choose < 0.95f) {
                    float r = 0.5f + 0.5f * nextSceneRandom(state), g = 0.5f + 0.5f * nextSceneRandom(state),
                          b = 0.5f + 0.5f * nextSceneRandom(state);
                    wpt_material m = makeMaterial(WPT_MATERIAL_METAL, r, g, b);
                    m.roughness = 0.5f * nextSceneRandom(state);
                    s.materials.push_back(m);
                } else {
                    wpt_material m = makeMaterial(WPT_MATERIAL_DIELECTRIC, 1.0f, 1.0f, 1.0f);
                    m.ior = 1.5f;
                    s.materials.push_back(m);
                }
            }
            s.spheres.push_back(sphere(cx, 0.2f, cz, 0.2f, material));
        }
    }
    // Instance materials: glass, Lambert, mirror (RTIOW's three large objects).
    uint32_t base = static_cast<uint32_t>(s.materials.size());
    wpt_material glass = makeMaterial(WPT_MATERIAL_DIELECTRIC, 1.0f, 1.0f, 1.0f);
    glass.ior = 1.5f;
    s.materials.push_back(glass);
    s.materials.push_back(makeMaterial(WPT_MATERIAL_DIFFUSE, 0.4f, 0.2f, 0.1f));
    s.materials.push_back(makeMaterial(WPT_MATERIAL_METAL, 0.7f, 0.6f, 0.5f));
    s.meshTriangles = displacedIcosphere(c.icosphereFrequency, c.seed, 0.06f);
    s.meshes.push_back(wpt_mesh{0, static_cast<uint32_t>(s.meshTriangles.size())});
    const float offsets[3][3] = {{0.0f, 1.0f, 0.0f}, {-4.0f, 1.0f, 0.0f}, {4.0f, 1.0f, 0.0f}};
    for (uint32_t i = 0; i < 3; ++i) {
        wpt_instance inst{};
        inst.mesh = 0;
        inst.material_override = base + i;
        inst.acceleration = WPT_ACCEL_BVH;
        inst.transform = makeScaledTransform(offsets[i][0], offsets[i][1], offsets[i][2], 1.0f);
        s.instances.push_back(inst);
    }
    s.blackBackground = false;
}

// Unit-scale Cornell box (half extent 1): five walls and an emissive quad as
// static triangles, a Lambertian and a dielectric sphere in the TLAS.
void cornellBoxScene(HostScene& s) {
    s.materials = {makeMaterial(WPT_MATERIAL_DIFFUSE, 0.73f, 0.73f, 0.73f),
        makeMaterial(WPT_MATERIAL_DIFFUSE, 0.65f, 0.05f, 0.05f),
        makeMaterial(WPT_MATERIAL_DIFFUSE, 0.12f, 0.45f, 0.15f),
        makeMaterial(WPT_MATERIAL_EMISSIVE, 1.0f, 1.0f, 1.0f, 17.0f, 12.0f, 4.0f),
        makeMaterial(WPT_MATERIAL_DIELECTRIC, 1.0f, 1.0f, 1.0f)};
    s.materials[4].ior = 1.5f;
    auto q = [&](std::array<float, 3> a, std::array<float, 3> b, std::array<float, 3> c, std::array<float, 3> d,
                 uint32_t m) { addQuad(s.staticTriangles, a.data(), b.data(), c.data(), d.data(), m); };
    q({-1, -1, 1}, {1, -1, 1}, {1, -1, -1}, {-1, -1, -1}, 0);   // floor
    q({-1, 1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, 1, 1}, 0);       // ceiling
    q({-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, 0);   // back
    q({-1, -1, 1}, {-1, -1, -1}, {-1, 1, -1}, {-1, 1, 1}, 1);   // left (red)
    q({1, -1, -1}, {1, -1, 1}, {1, 1, 1}, {1, 1, -1}, 2);       // right (green)
    q({-0.25f, 0.999f, -0.25f}, {0.25f, 0.999f, -0.25f}, {0.25f, 0.999f, 0.25f}, {-0.25f, 0.999f, 0.25f}, 3);
    s.spheres.push_back(sphere(-0.45f, -0.7f, -0.3f, 0.3f, 0));
    s.spheres.push_back(sphere(0.45f, -0.7f, 0.3f, 0.3f, 4));
    s.blackBackground = true;
}

// The unit Cornell box tessellated for the differentiable renderer: 64
// reflective triangles first (back wall 2x2 quads with texture coordinates
// over the wall, floor 4x2, ceiling 2x2, left and right walls 2x4), then the
// light as 2x2 quads = 8 emissive triangles; the two spheres as in the
// Cornell box.  Material 5 marks the back wall, whose albedo the
// differentiable renderer reads from its texture.
void diffCornellScene(HostScene& s) {
    s.materials = {makeMaterial(WPT_MATERIAL_DIFFUSE, 0.73f, 0.73f, 0.73f),
        makeMaterial(WPT_MATERIAL_DIFFUSE, 0.65f, 0.05f, 0.05f),
        makeMaterial(WPT_MATERIAL_DIFFUSE, 0.12f, 0.45f, 0.15f),
        makeMaterial(WPT_MATERIAL_EMISSIVE, 1.0f, 1.0f, 1.0f, 17.0f, 12.0f, 4.0f),
        makeMaterial(WPT_MATERIAL_DIELECTRIC, 1.0f, 1.0f, 1.0f),
        makeMaterial(WPT_MATERIAL_DIFFUSE, 0.73f, 0.73f, 0.73f)};
    s.materials[4].ior = 1.5f;
    // Quads a + U*i/nu + V*j/nv; U x V points into the room.
    auto grid = [&](std::array<float, 3> a, std::array<float, 3> u, std::array<float, 3> v, uint32_t nu, uint32_t nv,
                    uint32_t m, bool texcoords) {
        auto at = [&](uint32_t i, uint32_t j, float out[3]) {
            float fi = static_cast<float>(i) / static_cast<float>(nu), fj = static_cast<float>(j) / static_cast<float>(nv);
            for (int k = 0; k < 3; ++k) out[k] = a[k] + u[k] * fi + v[k] * fj;
        };
        for (uint32_t j = 0; j < nv; ++j) {
            for (uint32_t i = 0; i < nu; ++i) {
                float p00[3], p10[3], p11[3], p01[3];
                at(i, j, p00);
                at(i + 1, j, p10);
                at(i + 1, j + 1, p11);
                at(i, j + 1, p01);
                size_t first = s.staticTriangles.size();
                addQuad(s.staticTriangles, p00, p10, p11, p01, m);
                if (!texcoords) continue;
                // Texture coordinates across the wall: u along U, v along V.
                auto uv = [&](uint32_t ii, uint32_t jj, float out[2]) {
                    out[0] = static_cast<float>(ii) / static_cast<float>(nu);
                    out[1] = static_cast<float>(jj) / static_cast<float>(nv);
                };
                wpt_triangle& t0 = s.staticTriangles[first];
                wpt_triangle& t1 = s.staticTriangles[first + 1];
                uv(i, j, t0.uv0);
                uv(i + 1, j, t0.uv1);
                uv(i + 1, j + 1, t0.uv2);
                uv(i, j, t1.uv0);
                uv(i + 1, j + 1, t1.uv1);
                uv(i, j + 1, t1.uv2);
                t0.flags |= 2u;
                t1.flags |= 2u;
            }
        }
    };
    grid({-1, -1, -1}, {2, 0, 0}, {0, 2, 0}, 2, 2, 5, true);     // back wall (textured)
    grid({-1, -1, 1}, {2, 0, 0}, {0, 0, -2}, 4, 2, 0, false);    // floor
    grid({-1, 1, -1}, {2, 0, 0}, {0, 0, 2}, 2, 2, 0, false);     // ceiling
    grid({-1, -1, 1}, {0, 0, -2}, {0, 2, 0}, 2, 4, 1, false);    // left (red)
    grid({1, -1, -1}, {0, 0, 2}, {0, 2, 0}, 2, 4, 2, false);     // right (green)
    grid({-0.25f, 0.999f, -0.25f}, {0.5f, 0, 0}, {0, 0, 0.5f}, 2, 2, 3, false);  // light
    s.spheres.push_back(sphere(-0.45f, -0.7f, -0.3f, 0.3f, 0));
    s.spheres.push_back(sphere(0.45f, -0.7f, 0.3f, 0.3f, 4));
    s.blackBackground = true;
}

// Reference demo scene (scene_host.cpp:116-163).
void demoScene(HostScene& s) {
    s.spheres.push_back(sphere(0.0f, -100.5f, 0.0f, 100.0f, 3, WPT_ACCEL_DIRECT));
    s.spheres.push_back(sphere(1.35f, 0.08f, 0.0f, 0.5f, 2));
    const float c0[3] = {-1.85f, -0.42f, -0.5f}, c1[3] = {-0.85f, -0.42f, -0.5f}, c2[3] = {-0.85f, 0.58f, -0.5f},
                c3[3] = {-1.85f, 0.58f, -0.5f}, c4[3] = {-1.85f, -0.42f, 0.5f}, c5[3] = {-0.85f, -0.42f, 0.5f},
                c6[3] = {-0.85f, 0.58f, 0.5f}, c7[3] = {-1.85f, 0.58f, 0.5f};
    const float* cube[12][3] = {{c4, c5, c6}, {c4, c6, c7}, {c0, c2, c1}, {c0, c3, c2}, {c0, c4, c7}, {c0, c7, c3},
        {c1, c2, c6}, {c1, c6, c5}, {c3, c7, c6}, {c3, c6, c2}, {c0, c1, c5}, {c0, c5, c4}};
    for (auto& t : cube) s.flatTriangles.push_back(makeTriangle(t[0], t[1], t[2], 0));
    const float p0[3] = {-0.183f, -0.42f, -0.683f}, p1[3] = {0.683f, -0.42f, -0.183f}, p2[3] = {0.183f, -0.42f, 0.683f},
                p3[3] = {-0.683f, -0.42f, 0.183f}, top[3] = {0.0f, 0.73f, 0.0f};
    const float* pyramid[6][3] = {{p0, p1, top}, {p1, p2, top}, {p2, p3, top}, {p3, p0, top}, {p0, p2, p1}, {p0, p3, p2}};
    for (auto& t : pyramid) s.flatTriangles.push_back(makeTriangle(t[0], t[1], t[2], 1));
    s.materials = {makeMaterial(WPT_MATERIAL_DIFFUSE, 0.7f, 0.2f, 0.15f), makeMaterial(WPT_MATERIAL_METAL, 0.85f, 0.65f, 0.25f),
        makeMaterial(WPT_MATERIAL_DIELECTRIC, 1.0f, 1.0f, 1.0f), makeMaterial(WPT_MATERIAL_DIFFUSE, 0.75f, 0.75f, 0.75f)};
    s.materials[1].roughness = 0.08f;
    s.materials[2].ior = 1.5f;
}

// Reference procedural caustic prism (scene_host.cpp:26-43).
std::vector<wpt_triangle> causticPrism() {
    const float h = 1.25f;
    const float a0[3] = {-h, -0.82f, -0.68f}, b0[3] = {-h, -0.82f, 0.68f}, c0[3] = {-h, 0.38f, 0.0f};
    const float a1[3] = {h, -0.82f, -0.68f}, b1[3] = {h, -0.82f, 0.68f}, c1[3] = {h, 0.38f, 0.0f};
    return {makeTriangle(a0, a1, b1, 0), makeTriangle(a0, b1, b0, 0), makeTriangle(b0, b1, c1, 0),
        makeTriangle(b0, c1, c0, 0), makeTriangle(c0, c1, a1, 0), makeTriangle(c0, a1, a0, 0),
        makeTriangle(a0, c0, b0, 0), makeTriangle(a1, b1, c1, 0)};
}

// rotateTransform about +Y (scene.cuh:72-89) for the prism profile.
void rotateY(wpt_transform& t, float degrees) {
    double angle = static_cast<double>(degrees) * 3.14159265358979323846 / 180.0;
    float cs = static_cast<float>(std::cos(angle)), sn = static_cast<float>(std::sin(angle));
    auto rot = [&](const float v[3], float out[3], float c, float s) {
        float x = v[0], z = v[2];
        out[0] = c * x + s * z;
        out[1] = v[1];
        out[2] = -s * x + c * z;
    };
    for (int i = 0; i < 3; ++i) {
        float tmp[3];
        rot(t.local_to_world[i], tmp, cs, sn);
        std::memcpy(t.local_to_world[i], tmp, sizeof tmp);
    }
    // Inverse basis: world_to_local columns rotate by -angle applied on the left.
    float inv[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) inv[i][j] = t.world_to_local[i][j];
    float e[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; ++i) {
        float r[3];
        rot(e[i], r, cs, -sn);
        for (int k = 0; k < 3; ++k)
            t.world_to_local[i][k] = inv[0][k] * r[0] + inv[1][k] * r[1] + inv[2][k] * r[2];
    }
}

}  // namespace

wpt_scene_desc HostScene::desc() const {
    wpt_scene_desc d{};
    d.spheres = spheres.data();
    d.sphere_count = static_cast<uint32_t>(spheres.size());
    d.static_triangles = staticTriangles.data();
    d.static_triangle_count = static_cast<uint32_t>(staticTriangles.size());
    d.flat_triangles = flatTriangles.data();
    d.flat_triangle_count = static_cast<uint32_t>(flatTriangles.size());
    d.mesh_triangles = meshTriangles.data();
    d.mesh_triangle_count = static_cast<uint32_t>(meshTriangles.size());
    d.meshes = meshes.data();
    d.mesh_count = static_cast<uint32_t>(meshes.size());
    d.instances = instances.data();
    d.instance_count = static_cast<uint32_t>(instances.size());
    d.materials = materials.data();
    d.material_count = static_cast<uint32_t>(materials.size());
    d.black_background = blackBackground ? 1u : 0u;
    d.ignore_direct_light_occlusion = ignoreDirectLightOcclusion ? 1u : 0u;
    d.punctual_lights = punctualLights.empty() ? nullptr : punctualLights.data();
    d.punctual_light_count = static_cast<uint32_t>(punctualLights.size());
    d.textures = textures.empty() ? nullptr : textures.data();
    d.texture_count = static_cast<uint32_t>(textures.size());
    d.material_ext = materialExt.empty() ? nullptr : materialExt.data();
    if (!environment.empty()) {
        d.environment_rgb = environment.data();
        d.environment_width = environmentWidth;
        d.environment_height = environmentHeight;
        d.environment_intensity = environmentIntensity;
        d.environment_rotation_degrees = environmentRotationDegrees;
    }
    d.use_normal_maps = useNormalMaps ? 1u : 0u;
    d.normal_map_minimum_cosine = normalMapMinimumCosine;
    d.flat_accelerator = flatAccelerator;
    return d;
}

wpt_material_ext plainMaterialExt() {
    wpt_material_ext e{};
    wpt_texture_binding none{};
    none.texture = WPT_INVALID_INDEX;
    none.scale[0] = none.scale[1] = 1.0f;
    e.albedo_texture = e.metallic_roughness_texture = e.emissive_texture = e.normal_texture = none;
    e.thickness_texture = e.specular_glossiness_texture = e.specular_weight_texture = none;
    e.specular_color_texture = e.clearcoat_texture = e.clearcoat_roughness_texture = none;
    e.clearcoat_normal_texture = none;
    e.normal_scale = 1.0f;
    e.specular_weight = 1.0f;
    e.specular_color[0] = e.specular_color[1] = e.specular_color[2] = 1.0f;
    e.clearcoat_normal_scale = 1.0f;
    return e;
}

wpt_material makeMaterial(uint32_t type, float r, float g, float b, float er, float eg, float eb) {
    wpt_material m{};
    m.type = type;
    set3(m.albedo, r, g, b);
    set3(m.emission, er, eg, eb);
    m.roughness = 0.0f;
    m.ior = 1.0f;
    m.metallic = 0.0f;
    m.opacity = 1.0f;
    m.alpha_cutoff = 0.5f;
    m.alpha_masked = 0;
    set3(m.attenuation_color, 1.0f, 1.0f, 1.0f);
    m.attenuation_distance = 0.0f;
    m.volume_density = 1.0f;
    set3(m.specular, 1.0f, 1.0f, 1.0f);
    m.glossiness = 1.0f;
    return m;
}

wpt_triangle makeTriangle(const float v0[3], const float v1[3], const float v2[3], uint32_t material) {
    wpt_triangle t{};
    std::memcpy(t.v0, v0, sizeof t.v0);
    std::memcpy(t.v1, v1, sizeof t.v1);
    std::memcpy(t.v2, v2, sizeof t.v2);
    t.material = material;
    return t;
}

wpt_transform makeScaledTransform(float tx, float ty, float tz, float scale) {
    wpt_transform t{};
    float inv = 1.0f / scale;
    for (int i = 0; i < 3; ++i) {
        t.local_to_world[i][i] = scale;
        t.world_to_local[i][i] = inv;
    }
    set3(t.translation, tx, ty, tz);
    return t;
}

std::vector<wpt_triangle> displacedIcosphere(uint32_t n, uint32_t seed, float amplitude) {
    const double phi = (1.0 + std::sqrt(5.0)) * 0.5;
    const double base[12][3] = {{-1, phi, 0}, {1, phi, 0}, {-1, -phi, 0}, {1, -phi, 0}, {0, -1, phi}, {0, 1, phi},
        {0, -1, -phi}, {0, 1, -phi}, {phi, 0, -1}, {phi, 0, 1}, {-phi, 0, -1}, {-phi, 0, 1}};
    const int faces[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
        {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9}, {4, 9, 5},
        {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    if (n == 0) n = 1;
    struct Vertex {
        double p[3];
        double n[3];
    };
    std::vector<Vertex> vertices;
    std::unordered_map<uint64_t, uint32_t> index;
    // A lattice point is keyed by its corner weights in corner-index order so
    // that points shared between faces are computed once, bit-identically.
    auto vertexOf = [&](int a, int b, int c, uint32_t wa, uint32_t wb, uint32_t wc) -> uint32_t {
        std::array<std::pair<int, uint32_t>, 3> w = {{{a, wa}, {b, wb}, {c, wc}}};
        std::sort(w.begin(), w.end());
        uint64_t key = 0;
        for (auto& e : w) key = (key << 20) | (e.second ? (static_cast<uint64_t>(e.first) << 16 | e.second) : 0u);
        auto it = index.find(key);
        if (it != index.end()) return it->second;
        double p[3] = {0, 0, 0};
        for (auto& e : w)
            for (int k = 0; k < 3; ++k) p[k] += base[e.first][k] * static_cast<double>(e.second);
        double len = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (double& v : p) v /= len;
        double r = 1.0 + static_cast<double>(amplitude) * fbm(p[0], p[1], p[2], seed);
        Vertex v{};
        for (int k = 0; k < 3; ++k) v.p[k] = static_cast<double>(static_cast<float>(p[k] * r));
        uint32_t id = static_cast<uint32_t>(vertices.size());
        vertices.push_back(v);
        index.emplace(key, id);
        return id;
    };
    std::vector<std::array<uint32_t, 3>> tris;
    tris.reserve(20ull * n * n);
    for (const auto& f : faces) {
        auto at = [&](uint32_t i, uint32_t j) { return vertexOf(f[0], f[1], f[2], n - i - j, i, j); };
        for (uint32_t i = 0; i < n; ++i) {
            for (uint32_t j = 0; j + i < n; ++j) {
                tris.push_back({at(i, j), at(i + 1, j), at(i, j + 1)});
                if (i + j + 1 < n) tris.push_back({at(i + 1, j), at(i + 1, j + 1), at(i, j + 1)});
            }
        }
    }
    // Outward winding and area-weighted smooth normals.
    for (auto& t : tris) {
        const double* a = vertices[t[0]].p;
        const double* b = vertices[t[1]].p;
        const double* c = vertices[t[2]].p;
        double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        double nn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        double centroid[3] = {a[0] + b[0] + c[0], a[1] + b[1] + c[1], a[2] + b[2] + c[2]};
        if (nn[0] * centroid[0] + nn[1] * centroid[1] + nn[2] * centroid[2] < 0) {
            std::swap(t[1], t[2]);
            for (double& v : nn) v = -v;
        }
        for (uint32_t k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) vertices[t[k]].n[d] += nn[d];
    }
    std::vector<wpt_triangle> out;
    out.reserve(tris.size());
    for (const auto& t : tris) {
        wpt_triangle tri{};
        float* vs[3] = {tri.v0, tri.v1, tri.v2};
        float* ns[3] = {tri.n0, tri.n1, tri.n2};
        for (int k = 0; k < 3; ++k) {
            const Vertex& v = vertices[t[k]];
            double len = std::sqrt(v.n[0] * v.n[0] + v.n[1] * v.n[1] + v.n[2] * v.n[2]);
            for (int d = 0; d < 3; ++d) {
                vs[k][d] = static_cast<float>(v.p[d]);
                ns[k][d] = static_cast<float>(v.n[d] / len);
            }
        }
        tri.flags = WPT_TRIANGLE_HAS_NORMALS;
        tri.material = 0;
        out.push_back(tri);
    }
    return out;
}

namespace {

wpt_transform objectTransform(const SceneConfig& c) {
    wpt_transform t = makeScaledTransform(c.objectTranslation.x, c.objectTranslation.y, c.objectTranslation.z,
        c.objectScale);
    if (c.lightProfile == CornellLightProfile::Prism) rotateY(t, -20.0f);
    return t;
}

void scaleInstance(wpt_instance& inst, float scale, const float translation[3]) {
    // Vec3 * s and Vec3 / s (= Vec3 * (1 / s)) on the basis, s * t + offset
    // on the translation (scene_host.cpp:299-307, 358-366).
    float inverse = 1.0f / scale;
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k) {
            inst.transform.local_to_world[c][k] = scale * inst.transform.local_to_world[c][k];
            inst.transform.world_to_local[c][k] = inst.transform.world_to_local[c][k] * inverse;
        }
    for (int k = 0; k < 3; ++k) inst.transform.translation[k] = scale * inst.transform.translation[k] + translation[k];
}

void adoptTextures(HostScene& s, std::vector<std::vector<uint8_t>>& data,
    const std::vector<std::pair<uint32_t, uint32_t>>& sizes) {
    s.textureData = std::move(data);
    s.textures.clear();
    for (size_t i = 0; i < s.textureData.size(); ++i)
        s.textures.push_back(wpt_texture{sizes[i].first, sizes[i].second, s.textureData[i].data()});
}

// Material extensions of OBJ materials: their albedo textures.
std::vector<wpt_material_ext> objMaterialExt(const ObjScene& obj) {
    std::vector<wpt_material_ext> ext(obj.materials.size(), plainMaterialExt());
    for (size_t i = 0; i < ext.size(); ++i) ext[i].albedo_texture.texture = obj.albedoTexture[i];
    return ext;
}

// The Cornell room around its object (assembleCornellScene,
// scene_host.cpp:242-350): the procedural prism, the meshes of an OBJ file
// (one instance each) or a glTF scene (its node instances, scaled and
// moved, without mesh 0's when the backdrop is removed, and its punctual
// lights).  Object materials follow the room's five, optionally turned into
// dispersive glass; the prism keeps its own material.
bool cornellScene(HostScene& s, const SceneConfig& c, std::string* error) {
    std::vector<wpt_material> objectMaterials;
    std::vector<wpt_material_ext> objectExt;
    const float offset[3] = {c.objectTranslation.x, c.objectTranslation.y, c.objectTranslation.z};
    auto addMesh = [&](std::vector<wpt_triangle>& mesh) {
        uint32_t first = static_cast<uint32_t>(s.meshTriangles.size());
        for (wpt_triangle& t : mesh) {
            t.material = c.objectMaterialOverride >= 0 ? static_cast<uint32_t>(c.objectMaterialOverride) : t.material + 5;
            s.meshTriangles.push_back(t);
        }
        s.meshes.push_back(wpt_mesh{first, static_cast<uint32_t>(mesh.size())});
    };
    if (c.objectSource == CornellObjectSource::ProceduralPrism) {
        std::vector<wpt_triangle> prism = causticPrism();
        addMesh(prism);
        wpt_material glass = makeMaterial(WPT_MATERIAL_DIELECTRIC, 0.98f, 0.99f, 1.0f);
        glass.ior = 1.52f;
        s.materials.push_back(glass);
    } else if (c.objectSource == CornellObjectSource::Gltf) {
        GltfScene g;
        if (!loadGltfScene(c.objectFilename, g, error)) return false;
        for (std::vector<wpt_triangle>& mesh : g.meshes) addMesh(mesh);
        for (wpt_instance inst : g.instances) {
            if (c.removeBackdrop && inst.mesh == 0) continue;
            scaleInstance(inst, c.objectScale, offset);
            s.instances.push_back(inst);
        }
        for (wpt_punctual_light light : g.punctualLights) {
            for (int k = 0; k < 3; ++k) light.position[k] = c.objectScale * light.position[k] + offset[k];
            s.punctualLights.push_back(light);
        }
        objectMaterials = std::move(g.materials);
        objectExt = std::move(g.materialExt);
        adoptTextures(s, g.textureData, g.textureSizes);
    } else {
        ObjScene obj;
        if (!loadObjScene(c.objectFilename, obj, error)) return false;
        for (std::vector<wpt_triangle>& mesh : obj.meshes) addMesh(mesh);
        objectExt = objMaterialExt(obj);
        objectMaterials = std::move(obj.materials);
        adoptTextures(s, obj.textureData, obj.textureSizes);
    }
    if (c.convertObjectMaterialsToDielectric) {
        for (size_t i = 0; i < objectMaterials.size(); ++i) {  // scene_host.cpp:273-282
            wpt_material& m = objectMaterials[i];
            m.type = WPT_MATERIAL_DIELECTRIC;
            set3(m.albedo, 1.0f, 1.0f, 1.0f);
            set3(m.emission, 0.0f, 0.0f, 0.0f);
            m.roughness = 0.0f;
            m.ior = 1.52f;
            objectExt[i].dispersion = 0.02f;
            objectExt[i].albedo_texture.texture = WPT_INVALID_INDEX;
        }
    }
    if (c.objectSource != CornellObjectSource::Gltf) {
        for (uint32_t mesh = 0; mesh < s.meshes.size(); ++mesh) {
            wpt_instance inst{};
            inst.mesh = mesh;
            inst.material_override = WPT_INVALID_INDEX;
            inst.acceleration = WPT_ACCEL_BVH;
            inst.transform = objectTransform(c);
            s.instances.push_back(inst);
        }
    }
    bool prism = c.objectSource == CornellObjectSource::ProceduralPrism;
    s.materials.insert(s.materials.end(), objectMaterials.begin(), objectMaterials.end());
    // The light follows the dielectric switch first, then the profile
    // (scene_host.cpp:322-326).
    referenceCornellRoom(s, c, c.convertObjectMaterialsToDielectric, c.lightProfile == CornellLightProfile::Prism);
    s.materialExt.assign(5, plainMaterialExt());
    if (prism) {
        // The prism splits photons by channel (scene_host.cpp:251-255).
        s.materialExt.push_back(plainMaterialExt());
        s.materialExt.back().dispersion = 0.035f;
    }
    s.materialExt.insert(s.materialExt.end(), objectExt.begin(), objectExt.end());
    return true;
}

// A glTF scene under the sky (assembleGltfScene, scene_host.cpp:351-385):
// its instances and punctual lights scaled by the scene scale, and an
// optional 8 x 8 emissive quad above it.
bool gltfScene(HostScene& s, const SceneConfig& c, std::string* error) {
    GltfScene g;
    if (!loadGltfScene(c.gltfFilename, g, error)) return false;
    if (g.meshes.empty() || g.instances.empty()) {
        if (error) *error = "glTF scene contains no renderable mesh instances";
        return false;
    }
    for (const std::vector<wpt_triangle>& mesh : g.meshes) {
        s.meshes.push_back(wpt_mesh{static_cast<uint32_t>(s.meshTriangles.size()), static_cast<uint32_t>(mesh.size())});
        s.meshTriangles.insert(s.meshTriangles.end(), mesh.begin(), mesh.end());
    }
    const float zero[3] = {0.0f, 0.0f, 0.0f};
    for (wpt_instance inst : g.instances) {
        scaleInstance(inst, c.gltfScale, zero);
        s.instances.push_back(inst);
    }
    for (wpt_punctual_light light : g.punctualLights) {
        for (float& v : light.position) v = c.gltfScale * v;
        s.punctualLights.push_back(light);
    }
    s.materials = std::move(g.materials);
    s.materialExt = std::move(g.materialExt);
    adoptTextures(s, g.textureData, g.textureSizes);
    if (c.addSponzaTopLight) {
        uint32_t light = static_cast<uint32_t>(s.materials.size());
        s.materials.push_back(makeMaterial(WPT_MATERIAL_EMISSIVE, 1.0f, 1.0f, 1.0f, 80.0f, 80.0f, 80.0f));
        s.materialExt.push_back(plainMaterialExt());
        const float a[3] = {-4.0f, 9.2f, -4.0f}, b[3] = {4.0f, 9.2f, -4.0f}, cc[3] = {4.0f, 9.2f, 4.0f},
                    d[3] = {-4.0f, 9.2f, 4.0f};
        addQuad(s.staticTriangles, a, b, cc, d, light);
    }
    s.ignoreDirectLightOcclusion = c.ignoreSponzaLightOcclusion;
    return true;
}

// An OBJ object over a ground sphere (assembleObjScene,
// scene_host.cpp:391-418): the TLAS/BLAS policy instances the mesh, the
// others bake the transform into directly tested triangles.
bool objScene(HostScene& s, const SceneConfig& c, std::string* error) {
    ObjScene obj;
    if (!loadObjScene(c.objectFilename, obj, error)) return false;
    s.spheres.push_back(sphere(0.0f, -100.5f, 0.0f, 100.0f, 0, WPT_ACCEL_DIRECT));
    s.materials.push_back(makeMaterial(WPT_MATERIAL_DIFFUSE, 0.75f, 0.75f, 0.75f));
    s.materials.insert(s.materials.end(), obj.materials.begin(), obj.materials.end());
    if (!obj.textureData.empty()) {
        s.materialExt.assign(1, plainMaterialExt());
        std::vector<wpt_material_ext> ext = objMaterialExt(obj);
        s.materialExt.insert(s.materialExt.end(), ext.begin(), ext.end());
        adoptTextures(s, obj.textureData, obj.textureSizes);
    }
    std::vector<wpt_triangle> all;
    for (const std::vector<wpt_triangle>& mesh : obj.meshes) all.insert(all.end(), mesh.begin(), mesh.end());
    for (wpt_triangle& t : all) t.material += 1;
    if (c.acceleration == ObjAccelerationPolicy::TlasBlasBvh2) {
        s.meshTriangles = std::move(all);
        s.meshes.push_back(wpt_mesh{0, static_cast<uint32_t>(s.meshTriangles.size())});
        wpt_instance inst{};
        inst.mesh = 0;
        inst.material_override = WPT_INVALID_INDEX;
        inst.acceleration = WPT_ACCEL_BVH;
        inst.transform = makeScaledTransform(c.objectTranslation.x, c.objectTranslation.y, c.objectTranslation.z,
            c.objectScale);
        s.instances.push_back(inst);
    } else {
        const float tr[3] = {c.objectTranslation.x, c.objectTranslation.y, c.objectTranslation.z};
        for (wpt_triangle& t : all) {
            for (float* v : {t.v0, t.v1, t.v2})
                for (int k = 0; k < 3; ++k) v[k] = c.objectScale * v[k] + tr[k];
        }
        s.flatTriangles = std::move(all);
        // The reference's flat accelerators (scene.cu:241-294): NexusBVH's
        // BVH2 and BVH8, the median BVH and the spatial-split BVH.
        s.flatAccelerator = c.acceleration == ObjAccelerationPolicy::NexusBvh8 ? WPT_FLAT_WIDE8
            : c.acceleration == ObjAccelerationPolicy::MedianBvh ? WPT_FLAT_MEDIAN_BVH2
            : c.acceleration == ObjAccelerationPolicy::SpatialSplitBvh ? WPT_FLAT_SPATIAL_SPLIT_BVH2
                                                                       : WPT_FLAT_SAH_BVH2;
    }
    return true;
}

}  // namespace

namespace {
bool buildPreset(const SceneConfig& config, HostScene& out, std::string* error) {
    switch (config.preset) {
    case ScenePreset::Demo:
        demoScene(out);
        return true;
    case ScenePreset::CornellBox:
        cornellBoxScene(out);
        return true;
    case ScenePreset::DiffCornell:
        diffCornellScene(out);
        return true;
    case ScenePreset::Terrain:
        terrainScene(out, config);
        return true;
    case ScenePreset::SphereField:
        sphereFieldScene(out, config, true);
        return true;
    case ScenePreset::RtWeek:
        sphereFieldScene(out, config, false);
        return true;
    case ScenePreset::Cornell:
    case ScenePreset::Hurricane:
    case ScenePreset::Prism:
    case ScenePreset::Crystal:
        return cornellScene(out, config, error);
    case ScenePreset::Sponza:
        return gltfScene(out, config, error);
    case ScenePreset::Deer:
        return objScene(out, config, error);
    default:
        break;
    }
    if (error) *error = std::string("unsupported scene preset '") + presetName(config.preset) + "'";
    return false;
}
}  // namespace

bool createSceneFromConfig(const SceneConfig& config, HostScene& out, std::string* error) {
    out = HostScene{};
    if (!buildPreset(config, out, error)) return false;
    out.useNormalMaps = config.useNormalMaps;
    out.normalMapMinimumCosine = config.normalMapMinimumCosine;
    if (!config.environmentFilename.empty()) {
        // Any preset may light with a panorama (applySceneConfig,
        // scene_factory.cpp:92-104).
        if (!loadRadianceHdr(config.environmentFilename, out.environment, out.environmentWidth,
                out.environmentHeight, error))
            return false;
        out.environmentIntensity = config.environmentIntensity;
        out.environmentRotationDegrees = config.environmentRotationDegrees;
    }
    return true;
}

namespace {
wpt_camera lookAt(const float pos[3], const float target[3], float fovDegrees, uint32_t w, uint32_t h, bool guardUp) {
    // camera.cu:3-42: forward = normalize(target - position); right = normalize(cross(forward, up)); up = cross(right, forward)
    float f[3] = {target[0] - pos[0], target[1] - pos[1], target[2] - pos[2]};
    float fl = 1.0f / std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (float& v : f) v *= fl;
    float up[3] = {0.0f, 1.0f, 0.0f};
    if (guardUp && std::fabs(f[1]) > 0.999f) set3(up, 0.0f, 0.0f, 1.0f);
    float r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    float rl = 1.0f / std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (float& v : r) v *= rl;
    float u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    wpt_camera c{};
    std::memcpy(c.position, pos, sizeof c.position);
    std::memcpy(c.forward, f, sizeof f);
    std::memcpy(c.right, r, sizeof r);
    std::memcpy(c.up, u, sizeof u);
    c.vertical_fov = fovDegrees * 3.14159265359f / 180.0f;
    c.aspect_ratio = static_cast<float>(w) / static_cast<float>(h);
    return c;
}
}  // namespace

wpt_camera createSceneCamera(const SceneConfig& config, uint32_t w, uint32_t h) {
    switch (config.preset) {
    case ScenePreset::RtWeek:
    case ScenePreset::SphereField: {
        const float p[3] = {13.0f, 2.0f, 3.0f}, t[3] = {0.0f, 0.8f, 0.0f};
        return lookAt(p, t, 32.0f, w, h, true);
    }
    case ScenePreset::Cornell:
    case ScenePreset::Hurricane:
    case ScenePreset::Prism:
    case ScenePreset::Crystal: {
        const float p[3] = {0.0f, 1.2f, 6.5f}, t[3] = {0.0f, 0.8f, -0.6f};
        return lookAt(p, t, 42.0f, w, h, false);
    }
    case ScenePreset::CornellBox:
    case ScenePreset::DiffCornell: {
        const float p[3] = {0.0f, 0.0f, 3.4f}, t[3] = {0.0f, 0.0f, 0.0f};
        return lookAt(p, t, 40.0f, w, h, true);
    }
    case ScenePreset::Terrain: {
        const float p[3] = {-5.5f, 2.6f, 6.5f}, t[3] = {0.0f, 0.4f, 0.0f};
        return lookAt(p, t, 50.0f, w, h, true);
    }
    default: {
        const float p[3] = {-10.0f, 0.35f, 0.8f}, t[3] = {6.0f, 3.2f, 2.4f};
        return lookAt(p, t, 45.0f, w, h, true);
    }
    }
}

}  // namespace wpt
