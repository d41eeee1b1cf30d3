// oracle.cpp - TEST INFRASTRUCTURE ONLY.  A scalar C++ restatement of the
// reference wavefront path tracer's render path, used as the parity checker
// for the gfx950 kernels.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it; the product library never links it.
//
// Each function names the reference file:line it restates.  It is written in
// the reference's own structure (own-bounds BVH nodes re-tested on pop, hit
// attributes computed at every improvement, one path traced depth-first) and
// templated on the real type: R = float is the FP32 restatement ("oracle A"),
// R = double the FP64 variant ("oracle B").
//
// Shared dependency: elementary functions (sin, cos, tan, exp, log, pow) come
// from the product's deterministic wpt_math.h, because libm and the device
// library differ in the last ulp; that header is pinned separately against
// libm (tests/test_cpu_oracle.py).  Everything else here is independent code.
//
// Pinning summary (see tests/): RNG known answers and seed uniqueness; Fresnel
// and GGX closed forms from the reference's material tests; traversal vs the
// independent FP64 brute force on the reference's mt19937 ray set; Walker
// alias invariants; whole images against closed forms (an emissive
// enclosure, a mirror under an emitting plane, a Lambertian floor under a
// uniform emitting hemisphere).
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <random>
#include <thread>
#include <vector>

#include "../wavefront-path-tracer_amd/csrc/common/wpt_math.h"

namespace {

constexpr uint32_t kInvalid = 0xffffffffu;

// ---------------------------------------------------------------- vectors
template <class R>
struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    V3 operator-() const { return V3(-x, -y, -z); }
    V3 operator+(const V3& v) const { return V3(x + v.x, y + v.y, z + v.z); }
    V3 operator-(const V3& v) const { return V3(x - v.x, y - v.y, z - v.z); }
    V3 operator*(R s) const { return V3(x * s, y * s, z * s); }
    V3 operator/(R s) const {
        R inv = R(1) / s;
        return *this * inv;
    }
    V3& operator+=(const V3& v) {
        x += v.x;
        y += v.y;
        z += v.z;
        return *this;
    }
    V3& operator*=(R s) {
        x *= s;
        y *= s;
        z *= s;
        return *this;
    }
};
template <class R> V3<R> operator*(R s, const V3<R>& v) { return v * s; }
template <class R> R dot(const V3<R>& a, const V3<R>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> V3<R> cross(const V3<R>& a, const V3<R>& b) {
    return V3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> R lengthSquared(const V3<R>& v) { return dot(v, v); }
template <class R> R length(const V3<R>& v) { return std::sqrt(lengthSquared(v)); }
template <class R> V3<R> normalize(const V3<R>& v) {
    R inv = R(1) / std::sqrt(dot(v, v));
    return v * inv;
}
template <class R> V3<R> hadamard(const V3<R>& a, const V3<R>& b) { return V3<R>(a.x * b.x, a.y * b.y, a.z * b.z); }

template <class R> struct V2 { R x, y; };

template <class R> V3<R> vec(const float v[3]) { return V3<R>(R(v[0]), R(v[1]), R(v[2])); }

template <class R>
struct Ray {
    V3<R> o, d;
    V3<R> at(R t) const { return o + d * t; }
};

// ---------------------------------------------------------------- elementary functions
inline float fsin(float x) { return wpt::sinF(x); }
inline float fcos(float x) { return wpt::cosF(x); }
inline double fsin(double x) { double s, c; wpt::detail::sinCosD(x, s, c); return s; }
inline double fcos(double x) { double s, c; wpt::detail::sinCosD(x, s, c); return c; }
inline float ftan(float x) { return wpt::tanF(x); }
inline double ftan(double x) { return fsin(x) / fcos(x); }
inline float fexp(float x) { return wpt::expF(x); }
inline double fexp(double x) { return wpt::expD(x); }
inline float flog(float x) { return wpt::logF(x); }
inline double flog(double x) { return wpt::logD(x); }
inline float fpow(float x, float y) { return wpt::powF(x, y); }
inline double fpow(double x, double y) {
    if (y == 0.0) return 1.0;
    if (x == 0.0) return y > 0.0 ? 0.0 : INFINITY;
    if (x == 1.0) return 1.0;
    return wpt::expD(y * wpt::logD(x));
}

// ---------------------------------------------------------------- RNG (include/core/rng.cuh, src/core/rng.cu)
inline uint64_t splitMix64(uint64_t v) {
    v += 0x9e3779b97f4a7c15ull;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
    return v ^ (v >> 31);
}

struct RngState {
    uint64_t state;
    uint32_t strategy;
};

RngState makeRngState(uint64_t key, uint32_t strategy) {  // rng.cuh:16-29
    RngState r;
    r.strategy = strategy;
    if (strategy == WPT_RNG_XORSHIFT32) {
        uint32_t seed = static_cast<uint32_t>(key) * 747796405u + 2891336453u;
        r.state = seed == 0 ? 1u : seed;
    } else {
        r.state = splitMix64(key);
    }
    return r;
}

uint64_t pixelRngKey(uint32_t pixel, uint32_t sample, uint32_t strategy) {  // rng.cuh:32-41
    if (strategy == WPT_RNG_XORSHIFT32) return static_cast<uint64_t>(pixel ^ (sample * 0x9e3779b9u));
    return (static_cast<uint64_t>(sample) << 32) | pixel;
}

uint64_t photonRngKey(uint32_t photon, uint32_t seed, uint32_t strategy) {  // rng.cuh:43-51
    if (strategy == WPT_RNG_XORSHIFT32) return static_cast<uint64_t>(seed ^ (photon * 0x9e3779b9u));
    return (static_cast<uint64_t>(seed) << 32) | photon;
}

float randomFloat(RngState& s) {  // rng.cu:30-41
    uint32_t value;
    if (s.strategy == WPT_RNG_XORSHIFT32) {
        uint32_t x = static_cast<uint32_t>(s.state);
        x ^= x << 13;
        x ^= x >> 17;
        x ^= x << 5;
        s.state = x;
        value = x;
    } else {
        uint64_t old = s.state;
        s.state = old * 6364136223846793005ull + 1442695040888963407ull;
        uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = static_cast<uint32_t>(old >> 59u);
        value = (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    return static_cast<float>(value >> 8) * 0x1.0p-24f;
}

template <class R> R rnd(RngState& s) { return R(randomFloat(s)); }

// ---------------------------------------------------------------- scene tables
struct TriangleLight {
    V3<float> v0, v1, v2, normal;
    float area;
    uint32_t material;
    float selectionPdf;
};

struct AliasEntry {
    float probability;
    uint32_t alias;
};

struct Node {  // NXB::BVH2::Node as used by device_traversal.cuh
    float bmin[3], bmax[3];
    uint32_t left, right;
};

struct BvhView {
    const wpt_bvh2_node* nodes = nullptr;
    uint32_t count = 0;
};

}  // namespace

struct oracle_handle {
    wpt_scene_desc d;
    BvhView tlas, flat;
    std::vector<BvhView> blas;
    std::vector<uint32_t> directSpheres, directInstances, leaves;
    std::vector<uint32_t> staticLight;  // light index per static triangle
    std::vector<uint32_t> meshLight;    // per mesh triangle (mesh-local numbering)
    std::vector<uint32_t> instanceLightOffset;
    std::vector<TriangleLight> lights;
    std::vector<float> lightWeights;
    std::vector<AliasEntry> alias;
    // Environment map: texels (3 floats each), importance table, alias table.
    bool env = false;
    uint32_t envW = 0, envH = 0;
    float envIntensity = 1.0f, envRotation = 0.0f;
    std::vector<float> envRgb, envProb;
    std::vector<AliasEntry> envAlias;
    // Punctual lights: directional ones first (scene.cu:349-355).
    std::vector<wpt_punctual_light> punctual;
    uint32_t directional = 0;
    // Textures (RGBA8 rows) and every material's texture bindings (the
    // reference's defaults where the scene gives none).
    std::vector<std::vector<uint8_t>> tex;
    std::vector<uint32_t> texW, texH;
    std::vector<wpt_material_ext> ext;
};

namespace {

const wpt_material& mat(const oracle_handle& h, uint32_t i) { return h.d.materials[i]; }

// Luminance power registration (scene_host.cpp:60-106).
void registerLight(oracle_handle& h, const V3<float>& v0, const V3<float>& v1, const V3<float>& v2, uint32_t material) {
    V3<float> normal = normalize(cross(v1 - v0, v2 - v0));
    float area = 0.5f * length(cross(v1 - v0, v2 - v0));
    const wpt_material& m = mat(h, material);
    float power = area * (0.2126f * m.emission[0] + 0.7152f * m.emission[1] + 0.0722f * m.emission[2]);
    h.lights.push_back(TriangleLight{v0, v1, v2, normal, area, material, 0.0f});
    h.lightWeights.push_back(power);
}

V3<float> xform(const wpt_transform& t, const V3<float>& v) {  // transformVector, scene.cuh:38-41
    return vec<float>(t.local_to_world[0]) * v.x + vec<float>(t.local_to_world[1]) * v.y +
        vec<float>(t.local_to_world[2]) * v.z;
}

// Texel solid angle of row y, environment.cuh:31-35: the FP64 product
// (4 pi / w) sin(pi (y + 1/2) / h) sin(pi / 2h), rounded to FP32.
float envSolidAngle(uint32_t y, uint32_t w, uint32_t h) {
    const double pi = 3.14159265358979323846;
    return static_cast<float>((4.0 * pi / w) * fsin(pi * (y + 0.5) / h) * fsin(pi / (2.0 * h)));
}

// buildEnvironmentDistribution, src/scene/environment.cpp:10-64.  Weight of
// a texel: luminance under the separable [1/8 3/4 1/8] tent (u wraps, v
// clamps) times its solid angle, all in FP64; probabilities 0.99 weight
// share + 0.01 solid-angle share rounded to FP32, renormalized by their FP64
// sum; Walker alias over FP64 scaled masses with LIFO worklists.
void buildEnvironment(oracle_handle& h) {
    const wpt_scene_desc& d = h.d;
    if (!d.environment_rgb) return;
    const uint32_t w = d.environment_width, ht = d.environment_height, n = w * ht;
    h.env = true;
    h.envW = w;
    h.envH = ht;
    h.envIntensity = d.environment_intensity;
    h.envRotation = std::fmod(d.environment_rotation_degrees, 360.0f) * (3.14159265359f / 180.0f);
    h.envRgb.assign(d.environment_rgb, d.environment_rgb + 3ull * n);
    const double tent[3] = {0.125, 0.75, 0.125};
    std::vector<double> mass(n);
    double total = 0.0;
    for (uint32_t y = 0; y < ht; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            double lum = 0.0;
            for (int dy = -1; dy <= 1; ++dy) {
                int yy = std::min<int>(std::max<int>(static_cast<int>(y) + dy, 0), static_cast<int>(ht) - 1);
                for (int dx = -1; dx <= 1; ++dx) {
                    uint32_t xx = static_cast<uint32_t>((static_cast<int64_t>(x) + w + dx) % w);
                    const float* p = &h.envRgb[3ull * (static_cast<uint64_t>(yy) * w + xx)];
                    lum += tent[dy + 1] * tent[dx + 1] *
                        (0.2126 * double(p[0]) + 0.7152 * double(p[1]) + 0.0722 * double(p[2]));
                }
            }
            mass[y * w + x] = lum * envSolidAngle(y, w, ht);
            total += mass[y * w + x];
        }
    }
    h.envProb.resize(n);
    double sum = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        double uniform = envSolidAngle(i / w, w, ht) / 12.5663706143591729538;
        h.envProb[i] = static_cast<float>(total > 0 ? 0.99 * mass[i] / total + 0.01 * uniform : uniform);
        sum += h.envProb[i];
    }
    std::vector<uint32_t> small, large;
    for (uint32_t i = 0; i < n; ++i) {
        h.envProb[i] = static_cast<float>(h.envProb[i] / sum);
        mass[i] = double(h.envProb[i]) * n;
        (mass[i] < 1.0 ? small : large).push_back(i);
    }
    h.envAlias.assign(n, AliasEntry{1.0f, 0});
    while (!small.empty() && !large.empty()) {
        uint32_t lo = small.back(), hi = large.back();
        small.pop_back();
        large.pop_back();
        h.envAlias[lo] = AliasEntry{static_cast<float>(mass[lo]), hi};
        mass[hi] += mass[lo] - 1.0;
        (mass[hi] < 1.0 ? small : large).push_back(hi);
    }
    for (uint32_t i : small) h.envAlias[i] = AliasEntry{1.0f, i};
    for (uint32_t i : large) h.envAlias[i] = AliasEntry{1.0f, i};
}

// Reference TextureBinding / Material defaults (material.cuh:14-70).
wpt_material_ext defaultExt() {
    wpt_material_ext e{};
    wpt_texture_binding* all[] = {&e.albedo_texture, &e.metallic_roughness_texture, &e.emissive_texture,
        &e.normal_texture, &e.thickness_texture, &e.specular_glossiness_texture, &e.specular_weight_texture,
        &e.specular_color_texture, &e.clearcoat_texture, &e.clearcoat_roughness_texture, &e.clearcoat_normal_texture};
    for (wpt_texture_binding* b : all) {
        b->texture = kInvalid;
        b->scale[0] = b->scale[1] = 1.0f;
    }
    e.normal_scale = 1.0f;
    e.specular_weight = 1.0f;
    e.specular_color[0] = e.specular_color[1] = e.specular_color[2] = 1.0f;
    e.clearcoat_normal_scale = 1.0f;
    return e;
}

void buildTextures(oracle_handle& h) {
    const wpt_scene_desc& d = h.d;
    for (uint32_t i = 0; i < d.texture_count; ++i) {
        const wpt_texture& t = d.textures[i];
        h.tex.emplace_back(t.rgba, t.rgba + 4ull * t.width * t.height);
        h.texW.push_back(t.width);
        h.texH.push_back(t.height);
    }
    for (uint32_t m = 0; m < d.material_count; ++m) h.ext.push_back(d.material_ext ? d.material_ext[m] : defaultExt());
}

void buildTables(oracle_handle& h) {
    buildEnvironment(h);
    buildTextures(h);
    for (uint32_t i = 0; i < h.d.punctual_light_count; ++i)
        if (h.d.punctual_lights[i].type == WPT_LIGHT_DIRECTIONAL) h.punctual.push_back(h.d.punctual_lights[i]);
    h.directional = static_cast<uint32_t>(h.punctual.size());
    for (uint32_t i = 0; i < h.d.punctual_light_count; ++i)
        if (h.d.punctual_lights[i].type != WPT_LIGHT_DIRECTIONAL) h.punctual.push_back(h.d.punctual_lights[i]);
    const wpt_scene_desc& d = h.d;
    // Mesh light numbering, then per-instance world lights (scene_host.cpp:77-106).
    h.meshLight.assign(d.mesh_triangle_count, kInvalid);
    for (uint32_t m = 0; m < d.mesh_count; ++m) {
        uint32_t count = 0;
        for (uint32_t t = 0; t < d.meshes[m].triangle_count; ++t) {
            uint32_t gi = d.meshes[m].first_triangle + t;
            if (mat(h, d.mesh_triangles[gi].material).type == WPT_MATERIAL_EMISSIVE) h.meshLight[gi] = count++;
        }
    }
    uint32_t offset = 0;
    for (uint32_t i = 0; i < d.instance_count; ++i) {  // scene.cu:156-166
        h.instanceLightOffset.push_back(offset);
        const wpt_mesh& m = d.meshes[d.instances[i].mesh];
        for (uint32_t t = 0; t < m.triangle_count; ++t)
            if (h.meshLight[m.first_triangle + t] != kInvalid) ++offset;
    }
    for (uint32_t i = 0; i < d.instance_count; ++i) {
        const wpt_instance& inst = d.instances[i];
        const wpt_mesh& m = d.meshes[inst.mesh];
        for (uint32_t t = 0; t < m.triangle_count; ++t) {
            uint32_t gi = m.first_triangle + t;
            if (h.meshLight[gi] == kInvalid) continue;
            const wpt_triangle& tri = d.mesh_triangles[gi];
            V3<float> tr = vec<float>(inst.transform.translation);
            registerLight(h, xform(inst.transform, vec<float>(tri.v0)) + tr, xform(inst.transform, vec<float>(tri.v1)) + tr,
                xform(inst.transform, vec<float>(tri.v2)) + tr, tri.material);
        }
    }
    h.staticLight.assign(d.static_triangle_count, kInvalid);
    for (uint32_t i = 0; i < d.static_triangle_count; ++i) {
        const wpt_triangle& tri = d.static_triangles[i];
        if (mat(h, tri.material).type != WPT_MATERIAL_EMISSIVE) continue;
        h.staticLight[i] = static_cast<uint32_t>(h.lights.size());
        registerLight(h, vec<float>(tri.v0), vec<float>(tri.v1), vec<float>(tri.v2), tri.material);
    }
    // Walker alias (scene.cu:86-114).
    size_t n = h.lights.size();
    h.alias.assign(n, AliasEntry{0.0f, 0});
    if (n) {
        float total = 0.0f;
        for (float w : h.lightWeights) total += w;
        std::vector<float> scaled(n);
        std::vector<uint32_t> small, large;
        for (uint32_t i = 0; i < n; ++i) {
            h.lights[i].selectionPdf = h.lightWeights[i] / total;
            scaled[i] = h.lights[i].selectionPdf * n;
            (scaled[i] < 1.0f ? small : large).push_back(i);
        }
        while (!small.empty() && !large.empty()) {
            uint32_t low = small.back();
            small.pop_back();
            uint32_t high = large.back();
            large.pop_back();
            h.alias[low] = AliasEntry{scaled[low], high};
            scaled[high] = scaled[high] + scaled[low] - 1.0f;
            (scaled[high] < 1.0f ? small : large).push_back(high);
        }
        for (uint32_t i : small) h.alias[i] = AliasEntry{1.0f, i};
        for (uint32_t i : large) h.alias[i] = AliasEntry{1.0f, i};
    }
    // Partition (scene.cu:175-239).
    for (uint32_t i = 0; i < d.sphere_count; ++i) {
        if (d.spheres[i].acceleration == WPT_ACCEL_DIRECT) h.directSpheres.push_back(i);
        else h.leaves.push_back(0x80000000u | i);
    }
    for (uint32_t i = 0; i < d.instance_count; ++i) {
        if (d.instances[i].acceleration == WPT_ACCEL_DIRECT) {
            h.directInstances.push_back(i);
            continue;
        }
        if (d.meshes[d.instances[i].mesh].triangle_count == 0) continue;
        h.leaves.push_back(i);
    }
}

// ---------------------------------------------------------------- work counters
// Per-query work of the reference algorithm (node bound tests, triangle and
// sphere tests, instance entries), used for the roofline's byte model.
struct WorkCounts {
    uint64_t nodes = 0, triangles = 0, spheres = 0, instances = 0;
};
thread_local WorkCounts* tlWork = nullptr;
inline void countNode() { if (tlWork) ++tlWork->nodes; }
inline void countTriangle() { if (tlWork) ++tlWork->triangles; }
inline void countSphere() { if (tlWork) ++tlWork->spheres; }
inline void countInstance() { if (tlWork) ++tlWork->instances; }

// ---------------------------------------------------------------- intersection tests
template <class R>
bool intersectSphere(const Ray<R>& ray, const wpt_sphere& s, R& t) {  // device_traversal.cuh:24-53
    countSphere();
    V3<R> oc = ray.o - vec<R>(s.center);
    R a = dot(ray.d, ray.d);
    R b = R(2) * dot(oc, ray.d);
    R c = dot(oc, oc) - R(s.radius) * R(s.radius);
    R disc = b * b - R(4) * a * c;
    if (disc < R(0)) return false;
    R sq = std::sqrt(disc);
    R t0 = (-b - sq) / (R(2) * a);
    R t1 = (-b + sq) / (R(2) * a);
    const R tMin = R(1.0e-6f);
    if (t0 > tMin) {
        t = t0;
        return true;
    }
    if (t1 > tMin) {
        t = t1;
        return true;
    }
    return false;
}

template <class R>
bool intersectTriangle(const Ray<R>& ray, const wpt_triangle& tri, R& t, R& u, R& v) {  // :55-91
    countTriangle();
    V3<R> v0 = vec<R>(tri.v0), v1 = vec<R>(tri.v1), v2 = vec<R>(tri.v2);
    V3<R> e1 = v1 - v0;
    V3<R> e2 = v2 - v0;
    V3<R> p = cross(ray.d, e2);
    R det = dot(e1, p);
    R scale = std::fabs(e1.x) * (std::fabs(ray.d.y * e2.z) + std::fabs(ray.d.z * e2.y)) +
        std::fabs(e1.y) * (std::fabs(ray.d.z * e2.x) + std::fabs(ray.d.x * e2.z)) +
        std::fabs(e1.z) * (std::fabs(ray.d.x * e2.y) + std::fabs(ray.d.y * e2.x));
    if (!std::isfinite(det) || std::fabs(det) <= R(8) * std::numeric_limits<R>::epsilon() * scale) return false;
    R inv = R(1) / det;
    V3<R> ov = ray.o - v0;
    u = dot(ov, p) * inv;
    if (u < R(0) || u > R(1)) return false;
    V3<R> q = cross(ov, e1);
    v = dot(ray.d, q) * inv;
    if (v < R(0) || u + v > R(1)) return false;
    t = dot(e2, q) * inv;
    return t > R(1.0e-6f);
}

template <class R>
bool intersectAabb(const Ray<R>& ray, const V3<R>& inv, const wpt_bvh2_node& n, R closestT, R& nearT) {  // :182-196
    countNode();
    R tx0 = (R(n.bmin[0]) - ray.o.x) * inv.x;
    R tx1 = (R(n.bmax[0]) - ray.o.x) * inv.x;
    R ty0 = (R(n.bmin[1]) - ray.o.y) * inv.y;
    R ty1 = (R(n.bmax[1]) - ray.o.y) * inv.y;
    R tz0 = (R(n.bmin[2]) - ray.o.z) * inv.z;
    R tz1 = (R(n.bmax[2]) - ray.o.z) * inv.z;
    R tmin = std::fmax(std::fmin(tx0, tx1), std::fmax(std::fmin(ty0, ty1), std::fmin(tz0, tz1)));
    R tmax = std::fmin(std::fmax(tx0, tx1), std::fmin(std::fmax(ty0, ty1), std::fmax(tz0, tz1)));
    nearT = tmin;
    // Culled only beyond closestT * (1 + 2^-20): a box holding a hit at
    // exactly closestT can have a computed entry a few ulps above it, and
    // that hit must still take part in the tie order below.
    return tmax >= std::fmax(tmin, R(1.0e-6f)) && tmin <= closestT * R(1.00000095f);
}

// ---------------------------------------------------------------- closest hit (TraceHitState style)
template <class R>
struct HitState {
    R closestT = R(1e30f);
    int closestMaterial = -1;
    V3<R> geometricNormal, normal;
    V2<R> uv{R(0), R(0)}, uv1{R(0), R(0)};
    V3<R> tangent, bitangent, tangent1, bitangent1;
    uint32_t lightIndex = kInvalid;
    uint32_t sphereIndex = kInvalid;
    // identity of the current winner
    uint32_t kind = WPT_HIT_NONE, instance = kInvalid, prim = kInvalid;
    uint32_t rank = kInvalid;
    R u = R(0), v = R(0);
};

// Order of equally distant hits: by category in the reference's traceRay
// visit order (spheres first: device_traversal.cuh:789-800 lets a sphere
// displace an equally distant triangle; then direct-instance, static,
// TLAS-instance and flat triangles), then instance, then primitive.  Within
// one BVH the reference keeps whichever tie its (unpinned) tree visits
// first; a fixed order makes the result independent of the tree.
enum : uint32_t { kRankSphere = 0, kRankDirectInstance = 1, kRankStatic = 2, kRankTlasInstance = 3, kRankFlat = 4 };

template <class R>
bool nearerHit(R t, uint32_t rank, uint32_t instance, uint32_t prim, const HitState<R>& hs) {
    if (t != hs.closestT) return t < hs.closestT;
    if (rank != hs.rank) return rank < hs.rank;
    if (instance != hs.instance) return instance < hs.instance;
    return prim < hs.prim;
}

template <class R>
V3<R> triNormal(const wpt_triangle& tri, const Ray<R>& ray, R u, R v, bool faceForward) {  // :100-111
    V3<R> n;
    if (tri.flags & WPT_TRIANGLE_HAS_NORMALS)
        n = normalize((R(1) - u - v) * vec<R>(tri.n0) + u * vec<R>(tri.n1) + v * vec<R>(tri.n2));
    else
        n = normalize(cross(vec<R>(tri.v1) - vec<R>(tri.v0), vec<R>(tri.v2) - vec<R>(tri.v0)));
    return faceForward && dot(n, ray.d) > R(0) ? -n : n;
}

template <class R>
V3<R> triGeometricNormal(const wpt_triangle& tri, const Ray<R>& ray, bool faceForward) {  // :113-118
    V3<R> n = normalize(cross(vec<R>(tri.v1) - vec<R>(tri.v0), vec<R>(tri.v2) - vec<R>(tri.v0)));
    return faceForward && dot(n, ray.d) > R(0) ? -n : n;
}

template <class R>
V2<R> triUv(const wpt_triangle& tri, R u, R v) {  // :120-132
    if (!(tri.flags & WPT_TRIANGLE_HAS_TEXCOORDS)) return V2<R>{R(0), R(0)};
    R w = R(1) - u - v;
    return V2<R>{w * R(tri.uv0[0]) + u * R(tri.uv1[0]) + v * R(tri.uv2[0]), w * R(tri.uv0[1]) + u * R(tri.uv1[1]) + v * R(tri.uv2[1])};
}

template <class R>
V2<R> triUv1(const wpt_triangle& tri, R u, R v) {  // :120-132, second set
    if (!(tri.flags & WPT_TRIANGLE_HAS_TEXCOORDS1)) return V2<R>{R(0), R(0)};
    R w = R(1) - u - v;
    return V2<R>{w * R(tri.uv1_0[0]) + u * R(tri.uv1_1[0]) + v * R(tri.uv1_2[0]),
        w * R(tri.uv1_0[1]) + u * R(tri.uv1_1[1]) + v * R(tri.uv1_2[1])};
}

// Alpha mask at a candidate hit (:148-151): masked materials whose opacity
// (x albedo-texture alpha) is below the cutoff are passed through.
template <class R>
bool cutout(const oracle_handle& h, uint32_t m, const wpt_triangle& tri, R u, R v) {
    const wpt_material& x = mat(h, m);
    return x.alpha_masked && matOpacity(h, m, triUv(tri, u, v), triUv1(tri, u, v)) < R(x.alpha_cutoff);
}

// intersectTriangleHit, device_traversal.cuh:137-180.
template <class R>
bool triangleHit(const oracle_handle& h, const Ray<R>& ray, const wpt_triangle& tri, uint32_t triLight, HitState<R>& hs,
    uint32_t materialOverride, uint32_t kind, uint32_t rank, uint32_t instance, uint32_t prim) {
    R t, u, v;
    if (!intersectTriangle(ray, tri, t, u, v) || !nearerHit(t, rank, instance, prim, hs)) return false;
    V2<R> uv = triUv(tri, u, v);
    uint32_t m = materialOverride == kInvalid ? tri.material : materialOverride;
    if (cutout(h, m, tri, u, v)) return false;
    bool ff = mat(h, m).type != WPT_MATERIAL_DIELECTRIC;
    hs.closestT = t;
    hs.closestMaterial = static_cast<int>(m);
    hs.geometricNormal = triGeometricNormal(tri, ray, ff);
    hs.normal = triNormal(tri, ray, u, v, ff);
    hs.uv = uv;
    hs.uv1 = triUv1(tri, u, v);
    hs.tangent = vec<R>(tri.tangent);
    hs.bitangent = vec<R>(tri.bitangent);
    if (tri.flags & WPT_TRIANGLE_HAS_VERTEX_TANGENTS) {  // :159-173
        R w = R(1) - u - v;
        V3<R> tg = w * vec<R>(tri.vertex_tangent0) + u * vec<R>(tri.vertex_tangent1) + v * vec<R>(tri.vertex_tangent2);
        tg = tg - hs.normal * dot(hs.normal, tg);
        if (lengthSquared(tg) > R(0.00000001f)) {
            hs.tangent = normalize(tg);
            R sign = w * R(tri.vertex_tangent_sign[0]) + u * R(tri.vertex_tangent_sign[1]) +
                v * R(tri.vertex_tangent_sign[2]);
            hs.bitangent = (sign < R(0) ? R(-1) : R(1)) * normalize(cross(hs.normal, hs.tangent));
        }
    }
    hs.tangent1 = vec<R>(tri.tangent1);
    hs.bitangent1 = vec<R>(tri.bitangent1);
    hs.lightIndex = materialOverride == kInvalid ? triLight : kInvalid;
    hs.kind = kind;
    hs.rank = rank;
    hs.instance = instance;
    hs.prim = prim;
    hs.u = u;
    hs.v = v;
    return true;
}

template <class R>
void sphereHit(const oracle_handle& h, const Ray<R>& ray, uint32_t index, HitState<R>& hs) {  // :789-800
    R t;
    if (!intersectSphere(ray, h.d.spheres[index], t) || !nearerHit(t, kRankSphere, kInvalid, index, hs)) return;
    hs.closestT = t;
    hs.closestMaterial = static_cast<int>(h.d.spheres[index].material);
    hs.sphereIndex = index;
    hs.kind = WPT_HIT_SPHERE;
    hs.rank = kRankSphere;
    hs.instance = kInvalid;
    hs.prim = index;
    hs.u = hs.v = R(0);
}

// intersectNexusBvh2 over triangles, device_traversal.cuh:239-313.
template <class R>
void bvhTriangles(const oracle_handle& h, const Ray<R>& ray, const BvhView& bvh, const wpt_triangle* tris,
    const uint32_t* triLights, HitState<R>& hs, uint32_t materialOverride, uint32_t kind, uint32_t rank,
    uint32_t instance) {
    V3<R> inv(R(1) / ray.d.x, R(1) / ray.d.y, R(1) / ray.d.z);
    uint32_t stack[64];
    uint32_t size = 1;
    stack[0] = bvh.count - 1;
    while (size > 0) {
        const wpt_bvh2_node& node = bvh.nodes[stack[--size]];
        R nearT;
        if (!intersectAabb(ray, inv, node, hs.closestT, nearT)) continue;
        if (node.left == kInvalid) {
            uint32_t p = node.right;
            triangleHit(h, ray, tris[p], triLights ? triLights[p] : kInvalid, hs, materialOverride, kind, rank, instance,
                p);
        } else {
            R ln, rn;
            bool hl = intersectAabb(ray, inv, bvh.nodes[node.left], hs.closestT, ln);
            bool hr = intersectAabb(ray, inv, bvh.nodes[node.right], hs.closestT, rn);
            if (hl && hr) {
                uint32_t nearChild = ln < rn ? node.left : node.right;
                uint32_t farChild = ln < rn ? node.right : node.left;
                stack[size++] = farChild;
                stack[size++] = nearChild;
            } else if (hl) {
                stack[size++] = node.left;
            } else if (hr) {
                stack[size++] = node.right;
            }
        }
    }
}

template <class R>
V3<R> inverseXform(const wpt_transform& t, const V3<R>& v) {  // scene.cuh:43-46
    return vec<R>(t.world_to_local[0]) * v.x + vec<R>(t.world_to_local[1]) * v.y + vec<R>(t.world_to_local[2]) * v.z;
}

template <class R>
V3<R> xformVector(const wpt_transform& t, const V3<R>& v) {  // scene.cuh:38-41
    return vec<R>(t.local_to_world[0]) * v.x + vec<R>(t.local_to_world[1]) * v.y + vec<R>(t.local_to_world[2]) * v.z;
}

template <class R>
V3<R> xformNormal(const wpt_transform& t, const V3<R>& n) {  // scene.cuh:48-55
    return normalize(V3<R>(dot(vec<R>(t.world_to_local[0]), n), dot(vec<R>(t.world_to_local[1]), n),
        dot(vec<R>(t.world_to_local[2]), n)));
}

template <class R>
Ray<R> localRay(const wpt_transform& t, const Ray<R>& ray) {
    Ray<R> l;
    l.o = inverseXform(t, ray.o - vec<R>(t.translation));
    l.d = inverseXform(t, ray.d);
    return l;
}

// intersectMeshInstance, device_traversal.cuh:802-834.
template <class R>
void meshInstance(const oracle_handle& h, const Ray<R>& ray, uint32_t index, HitState<R>& hs, bool useBvh) {
    countInstance();
    const wpt_instance& inst = h.d.instances[index];
    const wpt_mesh& mesh = h.d.meshes[inst.mesh];
    const wpt_triangle* tris = h.d.mesh_triangles + mesh.first_triangle;
    const uint32_t* lights = h.meshLight.data() + mesh.first_triangle;
    Ray<R> local = localRay(inst.transform, ray);
    const uint32_t previousRank = hs.rank, previousInstance = hs.instance, previousPrim = hs.prim;
    const R previous = hs.closestT;
    const BvhView& bvh = h.blas[inst.mesh];
    const uint32_t rank = useBvh ? kRankTlasInstance : kRankDirectInstance;
    if (useBvh && bvh.count) {
        bvhTriangles(h, local, bvh, tris, lights, hs, inst.material_override, WPT_HIT_INSTANCE_TRIANGLE, rank, index);
    } else {
        for (uint32_t i = 0; i < mesh.triangle_count; ++i)
            triangleHit(h, local, tris[i], lights[i], hs, inst.material_override, WPT_HIT_INSTANCE_TRIANGLE, rank, index,
                i);
    }
    // Unchanged winner: nothing of this instance to transform.
    if (hs.closestT == previous && hs.rank == previousRank && hs.instance == previousInstance && hs.prim == previousPrim)
        return;
    hs.geometricNormal = xformNormal(inst.transform, hs.geometricNormal);
    hs.normal = xformNormal(inst.transform, hs.normal);
    hs.tangent = normalize(xformVector(inst.transform, hs.tangent));
    hs.bitangent = normalize(xformVector(inst.transform, hs.bitangent));
    hs.tangent1 = normalize(xformVector(inst.transform, hs.tangent1));
    hs.bitangent1 = normalize(xformVector(inst.transform, hs.bitangent1));
    if (hs.lightIndex != kInvalid) hs.lightIndex += h.instanceLightOffset[index];
    hs.sphereIndex = kInvalid;
}

template <class R>
void sceneTlas(const oracle_handle& h, const Ray<R>& ray, HitState<R>& hs) {  // :836-878
    if (!h.tlas.count) return;
    V3<R> inv(R(1) / ray.d.x, R(1) / ray.d.y, R(1) / ray.d.z);
    uint32_t stack[64];
    uint32_t size = 1;
    stack[0] = h.tlas.count - 1;
    while (size > 0) {
        const wpt_bvh2_node& node = h.tlas.nodes[stack[--size]];
        R nearT;
        if (!intersectAabb(ray, inv, node, hs.closestT, nearT)) continue;
        if (node.left == kInvalid) {
            uint32_t leaf = h.leaves[node.right];
            if (leaf & 0x80000000u) sphereHit(h, ray, leaf & 0x7fffffffu, hs);
            else meshInstance(h, ray, leaf, hs, true);
        } else {
            R ln, rn;
            bool hl = intersectAabb(ray, inv, h.tlas.nodes[node.left], hs.closestT, ln);
            bool hr = intersectAabb(ray, inv, h.tlas.nodes[node.right], hs.closestT, rn);
            if (hl && hr) {
                uint32_t nearChild = ln < rn ? node.left : node.right;
                uint32_t farChild = ln < rn ? node.right : node.left;
                stack[size++] = farChild;
                stack[size++] = nearChild;
            } else if (hl) {
                stack[size++] = node.left;
            } else if (hr) {
                stack[size++] = node.right;
            }
        }
    }
}

template <class R>
void directTriangles(const oracle_handle& h, const Ray<R>& ray, HitState<R>& hs) {  // :880-1033
    if (h.flat.count) {
        bvhTriangles(h, ray, h.flat, h.d.flat_triangles, nullptr, hs, kInvalid, WPT_HIT_FLAT_TRIANGLE, kRankFlat,
            kInvalid);
    } else {
        for (uint32_t i = 0; i < h.d.flat_triangle_count; ++i)
            triangleHit(h, ray, h.d.flat_triangles[i], kInvalid, hs, kInvalid, WPT_HIT_FLAT_TRIANGLE, kRankFlat, kInvalid,
                i);
    }
}

template <class R>
struct Hit {
    R t;
    V3<R> position, geometricNormal, normal;
    V2<R> uv, uv1;
    V3<R> tangent, bitangent, tangent1, bitangent1;
    uint32_t material, lightIndex;
};

// traceRay, device_traversal.cuh:1115-1184.
template <class R>
bool traceRay(const oracle_handle& h, const Ray<R>& ray, Hit<R>& hit, HitState<R>& hs) {
    for (uint32_t i = 0; i < h.directSpheres.size(); ++i) sphereHit(h, ray, h.directSpheres[i], hs);
    for (uint32_t i = 0; i < h.directInstances.size(); ++i) meshInstance(h, ray, h.directInstances[i], hs, false);
    for (uint32_t i = 0; i < h.d.static_triangle_count; ++i)
        if (triangleHit(h, ray, h.d.static_triangles[i], h.staticLight[i], hs, kInvalid, WPT_HIT_STATIC_TRIANGLE,
                kRankStatic, kInvalid, i))
            hs.sphereIndex = kInvalid;
    sceneTlas(h, ray, hs);
    if (h.d.flat_triangle_count > 0) {
        directTriangles(h, ray, hs);
        if (hs.kind == WPT_HIT_FLAT_TRIANGLE) hs.sphereIndex = kInvalid;
    }
    if (hs.sphereIndex != kInvalid) {
        const wpt_sphere& s = h.d.spheres[hs.sphereIndex];
        hs.normal = normalize(ray.at(hs.closestT) - vec<R>(s.center));
        hs.geometricNormal = hs.normal;
        hs.uv = hs.uv1 = V2<R>{R(0), R(0)};
        hs.tangent = hs.bitangent = hs.tangent1 = hs.bitangent1 = V3<R>();
        hs.lightIndex = kInvalid;
    }
    if (hs.closestMaterial < 0) return false;
    hit.t = hs.closestT;
    hit.position = ray.at(hs.closestT);
    hit.geometricNormal = hs.geometricNormal;
    hit.normal = hs.normal;
    hit.uv = hs.uv;
    hit.uv1 = hs.uv1;
    hit.tangent = hs.tangent;
    hit.bitangent = hs.bitangent;
    hit.tangent1 = hs.tangent1;
    hit.bitangent1 = hs.bitangent1;
    hit.material = static_cast<uint32_t>(hs.closestMaterial);
    hit.lightIndex = hs.lightIndex;
    return true;
}

// isOccluded, device_traversal.cuh:1035-1113.
template <class R>
bool isOccluded(const oracle_handle& h, const Ray<R>& ray, R maxDistance) {
    for (uint32_t i : h.directSpheres) {
        R t;
        if (intersectSphere(ray, h.d.spheres[i], t) && t < maxDistance) return true;
    }
    for (uint32_t i : h.directInstances) {
        HitState<R> hs;
        hs.closestT = maxDistance;
        meshInstance(h, ray, i, hs, false);
        if (hs.closestMaterial >= 0) return true;
    }
    for (uint32_t i = 0; i < h.d.static_triangle_count; ++i) {
        R t, u, v;
        const wpt_triangle& tri = h.d.static_triangles[i];
        if (intersectTriangle(ray, tri, t, u, v) && t < maxDistance && !cutout(h, tri.material, tri, u, v)) return true;
    }
    if (h.tlas.count) {
        V3<R> inv(R(1) / ray.d.x, R(1) / ray.d.y, R(1) / ray.d.z);
        uint32_t stack[64];
        uint32_t size = 1;
        stack[0] = h.tlas.count - 1;
        while (size > 0) {
            const wpt_bvh2_node& node = h.tlas.nodes[stack[--size]];
            R nearT;
            if (!intersectAabb(ray, inv, node, maxDistance, nearT)) continue;
            if (node.left == kInvalid) {
                uint32_t leaf = h.leaves[node.right];
                if (leaf & 0x80000000u) {
                    R t;
                    if (intersectSphere(ray, h.d.spheres[leaf & 0x7fffffffu], t) && t < maxDistance) return true;
                } else {
                    HitState<R> hs;
                    hs.closestT = maxDistance;
                    meshInstance(h, ray, leaf, hs, true);
                    if (hs.closestMaterial >= 0) return true;
                }
            } else {
                R ln, rn;
                bool hl = intersectAabb(ray, inv, h.tlas.nodes[node.left], maxDistance, ln);
                bool hr = intersectAabb(ray, inv, h.tlas.nodes[node.right], maxDistance, rn);
                if (hl && hr) {
                    uint32_t nearChild = ln < rn ? node.left : node.right;
                    uint32_t farChild = ln < rn ? node.right : node.left;
                    stack[size++] = farChild;
                    stack[size++] = nearChild;
                } else if (hl) {
                    stack[size++] = node.left;
                } else if (hr) {
                    stack[size++] = node.right;
                }
            }
        }
    }
    if (h.d.flat_triangle_count > 0) {
        HitState<R> hs;
        hs.closestT = maxDistance;
        directTriangles(h, ray, hs);
        if (hs.closestMaterial >= 0) return true;
    }
    return false;
}

// ---------------------------------------------------------------- shading
template <class R>
struct PathState {  // path_state.cuh:6-24
    Ray<R> ray;
    V3<R> throughput, radiance;
    uint32_t pixelIndex, depth;
    RngState rng;
    uint32_t mediumMaterial;
    R mediumDensity;
    R previousBsdfPdf;
    bool specularBounce, active;
    R etaScale;
    R clamp;
};

template <class R>
void addPathRadiance(PathState<R>& p, V3<R> c, bool emissionOrBackground = false) {  // path_state.cuh:33-43
    bool indirect = p.depth > (emissionOrBackground ? 1u : 0u);
    if (indirect && p.clamp > R(0)) {
        R magnitude = std::fabs(c.x) / R(3) + std::fabs(c.y) / R(3) + std::fabs(c.z) / R(3);
        if (magnitude > p.clamp) c *= p.clamp / magnitude;
    }
    p.radiance += c;
}

// ---------------------------------------------------------------- environment lookups
// include/scene/environment.cuh:17-70 and device_traversal.cuh:571-575.
inline float fatan2(float y, float x) { return wpt::atan2F(y, x); }
inline double fatan2(double y, double x) { return wpt::atan2D(y, x); }

template <class R>
V3<R> envRotate(const V3<R>& d, R angle) {  // right-handed turn about +Y
    R c = fcos(angle), s = fsin(angle);
    return V3<R>(c * d.x + s * d.z, d.y, -s * d.x + c * d.z);
}

// Panorama coordinates: +Y at v = 0, +X at u = 1/2, +Z at u = 3/4.
template <class R>
void envUv(const oracle_handle& h, const V3<R>& world, R& u, R& v) {
    const R pi = R(3.14159265359f);
    V3<R> d = envRotate(world, R(-h.envRotation));
    u = fatan2(d.z, d.x) / (R(2) * pi) + R(0.5f);
    u -= std::floor(u);
    v = std::fmin(R(1), std::fmax(R(0), fatan2(std::sqrt(d.x * d.x + d.z * d.z), d.y) / pi));
}

template <class R>
uint32_t envTexel(const oracle_handle& h, const V3<R>& dir) {
    R u, v;
    envUv(h, dir, u, v);
    uint32_t x = static_cast<uint32_t>(u * R(h.envW)), y = static_cast<uint32_t>(v * R(h.envH));
    return std::min(y, h.envH - 1) * h.envW + std::min(x, h.envW - 1);
}

template <class R>
R envPdf(const oracle_handle& h, const V3<R>& dir) {
    if (!h.env) return R(0);
    uint32_t i = envTexel(h, dir);
    return R(h.envProb[i]) / R(envSolidAngle(i / h.envW, h.envW, h.envH));
}

// Alias pick from a FP64 candidate and a coin, then a uniform longitude and
// a uniform cos(latitude) inside the texel (all in FP64; the longitude
// offset in FP32 as the reference writes it).
template <class R>
V3<R> envSample(const oracle_handle& h, double candidate, float coin, float lon, float lat, R& pdf) {
    const double pi = 3.14159265358979323846;
    uint32_t n = h.envW * h.envH;
    uint32_t i = std::min(static_cast<uint32_t>(candidate * n), n - 1);
    if (coin >= h.envAlias[i].probability) i = h.envAlias[i].alias;
    uint32_t x = i % h.envW, y = i / h.envW;
    float lonClamped = std::fmin(1.0f - 1.0e-7f, std::fmax(1.0e-7f, lon));
    double u = (static_cast<float>(x) + lonClamped) / static_cast<float>(h.envW);
    double t = std::fmin(1.0f - 1.0e-7f, std::fmax(1.0e-7f, lat));
    double c = (1.0 - t) * fcos(pi * y / h.envH) + t * fcos(pi * (y + 1.0) / h.envH);
    double sn = std::sqrt(std::fmax(0.0, 1.0 - c * c));
    double phi = (u - 0.5) * 2 * pi;
    V3<R> local(R(static_cast<float>(sn * fcos(phi))), R(static_cast<float>(c)), R(static_cast<float>(sn * fsin(phi))));
    pdf = R(h.envProb[i]) / R(envSolidAngle(y, h.envW, h.envH));
    return envRotate(local, R(h.envRotation));
}

// Bilinear reconstruction with texel centers at half-integers, u wrapping
// and v clamped, times the intensity.  Exact weights (the reference's
// texture unit quantizes them to 1/256).
template <class R>
V3<R> envRadiance(const oracle_handle& h, const V3<R>& dir) {
    R u, v;
    envUv(h, dir, u, v);
    R fx = u * R(h.envW) - R(0.5f), fy = v * R(h.envH) - R(0.5f);
    R x0 = std::floor(fx), y0 = std::floor(fy);
    R ax = fx - x0, ay = fy - y0;
    int w = static_cast<int>(h.envW), ht = static_cast<int>(h.envH);
    int ix = static_cast<int>(x0), iy = static_cast<int>(y0);
    auto texel = [&](int xx, int yy) {
        xx = ((xx % w) + w) % w;
        yy = std::min(std::max(yy, 0), ht - 1);
        const float* p = &h.envRgb[3ull * (static_cast<uint64_t>(yy) * w + xx)];
        return V3<R>(R(p[0]), R(p[1]), R(p[2]));
    };
    V3<R> a = texel(ix, iy), b = texel(ix + 1, iy), c = texel(ix, iy + 1), e = texel(ix + 1, iy + 1);
    R w00 = (R(1) - ax) * (R(1) - ay), w10 = ax * (R(1) - ay), w01 = (R(1) - ax) * ay, w11 = ax * ay;
    V3<R> r(a.x * w00 + b.x * w10 + c.x * w01 + e.x * w11, a.y * w00 + b.y * w10 + c.y * w01 + e.y * w11,
        a.z * w00 + b.z * w10 + c.z * w01 + e.z * w11);
    return r * R(h.envIntensity);
}

template <class R> R powerHeuristic(R a, R b) {
    R a2 = a * a, b2 = b * b;
    return a2 / (a2 + b2);
}

template <class R> V3<R> reflect(const V3<R>& i, const V3<R>& n) { return i - R(2) * dot(i, n) * n; }  // device_sampling.cuh:37-40

template <class R> V3<R> refract(const V3<R>& i, const V3<R>& n, R eta) {  // :42-48
    R c = std::fmin(dot(-i, n), R(1));
    V3<R> perp = eta * (i + c * n);
    V3<R> par = -std::sqrt(std::fmax(R(0), R(1) - lengthSquared(perp))) * n;
    return perp + par;
}

template <class R> R dielectricReflectance(R cosine, R eta) {  // :59-68
    if (eta == R(1)) return R(0);
    cosine = std::fmin(R(1), std::fmax(R(0), cosine));
    R s2 = eta * eta * std::fmax(R(0), R(1) - cosine * cosine);
    if (s2 >= R(1)) return R(1);
    R ct = std::sqrt(R(1) - s2);
    R perp = (eta * cosine - ct) / (eta * cosine + ct);
    R par = (cosine - eta * ct) / (cosine + eta * ct);
    return R(0.5f) * (perp * perp + par * par);
}

template <class R>
struct Interface {
    V3<R> normal, reflection, transmission;
    R eta, reflectance;
    bool frontFace;
};

template <class R>
Interface<R> dielectricInterface(const V3<R>& incident, const V3<R>& gn, V3<R> sn, R ior) {  // :71-93
    bool front = dot(incident, gn) < R(0);
    V3<R> facing = front ? gn : -gn;
    if (dot(sn, gn) < R(0)) sn = -sn;
    V3<R> n = front ? sn : -sn;
    R eta = front ? R(1) / ior : ior;
    if (dot(incident, n) >= R(0)) n = facing;
    R refl = dielectricReflectance(dot(-incident, n), eta);
    V3<R> r = reflect(incident, n);
    V3<R> t = refract(incident, n, eta);
    if (dot(r, facing) <= R(0) || (refl < R(1) && dot(t, facing) >= R(0))) {
        n = facing;
        refl = dielectricReflectance(dot(-incident, n), eta);
        r = reflect(incident, n);
        t = refract(incident, n, eta);
    }
    return Interface<R>{n, r, t, eta, refl, front};
}

template <class R>
V3<R> sampleCosineHemisphere(const V3<R>& n, RngState& rng) {  // :6-25
    const R twoPi = R(6.28318530718f);
    R u1 = rnd<R>(rng);
    R u2 = rnd<R>(rng);
    R radius = std::sqrt(u1);
    R phi = twoPi * u2;
    R x = radius * fcos(phi);
    R y = radius * fsin(phi);
    R z = std::sqrt(R(1) - u1);
    V3<R> helper = std::fabs(n.x) > R(0.1f) ? V3<R>(R(0), R(1), R(0)) : V3<R>(R(1), R(0), R(0));
    V3<R> tangent = normalize(cross(helper, n));
    V3<R> bitangent = cross(n, tangent);
    return normalize(tangent * x + bitangent * y + n * z);
}

template <class R>
V3<R> randomInUnitSphere(RngState& rng) {  // :27-35, draws taken x, y, z
    while (true) {
        R x = R(2) * rnd<R>(rng) - R(1);
        R y = R(2) * rnd<R>(rng) - R(1);
        R z = R(2) * rnd<R>(rng) - R(1);
        V3<R> p(x, y, z);
        if (lengthSquared(p) < R(1)) return p;
    }
}

// specular_glossiness.cuh:7-80
template <class R>
struct SG {
    V3<R> diffuse, f0;
    R roughness;
};
template <class R> R sgMax(const V3<R>& v) { return std::fmax(v.x, std::fmax(v.y, v.z)); }
template <class R> R sgProbability(const SG<R>& s) {
    R spec = sgMax(s.f0);
    R diff = sgMax(s.diffuse) * (R(1) - spec);
    return std::fmin(R(0.95f), std::fmax(R(0.05f), spec / std::fmax(R(1.0e-6f), spec + diff)));
}
template <class R> R sgAlpha(const SG<R>& s) {
    R r = std::fmax(R(0.02f), s.roughness);
    return r * r;
}
template <class R> R sgD(R nh, R alpha) {
    const R pi = R(3.14159265359f);
    R a2 = alpha * alpha;
    R den = (R(1) - nh * nh) + a2 * nh * nh;
    return a2 / (pi * den * den);
}
template <class R> R sgG1(R c, R alpha) {
    return R(2) * c / (c + std::sqrt(alpha * alpha + (R(1) - alpha * alpha) * c * c));
}
template <class R>
V3<R> evalSG(const SG<R>& s, const V3<R>& n, const V3<R>& wo, const V3<R>& wi, R& pdf) {
    const R invPi = R(0.31830988618f);
    pdf = R(0);
    R nv = dot(n, wo), nl = dot(n, wi);
    V3<R> sum = wo + wi;
    if (nv <= R(0) || nl <= R(0) || lengthSquared(sum) < R(1.0e-12f)) return V3<R>();
    V3<R> half = normalize(sum);
    R nh = std::fmin(R(1), std::fmax(R(0), dot(n, half)));
    R vh = std::fmin(R(1), std::fmax(R(1.0e-8f), dot(wo, half)));
    R alpha = sgAlpha(s);
    R d = sgD(nh, alpha);
    R prob = sgProbability(s);
    pdf = (R(1) - prob) * nl * invPi + prob * d * nh / (R(4) * vh);
    R schlick = fpow(R(1) - vh, R(5));
    V3<R> fresnel = s.f0 + (V3<R>(R(1), R(1), R(1)) - s.f0) * schlick;
    V3<R> diffuse = s.diffuse * ((R(1) - sgMax(s.f0)) * invPi);
    R g = sgG1(nv, alpha) * sgG1(nl, alpha);
    return diffuse + fresnel * (d * g / (R(4) * nv * nl));
}
template <class R>
V3<R> sgBasis(const V3<R>& n, R x, R y, R z) {
    V3<R> helper = std::fabs(n.x) > R(0.1f) ? V3<R>(R(0), R(1), R(0)) : V3<R>(R(1), R(0), R(0));
    V3<R> tangent = normalize(cross(helper, n));
    return normalize(x * tangent + y * cross(n, tangent) + z * n);
}
template <class R>
V3<R> sampleSG(const SG<R>& s, const V3<R>& n, const V3<R>& wo, R choice, R u, R v) {
    R phi = R(6.28318530718f) * v;
    if (choice < sgProbability(s)) {
        R alpha = sgAlpha(s);
        R c = std::sqrt((R(1) - u) / (R(1) + (alpha * alpha - R(1)) * u));
        R sn = std::sqrt(std::fmax(R(0), R(1) - c * c));
        V3<R> half = sgBasis(n, sn * fcos(phi), sn * fsin(phi), c);
        return R(2) * dot(wo, half) * half - wo;
    }
    R radius = std::sqrt(u);
    return sgBasis(n, radius * fcos(phi), radius * fsin(phi), std::sqrt(R(1) - u));
}

// ---------------------------------------------------------------- textures and materials
// device_traversal.cuh:325-525 (texture lookups, material terms, normal
// maps, glTF surface) and gltf_pbr.cuh.  Textures read normalized texels
// with bilinear weights and wrapping; the weights are exact (the
// reference's texture unit rounds them to 1/256).
template <class R> struct T4 { R x, y, z, w; };

template <class R>
T4<R> texel(const oracle_handle& h, uint32_t t, uint32_t x, uint32_t y) {
    const uint8_t* p = &h.tex[t][4ull * (static_cast<uint64_t>(y) * h.texW[t] + x)];
    return T4<R>{R(p[0]) / R(255.0f), R(p[1]) / R(255.0f), R(p[2]) / R(255.0f), R(p[3]) / R(255.0f)};
}

template <class R>
R unitWrap(R a) {
    R w = a - std::floor(a);
    return w >= R(0) && w <= R(1) ? w : R(0);  // non-finite coordinates read 0
}

template <class R>
T4<R> texBilinear(const oracle_handle& h, uint32_t t, R u, R v) {
    const int64_t W = h.texW[t], H = h.texH[t];
    R x = unitWrap(u) * R(static_cast<float>(W)) - R(0.5f);
    R y = unitWrap(v) * R(static_cast<float>(H)) - R(0.5f);
    R fx = std::floor(x), fy = std::floor(y);
    R ax = x - fx, ay = y - fy;
    int64_t ix = static_cast<int64_t>(fx), iy = static_cast<int64_t>(fy);
    auto wrap = [](int64_t i, int64_t n) { return static_cast<uint32_t>(((i % n) + n) % n); };
    T4<R> a = texel<R>(h, t, wrap(ix, W), wrap(iy, H)), b = texel<R>(h, t, wrap(ix + 1, W), wrap(iy, H));
    T4<R> c = texel<R>(h, t, wrap(ix, W), wrap(iy + 1, H)), e = texel<R>(h, t, wrap(ix + 1, W), wrap(iy + 1, H));
    R w00 = (R(1) - ax) * (R(1) - ay), w10 = ax * (R(1) - ay), w01 = (R(1) - ax) * ay, w11 = ax * ay;
    return T4<R>{a.x * w00 + b.x * w10 + c.x * w01 + e.x * w11, a.y * w00 + b.y * w10 + c.y * w01 + e.y * w11,
        a.z * w00 + b.z * w10 + c.z * w01 + e.z * w11, a.w * w00 + b.w * w10 + c.w * w01 + e.w * w11};
}

// getTextureUv + sampleTexture (:325-347): the image is read at 1 - v.
template <class R>
T4<R> sampleTex(const oracle_handle& h, const wpt_texture_binding& b, const V2<R>& uv0, const V2<R>& uv1,
    T4<R> fallback) {
    if (b.texture == kInvalid || b.texture >= h.tex.size()) return fallback;
    V2<R> uv = b.texcoord == 1 ? uv1 : uv0;
    if (b.has_transform) {
        uv.x *= R(b.scale[0]);
        uv.y *= R(b.scale[1]);
        R sn = fsin(R(b.rotation)), cs = fcos(R(b.rotation));
        uv = V2<R>{cs * uv.x - sn * uv.y + R(b.offset[0]), sn * uv.x + cs * uv.y + R(b.offset[1])};
    }
    return texBilinear(h, b.texture, uv.x, R(1) - uv.y);
}

template <class R> T4<R> white() { return T4<R>{R(1), R(1), R(1), R(1)}; }
template <class R> V3<R> srgb22(const T4<R>& t) {
    return V3<R>(fpow(t.x, R(2.2f)), fpow(t.y, R(2.2f)), fpow(t.z, R(2.2f)));
}
template <class R> R gltfSrgb1(R c) {
    return c <= R(0.04045f) ? c / R(12.92f) : fpow((c + R(0.055f)) / R(1.055f), R(2.4f));
}
template <class R> V3<R> gltfSrgb(const T4<R>& t) { return V3<R>(gltfSrgb1(t.x), gltfSrgb1(t.y), gltfSrgb1(t.z)); }

// The texture-carrying fields of reference Hit (geometry.cuh:68-82).
template <class R>
struct TexHit {
    V2<R> uv, uv1;
    V3<R> normal, geometricNormal, tangent, bitangent, tangent1, bitangent1;
};

template <class R>
V3<R> matAlbedo(const oracle_handle& h, uint32_t m, const TexHit<R>& th) {  // :359-363
    const wpt_material_ext& e = h.ext[m];
    T4<R> t = sampleTex(h, e.albedo_texture, th.uv, th.uv1, white<R>());
    return hadamard(vec<R>(mat(h, m).albedo), e.specular_glossiness_workflow ? gltfSrgb(t) : srgb22(t));
}

template <class R>
R matOpacity(const oracle_handle& h, uint32_t m, const V2<R>& uv, const V2<R>& uv1) {  // :365-375
    const wpt_material_ext& e = h.ext[m];
    if (e.albedo_texture.texture == kInvalid || e.albedo_texture.texture >= h.tex.size()) return R(mat(h, m).opacity);
    return R(mat(h, m).opacity) * sampleTex(h, e.albedo_texture, uv, uv1, white<R>()).w;
}

template <class R>
V3<R> matEmission(const oracle_handle& h, uint32_t m, const TexHit<R>& th) {  // :377-380
    return hadamard(vec<R>(mat(h, m).emission), srgb22(sampleTex(h, h.ext[m].emissive_texture, th.uv, th.uv1,
        white<R>())));
}

template <class R>
R matMetallic(const oracle_handle& h, uint32_t m, const TexHit<R>& th) {  // :382-385
    return R(mat(h, m).metallic) * sampleTex(h, h.ext[m].metallic_roughness_texture, th.uv, th.uv1, white<R>()).z;
}

template <class R>
R matRoughness(const oracle_handle& h, uint32_t m, const TexHit<R>& th) {  // :387-393
    const wpt_material_ext& e = h.ext[m];
    if (e.specular_glossiness_workflow)
        return R(1) - R(mat(h, m).glossiness) * sampleTex(h, e.specular_glossiness_texture, th.uv, th.uv1, white<R>()).w;
    return R(mat(h, m).roughness) * sampleTex(h, e.metallic_roughness_texture, th.uv, th.uv1, white<R>()).y;
}

template <class R>
R matThickness(const oracle_handle& h, uint32_t m, const TexHit<R>& th) {  // :443-449
    const wpt_texture_binding& b = h.ext[m].thickness_texture;
    if (b.texture == kInvalid || b.texture >= h.tex.size()) return R(1);
    return sampleTex(h, b, th.uv, th.uv1, white<R>()).x;
}

template <class R> bool finiteV(const V3<R>& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

// getMaterialNormal (:451-500) for a binding and its scale.
template <class R>
V3<R> mapNormal(const oracle_handle& h, const wpt_texture_binding& b, float scale, const TexHit<R>& th) {
    if (!h.d.use_normal_maps || b.texture == kInvalid || b.texture >= h.tex.size()) return th.normal;
    V3<R> tg = b.texcoord == 1 ? th.tangent1 : th.tangent;
    V3<R> bt = b.texcoord == 1 ? th.bitangent1 : th.bitangent;
    if (!finiteV(tg) || !finiteV(bt) || lengthSquared(tg) < R(0.00000001f) || lengthSquared(bt) < R(0.00000001f))
        return th.normal;
    V3<R> t = normalize(tg - th.normal * dot(th.normal, tg));
    V3<R> bb = normalize(cross(th.normal, t));
    if (!finiteV(t) || !finiteV(bb)) return th.normal;
    if (dot(bb, bt) < R(0)) bb = -bb;
    T4<R> x = sampleTex(h, b, th.uv, th.uv1, T4<R>{R(0.5f), R(0.5f), R(1), R(1)});
    V3<R> local((x.x * R(2) - R(1)) * R(scale), (x.y * R(2) - R(1)) * R(scale), x.z * R(2) - R(1));
    local = normalize(local);
    R minCos = R(h.d.normal_map_minimum_cosine);
    if (minCos > R(0) && local.z < minCos) {
        R mc = std::fmin(minCos, R(1));
        R len = std::sqrt(std::fmax(R(0), local.x * local.x + local.y * local.y));
        if (len > R(0)) {
            R k = std::sqrt(std::fmax(R(0), R(1) - mc * mc)) / len;
            local.x *= k;
            local.y *= k;
        } else {
            local.x = R(0);
            local.y = R(0);
        }
        local.z = mc;
    }
    V3<R> n = t * local.x + bb * local.y + th.normal * local.z;
    if (!finiteV(n) || lengthSquared(n) < R(0.00000001f)) return th.normal;
    return normalize(n);
}

// glTF metallic-roughness surface with specular and clearcoat (gltf_pbr.cuh).
template <class R>
struct GSurf {
    V3<R> color, normal, coatNormal;
    R metallic, roughness;
    V3<R> f0;
    R f90, coat, coatRoughness;
    V3<R> reflectionNormal;
};

template <class R>
V3<R> reflectionFrame(V3<R> n, V3<R> g, V3<R> view) {  // gltfReflectionNormal: 24 bisection steps
    R gv = dot(g, view);
    if (gv <= R(0)) return g;
    R target = std::fmin(R(0.01f), R(0.9f) * gv);
    V3<R> r = R(2) * dot(n, view) * n - view;
    if (dot(g, r) >= target) return n;
    R lo = R(0), hi = R(1);
    for (int i = 0; i < 24; ++i) {
        R t = (lo + hi) * R(0.5f);
        V3<R> c = normalize(n * (R(1) - t) + g * t);
        r = R(2) * dot(c, view) * c - view;
        if (dot(g, r) >= target) hi = t;
        else lo = t;
    }
    return normalize(n * (R(1) - hi) + g * hi);
}
template <class R> V3<R> specN(const GSurf<R>& s) { return lengthSquared(s.reflectionNormal) > R(0) ? s.reflectionNormal : s.normal; }
template <class R> R sat(R x) { return std::fmin(R(1), std::fmax(R(0), x)); }
template <class R> R schlick5(R c) {
    R x = R(1) - sat(c);
    return x * x * x * x * x;
}
template <class R> V3<R> dielF(const GSurf<R>& s, R c) {
    return s.f0 + (V3<R>(s.f90, s.f90, s.f90) - s.f0) * schlick5(c);
}
template <class R> V3<R> baseF(const GSurf<R>& s, R c) {
    V3<R> metal = s.color + (V3<R>(R(1), R(1), R(1)) - s.color) * schlick5(c);
    return dielF(s, c) * (R(1) - s.metallic) + metal * s.metallic;
}
template <class R> void lobeWeights(const GSurf<R>& s, const V3<R>& view, R& wd, R& ws, R& wc) {
    wc = s.coat * (R(0.04f) + R(0.96f) * schlick5(dot(s.coatNormal, view)));
    R spec = s.metallic + (R(1) - s.metallic) * sgMax(dielF(s, dot(specN(s), view)));
    wd = (R(1) - wc) * (R(1) - spec);
    ws = (R(1) - wc) * spec;
}
template <class R> R gAlpha(R r) { return std::fmax(R(1.e-4f), r * r); }
template <class R> R reflPdf(const V3<R>& n, const V3<R>& view, const V3<R>& light, R rough) {
    if (rough <= R(0) || dot(n, view) <= R(0) || dot(n, light) <= R(0) || lengthSquared(view + light) < R(1.e-12f))
        return R(0);
    V3<R> half = normalize(view + light);
    R vh = dot(view, half), nh = dot(n, half);
    return vh > R(0) && nh > R(0) ? sgD(sat(nh), gAlpha(rough)) * nh / (R(4) * vh) : R(0);
}
template <class R> R reflKernel(const V3<R>& n, const V3<R>& view, const V3<R>& light, R rough) {
    R nv = dot(n, view), nl = dot(n, light);
    if (rough <= R(0) || nv <= R(0) || nl <= R(0) || lengthSquared(view + light) < R(1.e-12f)) return R(0);
    R a = gAlpha(rough), nh = sat(dot(n, normalize(view + light)));
    return sgD(nh, a) * sgG1(nv, a) * sgG1(nl, a) / (R(4) * nv * nl);
}
template <class R> V3<R> reflSample(const V3<R>& n, const V3<R>& view, R rough, R u, R v) {
    if (rough <= R(0)) return R(2) * dot(n, view) * n - view;
    R a = gAlpha(rough);
    R c = std::sqrt((R(1) - u) / ((R(1) - u) + a * a * u));
    R sn = std::sqrt(std::fmax(R(0), R(1) - c * c)), phi = R(6.28318530718f) * v;
    V3<R> half = sgBasis(n, sn * fcos(phi), sn * fsin(phi), c);
    return R(2) * dot(view, half) * half - view;
}
template <class R>
V3<R> evalGltf(const GSurf<R>& s, const V3<R>& view, const V3<R>& light, R diffusePdf, R& pdf) {  // evaluateGltfPbrCos
    R wd, ws, wc;
    lobeWeights(s, view, wd, ws, wc);
    R nl = std::fmax(R(0), dot(s.normal, light));
    V3<R> rn = specN(s);
    pdf = wd * diffusePdf + ws * reflPdf(rn, view, light, s.roughness) + wc * reflPdf(s.coatNormal, view, light,
        s.coatRoughness);
    V3<R> base;
    if (lengthSquared(view + light) > R(1.e-12f)) {
        R vh = sat(dot(view, normalize(view + light)));
        R dw = (R(1) - s.metallic) * (R(1) - sgMax(dielF(s, vh)));
        if (dot(s.normal, view) > R(0)) base = s.color * (dw * nl * R(0.31830988618f));
        R sc = std::fmax(R(0), dot(rn, light));
        base += baseF(s, vh) * (reflKernel(rn, view, light, s.roughness) * sc);
    }
    R cc = std::fmax(R(0), dot(s.coatNormal, light));
    R cv = wc * reflKernel(s.coatNormal, view, light, s.coatRoughness) * cc;
    return base * (R(1) - wc) + V3<R>(cv, cv, cv);
}

// getGltfPbrSurface (:502-517) with correctGltfReflectionNormals.
template <class R>
GSurf<R> gltfAt(const oracle_handle& h, uint32_t m, const TexHit<R>& th, const V3<R>& shadedNormal,
    const V3<R>& view) {
    const wpt_material_ext& e = h.ext[m];
    R weight = R(e.specular_weight) * sampleTex(h, e.specular_weight_texture, th.uv, th.uv1, white<R>()).w;
    V3<R> color = hadamard(vec<R>(e.specular_color), gltfSrgb(sampleTex(h, e.specular_color_texture, th.uv, th.uv1,
        white<R>())));
    R ior = R(mat(h, m).ior);
    R f0 = (ior - R(1)) / (ior + R(1));
    f0 *= f0;
    GSurf<R> g;
    g.f0 = V3<R>(std::fmin(R(1), f0 * color.x) * weight, std::fmin(R(1), f0 * color.y) * weight,
        std::fmin(R(1), f0 * color.z) * weight);
    V3<R> cn = mapNormal(h, e.clearcoat_normal_texture, e.clearcoat_normal_scale, th);
    if (dot(cn, th.geometricNormal) < R(0)) cn = -cn;
    g.color = matAlbedo(h, m, th);
    g.normal = shadedNormal;
    g.coatNormal = cn;
    g.metallic = matMetallic(h, m, th);
    g.roughness = matRoughness(h, m, th);
    g.f90 = weight;
    g.coat = R(e.clearcoat) * sampleTex(h, e.clearcoat_texture, th.uv, th.uv1, white<R>()).x;
    g.coatRoughness = R(e.clearcoat_roughness) * sampleTex(h, e.clearcoat_roughness_texture, th.uv, th.uv1,
        white<R>()).y;
    g.reflectionNormal = reflectionFrame(g.normal, th.geometricNormal, view);
    if (g.coat > R(0)) g.coatNormal = reflectionFrame(g.coatNormal, th.geometricNormal, view);
    return g;
}

template <class R> V3<R> albedoOf(const wpt_material& m) { return vec<R>(m.albedo); }


// evaluateSurface / evaluateSurfaceCos (untextured), device_traversal.cuh:427-441, 519-527.
// Diffuse continuation density: half cosine, half environment when a
// usable map exists (diffuseEnvironmentProbability / diffuseContinuationPdf,
// device_traversal.cuh:403-418).
template <class R>
R envMixture(const oracle_handle& h) {
    return h.env && h.envIntensity > 0.0f && h.envW && h.envH ? R(0.5f) : R(0);
}

template <class R>
R diffusePdf(const oracle_handle& h, const V3<R>& n, const V3<R>& wi) {
    R p = envMixture<R>(h);
    R cosinePdf = std::fmax(R(0), dot(n, wi)) * R(0.31830988618f);
    return p > R(0) ? (R(1) - p) * cosinePdf + p * envPdf(h, wi) : cosinePdf;
}

// The material terms shadePaths reads at a hit (device_path_shading.cuh:79-87).
template <class R>
struct Vertex {
    uint32_t material;
    V3<R> n;        // shaded normal (normal-mapped, facing the geometric normal)
    V3<R> albedo;
    R roughness;
    uint32_t type;  // after the metallic switch
    bool gltf;
    GSurf<R> g;     // glTF surface, gltf only
};

// evaluateSurfaceCos (:519-527) over evaluateSurface (:427-441).
template <class R>
V3<R> evaluateSurfaceCos(const oracle_handle& h, const Vertex<R>& vx, const TexHit<R>& th, const V3<R>& wo,
    const V3<R>& wi, R& pdf) {
    if (vx.gltf) return evalGltf(vx.g, wo, wi, diffusePdf(h, vx.g.normal, wi), pdf);
    const wpt_material& m = mat(h, vx.material);
    const V3<R>& n = vx.n;
    V3<R> value;
    if (m.type == WPT_MATERIAL_SPECULAR_GLOSSINESS) {
        V3<R> f0 = hadamard(vec<R>(m.specular), gltfSrgb(sampleTex(h, h.ext[vx.material].specular_glossiness_texture,
            th.uv, th.uv1, white<R>())));
        value = evalSG(SG<R>{vx.albedo, f0, vx.roughness}, n, wo, wi, pdf);
    } else if (vx.type == WPT_MATERIAL_METAL) {
        if (vx.roughness <= R(0)) {
            pdf = R(0);
            value = V3<R>();
        } else {
            value = evalSG(SG<R>{V3<R>(), vx.albedo, vx.roughness}, n, wo, wi, pdf);
        }
    } else {
        pdf = diffusePdf(h, n, wi);
        value = vx.albedo * R(0.31830988618f);
    }
    return value * std::fmax(R(0), dot(n, wi));
}

template <class R>
TexHit<R> texHitOf(const Hit<R>& hit) {
    return TexHit<R>{hit.uv, hit.uv1, hit.normal, hit.geometricNormal, hit.tangent, hit.bitangent, hit.tangent1,
        hit.bitangent1};
}

template <class R>
V3<R> offsetRayOrigin(const V3<R>& p, const V3<R>& gn, const V3<R>& dir, R hitDistance) {  // ray.cuh:10-21
    R magnitude = std::fmax(std::fmax(R(1), std::fabs(hitDistance)),
        std::fmax(std::fabs(p.x), std::fmax(std::fabs(p.y), std::fabs(p.z))));
    R sign = dot(dir, gn) >= R(0) ? R(1) : R(-1);
    V3<R> offset = gn * (sign * R(32) * R(1.1920928955078125e-7f) * magnitude);
    V3<R> r = p + offset;
    const R inf = std::numeric_limits<R>::infinity();
    if (offset.x != R(0)) r.x = std::nextafter(r.x, offset.x > R(0) ? inf : -inf);
    if (offset.y != R(0)) r.y = std::nextafter(r.y, offset.y > R(0) ? inf : -inf);
    if (offset.z != R(0)) r.z = std::nextafter(r.z, offset.z > R(0) ? inf : -inf);
    return r;
}

struct Counters {
    uint64_t closest = 0, shadow = 0;
    WorkCounts closestWork, shadowWork;
};

// sampleDirectLight, device_traversal.cuh:529-569.
template <class R>
void sampleDirectLight(const oracle_handle& h, PathState<R>& path, const Hit<R>& hit, const Hit<R>& shaded,
    const Vertex<R>& vx, Counters& cnt) {
    uint32_t lightCount = static_cast<uint32_t>(h.lights.size());
    if (lightCount == 0) return;
    uint32_t candidate = std::min(static_cast<uint32_t>(randomFloat(path.rng) * lightCount), lightCount - 1);
    uint32_t li = randomFloat(path.rng) < h.alias[candidate].probability ? candidate : h.alias[candidate].alias;
    const TriangleLight& L = h.lights[li];
    V3<R> lv0(R(L.v0.x), R(L.v0.y), R(L.v0.z)), lv1(R(L.v1.x), R(L.v1.y), R(L.v1.z)), lv2(R(L.v2.x), R(L.v2.y), R(L.v2.z));
    V3<R> ln(R(L.normal.x), R(L.normal.y), R(L.normal.z));
    R rootU = std::sqrt(rnd<R>(path.rng));
    R v = rnd<R>(path.rng);
    V3<R> lp = (R(1) - rootU) * lv0 + rootU * (R(1) - v) * lv1 + rootU * v * lv2;
    V3<R> toLight = lp - hit.position;
    R d2 = lengthSquared(toLight);
    R dist = std::sqrt(d2);
    V3<R> dir = toLight / dist;
    R gc = std::fmax(R(0), dot(hit.geometricNormal, dir));
    R sc = std::fmax(R(0), dot(shaded.normal, dir));
    R lc = std::fmax(R(0), dot(ln, -dir));
    if (gc == R(0) || (!vx.gltf && sc == R(0)) || lc == R(0)) return;
    V3<R> so = offsetRayOrigin(hit.position, hit.geometricNormal, dir, hit.t);
    V3<R> st = offsetRayOrigin(lp, ln, -dir, dist);
    V3<R> seg = st - so;
    R sd = length(seg);
    if (!(sd > R(1.0e-6f))) return;
    Ray<R> shadow{so, seg / sd};
    if (!h.d.ignore_direct_light_occlusion) {
        ++cnt.shadow;
        tlWork = &cnt.shadowWork;
        bool blocked = isOccluded(h, shadow, sd - R(1.0e-6f));
        tlWork = nullptr;
        if (blocked) return;
    }
    const wpt_material& lm = mat(h, L.material);
    R area = R(L.area), pdfSel = R(L.selectionPdf);
    R geometryTerm = lc * area / (d2 * pdfSel);
    R lightPdf = pdfSel * d2 / (lc * area);
    R bsdfPdf;
    V3<R> bsdf = evaluateSurfaceCos(h, vx, texHitOf(hit), -path.ray.d, dir, bsdfPdf);
    R w = powerHeuristic(lightPdf, bsdfPdf);
    V3<R> direct = hadamard(bsdf, vec<R>(lm.emission)) * (geometryTerm * w);
    addPathRadiance(path, hadamard(path.throughput, direct));
}

// samplePunctualLight, punctual_light.cuh:24-50: direction toward the
// light, arriving intensity (inverse square, range window, spot cone) and
// distance.
template <class R>
V3<R> punctualSample(const wpt_punctual_light& L, const V3<R>& p, V3<R>& incident, R& distance) {
    distance = R(1.0e30f);
    if (L.type == WPT_LIGHT_DIRECTIONAL) {
        incident = vec<R>(L.intensity);
        return -vec<R>(L.direction);
    }
    V3<R> offset = vec<R>(L.position) - p;
    R d2 = lengthSquared(offset);
    if (!(d2 > R(0))) {
        incident = V3<R>();
        return V3<R>();
    }
    distance = std::sqrt(d2);
    V3<R> dir = offset / distance;
    R att = R(1) / d2;
    if (L.range > 0.0f) {
        R ratio = distance / R(L.range);
        att *= std::fmax(R(0), R(1) - ratio * ratio * ratio * ratio);
    }
    if (L.type == WPT_LIGHT_SPOT) {
        R c = dot(vec<R>(L.direction), -dir);
        R cone = std::fmin(R(1), std::fmax(R(0), (c - R(L.outer_cone_cos)) /
            std::fmax(R(1.0e-6f), R(L.inner_cone_cos) - R(L.outer_cone_cos))));
        att *= cone * cone;
    }
    incident = vec<R>(L.intensity) * att;
    return dir;
}

// addPunctualLight, device_traversal.cuh:653-674 (no MIS: a delta light).
template <class R>
void addPunctual(const oracle_handle& h, PathState<R>& path, const Hit<R>& hit, const Hit<R>& shaded,
    const Vertex<R>& vx, const wpt_punctual_light& L, R inverseSelectionPdf, Counters& cnt) {
    V3<R> incident;
    R distance;
    V3<R> dir = punctualSample(L, hit.position, incident, distance);
    if ((!vx.gltf && dot(shaded.normal, dir) <= R(0)) || dot(hit.geometricNormal, dir) <= R(0) ||
        lengthSquared(incident) == R(0))
        return;
    Ray<R> shadow{offsetRayOrigin(hit.position, hit.geometricNormal, dir, hit.t), dir};
    R reach = distance;
    if (L.type != WPT_LIGHT_DIRECTIONAL) {
        V3<R> segment = vec<R>(L.position) - shadow.o;
        reach = length(segment);
        if (!(reach > R(1.0e-6f))) return;
        shadow.d = segment / reach;
    }
    if (!h.d.ignore_direct_light_occlusion) {
        ++cnt.shadow;
        tlWork = &cnt.shadowWork;
        bool blocked = isOccluded(h, shadow, reach - R(1.0e-6f));
        tlWork = nullptr;
        if (blocked) return;
    }
    R ignored;
    V3<R> bsdf = evaluateSurfaceCos(h, vx, texHitOf(hit), -path.ray.d, dir, ignored);
    addPathRadiance(path, hadamard(path.throughput, hadamard(bsdf, incident) * inverseSelectionPdf));
}

// sampleDirectPunctualLight, device_traversal.cuh:676-683.
template <class R>
void samplePunctual(const oracle_handle& h, PathState<R>& path, const Hit<R>& hit, const Hit<R>& shaded,
    const Vertex<R>& vx, Counters& cnt) {
    for (uint32_t i = 0; i < h.directional; ++i) addPunctual(h, path, hit, shaded, vx, h.punctual[i], R(1), cnt);
    uint32_t n = static_cast<uint32_t>(h.punctual.size()) - h.directional;
    if (n == 0) return;
    uint32_t i = h.directional + std::min(static_cast<uint32_t>(randomFloat(path.rng) * n), n - 1);
    addPunctual(h, path, hit, shaded, vx, h.punctual[i], R(static_cast<float>(n)), cnt);
}

// The five draws of an environment sample (samplePathEnvironment,
// device_traversal.cuh:420-425): two for a FP64 alias candidate, coin,
// longitude, latitude.
template <class R>
V3<R> envDraw(const oracle_handle& h, RngState& rng, R& pdf) {
    float coarse = randomFloat(rng), fine = randomFloat(rng);
    double candidate = double(coarse) + double(fine) / 16777216.0;
    float coin = randomFloat(rng), lon = randomFloat(rng), lat = randomFloat(rng);
    return envSample<R>(h, candidate, coin, lon, lat, pdf);
}

// sampleDirectEnvironment, device_traversal.cuh:637-651.
template <class R>
void sampleDirectEnvironment(const oracle_handle& h, PathState<R>& path, const Hit<R>& hit, const Hit<R>& shaded,
    const Vertex<R>& vx, Counters& cnt) {
    if (!h.env || h.envIntensity == 0.0f) return;
    R pdf;
    V3<R> dir = envDraw<R>(h, path.rng, pdf);
    if ((!vx.gltf && dot(shaded.normal, dir) <= R(0)) || dot(hit.geometricNormal, dir) <= R(0) || !(pdf > R(0)))
        return;
    Ray<R> shadow{offsetRayOrigin(hit.position, hit.geometricNormal, dir, hit.t), dir};
    if (!h.d.ignore_direct_light_occlusion) {
        ++cnt.shadow;
        tlWork = &cnt.shadowWork;
        bool blocked = isOccluded(h, shadow, R(1.0e30f));
        tlWork = nullptr;
        if (blocked) return;
    }
    R bsdfPdf;
    V3<R> bsdf = evaluateSurfaceCos(h, vx, texHitOf(hit), -path.ray.d, dir, bsdfPdf);
    R w = powerHeuristic(pdf, bsdfPdf);
    addPathRadiance(path, hadamard(path.throughput, hadamard(bsdf, envRadiance(h, dir)) * (w / pdf)));
}

// Escape test of a dielectric branch toward the environment.
template <class R>
bool escapes(const oracle_handle& h, const Hit<R>& hit, const V3<R>& dir, Counters& cnt) {
    Ray<R> ray{offsetRayOrigin(hit.position, hit.geometricNormal, dir, hit.t), dir};
    ++cnt.shadow;
    tlWork = &cnt.shadowWork;
    bool blocked = isOccluded(h, ray, R(1.0e30f));
    tlWork = nullptr;
    return !blocked;
}

// ---------------------------------------------------------------- caustic photon map
// photon_mapping.cu:4-163 with render_driver.cu:180-234's grid.  A cell's
// photons are visited in photon order (the reference's linked lists are
// filled by atomic exchanges, in no fixed order).
template <class R>
struct OPhoton {
    V3<R> position, direction, flux;
};

template <class R>
struct PhotonMapT {
    std::vector<OPhoton<R>> photons;  // grouped by cell, photon order within a cell
    std::vector<uint32_t> start, end; // per cell
    uint32_t res = 64;
    R radius = R(0.06f);
    V3<R> lo, hi;
    uint32_t stored = 0;
};

template <class R>
bool photonCellOf(const PhotonMapT<R>& m, const V3<R>& p, int& x, int& y, int& z) {
    V3<R> ext = m.hi - m.lo, off = p - m.lo;
    R nx = off.x / ext.x, ny = off.y / ext.y, nz = off.z / ext.z;
    int r = static_cast<int>(m.res);
    auto clampi = [r](int v) { return std::min(std::max(v, 0), r - 1); };
    x = clampi(static_cast<int>(nx * R(static_cast<float>(m.res))));
    y = clampi(static_cast<int>(ny * R(static_cast<float>(m.res))));
    z = clampi(static_cast<int>(nz * R(static_cast<float>(m.res))));
    return !(nx < R(0) || ny < R(0) || nz < R(0) || nx > R(1) || ny > R(1) || nz > R(1));
}

// Bins stored photons (given in photon order) into the grid.
template <class R>
void binPhotons(PhotonMapT<R>& m, const std::vector<OPhoton<R>>& stored) {
    uint32_t cells = m.res * m.res * m.res;
    std::vector<std::vector<uint32_t>> bins(cells);
    for (uint32_t i = 0; i < stored.size(); ++i) {
        int x, y, z;
        if (!photonCellOf(m, stored[i].position, x, y, z)) continue;
        bins[(static_cast<uint32_t>(z) * m.res + static_cast<uint32_t>(y)) * m.res + static_cast<uint32_t>(x)].push_back(i);
    }
    m.start.assign(cells, 0);
    m.end.assign(cells, 0);
    m.photons.clear();
    for (uint32_t c = 0; c < cells; ++c) {
        m.start[c] = static_cast<uint32_t>(m.photons.size());
        for (uint32_t i : bins[c]) m.photons.push_back(stored[i]);
        m.end[c] = static_cast<uint32_t>(m.photons.size());
    }
    m.stored = static_cast<uint32_t>(m.photons.size());
}

// emitPhotons + tracePhotons for `count` photons; the survivors in photon order.
template <class R>
std::vector<OPhoton<R>> tracePhotonSet(const oracle_handle& h, uint32_t count, uint32_t maxDepth, uint32_t strategy,
    bool spectral) {
    std::vector<OPhoton<R>> out;
    uint32_t lights = static_cast<uint32_t>(h.lights.size());
    if (!lights) return out;
    for (uint32_t i = 0; i < count; ++i) {
        RngState rng = makeRngState(photonRngKey(i, 0x13579bdfu, strategy), strategy);
        uint32_t candidate = std::min(static_cast<uint32_t>(randomFloat(rng) * lights), lights - 1);
        uint32_t li = randomFloat(rng) < h.alias[candidate].probability ? candidate : h.alias[candidate].alias;
        const TriangleLight& L = h.lights[li];
        V3<R> v0(R(L.v0.x), R(L.v0.y), R(L.v0.z)), v1(R(L.v1.x), R(L.v1.y), R(L.v1.z)), v2(R(L.v2.x), R(L.v2.y), R(L.v2.z));
        V3<R> ln(R(L.normal.x), R(L.normal.y), R(L.normal.z));
        R rootU = std::sqrt(rnd<R>(rng));
        R v = rnd<R>(rng);
        V3<R> point = (R(1) - rootU) * v0 + rootU * ((R(1) - v) * v1 + v * v2);
        V3<R> dir = sampleCosineHemisphere(ln, rng);
        const wpt_material& em = mat(h, L.material);
        uint32_t channel = 3;
        V3<R> flux = vec<R>(em.emission) * (R(L.area) * R(3.14159265359f) / R(static_cast<float>(count)));
        if (spectral) {
            channel = std::min(static_cast<uint32_t>(randomFloat(rng) * 3.0f), 2u);
            R cf = R(em.emission[channel]) * (R(L.area) * R(3.14159265359f) * R(3) / R(static_cast<float>(count)));
            flux = V3<R>(channel == 0 ? cf : R(0), channel == 1 ? cf : R(0), channel == 2 ? cf : R(0));
        }
        Ray<R> ray{point + ln * R(0.001f), dir};
        uint32_t specular = 0;
        for (uint32_t depth = 0; depth < maxDepth; ++depth) {
            Hit<R> hit;
            HitState<R> hs;
            if (!traceRay(h, ray, hit, hs) || hit.material >= h.d.material_count) break;
            const wpt_material& m = mat(h, hit.material);
            V3<R> n = dot(hit.normal, -ray.d) >= R(0) ? hit.normal : -hit.normal;
            if (m.type == WPT_MATERIAL_DIFFUSE) {
                if (specular) out.push_back(OPhoton<R>{hit.position, -ray.d, flux});
                break;
            }
            V3<R> next;
            if (m.type == WPT_MATERIAL_METAL) {
                next = normalize(reflect(normalize(ray.d), n));
            } else if (m.type == WPT_MATERIAL_DIELECTRIC) {
                R ior = R(m.ior);
                float dispersion = h.ext[hit.material].dispersion;
                if (dispersion > 0.0f && channel < 3) ior += (R(static_cast<float>(channel)) - R(1)) * R(dispersion);
                Interface<R> b = dielectricInterface(normalize(ray.d), hit.geometricNormal, hit.normal, ior);
                bool reflected = b.reflectance >= R(1) || b.reflectance > rnd<R>(rng);
                next = reflected ? b.reflection : b.transmission;
            } else {
                next = sampleCosineHemisphere(n, rng);
            }
            if (m.type == WPT_MATERIAL_METAL || m.type == WPT_MATERIAL_DIELECTRIC) ++specular;
            flux = hadamard(flux, vec<R>(m.albedo));
            ray = Ray<R>{offsetRayOrigin(hit.position, hit.geometricNormal, next, hit.t), next};
        }
    }
    return out;
}

template <class R>
void photonGrid(const wpt_render_config& c, PhotonMapT<R>& m) {
    m.res = c.caustic_grid_resolution ? c.caustic_grid_resolution : 64u;
    m.radius = R(c.caustic_gather_radius > 0.0f ? c.caustic_gather_radius : 0.06f);
    bool box = c.caustic_grid_min[0] < c.caustic_grid_max[0] && c.caustic_grid_min[1] < c.caustic_grid_max[1] &&
        c.caustic_grid_min[2] < c.caustic_grid_max[2];
    m.lo = box ? vec<R>(c.caustic_grid_min) : V3<R>(R(-2.5f), R(-1.0f), R(-2.5f));
    m.hi = box ? vec<R>(c.caustic_grid_max) : V3<R>(R(2.5f), R(3.0f), R(2.5f));
}

// gatherCaustics, photon_mapping.cu:96-163.
template <class R>
V3<R> gatherPhotons(const PhotonMapT<R>& m, const V3<R>& p, const V3<R>& n) {
    const uint32_t kNearest = 64, kMinimum = 16;
    int cx, cy, cz;
    photonCellOf(m, p, cx, cy, cz);
    V3<R> ext = m.hi - m.lo;
    R search = std::fmax(m.radius * R(4), R(0.1f));
    int cellRadius = std::max(1, static_cast<int>(std::ceil(search * R(static_cast<float>(m.res)) /
        std::fmax(ext.x, R(0.0001f)))));
    R search2 = search * search;
    std::vector<R> dist;
    std::vector<uint32_t> idx;
    uint32_t far = 0;
    R far2 = R(0);
    int res = static_cast<int>(m.res);
    for (int z = cz - cellRadius; z <= cz + cellRadius; ++z)
        for (int y = cy - cellRadius; y <= cy + cellRadius; ++y)
            for (int x = cx - cellRadius; x <= cx + cellRadius; ++x) {
                if (x < 0 || y < 0 || z < 0 || x >= res || y >= res || z >= res) continue;
                uint32_t cell = (static_cast<uint32_t>(z) * m.res + static_cast<uint32_t>(y)) * m.res + static_cast<uint32_t>(x);
                for (uint32_t i = m.start[cell]; i < m.end[cell]; ++i) {
                    V3<R> d = m.photons[i].position - p;
                    R d2 = dot(d, d);
                    if (d2 > search2 || dot(n, m.photons[i].direction) <= R(0)) continue;
                    if (dist.size() < kNearest) {
                        dist.push_back(d2);
                        idx.push_back(i);
                        if (d2 > far2) {
                            far2 = d2;
                            far = static_cast<uint32_t>(dist.size() - 1);
                        }
                    } else if (d2 < far2) {  // replace the farthest, then find the new farthest
                        dist[far] = d2;
                        idx[far] = i;
                        far = 0;
                        far2 = dist[0];
                        for (uint32_t k = 1; k < kNearest; ++k)
                            if (dist[k] > far2) {
                                far2 = dist[k];
                                far = k;
                            }
                    }
                }
            }
    if (dist.empty()) return V3<R>();
    R g2 = search2;
    if (dist.size() >= kMinimum) {
        g2 = R(0);
        for (R d2 : dist) g2 = std::fmax(g2, d2);
        g2 = std::fmax(g2, m.radius * m.radius);
    }
    R g = std::sqrt(g2);
    V3<R> sum;
    for (size_t k = 0; k < dist.size(); ++k) {
        const OPhoton<R>& ph = m.photons[idx[k]];
        R cone = std::fmax(R(0), R(1) - std::sqrt(dist[k]) / g);
        sum += ph.flux * (cone * dot(n, ph.direction));
    }
    return sum * (R(3) / (R(3.14159265359f) * g2));
}

// continueGltfPbrSurface, device_traversal.cuh:727-787.
template <class R>
V3<R> gltfContinue(const oracle_handle& h, PathState<R>& path, const Hit<R>& hit, const GSurf<R>& g,
    uint32_t maxDepth, Counters& cnt) {
    V3<R> view = -path.ray.d;
    R wd, ws, wc;
    lobeWeights(g, view, wd, ws, wc);
    (void)wd;
    R choice = rnd<R>(path.rng);
    bool coat = choice < wc;
    bool spec = !coat && choice < wc + ws;
    bool delta = (coat && g.coatRoughness <= R(0)) || (spec && g.roughness <= R(0));
    bool baseOut = false, coatOut = false;
    V3<R> baseCoef = baseF(g, dot(specN(g), view)) * (R(1) - wc);
    if (h.env && path.depth + 1 < maxDepth) {  // delta lobes toward the map, integrated here
        for (int lobe = 0; lobe < 2; ++lobe) {
            bool isCoat = lobe == 1;
            R mass = isCoat ? wc : ws;
            R rough = isCoat ? g.coatRoughness : g.roughness;
            if (!(mass > R(0)) || rough > R(0)) continue;
            V3<R> n = isCoat ? g.coatNormal : specN(g);
            V3<R> dir = reflSample(n, view, R(0), R(0), R(0));
            if (dot(n, view) <= R(0) || dot(dir, hit.geometricNormal) <= R(0)) continue;
            bool out = escapes(h, hit, dir, cnt);
            if (isCoat) coatOut = out;
            else baseOut = out;
            if (out) {
                V3<R> coef = isCoat ? V3<R>(wc, wc, wc) : baseCoef;
                addPathRadiance(path, hadamard(path.throughput, hadamard(coef, envRadiance(h, dir))));
            }
        }
    }
    V3<R> dir;
    if (coat || spec) {
        R u = rnd<R>(path.rng), v = rnd<R>(path.rng);
        V3<R> n = coat ? g.coatNormal : specN(g);
        dir = reflSample(n, view, coat ? g.coatRoughness : g.roughness, u, v);
        if (dot(n, view) <= R(0)) {
            path.active = false;
            return dir;
        }
    } else {
        R mix = envMixture<R>(h);
        if (mix > R(0) && rnd<R>(path.rng) < mix) {
            R ignored;
            dir = envDraw<R>(h, path.rng, ignored);
        } else {
            dir = sampleCosineHemisphere(g.normal, path.rng);
        }
    }
    if (dot(dir, hit.geometricNormal) <= R(0)) {
        path.active = false;
        return dir;
    }
    if (delta) {
        R mass = coat ? wc : ws;
        V3<R> coef = coat ? V3<R>(wc, wc, wc) : baseCoef;
        path.throughput = hadamard(path.throughput, coef) / mass;
        path.specularBounce = true;
        path.previousBsdfPdf = R(0);
        if (coat ? coatOut : baseOut) path.active = false;
    } else {
        R pdf;
        V3<R> value = evalGltf(g, view, dir, diffusePdf(h, g.normal, dir), pdf);
        if (!(pdf > R(0))) {
            path.active = false;
            return dir;
        }
        path.throughput = hadamard(path.throughput, value) / pdf;
        path.previousBsdfPdf = pdf;
        path.specularBounce = false;
    }
    return dir;
}

// One path, shadePaths (device_path_shading.cuh:6-173) applied bounce by bounce.
template <class R>
V3<R> tracePath(const oracle_handle& h, PathState<R> path, uint32_t maxDepth, uint32_t rrStart, int debug,
    Counters& cnt, const PhotonMapT<R>* photons = nullptr) {
    const R invPi = R(0.31830988618f);
    for (uint32_t bounce = 0; bounce < maxDepth; ++bounce) {
        Hit<R> hit;
        HitState<R> hs;
        ++cnt.closest;
        tlWork = &cnt.closestWork;
        bool didHit = traceRay(h, path.ray, hit, hs);
        tlWork = nullptr;
        if (debug) {  // 1: intersection normals, 2: shading normals (:18-36)
            V3<R> color(R(0.05f), R(0.10f), R(0.25f));
            if (didHit) {
                const wpt_material_ext& e = h.ext[hit.material];
                V3<R> n = debug == 1 ? hit.normal : mapNormal(h, e.normal_texture, e.normal_scale, texHitOf(hit));
                color = R(0.5f) * (n + V3<R>(R(1), R(1), R(1)));
            }
            return color;
        }
        if (!didHit) {  // addEnvironmentOrSky, device_traversal.cuh:577-587
            if (h.env) {
                R w = path.depth > 0 && !path.specularBounce ? powerHeuristic(path.previousBsdfPdf, envPdf(h, path.ray.d))
                                                             : R(1);
                addPathRadiance(path, hadamard(path.throughput, envRadiance(h, path.ray.d)) * w, true);
            } else if (!h.d.black_background) {
                R t = R(0.5f) * (path.ray.d.y + R(1));
                V3<R> sky = (R(1) - t) * V3<R>(R(1), R(1), R(1)) + t * V3<R>(R(0.5f), R(0.7f), R(1.0f));
                addPathRadiance(path, hadamard(path.throughput, sky), true);
            }
            return path.radiance;
        }
        const wpt_material& m = mat(h, hit.material);
        if (path.mediumMaterial != kInvalid) {
            const wpt_material& medium = mat(h, path.mediumMaterial);
            if (medium.attenuation_distance > 0.0f) {
                R dist = hit.t * path.mediumDensity;
                R ad = R(medium.attenuation_distance);
                V3<R> sigma(-flog(std::fmax(R(medium.attenuation_color[0]), R(0.0001f))) / ad,
                    -flog(std::fmax(R(medium.attenuation_color[1]), R(0.0001f))) / ad,
                    -flog(std::fmax(R(medium.attenuation_color[2]), R(0.0001f))) / ad);
                path.throughput = hadamard(path.throughput, V3<R>(fexp(-sigma.x * dist), fexp(-sigma.y * dist), fexp(-sigma.z * dist)));
            }
        }
        if (m.type == WPT_MATERIAL_EMISSIVE) {
            R w = R(1);
            if (hit.lightIndex != kInvalid) {
                const TriangleLight& L = h.lights[hit.lightIndex];
                R lc = std::fmax(R(0), dot(V3<R>(R(L.normal.x), R(L.normal.y), R(L.normal.z)), -path.ray.d));
                if (lc == R(0)) {
                    w = R(0);
                } else if (path.depth > 0 && !path.specularBounce) {
                    R lightPdf = R(L.selectionPdf) * hit.t * hit.t / (lc * R(L.area));
                    w = powerHeuristic(path.previousBsdfPdf, lightPdf);
                }
            }
            addPathRadiance(path, hadamard(path.throughput, matEmission(h, hit.material, texHitOf(hit))) * w, true);
            return path.radiance;
        }
        const wpt_material_ext& ext = h.ext[hit.material];
        TexHit<R> th = texHitOf(hit);
        Hit<R> shaded = hit;
        shaded.normal = mapNormal(h, ext.normal_texture, ext.normal_scale, th);
        if (dot(shaded.normal, hit.geometricNormal) < R(0)) shaded.normal = -shaded.normal;
        V3<R> albedo = matAlbedo(h, hit.material, th);
        uint32_t type = (m.type == WPT_MATERIAL_DIFFUSE && matMetallic(h, hit.material, th) > R(0.5f))
            ? WPT_MATERIAL_METAL : m.type;
        R roughness = matRoughness(h, hit.material, th);
        Vertex<R> vx{hit.material, shaded.normal, albedo, roughness, type, ext.gltf_pbr_workflow != 0, GSurf<R>{}};
        V3<R> direction;
        auto glossy = [&]() {  // continueGlossySurface, device_traversal.cuh:706-725
            SG<R> s = m.type == WPT_MATERIAL_SPECULAR_GLOSSINESS
                ? SG<R>{albedo, hadamard(vec<R>(m.specular), gltfSrgb(sampleTex(h, ext.specular_glossiness_texture,
                      th.uv, th.uv1, white<R>()))), roughness}
                : SG<R>{V3<R>(), albedo, roughness};
            V3<R> wo = -path.ray.d;
            R choice = rnd<R>(path.rng);
            R u = rnd<R>(path.rng);
            R v = rnd<R>(path.rng);
            V3<R> dir = sampleSG(s, shaded.normal, wo, choice, u, v);
            R pdf;
            V3<R> f = evalSG(s, shaded.normal, wo, dir, pdf);
            if (!(pdf > R(0)) || dot(dir, hit.geometricNormal) <= R(0)) {
                path.active = false;
                return dir;
            }
            path.throughput = hadamard(path.throughput, f) * (std::fmax(R(0), dot(shaded.normal, dir)) / pdf);
            path.previousBsdfPdf = pdf;
            path.specularBounce = false;
            return dir;
        };
        if (vx.gltf) {
            vx.g = gltfAt(h, hit.material, th, shaded.normal, -path.ray.d);
            sampleDirectLight(h, path, hit, shaded, vx, cnt);
            samplePunctual(h, path, hit, shaded, vx, cnt);
            sampleDirectEnvironment(h, path, hit, shaded, vx, cnt);
            direction = gltfContinue(h, path, hit, vx.g, maxDepth, cnt);
        } else switch (type) {
        case WPT_MATERIAL_DIFFUSE: {
            sampleDirectLight(h, path, hit, shaded, vx, cnt);
            samplePunctual(h, path, hit, shaded, vx, cnt);
            sampleDirectEnvironment(h, path, hit, shaded, vx, cnt);
            if (photons)  // caustics gathered at the diffuse vertex (device_path_shading.cuh:99-102)
                addPathRadiance(path, hadamard(path.throughput, hadamard(albedo, gatherPhotons(*photons, hit.position,
                    shaded.normal))));
            // continueDiffuseSurface, device_traversal.cuh:685-704
            R mix = envMixture<R>(h);
            if (mix > R(0) && rnd<R>(path.rng) < mix) {
                R ignored;
                direction = envDraw<R>(h, path.rng, ignored);
            } else {
                direction = sampleCosineHemisphere(shaded.normal, path.rng);
            }
            R cosine = dot(shaded.normal, direction);
            R pdf = diffusePdf(h, shaded.normal, direction);
            if (!(pdf > R(0)) || cosine <= R(0) || dot(direction, hit.geometricNormal) <= R(0)) {
                path.active = false;
            } else {
                path.throughput = hadamard(path.throughput, albedo) * (cosine * invPi / pdf);
                path.previousBsdfPdf = pdf;
                path.specularBounce = false;
            }
            break;
        }
        case WPT_MATERIAL_SPECULAR_GLOSSINESS:
            sampleDirectLight(h, path, hit, shaded, vx, cnt);
            samplePunctual(h, path, hit, shaded, vx, cnt);
            sampleDirectEnvironment(h, path, hit, shaded, vx, cnt);
            direction = glossy();
            break;
        case WPT_MATERIAL_METAL:
            if (roughness > R(0)) {
                sampleDirectLight(h, path, hit, shaded, vx, cnt);
                samplePunctual(h, path, hit, shaded, vx, cnt);
                sampleDirectEnvironment(h, path, hit, shaded, vx, cnt);
                direction = glossy();
                break;
            }
            path.throughput = hadamard(path.throughput, albedo);
            {
                V3<R> reflected = reflect(normalize(path.ray.d), shaded.normal);
                direction = normalize(reflected + roughness * randomInUnitSphere<R>(path.rng));
                if (dot(direction, shaded.normal) <= R(0) || dot(direction, hit.geometricNormal) <= R(0))
                    direction = reflect(normalize(path.ray.d), hit.geometricNormal);
            }
            path.specularBounce = true;
            path.previousBsdfPdf = R(0);
            break;
        case WPT_MATERIAL_DIELECTRIC: {  // continueDielectricSurface, device_traversal.cuh:589-635
            path.throughput = hadamard(path.throughput, albedo);
            Interface<R> b = dielectricInterface(normalize(path.ray.d), hit.geometricNormal, shaded.normal, R(m.ior));
            R eta2 = b.eta * b.eta;
            // With a map, both branches that escape are integrated here.
            bool reflectionOut = false, transmissionOut = false;
            if (h.env && path.depth + 1 < maxDepth) {
                if (b.reflectance > R(0) && (reflectionOut = escapes(h, hit, b.reflection, cnt)))
                    addPathRadiance(path, hadamard(path.throughput, envRadiance(h, b.reflection)) * b.reflectance);
                if (b.reflectance < R(1) && (transmissionOut = escapes(h, hit, b.transmission, cnt)))
                    addPathRadiance(path, hadamard(path.throughput, envRadiance(h, b.transmission)) *
                        ((R(1) - b.reflectance) * eta2));
            }
            bool reflected = b.reflectance >= R(1) || b.reflectance > rnd<R>(path.rng);
            if (!reflected) {
                path.throughput *= eta2;
                path.etaScale = path.etaScale / eta2;
                if (b.frontFace) {
                    path.mediumMaterial = hit.material;
                    path.mediumDensity = R(m.volume_density) * matThickness(h, hit.material, th);
                } else {
                    path.mediumMaterial = kInvalid;
                    path.mediumDensity = R(1);
                }
            }
            if (reflected ? reflectionOut : transmissionOut) path.active = false;
            path.specularBounce = true;
            path.previousBsdfPdf = R(0);
            direction = reflected ? b.reflection : b.transmission;
            break;
        }
        default:
            break;
        }
        if (!path.active) return path.radiance;
        path.ray.o = offsetRayOrigin(hit.position, hit.geometricNormal, direction, hit.t);
        path.ray.d = direction;
        ++path.depth;
        if (path.depth >= maxDepth) return path.radiance;
        if (path.depth >= rrStart) {
            R survival = std::fmin(R(0.95f), std::fmax(R(0.05f), path.etaScale * std::fmax(path.throughput.x,
                std::fmax(path.throughput.y, path.throughput.z))));
            if (rnd<R>(path.rng) > survival) return path.radiance;
            path.throughput *= R(1) / survival;
        }
    }
    return path.radiance;
}

// generatePrimaryRays + Camera::generateRay (renderer.cu:10-42, camera.cuh:17-27).
template <class R>
PathState<R> primary(const wpt_camera& cam, uint32_t x, uint32_t y, uint32_t width, uint32_t height, uint32_t sample,
    uint32_t strategy, float clamp) {
    PathState<R> p{};
    p.pixelIndex = y * width + x;
    p.rng = makeRngState(pixelRngKey(p.pixelIndex, sample, strategy), strategy);
    R u = (R(x) + rnd<R>(p.rng)) / R(width);
    R v = (R(y) + rnd<R>(p.rng)) / R(height);
    R tanHalf = ftan(R(cam.vertical_fov) * R(0.5f));
    R px = (R(2) * u - R(1)) * R(cam.aspect_ratio) * tanHalf;
    R py = (R(1) - R(2) * v) * tanHalf;
    p.ray.o = vec<R>(cam.position);
    p.ray.d = normalize(vec<R>(cam.forward) + vec<R>(cam.right) * px + vec<R>(cam.up) * py);
    p.throughput = V3<R>(R(1), R(1), R(1));
    p.radiance = V3<R>();
    p.depth = 0;
    p.mediumMaterial = kInvalid;
    p.mediumDensity = R(1);
    p.previousBsdfPdf = R(0);
    p.specularBounce = true;
    p.active = true;
    p.etaScale = R(1);
    p.clamp = R(clamp);
    return p;
}

template <class R>
void renderRows(const oracle_handle& h, const wpt_camera& cam, const wpt_render_config& cfg, uint32_t x0, uint32_t y0,
    uint32_t w, uint32_t th, uint32_t s0, uint32_t sc, uint32_t first, uint32_t step, float* fb, Counters& cnt,
    const PhotonMapT<R>* photons) {
    int debug = cfg.intersection_debug ? 1 : (cfg.shading_normal_debug ? 2 : 0);
    for (uint32_t row = first; row < th; row += step) {
        uint32_t y = y0 + row;
        for (uint32_t x = x0; x < x0 + w; ++x) {
            float* px = fb + 3 * (static_cast<size_t>(y) * cfg.width + x);
            for (uint32_t s = s0; s < s0 + sc; ++s) {
                PathState<R> p = primary<R>(cam, x, y, cfg.width, cfg.height, s, cfg.rng_strategy, cfg.indirect_clamp);
                V3<R> L = tracePath(h, p, cfg.max_depth, cfg.rr_start_depth, debug, cnt, photons);
                px[0] += static_cast<float>(L.x);
                px[1] += static_cast<float>(L.y);
                px[2] += static_cast<float>(L.z);
            }
        }
    }
}


// ---------------------------------------------------------------- differentiable rendering
// Restatement of the differentiable estimator of wpt_diff (no reference
// counterpart): FP32 for parity with the device, FP64 for the host adjoint,
// and dual numbers (value, derivative along one parameter direction) for an
// independent forward-mode derivative.

struct Dual {
    double v, d;
    Dual(double a = 0.0, double b = 0.0) : v(a), d(b) {}
    Dual operator-() const { return Dual(-v, -d); }
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
inline Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
inline Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
inline Dual dsqrt(Dual a) {
    double s = std::sqrt(a.v);
    return Dual(s, a.d / (2.0 * s));
}
inline float dsqrt(float a) { return std::sqrt(a); }
inline double dsqrt(double a) { return std::sqrt(a); }
inline double value(Dual a) { return a.v; }
inline double value(double a) { return a; }
inline float value(float a) { return a; }

struct DiffSetup {
    uint32_t texW = 0, texH = 0, texturedMaterial = kInvalid;
    std::vector<uint32_t> staticSlot, lightSlot;
    uint32_t roughCount = 0, emitCount = 0, texFloats = 0;
    const float* params = nullptr;
    const float* direction = nullptr;  // dual seed, parameter layout
    uint32_t width = 0, height = 0, spp = 0, maxDepth = 0, strategy = 0, sampleStart = 0;
};

template <class P> P paramAt(const DiffSetup& s, size_t i) { return P(s.params[i]); }
template <> Dual paramAt<Dual>(const DiffSetup& s, size_t i) {
    return Dual(s.params[i], s.direction ? static_cast<double>(s.direction[i]) : 0.0);
}

DiffSetup diffSetup(const oracle_handle& h, const oracle_diff_config& c, const float* params) {
    DiffSetup s;
    bool textured = c.textured_material != kInvalid;
    s.texW = textured ? c.texture_width : 0;
    s.texH = textured ? c.texture_height : 0;
    s.texturedMaterial = textured ? c.textured_material : kInvalid;
    s.texFloats = s.texW * s.texH * 3;
    s.staticSlot.assign(h.d.static_triangle_count, kInvalid);
    s.lightSlot.assign(std::max<size_t>(h.lights.size(), 1), kInvalid);
    for (uint32_t i = 0; i < h.d.static_triangle_count; ++i) {
        uint32_t type = mat(h, h.d.static_triangles[i].material).type;
        if (type == WPT_MATERIAL_EMISSIVE) {
            s.staticSlot[i] = s.emitCount;
            if (h.staticLight[i] != kInvalid) s.lightSlot[h.staticLight[i]] = s.emitCount;
            ++s.emitCount;
        } else if (type != WPT_MATERIAL_DIELECTRIC) {
            s.staticSlot[i] = s.roughCount++;
        }
    }
    s.params = params;
    s.width = c.width;
    s.height = c.height;
    s.spp = c.samples_per_pixel;
    s.maxDepth = c.max_depth;
    s.strategy = c.rng_strategy;
    return s;
}

template <class R>
struct Footprint {
    uint32_t index[4];
    R weight[4];
};

template <class R>
Footprint<R> footprint(const DiffSetup& s, const V2<R>& uv) {
    R x = uv.x * R(s.texW) - R(0.5f);
    R y = uv.y * R(s.texH) - R(0.5f);
    R fx = std::floor(x), fy = std::floor(y);
    R tx = x - fx, ty = y - fy;
    int ix = static_cast<int>(fx), iy = static_cast<int>(fy);
    int mx = static_cast<int>(s.texW) - 1, my = static_cast<int>(s.texH) - 1;
    auto clampi = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
    int x0 = clampi(ix, mx), x1 = clampi(ix + 1, mx), y0 = clampi(iy, my), y1 = clampi(iy + 1, my);
    Footprint<R> f;
    f.index[0] = static_cast<uint32_t>(y0) * s.texW + static_cast<uint32_t>(x0);
    f.index[1] = static_cast<uint32_t>(y0) * s.texW + static_cast<uint32_t>(x1);
    f.index[2] = static_cast<uint32_t>(y1) * s.texW + static_cast<uint32_t>(x0);
    f.index[3] = static_cast<uint32_t>(y1) * s.texW + static_cast<uint32_t>(x1);
    f.weight[0] = (R(1) - tx) * (R(1) - ty);
    f.weight[1] = tx * (R(1) - ty);
    f.weight[2] = (R(1) - tx) * ty;
    f.weight[3] = tx * ty;
    return f;
}

// GGX lobe F0 D G / (4 cos_i cos_o) with alpha = roughness, and d/dalpha.
template <class R, class P>
void ggxLobe(const V3<R>& n, const V3<R>& wo, const V3<R>& wi, P alpha, P& value, P& dAlpha) {
    const R pi = R(3.14159265358979f);
    value = P(0);
    dAlpha = P(0);
    R ci = dot(n, wi), co = dot(n, wo);
    if (!(ci > R(0)) || !(co > R(0))) return;
    V3<R> hv = normalize(wi + wo);
    R nh = std::fmax(R(0), dot(n, hv));
    P c = P(nh * nh);
    P a2 = alpha * alpha;
    P k = c * (a2 - P(1)) + P(1);
    P d = a2 / (P(pi) * k * k);
    P dD = P(2) * alpha * (k - P(2) * a2 * c) / (P(pi) * k * k * k);
    auto g1 = [&](R mu, P& dG) {
        P m = P(mu);
        P s = dsqrt(a2 + (P(1) - a2) * m * m);
        P denom = m + s;
        dG = P(-2) * m * alpha * (P(1) - m * m) / (s * denom * denom);
        return P(2) * m / denom;
    };
    P dGi, dGo;
    P gi = g1(ci, dGi);
    P go = g1(co, dGo);
    P scale = P(R(0.04f)) / P(R(4) * ci * co);
    value = scale * (d * (gi * go));
    dAlpha = scale * (dD * (gi * go) + d * (dGi * go + gi * dGo));
}

// Record of one vertex for the FP64 reverse sweep.
struct DiffVertex {
    int kind = 0;  // 1 emitter, 2 diffuse surface, 3 dielectric
    V3<double> T;
    // emitter
    uint32_t eslot = kInvalid;
    V3<double> em;  // emission * MIS weight
    // surface
    bool textured = false;
    Footprint<double> fp{};
    uint32_t rslot = kInvalid;
    bool nee = false;
    V3<double> neeBase, neeF;  // contribution = T (neeF neeBase) e
    double neeE = 1.0, neeDAlpha = 0.0;
    uint32_t neeSlot = kInvalid;
    bool cont = false;
    V3<double> weight;  // continuation (or dielectric) throughput factor
    double cosOverPdf = 0.0, contDAlpha = 0.0;
};

// One path of the differentiable estimator; R geometry, P parameter values.
template <class R, class P>
V3<P> diffPathOracle(const oracle_handle& h, const DiffSetup& s, const wpt_camera& cam, uint32_t x, uint32_t y,
    uint32_t sample, std::vector<DiffVertex>* record) {
    const R invPi = R(0.31830988618f);
    PathState<R> path = primary<R>(cam, x, y, s.width, s.height, sample, s.strategy, 0.0f);
    V3<P> T(P(1), P(1), P(1));
    V3<P> acc(P(0), P(0), P(0));
    for (uint32_t depth = 0; depth < s.maxDepth; ++depth) {
        Hit<R> hit;
        HitState<R> hs;
        if (!traceRay(h, path.ray, hit, hs)) break;
        const wpt_material& m = mat(h, hit.material);
        DiffVertex vtx;
        if (record) {
            vtx.T = V3<double>(value(T.x), value(T.y), value(T.z));
        }
        if (m.type == WPT_MATERIAL_EMISSIVE) {
            R w = R(1);
            uint32_t slot = kInvalid;
            if (hit.lightIndex != kInvalid) {
                slot = s.lightSlot[hit.lightIndex];
                const TriangleLight& L = h.lights[hit.lightIndex];
                R lc = std::fmax(R(0), dot(V3<R>(R(L.normal.x), R(L.normal.y), R(L.normal.z)), -path.ray.d));
                if (lc == R(0)) {
                    w = R(0);
                } else if (depth > 0 && !path.specularBounce) {
                    R lightPdf = R(L.selectionPdf) * hit.t * hit.t / (lc * R(L.area));
                    w = powerHeuristic(path.previousBsdfPdf, lightPdf);
                }
            }
            V3<P> em(P(R(m.emission[0])), P(R(m.emission[1])), P(R(m.emission[2])));
            V3<P> term = hadamard(T, em) * P(w);
            P e = slot == kInvalid ? P(1) : paramAt<P>(s, s.texFloats + s.roughCount + slot);
            acc += term * e;
            if (record) {
                vtx.kind = 1;
                vtx.eslot = slot;
                vtx.em = V3<double>(double(R(m.emission[0])) * double(w), double(R(m.emission[1])) * double(w),
                    double(R(m.emission[2])) * double(w));
                record->push_back(vtx);
            }
            break;
        }
        V3<R> n = hit.normal;
        if (dot(n, hit.geometricNormal) < R(0)) n = -n;
        V3<R> direction;
        if (m.type == WPT_MATERIAL_DIELECTRIC) {
            V3<P> albedo(P(R(m.albedo[0])), P(R(m.albedo[1])), P(R(m.albedo[2])));
            T = hadamard(T, albedo);
            Interface<R> b = dielectricInterface(normalize(path.ray.d), hit.geometricNormal, n, R(m.ior));
            R eta2 = b.eta * b.eta;
            bool reflected = b.reflectance >= R(1) || b.reflectance > rnd<R>(path.rng);
            if (!reflected) T *= P(eta2);
            path.specularBounce = true;
            path.previousBsdfPdf = R(0);
            direction = reflected ? b.reflection : b.transmission;
            if (record) {
                vtx.kind = 3;
                double f = reflected ? 1.0 : double(eta2);
                vtx.weight = V3<double>(double(R(m.albedo[0])) * f, double(R(m.albedo[1])) * f, double(R(m.albedo[2])) * f);
                record->push_back(vtx);
            }
        } else {
            V3<R> wo = -path.ray.d;
            // Surface: albedo (texture or material), roughness slot.
            V3<P> albedo(P(R(m.albedo[0])), P(R(m.albedo[1])), P(R(m.albedo[2])));
            uint32_t rslot = hs.kind == WPT_HIT_STATIC_TRIANGLE ? s.staticSlot[hs.prim] : kInvalid;
            P alpha = rslot != kInvalid ? paramAt<P>(s, s.texFloats + rslot) : P(0);
            bool textured = hs.kind == WPT_HIT_STATIC_TRIANGLE && hit.material == s.texturedMaterial && s.texW > 0;
            Footprint<R> fp{};
            if (textured) {
                fp = footprint(s, hit.uv);
                albedo = V3<P>(P(0), P(0), P(0));
                for (int k = 0; k < 4; ++k) {
                    size_t base = 3u * static_cast<size_t>(fp.index[k]);
                    V3<P> t(paramAt<P>(s, base), paramAt<P>(s, base + 1), paramAt<P>(s, base + 2));
                    albedo = albedo + t * P(fp.weight[k]);
                }
            }
            auto bsdf = [&](const V3<R>& wi, P& dAlpha) {
                V3<P> f = albedo * P(invPi);
                dAlpha = P(0);
                if (rslot != kInvalid) {
                    P value;
                    ggxLobe<R, P>(n, wo, wi, alpha, value, dAlpha);
                    f = f + V3<P>(value, value, value);
                }
                return f;
            };
            if (record) {
                vtx.kind = 2;
                vtx.textured = textured;
                vtx.rslot = rslot;
                for (int k = 0; k < 4; ++k) {
                    vtx.fp.index[k] = fp.index[k];
                    vtx.fp.weight[k] = double(fp.weight[k]);
                }
            }
            uint32_t lightCount = static_cast<uint32_t>(h.lights.size());
            if (lightCount != 0) {
                uint32_t candidate = std::min(static_cast<uint32_t>(randomFloat(path.rng) * lightCount), lightCount - 1);
                uint32_t li = randomFloat(path.rng) < h.alias[candidate].probability ? candidate : h.alias[candidate].alias;
                const TriangleLight& L = h.lights[li];
                V3<R> lv0(R(L.v0.x), R(L.v0.y), R(L.v0.z)), lv1(R(L.v1.x), R(L.v1.y), R(L.v1.z)),
                    lv2(R(L.v2.x), R(L.v2.y), R(L.v2.z));
                V3<R> ln(R(L.normal.x), R(L.normal.y), R(L.normal.z));
                R rootU = std::sqrt(rnd<R>(path.rng));
                R v = rnd<R>(path.rng);
                V3<R> lp = (R(1) - rootU) * lv0 + rootU * (R(1) - v) * lv1 + rootU * v * lv2;
                V3<R> toLight = lp - hit.position;
                R d2 = lengthSquared(toLight);
                R dist = std::sqrt(d2);
                V3<R> wl = toLight / dist;
                R gc = std::fmax(R(0), dot(hit.geometricNormal, wl));
                R sc = std::fmax(R(0), dot(n, wl));
                R lc = std::fmax(R(0), dot(ln, -wl));
                if (gc != R(0) && sc != R(0) && lc != R(0)) {
                    V3<R> so = offsetRayOrigin(hit.position, hit.geometricNormal, wl, hit.t);
                    V3<R> st = offsetRayOrigin(lp, ln, -wl, dist);
                    V3<R> seg = st - so;
                    R sd = length(seg);
                    if (sd > R(1.0e-6f) && !isOccluded(h, Ray<R>{so, seg / sd}, sd - R(1.0e-6f))) {
                        R geometryTerm = lc * R(L.area) / (d2 * R(L.selectionPdf));
                        R lightPdf = R(L.selectionPdf) * d2 / (lc * R(L.area));
                        R bsdfPdf = sc * invPi;
                        R g = geometryTerm * powerHeuristic(lightPdf, bsdfPdf);
                        uint32_t slot = s.lightSlot[li];
                        P e = slot == kInvalid ? P(1) : paramAt<P>(s, s.texFloats + s.roughCount + slot);
                        P dAlpha;
                        V3<P> f = bsdf(wl, dAlpha);
                        const wpt_material& lm = mat(h, L.material);
                        V3<P> le(P(R(lm.emission[0])), P(R(lm.emission[1])), P(R(lm.emission[2])));
                        V3<P> base = hadamard(T, le) * P(g * sc);
                        V3<P> fb = hadamard(base, f);
                        acc += fb * e;
                        if (record) {
                            vtx.nee = true;
                            R gs = g * sc;
                            vtx.neeBase = V3<double>(double(R(lm.emission[0])) * double(gs),
                                double(R(lm.emission[1])) * double(gs), double(R(lm.emission[2])) * double(gs));
                            vtx.neeF = V3<double>(value(f.x), value(f.y), value(f.z));
                            vtx.neeE = value(e);
                            vtx.neeSlot = slot;
                            vtx.neeDAlpha = value(dAlpha);
                        }
                    }
                }
            }
            direction = sampleCosineHemisphere(n, path.rng);
            R cosine = dot(n, direction);
            R pdf = std::fmax(R(0), cosine) * invPi;
            bool continues = pdf > R(0) && cosine > R(0) && dot(direction, hit.geometricNormal) > R(0);
            if (continues) {
                P dAlpha;
                V3<P> f = bsdf(direction, dAlpha);
                R cosOverPdf = cosine / pdf;
                V3<P> weight = f * P(cosOverPdf);
                T = hadamard(T, weight);
                if (record) {
                    vtx.cont = true;
                    vtx.weight = V3<double>(value(weight.x), value(weight.y), value(weight.z));
                    vtx.cosOverPdf = double(cosOverPdf);
                    vtx.contDAlpha = value(dAlpha);
                }
                path.previousBsdfPdf = pdf;
                path.specularBounce = false;
            }
            if (record) record->push_back(vtx);
            if (!continues) break;
        }
        path.ray.o = offsetRayOrigin(hit.position, hit.geometricNormal, direction, hit.t);
        path.ray.d = direction;
        path.depth = depth + 1;
    }
    return acc;
}

// Reverse sweep over the recorded vertices: S_k is the radiance from vertex k
// on per unit throughput at k; gradients need no division.
void diffReverse(const std::vector<DiffVertex>& vs, const V3<double>& delta, const DiffSetup& s, double* grad) {
    V3<double> S;  // S_{k+1}
    const double invPi = double(0.31830988618f);
    double* gTex = grad;
    double* gRough = grad + s.texFloats;
    double* gEmit = grad + s.texFloats + s.roughCount;
    for (size_t i = vs.size(); i-- > 0;) {
        const DiffVertex& v = vs[i];
        V3<double> next = S;
        if (v.kind == 1) {
            double e = v.eslot == kInvalid ? 1.0 : s.params[s.texFloats + s.roughCount + v.eslot];
            if (v.eslot != kInvalid) gEmit[v.eslot] += dot(delta, hadamard(v.T, v.em));
            S = v.em * e;
        } else if (v.kind == 3) {
            S = hadamard(v.weight, next);
        } else {
            V3<double> texGrad;
            double roughGrad = 0.0;
            V3<double> own;
            if (v.nee) {
                V3<double> fb = hadamard(v.neeF, v.neeBase);
                own = fb * v.neeE;
                if (v.neeSlot != kInvalid) gEmit[v.neeSlot] += dot(delta, hadamard(v.T, fb));
                V3<double> db = hadamard(delta, hadamard(v.T, v.neeBase)) * v.neeE;
                texGrad += db * invPi;
                roughGrad += (db.x + db.y + db.z) * v.neeDAlpha;
            }
            if (v.cont) {
                V3<double> dn = hadamard(delta, hadamard(v.T, next));
                texGrad += dn * (v.cosOverPdf * invPi);
                roughGrad += (dn.x + dn.y + dn.z) * v.contDAlpha * v.cosOverPdf;
                own += hadamard(v.weight, next);
            }
            if (v.textured)
                for (int k = 0; k < 4; ++k) {
                    double* t = gTex + 3u * static_cast<size_t>(v.fp.index[k]);
                    t[0] += texGrad.x * v.fp.weight[k];
                    t[1] += texGrad.y * v.fp.weight[k];
                    t[2] += texGrad.z * v.fp.weight[k];
                }
            if (v.rslot != kInvalid) gRough[v.rslot] += roughGrad;
            S = own;
        }
    }
}

template <class F>
void parallelRows(uint32_t height, uint32_t threads, F&& body) {
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    threads = std::min(threads, std::max(1u, height));
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            for (uint32_t y = t; y < height; y += threads) body(t, y);
        });
    for (std::thread& t : pool) t.join();
}

}  // namespace

extern "C" {

int oracle_create(const oracle_scene* scene, oracle_handle** out) {
    oracle_handle* h = new oracle_handle();
    h->d = scene->desc;
    h->tlas = BvhView{scene->tlas_nodes, scene->tlas_node_count};
    h->flat = BvhView{scene->flat_nodes, scene->flat_node_count};
    for (uint32_t m = 0; m < scene->desc.mesh_count; ++m) {
        uint32_t count = scene->blas_node_count ? scene->blas_node_count[m] : 0;
        h->blas.push_back(BvhView{count ? scene->blas_nodes + scene->blas_node_offset[m] : nullptr, count});
    }
    buildTables(*h);
    *out = h;
    return 0;
}

void oracle_destroy(oracle_handle* h) { delete h; }

uint32_t oracle_env_tables(const oracle_handle* h, float* probabilities, float* alias2, uint32_t capacity) {
    uint32_t n = static_cast<uint32_t>(h->envProb.size());
    for (uint32_t i = 0; i < std::min(n, capacity); ++i) {
        if (probabilities) probabilities[i] = h->envProb[i];
        if (alias2) {
            alias2[2 * i] = h->envAlias[i].probability;
            std::memcpy(&alias2[2 * i + 1], &h->envAlias[i].alias, 4);
        }
    }
    return n;
}

// glTF lobes over a surface given as 23 floats: color[3], normal[3],
// coatNormal[3], metallic, roughness, f0[3], f90, coat, coatRoughness,
// reflectionNormal[3], unused[1].
static GSurf<float> gsurfFrom(const float* p) {
    GSurf<float> g;
    g.color = V3<float>(p[0], p[1], p[2]);
    g.normal = V3<float>(p[3], p[4], p[5]);
    g.coatNormal = V3<float>(p[6], p[7], p[8]);
    g.metallic = p[9];
    g.roughness = p[10];
    g.f0 = V3<float>(p[11], p[12], p[13]);
    g.f90 = p[14];
    g.coat = p[15];
    g.coatRoughness = p[16];
    g.reflectionNormal = V3<float>(p[17], p[18], p[19]);
    return g;
}

void oracle_gltf_eval(const float* surface, const float* view, const float* light, float diffusePdf, float* value,
    float* pdf, float* weights) {
    GSurf<float> g = gsurfFrom(surface);
    V3<float> v(view[0], view[1], view[2]), l(light[0], light[1], light[2]);
    V3<float> f = evalGltf(g, v, l, diffusePdf, *pdf);
    value[0] = f.x;
    value[1] = f.y;
    value[2] = f.z;
    lobeWeights(g, v, weights[0], weights[1], weights[2]);
}

void oracle_gltf_sample_reflection(const float* normal, const float* view, float roughness, float u, float v,
    float* out) {
    V3<float> d = reflSample(V3<float>(normal[0], normal[1], normal[2]), V3<float>(view[0], view[1], view[2]),
        roughness, u, v);
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}

void oracle_gltf_reflection_normal(const float* normal, const float* geometric, const float* view, float* out) {
    V3<float> d = reflectionFrame(V3<float>(normal[0], normal[1], normal[2]),
        V3<float>(geometric[0], geometric[1], geometric[2]), V3<float>(view[0], view[1], view[2]));
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}

// Material terms at a surface point (normal +z, tangent +x, bitangent +y):
// out = albedo[3], opacity, emission[3], metallic, roughness, specular f0[3],
// thickness, shading normal[3], and for glTF materials dielectricF0[3], F90,
// coat, coatRoughness, coatNormal[3] (29 floats).
void oracle_material_terms(const oracle_handle* h, uint32_t material, const float* uv, const float* uv1, float* out) {
    TexHit<float> th;
    th.uv = V2<float>{uv[0], uv[1]};
    th.uv1 = V2<float>{uv1[0], uv1[1]};
    th.normal = th.geometricNormal = V3<float>(0, 0, 1);
    th.tangent = th.tangent1 = V3<float>(1, 0, 0);
    th.bitangent = th.bitangent1 = V3<float>(0, 1, 0);
    const wpt_material_ext& e = h->ext[material];
    V3<float> a = matAlbedo(*h, material, th), em = matEmission(*h, material, th);
    V3<float> f0 = hadamard(vec<float>(mat(*h, material).specular), gltfSrgb(sampleTex(*h,
        e.specular_glossiness_texture, th.uv, th.uv1, white<float>())));
    V3<float> n = mapNormal(*h, e.normal_texture, e.normal_scale, th);
    float vals[18] = {a.x, a.y, a.z, matOpacity(*h, material, th.uv, th.uv1), em.x, em.y, em.z,
        matMetallic(*h, material, th), matRoughness(*h, material, th), f0.x, f0.y, f0.z,
        matThickness(*h, material, th), n.x, n.y, n.z, 0, 0};
    for (int i = 0; i < 16; ++i) out[i] = vals[i];
    if (e.gltf_pbr_workflow) {
        GSurf<float> g = gltfAt(*h, material, th, n, V3<float>(0, 0, 1));
        float gv[11] = {g.f0.x, g.f0.y, g.f0.z, g.f90, g.coat, g.coatRoughness, g.coatNormal.x, g.coatNormal.y,
            g.coatNormal.z, 0, 0};
        for (int i = 0; i < 9; ++i) out[16 + i] = gv[i];
    } else {
        for (int i = 16; i < 25; ++i) out[i] = 0.0f;
    }
}

uint32_t oracle_photon_map(const oracle_handle* h, const wpt_render_config* config, float* photons9, uint32_t capacity) {
    PhotonMapT<float> m;
    photonGrid(*config, m);
    binPhotons(m, tracePhotonSet<float>(*h, config->caustic_photons,
        config->caustic_max_depth ? config->caustic_max_depth : 8u, config->rng_strategy, config->caustic_spectral != 0));
    for (uint32_t i = 0; i < std::min(capacity, m.stored); ++i) {
        const OPhoton<float>& q = m.photons[i];
        float row[9] = {q.position.x, q.position.y, q.position.z, q.direction.x, q.direction.y, q.direction.z,
            q.flux.x, q.flux.y, q.flux.z};
        std::memcpy(photons9 + 9ull * i, row, sizeof row);
    }
    return m.stored;
}

void oracle_gather_photons(const wpt_render_config* config, const float* photons9, uint32_t n, const float* point,
    const float* normal, float* out) {
    PhotonMapT<float> m;
    photonGrid(*config, m);
    std::vector<OPhoton<float>> set(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = photons9 + 9 * i;
        set[i] = OPhoton<float>{V3<float>(q[0], q[1], q[2]), V3<float>(q[3], q[4], q[5]), V3<float>(q[6], q[7], q[8])};
    }
    binPhotons(m, set);
    V3<float> e = gatherPhotons(m, V3<float>(point[0], point[1], point[2]), V3<float>(normal[0], normal[1], normal[2]));
    out[0] = e.x;
    out[1] = e.y;
    out[2] = e.z;
}

void oracle_punctual_sample(const wpt_punctual_light* light, const float* point, float* direction, float* incident,
    float* distance) {
    V3<float> in;
    V3<float> d = punctualSample(*light, V3<float>(point[0], point[1], point[2]), in, *distance);
    direction[0] = d.x;
    direction[1] = d.y;
    direction[2] = d.z;
    incident[0] = in.x;
    incident[1] = in.y;
    incident[2] = in.z;
}

float oracle_env_solid_angle(uint32_t row, uint32_t width, uint32_t height) {
    return envSolidAngle(row, width, height);
}

int oracle_env_eval(const oracle_handle* h, const float* dirs, uint32_t n, float* uv, float* pdf, float* rgb,
    uint32_t* texel) {
    if (!h->env) return 1;
    for (uint32_t i = 0; i < n; ++i) {
        V3<float> d(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
        if (uv) envUv(*h, d, uv[2 * i], uv[2 * i + 1]);
        if (pdf) pdf[i] = envPdf(*h, d);
        if (rgb) {
            V3<float> r = envRadiance(*h, d);
            rgb[3 * i] = r.x;
            rgb[3 * i + 1] = r.y;
            rgb[3 * i + 2] = r.z;
        }
        if (texel) texel[i] = envTexel(*h, d);
    }
    return 0;
}

int oracle_env_sample(const oracle_handle* h, const double* candidates, const float* draws3, uint32_t n, float* dirs,
    float* pdf) {
    if (!h->env) return 1;
    for (uint32_t i = 0; i < n; ++i) {
        float p;
        V3<float> d = envSample<float>(*h, candidates[i], draws3[3 * i], draws3[3 * i + 1], draws3[3 * i + 2], p);
        dirs[3 * i] = d.x;
        dirs[3 * i + 1] = d.y;
        dirs[3 * i + 2] = d.z;
        pdf[i] = p;
    }
    return 0;
}

uint32_t oracle_list(const oracle_handle* h, uint32_t which, uint32_t* out, uint32_t capacity) {
    const std::vector<uint32_t>& v = which == 0 ? h->directSpheres : (which == 1 ? h->directInstances : h->leaves);
    if (out) std::memcpy(out, v.data(), sizeof(uint32_t) * std::min<size_t>(capacity, v.size()));
    return static_cast<uint32_t>(v.size());
}

uint32_t oracle_lights(const oracle_handle* h, float* lights16, float* alias2, uint32_t capacity) {
    uint32_t n = static_cast<uint32_t>(h->lights.size());
    for (uint32_t i = 0; i < std::min(n, capacity); ++i) {
        const TriangleLight& L = h->lights[i];
        if (lights16) {
            float* o = lights16 + 16 * i;
            const V3<float> v[4] = {L.v0, L.v1, L.v2, L.normal};
            for (int k = 0; k < 4; ++k) {
                o[3 * k] = v[k].x;
                o[3 * k + 1] = v[k].y;
                o[3 * k + 2] = v[k].z;
            }
            o[12] = L.area;
            std::memcpy(&o[13], &L.material, 4);
            o[14] = L.selectionPdf;
            o[15] = 0.0f;
        }
        if (alias2) {
            alias2[2 * i] = h->alias[i].probability;
            std::memcpy(&alias2[2 * i + 1], &h->alias[i].alias, 4);
        }
    }
    return n;
}

void oracle_trace(const oracle_handle* h, const float* rays, uint32_t n, wpt_hit* out) {
    for (uint32_t i = 0; i < n; ++i) {
        Ray<float> ray{V3<float>(rays[i], rays[n + i], rays[2 * n + i]), V3<float>(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i])};
        Hit<float> hit;
        HitState<float> hs;
        bool did = traceRay(*h, ray, hit, hs);
        wpt_hit r{};
        r.t = hs.closestT;
        r.u = hs.u;
        r.v = hs.v;
        r.kind = did ? hs.kind : WPT_HIT_NONE;
        r.instance = did ? hs.instance : kInvalid;
        r.primitive = did ? hs.prim : kInvalid;
        r.material = did ? hit.material : kInvalid;
        r.light_index = did ? hit.lightIndex : kInvalid;
        if (!did) r.t = 1e30f;
        out[i] = r;
    }
}

void oracle_occluded(const oracle_handle* h, const float* rays, const float* tmax, uint32_t n, uint8_t* out) {
    for (uint32_t i = 0; i < n; ++i) {
        Ray<float> ray{V3<float>(rays[i], rays[n + i], rays[2 * n + i]), V3<float>(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i])};
        out[i] = isOccluded(*h, ray, tmax[i]) ? 1 : 0;
    }
}

void oracle_brute_fp64(const oracle_handle* h, const float* rays, uint32_t n, double* tOut, uint32_t* matOut, float* nOut) {
    const wpt_scene_desc& d = h->d;
    for (uint32_t i = 0; i < n; ++i) {
        V3<double> o(rays[i], rays[n + i], rays[2 * n + i]);
        V3<double> dir(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        double best = 1e30;
        uint32_t bestMat = kInvalid;
        V3<double> bestN;
        for (uint32_t s = 0; s < d.sphere_count; ++s) {
            const wpt_sphere& sp = d.spheres[s];
            V3<double> oc = o - vec<double>(sp.center);
            double a = dot(dir, dir), b = dot(oc, dir), c = dot(oc, oc) - double(sp.radius) * sp.radius;
            double disc = b * b - a * c;
            if (disc < 0) continue;
            double root = std::sqrt(disc);
            double t = (-b - root) / a;
            if (t <= 0.001) t = (-b + root) / a;
            if (t <= 0.001 || t >= best) continue;
            best = t;
            bestMat = sp.material;
            bestN = normalize(o + dir * t - vec<double>(sp.center));
        }
        auto tri = [&](const wpt_triangle& tr, const wpt_transform* xf, uint32_t override) {
            V3<double> lo = o, ld = dir;
            if (xf) {
                lo = inverseXform(*xf, o - vec<double>(xf->translation));
                ld = inverseXform(*xf, dir);
            }
            V3<double> v0 = vec<double>(tr.v0), e1 = vec<double>(tr.v1) - v0, e2 = vec<double>(tr.v2) - v0;
            V3<double> p = cross(ld, e2);
            double det = dot(e1, p);
            if (std::fabs(det) < 1e-14) return;
            V3<double> off = lo - v0;
            double u = dot(off, p) / det;
            V3<double> q = cross(off, e1);
            double v = dot(ld, q) / det;
            double t = dot(e2, q) / det;
            if (u < 0 || v < 0 || u + v > 1 || t <= 0.001 || t >= best) return;
            uint32_t m = override == kInvalid ? tr.material : override;
            if (cutout(*h, m, tr, u, v)) return;
            V3<double> nn = (tr.flags & WPT_TRIANGLE_HAS_NORMALS)
                ? normalize(vec<double>(tr.n0) * (1 - u - v) + vec<double>(tr.n1) * u + vec<double>(tr.n2) * v)
                : normalize(cross(e1, e2));
            if (dot(nn, ld) > 0) nn = -nn;
            best = t;
            bestMat = m;
            bestN = xf ? xformNormal(*xf, nn) : nn;
        };
        for (uint32_t t = 0; t < d.static_triangle_count; ++t) tri(d.static_triangles[t], nullptr, kInvalid);
        for (uint32_t t = 0; t < d.flat_triangle_count; ++t) tri(d.flat_triangles[t], nullptr, kInvalid);
        for (uint32_t k = 0; k < d.instance_count; ++k) {
            const wpt_instance& inst = d.instances[k];
            const wpt_mesh& m = d.meshes[inst.mesh];
            for (uint32_t t = 0; t < m.triangle_count; ++t)
                tri(d.mesh_triangles[m.first_triangle + t], &inst.transform, inst.material_override);
        }
        tOut[i] = best;
        matOut[i] = bestMat;
        nOut[3 * i] = static_cast<float>(bestN.x);
        nOut[3 * i + 1] = static_cast<float>(bestN.y);
        nOut[3 * i + 2] = static_cast<float>(bestN.z);
    }
}

int oracle_render(const oracle_handle* h, const wpt_camera* camera, const wpt_render_config* config, uint32_t x,
    uint32_t y, uint32_t w, uint32_t th, uint32_t s0, uint32_t sc, uint32_t precision, uint32_t threads, float* fb,
    oracle_stats* stats) {
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    threads = std::min(threads, std::max(1u, th));
    std::vector<Counters> counters(threads);
    // The caustic map, as the renderer builds it on creation.
    PhotonMapT<float> map32;
    PhotonMapT<double> map64;
    bool gather = config->caustic_photons && config->caustic_gather;
    if (config->caustic_photons) {
        uint32_t depth = config->caustic_max_depth ? config->caustic_max_depth : 8u;
        if (precision == 64) {
            photonGrid(*config, map64);
            binPhotons(map64, tracePhotonSet<double>(*h, config->caustic_photons, depth, config->rng_strategy,
                config->caustic_spectral != 0));
        } else {
            photonGrid(*config, map32);
            binPhotons(map32, tracePhotonSet<float>(*h, config->caustic_photons, depth, config->rng_strategy,
                config->caustic_spectral != 0));
        }
    }
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; ++t) {
        pool.emplace_back([&, t] {
            if (precision == 64)
                renderRows<double>(*h, *camera, *config, x, y, w, th, s0, sc, t, threads, fb, counters[t],
                    gather ? &map64 : nullptr);
            else
                renderRows<float>(*h, *camera, *config, x, y, w, th, s0, sc, t, threads, fb, counters[t],
                    gather ? &map32 : nullptr);
        });
    }
    for (std::thread& t : pool) t.join();
    if (stats) {
        *stats = oracle_stats{};
        for (const Counters& c : counters) {
            stats->closest_queries += c.closest;
            stats->shadow_queries += c.shadow;
            stats->closest_node_tests += c.closestWork.nodes;
            stats->closest_triangle_tests += c.closestWork.triangles;
            stats->closest_sphere_tests += c.closestWork.spheres;
            stats->closest_instance_entries += c.closestWork.instances;
            stats->shadow_node_tests += c.shadowWork.nodes;
            stats->shadow_triangle_tests += c.shadowWork.triangles;
            stats->shadow_sphere_tests += c.shadowWork.spheres;
            stats->shadow_instance_entries += c.shadowWork.instances;
        }
        stats->paths = static_cast<uint64_t>(w) * th * sc;
        stats->caustic_photons_stored = precision == 64 ? map64.stored : map32.stored;
    }
    return 0;
}

uint32_t oracle_compact(const uint32_t* items, const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t capacity) {
    uint32_t count = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (!flags[i]) continue;
        if (count < capacity) out[count] = items[i];
        ++count;
    }
    return count;
}

uint32_t oracle_reference_rays(uint32_t seed, uint32_t nRandom, float* rays, uint32_t capacity) {
    std::vector<Ray<float>> r = {
        {V3<float>(0, 0, 0), V3<float>(0, 0, 1)}, {V3<float>(1, 0, 0), V3<float>(0, 0, 1)},
        {V3<float>(0, 0, 3), V3<float>(1, 0, 0)}, {V3<float>(0, 0, 0), V3<float>(0, -1, 0)},
        {V3<float>(0, 0, 0), V3<float>(0, 0, -1)}, {V3<float>(-3, 0, 0), V3<float>(0, 0, 1)},
        {V3<float>(3, 0, 0), V3<float>(0, 0, 1)}, {V3<float>(0, 0, 8), V3<float>(0, 0, -1)}};
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> position(-4, 4);
    std::uniform_real_distribution<float> direction(-1, 1);
    for (uint32_t i = 0; i < nRandom; ++i) {
        float ox = position(rng), oy = position(rng), oz = position(rng);
        float dx = direction(rng), dy = direction(rng), dz = direction(rng);
        r.push_back({V3<float>(ox, oy, oz), normalize(V3<float>(dx, dy, dz))});
    }
    uint32_t n = static_cast<uint32_t>(r.size());
    if (rays && capacity >= n) {
        for (uint32_t i = 0; i < n; ++i) {
            rays[i] = r[i].o.x;
            rays[n + i] = r[i].o.y;
            rays[2 * n + i] = r[i].o.z;
            rays[3 * n + i] = r[i].d.x;
            rays[4 * n + i] = r[i].d.y;
            rays[5 * n + i] = r[i].d.z;
        }
    }
    return n;
}

uint64_t oracle_rng_stream(uint32_t pixel, uint32_t sample, uint32_t strategy, uint32_t count, float* out) {
    RngState s = makeRngState(pixelRngKey(pixel, sample, strategy), strategy);
    uint64_t seeded = s.state;
    for (uint32_t i = 0; i < count; ++i) out[i] = randomFloat(s);
    return seeded;
}

void oracle_elementary(uint32_t op, const float* x, const float* y, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        switch (op) {
        case 0: out[i] = fsin(x[i]); break;
        case 1: out[i] = fcos(x[i]); break;
        case 2: out[i] = ftan(x[i]); break;
        case 3: out[i] = fexp(x[i]); break;
        case 4: out[i] = flog(x[i]); break;
        default: out[i] = fpow(x[i], y[i]); break;
        }
    }
}

float oracle_dielectric_reflectance(float cosine, float eta) { return dielectricReflectance<float>(cosine, eta); }

void oracle_offset_ray_origin(const float* position, const float* normal, const float* direction, float hit_distance,
    float* out) {
    V3<float> r = offsetRayOrigin(V3<float>{position[0], position[1], position[2]},
        V3<float>{normal[0], normal[1], normal[2]}, V3<float>{direction[0], direction[1], direction[2]}, hit_distance);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

void oracle_eval_specgloss(const float f0[3], const float diffuse[3], float roughness, const float n[3], const float wo[3],
    const float wi[3], float value[3], float* pdf) {
    SG<float> s{vec<float>(diffuse), vec<float>(f0), roughness};
    V3<float> r = evalSG(s, vec<float>(n), vec<float>(wo), vec<float>(wi), *pdf);
    value[0] = r.x;
    value[1] = r.y;
    value[2] = r.z;
}


uint32_t oracle_diff_layout(const oracle_handle* h, const oracle_diff_config* c, uint32_t* texture_floats,
    uint32_t* roughness_count, uint32_t* emission_count) {
    DiffSetup s = diffSetup(*h, *c, nullptr);
    if (texture_floats) *texture_floats = s.texFloats;
    if (roughness_count) *roughness_count = s.roughCount;
    if (emission_count) *emission_count = s.emitCount;
    return s.texFloats + s.roughCount + s.emitCount;
}

int oracle_diff_render(const oracle_handle* h, const wpt_camera* cam, const oracle_diff_config* c, const float* params,
    uint32_t sample_start, uint32_t precision, uint32_t threads, float* image) {
    DiffSetup s = diffSetup(*h, *c, params);
    parallelRows(c->height, threads, [&](uint32_t, uint32_t y) {
        for (uint32_t x = 0; x < c->width; ++x) {
            float* px = image + 3 * (static_cast<size_t>(y) * c->width + x);
            if (precision == 64) {
                V3<double> sum;
                for (uint32_t k = 0; k < c->samples_per_pixel; ++k)
                    sum += diffPathOracle<double, double>(*h, s, *cam, x, y, sample_start + k, nullptr);
                double inv = 1.0 / c->samples_per_pixel;
                px[0] = static_cast<float>(sum.x * inv);
                px[1] = static_cast<float>(sum.y * inv);
                px[2] = static_cast<float>(sum.z * inv);
            } else {
                V3<float> sum;
                for (uint32_t k = 0; k < c->samples_per_pixel; ++k)
                    sum += diffPathOracle<float, float>(*h, s, *cam, x, y, sample_start + k, nullptr);
                float inv = 1.0f / static_cast<float>(c->samples_per_pixel);
                px[0] = sum.x * inv;
                px[1] = sum.y * inv;
                px[2] = sum.z * inv;
            }
        }
    });
    return 0;
}

int oracle_diff_adjoint64(const oracle_handle* h, const wpt_camera* cam, const oracle_diff_config* c,
    const float* params, uint32_t sample_start, const float* image_grad, uint32_t threads, double* grad) {
    DiffSetup s = diffSetup(*h, *c, params);
    size_t count = static_cast<size_t>(s.texFloats) + s.roughCount + s.emitCount;
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    threads = std::min(threads, std::max(1u, c->height));
    std::vector<std::vector<double>> partial(threads, std::vector<double>(count, 0.0));
    parallelRows(c->height, threads, [&](uint32_t t, uint32_t y) {
        std::vector<DiffVertex> vs;
        for (uint32_t x = 0; x < c->width; ++x) {
            const float* g = image_grad + 3 * (static_cast<size_t>(y) * c->width + x);
            double inv = 1.0 / c->samples_per_pixel;
            V3<double> delta(g[0] * inv, g[1] * inv, g[2] * inv);
            for (uint32_t k = 0; k < c->samples_per_pixel; ++k) {
                vs.clear();
                diffPathOracle<double, double>(*h, s, *cam, x, y, sample_start + k, &vs);
                diffReverse(vs, delta, s, partial[t].data());
            }
        }
    });
    for (size_t i = 0; i < count; ++i) {
        double sum = 0.0;
        for (uint32_t t = 0; t < threads; ++t) sum += partial[t][i];
        grad[i] = sum;
    }
    return 0;
}

int oracle_diff_directional64(const oracle_handle* h, const wpt_camera* cam, const oracle_diff_config* c,
    const float* params, const float* direction, uint32_t sample_start, uint32_t threads, double* dimage) {
    DiffSetup s = diffSetup(*h, *c, params);
    s.direction = direction;
    parallelRows(c->height, threads, [&](uint32_t, uint32_t y) {
        for (uint32_t x = 0; x < c->width; ++x) {
            V3<Dual> sum(Dual(0), Dual(0), Dual(0));
            for (uint32_t k = 0; k < c->samples_per_pixel; ++k)
                sum += diffPathOracle<double, Dual>(*h, s, *cam, x, y, sample_start + k, nullptr);
            double* px = dimage + 3 * (static_cast<size_t>(y) * c->width + x);
            px[0] = sum.x.d / c->samples_per_pixel;
            px[1] = sum.y.d / c->samples_per_pixel;
            px[2] = sum.z.d / c->samples_per_pixel;
        }
    });
    return 0;
}

}  // extern "C"
