// Host scene build: binned-SAH BVH2 builder, AccelerationMode partition,
// TLAS leaves and the triangle-light alias table.
//
// Reference behaviour followed here:
//   partition + TLAS leaves ........ src/scene/scene.cu:175-239
//   BLAS only for meshes with a Bvh-mode instance ... scene.cu:123-173
//   flat-triangle BVH only when there are no meshes .. scene.cu:358-376
//   light registration ............. src/scene/scene_host.cpp:60-106
//   Walker alias table ............. src/scene/scene.cu:82-121
// NexusBVH (the reference's GPU H-PLOC builder) is not available; this
// builder produces the same kind of tree (binary, one primitive per leaf) with
// a host binned-SAH construction.
#include "scene_build.h"
#include "environment.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <thread>

namespace wpt {

namespace {

constexpr int kBins = 32;

struct BuildContext {
    const std::vector<Aabb>* bounds;
    std::vector<float> cx, cy, cz;
    std::vector<uint32_t> idx;
    std::vector<BvhNode>* nodes;
    unsigned parallelDepth;
};

inline float area(const Aabb& b) {
    float dx = b.mx[0] - b.mn[0];
    float dy = b.mx[1] - b.mn[1];
    float dz = b.mx[2] - b.mn[2];
    if (dx < 0.0f || dy < 0.0f || dz < 0.0f) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

inline Aabb emptyBox() {
    Aabb b;
    for (int a = 0; a < 3; ++a) {
        b.mn[a] = INFINITY;
        b.mx[a] = -INFINITY;
    }
    return b;
}

inline void grow(Aabb& b, const Aabb& o) {
    for (int a = 0; a < 3; ++a) {
        b.mn[a] = std::fmin(b.mn[a], o.mn[a]);
        b.mx[a] = std::fmax(b.mx[a], o.mx[a]);
    }
}

struct BuildResult {
    uint32_t pointer;
    Aabb bounds;
    uint32_t depth;
};

const float& centroid(const BuildContext& c, int axis, uint32_t prim) {
    return axis == 0 ? c.cx[prim] : (axis == 1 ? c.cy[prim] : c.cz[prim]);
}

// Picks a split position in [begin, end), reordering idx; returns mid.
uint32_t split(BuildContext& c, uint32_t begin, uint32_t end) {
    uint32_t n = end - begin;
    float cmin[3] = {INFINITY, INFINITY, INFINITY};
    float cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = begin; i < end; ++i) {
        uint32_t p = c.idx[i];
        float v[3] = {c.cx[p], c.cy[p], c.cz[p]};
        for (int a = 0; a < 3; ++a) {
            cmin[a] = std::fmin(cmin[a], v[a]);
            cmax[a] = std::fmax(cmax[a], v[a]);
        }
    }
    const std::vector<Aabb>& bounds = *c.bounds;
    float bestCost = INFINITY;
    int bestAxis = -1;
    int bestBin = -1;
    for (int axis = 0; axis < 3; ++axis) {
        float extent = cmax[axis] - cmin[axis];
        if (!(extent > 0.0f)) continue;
        float scale = static_cast<float>(kBins) * (1.0f - 1.0e-5f) / extent;
        Aabb binBox[kBins];
        uint32_t binCount[kBins];
        for (int b = 0; b < kBins; ++b) {
            binBox[b] = emptyBox();
            binCount[b] = 0;
        }
        for (uint32_t i = begin; i < end; ++i) {
            uint32_t p = c.idx[i];
            int b = static_cast<int>((centroid(c, axis, p) - cmin[axis]) * scale);
            b = std::min(std::max(b, 0), kBins - 1);
            ++binCount[b];
            grow(binBox[b], bounds[p]);
        }
        float rightArea[kBins];
        uint32_t rightCount[kBins];
        Aabb acc = emptyBox();
        uint32_t cnt = 0;
        for (int b = kBins - 1; b > 0; --b) {
            grow(acc, binBox[b]);
            cnt += binCount[b];
            rightArea[b] = area(acc);
            rightCount[b] = cnt;
        }
        acc = emptyBox();
        cnt = 0;
        for (int b = 0; b < kBins - 1; ++b) {
            grow(acc, binBox[b]);
            cnt += binCount[b];
            if (cnt == 0 || rightCount[b + 1] == 0) continue;
            float cost = area(acc) * static_cast<float>(cnt) + rightArea[b + 1] * static_cast<float>(rightCount[b + 1]);
            if (cost < bestCost) {
                bestCost = cost;
                bestAxis = axis;
                bestBin = b;
            }
        }
    }
    if (bestAxis < 0) return begin + n / 2;  // coincident centroids: keep order, split evenly
    float scale = static_cast<float>(kBins) * (1.0f - 1.0e-5f) / (cmax[bestAxis] - cmin[bestAxis]);
    float lo = cmin[bestAxis];
    int axis = bestAxis;
    auto* first = c.idx.data() + begin;
    auto* last = c.idx.data() + end;
    auto* midIt = std::stable_partition(first, last, [&](uint32_t p) {
        int b = static_cast<int>((centroid(c, axis, p) - lo) * scale);
        b = std::min(std::max(b, 0), kBins - 1);
        return b <= bestBin;
    });
    uint32_t mid = static_cast<uint32_t>(midIt - c.idx.data());
    if (mid == begin || mid == end) return begin + n / 2;
    return mid;
}

BuildResult buildRange(BuildContext& c, uint32_t begin, uint32_t end, uint32_t nodeIndex, unsigned level) {
    const std::vector<Aabb>& bounds = *c.bounds;
    uint32_t n = end - begin;
    if (n == 1) return BuildResult{kLeafBit | begin, bounds[c.idx[begin]], 0};
    uint32_t mid = split(c, begin, end);
    uint32_t leftCount = mid - begin;
    BuildResult left, right;
    if (level < c.parallelDepth && n > 32768) {
        auto future = std::async(std::launch::async, [&] { return buildRange(c, begin, mid, nodeIndex + 1, level + 1); });
        right = buildRange(c, mid, end, nodeIndex + leftCount, level + 1);
        left = future.get();
    } else {
        left = buildRange(c, begin, mid, nodeIndex + 1, level + 1);
        right = buildRange(c, mid, end, nodeIndex + leftCount, level + 1);
    }
    BvhNode& node = (*c.nodes)[nodeIndex];
    node.lminx = left.bounds.mn[0];
    node.lminy = left.bounds.mn[1];
    node.lminz = left.bounds.mn[2];
    node.lmaxx = left.bounds.mx[0];
    node.lmaxy = left.bounds.mx[1];
    node.lmaxz = left.bounds.mx[2];
    node.rminx = right.bounds.mn[0];
    node.rminy = right.bounds.mn[1];
    node.rminz = right.bounds.mn[2];
    node.rmaxx = right.bounds.mx[0];
    node.rmaxy = right.bounds.mx[1];
    node.rmaxz = right.bounds.mx[2];
    node.left = left.pointer;
    node.right = right.pointer;
    node.pad0 = node.pad1 = 0;
    Aabb u = left.bounds;
    grow(u, right.bounds);
    return BuildResult{nodeIndex, u, 1 + std::max(left.depth, right.depth)};
}

Aabb nodeChildBounds(const BvhNode& n, bool right) {
    Aabb b;
    if (!right) {
        b.mn[0] = n.lminx; b.mn[1] = n.lminy; b.mn[2] = n.lminz;
        b.mx[0] = n.lmaxx; b.mx[1] = n.lmaxy; b.mx[2] = n.lmaxz;
    } else {
        b.mn[0] = n.rminx; b.mn[1] = n.rminy; b.mn[2] = n.rminz;
        b.mx[0] = n.rmaxx; b.mx[1] = n.rmaxy; b.mx[2] = n.rmaxz;
    }
    return b;
}

Vec3 toVec(const float v[3]) { return Vec3(v[0], v[1], v[2]); }

constexpr uint32_t kTriExtraFlags = WPT_TRIANGLE_HAS_TEXCOORDS1 | WPT_TRIANGLE_HAS_VERTEX_TANGENTS;
static_assert(WPT_TRIANGLE_HAS_NORMALS == kTriHasNormals && WPT_TRIANGLE_HAS_TEXCOORDS == kTriHasTexcoords &&
    WPT_TRIANGLE_HAS_TEXCOORDS1 == kTriHasTexcoords1 && WPT_TRIANGLE_HAS_VERTEX_TANGENTS == kTriHasVertexTangents,
    "triangle flags match the C ABI");

bool nonzero3(const float v[3]) { return v[0] != 0.0f || v[1] != 0.0f || v[2] != 0.0f; }

// Shading record of a triangle; its second texture-coordinate set and
// tangent data, when present, go to `extra`.
TriShade toShade(const wpt_triangle& t, std::vector<TriExtra>& extra) {
    TriShade s{};
    s.v0 = toVec(t.v0);
    s.v1 = toVec(t.v1);
    s.v2 = toVec(t.v2);
    s.n0 = toVec(t.n0);
    s.n1 = toVec(t.n1);
    s.n2 = toVec(t.n2);
    s.uv0 = Vec2{t.uv0[0], t.uv0[1]};
    s.uv1 = Vec2{t.uv1[0], t.uv1[1]};
    s.uv2 = Vec2{t.uv2[0], t.uv2[1]};
    s.tangent = toVec(t.tangent);
    s.bitangent = toVec(t.bitangent);
    s.material = t.material;
    s.lightIndex = kInvalid;
    s.flags = t.flags & (kTriHasNormals | kTriHasTexcoords | kTriHasTexcoords1 | kTriHasVertexTangents);
    s.extra = kInvalid;
    if ((t.flags & kTriExtraFlags) || nonzero3(t.tangent1) || nonzero3(t.bitangent1)) {
        TriExtra e{};
        e.uv1_0 = Vec2{t.uv1_0[0], t.uv1_0[1]};
        e.uv1_1 = Vec2{t.uv1_1[0], t.uv1_1[1]};
        e.uv1_2 = Vec2{t.uv1_2[0], t.uv1_2[1]};
        e.vertexTangent0 = toVec(t.vertex_tangent0);
        e.vertexTangent1 = toVec(t.vertex_tangent1);
        e.vertexTangent2 = toVec(t.vertex_tangent2);
        e.sign0 = t.vertex_tangent_sign[0];
        e.sign1 = t.vertex_tangent_sign[1];
        e.sign2 = t.vertex_tangent_sign[2];
        e.tangent1 = toVec(t.tangent1);
        e.bitangent1 = toVec(t.bitangent1);
        s.extra = static_cast<uint32_t>(extra.size());
        extra.push_back(e);
    }
    return s;
}

float4 packVertex(const Vec3& v, uint32_t word) {
    float4 r;
    r.x = v.x;
    r.y = v.y;
    r.z = v.z;
    std::memcpy(&r.w, &word, sizeof word);
    return r;
}

// Positions in BVH slot order: (v0 | original index), (v1 | material), (v2 | flags).
std::vector<float4> slotGeometry(const std::vector<TriShade>& tris, const HostBvh& bvh) {
    std::vector<float4> g(static_cast<size_t>(bvh.primCount()) * 3);
    for (uint32_t slot = 0; slot < bvh.primCount(); ++slot) {
        uint32_t p = bvh.order[slot];
        const TriShade& t = tris[p];
        g[3 * slot + 0] = packVertex(t.v0, p);
        g[3 * slot + 1] = packVertex(t.v1, t.material);
        g[3 * slot + 2] = packVertex(t.v2, t.flags);
    }
    return g;
}

Material toMaterial(const wpt_material& m) {
    Material r{};
    r.type = m.type;
    r.albedo = toVec(m.albedo);
    r.emission = toVec(m.emission);
    r.roughness = m.roughness;
    r.ior = m.ior;
    r.metallic = m.metallic;
    r.opacity = m.opacity;
    r.alphaCutoff = m.alpha_cutoff;
    r.alphaMasked = m.alpha_masked ? 1u : 0u;
    r.attenuationColor = toVec(m.attenuation_color);
    r.attenuationDistance = m.attenuation_distance;
    r.volumeDensity = m.volume_density;
    r.specular = toVec(m.specular);
    r.glossiness = m.glossiness;
    r.ext = kInvalid;
    return r;
}

InstanceTransform toTransform(const wpt_transform& t) {
    InstanceTransform r;
    r.localToWorldX = toVec(t.local_to_world[0]);
    r.localToWorldY = toVec(t.local_to_world[1]);
    r.localToWorldZ = toVec(t.local_to_world[2]);
    r.worldToLocalX = toVec(t.world_to_local[0]);
    r.worldToLocalY = toVec(t.world_to_local[1]);
    r.worldToLocalZ = toVec(t.world_to_local[2]);
    r.translation = toVec(t.translation);
    return r;
}

// Luminance power of an emissive triangle, reference scene_host.cpp:67-73.
void registerLight(HostSceneData& out, const Vec3& v0, const Vec3& v1, const Vec3& v2, uint32_t material) {
    Vec3 n = cross(v1 - v0, v2 - v0);
    Vec3 normal = normalize(n);
    float a = 0.5f * length(cross(v1 - v0, v2 - v0));
    const Material& m = out.materials[material];
    float power = a * (0.2126f * m.emission.x + 0.7152f * m.emission.y + 0.0722f * m.emission.z);
    TriangleLight light{};
    light.v0 = v0;
    light.v1 = v1;
    light.v2 = v2;
    light.normal = normal;
    light.area = a;
    light.material = material;
    light.selectionPdf = 0.0f;
    out.lights.push_back(light);
    out.lightWeights.push_back(power);
}

// Walker alias construction with LIFO small/large worklists, reference
// src/scene/scene.cu:86-114 (the pop order determines the exact table).
void buildAlias(HostSceneData& out) {
    size_t n = out.lights.size();
    out.lightAlias.assign(n, LightAlias{0.0f, 0});
    if (n == 0) return;
    float total = 0.0f;
    for (float w : out.lightWeights) total += w;
    std::vector<float> scaled(n);
    std::vector<uint32_t> small, large;
    for (uint32_t i = 0; i < n; ++i) {
        out.lights[i].selectionPdf = out.lightWeights[i] / total;
        scaled[i] = out.lights[i].selectionPdf * static_cast<float>(n);
        (scaled[i] < 1.0f ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        uint32_t low = small.back();
        small.pop_back();
        uint32_t high = large.back();
        large.pop_back();
        out.lightAlias[low] = LightAlias{scaled[low], high};
        scaled[high] = scaled[high] + scaled[low] - 1.0f;
        (scaled[high] < 1.0f ? small : large).push_back(high);
    }
    for (uint32_t i : small) out.lightAlias[i] = LightAlias{1.0f, i};
    for (uint32_t i : large) out.lightAlias[i] = LightAlias{1.0f, i};
}

}  // namespace

Aabb triangleBounds(const Vec3& a, const Vec3& b, const Vec3& c) {
    Aabb r;
    r.mn[0] = std::fmin(a.x, std::fmin(b.x, c.x));
    r.mn[1] = std::fmin(a.y, std::fmin(b.y, c.y));
    r.mn[2] = std::fmin(a.z, std::fmin(b.z, c.z));
    r.mx[0] = std::fmax(a.x, std::fmax(b.x, c.x));
    r.mx[1] = std::fmax(a.y, std::fmax(b.y, c.y));
    r.mx[2] = std::fmax(a.z, std::fmax(b.z, c.z));
    return r;
}

uint32_t stackDepth(const HostSceneData& s) {
    return std::max(s.tlas.depth + s.maxBlasDepth + 2, s.flat.depth + 1);
}

HostBvh buildBvh2(const std::vector<Aabb>& primBounds, unsigned threads) {
    HostBvh bvh;
    bvh.primBounds = primBounds;
    uint32_t n = static_cast<uint32_t>(primBounds.size());
    if (n == 0) return bvh;
    BuildContext c;
    c.bounds = &bvh.primBounds;
    c.cx.resize(n);
    c.cy.resize(n);
    c.cz.resize(n);
    c.idx.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const Aabb& b = primBounds[i];
        c.cx[i] = (b.mn[0] + b.mx[0]) * 0.5f;
        c.cy[i] = (b.mn[1] + b.mx[1]) * 0.5f;
        c.cz[i] = (b.mn[2] + b.mx[2]) * 0.5f;
        c.idx[i] = i;
    }
    bvh.nodes.resize(n - 1);
    c.nodes = &bvh.nodes;
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    unsigned depth = 0;
    while ((1u << depth) < threads && depth < 6) ++depth;
    c.parallelDepth = threads > 1 ? depth + 1 : 0;
    BuildResult r = buildRange(c, 0, n, 0, 0);
    bvh.root = r.pointer;
    bvh.rootBounds = r.bounds;
    bvh.depth = r.depth;
    bvh.order = std::move(c.idx);
    return bvh;
}

std::vector<wpt_bvh2_node> exportBvh2(const HostBvh& bvh) {
    uint32_t n = bvh.primCount();
    std::vector<wpt_bvh2_node> out;
    if (n == 0) return out;
    out.resize(2 * static_cast<size_t>(n) - 1);
    auto nxbIndex = [n](uint32_t pointer) -> uint32_t {
        if (pointer & kLeafBit) return pointer & ~kLeafBit;
        return n + (n - 2 - pointer);
    };
    for (uint32_t slot = 0; slot < n; ++slot) {
        const Aabb& b = bvh.primBounds[bvh.order[slot]];
        wpt_bvh2_node& node = out[slot];
        std::memcpy(node.bmin, b.mn, sizeof node.bmin);
        std::memcpy(node.bmax, b.mx, sizeof node.bmax);
        node.left = 0xffffffffu;
        node.right = bvh.order[slot];
    }
    for (uint32_t i = 0; i + 1 < n; ++i) {
        const BvhNode& src = bvh.nodes[i];
        Aabb l = nodeChildBounds(src, false);
        Aabb r = nodeChildBounds(src, true);
        Aabb u = l;
        grow(u, r);
        wpt_bvh2_node& node = out[nxbIndex(i)];
        std::memcpy(node.bmin, u.mn, sizeof node.bmin);
        std::memcpy(node.bmax, u.mx, sizeof node.bmax);
        node.left = nxbIndex(src.left);
        node.right = nxbIndex(src.right);
    }
    return out;
}

// Environment importance distribution, reference buildEnvironmentDistribution
// (src/scene/environment.cpp:10-64): luminance averaged over the bilinear
// tent of each texel cell (separable [1/8, 3/4, 1/8], wrap in u, clamp in v)
// times the texel's solid angle; 99% of it mixed with 1% uniform solid angle
// so black texels keep support; normalized in double; Walker alias with
// LIFO small/large worklists.
std::string buildEnvironment(const wpt_scene_desc& desc, HostSceneData& out) {
    if (!desc.environment_rgb) return "";
    uint64_t count = static_cast<uint64_t>(desc.environment_width) * desc.environment_height;
    if (count == 0 || count > std::numeric_limits<uint32_t>::max()) return "environment has invalid dimensions";
    if (!std::isfinite(desc.environment_intensity) || desc.environment_intensity < 0.0f ||
        !std::isfinite(desc.environment_rotation_degrees))
        return "environment intensity or rotation is invalid";
    const uint32_t w = desc.environment_width, h = desc.environment_height;
    out.envTexels.resize(count);
    for (uint64_t i = 0; i < count; ++i) {
        const float* p = desc.environment_rgb + 3 * i;
        if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2]) || p[0] < 0 || p[1] < 0 || p[2] < 0)
            return "environment contains negative or non-finite radiance";
        out.envTexels[i] = float4{p[0], p[1], p[2], 1.0f};
    }
    std::vector<double> weights(count);
    double total = 0.0;
    const double filter[3] = {0.125, 0.75, 0.125};
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t x = i % w, y = i / w;
        uint32_t columns[3] = {x ? x - 1 : w - 1, x, x + 1 < w ? x + 1 : 0};
        uint32_t rows[3] = {y ? y - 1 : 0, y, y + 1 < h ? y + 1 : h - 1};
        double average = 0.0;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                const float4& p = out.envTexels[static_cast<uint64_t>(rows[r]) * w + columns[c]];
                average += filter[r] * filter[c] * (0.2126 * double(p.x) + 0.7152 * double(p.y) + 0.0722 * double(p.z));
            }
        weights[i] = average * environmentSolidAngle(y, w, h);
        total += weights[i];
    }
    const double fourPi = 12.5663706143591729538;
    out.envProbabilities.resize(count);
    out.envAlias.resize(count);
    double sum = 0.0;
    for (uint32_t i = 0; i < count; ++i) {
        double uniform = environmentSolidAngle(i / w, w, h) / fourPi;
        out.envProbabilities[i] = static_cast<float>(total > 0 ? 0.99 * weights[i] / total + 0.01 * uniform : uniform);
        sum += out.envProbabilities[i];
    }
    std::vector<uint32_t> small, large;
    for (uint32_t i = 0; i < count; ++i) {
        out.envProbabilities[i] /= sum;
        weights[i] = double(out.envProbabilities[i]) * count;
        (weights[i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        uint32_t low = small.back();
        small.pop_back();
        uint32_t high = large.back();
        large.pop_back();
        out.envAlias[low] = EnvAlias{static_cast<float>(weights[low]), high};
        weights[high] += weights[low] - 1.0;
        (weights[high] < 1.0 ? small : large).push_back(high);
    }
    for (uint32_t i : small) out.envAlias[i] = EnvAlias{1.0f, i};
    for (uint32_t i : large) out.envAlias[i] = EnvAlias{1.0f, i};
    out.envWidth = w;
    out.envHeight = h;
    out.envIntensity = desc.environment_intensity;
    out.envRotation = std::fmod(desc.environment_rotation_degrees, 360.0f) * (3.14159265359f / 180.0f);
    return "";
}

// Punctual lights, directional ones first in their given order, then the
// positional ones (the stable partition of scene.cu:349-355): NEE visits
// every directional light and one positional light chosen uniformly.
// RGBA8 textures, packed one 32-bit word per texel.
std::string buildTextures(const wpt_scene_desc& desc, HostSceneData& out) {
    if (desc.texture_count && !desc.textures) return "texture array is null";
    for (uint32_t i = 0; i < desc.texture_count; ++i) {
        const wpt_texture& t = desc.textures[i];
        uint64_t texels = static_cast<uint64_t>(t.width) * t.height;
        if (!t.width || !t.height || !t.rgba || texels > (1ull << 31)) return "texture has invalid dimensions or data";
        std::vector<uint32_t> packed(texels);
        for (uint64_t k = 0; k < texels; ++k) {
            const uint8_t* p = t.rgba + 4 * k;
            packed[k] = static_cast<uint32_t>(p[0]) | (static_cast<uint32_t>(p[1]) << 8) |
                (static_cast<uint32_t>(p[2]) << 16) | (static_cast<uint32_t>(p[3]) << 24);
        }
        out.textureTexels.push_back(std::move(packed));
        out.textureWidth.push_back(t.width);
        out.textureHeight.push_back(t.height);
    }
    return "";
}

std::vector<const wpt_texture_binding*> bindings(const wpt_material_ext& e) {
    return {&e.albedo_texture, &e.metallic_roughness_texture, &e.emissive_texture, &e.normal_texture,
        &e.thickness_texture, &e.specular_glossiness_texture, &e.specular_weight_texture, &e.specular_color_texture,
        &e.clearcoat_texture, &e.clearcoat_roughness_texture, &e.clearcoat_normal_texture};
}

// A material needs its extension record when a texture is bound or a glTF
// workflow changes its lookups; otherwise its constants are the exact result.
bool needsExt(const wpt_material_ext& e) {
    if (e.specular_glossiness_workflow || e.gltf_pbr_workflow) return true;
    for (const wpt_texture_binding* b : bindings(e))
        if (b->texture != WPT_INVALID_INDEX) return true;
    return false;
}

TexBinding toBinding(const wpt_texture_binding& b) {
    return TexBinding{b.texture, b.texcoord, Vec2{b.offset[0], b.offset[1]}, Vec2{b.scale[0], b.scale[1]}, b.rotation,
        b.has_transform ? 1u : 0u};
}

MaterialExt toMaterialExt(const wpt_material_ext& e) {
    MaterialExt r{};
    r.albedo = toBinding(e.albedo_texture);
    r.metallicRoughness = toBinding(e.metallic_roughness_texture);
    r.emissive = toBinding(e.emissive_texture);
    r.normal = toBinding(e.normal_texture);
    r.thickness = toBinding(e.thickness_texture);
    r.specularGlossiness = toBinding(e.specular_glossiness_texture);
    r.specularWeight = toBinding(e.specular_weight_texture);
    r.specularColor = toBinding(e.specular_color_texture);
    r.clearcoat = toBinding(e.clearcoat_texture);
    r.clearcoatRoughness = toBinding(e.clearcoat_roughness_texture);
    r.clearcoatNormal = toBinding(e.clearcoat_normal_texture);
    r.normalScale = e.normal_scale;
    r.specularGlossinessWorkflow = e.specular_glossiness_workflow ? 1u : 0u;
    r.gltfPbrWorkflow = e.gltf_pbr_workflow ? 1u : 0u;
    r.specularWeightFactor = e.specular_weight;
    r.specularColorFactor = toVec(e.specular_color);
    r.clearcoatFactor = e.clearcoat;
    r.clearcoatRoughnessFactor = e.clearcoat_roughness;
    r.clearcoatNormalScale = e.clearcoat_normal_scale;
    r.dispersion = e.dispersion;
    return r;
}

static_assert(WPT_LIGHT_POINT == kLightPoint && WPT_LIGHT_SPOT == kLightSpot &&
    WPT_LIGHT_DIRECTIONAL == kLightDirectional, "light types match the C ABI");

std::string buildPunctualLights(const wpt_scene_desc& desc, HostSceneData& out) {
    if (desc.punctual_light_count && !desc.punctual_lights) return "punctual light array is null";
    auto finite3 = [](const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); };
    for (int pass = 0; pass < 2; ++pass) {
        for (uint32_t i = 0; i < desc.punctual_light_count; ++i) {
            const wpt_punctual_light& l = desc.punctual_lights[i];
            if (l.type > WPT_LIGHT_DIRECTIONAL) return "punctual light has an unknown type";
            if (!finite3(l.position) || !finite3(l.direction) || !finite3(l.intensity) || !std::isfinite(l.range) ||
                l.range < 0.0f || !std::isfinite(l.inner_cone_cos) || !std::isfinite(l.outer_cone_cos))
                return "punctual light has a non-finite or negative parameter";
            if ((l.type == WPT_LIGHT_DIRECTIONAL) != (pass == 0)) continue;
            out.punctual.push_back(PunctualLight{l.type, Vec3(l.position[0], l.position[1], l.position[2]),
                Vec3(l.direction[0], l.direction[1], l.direction[2]), Vec3(l.intensity[0], l.intensity[1], l.intensity[2]),
                l.range, l.inner_cone_cos, l.outer_cone_cos});
        }
        if (pass == 0) out.directionalCount = static_cast<uint32_t>(out.punctual.size());
    }
    return "";
}

std::string buildHostScene(const wpt_scene_desc& desc, HostSceneData& out, unsigned threads) {
    if ((desc.sphere_count && !desc.spheres) || (desc.static_triangle_count && !desc.static_triangles) ||
        (desc.flat_triangle_count && !desc.flat_triangles) || (desc.mesh_triangle_count && !desc.mesh_triangles) ||
        (desc.mesh_count && !desc.meshes) || (desc.instance_count && !desc.instances) ||
        (desc.material_count && !desc.materials))
        return "scene description has a null array with a nonzero count";
    uint32_t materialCount = desc.material_count;
    auto badMaterial = [&](uint32_t m) { return m >= materialCount; };

    out = HostSceneData{};
    out.blackBackground = desc.black_background ? 1u : 0u;
    out.ignoreDirectLightOcclusion = desc.ignore_direct_light_occlusion ? 1u : 0u;
    std::string envError = buildEnvironment(desc, out);
    if (!envError.empty()) return envError;
    std::string lightError = buildPunctualLights(desc, out);
    if (!lightError.empty()) return lightError;
    std::string textureError = buildTextures(desc, out);
    if (!textureError.empty()) return textureError;
    for (uint32_t i = 0; i < materialCount; ++i) {
        const wpt_material& m = desc.materials[i];
        if (m.type > WPT_MATERIAL_SPECULAR_GLOSSINESS) return "material has an unknown type";
        Material r = toMaterial(m);
        // Dispersion (read by photon tracing only) lives on the material
        // itself, so that an otherwise plain dielectric keeps the lean path.
        if (desc.material_ext) {
            r.dispersion = desc.material_ext[i].dispersion;
            if (!(r.dispersion >= 0.0f) || !std::isfinite(r.dispersion)) return "material dispersion is invalid";
        }
        bool albedoTexture = false;
        if (desc.material_ext && needsExt(desc.material_ext[i])) {
            const wpt_material_ext& e = desc.material_ext[i];
            for (const wpt_texture_binding* b : bindings(e))
                if (b->texcoord > 1) return "texture binding has an unknown coordinate set";
            r.ext = static_cast<uint32_t>(out.materialExt.size());
            out.materialExt.push_back(toMaterialExt(e));
            albedoTexture = e.albedo_texture.texture < desc.texture_count;
        }
        out.materials.push_back(r);
        // kCut* policy of intersectTriangleHit: textured masks test per hit.
        out.materialCutout.push_back(!m.alpha_masked ? kCutNever
            : albedoTexture                        ? kCutTexture
            : m.opacity < m.alpha_cutoff           ? kCutAlways
                                                   : kCutNever);
    }
    out.useNormalMaps = desc.use_normal_maps ? 1u : 0u;
    out.normalMapMinimumCosine = desc.normal_map_minimum_cosine;

    for (uint32_t i = 0; i < desc.sphere_count; ++i) {
        const wpt_sphere& s = desc.spheres[i];
        if (badMaterial(s.material)) return "sphere material index out of range";
        if (s.acceleration > WPT_ACCEL_DIRECT) return "sphere has an unknown acceleration mode";
        float4 v;
        v.x = s.center[0];
        v.y = s.center[1];
        v.z = s.center[2];
        v.w = s.radius;
        out.spheres.push_back(v);
        out.sphereMaterial.push_back(s.material);
    }

    for (uint32_t i = 0; i < desc.static_triangle_count; ++i) {
        if (badMaterial(desc.static_triangles[i].material)) return "static triangle material index out of range";
        out.staticTriangles.push_back(toShade(desc.static_triangles[i], out.triExtra));
    }
    for (uint32_t i = 0; i < desc.flat_triangle_count; ++i) {
        if (badMaterial(desc.flat_triangles[i].material)) return "flat triangle material index out of range";
        out.flatTriangles.push_back(toShade(desc.flat_triangles[i], out.triExtra));
    }

    // Meshes: triangles and per-mesh light indices (scene_host.cpp:77-86).
    out.meshes.resize(desc.mesh_count);
    for (uint32_t m = 0; m < desc.mesh_count; ++m) {
        const wpt_mesh& mesh = desc.meshes[m];
        if (static_cast<uint64_t>(mesh.first_triangle) + mesh.triangle_count > desc.mesh_triangle_count)
            return "mesh triangle range out of bounds";
        uint32_t meshLights = 0;
        for (uint32_t t = 0; t < mesh.triangle_count; ++t) {
            const wpt_triangle& src = desc.mesh_triangles[mesh.first_triangle + t];
            if (badMaterial(src.material)) return "mesh triangle material index out of range";
            TriShade s = toShade(src, out.triExtra);
            s.lightIndex = out.materials[s.material].type == kEmissive ? meshLights++ : kInvalid;
            out.meshes[m].triangles.push_back(s);
        }
    }
    std::vector<bool> needsBvh(desc.mesh_count, false);
    for (uint32_t i = 0; i < desc.instance_count; ++i) {
        const wpt_instance& inst = desc.instances[i];
        if (inst.mesh >= desc.mesh_count) return "instance mesh index out of range";
        if (inst.material_override != WPT_INVALID_INDEX && badMaterial(inst.material_override))
            return "instance material override out of range";
        if (inst.acceleration > WPT_ACCEL_DIRECT) return "instance has an unknown acceleration mode";
        if (inst.acceleration == WPT_ACCEL_BVH) needsBvh[inst.mesh] = true;
    }
    if (desc.bvh_builder > WPT_BVH_BUILDER_DEVICE_PLOC) return "unknown BVH builder";
    if (desc.bvh_search_radius > 64) return "BVH search radius above 64";
    out.bvhBuilder = desc.bvh_builder;
    out.bvhSearchRadius = desc.bvh_search_radius ? desc.bvh_search_radius : 16u;
    uint32_t maxBlasDepth = 0;
    auto blasStart = std::chrono::steady_clock::now();
    for (uint32_t m = 0; m < desc.mesh_count; ++m) {
        HostMesh& mesh = out.meshes[m];
        if (!needsBvh[m] || mesh.triangles.empty()) continue;
        if (desc.bvh_builder == WPT_BVH_BUILDER_DEVICE_PLOC) {
            // Built, and counted in the depth, at upload.  The root box (the
            // union of the triangles' boxes, as the device tree's root)
            // places the instances in the TLAS.
            mesh.buildOnDevice = true;
            Aabb root = emptyBox();
            for (const TriShade& t : mesh.triangles) grow(root, triangleBounds(t.v0, t.v1, t.v2));
            mesh.bvh.rootBounds = root;
            continue;
        }
        std::vector<Aabb> bounds;
        bounds.reserve(mesh.triangles.size());
        for (const TriShade& t : mesh.triangles) bounds.push_back(triangleBounds(t.v0, t.v1, t.v2));
        mesh.bvh = buildBvh2(bounds, threads);
        mesh.geometry = slotGeometry(mesh.triangles, mesh.bvh);
        maxBlasDepth = std::max(maxBlasDepth, mesh.bvh.depth);
    }
    out.blasBuildMilliseconds =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - blasStart).count();
    out.maxBlasDepth = maxBlasDepth;

    // Instances and cumulative light offsets (scene.cu:156-166).
    uint32_t lightOffset = 0;
    for (uint32_t i = 0; i < desc.instance_count; ++i) {
        const wpt_instance& src = desc.instances[i];
        Instance inst{};
        inst.transform = toTransform(src.transform);
        inst.mesh = src.mesh;
        inst.lightOffset = lightOffset;
        inst.materialOverride = src.material_override;
        out.instances.push_back(inst);
        for (const TriShade& t : out.meshes[src.mesh].triangles)
            if (t.lightIndex != kInvalid) ++lightOffset;
    }

    // Lights: mesh instances first, then static triangles (scene_host.cpp:345-346).
    for (uint32_t i = 0; i < desc.instance_count; ++i) {
        const Instance& inst = out.instances[i];
        for (const TriShade& t : out.meshes[inst.mesh].triangles) {
            if (t.lightIndex == kInvalid) continue;
            Vec3 v0 = transformVector(inst.transform, t.v0) + inst.transform.translation;
            Vec3 v1 = transformVector(inst.transform, t.v1) + inst.transform.translation;
            Vec3 v2 = transformVector(inst.transform, t.v2) + inst.transform.translation;
            registerLight(out, v0, v1, v2, t.material);
        }
    }
    for (TriShade& t : out.staticTriangles) {
        if (out.materials[t.material].type != kEmissive) {
            t.lightIndex = kInvalid;
            continue;
        }
        t.lightIndex = static_cast<uint32_t>(out.lights.size());
        registerLight(out, t.v0, t.v1, t.v2, t.material);
    }
    buildAlias(out);
    for (const HostMesh& mesh : out.meshes)
        for (const TriShade& t : mesh.triangles) out.meshLightIndex.push_back(t.lightIndex);

    // TLAS over typed leaves: Bvh-mode spheres, then Bvh-mode nonempty instances.
    std::vector<Aabb> leafBounds;
    for (uint32_t i = 0; i < desc.sphere_count; ++i) {
        const wpt_sphere& s = desc.spheres[i];
        if (s.acceleration == WPT_ACCEL_DIRECT) {
            out.directSpheres.push_back(i);
            continue;
        }
        float r = std::fabs(s.radius);
        Aabb b;
        for (int a = 0; a < 3; ++a) {
            b.mn[a] = std::nextafter(s.center[a] - r, -INFINITY);
            b.mx[a] = std::nextafter(s.center[a] + r, INFINITY);
        }
        leafBounds.push_back(b);
        out.tlasLeavesOriginal.push_back(kTlasSphereBit | i);
    }
    for (uint32_t i = 0; i < desc.instance_count; ++i) {
        const wpt_instance& src = desc.instances[i];
        if (src.acceleration == WPT_ACCEL_DIRECT) {
            out.directInstances.push_back(i);
            continue;
        }
        const HostMesh& mesh = out.meshes[src.mesh];
        if (mesh.triangles.empty()) continue;
        const Aabb& local = mesh.bvh.rootBounds;
        const InstanceTransform& tr = out.instances[i].transform;
        Aabb world = emptyBox();
        for (uint32_t corner = 0; corner < 8; ++corner) {
            Vec3 p((corner & 1) ? local.mx[0] : local.mn[0], (corner & 2) ? local.mx[1] : local.mn[1],
                (corner & 4) ? local.mx[2] : local.mn[2]);
            Vec3 w = transformVector(tr, p) + tr.translation;
            world.mn[0] = std::fmin(world.mn[0], w.x);
            world.mn[1] = std::fmin(world.mn[1], w.y);
            world.mn[2] = std::fmin(world.mn[2], w.z);
            world.mx[0] = std::fmax(world.mx[0], w.x);
            world.mx[1] = std::fmax(world.mx[1], w.y);
            world.mx[2] = std::fmax(world.mx[2], w.z);
        }
        leafBounds.push_back(world);
        out.tlasLeavesOriginal.push_back(i);
    }
    out.tlas = buildBvh2(leafBounds, threads);
    for (uint32_t slot = 0; slot < out.tlas.primCount(); ++slot)
        out.tlasLeavesSlot.push_back(out.tlasLeavesOriginal[out.tlas.order[slot]]);

    // Flat triangles get a BVH only when the scene has no meshes.
    if (desc.mesh_count == 0 && !out.flatTriangles.empty()) {
        std::vector<Aabb> bounds;
        bounds.reserve(out.flatTriangles.size());
        for (const TriShade& t : out.flatTriangles) bounds.push_back(triangleBounds(t.v0, t.v1, t.v2));
        out.flat = buildBvh2(bounds, threads);
        out.flatGeometry = slotGeometry(out.flatTriangles, out.flat);
    }
    out.maxStackDepth = stackDepth(out);
    return std::string();
}

}  // namespace wpt
