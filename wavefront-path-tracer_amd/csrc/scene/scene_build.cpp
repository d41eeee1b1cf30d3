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
#include <functional>
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
    bool median = false;  // object median on the longest centroid axis instead of binned SAH
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
    if (c.median) {
        // Median BVH (the reference's host BVH, bvh.cpp:49-109, leaves of
        // one primitive here): the centroid bounds' longest axis, split at
        // the median centroid, ties kept in primitive order.
        float ex = cmax[0] - cmin[0], ey = cmax[1] - cmin[1], ez = cmax[2] - cmin[2];
        int axis = ex >= ey && ex >= ez ? 0 : (ey >= ez ? 1 : 2);
        uint32_t mid = begin + n / 2;
        std::nth_element(c.idx.begin() + begin, c.idx.begin() + mid, c.idx.begin() + end, [&](uint32_t a, uint32_t b) {
            float ca = centroid(c, axis, a), cb = centroid(c, axis, b);
            return ca < cb || (ca == cb && a < b);
        });
        return mid;
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
    uint32_t flat = s.flatWide.empty() ? s.flat.depth + 1 : s.flatWideStack;
    return std::max(s.tlas.depth + s.maxBlasDepth + 2, flat);
}

HostBvh buildBvh2(const std::vector<Aabb>& primBounds, unsigned threads, bool median) {
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
    c.median = median;
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
        const Aabb& b = bvh.slotBounds.empty() ? bvh.primBounds[bvh.order[slot]] : bvh.slotBounds[slot];
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


namespace {

// ---- spatial-split BVH ----------------------------------------------------
// Spatial splits (Stich, Friedrich and Dietrich 2009, the method behind the
// reference's SpatialSplitBvh policy, bvh_sbvh.cpp:94-219): besides object
// splits, a node may cut space with a plane; a triangle crossing it is
// referenced on both sides, each reference bounded by the part of the
// triangle inside its side.  Leaves hold one reference, so the tree keeps
// the single-primitive leaf format; a triangle may sit in several leaves.

struct SplitRef {
    uint32_t prim;
    Aabb box;  // bounds of the triangle's part inside this reference's region
};

struct SplitNode {
    Aabb box[2];
    int32_t child[2];  // >= 0: node, < 0: -(reference index) - 1
};

// Shared inputs of a spatial-split build.
struct SpatialParams {
    const std::vector<TriShade>* tris;
    float rootArea = 0.0f;
};

// A built subtree: its leaves in creation order and its interior nodes.
struct SpatialTree {
    std::vector<SplitRef> leaves;
    std::vector<SplitNode> nodes;
};

constexpr int kSpatialBins = 16;
constexpr uint32_t kSpatialMaxDepth = 28;  // deeper nodes take object splits only
constexpr float kSpatialOverlap = 1.0e-5f; // spatial splits only where object splits overlap this much (x root area)
constexpr uint32_t kSpatialParallelDepth = 5;  // subtrees above this depth build on their own threads
constexpr size_t kSpatialParallelMin = 4096;   // ... when they hold at least this many references

inline Aabb unionBox(const Aabb& a, const Aabb& b) {
    Aabb r = a;
    grow(r, b);
    return r;
}

inline Aabb intersectBox(const Aabb& a, const Aabb& b) {
    Aabb r;
    for (int k = 0; k < 3; ++k) {
        r.mn[k] = std::fmax(a.mn[k], b.mn[k]);
        r.mx[k] = std::fmin(a.mx[k], b.mx[k]);
    }
    return r;
}

inline bool emptyOrInverted(const Aabb& b) { return !(b.mn[0] <= b.mx[0] && b.mn[1] <= b.mx[1] && b.mn[2] <= b.mx[2]); }

// Bounds of triangle (a, b, c) clipped to `region` (Sutherland-Hodgman
// against the six planes), intersected with the region to absorb rounding.
Aabb clipTriangle(const Vec3& a, const Vec3& b, const Vec3& c, const Aabb& region) {
    float poly[9][3] = {{a.x, a.y, a.z}, {b.x, b.y, b.z}, {c.x, c.y, c.z}};
    int n = 3;
    for (int axis = 0; axis < 3 && n > 0; ++axis) {
        for (int side = 0; side < 2 && n > 0; ++side) {
            float plane = side == 0 ? region.mn[axis] : region.mx[axis];
            float out[9][3];
            int m = 0;
            for (int i = 0; i < n; ++i) {
                const float* p = poly[i];
                const float* q = poly[(i + 1) % n];
                float dp = side == 0 ? p[axis] - plane : plane - p[axis];
                float dq = side == 0 ? q[axis] - plane : plane - q[axis];
                if (dp >= 0.0f && m < 9) std::memcpy(out[m++], p, sizeof(float) * 3);
                if ((dp < 0.0f) != (dq < 0.0f) && m < 9) {
                    float t = dp / (dp - dq);
                    for (int k = 0; k < 3; ++k) out[m][k] = p[k] + (q[k] - p[k]) * t;
                    out[m][axis] = plane;
                    ++m;
                }
            }
            n = m;
            std::memcpy(poly, out, sizeof(float) * 3 * static_cast<size_t>(m));
        }
    }
    Aabb r = emptyBox();
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            r.mn[k] = std::fmin(r.mn[k], poly[i][k]);
            r.mx[k] = std::fmax(r.mx[k], poly[i][k]);
        }
    r = intersectBox(r, region);
    if (emptyOrInverted(r)) return r;
    // Clipped vertices are rounded: widen by a few ulps so the box surely
    // holds the triangle's part (overlap with the neighbour is harmless).
    for (int k = 0; k < 3; ++k) {
        r.mn[k] -= 4.0e-7f * std::fmax(1.0f, std::fabs(r.mn[k]));
        r.mx[k] += 4.0e-7f * std::fmax(1.0f, std::fabs(r.mx[k]));
    }
    return r;
}

struct SplitChoice {
    float cost = INFINITY;
    int axis = -1;
    bool spatial = false;
    float plane = 0.0f;  // spatial: the cutting plane; object: centroid threshold
};

// Object split: binned SAH over reference centroids.
SplitChoice objectSplit(const std::vector<SplitRef>& refs, const Aabb& centroids) {
    SplitChoice best;
    for (int axis = 0; axis < 3; ++axis) {
        float lo = centroids.mn[axis], extent = centroids.mx[axis] - lo;
        if (!(extent > 0.0f)) continue;
        Aabb box[kBins];
        uint32_t count[kBins] = {};
        for (int b = 0; b < kBins; ++b) box[b] = emptyBox();
        float scale = static_cast<float>(kBins) * (1.0f - 1.0e-5f) / extent;
        for (const SplitRef& r : refs) {
            float cc = (r.box.mn[axis] + r.box.mx[axis]) * 0.5f;
            int b = std::min(std::max(static_cast<int>((cc - lo) * scale), 0), kBins - 1);
            ++count[b];
            grow(box[b], r.box);
        }
        float rightArea[kBins];
        uint32_t rightCount[kBins];
        Aabb acc = emptyBox();
        uint32_t cnt = 0;
        for (int b = kBins - 1; b > 0; --b) {
            grow(acc, box[b]);
            cnt += count[b];
            rightArea[b] = area(acc);
            rightCount[b] = cnt;
        }
        acc = emptyBox();
        cnt = 0;
        for (int b = 0; b < kBins - 1; ++b) {
            grow(acc, box[b]);
            cnt += count[b];
            if (cnt == 0 || rightCount[b + 1] == 0) continue;
            float cost = area(acc) * static_cast<float>(cnt) + rightArea[b + 1] * static_cast<float>(rightCount[b + 1]);
            if (cost < best.cost) {
                best.cost = cost;
                best.axis = axis;
                best.spatial = false;
                best.plane = lo + static_cast<float>(b + 1) / scale;
            }
        }
    }
    return best;
}

// Spatial split: planes at bin boundaries of the node box; each reference
// is clipped into every bin it overlaps.
SplitChoice spatialSplit(const SpatialParams& sp, const std::vector<SplitRef>& refs, const Aabb& nodeBox) {
    SplitChoice best;
    for (int axis = 0; axis < 3; ++axis) {
        float lo = nodeBox.mn[axis], extent = nodeBox.mx[axis] - lo;
        if (!(extent > 0.0f)) continue;
        float width = extent / static_cast<float>(kSpatialBins);
        Aabb box[kSpatialBins];
        uint32_t enter[kSpatialBins] = {}, exit[kSpatialBins] = {};
        for (int b = 0; b < kSpatialBins; ++b) box[b] = emptyBox();
        for (const SplitRef& r : refs) {
            const TriShade& t = (*sp.tris)[r.prim];
            int first = std::min(std::max(static_cast<int>((r.box.mn[axis] - lo) / width), 0), kSpatialBins - 1);
            int last = std::min(std::max(static_cast<int>((r.box.mx[axis] - lo) / width), first), kSpatialBins - 1);
            for (int b = first; b <= last; ++b) {
                Aabb slab = r.box;
                slab.mn[axis] = std::fmax(slab.mn[axis], lo + width * static_cast<float>(b));
                slab.mx[axis] = std::fmin(slab.mx[axis], b == kSpatialBins - 1 ? nodeBox.mx[axis] : lo + width * static_cast<float>(b + 1));
                Aabb part = clipTriangle(t.v0, t.v1, t.v2, slab);
                if (!emptyOrInverted(part)) grow(box[b], part);
            }
            ++enter[first];
            ++exit[last];
        }
        float rightArea[kSpatialBins];
        uint32_t rightCount[kSpatialBins];
        Aabb acc = emptyBox();
        uint32_t cnt = 0;
        for (int b = kSpatialBins - 1; b > 0; --b) {
            grow(acc, box[b]);
            cnt += exit[b];
            rightArea[b] = area(acc);
            rightCount[b] = cnt;
        }
        acc = emptyBox();
        cnt = 0;
        for (int b = 0; b < kSpatialBins - 1; ++b) {
            grow(acc, box[b]);
            cnt += enter[b];
            if (cnt == 0 || rightCount[b + 1] == 0) continue;
            float cost = area(acc) * static_cast<float>(cnt) + rightArea[b + 1] * static_cast<float>(rightCount[b + 1]);
            if (cost < best.cost) {
                best.cost = cost;
                best.axis = axis;
                best.spatial = true;
                best.plane = lo + width * static_cast<float>(b + 1);
            }
        }
    }
    return best;
}

// Appends subtree `sub` to `out`; returns `pointer` (a pointer into `sub`)
// translated into `out`.
int32_t graftSubtree(SpatialTree& out, const SpatialTree& sub, int32_t pointer) {
    int32_t nodeOffset = static_cast<int32_t>(out.nodes.size());
    int32_t leafOffset = static_cast<int32_t>(out.leaves.size());
    auto shift = [&](int32_t p) { return p >= 0 ? p + nodeOffset : p - leafOffset; };
    out.leaves.insert(out.leaves.end(), sub.leaves.begin(), sub.leaves.end());
    for (const SplitNode& n : sub.nodes) out.nodes.push_back(SplitNode{{n.box[0], n.box[1]}, {shift(n.child[0]), shift(n.child[1])}});
    return shift(pointer);
}

// Builds the subtree over `refs` into `out`; returns its child pointer
// (node index, or -(leaf index) - 1) and writes its bounds.  `budget` is
// the most references the subtree may hold: a spatial split that would
// exceed it is not taken, and the children share what is left in
// proportion to their sizes, so the tree does not depend on the order in
// which subtrees are built (the upper levels build in parallel).
int32_t buildSpatial(const SpatialParams& sp, SpatialTree& out, std::vector<SplitRef>& refs, uint32_t depth,
    size_t budget, Aabb& bounds) {
    bounds = emptyBox();
    Aabb centroids = emptyBox();
    for (const SplitRef& r : refs) {
        grow(bounds, r.box);
        Aabb c;
        for (int k = 0; k < 3; ++k) c.mn[k] = c.mx[k] = (r.box.mn[k] + r.box.mx[k]) * 0.5f;
        grow(centroids, c);
    }
    if (refs.size() == 1) {
        out.leaves.push_back(refs[0]);
        return -static_cast<int32_t>(out.leaves.size());
    }
    std::vector<SplitRef> left, right;
    SplitChoice obj = depth < kSpatialMaxDepth ? objectSplit(refs, centroids) : SplitChoice{};
    SplitChoice best = obj;
    if (obj.axis >= 0 && depth < kSpatialMaxDepth && refs.size() < budget) {
        // Spatial splits pay only where the object split's children overlap.
        Aabb l = emptyBox(), r = emptyBox();
        for (const SplitRef& ref : refs) {
            float cc = (ref.box.mn[obj.axis] + ref.box.mx[obj.axis]) * 0.5f;
            grow(cc < obj.plane ? l : r, ref.box);
        }
        Aabb overlap = intersectBox(l, r);
        if (!emptyOrInverted(overlap) && area(overlap) > kSpatialOverlap * sp.rootArea) {
            SplitChoice sc = spatialSplit(sp, refs, bounds);
            if (sc.cost < best.cost) best = sc;
        }
    }
    if (best.axis >= 0 && best.spatial) {
        for (const SplitRef& ref : refs) {
            const TriShade& t = (*sp.tris)[ref.prim];
            if (ref.box.mn[best.axis] < best.plane) {
                Aabb region = ref.box;
                region.mx[best.axis] = std::fmin(region.mx[best.axis], best.plane);
                Aabb part = clipTriangle(t.v0, t.v1, t.v2, region);
                if (!emptyOrInverted(part)) left.push_back(SplitRef{ref.prim, part});
            }
            if (ref.box.mx[best.axis] > best.plane) {
                Aabb region = ref.box;
                region.mn[best.axis] = std::fmax(region.mn[best.axis], best.plane);
                Aabb part = clipTriangle(t.v0, t.v1, t.v2, region);
                if (!emptyOrInverted(part)) right.push_back(SplitRef{ref.prim, part});
            }
        }
        if (left.empty() || right.empty() || left.size() + right.size() < refs.size() ||
            left.size() + right.size() > budget) {
            left.clear();
            right.clear();
            best = obj;
        }
    }
    if (left.empty() && best.axis >= 0 && !best.spatial) {
        for (const SplitRef& ref : refs) {
            float cc = (ref.box.mn[best.axis] + ref.box.mx[best.axis]) * 0.5f;
            (cc < best.plane ? left : right).push_back(ref);
        }
        if (left.empty() || right.empty()) {
            left.clear();
            right.clear();
        }
    }
    if (left.empty()) {
        // Median split on the longest centroid axis (also below the spatial
        // depth limit: halving bounds the depth).
        float ex = centroids.mx[0] - centroids.mn[0], ey = centroids.mx[1] - centroids.mn[1],
              ez = centroids.mx[2] - centroids.mn[2];
        int axis = ex >= ey && ex >= ez ? 0 : (ey >= ez ? 1 : 2);
        size_t mid = refs.size() / 2;
        std::nth_element(refs.begin(), refs.begin() + static_cast<std::ptrdiff_t>(mid), refs.end(),
            [axis](const SplitRef& a, const SplitRef& b) {
                float ca = a.box.mn[axis] + a.box.mx[axis], cb = b.box.mn[axis] + b.box.mx[axis];
                return ca < cb || (ca == cb && a.prim < b.prim);
            });
        left.assign(refs.begin(), refs.begin() + static_cast<std::ptrdiff_t>(mid));
        right.assign(refs.begin() + static_cast<std::ptrdiff_t>(mid), refs.end());
    }
    std::vector<SplitRef>().swap(refs);
    size_t spare = budget - left.size() - right.size();
    size_t leftSpare = static_cast<size_t>(static_cast<double>(spare) * static_cast<double>(left.size()) /
        static_cast<double>(left.size() + right.size()));
    size_t leftBudget = left.size() + leftSpare, rightBudget = right.size() + (spare - leftSpare);
    int32_t index = static_cast<int32_t>(out.nodes.size());
    out.nodes.push_back(SplitNode{});
    Aabb lb, rb;
    int32_t lc, rc;
    if (depth < kSpatialParallelDepth && left.size() + right.size() >= kSpatialParallelMin) {
        SpatialTree lt, rt;
        std::thread worker([&] { lc = buildSpatial(sp, lt, left, depth + 1, leftBudget, lb); });
        rc = buildSpatial(sp, rt, right, depth + 1, rightBudget, rb);
        worker.join();
        lc = graftSubtree(out, lt, lc);
        rc = graftSubtree(out, rt, rc);
    } else {
        lc = buildSpatial(sp, out, left, depth + 1, leftBudget, lb);
        rc = buildSpatial(sp, out, right, depth + 1, rightBudget, rb);
    }
    SplitNode& node = out.nodes[static_cast<size_t>(index)];
    node.box[0] = lb;
    node.box[1] = rb;
    node.child[0] = lc;
    node.child[1] = rc;
    return index;
}

void setChild(BvhNode& n, bool right, const Aabb& b, uint32_t pointer) {
    float* v = right ? &n.rminx : &n.lminx;
    if (!right) {
        n.lminx = b.mn[0]; n.lminy = b.mn[1]; n.lminz = b.mn[2];
        n.lmaxx = b.mx[0]; n.lmaxy = b.mx[1]; n.lmaxz = b.mx[2];
        n.left = pointer;
    } else {
        n.rminx = b.mn[0]; n.rminy = b.mn[1]; n.rminz = b.mn[2];
        n.rmaxx = b.mx[0]; n.rmaxy = b.mx[1]; n.rmaxz = b.mx[2];
        n.right = pointer;
    }
    (void)v;
}

}  // namespace

HostBvh buildSpatialSplitBvh2(const std::vector<TriShade>& tris) {
    HostBvh bvh;
    uint32_t n = static_cast<uint32_t>(tris.size());
    for (const TriShade& t : tris) bvh.primBounds.push_back(triangleBounds(t.v0, t.v1, t.v2));
    if (n == 0) return bvh;
    SpatialParams sp;
    sp.tris = &tris;
    SpatialTree tree;
    std::vector<SplitRef> refs(n);
    Aabb all = emptyBox();
    for (uint32_t i = 0; i < n; ++i) {
        refs[i] = SplitRef{i, bvh.primBounds[i]};
        grow(all, bvh.primBounds[i]);
    }
    sp.rootArea = area(all);
    Aabb rootBounds;
    // At most one extra reference per triangle.
    int32_t root = buildSpatial(sp, tree, refs, 0, static_cast<size_t>(n) * 2, rootBounds);
    // Renumber depth first: a subtree over m leaves owns m - 1 consecutive
    // interior nodes, leaves take slots in the order the walk meets them.
    uint32_t leaves = static_cast<uint32_t>(tree.leaves.size());
    bvh.nodes.resize(leaves > 1 ? leaves - 1 : 0);
    bvh.order.reserve(leaves);
    bvh.slotBounds.reserve(leaves);
    uint32_t nextNode = 0;
    std::function<uint32_t(int32_t, uint32_t&)> emit = [&](int32_t ptr, uint32_t& depth) -> uint32_t {
        if (ptr < 0) {
            const SplitRef& r = tree.leaves[static_cast<size_t>(-ptr - 1)];
            uint32_t slot = static_cast<uint32_t>(bvh.order.size());
            bvh.order.push_back(r.prim);
            bvh.slotBounds.push_back(r.box);
            depth = 0;
            return kLeafBit | slot;
        }
        uint32_t self = nextNode++;
        const SplitNode sn = tree.nodes[static_cast<size_t>(ptr)];
        uint32_t dl = 0, dr = 0;
        uint32_t l = emit(sn.child[0], dl);
        uint32_t r = emit(sn.child[1], dr);
        BvhNode& node = bvh.nodes[self];
        setChild(node, false, sn.box[0], l);
        setChild(node, true, sn.box[1], r);
        node.pad0 = node.pad1 = 0;
        depth = 1 + std::max(dl, dr);
        return self;
    };
    uint32_t depth = 0;
    uint32_t rootPointer = emit(root, depth);
    bvh.root = rootPointer;
    bvh.rootBounds = rootBounds;
    bvh.depth = depth;
    return bvh;
}

std::vector<WideNode> collapseWide(const HostBvh& bvh, uint32_t* stackNeed) {
    std::vector<WideNode> wide;
    *stackNeed = 1;
    if (bvh.primCount() < 2) return wide;
    struct Item {
        uint32_t pointer;
        Aabb box;
    };
    // Children of binary node i.
    auto children = [&](uint32_t i, Item out[2]) {
        const BvhNode& n = bvh.nodes[i];
        out[0] = Item{n.left, nodeChildBounds(n, false)};
        out[1] = Item{n.right, nodeChildBounds(n, true)};
    };
    std::function<uint32_t(uint32_t, uint32_t)> build = [&](uint32_t binary, uint32_t stackSoFar) -> uint32_t {
        Item items[kWideChildren];
        int count = 0;
        Item two[2];
        children(binary, two);
        items[count++] = two[0];
        items[count++] = two[1];
        // Open the interior child of largest area until eight children.
        while (count < static_cast<int>(kWideChildren)) {
            int pick = -1;
            float best = -1.0f;
            for (int k = 0; k < count; ++k) {
                if (items[k].pointer & kLeafBit) continue;
                float a = area(items[k].box);
                if (a > best) {
                    best = a;
                    pick = k;
                }
            }
            if (pick < 0) break;
            children(items[pick].pointer, two);
            items[pick] = two[0];
            items[count++] = two[1];
        }
        uint32_t self = static_cast<uint32_t>(wide.size());
        wide.push_back(WideNode{});
        uint32_t need = stackSoFar + static_cast<uint32_t>(count) - 1u;
        *stackNeed = std::max(*stackNeed, need + 1u);
        uint32_t ptrs[kWideChildren];
        for (int k = 0; k < count; ++k)
            ptrs[k] = (items[k].pointer & kLeafBit) ? items[k].pointer : build(items[k].pointer, need);
        WideNode& w = wide[self];
        for (uint32_t k = 0; k < kWideChildren; ++k) {
            bool used = k < static_cast<uint32_t>(count);
            const Aabb b = used ? items[k].box : emptyBox();
            w.minX[k] = b.mn[0];
            w.minY[k] = b.mn[1];
            w.minZ[k] = b.mn[2];
            w.maxX[k] = b.mx[0];
            w.maxY[k] = b.mx[1];
            w.maxZ[k] = b.mx[2];
            w.child[k] = used ? ptrs[k] : kInvalid;
        }
        return self;
    };
    build(bvh.root, 0);
    return wide;
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
    if (desc.flat_accelerator > WPT_FLAT_WIDE8) return "unknown flat-triangle accelerator";
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

    // Flat triangles get a BVH only when the scene has no meshes, built by
    // the requested method (reference uploadDirectAcceleration,
    // scene.cu:241-294).
    if (desc.mesh_count == 0 && !out.flatTriangles.empty()) {
        out.flatAccelerator = desc.flat_accelerator;
        if (out.flatAccelerator == WPT_FLAT_SPATIAL_SPLIT_BVH2) {
            out.flat = buildSpatialSplitBvh2(out.flatTriangles);
        } else {
            std::vector<Aabb> bounds;
            bounds.reserve(out.flatTriangles.size());
            for (const TriShade& t : out.flatTriangles) bounds.push_back(triangleBounds(t.v0, t.v1, t.v2));
            out.flat = buildBvh2(bounds, threads, out.flatAccelerator == WPT_FLAT_MEDIAN_BVH2);
        }
        out.flatGeometry = slotGeometry(out.flatTriangles, out.flat);
        if (out.flatAccelerator == WPT_FLAT_WIDE8) out.flatWide = collapseWide(out.flat, &out.flatWideStack);
    }
    out.maxStackDepth = stackDepth(out);
    return std::string();
}

}  // namespace wpt
