// Closest-hit and shadow any-hit traversal over the device scene.
//
// Reference: include/renderer/device_traversal.cuh:24-91 (sphere/triangle
// tests), :182-222 (slab test), :239-313 (BLAS BVH2), :789-878 (sphere leaf,
// mesh instance, TLAS), :1035-1184 (isOccluded, traceRay).
//
// The node layout differs from the reference (64 B nodes holding both child
// boxes, or 256 B eight-wide nodes for flat triangles, see device_scene.h),
// and one loop serves closest-hit and any-hit queries.  The primitive tests
// are the reference's, in the same FP32 operation order, so each candidate's
// distance and barycentrics are bit-identical to the reference's.  The
// nearest hit is the candidate first in (distance, category, instance,
// primitive) order.  That is the reference's result whenever the nearest
// distance is held by one primitive, and across categories on exact ties;
// on an exact tie inside one tree (a ray through a shared edge) the
// reference keeps whichever triangle its tree visits first, while this
// order keeps the lower index, whatever the tree.
#pragma once

#include "../common/vec.h"
#include "../scene/device_scene.h"
#include "material.h"

namespace wpt {

constexpr float kNoHit = 1e30f;

// Winning primitive of a closest-hit query.
struct HitRecord {
    float t;
    float u, v;
    uint32_t kind;      // WPT_HIT_* values
    uint32_t instance;  // instance index for instance triangles
    uint32_t prim;      // sphere / triangle index in the caller's order
    uint32_t material;  // resolved (override applied)
    uint32_t rank;      // tie order of the hit's category (kRank*)
};

// Hits at exactly the same distance are ordered by category -- spheres,
// direct-instance triangles, static triangles, TLAS-instance triangles,
// flat triangles -- then by instance, then by primitive.  Across categories
// this is the reference's outcome (traceRay visits the categories in that
// order and a sphere displaces an equally distant triangle,
// device_traversal.cuh:789-800,1115-1184).  Within a BVH the reference keeps
// whichever tie its tree visits first; fixing the order instead makes the
// nearest hit independent of tree shape and visit order, so that any BVH
// (binary or wide, host- or device-built) gives the same hits.
constexpr uint32_t kRankSphere = 0, kRankDirectInstance = 1, kRankStatic = 2, kRankTlasInstance = 3, kRankFlat = 4;

// Boxes are culled only beyond tieReach(closestT): a box holding a hit at
// exactly closestT can have a computed entry distance a few ulps above it,
// and must still be opened for that hit to take part in the tie order.
WPT_HD float tieReach(float closestT) { return closestT * 1.00000095f; }

WPT_HD bool nearer(float t, uint32_t rank, uint32_t instance, uint32_t prim, const HitRecord& best) {
    if (t != best.t) return t < best.t;
    if (rank != best.rank) return rank < best.rank;
    if (instance != best.instance) return instance < best.instance;
    return prim < best.prim;
}

// Alpha mask of a candidate triangle hit (reference intersectTriangleHit,
// device_traversal.cuh:146-151): masked materials below their cutoff are
// passed through; textured ones test opacity x albedo alpha at the hit.
WPT_HD bool textureCutAway(const DevScene& s, uint32_t material, const TriShade& tri, float u, float v) {
    const Material& m = s.materials[material];
    return materialOpacity(s, m, triangleUv(tri, u, v), triangleUv1(s, tri, u, v)) < m.alphaCutoff;
}

// `Full` kernels serve every scene; the others only scenes without
// environment maps, punctual lights or texture bindings, and leave that code
// out (it costs registers on every path).
template <bool Full>
WPT_HD bool cutAway(const DevScene& s, uint32_t material, const TriShade* tris, uint32_t prim, float u, float v) {
    uint8_t policy = s.materialCutout[material];
    if (policy == kCutNever) return false;
    return policy == kCutAlways || (Full && textureCutAway(s, material, tris[prim], u, v));
}

constexpr uint32_t kHitNone = 0, kHitSphere = 1, kHitStatic = 2, kHitInstance = 3, kHitFlat = 4;

WPT_HD HitRecord emptyHit() {
    HitRecord h;
    h.t = kNoHit;
    h.u = h.v = 0.0f;
    h.kind = kHitNone;
    h.instance = kInvalid;
    h.prim = kInvalid;
    h.material = kInvalid;
    h.rank = kInvalid;
    return h;
}

// Quadratic with unnormalized b; nearest root beyond the cutoff.
WPT_HD bool intersectSphere(const Ray& ray, const Vec3& center, float radius, float& t) {
    Vec3 oc = ray.origin - center;
    float a = dot(ray.direction, ray.direction);
    float b = 2.0f * dot(oc, ray.direction);
    float c = dot(oc, oc) - radius * radius;
    float discriminant = b * b - 4.0f * a * c;
    if (discriminant < 0.0f) return false;
    float sqrtD = sqrtf(discriminant);
    float t0 = (-b - sqrtD) / (2.0f * a);
    float t1 = (-b + sqrtD) / (2.0f * a);
    if (t0 > kRayMinDistance) {
        t = t0;
        return true;
    }
    if (t1 > kRayMinDistance) {
        t = t1;
        return true;
    }
    return false;
}

// Moller-Trumbore with the determinant guard relative to its own products.
WPT_HD bool intersectTriangle(const Ray& ray, const Vec3& v0, const Vec3& v1, const Vec3& v2, float& t, float& u,
    float& v) {
    Vec3 edge1 = v1 - v0;
    Vec3 edge2 = v2 - v0;
    Vec3 p = cross(ray.direction, edge2);
    float determinant = dot(edge1, p);
    const Vec3& d = ray.direction;
    float scale = fabsf(edge1.x) * (fabsf(d.y * edge2.z) + fabsf(d.z * edge2.y)) +
        fabsf(edge1.y) * (fabsf(d.z * edge2.x) + fabsf(d.x * edge2.z)) +
        fabsf(edge1.z) * (fabsf(d.x * edge2.y) + fabsf(d.y * edge2.x));
    if (!isFinite(determinant) || fabsf(determinant) <= 8.0f * kFltEpsilon * scale) return false;
    float inverseDeterminant = 1.0f / determinant;
    Vec3 originToVertex = ray.origin - v0;
    u = dot(originToVertex, p) * inverseDeterminant;
    if (u < 0.0f || u > 1.0f) return false;
    Vec3 q = cross(originToVertex, edge1);
    v = dot(ray.direction, q) * inverseDeterminant;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(edge2, q) * inverseDeterminant;
    return t > kRayMinDistance;
}

struct RayInv {
    Ray ray;
    Vec3 inv;
};

WPT_HD RayInv makeRayInv(const Ray& r) {
    RayInv ri;
    ri.ray = r;
    ri.inv = Vec3(1.0f / r.direction.x, 1.0f / r.direction.y, 1.0f / r.direction.z);
    return ri;
}

WPT_HD bool slab(const RayInv& r, float mnx, float mny, float mnz, float mxx, float mxy, float mxz, float closestT,
    float& nearT) {
    float tx0 = (mnx - r.ray.origin.x) * r.inv.x;
    float tx1 = (mxx - r.ray.origin.x) * r.inv.x;
    float ty0 = (mny - r.ray.origin.y) * r.inv.y;
    float ty1 = (mxy - r.ray.origin.y) * r.inv.y;
    float tz0 = (mnz - r.ray.origin.z) * r.inv.z;
    float tz1 = (mxz - r.ray.origin.z) * r.inv.z;
    float tmin = fmaxf(fminf(tx0, tx1), fmaxf(fminf(ty0, ty1), fminf(tz0, tz1)));
    float tmax = fminf(fmaxf(tx0, tx1), fminf(fmaxf(ty0, ty1), fmaxf(tz0, tz1)));
    nearT = tmin;
    return tmax >= fmaxf(tmin, kRayMinDistance) && tmin <= tieReach(closestT);
}

WPT_HD Vec3 xyz(const float4& f) { return Vec3(f.x, f.y, f.z); }
WPT_HD uint32_t wbits(const float4& f) { return bitsOf(f.w); }

// Traversal stacks hold (node pointer, entry distance) pairs and expose
// `size`, push, pop and truncate.  This private-array stack serves host-side
// builds of the traversal code; the kernels use an LDS-resident stack with
// the same interface.  The scene build rejects trees whose combined
// TLAS+BLAS depth exceeds kMaxStackDepth.
constexpr uint32_t kMaxStackDepth = 64;
struct PrivateStack {
    static constexpr uint32_t kCapacity = kMaxStackDepth;
    uint32_t node[kCapacity];
    float tmin[kCapacity];
    uint32_t size = 0;
    WPT_HD void push(uint32_t n, float t) {
        node[size] = n;
        tmin[size] = t;
        ++size;
    }
    WPT_HD void pop(uint32_t& n, float& t) {
        --size;
        n = node[size];
        t = tmin[size];
    }
    WPT_HD void truncate(uint32_t base) { size = base; }
    // Entry k below the top (k < 8).
    WPT_HD uint2 peek(uint32_t k) const { return make_uint2(node[size - 1 - k], bitsOf(tmin[size - 1 - k])); }
    WPT_HD void poke(uint32_t k, uint2 e) {
        node[size - 1 - k] = e.x;
        tmin[size - 1 - k] = fromBits32(e.y);
    }
};

// Pops until an entry still closer than closestT; false when the traversal
// level (entries above `base`) is exhausted.
template <class Stack>
WPT_HD bool popLive(Stack& stack, uint32_t base, float closestT, uint32_t& cur) {
    while (stack.size > base) {
        float t;
        stack.pop(cur, t);
        if (t <= tieReach(closestT)) return true;
    }
    return false;
}

// Query kinds.  A closest-hit query keeps the nearest hit in the tie order
// above; an any-hit query (shadow rays) stops at the first blocker closer
// than its distance limit.  kEither picks one of the two per lane at run
// time, so that lanes of one wave running different kinds of query step
// through the same traversal loop together.
enum class Query { kClosest, kAny, kEither };

// Offers a candidate at distance t to the query.  Any-hit queries accept
// anything strictly closer than their limit (held in best.t); closest-hit
// queries apply the tie order.
WPT_HD bool wanted(bool any, float t, uint32_t rank, uint32_t instance, uint32_t prim, const HitRecord& best) {
    return any ? t < best.t : nearer(t, rank, instance, prim, best);
}

WPT_HD void record(HitRecord& best, float t, float u, float v, uint32_t kind, uint32_t instance, uint32_t prim,
    uint32_t material, uint32_t rank) {
    best.t = t;
    best.u = u;
    best.v = v;
    best.kind = kind;
    best.instance = instance;
    best.prim = prim;
    best.material = material;
    best.rank = rank;
}

// Triangles of one BVH (the flat-triangle BVH, world space).  Returns true
// when an any-hit query found its blocker.
template <bool Full, class Stack>
WPT_HD bool queryBvh(const DevScene& s, const RayInv& r, const Bvh& bvh, const float4* geometry, const TriShade* tris,
    uint32_t materialOverride, uint32_t kind, uint32_t rank, uint32_t instance, bool any, HitRecord& best,
    Stack& stack) {
    float nearT;
    if (!slab(r, bvh.rootMin[0], bvh.rootMin[1], bvh.rootMin[2], bvh.rootMax[0], bvh.rootMax[1], bvh.rootMax[2],
            best.t, nearT))
        return false;
    uint32_t base = stack.size;
    uint32_t cur = bvh.root;
    while (true) {
        if (cur & kLeafBit) {
            uint32_t slot = cur & ~kLeafBit;
            float4 a = geometry[3 * slot + 0];
            float4 b = geometry[3 * slot + 1];
            float4 c = geometry[3 * slot + 2];
            float t, u, v;
            if (intersectTriangle(r.ray, xyz(a), xyz(b), xyz(c), t, u, v) &&
                wanted(any, t, rank, instance, wbits(a), best)) {
                uint32_t material = materialOverride == kInvalid ? wbits(b) : materialOverride;
                if (!cutAway<Full>(s, material, tris, wbits(a), u, v)) {
                    record(best, t, u, v, kind, instance, wbits(a), material, rank);
                    if (any) {
                        stack.truncate(base);
                        return true;
                    }
                }
            }
        } else {
            const BvhNode& n = bvh.nodes[cur];
            float tl, tr;
            bool hl = slab(r, n.lminx, n.lminy, n.lminz, n.lmaxx, n.lmaxy, n.lmaxz, best.t, tl);
            bool hr = slab(r, n.rminx, n.rminy, n.rminz, n.rmaxx, n.rmaxy, n.rmaxz, best.t, tr);
            if (hl && hr) {
                bool leftNear = tl < tr;
                stack.push(leftNear ? n.right : n.left, leftNear ? tr : tl);
                cur = leftNear ? n.left : n.right;
                continue;
            }
            if (hl) {
                cur = n.left;
                continue;
            }
            if (hr) {
                cur = n.right;
                continue;
            }
        }
        if (!popLive(stack, base, best.t, cur)) return false;
    }
}

// Triangles under eight-wide nodes (the flat-triangle accelerator
// WPT_FLAT_WIDE8).  A node's children that the ray reaches are pushed so that
// the nearest ends on top (insertion from the top of the stack, at most
// eight entries, all in the stack's LDS part), then the walk pops as for
// binary nodes.  Returns true when an any-hit query found its blocker.
template <bool Full, class Stack>
WPT_HD bool queryWide(const DevScene& s, const RayInv& r, const Bvh& bvh, const WideNode* nodes, const float4* geometry,
    const TriShade* tris, uint32_t kind, uint32_t rank, bool any, HitRecord& best, Stack& stack) {
    float nearT;
    if (!slab(r, bvh.rootMin[0], bvh.rootMin[1], bvh.rootMin[2], bvh.rootMax[0], bvh.rootMax[1], bvh.rootMax[2],
            best.t, nearT))
        return false;
    uint32_t base = stack.size;
    uint32_t cur = bvh.primCount > 1 ? 0u : (kLeafBit | 0u);
    while (true) {
        if (cur & kLeafBit) {
            uint32_t slot = cur & ~kLeafBit;
            float4 a = geometry[3 * slot + 0];
            float4 b = geometry[3 * slot + 1];
            float4 c = geometry[3 * slot + 2];
            float t, u, v;
            if (intersectTriangle(r.ray, xyz(a), xyz(b), xyz(c), t, u, v) &&
                wanted(any, t, rank, kInvalid, wbits(a), best)) {
                uint32_t material = wbits(b);
                if (!cutAway<Full>(s, material, tris, wbits(a), u, v)) {
                    record(best, t, u, v, kind, kInvalid, wbits(a), material, rank);
                    if (any) {
                        stack.truncate(base);
                        return true;
                    }
                }
            }
        } else {
            const WideNode& n = nodes[cur];
            uint32_t pushed = 0;
            for (uint32_t k = 0; k < kWideChildren; ++k) {
                uint32_t child = n.child[k];
                if (child == kInvalid) break;
                float t;
                if (!slab(r, n.minX[k], n.minY[k], n.minZ[k], n.maxX[k], n.maxY[k], n.maxZ[k], best.t, t)) continue;
                stack.push(child, t);
                // Keep the pushed entries ordered, nearest on top.
                for (uint32_t j = 0; j < pushed; ++j) {
                    uint2 top = stack.peek(j), below = stack.peek(j + 1);
                    if (!(fromBits32(top.y) > fromBits32(below.y))) break;
                    stack.poke(j, below);
                    stack.poke(j + 1, top);
                }
                ++pushed;
            }
        }
        if (!popLive(stack, base, best.t, cur)) return false;
    }
}

// Brute force over original-order triangles (direct instances, static
// triangles, flat triangles without a BVH).  Returns true when an any-hit
// query found its blocker.
template <bool Full>
WPT_HD bool queryBrute(const DevScene& s, const Ray& ray, const TriShade* tris, uint32_t count,
    uint32_t materialOverride, uint32_t kind, uint32_t rank, uint32_t instance, bool any, HitRecord& best) {
    for (uint32_t i = 0; i < count; ++i) {
        const TriShade& tri = tris[i];
        float t, u, v;
        if (!intersectTriangle(ray, tri.v0, tri.v1, tri.v2, t, u, v) || !wanted(any, t, rank, instance, i, best))
            continue;
        uint32_t material = materialOverride == kInvalid ? tri.material : materialOverride;
        if (cutAway<Full>(s, material, tris, i, u, v)) continue;
        record(best, t, u, v, kind, instance, i, material, rank);
        if (any) return true;
    }
    return false;
}

WPT_HD Ray toLocal(const InstanceTransform& tr, const Ray& ray) {
    Ray local;
    local.origin = inverseTransformVector(tr, ray.origin - tr.translation);
    local.direction = inverseTransformVector(tr, ray.direction);
    return local;
}

// Sphere candidate: nearer root wins; on an exact tie the lower sphere index
// wins (reference device_traversal.cuh:789-800).  Returns true when an
// any-hit query found its blocker.
WPT_HD bool querySphere(const DevScene& s, const Ray& ray, uint32_t index, bool any, HitRecord& best) {
    float4 sp = s.spheres[index];
    float t;
    if (!intersectSphere(ray, xyz(sp), sp.w, t) || !wanted(any, t, kRankSphere, kInvalid, index, best)) return false;
    record(best, t, 0.0f, 0.0f, kHitSphere, kInvalid, index, s.sphereMaterial[index], kRankSphere);
    return any;
}

// One query in the reference's category order: direct spheres, direct
// instances, static triangles, TLAS, flat triangles (traceRay /
// isOccluded, device_traversal.cuh:1035-1184).  `maxDistance` bounds
// any-hit queries.  The result's kind is kHitNone when nothing was found;
// for an any-hit query any other kind means the ray is blocked.
//
// The TLAS and the BLASes of its instance leaves are walked by one loop: on
// an instance leaf whose BLAS root box is hit, the loop switches its node
// array, ray and stack floor to the BLAS; when the BLAS level is exhausted
// it switches back and resumes popping the TLAS.  The visit order is that of
// the nested walk (each BLAS finishes before the TLAS resumes), but lanes at
// different levels -- and, for kEither, lanes running different kinds of
// query -- run the same node-visit code in the same iteration.
//
// TlasWalk is the same query as a resumable walk: start() runs the
// categories before the TLAS and enters its root, step() makes one
// iteration of traceQuery's loop (one node or leaf), finish() runs the
// categories after the TLAS.  A kernel can thus interleave the walks of
// several rays in one lane (a lane whose ray ends starts its next one while
// the other lanes of its wave still walk theirs); each ray's visits, and so
// its result, are those of traceQuery.  traceQuery keeps its own loop: as
// start/step/finish its code is slower (the sphere field's closest-hit pass: 69 -> 80 ms
// per 64-sample frame).
enum class Walk { kTlas, kAfter, kDone };  // walking the TLAS; categories after it left; complete

template <bool Full, Query Q>
struct TlasWalk {
    Ray ray;
    RayInv r;
    HitRecord best;
    const BvhNode* nodes;
    const float4* geometry;
    uint32_t cur, instance, materialOverride, base, blasBase;
    bool anyHit;

    WPT_HD bool any() const { return Q == Query::kAny || (Q == Query::kEither && anyHit); }

    template <class Stack>
    WPT_HD Walk start(const DevScene& s, const Ray& query, float maxDistance, bool shadow, Stack& stack) {
        ray = query;
        anyHit = shadow;
        best = emptyHit();
        if (any()) best.t = maxDistance;
        for (uint32_t i = 0; i < s.directSphereCount; ++i)
            if (querySphere(s, ray, s.directSpheres[i], any(), best)) return Walk::kDone;
        for (uint32_t i = 0; i < s.directInstanceCount; ++i) {
            uint32_t index = s.directInstances[i];
            const Instance& inst = s.instances[index];
            const Blas& blas = s.blases[inst.mesh];
            if (queryBrute<Full>(s, toLocal(inst.transform, ray), blas.triangles, blas.triangleCount,
                    inst.materialOverride, kHitInstance, kRankDirectInstance, index, any(), best))
                return Walk::kDone;
        }
        if (queryBrute<Full>(s, ray, s.staticTriangles, s.staticCount, kInvalid, kHitStatic, kRankStatic, kInvalid,
                any(), best))
            return Walk::kDone;
        if (s.tlas.primCount == 0) return Walk::kAfter;
        r = makeRayInv(ray);
        float nearT;
        if (!slab(r, s.tlas.rootMin[0], s.tlas.rootMin[1], s.tlas.rootMin[2], s.tlas.rootMax[0], s.tlas.rootMax[1],
                s.tlas.rootMax[2], best.t, nearT))
            return Walk::kAfter;
        base = stack.size;
        blasBase = base;
        cur = s.tlas.root;
        nodes = s.tlas.nodes;
        geometry = nullptr;
        instance = kInvalid;  // kInvalid while in the TLAS
        materialOverride = kInvalid;
        return Walk::kTlas;
    }

    template <class Stack>
    WPT_HD Walk step(const DevScene& s, Stack& stack) {
        if (cur & kLeafBit) {
            uint32_t slot = cur & ~kLeafBit;
            if (instance != kInvalid) {
                float4 a = geometry[3 * slot + 0];
                float4 b = geometry[3 * slot + 1];
                float4 c = geometry[3 * slot + 2];
                float t, u, v;
                if (intersectTriangle(r.ray, xyz(a), xyz(b), xyz(c), t, u, v) &&
                    wanted(any(), t, kRankTlasInstance, instance, wbits(a), best)) {
                    uint32_t material = materialOverride == kInvalid ? wbits(b) : materialOverride;
                    if (s.materialCutout[material] == kCutNever ||
                        !cutAway<Full>(s, material, s.blases[s.instances[instance].mesh].triangles, wbits(a), u, v)) {
                        record(best, t, u, v, kHitInstance, instance, wbits(a), material, kRankTlasInstance);
                        if (any()) {
                            stack.truncate(base);
                            return Walk::kDone;
                        }
                    }
                }
            } else {
                uint32_t leaf = s.tlasLeaves[slot];
                if (leaf & kTlasSphereBit) {
                    if (querySphere(s, ray, leaf & ~kTlasSphereBit, any(), best)) {
                        stack.truncate(base);
                        return Walk::kDone;
                    }
                } else {
                    const Instance& inst = s.instances[leaf];
                    const Blas& blas = s.blases[inst.mesh];
                    Ray local = toLocal(inst.transform, ray);
                    if (blas.bvh.primCount != 0) {
                        RayInv rl = makeRayInv(local);
                        float nearT;
                        if (slab(rl, blas.bvh.rootMin[0], blas.bvh.rootMin[1], blas.bvh.rootMin[2], blas.bvh.rootMax[0],
                                blas.bvh.rootMax[1], blas.bvh.rootMax[2], best.t, nearT)) {
                            r = rl;
                            instance = leaf;
                            nodes = blas.bvh.nodes;
                            geometry = blas.geometry;
                            materialOverride = inst.materialOverride;
                            blasBase = stack.size;
                            cur = blas.bvh.root;
                            return Walk::kTlas;
                        }
                    } else if (queryBrute<Full>(s, local, blas.triangles, blas.triangleCount, inst.materialOverride,
                                   kHitInstance, kRankTlasInstance, leaf, any(), best)) {
                        stack.truncate(base);
                        return Walk::kDone;
                    }
                }
            }
        } else {
            const BvhNode& n = nodes[cur];
            float tl, tr;
            bool hl = slab(r, n.lminx, n.lminy, n.lminz, n.lmaxx, n.lmaxy, n.lmaxz, best.t, tl);
            bool hr = slab(r, n.rminx, n.rminy, n.rminz, n.rmaxx, n.rmaxy, n.rmaxz, best.t, tr);
            if (hl && hr) {
                bool leftNear = tl < tr;
                stack.push(leftNear ? n.right : n.left, leftNear ? tr : tl);
                cur = leftNear ? n.left : n.right;
                return Walk::kTlas;
            }
            if (hl) {
                cur = n.left;
                return Walk::kTlas;
            }
            if (hr) {
                cur = n.right;
                return Walk::kTlas;
            }
        }
        if (instance != kInvalid) {
            if (popLive(stack, blasBase, best.t, cur)) return Walk::kTlas;
            instance = kInvalid;
            nodes = s.tlas.nodes;
            r = makeRayInv(ray);
        }
        return popLive(stack, base, best.t, cur) ? Walk::kTlas : Walk::kAfter;
    }

    template <class Stack>
    WPT_HD void finish(const DevScene& s, Stack& stack) {
        if (s.flatCount == 0) return;
        if (s.flatWide) {
            RayInv rf = makeRayInv(ray);
            queryWide<Full>(s, rf, s.flat, s.flatWide, s.flatGeometry, s.flatTriangles, kHitFlat, kRankFlat, any(), best,
                stack);
        } else if (s.flat.primCount != 0) {
            RayInv rf = makeRayInv(ray);
            queryBvh<Full>(s, rf, s.flat, s.flatGeometry, s.flatTriangles, kInvalid, kHitFlat, kRankFlat, kInvalid,
                any(), best, stack);
        } else {
            queryBrute<Full>(s, ray, s.flatTriangles, s.flatCount, kInvalid, kHitFlat, kRankFlat, kInvalid, any(),
                best);
        }
    }
};

template <bool Full, Query Q, class Stack>
WPT_HD HitRecord traceQuery(const DevScene& s, const Ray& ray, float maxDistance, bool anyHit, Stack& stack) {
    const bool any = Q == Query::kAny || (Q == Query::kEither && anyHit);
    HitRecord best = emptyHit();
    if (any) best.t = maxDistance;
    for (uint32_t i = 0; i < s.directSphereCount; ++i)
        if (querySphere(s, ray, s.directSpheres[i], any, best)) return best;
    for (uint32_t i = 0; i < s.directInstanceCount; ++i) {
        uint32_t index = s.directInstances[i];
        const Instance& inst = s.instances[index];
        const Blas& blas = s.blases[inst.mesh];
        if (queryBrute<Full>(s, toLocal(inst.transform, ray), blas.triangles, blas.triangleCount, inst.materialOverride,
                kHitInstance, kRankDirectInstance, index, any, best))
            return best;
    }
    if (queryBrute<Full>(s, ray, s.staticTriangles, s.staticCount, kInvalid, kHitStatic, kRankStatic, kInvalid, any,
            best))
        return best;
    if (s.tlas.primCount != 0) {
        RayInv r = makeRayInv(ray);
        float nearT;
        if (slab(r, s.tlas.rootMin[0], s.tlas.rootMin[1], s.tlas.rootMin[2], s.tlas.rootMax[0], s.tlas.rootMax[1],
                s.tlas.rootMax[2], best.t, nearT)) {
            const uint32_t base = stack.size;
            uint32_t blasBase = base;
            uint32_t cur = s.tlas.root;
            const BvhNode* nodes = s.tlas.nodes;
            const float4* geometry = nullptr;
            uint32_t instance = kInvalid;  // kInvalid while in the TLAS
            uint32_t materialOverride = kInvalid;
            while (true) {
                if (cur & kLeafBit) {
                    uint32_t slot = cur & ~kLeafBit;
                    if (instance != kInvalid) {
                        float4 a = geometry[3 * slot + 0];
                        float4 b = geometry[3 * slot + 1];
                        float4 c = geometry[3 * slot + 2];
                        float t, u, v;
                        if (intersectTriangle(r.ray, xyz(a), xyz(b), xyz(c), t, u, v) &&
                            wanted(any, t, kRankTlasInstance, instance, wbits(a), best)) {
                            uint32_t material = materialOverride == kInvalid ? wbits(b) : materialOverride;
                            if (s.materialCutout[material] == kCutNever ||
                                !cutAway<Full>(s, material, s.blases[s.instances[instance].mesh].triangles, wbits(a), u, v)) {
                                record(best, t, u, v, kHitInstance, instance, wbits(a), material, kRankTlasInstance);
                                if (any) {
                                    stack.truncate(base);
                                    return best;
                                }
                            }
                        }
                    } else {
                        uint32_t leaf = s.tlasLeaves[slot];
                        if (leaf & kTlasSphereBit) {
                            if (querySphere(s, ray, leaf & ~kTlasSphereBit, any, best)) {
                                stack.truncate(base);
                                return best;
                            }
                        } else {
                            const Instance& inst = s.instances[leaf];
                            const Blas& blas = s.blases[inst.mesh];
                            Ray local = toLocal(inst.transform, ray);
                            if (blas.bvh.primCount != 0) {
                                RayInv rl = makeRayInv(local);
                                if (slab(rl, blas.bvh.rootMin[0], blas.bvh.rootMin[1], blas.bvh.rootMin[2],
                                        blas.bvh.rootMax[0], blas.bvh.rootMax[1], blas.bvh.rootMax[2], best.t, nearT)) {
                                    r = rl;
                                    instance = leaf;
                                    nodes = blas.bvh.nodes;
                                    geometry = blas.geometry;
                                    materialOverride = inst.materialOverride;
                                    blasBase = stack.size;
                                    cur = blas.bvh.root;
                                    continue;
                                }
                            } else if (queryBrute<Full>(s, local, blas.triangles, blas.triangleCount,
                                           inst.materialOverride, kHitInstance, kRankTlasInstance, leaf, any, best)) {
                                stack.truncate(base);
                                return best;
                            }
                        }
                    }
                } else {
                    const BvhNode& n = nodes[cur];
                    float tl, tr;
                    bool hl = slab(r, n.lminx, n.lminy, n.lminz, n.lmaxx, n.lmaxy, n.lmaxz, best.t, tl);
                    bool hr = slab(r, n.rminx, n.rminy, n.rminz, n.rmaxx, n.rmaxy, n.rmaxz, best.t, tr);
                    if (hl && hr) {
                        bool leftNear = tl < tr;
                        stack.push(leftNear ? n.right : n.left, leftNear ? tr : tl);
                        cur = leftNear ? n.left : n.right;
                        continue;
                    }
                    if (hl) {
                        cur = n.left;
                        continue;
                    }
                    if (hr) {
                        cur = n.right;
                        continue;
                    }
                }
                if (instance != kInvalid) {
                    if (popLive(stack, blasBase, best.t, cur)) continue;
                    instance = kInvalid;
                    nodes = s.tlas.nodes;
                    r = makeRayInv(ray);
                }
                if (!popLive(stack, base, best.t, cur)) break;
            }
        }
    }
    if (s.flatCount != 0) {
        if (s.flatWide) {
            RayInv r = makeRayInv(ray);
            queryWide<Full>(s, r, s.flat, s.flatWide, s.flatGeometry, s.flatTriangles, kHitFlat, kRankFlat, any, best,
                stack);
        } else if (s.flat.primCount != 0) {
            RayInv r = makeRayInv(ray);
            queryBvh<Full>(s, r, s.flat, s.flatGeometry, s.flatTriangles, kInvalid, kHitFlat, kRankFlat, kInvalid, any,
                best, stack);
        } else {
            queryBrute<Full>(s, ray, s.flatTriangles, s.flatCount, kInvalid, kHitFlat, kRankFlat, kInvalid, any, best);
        }
    }
    return best;
}

template <bool Full = true, class Stack>
WPT_HD HitRecord traceClosest(const DevScene& s, const Ray& ray, Stack& stack) {
    return traceQuery<Full, Query::kClosest>(s, ray, kNoHit, false, stack);
}

// Shadow query: any blocker with t < maxDistance.
template <bool Full = true, class Stack>
WPT_HD bool traceOccluded(const DevScene& s, const Ray& ray, float maxDistance, Stack& stack) {
    return traceQuery<Full, Query::kAny>(s, ray, maxDistance, true, stack).kind != kHitNone;
}

}  // namespace wpt
