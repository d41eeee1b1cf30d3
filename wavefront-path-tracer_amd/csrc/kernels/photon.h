// Photon map for caustics: the photons that reached a diffuse surface after
// at least one specular bounce, binned in a uniform grid, and the gather
// that turns them into irradiance at a shading point.
//
// Reference: src/renderer/photon_mapping.cu:4-163 (emitPhotons,
// tracePhotons, buildPhotonGrid, gatherCaustics), render_driver.cu:180-234
// (grid bounds and resolution, photon seed).
//
// The reference bins photons into per-cell linked lists with atomic
// exchanges, so the order in which a cell's photons are visited (and with
// it the summation order of the gather) varies from run to run.  Here the
// photons are sorted by (cell, photon index) and each cell is a contiguous
// range visited in photon order, which makes the gather deterministic.
#pragma once

#include "../scene/device_scene.h"

namespace wpt {

// A stored photon: where it landed, the direction back toward where it came
// from, and its flux.
struct Photon {
    Vec3 position;
    Vec3 direction;
    Vec3 flux;
    uint32_t valid;
    uint32_t index;  // photon number, the secondary sort key
    uint32_t pad;
};

// Grid over [minimum, maximum] with resolution^3 cells; photons sorted by
// cell, cell c holding photons[cellStart[c] .. cellStart[c + 1]) (cellStart
// has resolution^3 + 1 entries).  occupiedMin/Max bound the cell
// coordinates of every non-empty cell (min > max: no photon stored); cells
// outside that box are empty.
struct PhotonGridView {
    const Photon* photons;
    const uint32_t* cellStart;
    uint32_t resolution;
    float radius;  // gather radius floor
    Vec3 minimum;
    Vec3 maximum;
    uint32_t occupiedMin[3];
    uint32_t occupiedMax[3];
};

// Cell of a position, clamped per axis; `inside` is false outside the box
// (buildPhotonGrid, photon_mapping.cu:81-93).
WPT_HD void photonCell(const PhotonGridView& g, const Vec3& p, int& x, int& y, int& z, bool& inside) {
    Vec3 extent = g.maximum - g.minimum;
    Vec3 offset = p - g.minimum;
    Vec3 n(offset.x / extent.x, offset.y / extent.y, offset.z / extent.z);
    int r = static_cast<int>(g.resolution);
    x = static_cast<int>(n.x * static_cast<float>(g.resolution));
    y = static_cast<int>(n.y * static_cast<float>(g.resolution));
    z = static_cast<int>(n.z * static_cast<float>(g.resolution));
    x = x < 0 ? 0 : (x > r - 1 ? r - 1 : x);
    y = y < 0 ? 0 : (y > r - 1 ? r - 1 : y);
    z = z < 0 ? 0 : (z > r - 1 ? r - 1 : z);
    inside = !(n.x < 0.0f || n.y < 0.0f || n.z < 0.0f || n.x > 1.0f || n.y > 1.0f || n.z > 1.0f);
}

// Irradiance estimate at a point (gatherCaustics, photon_mapping.cu:96-163):
// the up to 64 nearest photons within max(4 radius, 0.1) that arrive from
// the normal's side, cone-filtered over the farthest kept photon's distance
// (at least `radius`; the whole search radius below 16 photons).
WPT_HD Vec3 gatherCaustics(const PhotonGridView& g, const Vec3& position, const Vec3& normal) {
    constexpr uint32_t kNearest = 64;
    constexpr uint32_t kMinimum = 16;
    if (g.occupiedMin[0] > g.occupiedMax[0]) return Vec3(0.0f, 0.0f, 0.0f);
    int cx, cy, cz;
    bool inside;
    photonCell(g, position, cx, cy, cz, inside);
    Vec3 extent = g.maximum - g.minimum;
    float searchRadius = fmaxf(g.radius * 4.0f, 0.1f);
    int cellRadius = static_cast<int>(ceilf(searchRadius * static_cast<float>(g.resolution) / fmaxf(extent.x, 0.0001f)));
    cellRadius = cellRadius > 1 ? cellRadius : 1;
    float searchRadiusSquared = searchRadius * searchRadius;
    float distances[kNearest];
    uint32_t indices[kNearest];
    uint32_t count = 0, farthest = 0;
    float farthestSquared = 0.0f;
    // Once the slots are full, the largest distance of each group of eight
    // slots and its first slot (in registers: constant indices only).  The
    // farthest slot is then the first group maximum's slot, the slot a
    // rescan of all 64 for the first strictly larger distance finds, and a
    // replacement reads only its own group.
    constexpr uint32_t kGroups = 8, kGroupSize = kNearest / kGroups;
    float groupMax[kGroups];
    uint32_t groupArg[kGroups];
    auto scanGroup = [&](uint32_t group, float& m, uint32_t& a) {
        a = group * kGroupSize;
        m = distances[a];
        for (uint32_t j = 1; j < kGroupSize; ++j) {
            float x = distances[group * kGroupSize + j];
            if (x > m) {
                m = x;
                a = group * kGroupSize + j;
            }
        }
    };
    auto firstMaximum = [&] {
        farthestSquared = groupMax[0];
        farthest = groupArg[0];
#pragma unroll
        for (uint32_t k = 1; k < kGroups; ++k)
            if (groupMax[k] > farthestSquared) {
                farthestSquared = groupMax[k];
                farthest = groupArg[k];
            }
    };
    // The search cube clipped to the occupied cells: the cells left out are
    // empty, so the photons visited, and their order, are unchanged.
    int lo[3], hi[3];
    const int c[3] = {cx, cy, cz};
    for (int a = 0; a < 3; ++a) {
        int omin = static_cast<int>(g.occupiedMin[a]), omax = static_cast<int>(g.occupiedMax[a]);
        lo[a] = c[a] - cellRadius > omin ? c[a] - cellRadius : omin;
        hi[a] = c[a] + cellRadius < omax ? c[a] + cellRadius : omax;
    }
    // Cells whose box lies beyond the search radius hold no photon that
    // passes the distance test below, so skipping them leaves the photons
    // kept, and their order, unchanged.  The boxes are widened by 1% of a
    // cell plus a few ulps of the grid's coordinates, and the radius by a
    // relative 1e-4, so that rounding never skips a cell it should visit.
    const float res = static_cast<float>(g.resolution);
    const float mins[3] = {g.minimum.x, g.minimum.y, g.minimum.z};
    const float sizes[3] = {extent.x / res, extent.y / res, extent.z / res};
    const float query[3] = {position.x, position.y, position.z};
    float margin[3];
    for (int a = 0; a < 3; ++a) margin[a] = 0.01f * sizes[a] + 4e-6f * (fabsf(mins[a]) + fabsf(sizes[a] * res));
    const float cullSquared = searchRadiusSquared * 1.0001f;
    auto gap = [&](int a, int i) {
        float lower = mins[a] + sizes[a] * static_cast<float>(i) - margin[a];
        float upper = mins[a] + sizes[a] * static_cast<float>(i + 1) + margin[a];
        return query[a] < lower ? lower - query[a] : (query[a] > upper ? query[a] - upper : 0.0f);
    };
    // The cells of one row (z, y) that the test keeps form one run of x (the
    // gap grows away from the query's cell), and their photons one run of
    // the sorted order: a row is a single loop over photons.
    for (int z = lo[2]; z <= hi[2]; ++z) {
        float gz = gap(2, z);
        for (int y = lo[1]; y <= hi[1]; ++y) {
            float gy = gap(1, y);
            float gyz = gy * gy + gz * gz;
            if (gyz > cullSquared) continue;
            int first = hi[0] + 1, last = lo[0] - 1;
            for (int x = lo[0]; x <= hi[0]; ++x) {
                float gx = gap(0, x);
                if (gx * gx + gyz > cullSquared) continue;
                first = x < first ? x : first;
                last = x;
            }
            if (first > last) continue;
            uint32_t row = (static_cast<uint32_t>(z) * g.resolution + static_cast<uint32_t>(y)) * g.resolution;
            uint32_t begin = g.cellStart[row + static_cast<uint32_t>(first)];
            uint32_t end = g.cellStart[row + static_cast<uint32_t>(last) + 1u];
            // Photons in groups of kBatch: the group's loads are issued
            // together (indices past the run repeat its last photon and are
            // not visited), then the photons are visited in order.
            constexpr uint32_t kBatch = 8;
            for (uint32_t i = begin; i < end; i += kBatch) {
                Vec3 at[kBatch], from[kBatch];
#pragma unroll
                for (uint32_t k = 0; k < kBatch; ++k) {
                    const Photon& photon = g.photons[i + k < end ? i + k : end - 1u];
                    at[k] = photon.position;
                    from[k] = photon.direction;
                }
#pragma unroll
                for (uint32_t k = 0; k < kBatch; ++k) {
                    if (i + k >= end) break;
                    Vec3 d = at[k] - position;
                    float distanceSquared = dot(d, d);
                    float cosine = dot(normal, from[k]);
                    if (distanceSquared > searchRadiusSquared || cosine <= 0.0f) continue;
                    if (count < kNearest) {
                        distances[count] = distanceSquared;
                        indices[count] = i + k;
                        if (distanceSquared > farthestSquared) {
                            farthestSquared = distanceSquared;
                            farthest = count;
                        }
                        if (++count == kNearest) {
#pragma unroll
                            for (uint32_t q = 0; q < kGroups; ++q) scanGroup(q, groupMax[q], groupArg[q]);
                        }
                    } else if (distanceSquared < farthestSquared) {
                        distances[farthest] = distanceSquared;
                        indices[farthest] = i + k;
                        uint32_t group = farthest / kGroupSize;
                        float m;
                        uint32_t a;
                        scanGroup(group, m, a);
#pragma unroll
                        for (uint32_t q = 0; q < kGroups; ++q)
                            if (q == group) {
                                groupMax[q] = m;
                                groupArg[q] = a;
                            }
                        firstMaximum();
                    }
                }
            }
        }
    }
    if (count == 0) return Vec3(0.0f, 0.0f, 0.0f);
    float gatherSquared = searchRadiusSquared;
    if (count >= kMinimum) {
        gatherSquared = 0.0f;
        for (uint32_t k = 0; k < count; ++k) gatherSquared = fmaxf(gatherSquared, distances[k]);
        gatherSquared = fmaxf(gatherSquared, g.radius * g.radius);
    }
    float gatherRadius = sqrtf(gatherSquared);
    Vec3 result(0.0f, 0.0f, 0.0f);
    for (uint32_t k = 0; k < count; ++k) {
        const Photon& photon = g.photons[indices[k]];
        float distance = sqrtf(distances[k]);
        float cone = fmaxf(0.0f, 1.0f - distance / gatherRadius);
        result += photon.flux * (cone * dot(normal, photon.direction));
    }
    return result * (3.0f / (3.14159265359f * gatherSquared));
}

}  // namespace wpt
