// C ABI of libwpt.so: scene upload, renderer resources, hipGraph capture,
// persistent launches, the task seam and traversal/compaction entry points.
//
// Reference seams: uploadScene (src/scene/scene.cu:336-420), RenderDriver
// (src/renderer/render_driver.cu:105-482), RenderTaskRenderer::render
// (src/renderer/render_task.cu:26-114), render session
// (src/renderer/render_session.cu:16-48).  Errors are returned as status
// codes with a thread-local message instead of exiting the process.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../../include/wpt.h"
#include "last_error.h"
#include "../kernels/bvh_build.h"
#include "../kernels/diff_kernels.h"
#include "../kernels/wavefront.h"
#include "../scene/scene_build.h"

using namespace wpt;

namespace {

thread_local std::string gLastError;

int fail(int code, const std::string& message) {
    gLastError = message;
    return code;
}

}  // namespace

// Host-layer entry points report errors through the same message.
int wpt::setLastError(int code, const std::string& message) { return fail(code, message); }

namespace {

int hipFail(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? WPT_ERROR_OUT_OF_MEMORY : WPT_ERROR_DEVICE,
        std::string(what) + ": " + hipGetErrorString(e));
}

#define WPT_HIP(call)                                  \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return hipFail(e_, #call); \
    } while (0)

struct DeviceArena {
    std::vector<void*> blocks;
    ~DeviceArena() {
        for (void* p : blocks) (void)hipFree(p);
    }
    template <class T>
    hipError_t upload(const std::vector<T>& v, const T** out) {
        *out = nullptr;
        if (v.empty()) return hipSuccess;
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, sizeof(T) * v.size());
        if (e != hipSuccess) return e;
        blocks.push_back(p);
        e = hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
        *out = static_cast<const T*>(p);
        return e;
    }
    hipError_t alloc(size_t bytes, void** out) {
        *out = nullptr;
        if (bytes == 0) return hipSuccess;
        hipError_t e = hipMalloc(out, bytes);
        if (e == hipSuccess) blocks.push_back(*out);
        return e;
    }
};

Bvh deviceBvhView(const HostBvh& h, const BvhNode* nodes) {
    Bvh b{};
    b.nodes = nodes;
    b.nodeCount = static_cast<uint32_t>(h.nodes.size());
    b.primCount = h.primCount();
    b.root = h.primCount() > 1 ? 0u : (kLeafBit | 0u);
    for (int a = 0; a < 3; ++a) {
        b.rootMin[a] = h.rootBounds.mn[a];
        b.rootMax[a] = h.rootBounds.mx[a];
    }
    return b;
}

}  // namespace

// A BLAS built on the device: where its nodes and slot order live.
struct DeviceBlas {
    const BvhNode* nodes = nullptr;
    const uint32_t* order = nullptr;
    uint32_t primCount = 0;
    DeviceBvhBuild build{};
};

struct wpt_scene {
    HostSceneData host;
    std::vector<DeviceBlas> deviceBlas;  // per mesh; primCount 0: host-built or none
    bool onDevice = false;
    DevScene dev{};
    DeviceArena arena;
    uint2* abiSpill = nullptr;  // traversal stack spill of the C ABI trace calls
};

struct wpt_renderer {
    wpt_scene* scene = nullptr;
    wpt_render_config config{};
    uint64_t fingerprint = 0;
    uint32_t capacity = 0;
    std::vector<RenderTile> tiles;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t start = nullptr, end = nullptr;
    LaunchContext ctx{};
    DeviceArena arena;
    RenderTile uploaded{};  // tile and pass size last copied to the device
    bool tileValid = false;
    uint32_t tileCapacity = 0;  // pixels of the largest tile
    double lastMs = 0.0;
    uint64_t lastSamples = 0;
    double photonMs = 0.0;       // caustic photon map build
    uint32_t photonsStored = 0;  // photons kept in the grid
    // Regenerate mode: the batch consumer over the renderer's tiles.
    BatchContext regen{};
    uint32_t regenBlocks = 0;
    std::vector<BatchTask> regenTasks;
    ~wpt_renderer() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (start) (void)hipEventDestroy(start);
        if (end) (void)hipEventDestroy(end);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

__global__ void kSetTile(RenderTile* dst, RenderTile tile) { *dst = tile; }
__global__ void kSetSample(WavefrontCounters* c, uint32_t sample) { c->sampleIndex = sample; }
__global__ void kClearStats(WavefrontCounters* c) {
    c->closestQueries = 0;
    c->shadowQueries = 0;
    for (int i = 0; i < 64; ++i) c->queueCounts[i] = 0;
}
__global__ void kClearRect(float* fb, uint32_t width, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t n = static_cast<size_t>(w) * h * 3;
    if (i >= n) return;
    uint32_t px = static_cast<uint32_t>(i / 3);
    uint32_t c = static_cast<uint32_t>(i % 3);
    uint32_t y = y0 + px / w;
    uint32_t x = x0 + px % w;
    fb[(static_cast<size_t>(y) * width + x) * 3 + c] = 0.0f;
}

int validateConfig(const wpt_render_config& c) {
    if (!std::isfinite(c.indirect_clamp) || c.indirect_clamp < 0.0f)
        return fail(WPT_ERROR_INVALID_ARGUMENT, "Indirect radiance clamp must be finite and nonnegative");
    if (c.width == 0 || c.height == 0) return fail(WPT_ERROR_INVALID_ARGUMENT, "Render dimensions must be greater than zero");
    if (c.samples_per_pixel == 0 || c.max_depth == 0)
        return fail(WPT_ERROR_INVALID_ARGUMENT, "Samples per pixel and maximum depth must be greater than zero");
    if (c.block_size == 0 || c.block_size > 512 || c.block_size % 64 != 0)
        return fail(WPT_ERROR_INVALID_ARGUMENT, "Block size must be a multiple of 64 between 64 and 512");
    if (c.tiled && (c.tile_width == 0 || c.tile_height == 0))
        return fail(WPT_ERROR_INVALID_ARGUMENT, "Tile dimensions must be greater than zero");
    if (c.rng_strategy > WPT_RNG_PCG32) return fail(WPT_ERROR_INVALID_ARGUMENT, "Unknown RNG strategy");
    if (c.mode > WPT_MODE_REGENERATE) return fail(WPT_ERROR_INVALID_ARGUMENT, "Unknown scheduler mode");
    if (c.max_depth > 0x7fffffffu) return fail(WPT_ERROR_INVALID_ARGUMENT, "Maximum depth out of range");
    uint64_t pixels = static_cast<uint64_t>(c.width) * c.height;
    uint64_t cap = c.tiled ? static_cast<uint64_t>(c.tile_width) * c.tile_height : pixels;
    if (pixels > std::numeric_limits<uint32_t>::max() || cap >= (1ull << 30))
        return fail(WPT_ERROR_INVALID_ARGUMENT, "Render dimensions exceed the supported queue index range");
    // Staged and persistent paths keep their sample slot above the pixel index.
    if (c.mode != WPT_MODE_REGENERATE && pixels > kPixelMask + 1ull)
        return fail(WPT_ERROR_INVALID_ARGUMENT, "Render dimensions exceed 2^26 pixels for the wavefront schedulers");
    return WPT_OK;
}

std::vector<RenderTile> makeTiles(uint32_t width, uint32_t height, uint32_t tw, uint32_t th) {
    std::vector<RenderTile> tiles;
    for (uint32_t y = 0; y < height; y += th)
        for (uint32_t x = 0; x < width; x += tw)
            tiles.push_back(RenderTile{x, y, std::min(tw, width - x), std::min(th, height - y), 0u, 1u});
    return tiles;
}

int ensureGraph(wpt_renderer* r) {
    if (r->config.mode != WPT_MODE_GRAPH || r->exec) return WPT_OK;
    WPT_HIP(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueueStagedSample(r->ctx, r->stream);
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(r->stream, &graph);
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return hipFail(e, "enqueueStagedSample");
    }
    if (e2 != hipSuccess) return hipFail(e2, "hipStreamEndCapture");
    r->graph = graph;
    WPT_HIP(hipGraphInstantiate(&r->exec, r->graph, nullptr, nullptr, 0));
    return WPT_OK;
}

// One pass (tile.samples samples) of a tile by the staged or persistent
// scheduler; the tile descriptor is copied to the device when it changes.
int launchTilePass(wpt_renderer* r, const RenderTile& tile) {
    const RenderTile& u = r->uploaded;
    if (!r->tileValid || u.x != tile.x || u.y != tile.y || u.width != tile.width || u.height != tile.height ||
        u.advanceSample != tile.advanceSample || u.samples != tile.samples) {
        hipLaunchKernelGGL(kSetTile, dim3(1), dim3(1), 0, r->stream, const_cast<RenderTile*>(r->ctx.tile), tile);
        WPT_HIP(hipGetLastError());
        r->uploaded = tile;
        r->tileValid = true;
    }
    if (r->config.mode == WPT_MODE_GRAPH) {
        WPT_HIP(hipGraphLaunch(r->exec, r->stream));
    } else {
        WPT_HIP(enqueuePersistentSample(r->ctx, r->stream));
    }
    return WPT_OK;
}

// Regenerate mode: one batch task per tile over samples [first, first +
// count), accumulating into the framebuffer (or overwriting, for a task).
BatchTask regenTask(const wpt_renderer* r, const RenderTile& tile, uint32_t first, uint32_t count, bool accumulate) {
    const LaunchContext& c = r->ctx;
    BatchTask t{};
    t.camera = c.camera;
    t.shade = c.shade;
    t.tile = tile;
    t.scene = 0;
    t.width = c.width;
    t.height = c.height;
    t.sampleStart = first;
    t.sampleCount = count;
    t.pixelCount = count ? tile.width * tile.height : 0u;
    t.outStride = c.width;
    t.out = 3ull * (static_cast<unsigned long long>(tile.y) * c.width + tile.x);
    t.accumulate = accumulate ? 1u : 0u;
    return t;
}

int regenRun(wpt_renderer* r, const std::vector<BatchTask>& tasks) {
    BatchContext ctx = r->regen;
    ctx.taskCount = static_cast<uint32_t>(tasks.size());
    r->regenTasks = tasks;  // the copy below may read the host vector after returning
    WPT_HIP(hipMemcpyAsync(const_cast<BatchTask*>(ctx.tasks), r->regenTasks.data(), sizeof(BatchTask) * tasks.size(),
        hipMemcpyHostToDevice, r->stream));
    WPT_HIP(enqueueBatch(ctx, r->regenBlocks, r->config.block_size, r->stream));
    return WPT_OK;
}

}  // namespace

extern "C" {

WPT_API int wpt_abi_version(void) { return WPT_ABI_VERSION; }

WPT_API const char* wpt_last_error(void) { return gLastError.c_str(); }

WPT_API int wpt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

static int createScene(const wpt_scene_desc* desc, wpt_scene** out, bool upload) {
    if (!desc || !out) return fail(WPT_ERROR_INVALID_ARGUMENT, "null scene description or output");
    *out = nullptr;
    wpt_scene* s = new (std::nothrow) wpt_scene();
    if (!s) return fail(WPT_ERROR_OUT_OF_MEMORY, "host allocation failed");
    std::string err = buildHostScene(*desc, s->host);
    if (!err.empty()) {
        delete s;
        return fail(WPT_ERROR_INVALID_ARGUMENT, err);
    }
    if (s->host.maxStackDepth > kMaxStackDepth) {
        delete s;
        return fail(WPT_ERROR_UNSUPPORTED, "acceleration structure deeper than the traversal stack");
    }
    if (!upload) {
        for (const HostMesh& m : s->host.meshes)
            if (m.buildOnDevice) {
                delete s;
                return fail(WPT_ERROR_UNSUPPORTED, "the device BVH builder needs a device scene");
            }
    }
    s->deviceBlas.resize(s->host.meshes.size());
    if (upload) {
        int status = WPT_OK;
        auto check = [&](hipError_t e, const char* what) {
            if (e != hipSuccess && status == WPT_OK) status = hipFail(e, what);
        };
        HostSceneData& h = s->host;
        DevScene& d = s->dev;
        DeviceArena& a = s->arena;
        const float4* f4 = nullptr;
        check(a.upload(h.spheres, &d.spheres), "spheres");
        check(a.upload(h.sphereMaterial, &d.sphereMaterial), "sphere materials");
        d.sphereCount = static_cast<uint32_t>(h.spheres.size());
        check(a.upload(h.directSpheres, &d.directSpheres), "direct spheres");
        d.directSphereCount = static_cast<uint32_t>(h.directSpheres.size());
        check(a.upload(h.directInstances, &d.directInstances), "direct instances");
        d.directInstanceCount = static_cast<uint32_t>(h.directInstances.size());
        check(a.upload(h.staticTriangles, &d.staticTriangles), "static triangles");
        d.staticCount = static_cast<uint32_t>(h.staticTriangles.size());
        const BvhNode* nodes = nullptr;
        check(a.upload(h.tlas.nodes, &nodes), "tlas nodes");
        d.tlas = deviceBvhView(h.tlas, nodes);
        check(a.upload(h.tlasLeavesSlot, &d.tlasLeaves), "tlas leaves");
        check(a.upload(h.instances, &d.instances), "instances");
        d.instanceCount = static_cast<uint32_t>(h.instances.size());
        std::vector<Blas> blases(h.meshes.size());
        double deviceBuildMs = 0.0;
        for (size_t m = 0; m < h.meshes.size(); ++m) {
            const HostMesh& mesh = h.meshes[m];
            Blas& b = blases[m];
            check(a.upload(mesh.triangles, &b.triangles), "mesh triangles");
            b.triangleCount = static_cast<uint32_t>(mesh.triangles.size());
            if (mesh.buildOnDevice && status == WPT_OK) {
                // Device construction straight into the scene's arena.
                uint32_t n = b.triangleCount;
                void *nodes = nullptr, *geometry = nullptr, *order = nullptr;
                check(a.alloc(sizeof(BvhNode) * (n - 1), &nodes), "blas nodes");
                check(a.alloc(sizeof(float4) * 3 * static_cast<size_t>(n), &geometry), "blas geometry");
                check(a.alloc(sizeof(uint32_t) * n, &order), "blas order");
                DeviceBlas& rec = s->deviceBlas[m];
                if (status == WPT_OK)
                    check(buildTriangleBvh(b.triangles, n, h.bvhSearchRadius, static_cast<BvhNode*>(nodes),
                              static_cast<float4*>(geometry), static_cast<uint32_t*>(order), nullptr, &rec.build),
                        "device BVH build");
                rec.nodes = static_cast<const BvhNode*>(nodes);
                rec.order = static_cast<const uint32_t*>(order);
                rec.primCount = n;
                b.bvh = Bvh{};
                b.bvh.nodes = rec.nodes;
                b.bvh.nodeCount = n - 1;
                b.bvh.primCount = n;
                b.bvh.root = n > 1 ? 0u : (kLeafBit | 0u);
                for (int k = 0; k < 3; ++k) {
                    b.bvh.rootMin[k] = rec.build.rootMin[k];
                    b.bvh.rootMax[k] = rec.build.rootMax[k];
                }
                b.geometry = static_cast<const float4*>(geometry);
                h.maxBlasDepth = std::max(h.maxBlasDepth, rec.build.depth);
                deviceBuildMs += rec.build.milliseconds;
                continue;
            }
            const BvhNode* bn = nullptr;
            check(a.upload(mesh.bvh.nodes, &bn), "blas nodes");
            b.bvh = deviceBvhView(mesh.bvh, bn);
            check(a.upload(mesh.geometry, &f4), "blas geometry");
            b.geometry = f4;
        }
        if (h.bvhBuilder == WPT_BVH_BUILDER_DEVICE_PLOC) {
            h.blasBuildMilliseconds = deviceBuildMs;
            h.maxStackDepth = stackDepth(h);
            if (status == WPT_OK && h.maxStackDepth > kMaxStackDepth)
                status = fail(WPT_ERROR_UNSUPPORTED, "device-built BVH deeper than the traversal stack");
        }
        check(a.upload(blases, &d.blases), "blases");
        d.blasCount = static_cast<uint32_t>(blases.size());
        check(a.upload(h.flat.nodes, &nodes), "flat nodes");
        d.flat = deviceBvhView(h.flat, nodes);
        check(a.upload(h.flatWide, &d.flatWide), "flat wide nodes");
        check(a.upload(h.flatGeometry, &f4), "flat geometry");
        d.flatGeometry = f4;
        check(a.upload(h.flatTriangles, &d.flatTriangles), "flat triangles");
        d.flatCount = static_cast<uint32_t>(h.flatTriangles.size());
        check(a.upload(h.materials, &d.materials), "materials");
        d.materialCount = static_cast<uint32_t>(h.materials.size());
        check(a.upload(h.materialCutout, &d.materialCutout), "material cutout");
        check(a.upload(h.lights, &d.lights), "lights");
        check(a.upload(h.lightAlias, &d.lightAlias), "light alias");
        d.lightCount = static_cast<uint32_t>(h.lights.size());
        d.blackBackground = h.blackBackground;
        d.ignoreDirectLightOcclusion = h.ignoreDirectLightOcclusion;
        d.environment = EnvironmentMap{};
        if (!h.envTexels.empty()) {
            check(a.upload(h.envTexels, &d.environment.texels), "environment texels");
            check(a.upload(h.envProbabilities, &d.environment.probabilities), "environment probabilities");
            check(a.upload(h.envAlias, &d.environment.alias), "environment alias");
            d.environment.width = h.envWidth;
            d.environment.height = h.envHeight;
            d.environment.intensity = h.envIntensity;
            d.environment.rotation = h.envRotation;
        }
        check(a.upload(h.punctual, &d.punctual), "punctual lights");
        d.punctualCount = static_cast<uint32_t>(h.punctual.size());
        d.directionalCount = h.directionalCount;
        std::vector<DevTexture> textures(h.textureTexels.size());
        for (size_t t = 0; t < textures.size(); ++t) {
            check(a.upload(h.textureTexels[t], &textures[t].texels), "texture texels");
            textures[t].width = h.textureWidth[t];
            textures[t].height = h.textureHeight[t];
        }
        check(a.upload(textures, &d.textures), "textures");
        d.textureCount = static_cast<uint32_t>(textures.size());
        check(a.upload(h.materialExt, &d.materialExt), "material texture bindings");
        check(a.upload(h.triExtra, &d.triExtra), "triangle texture data");
        d.useNormalMaps = h.useNormalMaps;
        d.normalMapMinimumCosine = h.normalMapMinimumCosine;
        d.fullFeatures = h.fullFeatures() ? 1u : 0u;
        if (status != WPT_OK) {
            delete s;
            return status;
        }
        s->onDevice = true;
    }
    *out = s;
    return WPT_OK;
}

WPT_API int wpt_scene_create(const wpt_scene_desc* desc, wpt_scene** out) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(WPT_ERROR_DEVICE, "no gfx device visible: the wavefront renderer needs a GPU");
    return createScene(desc, out, true);
}

WPT_API int wpt_scene_create_host(const wpt_scene_desc* desc, wpt_scene** out) { return createScene(desc, out, false); }

WPT_API int wpt_scene_destroy(wpt_scene* scene) {
    delete scene;
    return WPT_OK;
}

WPT_API int wpt_scene_info_get(const wpt_scene* s, wpt_scene_info* info) {
    if (!s || !info) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    const HostSceneData& h = s->host;
    info->sphere_count = static_cast<uint32_t>(h.spheres.size());
    info->direct_sphere_count = static_cast<uint32_t>(h.directSpheres.size());
    info->instance_count = static_cast<uint32_t>(h.instances.size());
    info->direct_instance_count = static_cast<uint32_t>(h.directInstances.size());
    info->tlas_leaf_count = static_cast<uint32_t>(h.tlasLeavesOriginal.size());
    info->tlas_node_count = static_cast<uint32_t>(exportBvh2(h.tlas).size());
    uint32_t blases = 0;
    for (const HostMesh& m : h.meshes) blases += m.bvh.primCount() != 0 || m.buildOnDevice;
    info->blas_count = blases;
    info->bvh_builder = h.bvhBuilder;
    info->max_blas_depth = h.maxBlasDepth;
    info->blas_build_milliseconds = h.blasBuildMilliseconds;
    info->light_count = static_cast<uint32_t>(h.lights.size());
    info->static_triangle_count = static_cast<uint32_t>(h.staticTriangles.size());
    info->flat_triangle_count = static_cast<uint32_t>(h.flatTriangles.size());
    info->flat_node_count = h.flat.primCount() ? 2 * h.flat.primCount() - 1 : 0;
    info->flat_accelerator = h.flatAccelerator;
    info->flat_leaf_slots = h.flat.primCount();
    info->flat_wide_node_count = static_cast<uint32_t>(h.flatWide.size());
    info->max_stack_depth = h.maxStackDepth;
    info->material_count = static_cast<uint32_t>(h.materials.size());
    return WPT_OK;
}

WPT_API int wpt_scene_get_list(const wpt_scene* s, uint32_t which, uint32_t* out, uint32_t capacity, uint32_t* count) {
    if (!s || !count) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    const std::vector<uint32_t>* list = nullptr;
    if (which == 0) list = &s->host.directSpheres;
    else if (which == 1) list = &s->host.directInstances;
    else if (which == 2) list = &s->host.tlasLeavesOriginal;
    else return fail(WPT_ERROR_INVALID_ARGUMENT, "unknown list");
    *count = static_cast<uint32_t>(list->size());
    if (out) std::memcpy(out, list->data(), sizeof(uint32_t) * std::min<size_t>(capacity, list->size()));
    return WPT_OK;
}

WPT_API int wpt_scene_export_bvh2(const wpt_scene* s, uint32_t bvh, wpt_bvh2_node* out, uint32_t capacity,
    uint32_t* nodeCount) {
    if (!s || !nodeCount) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    const HostBvh* h = nullptr;
    HostBvh fetched;
    if (bvh == 0) h = &s->host.tlas;
    else if (bvh == 1) h = &s->host.flat;
    else if (bvh - 2 < s->host.meshes.size()) h = &s->host.meshes[bvh - 2].bvh;
    else return fail(WPT_ERROR_INVALID_ARGUMENT, "unknown BVH");
    if (bvh >= 2 && s->deviceBlas[bvh - 2].primCount) {
        // A device-built BLAS: nodes and slot order back from the device,
        // leaf bounds recomputed from the triangles as the builder did.
        const DeviceBlas& d = s->deviceBlas[bvh - 2];
        const HostMesh& mesh = s->host.meshes[bvh - 2];
        fetched.nodes.resize(d.primCount - 1);
        fetched.order.resize(d.primCount);
        if (d.primCount > 1)
            WPT_HIP(hipMemcpy(fetched.nodes.data(), d.nodes, sizeof(BvhNode) * (d.primCount - 1),
                hipMemcpyDeviceToHost));
        WPT_HIP(hipMemcpy(fetched.order.data(), d.order, sizeof(uint32_t) * d.primCount, hipMemcpyDeviceToHost));
        for (const TriShade& t : mesh.triangles) fetched.primBounds.push_back(triangleBounds(t.v0, t.v1, t.v2));
        fetched.root = d.primCount > 1 ? 0u : (kLeafBit | 0u);
        fetched.depth = d.build.depth;
        for (int k = 0; k < 3; ++k) {
            fetched.rootBounds.mn[k] = d.build.rootMin[k];
            fetched.rootBounds.mx[k] = d.build.rootMax[k];
        }
        h = &fetched;
    }
    std::vector<wpt_bvh2_node> nodes = exportBvh2(*h);
    *nodeCount = static_cast<uint32_t>(nodes.size());
    if (out) std::memcpy(out, nodes.data(), sizeof(wpt_bvh2_node) * std::min<size_t>(capacity, nodes.size()));
    return WPT_OK;
}

WPT_API int wpt_scene_export_environment(const wpt_scene* s, float* probabilities, float* alias2, uint32_t capacity,
    uint32_t* count) {
    if (!s || !count) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    const HostSceneData& h = s->host;
    *count = static_cast<uint32_t>(h.envProbabilities.size());
    uint32_t n = std::min<uint32_t>(capacity, *count);
    for (uint32_t i = 0; i < n; ++i) {
        if (probabilities) probabilities[i] = h.envProbabilities[i];
        if (alias2) {
            alias2[2 * i] = h.envAlias[i].probability;
            std::memcpy(&alias2[2 * i + 1], &h.envAlias[i].alias, 4);
        }
    }
    return WPT_OK;
}

WPT_API int wpt_scene_export_lights(const wpt_scene* s, float* lights16, float* alias2, uint32_t capacity,
    uint32_t* count) {
    if (!s || !count) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    const HostSceneData& h = s->host;
    *count = static_cast<uint32_t>(h.lights.size());
    uint32_t n = std::min<uint32_t>(capacity, *count);
    for (uint32_t i = 0; i < n; ++i) {
        const TriangleLight& l = h.lights[i];
        if (lights16) {
            float* o = lights16 + 16 * i;
            const Vec3 v[4] = {l.v0, l.v1, l.v2, l.normal};
            for (int k = 0; k < 4; ++k) {
                o[3 * k] = v[k].x;
                o[3 * k + 1] = v[k].y;
                o[3 * k + 2] = v[k].z;
            }
            o[12] = l.area;
            std::memcpy(&o[13], &l.material, 4);
            o[14] = l.selectionPdf;
            o[15] = 0.0f;
        }
        if (alias2) {
            alias2[2 * i] = h.lightAlias[i].probability;
            std::memcpy(&alias2[2 * i + 1], &h.lightAlias[i].alias, 4);
        }
    }
    return WPT_OK;
}

WPT_API int wpt_scene_export_light_indices(const wpt_scene* s, uint32_t* instanceOffset, uint32_t* staticIndex,
    uint32_t* meshIndex) {
    if (!s) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    const HostSceneData& h = s->host;
    if (instanceOffset)
        for (size_t i = 0; i < h.instances.size(); ++i) instanceOffset[i] = h.instances[i].lightOffset;
    if (staticIndex)
        for (size_t i = 0; i < h.staticTriangles.size(); ++i) staticIndex[i] = h.staticTriangles[i].lightIndex;
    if (meshIndex) std::memcpy(meshIndex, h.meshLightIndex.data(), sizeof(uint32_t) * h.meshLightIndex.size());
    return WPT_OK;
}

static int ensureAbiSpill(wpt_scene* s) {
    if (s->abiSpill) return WPT_OK;
    void* p = nullptr;
    size_t bytes = sizeof(uint2) * (kMaxStackDepth - kLdsStackEntries) * static_cast<size_t>(abiTraceThreads());
    WPT_HIP(s->arena.alloc(bytes, &p));
    s->abiSpill = static_cast<uint2*>(p);
    return WPT_OK;
}

WPT_API int wpt_trace_closest(const wpt_scene* cs, const float* rays, uint32_t n, wpt_hit* hits, void* stream) {
    wpt_scene* s = const_cast<wpt_scene*>(cs);
    if (!s || !s->onDevice) return fail(WPT_ERROR_INVALID_ARGUMENT, "scene is not resident on a device");
    if (n && (!rays || !hits)) return fail(WPT_ERROR_INVALID_ARGUMENT, "null ray or hit buffer");
    int status = ensureAbiSpill(s);
    if (status != WPT_OK) return status;
    WPT_HIP(launchTraceClosest(s->dev, rays, n, hits, s->abiSpill, static_cast<hipStream_t>(stream)));
    return WPT_OK;
}

WPT_API int wpt_trace_occluded(const wpt_scene* cs, const float* rays, const float* tmax, uint32_t n, uint8_t* out,
    void* stream) {
    wpt_scene* s = const_cast<wpt_scene*>(cs);
    if (!s || !s->onDevice) return fail(WPT_ERROR_INVALID_ARGUMENT, "scene is not resident on a device");
    if (n && (!rays || !tmax || !out)) return fail(WPT_ERROR_INVALID_ARGUMENT, "null ray, distance or output buffer");
    int status = ensureAbiSpill(s);
    if (status != WPT_OK) return status;
    WPT_HIP(launchTraceOccluded(s->dev, rays, tmax, n, out, s->abiSpill, static_cast<hipStream_t>(stream)));
    return WPT_OK;
}

WPT_API int wpt_compact(const uint32_t* items, const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t capacity,
    uint32_t* count, void* stream) {
    if (!count || (n && (!items || !flags)) || (capacity && !out))
        return fail(WPT_ERROR_INVALID_ARGUMENT, "null compaction buffer");
    if (n >= (1u << 30)) return fail(WPT_ERROR_INVALID_ARGUMENT, "compaction input too large");
    uint32_t words = compactScratchWords(n);
    void* temp = nullptr;
    WPT_HIP(hipMalloc(&temp, sizeof(uint32_t) * words));
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e = launchCompact(items, flags, n, out, capacity, count, static_cast<uint32_t*>(temp), words, st);
    hipError_t e2 = hipStreamSynchronize(st);
    (void)hipFree(temp);
    if (e != hipSuccess) return hipFail(e, "launchCompact");
    if (e2 != hipSuccess) return hipFail(e2, "hipStreamSynchronize");
    return WPT_OK;
}

// ---- caustic photon maps -----------------------------------------------
namespace {

// Caustic settings of a render configuration (config.hpp:18-22).
int validateCaustics(const wpt_render_config& c) {
    if (c.caustic_photons > 0x7fffffffu) return fail(WPT_ERROR_INVALID_ARGUMENT, "too many caustic photons");
    if (!(c.caustic_gather_radius >= 0.0f) || !std::isfinite(c.caustic_gather_radius) ||
        c.caustic_grid_resolution > 1024)
        return fail(WPT_ERROR_INVALID_ARGUMENT, "invalid caustic gather radius or grid resolution");
    if (c.rng_strategy > WPT_RNG_PCG32) return fail(WPT_ERROR_INVALID_ARGUMENT, "unknown RNG strategy");
    return WPT_OK;
}

struct PhotonMapResult {
    const PhotonGridView* view;  // device copy of the gather's view
    double milliseconds;         // trace + grid build
    uint32_t stored;             // photons kept inside the grid
};

// Builds the caustic photon map of `scene` under the caustic fields of `c`
// (RenderDriver::buildPhotonMap, render_driver.cu:180-234) into buffers of
// `arena`, on `stream`; waits for it.  `spill` holds the traversal stacks of
// `spillThreads` threads (at least 256).
int buildPhotonMap(const DevScene& scene, const wpt_render_config& c, DeviceArena& arena, uint2* spill,
    uint32_t spillThreads, hipStream_t stream, hipEvent_t start, hipEvent_t end, PhotonMapResult* out) {
    PhotonSettings ps{c.caustic_photons, c.caustic_max_depth ? c.caustic_max_depth : 8u, 0x13579bdfu, c.rng_strategy,
        c.caustic_spectral ? 1u : 0u};
    PhotonGridView view{};
    view.resolution = c.caustic_grid_resolution ? c.caustic_grid_resolution : 64u;
    view.radius = c.caustic_gather_radius > 0.0f ? c.caustic_gather_radius : 0.06f;
    bool box = c.caustic_grid_min[0] < c.caustic_grid_max[0] && c.caustic_grid_min[1] < c.caustic_grid_max[1] &&
        c.caustic_grid_min[2] < c.caustic_grid_max[2];
    view.minimum = box ? Vec3(c.caustic_grid_min[0], c.caustic_grid_min[1], c.caustic_grid_min[2])
                       : Vec3(-2.5f, -1.0f, -2.5f);
    view.maximum = box ? Vec3(c.caustic_grid_max[0], c.caustic_grid_max[1], c.caustic_grid_max[2])
                       : Vec3(2.5f, 3.0f, 2.5f);
    uint32_t cells = view.resolution * view.resolution * view.resolution;
    PhotonMapBuffers pb{};
    size_t n = c.caustic_photons;
    void* p = nullptr;
    WPT_HIP(arena.alloc(sizeof(Photon) * n, &p));
    pb.photons = static_cast<Photon*>(p);
    WPT_HIP(arena.alloc(sizeof(Photon) * n, &p));
    pb.sorted = static_cast<Photon*>(p);
    WPT_HIP(arena.alloc(sizeof(uint64_t) * n, &p));
    pb.keys = static_cast<uint64_t*>(p);
    WPT_HIP(arena.alloc(sizeof(uint64_t) * n, &p));
    pb.sortedKeys = static_cast<uint64_t*>(p);
    WPT_HIP(photonSortBytes(c.caustic_photons, cells, &pb.tempBytes));
    WPT_HIP(arena.alloc(pb.tempBytes > 0 ? pb.tempBytes : 1, &p));
    pb.temp = p;
    WPT_HIP(arena.alloc(sizeof(uint32_t) * (static_cast<size_t>(cells) + 1), &p));
    pb.cellStart = static_cast<uint32_t*>(p);
    WPT_HIP(arena.alloc(sizeof(uint32_t) * 7, &p));
    pb.stored = static_cast<uint32_t*>(p);
    pb.occupied = pb.stored + 1;
    view.photons = pb.sorted;
    view.cellStart = pb.cellStart;
    WPT_HIP(hipEventRecord(start, stream));
    WPT_HIP(enqueuePhotonMap(scene, ps, pb, view, spill, spillThreads, stream));
    WPT_HIP(hipEventRecord(end, stream));
    WPT_HIP(hipEventSynchronize(end));
    float ms = 0.0f;
    WPT_HIP(hipEventElapsedTime(&ms, start, end));
    uint32_t counters[7];
    WPT_HIP(hipMemcpy(counters, pb.stored, sizeof counters, hipMemcpyDeviceToHost));
    for (int a = 0; a < 3; ++a) {
        view.occupiedMin[a] = counters[1 + a];
        view.occupiedMax[a] = counters[4 + a];
    }
    // The gather's view, with the occupied cell box, once the map exists.
    WPT_HIP(arena.alloc(sizeof(PhotonGridView), &p));
    WPT_HIP(hipMemcpy(p, &view, sizeof(PhotonGridView), hipMemcpyHostToDevice));
    out->view = static_cast<const PhotonGridView*>(p);
    out->milliseconds = ms;
    out->stored = counters[0];
    return WPT_OK;
}

}  // namespace

// ---- compaction without allocation ---------------------------------------
// The scratch of wpt_compact_noalloc (chunk ticket + one look-back flag per
// 256-item chunk) lives in a fixed device array instead of a per-call
// allocation.  Every device holds its own copy, at its own address; host
// code reaches it only through the address hipGetSymbolAddress reports for
// the current device, never through the array's name.
constexpr uint32_t kNoAllocMaxItems = 1u << 24;
// Default paths per pass of the staged and cooperative schedulers (64
// samples of a 1920x1080 frame): about 260 B each of queues, hit and shadow
// records and sample planes, 35 GB of the device's 288 GB.
constexpr uint64_t kDefaultPassPaths = 1ull << 27;
// Staged kernels launch at most this many times the trace kernel's resident
// workgroups (the dispatcher still balances the workgroups).
#ifndef WPT_STAGED_GRID_FACTOR
#define WPT_STAGED_GRID_FACTOR 8
#endif
constexpr uint32_t kStagedGridFactor = WPT_STAGED_GRID_FACTOR;
// Most radiance events per vertex that event shading stores (scenes with
// more directional lights trace their shadow rays inline).
constexpr uint32_t kMaxVertexEvents = 12;

// Radiance events one vertex of `s` can store: the area-light sample, every
// directional light, one positional light, the environment sample, then two
// escape tests (dielectric, glTF delta lobes) or the caustic gather.
uint32_t vertexEventsOf(const DevScene& s) {
    return (s.lightCount ? 1u : 0u) + s.directionalCount + (s.punctualCount > s.directionalCount ? 1u : 0u) +
        (s.environment.texels ? 1u : 0u) + 2u;
}

uint32_t eventWordsPerColumn(uint32_t maxEvents) { return kEventHeaderFields + maxEvents * kEventItemFields; }

// The staged event ray and gather lists: a queue segment's paths append at
// most maxEvents rays and one gather each to the same segment of the lists.
hipError_t allocEventRays(DeviceArena& arena, uint64_t segment, uint32_t maxEvents, EventBuffer* events) {
    uint64_t raySegment = segment * maxEvents;
    if (raySegment >= (1ull << 32) / kSegments) return hipErrorInvalidValue;
    size_t rays = static_cast<size_t>(raySegment) * kSegments, gathers = static_cast<size_t>(segment) * kSegments;
    void* p = nullptr;
    hipError_t e = arena.alloc((sizeof(uint32_t) + sizeof(uint8_t)) * (rays + gathers), &p);
    if (e != hipSuccess) return e;
    events->raySlot = static_cast<uint32_t*>(p);
    events->gatherSlot = events->raySlot + rays;
    events->rayEvent = reinterpret_cast<uint8_t*>(events->gatherSlot + gathers);
    events->gatherEvent = events->rayEvent + rays;
    events->raySegment = static_cast<uint32_t>(raySegment);
    events->gatherSegment = static_cast<uint32_t>(segment);
    return hipSuccess;
}
constexpr uint32_t kNoAllocScratchWords = 1u + kNoAllocMaxItems / 256u;
static __device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
uint32_t* gScratchAddress[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device.
int currentScratch(uint32_t** out) {
    int device = 0;
    WPT_HIP(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) return fail(WPT_ERROR_UNSUPPORTED, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchAddress[device]) {
        void* p = nullptr;
        WPT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        gScratchAddress[device] = static_cast<uint32_t*>(p);
    }
    *out = gScratchAddress[device];
    return WPT_OK;
}

}  // namespace

WPT_API int wpt_compact_noalloc(const uint32_t* items, const uint8_t* flags, uint32_t n, uint32_t* out,
    uint32_t capacity, uint32_t* count, void* stream) {
    if (!count || (n && (!items || !flags)) || (capacity && !out))
        return fail(WPT_ERROR_INVALID_ARGUMENT, "null compaction buffer");
    if (n > kNoAllocMaxItems)
        return fail(WPT_ERROR_INVALID_ARGUMENT, "compaction input exceeds the 16777216 items of the fixed scratch");
    uint32_t words = compactScratchWords(n);
    if (words > kNoAllocScratchWords) return fail(WPT_ERROR_INVALID_ARGUMENT, "compaction scratch too small");
    uint32_t* base = nullptr;
    int status = currentScratch(&base);
    if (status != WPT_OK) return status;
    WPT_HIP(launchCompact(items, flags, n, out, capacity, count, base, words, static_cast<hipStream_t>(stream)));
    return WPT_OK;
}

WPT_API int wpt_renderer_create(wpt_scene* scene, const wpt_camera* camera, const wpt_render_config* config,
    uint64_t fingerprint, wpt_renderer** out) {
    if (!scene || !camera || !config || !out) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (!scene->onDevice) return fail(WPT_ERROR_INVALID_ARGUMENT, "scene is not resident on a device");
    int v = validateConfig(*config);
    if (v != WPT_OK) return v;
    wpt_renderer* r = new (std::nothrow) wpt_renderer();
    if (!r) return fail(WPT_ERROR_OUT_OF_MEMORY, "host allocation failed");
    r->scene = scene;
    r->config = *config;
    r->fingerprint = fingerprint;
    const wpt_render_config& c = r->config;
    uint32_t tilePixels = c.tiled ? c.tile_width * c.tile_height : c.width * c.height;
    // Scenes whose shading adds radiance after an occlusion test (environment,
    // punctual lights, caustics) shade through ordered events in the staged
    // scheduler, at most `vertexEvents` per vertex: the area-light sample, every
    // directional light, one positional light, the environment sample, then
    // two escape tests (dielectric, glTF delta lobes) or the caustic gather.
    const DevScene& sceneView = scene->dev;
    bool inlineShadows = shadesWithInlineShadows(sceneView) || (c.caustic_photons && c.caustic_gather);
    uint32_t vertexEvents = vertexEventsOf(sceneView);
    bool eventShading = c.mode == WPT_MODE_GRAPH && inlineShadows && vertexEvents <= kMaxVertexEvents;
    // Samples per pass of the staged and cooperative schedulers: by default
    // as many as keep a pass within kDefaultPassPaths paths (64 at 1080p;
    // half that with event shading, whose events take about 200 B more per
    // path).
    uint64_t passPaths = eventShading ? kDefaultPassPaths / 2 : kDefaultPassPaths;
    uint32_t slots = c.samples_per_pass ? c.samples_per_pass
                                        : static_cast<uint32_t>(std::max<uint64_t>(1u, passPaths / tilePixels));
    if (c.mode == WPT_MODE_REGENERATE) slots = 1u;
    slots = std::max(1u, std::min(std::min(slots, kMaxSlots), c.samples_per_pixel));
    uint64_t cap64 = static_cast<uint64_t>(tilePixels) * slots;
    if (cap64 >= (1ull << 31)) {
        delete r;
        return fail(WPT_ERROR_INVALID_ARGUMENT, "tile pixels x samples per pass out of range");
    }
    uint32_t cap = static_cast<uint32_t>(cap64);
    r->capacity = cap;
    r->tileCapacity = tilePixels;
    r->tiles = c.tiled ? makeTiles(c.width, c.height, c.tile_width, c.tile_height)
                       : makeTiles(c.width, c.height, c.width, c.height);
    auto cleanup = [&](int status) {
        delete r;
        return status;
    };
    hipError_t e;
#define R_HIP(call)                                            \
    do {                                                       \
        e = (call);                                            \
        if (e != hipSuccess) return cleanup(hipFail(e, #call)); \
    } while (0)
    R_HIP(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    R_HIP(hipEventCreate(&r->start));
    R_HIP(hipEventCreate(&r->end));
    LaunchContext& ctx = r->ctx;
    ctx.scene = scene->dev;
    // Caustic photons gathered at diffuse vertices need the full-feature
    // kernels (render_driver.cu:180-234 builds the map per driver).
    if (c.caustic_photons) {
        int vs = validateCaustics(c);
        if (vs != WPT_OK) return cleanup(vs);
        if (c.caustic_gather) ctx.scene.fullFeatures = 1;
    }
    ctx.camera.position = Vec3(camera->position[0], camera->position[1], camera->position[2]);
    ctx.camera.forward = Vec3(camera->forward[0], camera->forward[1], camera->forward[2]);
    ctx.camera.right = Vec3(camera->right[0], camera->right[1], camera->right[2]);
    ctx.camera.up = Vec3(camera->up[0], camera->up[1], camera->up[2]);
    ctx.camera.verticalFov = camera->vertical_fov;
    ctx.camera.aspectRatio = camera->aspect_ratio;
    ctx.shade.maxDepth = c.max_depth;
    ctx.shade.rrStart = c.rr_start_depth;
    ctx.shade.rngStrategy = c.rng_strategy;
    ctx.shade.clamp = c.indirect_clamp;
    ctx.shade.intersectionDebug = c.intersection_debug;
    ctx.shade.shadingNormalDebug = c.shading_normal_debug;
    ctx.width = c.width;
    ctx.height = c.height;
    ctx.blockSize = c.block_size;
    ctx.gridBlocks = (cap + c.block_size - 1) / c.block_size;  // capped below by the resident grid
    ctx.resolveBlocks = (tilePixels + c.block_size - 1) / c.block_size;
    ctx.profileQueues = 1;
    ctx.capacity = cap;
    ctx.slots = slots;
    // Primary, trace, shade and shadow kernels walk their items in grid
    // stride over gridBlocks workgroups, at most kStagedGridFactor times the
    // trace kernel's resident grid (which bounds the traversal spill area);
    // the persistent kernel runs a resident grid.
    Occupancy occ{};
    R_HIP(kernelOccupancy(c.block_size, ctx.scene.fullFeatures != 0, &occ, vertexEvents <= kMaxVertexEvents));
    int device = 0, cus = 0, coop = 0;
    R_HIP(hipGetDevice(&device));
    R_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    if (occ.trace <= 0 || occ.shade <= 0 || occ.shadow <= 0)
        return cleanup(fail(WPT_ERROR_UNSUPPORTED, "block size cannot launch the wavefront kernels"));
    uint32_t cu = static_cast<uint32_t>(cus);
    ctx.gridBlocks = std::max(1u, std::min(ctx.gridBlocks, kStagedGridFactor * static_cast<uint32_t>(occ.trace) * cu));
    ctx.persistentBlocks = 1;
    if (c.mode == WPT_MODE_PERSISTENT) {
        if (occ.persistent <= 0) return cleanup(fail(WPT_ERROR_UNSUPPORTED, "block size cannot launch the persistent kernel"));
        R_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device));
        if (!coop) return cleanup(fail(WPT_ERROR_UNSUPPORTED, "device lacks cooperative launch"));
        uint32_t resident = static_cast<uint32_t>(occ.persistent) * cu;
        uint32_t blocks = c.persistent_blocks ? std::min(c.persistent_blocks, resident) : resident;
        ctx.persistentBlocks = std::max(1u, std::min(blocks, ctx.gridBlocks));
    }
    // Segment capacity: workgroup b appends to segment b % kSegments, so a
    // segment receives at most (its workgroups) x (items per workgroup); a
    // fresh pass fills segment s with the paths of virtual workgroups
    // s, s + kSegments, ... (primarySlot).
    uint32_t B = c.block_size;
    uint64_t stagedRounds = (static_cast<uint64_t>(cap) + static_cast<uint64_t>(ctx.gridBlocks) * B - 1) /
        (static_cast<uint64_t>(ctx.gridBlocks) * B);
    uint64_t virtualBlocks = (static_cast<uint64_t>(cap) + B - 1) / B;
    uint64_t stagedSegment = std::max(static_cast<uint64_t>((ctx.gridBlocks + kSegments - 1) / kSegments) * stagedRounds,
                                      (virtualBlocks + kSegments - 1) / kSegments) * B;
    uint32_t P = ctx.persistentBlocks;
    uint64_t rounds = (static_cast<uint64_t>(cap) + static_cast<uint64_t>(P) * B - 1) / (static_cast<uint64_t>(P) * B);
    uint64_t persistentSegment = static_cast<uint64_t>((P + kSegments - 1) / kSegments) * rounds * B;
    uint64_t segment = std::max(stagedSegment, persistentSegment);
    uint64_t qcap = segment * kSegments;
    if (qcap >= (1ull << 31)) return cleanup(fail(WPT_ERROR_INVALID_ARGUMENT, "queue capacity out of range"));
    void* p = nullptr;
    bool queues = c.mode != WPT_MODE_REGENERATE;
    for (int q = 0; queues && q < 2; ++q) {
        R_HIP(r->arena.alloc(sizeof(uint32_t) * static_cast<size_t>(kQueueFieldCount) * qcap, &p));
        ctx.queue[q].words = static_cast<uint32_t*>(p);
        ctx.queue[q].capacity = static_cast<uint32_t>(qcap);
        ctx.queue[q].segmentCapacity = static_cast<uint32_t>(segment);
    }
    if (c.mode == WPT_MODE_GRAPH) {
        R_HIP(r->arena.alloc(sizeof(float) * 3 * static_cast<size_t>(qcap) + sizeof(uint32_t) * 4 * static_cast<size_t>(qcap), &p));
        float* f = static_cast<float*>(p);
        ctx.hits.t = f;
        ctx.hits.u = f + qcap;
        ctx.hits.v = f + 2 * static_cast<size_t>(qcap);
        uint32_t* u = reinterpret_cast<uint32_t*>(f + 3 * static_cast<size_t>(qcap));
        ctx.hits.kind = u;
        ctx.hits.instance = u + qcap;
        ctx.hits.prim = u + 2 * static_cast<size_t>(qcap);
        ctx.hits.material = u + 3 * static_cast<size_t>(qcap);
        if (!inlineShadows) {  // deferred area-light shadow rays
            R_HIP(r->arena.alloc(sizeof(uint32_t) * static_cast<size_t>(kShadowFieldCount) * qcap, &p));
            ctx.shadow.words = static_cast<uint32_t*>(p);
            ctx.shadow.capacity = static_cast<uint32_t>(qcap);
            ctx.shadow.segmentCapacity = static_cast<uint32_t>(segment);
        }
        if (eventShading) {
            size_t words = static_cast<size_t>(eventWordsPerColumn(vertexEvents)) * qcap;
            R_HIP(r->arena.alloc(sizeof(uint32_t) * words, &p));
            ctx.events.words = static_cast<uint32_t*>(p);
            ctx.events.capacity = static_cast<uint32_t>(qcap);
            ctx.events.maxEvents = vertexEvents;
            R_HIP(allocEventRays(r->arena, segment, vertexEvents, &ctx.events));
        }
    }
    if (c.mode == WPT_MODE_PERSISTENT && ctx.scene.fullFeatures && vertexEvents <= kMaxVertexEvents) {
        // One event column per thread of the cooperative grid.
        uint32_t threads = ctx.persistentBlocks * c.block_size;
        R_HIP(r->arena.alloc(sizeof(uint32_t) * static_cast<size_t>(eventWordsPerColumn(vertexEvents)) * threads, &p));
        ctx.events.words = static_cast<uint32_t*>(p);
        ctx.events.capacity = threads;
        ctx.events.maxEvents = vertexEvents;
    }
    R_HIP(r->arena.alloc(sizeof(float) * 3 * static_cast<size_t>(c.width) * c.height, &p));
    ctx.framebuffer = static_cast<float*>(p);
    if (c.mode != WPT_MODE_REGENERATE) {
        R_HIP(r->arena.alloc(sizeof(float) * 3 * static_cast<size_t>(c.width) * c.height * slots, &p));
        ctx.slotRadiance = static_cast<float*>(p);
    }
    R_HIP(r->arena.alloc(sizeof(WavefrontCounters), &p));
    ctx.counters = static_cast<WavefrontCounters*>(p);
    R_HIP(hipMemset(p, 0, sizeof(WavefrontCounters)));
    R_HIP(r->arena.alloc(sizeof(RenderTile), &p));
    ctx.tile = static_cast<const RenderTile*>(p);
    R_HIP(hipMemset(ctx.framebuffer, 0, sizeof(float) * 3 * static_cast<size_t>(c.width) * c.height));
    if (c.mode == WPT_MODE_REGENERATE) {
        int perCu = 0;
        R_HIP(batchOccupancy(c.block_size, ctx.scene.fullFeatures != 0, &perCu));
        if (perCu <= 0) return cleanup(fail(WPT_ERROR_UNSUPPORTED, "block size cannot launch the regenerating kernel"));
        uint32_t resident = static_cast<uint32_t>(perCu) * cu;
        r->regenBlocks = c.persistent_blocks ? std::min(c.persistent_blocks, resident) : resident;
        BatchContext& b = r->regen;
        R_HIP(r->arena.alloc(sizeof(DevScene), &p));
        b.scenes = static_cast<const DevScene*>(p);
        R_HIP(hipMemcpy(p, &ctx.scene, sizeof(DevScene), hipMemcpyHostToDevice));
        size_t tasks = r->tiles.size();
        R_HIP(r->arena.alloc(sizeof(BatchTask) * tasks, &p));
        b.tasks = static_cast<const BatchTask*>(p);
        R_HIP(r->arena.alloc(sizeof(BatchTaskState) * tasks, &p));
        b.state = static_cast<BatchTaskState*>(p);
        R_HIP(r->arena.alloc(sizeof(BatchCounters), &p));
        b.counters = static_cast<BatchCounters*>(p);
        R_HIP(hipMemset(p, 0, sizeof(BatchCounters)));
        b.out = ctx.framebuffer;
        b.full = ctx.scene.fullFeatures;
        if (b.full && vertexEvents <= kMaxVertexEvents) {  // one event column per thread of the launch
            uint32_t threads = r->regenBlocks * c.block_size;
            R_HIP(r->arena.alloc(sizeof(uint32_t) * static_cast<size_t>(eventWordsPerColumn(vertexEvents)) * threads,
                &p));
            b.events.words = static_cast<uint32_t*>(p);
            b.events.capacity = threads;
            b.events.maxEvents = vertexEvents;
        }
    }
    // One spill column per thread index of any grid.
    ctx.spillThreads = std::max(std::max(ctx.gridBlocks, ctx.persistentBlocks), r->regenBlocks) * c.block_size;
    R_HIP(r->arena.alloc(sizeof(uint2) * (kMaxStackDepth - kLdsStackEntries) * static_cast<size_t>(ctx.spillThreads), &p));
    ctx.spill = static_cast<uint2*>(p);
    r->regen.spill = ctx.spill;
    r->regen.spillThreads = ctx.spillThreads;
    if (c.caustic_photons) {
        PhotonMapResult map{};
        int ps = buildPhotonMap(ctx.scene, c, r->arena, ctx.spill, ctx.spillThreads, r->stream, r->start, r->end, &map);
        if (ps != WPT_OK) return cleanup(ps);
        r->photonMs = map.milliseconds;
        r->photonsStored = map.stored;
        if (c.caustic_gather) ctx.shade.photons = map.view;
    }
    R_HIP(hipDeviceSynchronize());
#undef R_HIP
    *out = r;
    return WPT_OK;
}

WPT_API int wpt_renderer_destroy(wpt_renderer* r) {
    delete r;
    return WPT_OK;
}

WPT_API int wpt_renderer_reset(wpt_renderer* r) {
    if (!r) return fail(WPT_ERROR_INVALID_ARGUMENT, "null renderer");
    const wpt_render_config& c = r->config;
    size_t n = 3 * static_cast<size_t>(c.width) * c.height;
    hipLaunchKernelGGL(kClearRect, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, r->stream,
        r->ctx.framebuffer, c.width, 0u, 0u, c.width, c.height);
    hipLaunchKernelGGL(kSetSample, dim3(1), dim3(1), 0, r->stream, r->ctx.counters, 0u);
    WPT_HIP(hipGetLastError());
    WPT_HIP(hipStreamSynchronize(r->stream));
    return WPT_OK;
}

WPT_API int wpt_renderer_render_samples(wpt_renderer* r, uint32_t first, uint32_t count) {
    if (!r) return fail(WPT_ERROR_INVALID_ARGUMENT, "null renderer");
    if (r->config.mode == WPT_MODE_REGENERATE) {
        std::vector<BatchTask> tasks;
        for (const RenderTile& t : r->tiles) tasks.push_back(regenTask(r, t, first, count, true));
        WPT_HIP(hipEventRecord(r->start, r->stream));
        int s = regenRun(r, tasks);
        if (s != WPT_OK) return s;
        WPT_HIP(hipEventRecord(r->end, r->stream));
        hipLaunchKernelGGL(kSetSample, dim3(1), dim3(1), 0, r->stream, r->ctx.counters, first + count);
        WPT_HIP(hipGetLastError());
        WPT_HIP(hipEventSynchronize(r->end));
        float ms = 0.0f;
        WPT_HIP(hipEventElapsedTime(&ms, r->start, r->end));
        r->lastMs = ms;
        r->lastSamples = static_cast<uint64_t>(count) * r->config.width * r->config.height;
        return WPT_OK;
    }
    int status = ensureGraph(r);
    if (status != WPT_OK) return status;
    hipLaunchKernelGGL(kClearStats, dim3(1), dim3(1), 0, r->stream, r->ctx.counters);
    hipLaunchKernelGGL(kSetSample, dim3(1), dim3(1), 0, r->stream, r->ctx.counters, first);
    WPT_HIP(hipGetLastError());
    WPT_HIP(hipEventRecord(r->start, r->stream));
    for (uint32_t s = 0; s < count; s += r->ctx.slots) {
        for (uint32_t t = 0; t < r->tiles.size(); ++t) {
            RenderTile tile = r->tiles[t];
            tile.advanceSample = t + 1 == r->tiles.size() ? 1u : 0u;
            tile.samples = std::min(r->ctx.slots, count - s);
            status = launchTilePass(r, tile);
            if (status != WPT_OK) return status;
        }
    }
    WPT_HIP(hipEventRecord(r->end, r->stream));
    WPT_HIP(hipEventSynchronize(r->end));
    float ms = 0.0f;
    WPT_HIP(hipEventElapsedTime(&ms, r->start, r->end));
    r->lastMs = ms;
    r->lastSamples = static_cast<uint64_t>(count) * r->config.width * r->config.height;
    return WPT_OK;
}

WPT_API int wpt_renderer_read_framebuffer(wpt_renderer* r, float* host, uint32_t* sampleCount) {
    if (!r || !host) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    WPT_HIP(hipStreamSynchronize(r->stream));
    size_t n = 3 * static_cast<size_t>(r->config.width) * r->config.height;
    WPT_HIP(hipMemcpy(host, r->ctx.framebuffer, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (sampleCount) {
        WavefrontCounters c;
        WPT_HIP(hipMemcpy(&c, r->ctx.counters, sizeof c, hipMemcpyDeviceToHost));
        *sampleCount = c.sampleIndex;
    }
    return WPT_OK;
}

WPT_API int wpt_renderer_load_framebuffer(wpt_renderer* r, const float* host, uint32_t sampleCount) {
    if (!r || !host) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    WPT_HIP(hipStreamSynchronize(r->stream));
    size_t n = 3 * static_cast<size_t>(r->config.width) * r->config.height;
    WPT_HIP(hipMemcpy(r->ctx.framebuffer, host, sizeof(float) * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(kSetSample, dim3(1), dim3(1), 0, r->stream, r->ctx.counters, sampleCount);
    WPT_HIP(hipGetLastError());
    WPT_HIP(hipStreamSynchronize(r->stream));
    return WPT_OK;
}

WPT_API int wpt_render_frame(wpt_renderer* r, float* host, uint32_t* sampleCount) {
    if (!r || !host) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    int s = wpt_renderer_reset(r);
    if (s != WPT_OK) return s;
    s = wpt_renderer_render_samples(r, 0, r->config.samples_per_pixel);
    if (s != WPT_OK) return s;
    return wpt_renderer_read_framebuffer(r, host, sampleCount);
}

WPT_API int wpt_render_task(wpt_renderer* r, const wpt_task* task, float* hostTile) {
    if (!r || !task || !hostTile) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    const wpt_render_config& c = r->config;
    if (task->job_fingerprint != r->fingerprint)
        return fail(WPT_ERROR_INVALID_TASK, "Render task fingerprint does not match the prepared renderer");
    if (task->width == 0 || task->height == 0 || task->x >= c.width || task->y >= c.height ||
        task->width > c.width - task->x || task->height > c.height - task->y ||
        static_cast<uint64_t>(task->width) * task->height > r->tileCapacity)
        return fail(WPT_ERROR_INVALID_TASK, "Render task tile is outside the prepared renderer bounds");
    if (task->sample_count == 0 || task->sample_start > c.samples_per_pixel ||
        task->sample_count > c.samples_per_pixel - task->sample_start)
        return fail(WPT_ERROR_INVALID_TASK, "Render task sample range is outside the prepared renderer bounds");
    RenderTile tile{task->x, task->y, task->width, task->height, 1u, 1u};
    WPT_HIP(hipEventRecord(r->start, r->stream));
    if (c.mode == WPT_MODE_REGENERATE) {
        // Sums written over the tile: the same values as clearing it first.
        int s = regenRun(r, {regenTask(r, tile, task->sample_start, task->sample_count, false)});
        if (s != WPT_OK) return s;
    } else {
        int status = ensureGraph(r);
        if (status != WPT_OK) return status;
        size_t n = static_cast<size_t>(task->width) * task->height * 3;
        hipLaunchKernelGGL(kClearRect, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, r->stream,
            r->ctx.framebuffer, c.width, task->x, task->y, task->width, task->height);
        hipLaunchKernelGGL(kClearStats, dim3(1), dim3(1), 0, r->stream, r->ctx.counters);
        hipLaunchKernelGGL(kSetSample, dim3(1), dim3(1), 0, r->stream, r->ctx.counters, task->sample_start);
        WPT_HIP(hipGetLastError());
        for (uint32_t s = 0; s < task->sample_count; s += r->ctx.slots) {
            tile.samples = std::min(r->ctx.slots, task->sample_count - s);
            status = launchTilePass(r, tile);
            if (status != WPT_OK) return status;
        }
    }
    WPT_HIP(hipEventRecord(r->end, r->stream));
    WPT_HIP(hipStreamSynchronize(r->stream));
    float ms = 0.0f;
    WPT_HIP(hipEventElapsedTime(&ms, r->start, r->end));
    r->lastMs = ms;
    r->lastSamples = static_cast<uint64_t>(task->sample_count) * task->width * task->height;
    size_t pitch = sizeof(float) * 3 * c.width;
    const float* src = r->ctx.framebuffer + (static_cast<size_t>(task->y) * c.width + task->x) * 3;
    WPT_HIP(hipMemcpy2D(hostTile, sizeof(float) * 3 * task->width, src, pitch, sizeof(float) * 3 * task->width,
        task->height, hipMemcpyDeviceToHost));
    return WPT_OK;
}

WPT_API int wpt_renderer_profile_kernels(wpt_renderer* r, uint32_t first, uint32_t count, double* kernelMs,
    uint32_t* launches) {
    if (!r || !kernelMs || !launches) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    if (r->config.mode == WPT_MODE_REGENERATE) {
        // One launch over all tiles and samples, timed as the persistent slot.
        int s = wpt_renderer_render_samples(r, first, count);
        if (s != WPT_OK) return s;
        kernelMs[4] += r->lastMs;
        launches[4] += 1;
        return WPT_OK;
    }
    hipLaunchKernelGGL(kClearStats, dim3(1), dim3(1), 0, r->stream, r->ctx.counters);
    hipLaunchKernelGGL(kSetSample, dim3(1), dim3(1), 0, r->stream, r->ctx.counters, first);
    WPT_HIP(hipGetLastError());
    bool persistent = r->config.mode == WPT_MODE_PERSISTENT;
    for (uint32_t s = 0; s < count; s += r->ctx.slots) {
        for (uint32_t t = 0; t < r->tiles.size(); ++t) {
            RenderTile tile = r->tiles[t];
            tile.advanceSample = t + 1 == r->tiles.size() ? 1u : 0u;
            tile.samples = std::min(r->ctx.slots, count - s);
            hipLaunchKernelGGL(kSetTile, dim3(1), dim3(1), 0, r->stream, const_cast<RenderTile*>(r->ctx.tile), tile);
            WPT_HIP(hipGetLastError());
            WPT_HIP(profileSample(r->ctx, r->stream, persistent, kernelMs, launches));
        }
    }
    r->tileValid = false;
    return WPT_OK;
}

WPT_API int wpt_renderer_launch_info(const wpt_renderer* r, wpt_launch_info* info) {
    if (!r || !info) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    const wpt_render_config& c = r->config;
    const LaunchContext& ctx = r->ctx;
    bool full = ctx.scene.fullFeatures != 0;
    *info = wpt_launch_info{};
    info->mode = c.mode;
    info->block_size = c.block_size;
    hipFuncAttributes attr{};
    int perCu = 0;
    size_t lds = c.mode == WPT_MODE_GRAPH ? traversalLdsBytes(c.block_size) : stackLdsBytes(c.block_size);
    if (c.mode == WPT_MODE_REGENERATE) {
        WPT_HIP(batchAttributes(full, &attr));
        WPT_HIP(batchOccupancy(c.block_size, full, &perCu));
        info->blocks = r->regenBlocks;
    } else {
        Occupancy occ{};
        bool events = ctx.events.words != nullptr;
        WPT_HIP(kernelOccupancy(c.block_size, full, &occ, events));
        WPT_HIP(kernelAttributes(c.mode == WPT_MODE_PERSISTENT, full, &attr, events));
        perCu = c.mode == WPT_MODE_PERSISTENT ? occ.persistent : occ.trace;
        info->blocks = c.mode == WPT_MODE_PERSISTENT ? ctx.persistentBlocks : 0u;
    }
    int device = 0, cus = 0, ldsCu = 0;
    WPT_HIP(hipGetDevice(&device));
    WPT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    WPT_HIP(hipDeviceGetAttribute(&ldsCu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device));
    info->blocks_per_cu = static_cast<uint32_t>(perCu);
    info->compute_units = static_cast<uint32_t>(cus);
    info->registers = static_cast<uint32_t>(attr.numRegs);
    info->scratch_bytes = static_cast<uint32_t>(attr.localSizeBytes);
    info->lds_bytes_per_block = static_cast<uint32_t>(attr.sharedSizeBytes + lds);
    info->lds_bytes_per_cu = static_cast<uint32_t>(ldsCu);
    info->waves_per_simd = static_cast<uint32_t>(perCu) * ((c.block_size + 63u) / 64u) / 4u;
    return WPT_OK;
}

WPT_API int wpt_renderer_last_stats(const wpt_renderer* r, wpt_stats* stats) {
    if (!r || !stats) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    WavefrontCounters c;
    WPT_HIP(hipStreamSynchronize(r->stream));
    WPT_HIP(hipMemcpy(&c, r->ctx.counters, sizeof c, hipMemcpyDeviceToHost));
    std::memset(stats, 0, sizeof *stats);
    if (r->config.mode == WPT_MODE_REGENERATE) {
        // No queues: per-bounce counts are not recorded.
        BatchCounters b;
        WPT_HIP(hipMemcpy(&b, r->regen.counters, sizeof b, hipMemcpyDeviceToHost));
        c.closestQueries = b.closestQueries;
        c.shadowQueries = b.shadowQueries;
        std::memset(c.queueCounts, 0, sizeof c.queueCounts);
    }
    stats->closest_queries = c.closestQueries;
    stats->shadow_queries = c.shadowQueries;
    stats->samples = r->lastSamples;
    stats->trace_milliseconds = r->lastMs;
    uint32_t depth = std::min<uint32_t>(64u, r->config.max_depth);
    for (uint32_t i = 0; i < depth; ++i) stats->queue_counts[i] = c.queueCounts[i];
    stats->queue_count_depth = depth;
    stats->caustic_photons_stored = r->photonsStored;
    stats->caustic_build_milliseconds = r->photonMs;
    return WPT_OK;
}

}  // extern "C"

// ---- batch ---------------------------------------------------------------

struct wpt_batch {
    std::vector<wpt_scene*> scenes;
    wpt_batch_config config{};
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, end = nullptr;
    DevScene* scenesDev = nullptr;
    BatchCounters* counters = nullptr;
    BatchTask* tasks = nullptr;
    BatchTaskState* state = nullptr;
    uint32_t taskCapacity = 0;
    float* out = nullptr;
    size_t outCapacity = 0;  // floats
    uint2* spill = nullptr;
    uint32_t spillThreads = 0;
    uint32_t* events = nullptr;  // full-feature kernel: one event column per thread
    uint32_t eventThreads = 0, eventMax = 0;
    bool full = false;  // some scene needs the full-feature kernel
    int cus = 0;
    int wallClockKhz = 0;
    wpt_batch_stats last{};
    // Per scene: the caustic photon map its tasks gather at diffuse vertices
    // (null: none) and the buffers behind it.
    std::vector<const PhotonGridView*> photons;
    std::vector<std::unique_ptr<DeviceArena>> photonArena;
    // Staged scheduler: wavefront buffers for passes of `stagedCapacity`
    // paths, shared by every scene's passes (allocated on first use).
    std::unique_ptr<DeviceArena> stagedArena;
    LaunchContext staged{};
    uint32_t stagedCapacity = 0, stagedBlock = 0, stagedEvents = 0;
    PassPiece* pieces = nullptr;
    size_t pieceCapacity = 0;
    ~wpt_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : {static_cast<void*>(scenesDev), static_cast<void*>(counters), static_cast<void*>(tasks),
                 static_cast<void*>(state), static_cast<void*>(out), static_cast<void*>(spill),
                 static_cast<void*>(events), static_cast<void*>(pieces)})
            if (p) (void)hipFree(p);
        if (start) (void)hipEventDestroy(start);
        if (end) (void)hipEventDestroy(end);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int validateBatchConfig(const wpt_batch_config& c) {
    if (c.block_size == 0 || c.block_size > 512 || c.block_size % 64 != 0)
        return fail(WPT_ERROR_INVALID_ARGUMENT, "Block size must be a multiple of 64 between 64 and 512");
    if (c.scheduler > WPT_BATCH_REGENERATE) return fail(WPT_ERROR_INVALID_ARGUMENT, "unknown batch scheduler");
    if (c.pass_paths > (1u << 26)) return fail(WPT_ERROR_INVALID_ARGUMENT, "staged batch passes hold at most 2^26 paths");
    return WPT_OK;
}

// Grows a device buffer to hold `count` elements; contents are not kept.
template <class T>
int growDevice(T** p, size_t count, size_t* capacity) {
    if (count <= *capacity) return WPT_OK;
    if (*p) WPT_HIP(hipFree(*p));
    *p = nullptr;
    *capacity = 0;
    WPT_HIP(hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * count));
    *capacity = count;
    return WPT_OK;
}

// Resident workgroups of the batch kernel at `blockSize`, and the grid a
// configuration asks for.
int batchGrid(wpt_batch* b, const wpt_batch_config& c, uint32_t* blocks) {
    int perCu = 0;
    WPT_HIP(batchOccupancy(c.block_size, b->full, &perCu));
    if (perCu <= 0) return fail(WPT_ERROR_UNSUPPORTED, "block size cannot launch the batch kernel");
    uint32_t resident = static_cast<uint32_t>(perCu) * static_cast<uint32_t>(b->cus);
    *blocks = c.blocks ? std::min(c.blocks, resident) : resident;
    return WPT_OK;
}

// Validates the tasks and uploads them; returns the float offsets of each
// task's sums in b->out and the total.
int stageBatch(wpt_batch* b, const wpt_batch_task* tasks, uint32_t n, std::vector<size_t>& offsets, uint64_t* samples) {
    std::vector<BatchTask> host(n);
    offsets.assign(n, 0);
    size_t total = 0;
    *samples = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const wpt_batch_task& t = tasks[i];
        if (t.scene >= b->scenes.size()) return fail(WPT_ERROR_INVALID_TASK, "Batch task " + std::to_string(i) + ": scene index out of range");
        if (t.width == 0 || t.height == 0 || static_cast<uint64_t>(t.width) * t.height > std::numeric_limits<uint32_t>::max())
            return fail(WPT_ERROR_INVALID_TASK, "Batch task " + std::to_string(i) + ": invalid frame dimensions");
        if (t.tile_width == 0 || t.tile_height == 0 || t.x >= t.width || t.y >= t.height ||
            t.tile_width > t.width - t.x || t.tile_height > t.height - t.y)
            return fail(WPT_ERROR_INVALID_TASK, "Batch task " + std::to_string(i) + ": tile is outside the frame");
        if (t.max_depth == 0 || t.max_depth > 0x7fffffffu)
            return fail(WPT_ERROR_INVALID_TASK, "Batch task " + std::to_string(i) + ": maximum depth out of range");
        if (t.rng_strategy > WPT_RNG_PCG32) return fail(WPT_ERROR_INVALID_TASK, "Batch task " + std::to_string(i) + ": unknown RNG strategy");
        if (!std::isfinite(t.indirect_clamp) || t.indirect_clamp < 0.0f)
            return fail(WPT_ERROR_INVALID_TASK, "Batch task " + std::to_string(i) + ": indirect clamp must be finite and nonnegative");
        if (static_cast<uint64_t>(t.sample_start) + t.sample_count > 0xffffffffull)
            return fail(WPT_ERROR_INVALID_TASK, "Batch task " + std::to_string(i) + ": sample range out of range");
        BatchTask& d = host[i];
        d.camera.position = Vec3(t.camera.position[0], t.camera.position[1], t.camera.position[2]);
        d.camera.forward = Vec3(t.camera.forward[0], t.camera.forward[1], t.camera.forward[2]);
        d.camera.right = Vec3(t.camera.right[0], t.camera.right[1], t.camera.right[2]);
        d.camera.up = Vec3(t.camera.up[0], t.camera.up[1], t.camera.up[2]);
        d.camera.verticalFov = t.camera.vertical_fov;
        d.camera.aspectRatio = t.camera.aspect_ratio;
        d.shade = ShadeParams{t.max_depth, t.rr_start_depth, t.rng_strategy, t.indirect_clamp, 0u, 0u,
            b->photons[t.scene]};
        d.tile = RenderTile{t.x, t.y, t.tile_width, t.tile_height, 0u, 1u};
        d.scene = t.scene;
        d.width = t.width;
        d.height = t.height;
        d.sampleStart = t.sample_start;
        d.sampleCount = t.sample_count;
        uint32_t pixels = t.tile_width * t.tile_height;
        d.pixelCount = t.sample_count ? pixels : 0u;
        d.outStride = t.tile_width;
        d.accumulate = 0;
        d.out = total;
        offsets[i] = total;
        total += 3ull * pixels;
        *samples += static_cast<uint64_t>(pixels) * t.sample_count;
    }
    size_t cap = b->taskCapacity;
    int s = growDevice(&b->tasks, std::max<size_t>(n, 1), &cap);
    if (s != WPT_OK) return s;
    cap = b->taskCapacity;
    s = growDevice(&b->state, std::max<size_t>(n, 1), &cap);
    if (s != WPT_OK) return s;
    b->taskCapacity = static_cast<uint32_t>(cap);
    s = growDevice(&b->out, std::max<size_t>(total, 1), &b->outCapacity);
    if (s != WPT_OK) return s;
    if (n) WPT_HIP(hipMemcpy(b->tasks, host.data(), sizeof(BatchTask) * n, hipMemcpyHostToDevice));
    // Tasks without samples still report zero sums.
    WPT_HIP(hipMemsetAsync(b->out, 0, sizeof(float) * std::max<size_t>(total, 1), b->stream));
    return WPT_OK;
}

// Traversal spill columns for `threads` threads.
int ensureBatchSpill(wpt_batch* b, uint32_t threads) {
    if (threads <= b->spillThreads) return WPT_OK;
    if (b->spill) WPT_HIP(hipFree(b->spill));
    b->spill = nullptr;
    b->spillThreads = 0;
    WPT_HIP(hipMalloc(reinterpret_cast<void**>(&b->spill), sizeof(uint2) * kSpillEntriesTotal * static_cast<size_t>(threads)));
    b->spillThreads = threads;
    return WPT_OK;
}

// Event columns for `threads` threads when a scene needs the full-feature
// kernel (none when a scene would store more than kMaxVertexEvents).
int ensureBatchEvents(wpt_batch* b, uint32_t threads) {
    uint32_t maxEvents = 0;
    for (const wpt_scene* s : b->scenes) maxEvents = std::max(maxEvents, vertexEventsOf(s->dev));
    if (!b->full || maxEvents > kMaxVertexEvents) return WPT_OK;
    if (threads <= b->eventThreads && maxEvents <= b->eventMax) return WPT_OK;
    if (b->events) WPT_HIP(hipFree(b->events));
    b->events = nullptr;
    b->eventThreads = b->eventMax = 0;
    WPT_HIP(hipMalloc(reinterpret_cast<void**>(&b->events),
        sizeof(uint32_t) * static_cast<size_t>(eventWordsPerColumn(maxEvents)) * threads));
    b->eventThreads = threads;
    b->eventMax = maxEvents;
    return WPT_OK;
}

// One launch over the staged tasks under configuration `c`; device time in *ms.
int runBatch(wpt_batch* b, uint32_t n, const wpt_batch_config& c, float* ms) {
    uint32_t blocks = 0;
    int s = batchGrid(b, c, &blocks);
    if (s != WPT_OK) return s;
    s = ensureBatchSpill(b, blocks * c.block_size);
    if (s != WPT_OK) return s;
    s = ensureBatchEvents(b, blocks * c.block_size);
    if (s != WPT_OK) return s;
    BatchContext ctx{};
    ctx.scenes = b->scenesDev;
    ctx.tasks = b->tasks;
    ctx.state = b->state;
    ctx.counters = b->counters;
    ctx.taskCount = n;
    ctx.out = b->out;
    ctx.spill = b->spill;
    ctx.spillThreads = b->spillThreads;
    ctx.full = b->full ? 1u : 0u;
    if (b->full && b->events) ctx.events = EventBuffer{b->events, b->eventThreads, b->eventMax};
    WPT_HIP(hipEventRecord(b->start, b->stream));
    WPT_HIP(enqueueBatch(ctx, blocks, c.block_size, b->stream));
    WPT_HIP(hipEventRecord(b->end, b->stream));
    WPT_HIP(hipStreamSynchronize(b->stream));
    WPT_HIP(hipEventElapsedTime(ms, b->start, b->end));
    b->last.block_size = c.block_size;
    b->last.blocks = blocks;
    return WPT_OK;
}


// ---- staged batches
// Paths per staged batch pass (2^26: 32 samples of a 1920x1080 tile; about
// 31 GB of queues, records and events when a scene needs events; the 256-task
// benchmark batch: 2.72 s at 2^25, 2.59 s at 2^26).
#ifndef WPT_STAGED_BATCH_PATHS_LOG2
#define WPT_STAGED_BATCH_PATHS_LOG2 26
#endif
constexpr uint32_t kStagedBatchPaths = 1u << WPT_STAGED_BATCH_PATHS_LOG2;

// Allocates the shared wavefront buffers of staged batch passes: queues, hit
// records, deferred shadow requests, radiance events (when a scene needs
// them), per-path radiance, counters and the traversal spill area.
int ensureStagedBatch(wpt_batch* b, uint32_t blockSize, uint32_t paths) {
    uint32_t maxEvents = 0;
    bool anyEvents = false, full = false;
    for (size_t i = 0; i < b->scenes.size(); ++i) {
        const DevScene& d = b->scenes[i]->dev;
        full = full || d.fullFeatures || b->photons[i];
        if (shadesWithInlineShadows(d) || b->photons[i]) {
            anyEvents = true;
            maxEvents = std::max(maxEvents, vertexEventsOf(d));
        }
    }
    if (maxEvents > kMaxVertexEvents) maxEvents = 0;  // such scenes trace inline
    uint32_t wantEvents = anyEvents ? maxEvents : 0u;
    if (b->stagedArena && b->stagedBlock == blockSize && b->stagedEvents >= wantEvents && b->stagedCapacity >= paths)
        return WPT_OK;
    b->stagedArena.reset();  // free the old buffers first
    b->stagedArena.reset(new (std::nothrow) DeviceArena());
    if (!b->stagedArena) return fail(WPT_ERROR_OUT_OF_MEMORY, "host allocation failed");
    b->staged = LaunchContext{};
    LaunchContext& ctx = b->staged;
    DeviceArena& arena = *b->stagedArena;
    const uint32_t cap = std::max(paths, 1u), B = blockSize;
    Occupancy lean{}, rich{};
    WPT_HIP(kernelOccupancy(B, false, &lean));
    if (full) WPT_HIP(kernelOccupancy(B, true, &rich));
    int trace = std::max(lean.trace, full ? rich.trace : 0);
    if (trace <= 0 || lean.shade <= 0 || (full && rich.shade <= 0))
        return fail(WPT_ERROR_UNSUPPORTED, "block size cannot launch the wavefront kernels");
    ctx.blockSize = B;
    ctx.capacity = cap;
    ctx.gridBlocks = std::max(1u, std::min((cap + B - 1) / B, kStagedGridFactor * static_cast<uint32_t>(trace) *
        static_cast<uint32_t>(b->cus)));
    uint64_t rounds = (static_cast<uint64_t>(cap) + static_cast<uint64_t>(ctx.gridBlocks) * B - 1) /
        (static_cast<uint64_t>(ctx.gridBlocks) * B);
    uint64_t virtualBlocks = (static_cast<uint64_t>(cap) + B - 1) / B;
    uint64_t segment = std::max(static_cast<uint64_t>((ctx.gridBlocks + kSegments - 1) / kSegments) * rounds,
                                (virtualBlocks + kSegments - 1) / kSegments) * B;
    uint64_t qcap = segment * kSegments;
    if (qcap >= (1ull << 31)) return fail(WPT_ERROR_INVALID_ARGUMENT, "queue capacity out of range");
    void* p = nullptr;
    for (int q = 0; q < 2; ++q) {
        WPT_HIP(arena.alloc(sizeof(uint32_t) * static_cast<size_t>(kQueueFieldCount) * qcap, &p));
        ctx.queue[q].words = static_cast<uint32_t*>(p);
        ctx.queue[q].capacity = static_cast<uint32_t>(qcap);
        ctx.queue[q].segmentCapacity = static_cast<uint32_t>(segment);
    }
    WPT_HIP(arena.alloc(sizeof(float) * 3 * static_cast<size_t>(qcap) + sizeof(uint32_t) * 4 * static_cast<size_t>(qcap), &p));
    float* f = static_cast<float*>(p);
    ctx.hits.t = f;
    ctx.hits.u = f + qcap;
    ctx.hits.v = f + 2 * static_cast<size_t>(qcap);
    uint32_t* u = reinterpret_cast<uint32_t*>(f + 3 * static_cast<size_t>(qcap));
    ctx.hits.kind = u;
    ctx.hits.instance = u + qcap;
    ctx.hits.prim = u + 2 * static_cast<size_t>(qcap);
    ctx.hits.material = u + 3 * static_cast<size_t>(qcap);
    WPT_HIP(arena.alloc(sizeof(uint32_t) * static_cast<size_t>(kShadowFieldCount) * qcap, &p));
    ctx.shadow.words = static_cast<uint32_t*>(p);
    ctx.shadow.capacity = static_cast<uint32_t>(qcap);
    ctx.shadow.segmentCapacity = static_cast<uint32_t>(segment);
    if (wantEvents) {
        WPT_HIP(arena.alloc(sizeof(uint32_t) * static_cast<size_t>(eventWordsPerColumn(wantEvents)) * qcap, &p));
        ctx.events.words = static_cast<uint32_t*>(p);
        ctx.events.capacity = static_cast<uint32_t>(qcap);
        ctx.events.maxEvents = wantEvents;
        WPT_HIP(allocEventRays(arena, segment, wantEvents, &ctx.events));
    }
    WPT_HIP(arena.alloc(sizeof(float) * 3 * static_cast<size_t>(cap), &p));
    ctx.slotRadiance = static_cast<float*>(p);
    WPT_HIP(arena.alloc(sizeof(WavefrontCounters), &p));
    ctx.counters = static_cast<WavefrontCounters*>(p);
    WPT_HIP(hipMemset(p, 0, sizeof(WavefrontCounters)));
    ctx.spillThreads = ctx.gridBlocks * B;
    WPT_HIP(arena.alloc(sizeof(uint2) * kSpillEntriesTotal * static_cast<size_t>(ctx.spillThreads), &p));
    ctx.spill = static_cast<uint2*>(p);
    ctx.profileQueues = 0;
    ctx.slots = 1;
    b->stagedCapacity = cap;
    b->stagedBlock = B;
    b->stagedEvents = wantEvents;
    return WPT_OK;
}

// Tasks that one staged pass can share: same scene, camera, frame and path
// parameters.
bool sameStagedGroup(const wpt_batch_task& a, const wpt_batch_task& c) {
    return a.scene == c.scene && std::memcmp(&a.camera, &c.camera, sizeof a.camera) == 0 && a.width == c.width &&
        a.height == c.height && a.max_depth == c.max_depth && a.rr_start_depth == c.rr_start_depth &&
        a.rng_strategy == c.rng_strategy && bitsOf(a.indirect_clamp) == bitsOf(c.indirect_clamp);
}

// Every task through staged passes: the tasks of one group fill passes of up
// to stagedCapacity paths (a task's samples split across passes in order),
// each pass ends by adding its pieces' samples to the tasks' sums.  *ms is
// the device time; latencyMs (optional) the end of each task's last pass.
int runStagedBatch(wpt_batch* b, const wpt_batch_task* tasks, uint32_t n, const std::vector<size_t>& offsets,
    float* ms, double* latencyMs) {
    const uint32_t passLimit = b->config.pass_paths ? b->config.pass_paths : kStagedBatchPaths;
    for (uint32_t i = 0; i < n; ++i)
        if (tasks[i].sample_count && static_cast<uint64_t>(tasks[i].tile_width) * tasks[i].tile_height > passLimit)
            return fail(WPT_ERROR_UNSUPPORTED, "Batch task " + std::to_string(i) +
                ": tile larger than a staged pass; use the regenerating batch scheduler");
    // Passes hold at most passLimit paths; a smaller batch needs no more
    // than its own paths.
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; ++i) total += static_cast<uint64_t>(tasks[i].tile_width) * tasks[i].tile_height * tasks[i].sample_count;
    int s = ensureStagedBatch(b, b->config.block_size, static_cast<uint32_t>(std::min<uint64_t>(total, passLimit)));
    if (s != WPT_OK) return s;
    const uint32_t cap = std::min(b->stagedCapacity, passLimit);
    // Groups in order of their first task; passes in group order.
    std::vector<int> group(n, -1);
    std::vector<uint32_t> leaders;
    for (uint32_t i = 0; i < n; ++i) {
        for (size_t g = 0; g < leaders.size() && group[i] < 0; ++g)
            if (sameStagedGroup(tasks[leaders[g]], tasks[i])) group[i] = static_cast<int>(g);
        if (group[i] < 0) {
            group[i] = static_cast<int>(leaders.size());
            leaders.push_back(i);
        }
    }
    struct Pass {
        uint32_t leader;
        std::vector<PassPiece> pieces;
        std::vector<uint32_t> finishing;  // tasks whose last samples are in this pass
        uint32_t paths = 0, pixels = 0;
    };
    std::vector<Pass> passes;
    for (size_t g = 0; g < leaders.size(); ++g) {
        Pass cur{leaders[g], {}, {}, 0, 0};
        for (uint32_t i = 0; i < n; ++i) {
            if (group[i] != static_cast<int>(g)) continue;
            const wpt_batch_task& t = tasks[i];
            uint32_t pixels = t.tile_width * t.tile_height, done = 0;
            while (done < t.sample_count) {
                uint32_t fit = (cap - cur.paths) / pixels;
                if (fit == 0) {
                    passes.push_back(std::move(cur));
                    cur = Pass{leaders[g], {}, {}, 0, 0};
                    continue;
                }
                uint32_t k = std::min(fit, t.sample_count - done);
                PassPiece piece{};
                piece.tile = RenderTile{t.x, t.y, t.tile_width, t.tile_height, 0u, 1u};
                piece.sampleStart = t.sample_start + done;
                piece.samples = k;
                piece.pathBase = cur.paths;
                piece.pixelBase = cur.pixels;
                piece.outStride = t.tile_width;
                piece.out = offsets[i];
                cur.pieces.push_back(piece);
                cur.paths += k * pixels;
                cur.pixels += pixels;
                done += k;
                if (done == t.sample_count) cur.finishing.push_back(i);
            }
        }
        if (!cur.pieces.empty()) passes.push_back(std::move(cur));
    }
    size_t most = 1;
    for (const Pass& pass : passes) most = std::max(most, pass.pieces.size());
    s = growDevice(&b->pieces, most, &b->pieceCapacity);
    if (s != WPT_OK) return s;
    std::vector<hipEvent_t> ends(passes.size(), nullptr);
    auto release = [&] {
        for (hipEvent_t e : ends)
            if (e) (void)hipEventDestroy(e);
    };
    WPT_HIP(hipMemsetAsync(b->staged.counters, 0, sizeof(WavefrontCounters), b->stream));
    WPT_HIP(hipEventRecord(b->start, b->stream));
    for (size_t k = 0; k < passes.size(); ++k) {
        const Pass& pass = passes[k];
        const wpt_batch_task& t = tasks[pass.leader];
        hipError_t e = hipMemcpyAsync(b->pieces, pass.pieces.data(), sizeof(PassPiece) * pass.pieces.size(),
            hipMemcpyHostToDevice, b->stream);
        LaunchContext ctx = b->staged;
        ctx.scene = b->scenes[t.scene]->dev;
        const PhotonGridView* photons = b->photons[t.scene];
        if (photons) ctx.scene.fullFeatures = 1;  // the gather runs in the full-feature kernels
        if (!(shadesWithInlineShadows(ctx.scene) || photons) || vertexEventsOf(ctx.scene) > kMaxVertexEvents)
            ctx.events = EventBuffer{};
        ctx.camera.position = Vec3(t.camera.position[0], t.camera.position[1], t.camera.position[2]);
        ctx.camera.forward = Vec3(t.camera.forward[0], t.camera.forward[1], t.camera.forward[2]);
        ctx.camera.right = Vec3(t.camera.right[0], t.camera.right[1], t.camera.right[2]);
        ctx.camera.up = Vec3(t.camera.up[0], t.camera.up[1], t.camera.up[2]);
        ctx.camera.verticalFov = t.camera.vertical_fov;
        ctx.camera.aspectRatio = t.camera.aspect_ratio;
        ctx.shade = ShadeParams{t.max_depth, t.rr_start_depth, t.rng_strategy, t.indirect_clamp, 0u, 0u, photons};
        ctx.width = t.width;
        ctx.height = t.height;
        ctx.pieces = b->pieces;
        ctx.pieceCount = static_cast<uint32_t>(pass.pieces.size());
        ctx.passPaths = pass.paths;
        ctx.passPixels = pass.pixels;
        ctx.batchOut = b->out;
        if (e == hipSuccess) e = enqueueStagedBatchPass(ctx, b->stream);
        if (e == hipSuccess) e = hipEventCreate(&ends[k]);
        if (e == hipSuccess) e = hipEventRecord(ends[k], b->stream);
        if (e != hipSuccess) {
            release();
            return hipFail(e, "staged batch pass");
        }
    }
    hipError_t e = hipEventRecord(b->end, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, b->start, b->end);
    for (size_t k = 0; e == hipSuccess && latencyMs && k < passes.size(); ++k) {
        float at = 0.0f;
        e = hipEventElapsedTime(&at, b->start, ends[k]);
        for (uint32_t i : passes[k].finishing) latencyMs[i] = at;
    }
    release();
    if (e != hipSuccess) return hipFail(e, "staged batch");
    if (latencyMs)
        for (uint32_t i = 0; i < n; ++i)
            if (tasks[i].sample_count == 0) latencyMs[i] = 0.0;
    WavefrontCounters counters{};
    WPT_HIP(hipMemcpy(&counters, b->staged.counters, sizeof counters, hipMemcpyDeviceToHost));
    b->last.closest_queries = counters.closestQueries;
    b->last.shadow_queries = counters.shadowQueries;
    b->last.block_size = b->config.block_size;
    b->last.blocks = b->staged.gridBlocks;
    return WPT_OK;
}

}  // namespace

extern "C" {

WPT_API int wpt_batch_create(wpt_scene* const* scenes, uint32_t sceneCount, const wpt_batch_config* config,
    wpt_batch** out) {
    if (!out || !config || (sceneCount && !scenes)) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (sceneCount == 0) return fail(WPT_ERROR_INVALID_ARGUMENT, "a batch needs at least one scene");
    int v = validateBatchConfig(*config);
    if (v != WPT_OK) return v;
    std::vector<DevScene> dev(sceneCount);
    for (uint32_t i = 0; i < sceneCount; ++i) {
        if (!scenes[i]) return fail(WPT_ERROR_INVALID_ARGUMENT, "null scene");
        if (!scenes[i]->onDevice) return fail(WPT_ERROR_INVALID_ARGUMENT, "scene is not on a device");
        dev[i] = scenes[i]->dev;
    }
    wpt_batch* b = new (std::nothrow) wpt_batch();
    if (!b) return fail(WPT_ERROR_OUT_OF_MEMORY, "host allocation failed");
    b->scenes.assign(scenes, scenes + sceneCount);
    b->photons.assign(sceneCount, nullptr);
    b->photonArena.resize(sceneCount);
    for (uint32_t i = 0; i < sceneCount; ++i) b->full = b->full || scenes[i]->dev.fullFeatures;
    b->config = *config;
    auto cleanup = [&](int status) {
        delete b;
        return status;
    };
    hipError_t e;
#define B_HIP(call)                                            \
    do {                                                       \
        e = (call);                                            \
        if (e != hipSuccess) return cleanup(hipFail(e, #call)); \
    } while (0)
    int device = 0;
    B_HIP(hipGetDevice(&device));
    B_HIP(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device));
    B_HIP(hipDeviceGetAttribute(&b->wallClockKhz, hipDeviceAttributeWallClockRate, device));
    B_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    B_HIP(hipEventCreate(&b->start));
    B_HIP(hipEventCreate(&b->end));
    B_HIP(hipMalloc(reinterpret_cast<void**>(&b->scenesDev), sizeof(DevScene) * sceneCount));
    B_HIP(hipMemcpy(b->scenesDev, dev.data(), sizeof(DevScene) * sceneCount, hipMemcpyHostToDevice));
    B_HIP(hipMalloc(reinterpret_cast<void**>(&b->counters), sizeof(BatchCounters)));
#undef B_HIP
    uint32_t blocks = 0;
    int s = batchGrid(b, b->config, &blocks);
    if (s != WPT_OK) return cleanup(s);
    *out = b;
    return WPT_OK;
}

WPT_API int wpt_batch_set_caustics(wpt_batch* b, uint32_t scene, const wpt_render_config* config,
    uint32_t* photonsStored, double* buildMs) {
    if (!b || !config) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    if (scene >= b->scenes.size()) return fail(WPT_ERROR_INVALID_ARGUMENT, "scene index out of range");
    int s = validateCaustics(*config);
    if (s != WPT_OK) return s;
    WPT_HIP(hipStreamSynchronize(b->stream));
    b->photons[scene] = nullptr;
    b->photonArena[scene].reset();
    uint32_t stored = 0;
    double ms = 0.0;
    if (config->caustic_photons) {
        uint32_t blocks = 0;
        s = batchGrid(b, b->config, &blocks);
        if (s != WPT_OK) return s;
        s = ensureBatchSpill(b, std::max(blocks * b->config.block_size, 256u));
        if (s != WPT_OK) return s;
        auto arena = std::make_unique<DeviceArena>();
        PhotonMapResult map{};
        s = buildPhotonMap(b->scenes[scene]->dev, *config, *arena, b->spill, b->spillThreads, b->stream, b->start,
            b->end, &map);
        if (s != WPT_OK) return s;
        b->photonArena[scene] = std::move(arena);
        if (config->caustic_gather) b->photons[scene] = map.view;
        stored = map.stored;
        ms = map.milliseconds;
    }
    // Gathering needs the full-feature kernel.
    bool full = false;
    for (size_t i = 0; i < b->scenes.size(); ++i) full = full || b->scenes[i]->dev.fullFeatures || b->photons[i];
    b->full = full;
    if (photonsStored) *photonsStored = stored;
    if (buildMs) *buildMs = ms;
    return WPT_OK;
}

WPT_API int wpt_batch_destroy(wpt_batch* b) {
    delete b;
    return WPT_OK;
}

WPT_API int wpt_batch_render(wpt_batch* b, const wpt_batch_task* tasks, uint32_t n, float* const* outs,
    double* latencyMs) {
    if (!b || (n && (!tasks || !outs))) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    for (uint32_t i = 0; i < n; ++i)
        if (!outs[i]) return fail(WPT_ERROR_INVALID_ARGUMENT, "null task output");
    std::vector<size_t> offsets;
    uint64_t samples = 0;
    int s = stageBatch(b, tasks, n, offsets, &samples);
    if (s != WPT_OK) return s;
    float ms = 0.0f;
    const bool staged = b->config.scheduler == WPT_BATCH_STAGED;
    if (staged) {
        s = runStagedBatch(b, tasks, n, offsets, &ms, latencyMs);
        if (s != WPT_OK) return s;
    } else {
        s = runBatch(b, n, b->config, &ms);
        if (s != WPT_OK) return s;
    }
    BatchCounters counters{};
    if (!staged) {
        WPT_HIP(hipMemcpy(&counters, b->counters, sizeof counters, hipMemcpyDeviceToHost));
        b->last.closest_queries = counters.closestQueries;
        b->last.shadow_queries = counters.shadowQueries;
    }
    b->last.samples = samples;
    b->last.milliseconds = ms;
    size_t total = 0;
    for (uint32_t i = 0; i < n; ++i) total += 3ull * tasks[i].tile_width * tasks[i].tile_height;
    std::vector<float> host(total);
    if (total) WPT_HIP(hipMemcpy(host.data(), b->out, sizeof(float) * total, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i)
        std::memcpy(outs[i], host.data() + offsets[i], sizeof(float) * 3ull * tasks[i].tile_width * tasks[i].tile_height);
    if (latencyMs && n && !staged) {
        std::vector<BatchTaskState> state(n);
        WPT_HIP(hipMemcpy(state.data(), b->state, sizeof(BatchTaskState) * n, hipMemcpyDeviceToHost));
        double khz = b->wallClockKhz > 0 ? static_cast<double>(b->wallClockKhz) : 1.0;
        for (uint32_t i = 0; i < n; ++i)
            latencyMs[i] = state[i].finishClock > counters.startClock
                               ? static_cast<double>(state[i].finishClock - counters.startClock) / khz
                               : 0.0;
    }
    return WPT_OK;
}

WPT_API int wpt_batch_last_stats(const wpt_batch* b, wpt_batch_stats* stats) {
    if (!b || !stats) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    *stats = b->last;
    return WPT_OK;
}

WPT_API int wpt_batch_autotune(wpt_batch* b, const wpt_batch_task* tasks, uint32_t n, wpt_batch_config* chosen) {
    if (!b || (n && !tasks)) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    std::vector<size_t> offsets;
    uint64_t samples = 0;
    int s = stageBatch(b, tasks, n, offsets, &samples);
    if (s != WPT_OK) return s;
    wpt_batch_config best = b->config;
    float bestMs = std::numeric_limits<float>::infinity();
    if (b->config.scheduler == WPT_BATCH_STAGED) {
        // Staged passes: the workgroup size of the pass kernels.
        for (uint32_t block : {128u, 256u, 512u}) {
            wpt_batch_config c{block, 0u, WPT_BATCH_STAGED, b->config.pass_paths};
            wpt_batch_config saved = b->config;
            b->config = c;
            float ms = 0.0f;
            s = runStagedBatch(b, tasks, n, offsets, &ms, nullptr);
            b->config = saved;
            if (s != WPT_OK) return s;
            if (ms < bestMs) {
                bestMs = ms;
                best = c;
            }
        }
    } else {
        // One persistent launch: workgroup size and grid.
        for (uint32_t block : {64u, 128u, 256u, 512u}) {
            int perCu = 0;
            WPT_HIP(batchOccupancy(block, b->full, &perCu));
            if (perCu <= 0) continue;
            uint32_t resident = static_cast<uint32_t>(perCu) * static_cast<uint32_t>(b->cus);
            for (uint32_t quarter = 1; quarter <= 4; ++quarter) {
                wpt_batch_config c{block, std::max(1u, resident * quarter / 4u), WPT_BATCH_REGENERATE, b->config.pass_paths};
                float ms = 0.0f;
                s = runBatch(b, n, c, &ms);
                if (s != WPT_OK) return s;
                if (ms < bestMs) {
                    bestMs = ms;
                    best = c;
                }
            }
        }
    }
    b->config = best;
    if (chosen) *chosen = best;
    return WPT_OK;
}

}  // extern "C"

// ---- differentiable rendering ---------------------------------------------

struct wpt_diff {
    wpt_scene* scene = nullptr;
    wpt_diff_config config{};
    hipStream_t stream = nullptr;
    hipEvent_t marks[4] = {nullptr, nullptr, nullptr, nullptr};
    DeviceArena arena;
    DiffContext ctx{};
    uint32_t forwardBlocks = 0, adjointBlocks = 0;
    uint32_t textureFloats = 0, roughnessCount = 0, emissionCount = 0;
    // Parameters of the kept path radiance; cleared when wpt_diff_adam
    // changes them, so an adjoint cannot replay paths of older values.
    const float* forwardParams = nullptr;
    wpt_diff_stats stats{};
    ~wpt_diff() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipEvent_t e : marks)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

WPT_API int wpt_diff_create(wpt_scene* scene, const wpt_camera* camera, const wpt_diff_config* config, wpt_diff** out) {
    if (!scene || !camera || !config || !out) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (!scene->onDevice) return fail(WPT_ERROR_INVALID_ARGUMENT, "scene is not on a device");
    if (scene->dev.fullFeatures)
        return fail(WPT_ERROR_UNSUPPORTED,
            "the differentiable renderer supports neither environment maps, punctual lights nor texture bindings");
    const wpt_diff_config& c = *config;
    if (c.width == 0 || c.height == 0 || c.samples_per_pixel == 0 || c.max_depth == 0)
        return fail(WPT_ERROR_INVALID_ARGUMENT, "image size, samples per pixel and depth must be greater than zero");
    if (c.block_size == 0 || c.block_size > 512 || c.block_size % 64 != 0)
        return fail(WPT_ERROR_INVALID_ARGUMENT, "Block size must be a multiple of 64 between 64 and 512");
    if (c.rng_strategy > WPT_RNG_PCG32) return fail(WPT_ERROR_INVALID_ARGUMENT, "Unknown RNG strategy");
    uint64_t paths = static_cast<uint64_t>(c.width) * c.height * c.samples_per_pixel;
    if (paths >= (1ull << 31)) return fail(WPT_ERROR_INVALID_ARGUMENT, "too many paths per forward pass");
    bool textured = c.textured_material != WPT_INVALID_INDEX;
    if (textured && (c.texture_width == 0 || c.texture_height == 0 || c.textured_material >= scene->host.materials.size()))
        return fail(WPT_ERROR_INVALID_ARGUMENT, "texture needs a size and an existing material");
    if (textured && static_cast<uint64_t>(c.texture_width) * c.texture_height * 3 >= (1ull << 31))
        return fail(WPT_ERROR_INVALID_ARGUMENT, "texture too large");
    wpt_diff* d = new (std::nothrow) wpt_diff();
    if (!d) return fail(WPT_ERROR_OUT_OF_MEMORY, "host allocation failed");
    d->scene = scene;
    d->config = c;
    auto cleanup = [&](int status) {
        delete d;
        return status;
    };
    // Parameter slots of the static triangles and lights.
    const HostSceneData& h = scene->host;
    std::vector<uint32_t> staticSlot(h.staticTriangles.size(), kInvalid);
    std::vector<uint32_t> lightSlot(std::max<size_t>(h.lights.size(), 1), kInvalid);
    for (size_t i = 0; i < h.staticTriangles.size(); ++i) {
        const TriShade& t = h.staticTriangles[i];
        uint32_t type = h.materials[t.material].type;
        if (type == kEmissive) {
            staticSlot[i] = d->emissionCount;
            if (t.lightIndex != kInvalid) lightSlot[t.lightIndex] = d->emissionCount;
            ++d->emissionCount;
        } else if (type != kDielectric) {
            staticSlot[i] = d->roughnessCount++;
        }
    }
    d->textureFloats = textured ? c.texture_width * c.texture_height * 3 : 0;
    hipError_t e;
#define D_HIP(call)                                            \
    do {                                                       \
        e = (call);                                            \
        if (e != hipSuccess) return cleanup(hipFail(e, #call)); \
    } while (0)
    D_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    for (hipEvent_t& ev : d->marks) D_HIP(hipEventCreate(&ev));
    DiffContext& x = d->ctx;
    x.scene = scene->dev;
    x.camera.position = Vec3(camera->position[0], camera->position[1], camera->position[2]);
    x.camera.forward = Vec3(camera->forward[0], camera->forward[1], camera->forward[2]);
    x.camera.right = Vec3(camera->right[0], camera->right[1], camera->right[2]);
    x.camera.up = Vec3(camera->up[0], camera->up[1], camera->up[2]);
    x.camera.verticalFov = camera->vertical_fov;
    x.camera.aspectRatio = camera->aspect_ratio;
    x.layout.texWidth = textured ? c.texture_width : 0;
    x.layout.texHeight = textured ? c.texture_height : 0;
    x.layout.texturedMaterial = textured ? c.textured_material : kInvalid;
    x.layout.roughnessCount = d->roughnessCount;
    x.layout.emissionCount = d->emissionCount;
    const uint32_t* p = nullptr;
    D_HIP(d->arena.upload(staticSlot, &p));
    x.layout.staticSlot = p;
    D_HIP(d->arena.upload(lightSlot, &p));
    x.layout.lightSlot = p;
    x.width = c.width;
    x.height = c.height;
    x.spp = c.samples_per_pixel;
    x.maxDepth = c.max_depth;
    x.rngStrategy = c.rng_strategy;
    void* mem = nullptr;
    D_HIP(d->arena.alloc(sizeof(float) * 3 * static_cast<size_t>(paths), &mem));
    x.pathRadiance = static_cast<float*>(mem);
    D_HIP(d->arena.alloc(sizeof(uint32_t), &mem));
    x.cursor = static_cast<uint32_t*>(mem);
    int forwardPerCu = 0, adjointPerCu = 0, device = 0, cus = 0;
    D_HIP(diffOccupancy(c.block_size, &forwardPerCu, &adjointPerCu));
    if (forwardPerCu <= 0 || adjointPerCu <= 0)
        return cleanup(fail(WPT_ERROR_UNSUPPORTED, "block size cannot launch the differentiable kernels"));
    D_HIP(hipGetDevice(&device));
    D_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    uint64_t pathBlocks = (paths + c.block_size - 1) / c.block_size;
    d->forwardBlocks = static_cast<uint32_t>(std::min<uint64_t>(pathBlocks, static_cast<uint64_t>(forwardPerCu) * cus));
    d->adjointBlocks = static_cast<uint32_t>(std::min<uint64_t>(pathBlocks, static_cast<uint64_t>(adjointPerCu) * cus));
    x.spillThreads = std::max(d->forwardBlocks, d->adjointBlocks) * c.block_size;
    D_HIP(d->arena.alloc(sizeof(uint2) * kSpillEntriesTotal * static_cast<size_t>(x.spillThreads), &mem));
    x.spill = static_cast<uint2*>(mem);
#undef D_HIP
    d->stats.paths = paths;
    *out = d;
    return WPT_OK;
}

WPT_API int wpt_diff_destroy(wpt_diff* d) {
    delete d;
    return WPT_OK;
}

WPT_API int wpt_diff_get_layout(const wpt_diff* d, wpt_diff_layout* layout) {
    if (!d || !layout) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    layout->texture_floats = d->textureFloats;
    layout->roughness_count = d->roughnessCount;
    layout->emission_count = d->emissionCount;
    layout->parameter_count = d->textureFloats + d->roughnessCount + d->emissionCount;
    return WPT_OK;
}

WPT_API int wpt_diff_forward(wpt_diff* d, const float* params, uint32_t sampleStart, float* image) {
    if (!d || !params || !image) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    if (static_cast<uint64_t>(sampleStart) + d->config.samples_per_pixel > 0xffffffffull)
        return fail(WPT_ERROR_INVALID_ARGUMENT, "sample range out of range");
    DiffContext c = d->ctx;
    c.params.texture = params;
    c.params.roughness = params + d->textureFloats;
    c.params.emission = params + d->textureFloats + d->roughnessCount;
    c.sampleStart = sampleStart;
    c.image = image;
    WPT_HIP(hipEventRecord(d->marks[0], d->stream));
    WPT_HIP(enqueueDiffForward(c, d->forwardBlocks, d->config.block_size, d->stream));
    WPT_HIP(hipEventRecord(d->marks[1], d->stream));
    WPT_HIP(hipStreamSynchronize(d->stream));
    float ms = 0.0f;
    WPT_HIP(hipEventElapsedTime(&ms, d->marks[0], d->marks[1]));
    d->stats.forward_milliseconds = ms;
    d->forwardParams = params;
    d->ctx.sampleStart = sampleStart;
    return WPT_OK;
}

WPT_API int wpt_diff_adjoint(wpt_diff* d, const float* params, const float* imageGrad, float* paramGrad) {
    if (!d || !params || !imageGrad || !paramGrad) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    if (params != d->forwardParams)
        return fail(WPT_ERROR_INVALID_ARGUMENT, "the adjoint replays the last forward: run wpt_diff_forward with these parameters first");
    DiffContext c = d->ctx;
    c.params.texture = params;
    c.params.roughness = params + d->textureFloats;
    c.params.emission = params + d->textureFloats + d->roughnessCount;
    c.imageGrad = imageGrad;
    c.grads.texture = paramGrad;
    c.grads.roughness = paramGrad + d->textureFloats;
    c.grads.emission = paramGrad + d->textureFloats + d->roughnessCount;
    size_t count = static_cast<size_t>(d->textureFloats) + d->roughnessCount + d->emissionCount;
    WPT_HIP(hipEventRecord(d->marks[2], d->stream));
    WPT_HIP(enqueueDiffAdjoint(c, d->adjointBlocks, d->config.block_size, count, d->stream));
    WPT_HIP(hipEventRecord(d->marks[3], d->stream));
    WPT_HIP(hipStreamSynchronize(d->stream));
    float ms = 0.0f;
    WPT_HIP(hipEventElapsedTime(&ms, d->marks[2], d->marks[3]));
    d->stats.adjoint_milliseconds = ms;
    return WPT_OK;
}

WPT_API int wpt_diff_adam(wpt_diff* d, float* params, const float* grad, float* m, float* v, uint32_t step, float lr,
    float beta1, float beta2, float epsilon) {
    if (!d || !params || !grad || !m || !v) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    if (step == 0) return fail(WPT_ERROR_INVALID_ARGUMENT, "Adam steps count from 1");
    AdamStep a{};
    a.params = params;
    a.grad = grad;
    a.m = m;
    a.v = v;
    a.count = static_cast<size_t>(d->textureFloats) + d->roughnessCount + d->emissionCount;
    a.textureFloats = d->textureFloats;
    a.roughnessCount = d->roughnessCount;
    a.lr = lr;
    a.beta1 = beta1;
    a.beta2 = beta2;
    a.epsilon = epsilon;
    a.bias1 = static_cast<float>(1.0 - std::pow(static_cast<double>(beta1), static_cast<double>(step)));
    a.bias2 = static_cast<float>(1.0 - std::pow(static_cast<double>(beta2), static_cast<double>(step)));
    a.textureRange[0] = 0.0f;
    a.textureRange[1] = 1.0f;
    a.roughnessRange[0] = 0.02f;
    a.roughnessRange[1] = 1.0f;
    a.emissionRange[0] = 0.0f;
    a.emissionRange[1] = std::numeric_limits<float>::infinity();
    // The kept paths belong to the old values from here on.
    if (params == d->forwardParams) d->forwardParams = nullptr;
    WPT_HIP(enqueueAdam(a, d->stream));
    WPT_HIP(hipStreamSynchronize(d->stream));
    return WPT_OK;
}

WPT_API int wpt_diff_last_stats(const wpt_diff* d, wpt_diff_stats* stats) {
    if (!d || !stats) return fail(WPT_ERROR_INVALID_ARGUMENT, "null argument");
    *stats = d->stats;
    return WPT_OK;
}

}  // extern "C"
