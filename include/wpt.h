/*
 * wpt.h - C ABI of the MI355X wavefront path tracer (libwpt.so).
 *
 * Plain C structs, plain pointers and sizes.  Every call returns a status
 * code (WPT_OK == 0) instead of exiting the process; wpt_last_error() gives
 * the message of the last failure on the calling thread.  Handles are opaque
 * and owned by the caller, who destroys a scene only after every renderer
 * created from it.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   uploadScene / destroyDeviceScene   src/scene/scene.cu:336-420, include/scene/scene.cuh:241
 *   RenderDriver(...)::run             include/renderer/render_driver.cuh:52-59, src/renderer/render_driver.cu:451-482
 *   RenderTaskRenderer::render         include/renderer/render_task.cuh:21-29, src/renderer/render_task.cu:26-114
 *   traceRay / isOccluded              include/renderer/device_traversal.cuh:1035-1184
 *   appendWarpCompacted                include/renderer/queue_compaction.cuh:8-29
 */
#ifndef WPT_H
#define WPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WPT_API __attribute__((visibility("default")))

#define WPT_ABI_VERSION 2

/* Status codes. */
#define WPT_OK 0
#define WPT_ERROR_INVALID_ARGUMENT 1
#define WPT_ERROR_DEVICE 2
#define WPT_ERROR_OUT_OF_MEMORY 3
#define WPT_ERROR_UNSUPPORTED 4
#define WPT_ERROR_INVALID_TASK 5

/* AccelerationMode, reference include/scene/scene.cuh:21. */
#define WPT_ACCEL_BVH 0u
#define WPT_ACCEL_DIRECT 1u

/* MaterialType, reference include/scene/material.cuh:18-24. */
#define WPT_MATERIAL_DIFFUSE 0u
#define WPT_MATERIAL_METAL 1u
#define WPT_MATERIAL_DIELECTRIC 2u
#define WPT_MATERIAL_EMISSIVE 3u
#define WPT_MATERIAL_SPECULAR_GLOSSINESS 4u

/* RngStrategy, reference include/core/rng_types.hpp:5-8. */
#define WPT_RNG_XORSHIFT32 0u
#define WPT_RNG_PCG32 1u

/* Scheduler: the hipGraph-captured staged wavefront or the persistent
 * cooperative kernel (reference render_driver.cu:264-348), or a persistent
 * kernel with path regeneration and no barrier between bounces (no
 * reference counterpart): each lane renders its pixel's samples one bounce
 * per iteration and starts the next sample when a path ends.  All three
 * produce bit-identical framebuffers. */
#define WPT_MODE_GRAPH 0u
#define WPT_MODE_PERSISTENT 1u
#define WPT_MODE_REGENERATE 2u

/* Closest-hit categories, in the reference's test order
 * (device_traversal.cuh:1129-1149). */
#define WPT_HIT_NONE 0u
#define WPT_HIT_SPHERE 1u
#define WPT_HIT_STATIC_TRIANGLE 2u
#define WPT_HIT_INSTANCE_TRIANGLE 3u
#define WPT_HIT_FLAT_TRIANGLE 4u

#define WPT_INVALID_INDEX 0xffffffffu

/* Sphere + placement; reference Sphere (geometry.cuh:6-11) and SceneSphere. */
typedef struct wpt_sphere {
    float center[3];
    float radius;
    uint32_t material;
    uint32_t acceleration; /* WPT_ACCEL_BVH | WPT_ACCEL_DIRECT */
} wpt_sphere;

/* Triangle; the geometry and shading subset of reference Triangle
 * (geometry.cuh:13-49).  flags: bit0 hasVertexNormals, bit1 hasTexcoords. */
typedef struct wpt_triangle {
    float v0[3], v1[3], v2[3];
    float n0[3], n1[3], n2[3];
    float uv0[2], uv1[2], uv2[2];
    float tangent[3];
    float bitangent[3];
    uint32_t material;
    uint32_t flags;
    /* Second texture-coordinate set, per-vertex tangents with handedness
     * signs, and the second set's face tangent frame (Triangle's uv1_*,
     * vertexTangent*, vertexTangentSign*, tangent1, bitangent1). */
    float uv1_0[2], uv1_1[2], uv1_2[2];
    float vertex_tangent0[3], vertex_tangent1[3], vertex_tangent2[3];
    float vertex_tangent_sign[3];
    float tangent1[3];
    float bitangent1[3];
} wpt_triangle;

#define WPT_TRIANGLE_HAS_NORMALS 1u
#define WPT_TRIANGLE_HAS_TEXCOORDS 2u
#define WPT_TRIANGLE_HAS_TEXCOORDS1 4u
#define WPT_TRIANGLE_HAS_VERTEX_TANGENTS 8u

/* Mesh = a range of wpt_scene_desc.mesh_triangles (reference MeshAsset). */
typedef struct wpt_mesh {
    uint32_t first_triangle;
    uint32_t triangle_count;
} wpt_mesh;

/* Affine instance transform with its explicit inverse basis
 * (reference InstanceTransform, scene.cuh:29-55). */
typedef struct wpt_transform {
    float local_to_world[3][3]; /* columns X, Y, Z */
    float world_to_local[3][3]; /* columns X, Y, Z */
    float translation[3];
} wpt_transform;

/* Reference SceneInstance (scene.cuh:104-109). */
typedef struct wpt_instance {
    uint32_t mesh;
    uint32_t material_override; /* WPT_INVALID_INDEX: keep triangle materials */
    uint32_t acceleration;      /* WPT_ACCEL_BVH | WPT_ACCEL_DIRECT */
    uint32_t reserved;
    wpt_transform transform;
} wpt_instance;

/* Reference Material (material.cuh:31-70), untextured subset. */
typedef struct wpt_material {
    uint32_t type;
    float albedo[3];
    float emission[3];
    float roughness;
    float ior;
    float metallic;
    float opacity;
    float alpha_cutoff;
    uint32_t alpha_masked;
    float attenuation_color[3];
    float attenuation_distance;
    float volume_density;
    float specular[3];
    float glossiness;
} wpt_material;

/* RGBA8 texture (reference ObjTexture, obj_loader.hpp:9-13), rows top to
 * bottom.  Read as normalized floats with bilinear filtering and wrapping
 * in both directions; a lookup at (u, v) reads image position (u, 1 - v). */
typedef struct wpt_texture {
    uint32_t width, height;
    const uint8_t* rgba; /* width * height * 4 bytes */
} wpt_texture;

/* Reference TextureBinding (material.cuh:14-21). */
typedef struct wpt_texture_binding {
    uint32_t texture;  /* index into wpt_scene_desc.textures; WPT_INVALID_INDEX: none */
    uint32_t texcoord; /* 0: first set, 1: second set */
    float offset[2];
    float scale[2];
    float rotation;    /* radians */
    uint32_t has_transform;
} wpt_texture_binding;

/* Texture bindings and the glTF workflow fields of one material (the rest
 * of reference Material, material.cuh:31-70).  Tints multiply the texel:
 * albedo x albedo texture (sRGB 2.2, or the glTF sRGB curve in the
 * specular-glossiness workflow; alpha x opacity is the coverage), emission
 * x emissive texture, metallic x blue and roughness x green of the
 * metallic-roughness texture, specular x specular-glossiness rgb and
 * glossiness x its alpha.  The glTF workflow adds the KHR_materials_specular
 * and KHR_materials_clearcoat lobes. */
typedef struct wpt_material_ext {
    wpt_texture_binding albedo_texture;
    wpt_texture_binding metallic_roughness_texture;
    wpt_texture_binding emissive_texture;
    wpt_texture_binding normal_texture;
    wpt_texture_binding thickness_texture;
    wpt_texture_binding specular_glossiness_texture;
    wpt_texture_binding specular_weight_texture;
    wpt_texture_binding specular_color_texture;
    wpt_texture_binding clearcoat_texture;
    wpt_texture_binding clearcoat_roughness_texture;
    wpt_texture_binding clearcoat_normal_texture;
    float normal_scale;
    uint32_t specular_glossiness_workflow;
    uint32_t gltf_pbr_workflow;
    float specular_weight;
    float specular_color[3];
    float clearcoat;
    float clearcoat_roughness;
    float clearcoat_normal_scale;
    float dispersion;
} wpt_material_ext;

/* Punctual light (reference PunctualLight, punctual_light.cuh:6-16; glTF
 * KHR_lights_punctual).  Directional lights shine along `direction`;
 * positional ones fall off with the inverse square, a smooth range window
 * (range 0: unlimited) and, for spots, a squared smoothstep between the
 * cone cosines. */
#define WPT_LIGHT_POINT 0u
#define WPT_LIGHT_SPOT 1u
#define WPT_LIGHT_DIRECTIONAL 2u

typedef struct wpt_punctual_light {
    uint32_t type;
    float position[3];
    float direction[3];  /* direction of emission */
    float intensity[3];  /* linear RGB x intensity */
    float range;
    float inner_cone_cos;
    float outer_cone_cos;
} wpt_punctual_light;

typedef struct wpt_scene_desc {
    const wpt_sphere* spheres;
    uint32_t sphere_count;
    const wpt_triangle* static_triangles; /* world space, tested without a BVH */
    uint32_t static_triangle_count;
    const wpt_triangle* flat_triangles;   /* world space, own BVH2 (reference scene.triangles) */
    uint32_t flat_triangle_count;
    const wpt_triangle* mesh_triangles;   /* mesh-local space */
    uint32_t mesh_triangle_count;
    const wpt_mesh* meshes;
    uint32_t mesh_count;
    const wpt_instance* instances;
    uint32_t instance_count;
    const wpt_material* materials;
    uint32_t material_count;
    uint32_t black_background;            /* reference HostScene.blackBackground */
    uint32_t ignore_direct_light_occlusion;
    /* Environment map (reference HostEnvironment, environment.cpp:10-64, and
     * uploadEnvironmentMap, scene.cu:308-334): equirectangular RGB radiance,
     * environment_width x environment_height, row 0 at +Y, column width/2 at
     * +X, column 3*width/4 at +Z.  NULL: no map (sky or black background).
     * Radiance is scaled by environment_intensity; the map turns about +Y by
     * environment_rotation_degrees. */
    const float* environment_rgb;
    uint32_t environment_width, environment_height;
    float environment_intensity;
    float environment_rotation_degrees;
    /* Punctual lights in any order; the upload moves directional lights
     * first, keeping the order within each kind (scene.cu:349-355). */
    const wpt_punctual_light* punctual_lights;
    uint32_t punctual_light_count;
    /* Textures and per-material texture bindings (NULL: every material
     * untextured; else material_count entries), and the normal-map switch
     * and minimum tangent-space cosine (reference Scene.useNormalMaps,
     * normalMapMinimumCosine; config defaults true and 0). */
    const wpt_texture* textures;
    uint32_t texture_count;
    const wpt_material_ext* material_ext;
    uint32_t use_normal_maps;
    float normal_map_minimum_cosine;
    /* Builder of the mesh BLASes: WPT_BVH_BUILDER_HOST_SAH (0, binned SAH on
     * the host) or WPT_BVH_BUILDER_DEVICE_PLOC (clustering on the GPU; the
     * reference builds on the GPU with NexusBVH, scene.cu:123-239), with
     * bvh_search_radius its neighbour radius (0: 16; at most 64).  The
     * device builder needs a device scene (wpt_scene_create). */
    uint32_t bvh_builder;
    uint32_t bvh_search_radius;
    /* Accelerator of flat_triangles (built when the scene has no meshes;
     * reference ObjAccelerationPolicy, scene.cu:241-294): WPT_FLAT_*. */
    uint32_t flat_accelerator;
} wpt_scene_desc;

#define WPT_BVH_BUILDER_HOST_SAH 0u
#define WPT_BVH_BUILDER_DEVICE_PLOC 1u

/* Flat-triangle accelerators.  Every one gives the same hits (exact
 * distance ties are ordered by category, instance and primitive). */
#define WPT_FLAT_SAH_BVH2 0u           /* binned SAH, one triangle per leaf (the nexus_bvh2 policy's tree kind) */
#define WPT_FLAT_MEDIAN_BVH2 1u        /* object median on the longest centroid axis (median_bvh, bvh.cpp:49-109) */
#define WPT_FLAT_SPATIAL_SPLIT_BVH2 2u /* spatial splits, triangles referenced on both sides of a cut
                                          (spatial_split_bvh, bvh_sbvh.cpp:94-219) */
#define WPT_FLAT_WIDE8 3u              /* eight-wide nodes collapsed from the SAH tree (nexus_bvh8,
                                          device_traversal.cuh:893-963) */

/* Reference Camera (include/scene/camera.cuh:7-27). */
typedef struct wpt_camera {
    float position[3];
    float forward[3];
    float right[3];
    float up[3];
    float vertical_fov; /* radians */
    float aspect_ratio;
} wpt_camera;

/* Reference RenderConfig render fields (include/config.hpp:11-37). */
typedef struct wpt_render_config {
    uint32_t width, height;
    uint32_t tiled;
    uint32_t tile_width, tile_height;
    uint32_t max_depth;
    uint32_t rr_start_depth;
    uint32_t samples_per_pixel;
    uint32_t block_size;            /* multiple of 64, at most 512 */
    uint32_t rng_strategy;          /* WPT_RNG_* */
    float indirect_clamp;           /* 0 disables */
    uint32_t mode;                  /* WPT_MODE_* */
    uint32_t persistent_blocks;     /* 0: occupancy-derived */
    uint32_t intersection_debug;
    uint32_t shading_normal_debug;
    uint32_t profile_queues;
    /* Photon-mapped caustics, built once when the renderer is created
     * (reference RenderDriver::buildPhotonMap, render_driver.cu:180-234, and
     * config.hpp:18-22).  0 photons: off.  The photons start on the
     * triangle lights and are kept where they reach a diffuse surface after
     * at least one mirror or dielectric bounce; diffuse vertices then add
     * albedo x the gathered irradiance (caustic_gather).  Spectral photons
     * carry one RGB channel each, for dispersive dielectrics.  A zero grid
     * box or resolution means the reference's (-2.5,-1,-2.5)..(2.5,3,2.5)
     * and 64. */
    uint32_t caustic_photons;
    uint32_t caustic_max_depth;      /* 0: 8 */
    float caustic_gather_radius;     /* 0: 0.06 */
    uint32_t caustic_gather;
    uint32_t caustic_spectral;
    uint32_t caustic_grid_resolution;
    float caustic_grid_min[3];
    float caustic_grid_max[3];
    /* Samples the staged and cooperative persistent schedulers render per
     * pass of a tile (0: as many as keep a pass within 2^27 paths, 64 at
     * 1920x1080, half that for scenes shaded with radiance events; at most
     * 64; never more than samples_per_pixel).  Path memory grows with it,
     * about 260 B per path of a pass (about 435 B with radiance events for
     * a scene whose vertices store up to 4 of them).  Each pixel
     * still adds its samples to the framebuffer in sample order, so images
     * do not depend on it. */
    uint32_t samples_per_pass;
} wpt_render_config;

/* Reference RenderTask / RenderTaskId / DistributedTile
 * (include/distributed/task.hpp:6-29). */
typedef struct wpt_task {
    uint64_t job_fingerprint;
    uint32_t tile_ordinal;
    uint32_t sample_start;
    uint32_t sample_count;
    uint32_t x, y, width, height;
} wpt_task;

/* Closest-hit record: a defined identity for every hit (the reference Hit
 * carries none).  kind: WPT_HIT_*; instance: instance index for
 * WPT_HIT_INSTANCE_TRIANGLE else WPT_INVALID_INDEX; primitive: sphere index,
 * static/flat triangle index, or mesh-local triangle index. */
typedef struct wpt_hit {
    float t;
    float u, v;
    uint32_t kind;
    uint32_t instance;
    uint32_t primitive;
    uint32_t material;
    uint32_t light_index;
} wpt_hit;

/* Per-render counters (reference --profile-queues, render_driver.cu:393-400). */
typedef struct wpt_stats {
    uint64_t closest_queries;
    uint64_t shadow_queries;
    uint64_t samples;          /* pixels x spp rendered */
    double trace_milliseconds; /* device time of the last render call */
    uint32_t queue_counts[64]; /* last launch, per bounce */
    uint32_t queue_count_depth;
    uint32_t caustic_photons_stored;    /* photons in the caustic grid */
    double caustic_build_milliseconds;  /* photon trace + grid build at creation */
} wpt_stats;

typedef struct wpt_scene wpt_scene;
typedef struct wpt_renderer wpt_renderer;

/* ---- library ---------------------------------------------------------- */
WPT_API int wpt_abi_version(void);
WPT_API const char* wpt_last_error(void);
/* Number of visible gfx devices (0 without a GPU); never fails. */
WPT_API int wpt_device_count(void);

/* ---- scene: reference uploadScene / destroyDeviceScene ---------------- */
/* Partitions spheres/instances by AccelerationMode, builds BLAS per mesh with
 * a BVH instance and the TLAS over typed leaves (scene.cu:123-239), builds the
 * triangle-light alias table (scene.cu:82-121) and uploads to the device. */
WPT_API int wpt_scene_create(const wpt_scene_desc* desc, wpt_scene** out);
/* Same as wpt_scene_create but host-only: no device upload.  Used to inspect
 * partitions and BVHs on machines without a GPU. */
WPT_API int wpt_scene_create_host(const wpt_scene_desc* desc, wpt_scene** out);
WPT_API int wpt_scene_destroy(wpt_scene* scene);

/* Scene inspection (partition lists, light table, BVH export). */
typedef struct wpt_scene_info {
    uint32_t sphere_count, direct_sphere_count;
    uint32_t instance_count, direct_instance_count;
    uint32_t tlas_leaf_count, tlas_node_count;
    uint32_t blas_count;
    uint32_t light_count;
    uint32_t static_triangle_count, flat_triangle_count;
    uint32_t flat_node_count;
    uint32_t material_count;
    uint32_t bvh_builder;           /* WPT_BVH_BUILDER_* of the mesh BLASes */
    uint32_t max_blas_depth;        /* longest root-to-leaf path of any BLAS */
    double blas_build_milliseconds; /* BLAS construction (host wall or device) */
    uint32_t flat_accelerator;      /* WPT_FLAT_* of the flat-triangle tree */
    uint32_t flat_leaf_slots;       /* leaf slots (above the triangle count after spatial splits) */
    uint32_t flat_wide_node_count;  /* eight-wide nodes (WPT_FLAT_WIDE8) */
    uint32_t max_stack_depth;       /* deepest traversal stack any query can need */
} wpt_scene_info;
WPT_API int wpt_scene_info_get(const wpt_scene* scene, wpt_scene_info* info);
/* which: 0 direct spheres, 1 direct instances, 2 TLAS leaves packed as
 * (type << 31 | index) with type 1 = sphere, 0 = mesh instance. */
WPT_API int wpt_scene_get_list(const wpt_scene* scene, uint32_t which, uint32_t* out, uint32_t capacity,
                               uint32_t* count);
/* BVH2 in the reference's node convention (NXB::BVH2::Node: float bmin[3],
 * bmax[3], uint32 left, right; leaf iff left == 0xffffffff with right = the
 * primitive index; root = node_count - 1).  bvh: 0 TLAS, 1 flat triangles,
 * 2 + m BLAS of mesh m (node_count 0 when the mesh has no BLAS). */
typedef struct wpt_bvh2_node {
    float bmin[3];
    float bmax[3];
    uint32_t left, right;
} wpt_bvh2_node;
WPT_API int wpt_scene_export_bvh2(const wpt_scene* scene, uint32_t bvh, wpt_bvh2_node* out, uint32_t capacity,
                                  uint32_t* node_count);
/* Triangle lights (TriangleLight, geometry.cuh:51-59) as 16 floats each:
 * v0,v1,v2,normal (12), area, material (bits), selectionPdf, pad; and the
 * alias table as (probability, alias bits) pairs. */
WPT_API int wpt_scene_export_lights(const wpt_scene* scene, float* lights16, float* alias2, uint32_t capacity,
                                    uint32_t* count);
/* Environment sampling tables (reference HostEnvironment probabilities and
 * alias, environment.cpp:10-64): width*height floats, and width*height
 * (probability, index bits) pairs; returns WPT_OK with *count = 0 when the
 * scene has no map. */
WPT_API int wpt_scene_export_environment(const wpt_scene* scene, float* probabilities, float* alias2,
                                         uint32_t capacity, uint32_t* count);
/* Light offset per instance (reference MeshInstance.lightOffset) and light
 * index per static/mesh triangle as assigned by the scene build. */
WPT_API int wpt_scene_export_light_indices(const wpt_scene* scene, uint32_t* instance_light_offset,
                                           uint32_t* static_light_index, uint32_t* mesh_light_index);

/* ---- traversal: reference traceRay / isOccluded ----------------------- */
/* Device pointers; rays are SoA: origin x,y,z then direction x,y,z, each an
 * array of n floats (rays6 = 6 consecutive arrays of n).  stream may be NULL. */
WPT_API int wpt_trace_closest(const wpt_scene* scene, const float* d_rays6, uint32_t n, wpt_hit* d_hits,
                              void* stream);
WPT_API int wpt_trace_occluded(const wpt_scene* scene, const float* d_rays6, const float* d_tmax, uint32_t n,
                               uint8_t* d_occluded, void* stream);

/* ---- queue compaction: reference appendWarpCompacted ------------------ */
/* Order-preserving wave64 compaction: out[k] = items[i] for the k-th i with
 * flags[i] != 0, writes bounded by capacity; *d_count receives the number of
 * active items (may exceed capacity, as in the reference).  Device pointers. */
WPT_API int wpt_compact(const uint32_t* d_items, const uint8_t* d_flags, uint32_t n, uint32_t* d_out,
                        uint32_t capacity, uint32_t* d_count, void* stream);

/* wpt_compact without any device allocation, for callers that cannot
 * allocate (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same contract and
 * output as wpt_compact for n <= 16777216 items (larger inputs are rejected);
 * its scratch is a fixed 256 KiB device array that each GPU holds until the
 * process ends, so calls that overlap on one GPU race.  Asynchronous: the
 * work is queued on `stream` and nothing waits for it. */
WPT_API int wpt_compact_noalloc(const uint32_t* d_items, const uint8_t* d_flags, uint32_t n, uint32_t* d_out,
                                uint32_t capacity, uint32_t* d_count, void* stream);

/* ---- renderer: reference RenderDriver / RenderTaskRenderer ------------ */
WPT_API int wpt_renderer_create(wpt_scene* scene, const wpt_camera* camera, const wpt_render_config* config,
                                uint64_t job_fingerprint, wpt_renderer** out);
WPT_API int wpt_renderer_destroy(wpt_renderer* renderer);
/* Clears the framebuffer and sample counter (resetRenderSession). */
WPT_API int wpt_renderer_reset(wpt_renderer* renderer);
/* Renders samples [first, first+count) over all tiles, accumulating into the
 * device framebuffer (renderSamples, render_driver.cu:354-401). */
WPT_API int wpt_renderer_render_samples(wpt_renderer* renderer, uint32_t first_sample, uint32_t count);
/* Copies the unnormalized framebuffer (w*h*3 floats) to host memory and
 * reports the accumulated sample count (writeFinalImage's source data). */
WPT_API int wpt_renderer_read_framebuffer(wpt_renderer* renderer, float* host_rgb, uint32_t* sample_count);
/* Restores a framebuffer of unnormalized sums (w*h*3 floats) and the sample
 * counter, so that rendering resumes at sample_count (the device half of
 * loadRenderSessionCheckpoint, render_checkpoint.cpp:187-192). */
WPT_API int wpt_renderer_load_framebuffer(wpt_renderer* renderer, const float* host_rgb, uint32_t sample_count);
/* Full render: reset + all spp + read back (RenderDriver::run). */
WPT_API int wpt_render_frame(wpt_renderer* renderer, float* host_rgb, uint32_t* sample_count);
/* One task: validates fingerprint/tile/sample range, clears the tile,
 * renders the sample range, returns unnormalized tile sums (w*h*3). */
WPT_API int wpt_render_task(wpt_renderer* renderer, const wpt_task* task, float* host_tile_rgb);
WPT_API int wpt_renderer_last_stats(const wpt_renderer* renderer, wpt_stats* stats);
/* Renders samples [first, first+count) with every kernel launch bracketed by
 * HIP events on the renderer's stream (no graph), accumulating device time
 * per kernel into kernel_ms[7] and launch counts into launches[7]: 0 primary
 * rays, 1 queue reset, 2 closest-hit trace, 3 shade, 4 persistent, 5 shadow
 * rays (with radiance events: the event rays and the pass adding the
 * unblocked contributions), 6 pass resolve (a pass's samples added to the
 * framebuffer in sample order). */
WPT_API int wpt_renderer_profile_kernels(wpt_renderer* renderer, uint32_t first_sample, uint32_t count,
                                         double* kernel_ms, uint32_t* launches);

/* Launch shape and resources of the renderer's dominant kernel (the
 * regenerating kernel, the cooperative persistent kernel, or the staged
 * closest-hit trace), as the occupancy calculation sees them. */
typedef struct wpt_launch_info {
    uint32_t mode;                 /* WPT_MODE_* of the renderer */
    uint32_t block_size;           /* threads per workgroup */
    uint32_t blocks;               /* workgroups of a persistent launch (0: one item per thread) */
    uint32_t blocks_per_cu;        /* resident workgroups per compute unit */
    uint32_t compute_units;
    uint32_t registers;            /* VGPRs per lane */
    uint32_t scratch_bytes;        /* private (spill) bytes per lane */
    uint32_t lds_bytes_per_block;  /* static + dynamic LDS of one workgroup */
    uint32_t lds_bytes_per_cu;     /* LDS available on a compute unit */
    uint32_t waves_per_simd;       /* resident waves per SIMD at blocks_per_cu */
} wpt_launch_info;
WPT_API int wpt_renderer_launch_info(const wpt_renderer* renderer, wpt_launch_info* info);

/* ---- batch: many tasks of several scenes in one persistent launch -------
 * No reference counterpart: the reference renders its tasks one at a time
 * (RenderTaskRenderer::render, render_task.cu:26-114); a batch takes the
 * same tasks, of uneven tile size, sample range and scene, in one launch. */
typedef struct wpt_batch wpt_batch;

typedef struct wpt_batch_task {
    uint32_t scene;            /* index into the scenes of wpt_batch_create */
    wpt_camera camera;
    uint32_t width, height;    /* frame the tile belongs to (pixel ids, RNG keys) */
    uint32_t x, y, tile_width, tile_height;
    uint32_t sample_start, sample_count;
    uint32_t max_depth, rr_start_depth, rng_strategy;
    float indirect_clamp;
} wpt_batch_task;

typedef struct wpt_batch_config {
    uint32_t block_size;       /* multiple of 64, at most 512 */
    uint32_t blocks;           /* persistent grid; 0 = every resident workgroup */
    /* WPT_BATCH_STAGED (0): the tasks of each scene (and camera, frame and
     * path parameters) run as staged wavefront passes of up to 2^26 paths
     * over many tasks at once; WPT_BATCH_REGENERATE: one persistent launch
     * over every task (per-task latency from the kernel's clock).  Outputs
     * are identical. */
    uint32_t scheduler;
    /* Staged scheduler: paths per pass (0: 2^26; at most 2^26).  A task's
     * tile must not hold more pixels than a pass holds paths. */
    uint32_t pass_paths;
} wpt_batch_config;

#define WPT_BATCH_STAGED 0u
#define WPT_BATCH_REGENERATE 1u

typedef struct wpt_batch_stats {
    uint64_t closest_queries;
    uint64_t shadow_queries;
    uint64_t samples;          /* pixel samples rendered */
    double milliseconds;       /* device time of the launch */
    uint32_t block_size, blocks;
} wpt_batch_stats;

/* The scenes must outlive the batch. */
WPT_API int wpt_batch_create(wpt_scene* const* scenes, uint32_t scene_count, const wpt_batch_config* config,
                             wpt_batch** out);
WPT_API int wpt_batch_destroy(wpt_batch* batch);
/* Caustics of one batch scene (reference RenderDriver::buildPhotonMap,
 * render_driver.cu:180-234, with the photon fields of RenderConfig,
 * config.hpp:18-22): builds the scene's photon map from the caustic_* fields
 * and rng_strategy of `config`, exactly as wpt_renderer_create does for a
 * renderer with that configuration; with caustic_gather set, every task of
 * that scene then gathers it at diffuse vertices, so its sums equal
 * wpt_render_task's on such a renderer.  caustic_photons 0 removes the map.
 * photons_stored and build_ms (may be NULL) receive the photons kept in the
 * grid and the build's device time. */
WPT_API int wpt_batch_set_caustics(wpt_batch* batch, uint32_t scene, const wpt_render_config* config,
                                   uint32_t* photons_stored, double* build_ms);
/* Renders every task (staged passes, or one persistent launch).  outs[i]
 * (host memory) receives task i's unnormalized tile sums,
 * tile_width*tile_height*3 floats row-major, bit identical to
 * wpt_render_task on the same task.  task_latency_ms, when not NULL,
 * receives per task the time from the render's start to the end of its last
 * samples. */
WPT_API int wpt_batch_render(wpt_batch* batch, const wpt_batch_task* tasks, uint32_t n, float* const* outs,
                             double* task_latency_ms);
WPT_API int wpt_batch_last_stats(const wpt_batch* batch, wpt_batch_stats* stats);
/* Autotuner within the batch's scheduler: renders the tasks under each
 * candidate configuration -- staged: pass workgroups of 128, 256 or 512
 * lanes; persistent: workgroups of 64, 128, 256 or 512 lanes times a grid of
 * 1/4, 1/2, 3/4 or all of the resident workgroups -- keeps the fastest for
 * later renders and reports it in *chosen. */
WPT_API int wpt_batch_autotune(wpt_batch* batch, const wpt_batch_task* tasks, uint32_t n, wpt_batch_config* chosen);

/* ---- differentiable rendering: forward + path-replay adjoint -----------
 * No reference counterpart.  Parameters form one float vector laid out as
 *   [ texture: texture_height x texture_width x RGB (row j at v = (j+0.5)/h) |
 *     roughness: one GGX alpha per reflective static triangle, in order |
 *     emission: one scale per emissive static triangle, in order ].
 * The texture is the albedo of the static triangles of `textured_material`
 * (bilinear, clamp to edge).  Reflective static triangles are
 * albedo/pi + a GGX lobe (F0 0.04) sampled by cosine; other diffuse surfaces
 * are Lambertian; light selection and MIS do not depend on the parameters;
 * no Russian roulette, no clamp.  Image: width x height x RGB, the mean over
 * samples_per_pixel paths per pixel. */
typedef struct wpt_diff wpt_diff;

typedef struct wpt_diff_config {
    uint32_t width, height;
    uint32_t samples_per_pixel;
    uint32_t max_depth;
    uint32_t rng_strategy;
    uint32_t texture_width, texture_height;
    uint32_t textured_material;  /* WPT_INVALID_INDEX: no texture */
    uint32_t block_size;         /* multiple of 64, at most 512 */
} wpt_diff_config;

typedef struct wpt_diff_layout {
    uint32_t texture_floats;
    uint32_t roughness_count;
    uint32_t emission_count;
    uint32_t parameter_count;
} wpt_diff_layout;

typedef struct wpt_diff_stats {
    double forward_milliseconds;   /* device time of the last forward */
    double adjoint_milliseconds;   /* device time of the last adjoint */
    uint64_t paths;
} wpt_diff_stats;

/* The scene must be on the device and outlive the object. */
WPT_API int wpt_diff_create(wpt_scene* scene, const wpt_camera* camera, const wpt_diff_config* config,
                            wpt_diff** out);
WPT_API int wpt_diff_destroy(wpt_diff* diff);
WPT_API int wpt_diff_get_layout(const wpt_diff* diff, wpt_diff_layout* layout);
/* Renders samples [sample_start, sample_start + spp) of every pixel with the
 * parameters at d_params (device); writes the mean image to d_image
 * (device) and keeps every path's radiance for the adjoint. */
WPT_API int wpt_diff_forward(wpt_diff* diff, const float* d_params, uint32_t sample_start, float* d_image);
/* Path-replay adjoint of the last forward (same d_params, unchanged since;
 * an adjoint after a wpt_diff_adam step on those parameters is rejected until
 * a new forward has run):
 * d_param_grad (device, parameter_count floats) = dLoss/dparams given
 * d_image_grad = dLoss/dimage (device). */
WPT_API int wpt_diff_adjoint(wpt_diff* diff, const float* d_params, const float* d_image_grad, float* d_param_grad);
/* One Adam step (t = step, from 1) on device vectors, then each section is
 * clamped: texture [0, 1], roughness [0.02, 1], emission [0, inf). */
WPT_API int wpt_diff_adam(wpt_diff* diff, float* d_params, const float* d_grad, float* d_m, float* d_v, uint32_t step,
                          float learning_rate, float beta1, float beta2, float epsilon);
WPT_API int wpt_diff_last_stats(const wpt_diff* diff, wpt_diff_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* WPT_H */
