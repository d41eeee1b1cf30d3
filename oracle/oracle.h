/*
 * oracle.h - CPU restatement of the reference render path, used only as the
 * checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * It is never linked into the product library.
 *
 * It takes the same scene description as the product's C ABI (include/wpt.h)
 * plus the product's BVHs exported in the reference's node convention, and
 * re-derives everything else (AccelerationMode partition, light table and
 * alias, hit attributes, shading) following the reference sources.
 */
#ifndef WPT_ORACLE_H
#define WPT_ORACLE_H

#include <stdint.h>

#include "../include/wpt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene {
    wpt_scene_desc desc;
    /* BVHs in the reference convention (wpt_bvh2_node), from the product. */
    const wpt_bvh2_node* tlas_nodes;
    uint32_t tlas_node_count;
    const wpt_bvh2_node* flat_nodes;
    uint32_t flat_node_count;
    const wpt_bvh2_node* blas_nodes;   /* all meshes concatenated */
    const uint32_t* blas_node_offset;  /* per mesh */
    const uint32_t* blas_node_count;   /* per mesh, 0 = no BLAS */
} oracle_scene;

typedef struct oracle_stats {
    uint64_t closest_queries;
    uint64_t shadow_queries;
    uint64_t paths;
    /* Work of the reference algorithm, per query type: node bound tests,
     * triangle tests, sphere tests and mesh-instance entries. */
    uint64_t closest_node_tests, closest_triangle_tests, closest_sphere_tests, closest_instance_entries;
    uint64_t shadow_node_tests, shadow_triangle_tests, shadow_sphere_tests, shadow_instance_entries;
    uint32_t caustic_photons_stored; /* photons in the caustic grid (config.caustic_photons > 0) */
} oracle_stats;

typedef struct oracle_handle oracle_handle;

/* Builds the oracle's own scene tables (partition, lights, alias). */
int oracle_create(const oracle_scene* scene, oracle_handle** out);

/* Environment map of the scene (FP32): importance and alias tables
 * (alias2: probability, alias index bits per texel), texel solid angle, and
 * per-direction uv, pdf, radiance and texel index (null outputs skipped).
 * oracle_env_sample draws directions from FP64 alias candidates and
 * (coin, longitude, latitude) triples.  Both return 1 without a map. */
/* glTF metallic-roughness lobes (FP32, gltf_pbr.cuh).  `surface` holds 20
 * floats: color[3], normal[3], coatNormal[3], metallic, roughness, F0[3],
 * F90, coat, coatRoughness, reflectionNormal[3] (zero: the normal).
 * oracle_gltf_eval returns f x cos, the mixture pdf and the diffuse,
 * specular and coat weights. */
void oracle_gltf_eval(const float* surface, const float* view, const float* light, float diffuse_pdf, float* value,
    float* pdf, float* weights);
void oracle_gltf_sample_reflection(const float* normal, const float* view, float roughness, float u, float v,
    float* out);
void oracle_gltf_reflection_normal(const float* normal, const float* geometric, const float* view, float* out);
/* Material terms of `material` at texture coordinates (uv, uv1) on a
 * surface with normal +z, tangent +x, bitangent +y: albedo[3], opacity,
 * emission[3], metallic, roughness, specular-glossiness F0[3], thickness,
 * mapped normal[3]; for glTF materials dielectric F0[3], F90, coat, coat
 * roughness, coat normal[3] (25 floats). */
void oracle_material_terms(const oracle_handle* h, uint32_t material, const float* uv, const float* uv1, float* out);

/* Caustic gather (FP32) over a given photon set: `photons9` holds n rows
 * of position[3], direction[3], flux[3] in photon order; they are binned in
 * the grid of `config` (resolution, box, radius) and gathered at `point`
 * with `normal`. */
void oracle_gather_photons(const wpt_render_config* config, const float* photons9, uint32_t n, const float* point,
    const float* normal, float* out);
/* The caustic map of `config` (FP32): the stored photons in grid order as
 * rows of position, direction, flux; returns their count. */
uint32_t oracle_photon_map(const oracle_handle* h, const wpt_render_config* config, float* photons9, uint32_t capacity);

/* samplePunctualLight (FP32): direction toward the light, arriving
 * intensity and distance (1e30 for directional lights) at `point`. */
void oracle_punctual_sample(const wpt_punctual_light* light, const float* point, float* direction, float* incident,
    float* distance);

uint32_t oracle_env_tables(const oracle_handle* h, float* probabilities, float* alias2, uint32_t capacity);
float oracle_env_solid_angle(uint32_t row, uint32_t width, uint32_t height);
int oracle_env_eval(const oracle_handle* h, const float* dirs, uint32_t n, float* uv, float* pdf, float* rgb,
    uint32_t* texel);
int oracle_env_sample(const oracle_handle* h, const double* candidates, const float* draws3, uint32_t n, float* dirs,
    float* pdf);
void oracle_destroy(oracle_handle* h);

/* Partition lists: which 0 direct spheres, 1 direct instances, 2 TLAS leaves
 * (type << 31 | index, type 1 = sphere). */
uint32_t oracle_list(const oracle_handle* h, uint32_t which, uint32_t* out, uint32_t capacity);
/* Light table as 16 floats per light and the alias table as 2 words. */
uint32_t oracle_lights(const oracle_handle* h, float* lights16, float* alias2, uint32_t capacity);

/* FP32 restatement of traceRay / isOccluded.  rays6: SoA (ox,oy,oz,dx,dy,dz). */
void oracle_trace(const oracle_handle* h, const float* rays6, uint32_t n, wpt_hit* out);
void oracle_occluded(const oracle_handle* h, const float* rays6, const float* tmax, uint32_t n, uint8_t* out);

/* Independent FP64 brute force over every primitive, ignoring all placement
 * policy and acceleration structures: nearest t > 1e-3 (the tolerance model
 * of the reference's scene_acceleration test), material and shading normal. */
void oracle_brute_fp64(const oracle_handle* h, const float* rays6, uint32_t n, double* t, uint32_t* material,
                       float* normal3);

/* Renders samples [sample_start, sample_start + sample_count) of the pixels
 * of one tile, adding unnormalized radiance into framebuffer (width*height*3).
 * precision 32 = FP32 restatement, 64 = FP64 variant.  threads 0 = all. */
int oracle_render(const oracle_handle* h, const wpt_camera* camera, const wpt_render_config* config, uint32_t x,
                  uint32_t y, uint32_t w, uint32_t tile_h, uint32_t sample_start, uint32_t sample_count,
                  uint32_t precision, uint32_t threads, float* framebuffer, oracle_stats* stats);

/* Sequential stable filter (the order-preserving compaction contract). */
uint32_t oracle_compact(const uint32_t* items, const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t capacity);

/* The ray set of the reference's scene_acceleration test: 8 fixed rays then
 * n_random rays from std::mt19937(seed) with origins U(-4,4) and normalized
 * U(-1,1) directions (tests/scene_acceleration_tests.cu:184-201). */
uint32_t oracle_reference_rays(uint32_t seed, uint32_t n_random, float* rays6, uint32_t capacity);

/* Reference RNG known answers: the first `count` floats of the stream for
 * (pixel, sample) under a strategy, and the seeded state. */
uint64_t oracle_rng_stream(uint32_t pixel, uint32_t sample, uint32_t strategy, uint32_t count, float* out);

/* Elementary functions the render path uses (op: 0 sin, 1 cos, 2 tan, 3 exp,
 * 4 log, 5 pow(x, y)), as computed by both the kernels and this oracle. */
void oracle_elementary(uint32_t op, const float* x, const float* y, uint32_t n, float* out);

/* Reference BSDF helpers for known-answer tests. */
float oracle_dielectric_reflectance(float cosine, float eta);
/* offsetRayOrigin (ray.cuh:10-21): the continuation / shadow ray origin
 * pushed off a surface point along the geometric normal, to the side of
 * `direction`. */
void oracle_offset_ray_origin(const float* position, const float* normal, const float* direction, float hit_distance,
                              float* out);
void oracle_eval_specgloss(const float f0[3], const float diffuse[3], float roughness, const float n[3],
                           const float wo[3], const float wi[3], float value[3], float* pdf);

/* Differentiable estimator of wpt_diff (no reference counterpart), with
 * the parameter vector laid out as texture | roughness | emission. */
typedef struct oracle_diff_config {
    uint32_t width, height, samples_per_pixel, max_depth, rng_strategy;
    uint32_t texture_width, texture_height, textured_material;
} oracle_diff_config;

uint32_t oracle_diff_layout(const oracle_handle* h, const oracle_diff_config* c, uint32_t* texture_floats,
                            uint32_t* roughness_count, uint32_t* emission_count);
/* Mean image (w*h*3) in FP32 (precision 32) or FP64 (64). */
int oracle_diff_render(const oracle_handle* h, const wpt_camera* cam, const oracle_diff_config* c, const float* params,
                       uint32_t sample_start, uint32_t precision, uint32_t threads, float* image);
/* FP64 reverse-mode adjoint over recorded path vertices: grad = dLoss/dparams
 * for dLoss/dimage = image_grad. */
int oracle_diff_adjoint64(const oracle_handle* h, const wpt_camera* cam, const oracle_diff_config* c,
                          const float* params, uint32_t sample_start, const float* image_grad, uint32_t threads,
                          double* grad);
/* FP64 forward-mode (dual number) derivative of the image along a parameter
 * direction. */
int oracle_diff_directional64(const oracle_handle* h, const wpt_camera* cam, const oracle_diff_config* c,
                              const float* params, const float* direction, uint32_t sample_start, uint32_t threads,
                              double* dimage);

#ifdef __cplusplus
}
#endif

#endif
