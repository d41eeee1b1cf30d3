/*
 * wpt_host.h - C ABI of the host layer in libwpt.so: scene presets and
 * manifests, cameras, the command line, fingerprints and the task table.
 *
 * Reference interfaces replaced:
 *   createSceneFromConfig / createSceneCamera   include/scene/scene_factory.hpp:9-11
 *   scenePreset / loadSceneManifest             src/config.cpp:91-121, src/scene/scene_manifest.cpp:79-224
 *   parseCommandLine                            src/config.cpp:136-259
 *   computeRenderFingerprint                    src/renderer/render_fingerprint.cpp:152-157
 *   createRenderTasks / computeRenderTaskKey    src/distributed/task.cpp:33-75
 *   render checkpoint files                     src/renderer/render_checkpoint.cpp:17-64,163-195
 *   DistributedCoordinator (task ledger)        include/distributed/coordinator.hpp:29-79
 *   write/readDistributedCheckpoint             src/distributed/checkpoint.cpp:86-221
 *   DistributedCheckpointWriter                 include/distributed/checkpoint_writer.hpp:11-40
 *   loadObjScene, loadEnvironmentMap            src/scene/obj_loader.cpp:47-198, src/scene/environment.cpp:66-82
 *   main (local render, coordinator role)       src/main.cu:11-53, src/distributed/coordinator_app.cpp:56-180
 */
#ifndef WPT_HOST_H
#define WPT_HOST_H

#include "wpt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wpt_host_scene wpt_host_scene;

/* A scene from a preset name ("cornell-box", "sphere-field", "demo",
 * "rtweek", "prism", ...) or a manifest path ending in ".json".  seed and
 * icosphere_frequency override the procedural parameters when nonzero. */
WPT_API int wpt_host_scene_create(const char* preset_or_manifest, uint32_t seed, uint32_t icosphere_frequency,
                                  wpt_host_scene** out);
WPT_API int wpt_host_scene_get_desc(const wpt_host_scene* scene, wpt_scene_desc* out);
WPT_API int wpt_host_scene_get_camera(const wpt_host_scene* scene, uint32_t width, uint32_t height, wpt_camera* out);
WPT_API int wpt_host_scene_destroy(wpt_host_scene* scene);

/* Parses a reference-style command line (argv[0] is the program name) into
 * the render configuration and the scene's fingerprint.  Returns WPT_OK when
 * the program should render, WPT_ERROR_INVALID_ARGUMENT otherwise. */
WPT_API int wpt_host_parse_command_line(int argc, const char* const* argv, wpt_render_config* render,
                                        char* scene_name, uint32_t scene_name_capacity, uint64_t* fingerprint);

/* The scene's identity for dataset grouping: the fingerprint of a preset or
 * manifest without its display name (computeSceneFingerprint,
 * src/renderer/render_fingerprint.cpp:157-161).  Asset and environment
 * files are hashed by content. */
WPT_API int wpt_host_scene_fingerprint(const char* preset_or_manifest, uint64_t* out);

/* Task table: tiles in row-major order, each split into sample ranges. */
WPT_API uint32_t wpt_host_create_tasks(uint32_t width, uint32_t height, uint32_t tile_width, uint32_t tile_height,
                                       uint32_t samples_per_pixel, uint32_t samples_per_task, uint64_t fingerprint,
                                       wpt_task* out, uint32_t capacity);
WPT_API uint64_t wpt_host_task_key(uint64_t fingerprint, const wpt_task* task);

/* Decodes a Radiance .hdr file (the reference's loadEnvironmentMap over
 * stb_image, src/scene/environment.cpp:66-82) into linear RGB floats, rows
 * top to bottom.  With rgb NULL only the size is returned; else capacity
 * (floats) must hold width*height*3.  Fails on unreadable, non-Radiance or
 * truncated files. */
WPT_API int wpt_host_load_hdr(const char* filename, float* rgb, uint64_t capacity, uint32_t* width,
                              uint32_t* height);

/* Decodes an image file held in memory to 8-bit RGBA, rows top to bottom,
 * as the reference's glTF texture import asks stb_image for it
 * (loadTexture, src/scene/gltf_loader.cpp:273-305, STBI_rgb_alpha).  PNG
 * of every colour type, bit depth and interlacing; baseline and progressive
 * JPEG; other formats fail.  With rgba NULL only the size is returned; else
 * capacity (bytes) must hold width*height*4. */
WPT_API int wpt_host_decode_image(const uint8_t* data, uint64_t size, uint8_t* rgba, uint64_t capacity,
                                  uint32_t* width, uint32_t* height);

/* ---- render checkpoints ------------------------------------------------
 * Header {"PTCHKPT1", version 3, width, height, completed samples,
 * fingerprint} then w*h*3 floats of unnormalized sums; written to
 * "<filename>.tmp" and renamed.  read fails (WPT_ERROR_INVALID_ARGUMENT,
 * with the reason in wpt_last_error) on a missing, truncated or foreign
 * file, another resolution or another fingerprint. */
WPT_API int wpt_host_render_checkpoint_write(const char* filename, const float* rgb, uint32_t width, uint32_t height,
                                             uint32_t sample_index, uint64_t fingerprint);
WPT_API int wpt_host_render_checkpoint_read(const char* filename, uint32_t width, uint32_t height,
                                            uint64_t fingerprint, float* rgb, uint32_t* sample_index);

/* ---- task ledger ------------------------------------------------------
 * The coordinator's task state on one host: leases with expiry, merging of
 * tile sums, snapshots to "PTDIST1" ledger checkpoints.  Thread-safe. */
typedef struct wpt_ledger wpt_ledger;
typedef struct wpt_ledger_writer wpt_ledger_writer;

typedef struct wpt_ledger_job {
    uint64_t job_fingerprint;
    uint32_t width, height, tile_width, tile_height;
    uint32_t samples_per_pixel, samples_per_task;
} wpt_ledger_job;

#define WPT_LEDGER_ACCEPTED 0
#define WPT_LEDGER_DUPLICATE 1
#define WPT_LEDGER_REJECTED 2

#define WPT_TASK_PENDING 0
#define WPT_TASK_LEASED 1
#define WPT_TASK_COMPLETED 2

/* Tasks whose job_fingerprint differs from the job's are dropped. */
WPT_API int wpt_ledger_create(const wpt_ledger_job* job, const wpt_task* tasks, uint32_t task_count,
                              wpt_ledger** out);
WPT_API int wpt_ledger_destroy(wpt_ledger* ledger);
/* Leases the oldest pending task to worker_id until now_ms + lease_ms:
 * returns 1 with *task filled, 0 when nothing is pending (or the worker id
 * is empty), or an error code below 0. */
WPT_API int wpt_ledger_lease_next(wpt_ledger* ledger, const char* worker_id, uint64_t now_ms, uint64_t lease_ms,
                                  wpt_task* task, uint64_t* expires_at_ms);
/* Merges the unnormalized sums of a task's tile (width*height*3 floats):
 * returns WPT_LEDGER_ACCEPTED / DUPLICATE / REJECTED. */
WPT_API int wpt_ledger_commit(wpt_ledger* ledger, const wpt_task* task, uint32_t sample_count,
                              const float* radiance, uint64_t radiance_floats);
WPT_API uint32_t wpt_ledger_renew_leases(wpt_ledger* ledger, const char* worker_id, uint64_t now_ms,
                                         uint64_t lease_ms);
WPT_API uint32_t wpt_ledger_release_leases(wpt_ledger* ledger, const char* worker_id);
WPT_API uint32_t wpt_ledger_requeue_expired(wpt_ledger* ledger, uint64_t now_ms);
/* WPT_TASK_* of the task with this (tile_ordinal, sample_start,
 * sample_count); pending for an unknown one. */
WPT_API int wpt_ledger_task_state(const wpt_ledger* ledger, const wpt_task* task);
WPT_API int wpt_ledger_counts(const wpt_ledger* ledger, uint64_t* tasks, uint64_t* pending, uint64_t* completed);
/* Accumulated sums (width*height*3 floats) and per-pixel sample counts. */
WPT_API int wpt_ledger_read(const wpt_ledger* ledger, float* radiance, uint32_t* sample_counts);
WPT_API int wpt_ledger_write_checkpoint(const wpt_ledger* ledger, const char* filename, uint64_t sequence);
/* Restores completed tasks and sums from a checkpoint of the same job; every
 * other task becomes pending. */
WPT_API int wpt_ledger_restore_checkpoint(wpt_ledger* ledger, const char* filename, uint64_t* sequence);
/* Background checkpoint writer keeping only the newest waiting snapshot.
 * destroy writes what is still queued before it returns. */
WPT_API int wpt_ledger_writer_create(const char* filename, wpt_ledger_writer** out);
WPT_API int wpt_ledger_writer_enqueue(wpt_ledger_writer* writer, const wpt_ledger* ledger, uint64_t sequence);
WPT_API int wpt_ledger_writer_flush(wpt_ledger_writer* writer);
WPT_API int wpt_ledger_writer_destroy(wpt_ledger_writer* writer);

/* SHA-256 of a byte buffer as 64 lowercase hex characters + NUL. */
WPT_API void wpt_host_sha256(const void* data, size_t bytes, char out[65]);

/* The pathtracer command line as a library call (returns the exit status).
 * --coordinator renders the job's task table through a local task ledger
 * with ledger checkpoints (and --resume from one); --worker, which needs a
 * networked coordinator, is rejected. */
WPT_API int wpt_host_main(int argc, char** argv);

#ifdef __cplusplus
}
#endif

#endif /* WPT_HOST_H */
