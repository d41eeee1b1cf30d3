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
 *   main (local render)                         src/main.cu:11-53
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

/* Task table: tiles in row-major order, each split into sample ranges. */
WPT_API uint32_t wpt_host_create_tasks(uint32_t width, uint32_t height, uint32_t tile_width, uint32_t tile_height,
                                       uint32_t samples_per_pixel, uint32_t samples_per_task, uint64_t fingerprint,
                                       wpt_task* out, uint32_t capacity);
WPT_API uint64_t wpt_host_task_key(uint64_t fingerprint, const wpt_task* task);

/* SHA-256 of a byte buffer as 64 lowercase hex characters + NUL. */
WPT_API void wpt_host_sha256(const void* data, size_t bytes, char out[65]);

/* The pathtracer command line as a library call (returns the exit status). */
WPT_API int wpt_host_main(int argc, char** argv);

#ifdef __cplusplus
}
#endif

#endif /* WPT_HOST_H */
