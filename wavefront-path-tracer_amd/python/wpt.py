"""ctypes binding of libwpt.so (include/wpt.h, include/wpt_host.h).

Used by the tests, the benchmark and the graft entry points.  Device buffers
for the traversal/compaction entry points come from PyTorch tensors; PyTorch
is plumbing only, all rendering work runs in the gfx950 kernels of libwpt.so.
The library must exist (built by ``make`` in this package); nothing here falls
back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Iterable, List, Optional, Sequence

import numpy as np

PACKAGE_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("WPT_LIBRARY") or os.path.join(PACKAGE_DIR, "lib", "libwpt.so")

OK = 0
ACCEL_BVH, ACCEL_DIRECT = 0, 1
DIFFUSE, METAL, DIELECTRIC, EMISSIVE, SPECULAR_GLOSSINESS = 0, 1, 2, 3, 4
RNG_XORSHIFT32, RNG_PCG32 = 0, 1
MODE_GRAPH, MODE_PERSISTENT, MODE_REGENERATE = 0, 1, 2
HIT_NONE, HIT_SPHERE, HIT_STATIC, HIT_INSTANCE, HIT_FLAT = 0, 1, 2, 3, 4
INVALID = 0xFFFFFFFF
TRI_HAS_NORMALS, TRI_HAS_TEXCOORDS, TRI_HAS_TEXCOORDS1, TRI_HAS_VERTEX_TANGENTS = 1, 2, 4, 8

F3 = C.c_float * 3
F2 = C.c_float * 2


class Sphere(C.Structure):
    _fields_ = [("center", F3), ("radius", C.c_float), ("material", C.c_uint32), ("acceleration", C.c_uint32)]


class Triangle(C.Structure):
    _fields_ = [("v0", F3), ("v1", F3), ("v2", F3), ("n0", F3), ("n1", F3), ("n2", F3),
                ("uv0", F2), ("uv1", F2), ("uv2", F2), ("tangent", F3), ("bitangent", F3),
                ("material", C.c_uint32), ("flags", C.c_uint32),
                ("uv1_0", F2), ("uv1_1", F2), ("uv1_2", F2), ("vertex_tangent0", F3), ("vertex_tangent1", F3),
                ("vertex_tangent2", F3), ("vertex_tangent_sign", F3), ("tangent1", F3), ("bitangent1", F3)]


class Mesh(C.Structure):
    _fields_ = [("first_triangle", C.c_uint32), ("triangle_count", C.c_uint32)]


class Transform(C.Structure):
    _fields_ = [("local_to_world", F3 * 3), ("world_to_local", F3 * 3), ("translation", F3)]


class Instance(C.Structure):
    _fields_ = [("mesh", C.c_uint32), ("material_override", C.c_uint32), ("acceleration", C.c_uint32),
                ("reserved", C.c_uint32), ("transform", Transform)]


class Material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("albedo", F3), ("emission", F3), ("roughness", C.c_float),
                ("ior", C.c_float), ("metallic", C.c_float), ("opacity", C.c_float),
                ("alpha_cutoff", C.c_float), ("alpha_masked", C.c_uint32), ("attenuation_color", F3),
                ("attenuation_distance", C.c_float), ("volume_density", C.c_float), ("specular", F3),
                ("glossiness", C.c_float)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba", C.c_void_p)]


class TextureBinding(C.Structure):
    _fields_ = [("texture", C.c_uint32), ("texcoord", C.c_uint32), ("offset", F2), ("scale", F2),
                ("rotation", C.c_float), ("has_transform", C.c_uint32)]


TEXTURE_SLOTS = ("albedo_texture", "metallic_roughness_texture", "emissive_texture", "normal_texture",
                 "thickness_texture", "specular_glossiness_texture", "specular_weight_texture",
                 "specular_color_texture", "clearcoat_texture", "clearcoat_roughness_texture",
                 "clearcoat_normal_texture")


class MaterialExt(C.Structure):
    _fields_ = [(name, TextureBinding) for name in TEXTURE_SLOTS] + [
        ("normal_scale", C.c_float), ("specular_glossiness_workflow", C.c_uint32),
        ("gltf_pbr_workflow", C.c_uint32), ("specular_weight", C.c_float), ("specular_color", F3),
        ("clearcoat", C.c_float), ("clearcoat_roughness", C.c_float), ("clearcoat_normal_scale", C.c_float),
        ("dispersion", C.c_float)]


def texture_binding(texture=INVALID, texcoord=0, offset=(0.0, 0.0), scale=(1.0, 1.0), rotation=0.0,
                    transform=False) -> TextureBinding:
    return TextureBinding(texture, texcoord, F2(*offset), F2(*scale), rotation, 1 if transform else 0)


def make_material_ext(normal_scale=1.0, specular_glossiness_workflow=False, gltf_pbr_workflow=False,
                      specular_weight=1.0, specular_color=(1.0, 1.0, 1.0), clearcoat=0.0, clearcoat_roughness=0.0,
                      clearcoat_normal_scale=1.0, dispersion=0.0, **bindings) -> MaterialExt:
    """Texture bindings (keyword per TEXTURE_SLOTS name, a TextureBinding or
    a texture index) and glTF workflow fields, with the reference's defaults."""
    e = MaterialExt()
    for name in TEXTURE_SLOTS:
        b = bindings.pop(name, None)
        if b is None:
            b = texture_binding()
        elif isinstance(b, int):
            b = texture_binding(b)
        setattr(e, name, b)
    if bindings:
        raise TypeError(f"unknown texture slots {sorted(bindings)}")
    e.normal_scale = normal_scale
    e.specular_glossiness_workflow = 1 if specular_glossiness_workflow else 0
    e.gltf_pbr_workflow = 1 if gltf_pbr_workflow else 0
    e.specular_weight = specular_weight
    e.specular_color = vec3(specular_color)
    e.clearcoat = clearcoat
    e.clearcoat_roughness = clearcoat_roughness
    e.clearcoat_normal_scale = clearcoat_normal_scale
    e.dispersion = dispersion
    return e


LIGHT_POINT, LIGHT_SPOT, LIGHT_DIRECTIONAL = 0, 1, 2


class PunctualLight(C.Structure):
    _fields_ = [("type", C.c_uint32), ("position", F3), ("direction", F3), ("intensity", F3), ("range", C.c_float),
                ("inner_cone_cos", C.c_float), ("outer_cone_cos", C.c_float)]


def make_punctual_light(kind=LIGHT_POINT, position=(0, 0, 0), direction=(0, 0, -1), intensity=(1, 1, 1), range_=0.0,
                        inner_cone_cos=1.0, outer_cone_cos=0.70710678118) -> PunctualLight:
    return PunctualLight(kind, vec3(position), vec3(direction), vec3(intensity), range_, inner_cone_cos,
                         outer_cone_cos)


class SceneDesc(C.Structure):
    _fields_ = [("spheres", C.POINTER(Sphere)), ("sphere_count", C.c_uint32),
                ("static_triangles", C.POINTER(Triangle)), ("static_triangle_count", C.c_uint32),
                ("flat_triangles", C.POINTER(Triangle)), ("flat_triangle_count", C.c_uint32),
                ("mesh_triangles", C.POINTER(Triangle)), ("mesh_triangle_count", C.c_uint32),
                ("meshes", C.POINTER(Mesh)), ("mesh_count", C.c_uint32),
                ("instances", C.POINTER(Instance)), ("instance_count", C.c_uint32),
                ("materials", C.POINTER(Material)), ("material_count", C.c_uint32),
                ("black_background", C.c_uint32), ("ignore_direct_light_occlusion", C.c_uint32),
                ("environment_rgb", C.POINTER(C.c_float)), ("environment_width", C.c_uint32),
                ("environment_height", C.c_uint32), ("environment_intensity", C.c_float),
                ("environment_rotation_degrees", C.c_float),
                ("punctual_lights", C.POINTER(PunctualLight)), ("punctual_light_count", C.c_uint32),
                ("textures", C.POINTER(Texture)), ("texture_count", C.c_uint32),
                ("material_ext", C.POINTER(MaterialExt)), ("use_normal_maps", C.c_uint32),
                ("normal_map_minimum_cosine", C.c_float), ("bvh_builder", C.c_uint32),
                ("bvh_search_radius", C.c_uint32), ("flat_accelerator", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("position", F3), ("forward", F3), ("right", F3), ("up", F3),
                ("vertical_fov", C.c_float), ("aspect_ratio", C.c_float)]


class RenderConfig(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("tiled", C.c_uint32),
                ("tile_width", C.c_uint32), ("tile_height", C.c_uint32), ("max_depth", C.c_uint32),
                ("rr_start_depth", C.c_uint32), ("samples_per_pixel", C.c_uint32),
                ("block_size", C.c_uint32), ("rng_strategy", C.c_uint32), ("indirect_clamp", C.c_float),
                ("mode", C.c_uint32), ("persistent_blocks", C.c_uint32), ("intersection_debug", C.c_uint32),
                ("shading_normal_debug", C.c_uint32), ("profile_queues", C.c_uint32),
                ("caustic_photons", C.c_uint32), ("caustic_max_depth", C.c_uint32),
                ("caustic_gather_radius", C.c_float), ("caustic_gather", C.c_uint32),
                ("caustic_spectral", C.c_uint32), ("caustic_grid_resolution", C.c_uint32),
                ("caustic_grid_min", F3), ("caustic_grid_max", F3), ("samples_per_pass", C.c_uint32)]


class Task(C.Structure):
    _fields_ = [("job_fingerprint", C.c_uint64), ("tile_ordinal", C.c_uint32), ("sample_start", C.c_uint32),
                ("sample_count", C.c_uint32), ("x", C.c_uint32), ("y", C.c_uint32),
                ("width", C.c_uint32), ("height", C.c_uint32)]


class LedgerJob(C.Structure):
    _fields_ = [("job_fingerprint", C.c_uint64), ("width", C.c_uint32), ("height", C.c_uint32),
                ("tile_width", C.c_uint32), ("tile_height", C.c_uint32), ("samples_per_pixel", C.c_uint32),
                ("samples_per_task", C.c_uint32)]


LEDGER_ACCEPTED, LEDGER_DUPLICATE, LEDGER_REJECTED = 0, 1, 2
BVH_HOST_SAH, BVH_DEVICE_PLOC = 0, 1
FLAT_SAH_BVH2, FLAT_MEDIAN_BVH2, FLAT_SPATIAL_SPLIT_BVH2, FLAT_WIDE8 = 0, 1, 2, 3
TASK_PENDING, TASK_LEASED, TASK_COMPLETED = 0, 1, 2


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("u", C.c_float), ("v", C.c_float), ("kind", C.c_uint32),
                ("instance", C.c_uint32), ("primitive", C.c_uint32), ("material", C.c_uint32),
                ("light_index", C.c_uint32)]


HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("kind", "<u4"), ("instance", "<u4"),
                      ("primitive", "<u4"), ("material", "<u4"), ("light_index", "<u4")])


class Stats(C.Structure):
    _fields_ = [("closest_queries", C.c_uint64), ("shadow_queries", C.c_uint64), ("samples", C.c_uint64),
                ("trace_milliseconds", C.c_double), ("queue_counts", C.c_uint32 * 64),
                ("queue_count_depth", C.c_uint32), ("caustic_photons_stored", C.c_uint32),
                ("caustic_build_milliseconds", C.c_double)]


class LaunchInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("mode", "block_size", "blocks", "blocks_per_cu", "compute_units",
                                          "registers", "scratch_bytes", "lds_bytes_per_block", "lds_bytes_per_cu",
                                          "waves_per_simd")]


class SceneInfo(C.Structure):
    _fields_ = [(name, C.c_uint32) for name in (
        "sphere_count", "direct_sphere_count", "instance_count", "direct_instance_count", "tlas_leaf_count",
        "tlas_node_count", "blas_count", "light_count", "static_triangle_count", "flat_triangle_count",
        "flat_node_count", "material_count", "bvh_builder", "max_blas_depth")] + [
        ("blas_build_milliseconds", C.c_double)] + [(name, C.c_uint32) for name in (
        "flat_accelerator", "flat_leaf_slots", "flat_wide_node_count", "max_stack_depth")]


class BatchTask(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("camera", Camera), ("width", C.c_uint32), ("height", C.c_uint32),
                ("x", C.c_uint32), ("y", C.c_uint32), ("tile_width", C.c_uint32), ("tile_height", C.c_uint32),
                ("sample_start", C.c_uint32), ("sample_count", C.c_uint32), ("max_depth", C.c_uint32),
                ("rr_start_depth", C.c_uint32), ("rng_strategy", C.c_uint32), ("indirect_clamp", C.c_float)]


BATCH_STAGED, BATCH_REGENERATE = 0, 1


class BatchConfig(C.Structure):
    _fields_ = [("block_size", C.c_uint32), ("blocks", C.c_uint32), ("scheduler", C.c_uint32),
                ("pass_paths", C.c_uint32)]


class BatchStats(C.Structure):
    _fields_ = [("closest_queries", C.c_uint64), ("shadow_queries", C.c_uint64), ("samples", C.c_uint64),
                ("milliseconds", C.c_double), ("block_size", C.c_uint32), ("blocks", C.c_uint32)]


class DiffConfig(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("samples_per_pixel", C.c_uint32),
                ("max_depth", C.c_uint32), ("rng_strategy", C.c_uint32), ("texture_width", C.c_uint32),
                ("texture_height", C.c_uint32), ("textured_material", C.c_uint32), ("block_size", C.c_uint32)]


class DiffLayout(C.Structure):
    _fields_ = [("texture_floats", C.c_uint32), ("roughness_count", C.c_uint32), ("emission_count", C.c_uint32),
                ("parameter_count", C.c_uint32)]


class DiffStats(C.Structure):
    _fields_ = [("forward_milliseconds", C.c_double), ("adjoint_milliseconds", C.c_double), ("paths", C.c_uint64)]


class Bvh2Node(C.Structure):
    _fields_ = [("bmin", F3), ("bmax", F3), ("left", C.c_uint32), ("right", C.c_uint32)]


_lib = None


def library() -> C.CDLL:
    """Loads libwpt.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {PACKAGE_DIR}`")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    u32, u64 = C.c_uint32, C.c_uint64
    sigs = {
        "wpt_abi_version": ([], C.c_int),
        "wpt_last_error": ([], C.c_char_p),
        "wpt_device_count": ([], C.c_int),
        "wpt_scene_create": ([C.POINTER(SceneDesc), C.POINTER(P)], C.c_int),
        "wpt_scene_create_host": ([C.POINTER(SceneDesc), C.POINTER(P)], C.c_int),
        "wpt_scene_destroy": ([P], C.c_int),
        "wpt_scene_info_get": ([P, C.POINTER(SceneInfo)], C.c_int),
        "wpt_scene_get_list": ([P, u32, P, u32, C.POINTER(u32)], C.c_int),
        "wpt_scene_export_bvh2": ([P, u32, P, u32, C.POINTER(u32)], C.c_int),
        "wpt_scene_export_lights": ([P, P, P, u32, C.POINTER(u32)], C.c_int),
        "wpt_scene_export_environment": ([P, P, P, u32, C.POINTER(u32)], C.c_int),
        "wpt_scene_export_light_indices": ([P, P, P, P], C.c_int),
        "wpt_trace_closest": ([P, P, u32, P, P], C.c_int),
        "wpt_trace_occluded": ([P, P, P, u32, P, P], C.c_int),
        "wpt_compact": ([P, P, u32, P, u32, P, P], C.c_int),
        "wpt_compact_noalloc": ([P, P, u32, P, u32, P, P], C.c_int),
        "wpt_renderer_create": ([P, C.POINTER(Camera), C.POINTER(RenderConfig), u64, C.POINTER(P)], C.c_int),
        "wpt_renderer_destroy": ([P], C.c_int),
        "wpt_renderer_reset": ([P], C.c_int),
        "wpt_renderer_render_samples": ([P, u32, u32], C.c_int),
        "wpt_renderer_read_framebuffer": ([P, P, C.POINTER(u32)], C.c_int),
        "wpt_render_frame": ([P, P, C.POINTER(u32)], C.c_int),
        "wpt_render_task": ([P, C.POINTER(Task), P], C.c_int),
        "wpt_renderer_last_stats": ([P, C.POINTER(Stats)], C.c_int),
        "wpt_renderer_profile_kernels": ([P, u32, u32, P, P], C.c_int),
        "wpt_renderer_launch_info": ([P, C.POINTER(LaunchInfo)], C.c_int),
        "wpt_batch_create": ([C.POINTER(P), u32, C.POINTER(BatchConfig), C.POINTER(P)], C.c_int),
        "wpt_batch_destroy": ([P], C.c_int),
        "wpt_batch_set_caustics": ([P, u32, C.POINTER(RenderConfig), C.POINTER(u32), C.POINTER(C.c_double)], C.c_int),
        "wpt_batch_render": ([P, C.POINTER(BatchTask), u32, C.POINTER(P), P], C.c_int),
        "wpt_batch_last_stats": ([P, C.POINTER(BatchStats)], C.c_int),
        "wpt_batch_autotune": ([P, C.POINTER(BatchTask), u32, C.POINTER(BatchConfig)], C.c_int),
        "wpt_diff_create": ([P, C.POINTER(Camera), C.POINTER(DiffConfig), C.POINTER(P)], C.c_int),
        "wpt_diff_destroy": ([P], C.c_int),
        "wpt_diff_get_layout": ([P, C.POINTER(DiffLayout)], C.c_int),
        "wpt_diff_forward": ([P, P, u32, P], C.c_int),
        "wpt_diff_adjoint": ([P, P, P, P], C.c_int),
        "wpt_diff_adam": ([P, P, P, P, P, u32, C.c_float, C.c_float, C.c_float, C.c_float], C.c_int),
        "wpt_diff_last_stats": ([P, C.POINTER(DiffStats)], C.c_int),
        "wpt_host_scene_create": ([C.c_char_p, u32, u32, C.POINTER(P)], C.c_int),
        "wpt_host_scene_get_desc": ([P, C.POINTER(SceneDesc)], C.c_int),
        "wpt_host_scene_get_camera": ([P, u32, u32, C.POINTER(Camera)], C.c_int),
        "wpt_host_scene_destroy": ([P], C.c_int),
        "wpt_host_parse_command_line": ([C.c_int, C.POINTER(C.c_char_p), C.POINTER(RenderConfig), C.c_char_p, u32,
                                         C.POINTER(u64)], C.c_int),
        "wpt_host_create_tasks": ([u32, u32, u32, u32, u32, u32, u64, P, u32], u32),
        "wpt_host_task_key": ([u64, C.POINTER(Task)], u64),
        "wpt_host_scene_fingerprint": ([C.c_char_p, C.POINTER(u64)], C.c_int),
        "wpt_renderer_load_framebuffer": ([P, P, u32], C.c_int),
        "wpt_host_load_hdr": ([C.c_char_p, P, u64, C.POINTER(u32), C.POINTER(u32)], C.c_int),
        "wpt_host_decode_image": ([P, u64, P, u64, C.POINTER(u32), C.POINTER(u32)], C.c_int),
        "wpt_host_render_checkpoint_write": ([C.c_char_p, P, u32, u32, u32, u64], C.c_int),
        "wpt_host_render_checkpoint_read": ([C.c_char_p, u32, u32, u64, P, C.POINTER(u32)], C.c_int),
        "wpt_ledger_create": ([C.POINTER(LedgerJob), P, u32, C.POINTER(P)], C.c_int),
        "wpt_ledger_destroy": ([P], C.c_int),
        "wpt_ledger_lease_next": ([P, C.c_char_p, u64, u64, C.POINTER(Task), C.POINTER(u64)], C.c_int),
        "wpt_ledger_commit": ([P, C.POINTER(Task), u32, P, u64], C.c_int),
        "wpt_ledger_renew_leases": ([P, C.c_char_p, u64, u64], u32),
        "wpt_ledger_release_leases": ([P, C.c_char_p], u32),
        "wpt_ledger_requeue_expired": ([P, u64], u32),
        "wpt_ledger_task_state": ([P, C.POINTER(Task)], C.c_int),
        "wpt_ledger_counts": ([P, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)], C.c_int),
        "wpt_ledger_read": ([P, P, P], C.c_int),
        "wpt_ledger_write_checkpoint": ([P, C.c_char_p, u64], C.c_int),
        "wpt_ledger_restore_checkpoint": ([P, C.c_char_p, C.POINTER(u64)], C.c_int),
        "wpt_ledger_writer_create": ([C.c_char_p, C.POINTER(P)], C.c_int),
        "wpt_ledger_writer_enqueue": ([P, P, u64], C.c_int),
        "wpt_ledger_writer_flush": ([P], C.c_int),
        "wpt_ledger_writer_destroy": ([P], C.c_int),
        "wpt_host_sha256": ([P, C.c_size_t, C.c_char_p], None),
        "wpt_host_main": ([C.c_int, C.POINTER(C.c_char_p)], C.c_int),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


class WptError(RuntimeError):
    def __init__(self, status: int, where: str):
        msg = library().wpt_last_error().decode(errors="replace")
        super().__init__(f"{where} failed with status {status}: {msg}")
        self.status = status


def check(status: int, where: str) -> None:
    if status != OK:
        raise WptError(status, where)


def device_count() -> int:
    return int(library().wpt_device_count())


# ---------------------------------------------------------------- scene construction
def vec3(v) -> F3:
    return F3(*[float(x) for x in v])


def make_material(kind=DIFFUSE, albedo=(0.5, 0.5, 0.5), emission=(0, 0, 0), roughness=0.0, ior=1.0, metallic=0.0,
                  opacity=1.0, alpha_cutoff=0.5, alpha_masked=False, attenuation_color=(1, 1, 1),
                  attenuation_distance=0.0, volume_density=1.0, specular=(1, 1, 1), glossiness=1.0) -> Material:
    return Material(kind, vec3(albedo), vec3(emission), roughness, ior, metallic, opacity, alpha_cutoff,
                    1 if alpha_masked else 0, vec3(attenuation_color), attenuation_distance, volume_density,
                    vec3(specular), glossiness)


def make_triangle(v0, v1, v2, material, normals=None, uvs=None, tangent=(0, 0, 0), bitangent=(0, 0, 0), uvs1=None,
                  vertex_tangents=None, tangent_signs=(1.0, 1.0, 1.0), tangent1=(0, 0, 0),
                  bitangent1=(0, 0, 0)) -> Triangle:
    t = Triangle()
    t.v0, t.v1, t.v2 = vec3(v0), vec3(v1), vec3(v2)
    flags = 0
    if normals is not None:
        t.n0, t.n1, t.n2 = vec3(normals[0]), vec3(normals[1]), vec3(normals[2])
        flags |= TRI_HAS_NORMALS
    if uvs is not None:
        t.uv0, t.uv1, t.uv2 = F2(*uvs[0]), F2(*uvs[1]), F2(*uvs[2])
        flags |= TRI_HAS_TEXCOORDS
    if uvs1 is not None:
        t.uv1_0, t.uv1_1, t.uv1_2 = F2(*uvs1[0]), F2(*uvs1[1]), F2(*uvs1[2])
        flags |= TRI_HAS_TEXCOORDS1
    if vertex_tangents is not None:
        t.vertex_tangent0, t.vertex_tangent1, t.vertex_tangent2 = (vec3(v) for v in vertex_tangents)
        t.vertex_tangent_sign = vec3(tangent_signs)
        flags |= TRI_HAS_VERTEX_TANGENTS
    t.tangent1, t.bitangent1 = vec3(tangent1), vec3(bitangent1)
    t.tangent, t.bitangent = vec3(tangent), vec3(bitangent)
    t.material = material
    t.flags = flags
    return t


def scaled_transform(translation=(0, 0, 0), scale=1.0, axis_scales=None) -> Transform:
    sx, sy, sz = axis_scales if axis_scales is not None else (scale, scale, scale)
    t = Transform()
    for i, s in enumerate((sx, sy, sz)):
        t.local_to_world[i][i] = s
        t.world_to_local[i][i] = np.float32(1.0) / np.float32(s)
    t.translation = vec3(translation)
    return t


class SceneBuilder:
    """Owns the arrays behind a wpt_scene_desc assembled from Python."""

    def __init__(self):
        self.spheres: List[Sphere] = []
        self.static_triangles: List[Triangle] = []
        self.flat_triangles: List[Triangle] = []
        self.flat_accelerator = FLAT_SAH_BVH2
        self.mesh_triangles: List[Triangle] = []
        self.meshes: List[Mesh] = []
        self.instances: List[Instance] = []
        self.materials: List[Material] = []
        self.black_background = False
        self.ignore_direct_light_occlusion = False
        self.environment_rgb = None  # float32 (height, width, 3), row 0 at +Y
        self.environment_intensity = 1.0
        self.environment_rotation_degrees = 0.0
        self.punctual_lights: List[PunctualLight] = []
        self.textures: List[np.ndarray] = []
        self.material_exts: List[Optional[MaterialExt]] = []
        self.use_normal_maps = True
        self.normal_map_minimum_cosine = 0.0
        self._keep = []

    def texture(self, rgba) -> int:
        """An RGBA8 texture: uint8 (height, width, 4), rows top to bottom."""
        arr = np.ascontiguousarray(rgba, dtype=np.uint8)
        if arr.ndim != 3 or arr.shape[2] != 4:
            raise ValueError("texture must be (height, width, 4) uint8")
        self.textures.append(arr)
        return len(self.textures) - 1

    def punctual_light(self, light: PunctualLight) -> int:
        self.punctual_lights.append(light)
        return len(self.punctual_lights) - 1

    def environment(self, rgb, intensity=1.0, rotation_degrees=0.0):
        """Equirectangular HDR map: float (height, width, 3) linear radiance."""
        arr = np.ascontiguousarray(rgb, dtype=np.float32)
        if arr.ndim != 3 or arr.shape[2] != 3:
            raise ValueError("environment must be (height, width, 3)")
        self.environment_rgb = arr
        self.environment_intensity = float(intensity)
        self.environment_rotation_degrees = float(rotation_degrees)

    def sphere(self, center, radius, material, acceleration=ACCEL_BVH):
        self.spheres.append(Sphere(vec3(center), radius, material, acceleration))
        return len(self.spheres) - 1

    def mesh(self, triangles: Sequence[Triangle]) -> int:
        self.meshes.append(Mesh(len(self.mesh_triangles), len(triangles)))
        self.mesh_triangles.extend(triangles)
        return len(self.meshes) - 1

    def instance(self, mesh, transform=None, material_override=INVALID, acceleration=ACCEL_BVH):
        self.instances.append(Instance(mesh, material_override, acceleration, 0, transform or scaled_transform()))
        return len(self.instances) - 1

    def material(self, m: Material, ext: Optional[MaterialExt] = None) -> int:
        self.materials.append(m)
        while len(self.material_exts) < len(self.materials) - 1:
            self.material_exts.append(None)
        self.material_exts.append(ext)
        return len(self.materials) - 1

    def _array(self, ctype, items):
        if not items:
            return None, 0
        arr = (ctype * len(items))(*items)
        self._keep.append(arr)
        return C.cast(arr, C.POINTER(ctype)), len(items)

    def desc(self) -> SceneDesc:
        d = SceneDesc()
        d.spheres, d.sphere_count = self._array(Sphere, self.spheres)
        d.static_triangles, d.static_triangle_count = self._array(Triangle, self.static_triangles)
        d.flat_triangles, d.flat_triangle_count = self._array(Triangle, self.flat_triangles)
        d.flat_accelerator = self.flat_accelerator
        d.mesh_triangles, d.mesh_triangle_count = self._array(Triangle, self.mesh_triangles)
        d.meshes, d.mesh_count = self._array(Mesh, self.meshes)
        d.instances, d.instance_count = self._array(Instance, self.instances)
        d.materials, d.material_count = self._array(Material, self.materials)
        d.black_background = 1 if self.black_background else 0
        d.ignore_direct_light_occlusion = 1 if self.ignore_direct_light_occlusion else 0
        d.punctual_lights, d.punctual_light_count = self._array(PunctualLight, self.punctual_lights)
        if self.textures:
            views = []
            for t in self.textures:
                self._keep.append(t)
                views.append(Texture(t.shape[1], t.shape[0], t.ctypes.data))
            d.textures, d.texture_count = self._array(Texture, views)
        if any(e is not None for e in self.material_exts):
            exts = [self.material_exts[i] if i < len(self.material_exts) and self.material_exts[i] is not None
                    else make_material_ext() for i in range(len(self.materials))]
            d.material_ext, _ = self._array(MaterialExt, exts)
        d.use_normal_maps = 1 if self.use_normal_maps else 0
        d.normal_map_minimum_cosine = self.normal_map_minimum_cosine
        if self.environment_rgb is not None:
            env = self.environment_rgb
            self._keep.append(env)
            d.environment_rgb = env.ctypes.data_as(C.POINTER(C.c_float))
            d.environment_height, d.environment_width = env.shape[0], env.shape[1]
            d.environment_intensity = self.environment_intensity
            d.environment_rotation_degrees = self.environment_rotation_degrees
        self._desc = d
        return d


class HostScene:
    """A preset or manifest scene assembled by the C++ host layer."""

    def __init__(self, name: str, seed: int = 0, icosphere_frequency: int = 0):
        lib = library()
        self._h = C.c_void_p()
        check(lib.wpt_host_scene_create(name.encode(), seed, icosphere_frequency, C.byref(self._h)),
              f"wpt_host_scene_create({name})")
        self._desc = SceneDesc()
        check(lib.wpt_host_scene_get_desc(self._h, C.byref(self._desc)), "wpt_host_scene_get_desc")

    def desc(self) -> SceneDesc:
        return self._desc

    def camera(self, width: int, height: int) -> Camera:
        cam = Camera()
        check(library().wpt_host_scene_get_camera(self._h, width, height, C.byref(cam)), "wpt_host_scene_get_camera")
        return cam

    def close(self):
        if self._h:
            library().wpt_host_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """A scene handle: device resident (default) or host-only (inspection)."""

    def __init__(self, desc: SceneDesc, device: bool = True, owner=None):
        lib = library()
        self._owner = owner  # keeps the description's arrays alive
        self._desc = desc
        self._h = C.c_void_p()
        fn = lib.wpt_scene_create if device else lib.wpt_scene_create_host
        check(fn(C.byref(desc), C.byref(self._h)), "wpt_scene_create" if device else "wpt_scene_create_host")
        self.device = device

    @property
    def handle(self):
        return self._h

    def info(self) -> SceneInfo:
        info = SceneInfo()
        check(library().wpt_scene_info_get(self._h, C.byref(info)), "wpt_scene_info_get")
        return info

    def list(self, which: int) -> np.ndarray:
        n = C.c_uint32()
        check(library().wpt_scene_get_list(self._h, which, None, 0, C.byref(n)), "wpt_scene_get_list")
        out = np.zeros(n.value, dtype=np.uint32)
        check(library().wpt_scene_get_list(self._h, which, out.ctypes.data, n.value, C.byref(n)), "wpt_scene_get_list")
        return out

    def export_bvh2(self, which: int) -> np.ndarray:
        n = C.c_uint32()
        check(library().wpt_scene_export_bvh2(self._h, which, None, 0, C.byref(n)), "wpt_scene_export_bvh2")
        out = (Bvh2Node * max(1, n.value))()
        check(library().wpt_scene_export_bvh2(self._h, which, C.cast(out, C.c_void_p), n.value, C.byref(n)),
              "wpt_scene_export_bvh2")
        return out, n.value

    def lights(self):
        n = C.c_uint32()
        check(library().wpt_scene_export_lights(self._h, None, None, 0, C.byref(n)), "wpt_scene_export_lights")
        lights = np.zeros((n.value, 16), dtype=np.float32)
        alias = np.zeros((n.value, 2), dtype=np.float32)
        check(library().wpt_scene_export_lights(self._h, lights.ctypes.data, alias.ctypes.data, n.value, C.byref(n)),
              "wpt_scene_export_lights")
        return lights, alias

    def environment_tables(self):
        """(probabilities, alias probability, alias index) of the scene's map."""
        n = C.c_uint32()
        check(library().wpt_scene_export_environment(self._h, None, None, 0, C.byref(n)),
              "wpt_scene_export_environment")
        prob = np.zeros(n.value, dtype=np.float32)
        alias = np.zeros((n.value, 2), dtype=np.float32)
        check(library().wpt_scene_export_environment(self._h, prob.ctypes.data, alias.ctypes.data, n.value,
                                                      C.byref(n)), "wpt_scene_export_environment")
        return prob, alias[:, 0].copy(), alias[:, 1].copy().view(np.uint32)

    def light_indices(self):
        d = self._desc
        inst = np.zeros(d.instance_count, dtype=np.uint32)
        static = np.zeros(d.static_triangle_count, dtype=np.uint32)
        mesh = np.zeros(d.mesh_triangle_count, dtype=np.uint32)
        check(library().wpt_scene_export_light_indices(self._h, inst.ctypes.data, static.ctypes.data,
                                                        mesh.ctypes.data), "wpt_scene_export_light_indices")
        return inst, static, mesh

    def trace_closest(self, rays6, stream=None):
        """rays6: float32 CUDA tensor (6, n). Returns a uint8 tensor of n wpt_hit records."""
        import torch
        n = rays6.shape[1]
        hits = torch.empty((max(n, 1), HIT_DTYPE.itemsize), dtype=torch.uint8, device=rays6.device)
        check(library().wpt_trace_closest(self._h, C.c_void_p(rays6.data_ptr()), n, C.c_void_p(hits.data_ptr()),
                                          C.c_void_p(stream or 0)), "wpt_trace_closest")
        return hits[:n]

    def trace_occluded(self, rays6, tmax, stream=None):
        import torch
        n = rays6.shape[1]
        out = torch.empty(max(n, 1), dtype=torch.uint8, device=rays6.device)
        check(library().wpt_trace_occluded(self._h, C.c_void_p(rays6.data_ptr()), C.c_void_p(tmax.data_ptr()), n,
                                           C.c_void_p(out.data_ptr()), C.c_void_p(stream or 0)), "wpt_trace_occluded")
        return out[:n]

    def close(self):
        if self._h:
            library().wpt_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hits_to_numpy(hits_u8) -> np.ndarray:
    return hits_u8.cpu().numpy().view(HIT_DTYPE).reshape(-1)


def compact(items, flags, capacity):
    """Order-preserving compaction on the device (wpt_compact)."""
    import torch
    n = items.numel()
    out = torch.full((max(capacity, 1),), 0xCDCDCDCD - (1 << 32), dtype=torch.int32, device=items.device)
    count = torch.zeros(1, dtype=torch.int32, device=items.device)
    check(library().wpt_compact(C.c_void_p(items.data_ptr()), C.c_void_p(flags.data_ptr()), n,
                                C.c_void_p(out.data_ptr()), capacity, C.c_void_p(count.data_ptr()), None),
          "wpt_compact")
    return out, int(count.item())


def render_config(width, height, spp=1, max_depth=8, rr_start=3, block_size=256, rng=RNG_PCG32, clamp=10.0,
                  mode=MODE_GRAPH, tile=None, persistent_blocks=0, intersection_debug=False,
                  profile_queues=True, caustic_photons=0, caustic_max_depth=8, caustic_radius=0.06,
                  caustic_gather=True, caustic_spectral=False, caustic_grid=None, samples_per_pass=0) -> RenderConfig:
    rc = RenderConfig()
    rc.width, rc.height = width, height
    rc.tiled = 1 if tile else 0
    rc.tile_width, rc.tile_height = tile if tile else (width, height)
    rc.max_depth, rc.rr_start_depth, rc.samples_per_pixel = max_depth, rr_start, spp
    rc.block_size, rc.rng_strategy, rc.indirect_clamp, rc.mode = block_size, rng, clamp, mode
    rc.persistent_blocks = persistent_blocks
    rc.intersection_debug = 1 if intersection_debug else 0
    rc.shading_normal_debug = 0
    rc.profile_queues = 1 if profile_queues else 0
    rc.caustic_photons, rc.caustic_max_depth, rc.caustic_gather_radius = caustic_photons, caustic_max_depth, caustic_radius
    rc.caustic_gather = 1 if caustic_gather else 0
    rc.caustic_spectral = 1 if caustic_spectral else 0
    rc.samples_per_pass = samples_per_pass
    if caustic_grid is not None:  # (resolution, (min xyz), (max xyz))
        rc.caustic_grid_resolution = caustic_grid[0]
        rc.caustic_grid_min, rc.caustic_grid_max = vec3(caustic_grid[1]), vec3(caustic_grid[2])
    return rc


class Renderer:
    def __init__(self, scene: Scene, camera: Camera, config: RenderConfig, fingerprint: int = 0):
        self.scene, self.camera, self.config = scene, camera, config
        self.fingerprint = fingerprint
        self._h = C.c_void_p()
        check(library().wpt_renderer_create(scene.handle, C.byref(camera), C.byref(config), fingerprint,
                                            C.byref(self._h)), "wpt_renderer_create")

    def reset(self):
        check(library().wpt_renderer_reset(self._h), "wpt_renderer_reset")

    def render_samples(self, first: int, count: int):
        check(library().wpt_renderer_render_samples(self._h, first, count), "wpt_renderer_render_samples")

    def framebuffer(self):
        c = self.config
        out = np.zeros((c.height, c.width, 3), dtype=np.float32)
        n = C.c_uint32()
        check(library().wpt_renderer_read_framebuffer(self._h, out.ctypes.data, C.byref(n)),
              "wpt_renderer_read_framebuffer")
        return out, n.value

    def load_framebuffer(self, rgb: np.ndarray, sample_count: int):
        """Restore sums and the sample counter, to resume at sample_count."""
        rgb = np.ascontiguousarray(rgb, np.float32)
        check(library().wpt_renderer_load_framebuffer(self._h, rgb.ctypes.data, sample_count),
              "wpt_renderer_load_framebuffer")

    def render_frame(self):
        c = self.config
        out = np.zeros((c.height, c.width, 3), dtype=np.float32)
        n = C.c_uint32()
        check(library().wpt_render_frame(self._h, out.ctypes.data, C.byref(n)), "wpt_render_frame")
        return out, n.value

    def render_task(self, task: Task):
        out = np.zeros((task.height, task.width, 3), dtype=np.float32)
        check(library().wpt_render_task(self._h, C.byref(task), out.ctypes.data), "wpt_render_task")
        return out

    def profile_kernels(self, first: int, count: int):
        """Per-kernel device milliseconds and launch counts (event-bracketed)."""
        ms = np.zeros(7, dtype=np.float64)
        launches = np.zeros(7, dtype=np.uint32)
        check(library().wpt_renderer_profile_kernels(self._h, first, count, ms.ctypes.data, launches.ctypes.data),
              "wpt_renderer_profile_kernels")
        return ms, launches

    def launch_info(self) -> dict:
        """Launch shape and compiled resources of the dominant kernel."""
        info = LaunchInfo()
        check(library().wpt_renderer_launch_info(self._h, C.byref(info)), "wpt_renderer_launch_info")
        return {n: int(getattr(info, n)) for n, _ in LaunchInfo._fields_}

    def stats(self) -> Stats:
        s = Stats()
        check(library().wpt_renderer_last_stats(self._h, C.byref(s)), "wpt_renderer_last_stats")
        return s

    def close(self):
        if self._h:
            library().wpt_renderer_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_task(scene_index, camera, width, height, x, y, tile_w, tile_h, sample_start, sample_count, max_depth=8,
               rr_start=3, rng=RNG_PCG32, clamp=10.0) -> BatchTask:
    return BatchTask(scene_index, camera, width, height, x, y, tile_w, tile_h, sample_start, sample_count, max_depth,
                     rr_start, rng, clamp)


class Batch:
    """Many tasks of several device scenes rendered together (wpt_batch_*):
    staged wavefront passes over many tasks (BATCH_STAGED) or one persistent
    launch (BATCH_REGENERATE).  The scenes must outlive the batch."""

    def __init__(self, scenes: Sequence["Scene"], block_size: int = 256, blocks: int = 0,
                 scheduler: int = BATCH_STAGED, pass_paths: int = 0):
        self._scenes = list(scenes)
        handles = (C.c_void_p * len(self._scenes))(*[sc.handle for sc in self._scenes])
        self._h = C.c_void_p()
        cfg = BatchConfig(block_size, blocks, scheduler, pass_paths)
        check(library().wpt_batch_create(handles, len(self._scenes), C.byref(cfg), C.byref(self._h)),
              "wpt_batch_create")

    def set_caustics(self, scene_index: int, config: RenderConfig):
        """Builds scene `scene_index`'s caustic photon map from the caustic
        fields of a render configuration (wpt_batch_set_caustics); returns
        (photons stored, build ms)."""
        stored = C.c_uint32()
        ms = C.c_double()
        check(library().wpt_batch_set_caustics(self._h, scene_index, C.byref(config), C.byref(stored), C.byref(ms)),
              "wpt_batch_set_caustics")
        return int(stored.value), float(ms.value)

    @staticmethod
    def _tasks(tasks):
        return (BatchTask * max(1, len(tasks)))(*tasks)

    def render(self, tasks: Sequence[BatchTask], latencies: bool = False):
        """Returns a list of (tile_height, tile_width, 3) float32 sums, and the
        per-task latencies in ms when asked."""
        arr = self._tasks(tasks)
        outs = [np.zeros((t.tile_height, t.tile_width, 3), np.float32) for t in tasks]
        ptrs = (C.c_void_p * max(1, len(tasks)))(*[o.ctypes.data for o in outs])
        lat = np.zeros(max(1, len(tasks)), np.float64)
        check(library().wpt_batch_render(self._h, arr, len(tasks), ptrs, lat.ctypes.data if latencies else None),
              "wpt_batch_render")
        return (outs, lat[:len(tasks)]) if latencies else outs

    def autotune(self, tasks: Sequence[BatchTask]) -> BatchConfig:
        arr = self._tasks(tasks)
        chosen = BatchConfig()
        check(library().wpt_batch_autotune(self._h, arr, len(tasks), C.byref(chosen)), "wpt_batch_autotune")
        return chosen

    def stats(self) -> BatchStats:
        st = BatchStats()
        check(library().wpt_batch_last_stats(self._h, C.byref(st)), "wpt_batch_last_stats")
        return st

    def close(self):
        if self._h:
            library().wpt_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DiffRenderer:
    """Differentiable renderer (wpt_diff_*): forward image and path-replay
    adjoint over one parameter vector [texture | roughness | emission].
    Device buffers are torch tensors on cuda:0; every call synchronizes.
    torch ships its own HIP runtime: initialise torch's device (any CUDA
    tensor) before this library's first device call in the process."""

    def __init__(self, scene: "Scene", camera: Camera, width: int, height: int, spp: int, max_depth: int,
                 texture_size=(0, 0), textured_material: int = INVALID, rng: int = RNG_PCG32, block_size: int = 256):
        self.scene, self.camera = scene, camera
        self.width, self.height, self.spp = width, height, spp
        cfg = DiffConfig(width, height, spp, max_depth, rng, texture_size[0], texture_size[1], textured_material,
                         block_size)
        self._h = C.c_void_p()
        check(library().wpt_diff_create(scene.handle, C.byref(camera), C.byref(cfg), C.byref(self._h)),
              "wpt_diff_create")
        self.layout = DiffLayout()
        check(library().wpt_diff_get_layout(self._h, C.byref(self.layout)), "wpt_diff_get_layout")

    @staticmethod
    def _ptr(t):
        import torch
        assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
        return C.c_void_p(t.data_ptr())

    def forward(self, params, sample_start: int = 0):
        import torch
        torch.cuda.synchronize()
        image = torch.empty((self.height, self.width, 3), dtype=torch.float32, device=params.device)
        check(library().wpt_diff_forward(self._h, self._ptr(params), sample_start, self._ptr(image)),
              "wpt_diff_forward")
        return image

    def adjoint(self, params, image_grad):
        import torch
        torch.cuda.synchronize()
        grad = torch.empty_like(params)
        check(library().wpt_diff_adjoint(self._h, self._ptr(params), self._ptr(image_grad.contiguous()),
                                         self._ptr(grad)), "wpt_diff_adjoint")
        return grad

    def adam(self, params, grad, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        import torch
        torch.cuda.synchronize()
        check(library().wpt_diff_adam(self._h, self._ptr(params), self._ptr(grad), self._ptr(m), self._ptr(v), step,
                                      lr, beta1, beta2, eps), "wpt_diff_adam")

    def stats(self) -> DiffStats:
        st = DiffStats()
        check(library().wpt_diff_last_stats(self._h, C.byref(st)), "wpt_diff_last_stats")
        return st

    def close(self):
        if self._h:
            library().wpt_diff_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def diff_render_function():
    """torch.autograd.Function over DiffRenderer: image = f(params), the
    backward is the path-replay adjoint.  For checking the adjoint inside
    torch; the renderer itself does not need torch."""
    import torch

    class DiffRender(torch.autograd.Function):
        @staticmethod
        def forward(ctx, params, renderer, sample_start):
            params = params.detach().contiguous()
            ctx.renderer = renderer
            ctx.save_for_backward(params)
            return renderer.forward(params, sample_start)

        @staticmethod
        def backward(ctx, image_grad):
            (params,) = ctx.saved_tensors
            return ctx.renderer.adjoint(params, image_grad.contiguous()), None, None

    return DiffRender


def create_tasks(width, height, tile_w, tile_h, spp, per_task, fingerprint) -> List[Task]:
    lib = library()
    n = lib.wpt_host_create_tasks(width, height, tile_w, tile_h, spp, per_task, fingerprint, None, 0)
    arr = (Task * max(1, n))()
    lib.wpt_host_create_tasks(width, height, tile_w, tile_h, spp, per_task, fingerprint, C.cast(arr, C.c_void_p), n)
    return [arr[i] for i in range(n)]


def load_hdr(filename) -> np.ndarray:
    """A Radiance .hdr file as float32 (height, width, 3)."""
    w, h = C.c_uint32(), C.c_uint32()
    check(library().wpt_host_load_hdr(str(filename).encode(), None, 0, C.byref(w), C.byref(h)), "load_hdr")
    out = np.zeros((h.value, w.value, 3), np.float32)
    check(library().wpt_host_load_hdr(str(filename).encode(), out.ctypes.data, out.size, C.byref(w), C.byref(h)),
          "load_hdr")
    return out


def decode_image(data: bytes) -> np.ndarray:
    """An encoded image (PNG) as uint8 (height, width, 4) RGBA."""
    buf = np.frombuffer(bytes(data), np.uint8)
    w, h = C.c_uint32(), C.c_uint32()
    check(library().wpt_host_decode_image(buf.ctypes.data, buf.size, None, 0, C.byref(w), C.byref(h)), "decode_image")
    out = np.zeros((h.value, w.value, 4), np.uint8)
    check(library().wpt_host_decode_image(buf.ctypes.data, buf.size, out.ctypes.data, out.size, C.byref(w),
                                          C.byref(h)), "decode_image")
    return out


def write_render_checkpoint(filename: str, rgb: np.ndarray, sample_index: int, fingerprint: int) -> None:
    """A "PTCHKPT1" render checkpoint of an (h, w, 3) float32 sum image."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w = rgb.shape[:2]
    check(library().wpt_host_render_checkpoint_write(str(filename).encode(), rgb.ctypes.data, w, h, sample_index,
                                                     fingerprint), "render checkpoint write")


def read_render_checkpoint(filename: str, width: int, height: int, fingerprint: int):
    """(sums (h, w, 3), completed samples); WptError on any mismatch."""
    out = np.zeros((height, width, 3), np.float32)
    n = C.c_uint32()
    check(library().wpt_host_render_checkpoint_read(str(filename).encode(), width, height, fingerprint,
                                                    out.ctypes.data, C.byref(n)), "render checkpoint read")
    return out, n.value


class Ledger:
    """Task ledger of one job: leases, merged tile sums, checkpoints."""

    def __init__(self, fingerprint, width, height, tile_w, tile_h, spp, per_task, tasks=None):
        self.job = LedgerJob(fingerprint, width, height, tile_w, tile_h, spp, per_task)
        if tasks is None:
            tasks = create_tasks(width, height, tile_w, tile_h, spp, per_task, fingerprint)
        self.tasks = list(tasks)
        arr = (Task * max(1, len(self.tasks)))(*self.tasks)
        h = C.c_void_p()
        check(library().wpt_ledger_create(C.byref(self.job), C.cast(arr, C.c_void_p), len(self.tasks), C.byref(h)),
              "ledger create")
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None):
            library().wpt_ledger_destroy(self.handle)
            self.handle = None

    def lease_next(self, worker: str, now_ms: int, lease_ms: int):
        t, expires = Task(), C.c_uint64()
        r = library().wpt_ledger_lease_next(self.handle, worker.encode(), now_ms, lease_ms, C.byref(t),
                                            C.byref(expires))
        if r < 0:
            raise WptError(-r, "ledger lease")
        return (t, expires.value) if r == 1 else None

    def commit(self, task: Task, sample_count: int, radiance: np.ndarray) -> int:
        rad = np.ascontiguousarray(radiance, np.float32)
        return library().wpt_ledger_commit(self.handle, C.byref(task), sample_count, rad.ctypes.data, rad.size)

    def renew(self, worker, now_ms, lease_ms):
        return library().wpt_ledger_renew_leases(self.handle, worker.encode(), now_ms, lease_ms)

    def release(self, worker):
        return library().wpt_ledger_release_leases(self.handle, worker.encode())

    def requeue_expired(self, now_ms):
        return library().wpt_ledger_requeue_expired(self.handle, now_ms)

    def state(self, task: Task) -> int:
        return library().wpt_ledger_task_state(self.handle, C.byref(task))

    def counts(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(library().wpt_ledger_counts(self.handle, C.byref(a), C.byref(b), C.byref(c)), "ledger counts")
        return a.value, b.value, c.value

    def image(self):
        rad = np.zeros((self.job.height, self.job.width, 3), np.float32)
        cnt = np.zeros((self.job.height, self.job.width), np.uint32)
        check(library().wpt_ledger_read(self.handle, rad.ctypes.data, cnt.ctypes.data), "ledger read")
        return rad, cnt

    def write_checkpoint(self, filename, sequence):
        check(library().wpt_ledger_write_checkpoint(self.handle, str(filename).encode(), sequence),
              "ledger checkpoint write")

    def restore_checkpoint(self, filename) -> int:
        seq = C.c_uint64()
        check(library().wpt_ledger_restore_checkpoint(self.handle, str(filename).encode(), C.byref(seq)),
              "ledger checkpoint restore")
        return seq.value


class LedgerWriter:
    """Background ledger checkpoint writer (newest waiting snapshot wins)."""

    def __init__(self, filename):
        h = C.c_void_p()
        check(library().wpt_ledger_writer_create(str(filename).encode(), C.byref(h)), "ledger writer create")
        self.handle = h

    def enqueue(self, ledger: Ledger, sequence: int):
        check(library().wpt_ledger_writer_enqueue(self.handle, ledger.handle, sequence), "ledger writer enqueue")

    def flush(self):
        check(library().wpt_ledger_writer_flush(self.handle), "ledger writer flush")

    def close(self):
        if self.handle:
            h, self.handle = self.handle, None
            check(library().wpt_ledger_writer_destroy(h), "ledger writer close")

    def __del__(self):
        if getattr(self, "handle", None):
            self.close()


def scene_fingerprint(preset_or_manifest: str) -> int:
    """Dataset identity of a preset or manifest (no display name)."""
    out = C.c_uint64()
    check(library().wpt_host_scene_fingerprint(preset_or_manifest.encode(), C.byref(out)), "scene_fingerprint")
    return out.value


def task_key(fingerprint: int, task: Task) -> int:
    return int(library().wpt_host_task_key(fingerprint, C.byref(task)))


def parse_command_line(args: Iterable[str]):
    argv = ["pathtracer"] + list(args)
    arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
    rc = RenderConfig()
    name = C.create_string_buffer(128)
    fp = C.c_uint64()
    status = library().wpt_host_parse_command_line(len(argv), arr, C.byref(rc), name, 128, C.byref(fp))
    return status, rc, name.value.decode(), fp.value


def sha256(buf: np.ndarray) -> str:
    out = C.create_string_buffer(65)
    library().wpt_host_sha256(buf.ctypes.data, buf.nbytes, out)
    return out.value.decode()
