// File loaders of the host layer: Wavefront OBJ/MTL meshes and Radiance HDR
// environment maps, producing the scene description's arrays.
//
// OBJ follows the reference's loadObjScene (src/scene/obj_loader.cpp:47-198,
// over tinyobjloader with triangulation): one mesh per object/group with
// faces, material 0 the reference's default (Lambertian 0.65 0.3 0.18) and
// MTL materials from 1 in file order, Lambertian or emissive (Ke != 0) with
// Kd, Ke and Ni; every triangle gets vertex normals, the file's when all
// three are given, else the normalized sum of the unnormalized face normals
// around the vertex; texture coordinates when all three are given.
// Polygons are split as fans from their first vertex.  Image textures
// (map_Kd, relative to the OBJ file, shared between materials naming the
// same file) are decoded to RGBA8 albedo textures; an unreadable file is an
// error, an image in a format this build does not decode (JPEG) leaves the
// material with its Kd and a warning.
//
// HDR follows stb_image's Radiance decoder, which the reference's
// loadEnvironmentMap uses (src/scene/environment.cpp:66-82): "#?RADIANCE" or
// "#?RGBE" header, FORMAT=32-bit_rle_rgbe, "-Y h +X w" rows top to bottom,
// flat or new-style run-length scanlines, and texel (r, g, b, e) ->
// (r, g, b) * 2^(e - 136) (zero when e = 0).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/wpt.h"

namespace wpt {

struct ObjScene {
    std::vector<std::vector<wpt_triangle>> meshes;
    std::vector<wpt_material> materials;   // [0] default, then the MTL's
    std::vector<std::string> albedoMaps;   // per material; empty: none
    // Decoded map_Kd images (one per distinct file name) and each
    // material's texture index (WPT_INVALID_INDEX: none).
    std::vector<std::vector<uint8_t>> textureData;
    std::vector<std::pair<uint32_t, uint32_t>> textureSizes;  // width, height
    std::vector<uint32_t> albedoTexture;
};

bool loadObjScene(const std::string& filename, ObjScene& out, std::string* error);

// glTF 2.0 (src/scene/gltf_loader.cpp:307-503): one mesh asset per indexed
// triangle primitive with FLOAT positions, normals, tangents and two
// texture-coordinate sets (v flipped to 1 - v) and face tangent frames per
// set; instances and KHR_lights_punctual lights from the default scene's
// node trees (matrix or TRS, composed in double precision); material 0 the
// reference's default (Lambertian 0.7) and the file's materials from 1 with
// the metallic-roughness, specular-glossiness, transmission/volume, ior,
// specular and clearcoat fields and their texture bindings; images decoded
// to RGBA8 in file order (texture bindings name images).  Malformed input is
// an error (the reference exits or throws).
struct GltfScene {
    std::vector<std::vector<wpt_triangle>> meshes;
    std::vector<wpt_instance> instances;  // mesh = index into meshes
    std::vector<wpt_material> materials;
    std::vector<wpt_material_ext> materialExt;
    std::vector<std::vector<uint8_t>> textureData;
    std::vector<std::pair<uint32_t, uint32_t>> textureSizes;  // width, height
    std::vector<wpt_punctual_light> punctualLights;
    uint32_t specularGlossinessMaterials = 0;
};

bool loadGltfScene(const std::string& filename, GltfScene& out, std::string* error);

bool loadRadianceHdr(const std::string& filename, std::vector<float>& rgb, uint32_t& width, uint32_t& height,
    std::string* error);

}  // namespace wpt
