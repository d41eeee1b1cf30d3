/* Checker for the product's PNG and JPEG decoder: stb_image, the decoder behind the
 * reference's glTF texture import (src/scene/gltf_loader.cpp:273-305,
 * stbi_load_from_memory with STBI_rgb_alpha), compiled from the copy the
 * reference vendors (extern/stb/stb_image.h) into oracle/_ref by this
 * directory's Makefile.  Test infrastructure only.
 *
 * wpt_ref_stbi_rgba: decodes `size` bytes to RGBA8; writes width and height
 * and, when out is not NULL and capacity suffices, the pixels.  Returns
 * the pixel byte count, or -1 when stb_image rejects the data.
 *
 * wpt_ref_stbi_rgbf: the same for a Radiance .hdr file as linear RGB
 * floats (stbi_loadf with 3 channels, as the reference's
 * loadEnvironmentMap, src/scene/environment.cpp:66-82); returns the float
 * count. */
#define STB_IMAGE_IMPLEMENTATION
#define STBI_ONLY_PNG
#define STBI_ONLY_JPEG
#define STBI_ONLY_HDR
#define STBI_NO_STDIO
#include "stb_image.h"

#include <string.h>

long long wpt_ref_stbi_rgba(const unsigned char* data, int size, unsigned char* out, long long capacity, int* width,
                            int* height) {
    int channels = 0;
    unsigned char* pixels = stbi_load_from_memory(data, size, width, height, &channels, 4);
    if (!pixels) return -1;
    long long bytes = (long long)(*width) * (*height) * 4;
    if (out && capacity >= bytes) memcpy(out, pixels, (size_t)bytes);
    stbi_image_free(pixels);
    return bytes;
}

long long wpt_ref_stbi_rgbf(const unsigned char* data, int size, float* out, long long capacity, int* width,
                            int* height) {
    int channels = 0;
    float* pixels = stbi_loadf_from_memory(data, size, width, height, &channels, 3);
    if (!pixels) return -1;
    long long floats = (long long)(*width) * (*height) * 3;
    if (out && capacity >= floats) memcpy(out, pixels, (size_t)floats * sizeof(float));
    stbi_image_free(pixels);
    return floats;
}
