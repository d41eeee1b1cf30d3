// Image decoding for glTF and OBJ textures: PNG and JPEG files to 8-bit RGBA, rows top to
// bottom, as the reference's loadTexture asks stb_image for them
// (src/scene/gltf_loader.cpp:273-305, stbi_load_from_memory with
// STBI_rgb_alpha).
//
// PNG (ISO/IEC 15948): zlib/deflate streams (RFC 1950/1951), the five
// scanline filters, Adam7 interlacing, every colour type and bit depth.
// Conversions to RGBA8 are stb_image's: grey g -> (g, g, g, 255) with
// 1/2/4-bit greys scaled by 255/3/17, grey+alpha -> (g, g, g, a), palette
// entries with tRNS alphas (255 without), a tRNS colour key turns matching
// pixels' alpha to 0 (keys of sub-byte greys scaled like the pixels),
// 16-bit samples keep their high byte.  Chunk CRCs and the zlib Adler-32
// are not verified, and extra image data after the last row is ignored, as
// in stb_image.
//
// JPEG (ITU T.81): baseline and progressive Huffman coding, 8-bit samples,
// greyscale, YCbCr/RGB and CMYK/YCCK, any sampling factors, restart
// intervals; the pixels are stb_image's: its integer IDCT, its centred 2x
// upsampling filters and its fixed-point YCbCr conversion.  Lossless,
// hierarchical, arithmetic-coded and 12-bit JPEG fail, as in stb_image.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace wpt {

bool decodeImage(const uint8_t* data, size_t size, std::vector<uint8_t>& rgba, uint32_t& width, uint32_t& height,
    std::string* error);

// zlib stream (RFC 1950 header, deflate body) to bytes.  Decoding stops
// once expectedSize bytes exist (the rest of the stream is not read), so the
// output never exceeds expectedSize; a shorter output is the caller's to
// reject.
bool inflateZlib(const uint8_t* data, size_t size, std::vector<uint8_t>& out, size_t expectedSize,
    std::string* error);

}  // namespace wpt
