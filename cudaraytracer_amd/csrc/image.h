// LDR image decoding for diffuse textures.  The reference loads each texture
// with stb_image 2.30's stbi_load(path, ..., 4) (Src/Core/texture.cu:207-214):
// RGBA8, row 0 = the top of the image.  The decoders here produce the same
// bytes for PNG (every colour type and bit depth, Adam7, tRNS), JPEG (baseline
// and progressive Huffman, 1/3/4 components, any sampling factors, restart
// intervals) and binary PNM (P5/P6, maxval <= 255).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace rt {

struct LdrImage {
    int w = 0, h = 0;
    std::vector<uint8_t> rgba;  // 4 * w * h bytes, row-major, row 0 = top
};

// Decode an in-memory PNG, JPEG or PNM file.  Throws std::runtime_error.
void decode_ldr_image(const uint8_t* data, size_t size, LdrImage& out);
// Read and decode a file.  Throws std::runtime_error naming the file.
void load_ldr_image(const std::string& path, LdrImage& out);

// 8-bit RGBA PNG, row 0 = the top of the image (what stbi_write_png writes for
// the reference's screenshot, Src/Utils/ui.cpp:370-380; the pixels, not the
// compressed bytes, are the same).  Throws std::runtime_error.
void write_png_rgba8(const std::string& path, int w, int h, const uint8_t* rgba);
// Colour PFM (little-endian float RGB) of an RGBA float image whose row 0 is
// the bottom row, which is PFM's own row order.  Throws std::runtime_error.
void write_pfm_rgb(const std::string& path, int w, int h, const float* rgba);

}  // namespace rt
