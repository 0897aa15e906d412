// Diffuse (LDR) texture lookup: the albedo of a textured material at a hit's
// uv (reference: Texture::Sample, Src/Core/texture.cu:242-249, used for the
// base colour by the diffuse, GGX and subsurface lobes, material.cu:68-76,
// 205-213, 243-248, 280-285, 326-338).
#pragma once

#include "device.h"
#include "rt_math.h"

namespace rt {

RT_HD vec3 texel_rgb(uint32_t t) {
    return v3((float)(t & 255u) / 255.f, (float)((t >> 8) & 255u) / 255.f, (float)((t >> 16) & 255u) / 255.f);
}

// u and v are clamped to [0, 0.999999] and v is flipped (OBJ texture
// coordinates start at the bottom-left, image rows at the top), then the
// RGBA8 texels, read as normalised floats, are filtered bilinearly with the
// addressing of a linear-filtered, normalised-coordinate texture: U wraps,
// V clamps, texel centres at half-integers.  Weights are FP32 (the
// reference's texture unit rounds them to 8 fractional bits).
RT_HD vec3 texture_albedo(const DevScene& sc, int tex, float u, float v) {
    Global<int32_t>* info = &gref(sc.tex_info + 4 * tex);
    const uint32_t* texels = sc.tex_texels + (uint32_t)info[0];
    const int w = info[1], h = info[2];
    // Clamp (mathutil.cu:72-79); a NaN coordinate, which the reference leaves
    // to the texture unit, reads as 0.
    u = u >= 0.0f ? (u > 0.999999f ? 0.999999f : u) : 0.0f;
    v = v >= 0.0f ? (v > 0.999999f ? 0.999999f : v) : 0.0f;
    float x = u * (float)w - 0.5f;
    float y = (1.0f - v) * (float)h - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y);
    float ax = x - fx0, ay = y - fy0;
    int ix0 = (int)fx0, iy0 = (int)fy0;
    int ix1 = ix0 + 1, iy1 = iy0 + 1;
    ix0 %= w; if (ix0 < 0) ix0 += w;
    ix1 %= w; if (ix1 < 0) ix1 += w;
    iy0 = clampi(0, h - 1, iy0);
    iy1 = clampi(0, h - 1, iy1);
    vec3 t00 = texel_rgb(gref(texels + (size_t)iy0 * w + ix0)), t10 = texel_rgb(gref(texels + (size_t)iy0 * w + ix1));
    vec3 t01 = texel_rgb(gref(texels + (size_t)iy1 * w + ix0)), t11 = texel_rgb(gref(texels + (size_t)iy1 * w + ix1));
    vec3 top = t00 * (1.f - ax) + t10 * ax;
    vec3 bot = t01 * (1.f - ax) + t11 * ax;
    return top * (1.f - ay) + bot * ay;
}

}  // namespace rt
