// Environment light: emission lookup and importance sampling by CDF binary
// search (reference: Src/Core/texture.cu:17-166, Src/Core/light.cu:91-155) or,
// as an opt-in fast mode, by alias tables over the same texel distribution.
#pragma once

#include "device.h"
#include "rt_math.h"
#include "shade.h"

namespace rt {

// Bilinear lookup at normalised (u, v): U wraps, V clamps, texel centres at
// half-integers (the addressing of a linear-filtered, normalised-coordinate
// texture).  Weights are FP32.
RT_HD vec3 env_texel(const DevScene& sc, int x, int y) {
    Global<float>* p = &gref(sc.env_rgb + 3 * ((size_t)y * sc.env_w + x));
    return v3(p[0], p[1], p[2]);
}
RT_HD vec3 env_bilinear(const DevScene& sc, float u, float v) {
    float x = u * (float)sc.env_w - 0.5f;
    float y = v * (float)sc.env_h - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y);
    float ax = x - fx0, ay = y - fy0;
    int ix0 = (int)fx0, iy0 = (int)fy0;
    int ix1 = ix0 + 1, iy1 = iy0 + 1;
    ix0 %= sc.env_w; if (ix0 < 0) ix0 += sc.env_w;
    ix1 %= sc.env_w; if (ix1 < 0) ix1 += sc.env_w;
    iy0 = clampi(0, sc.env_h - 1, iy0);
    iy1 = clampi(0, sc.env_h - 1, iy1);
    vec3 t00 = env_texel(sc, ix0, iy0), t10 = env_texel(sc, ix1, iy0);
    vec3 t01 = env_texel(sc, ix0, iy1), t11 = env_texel(sc, ix1, iy1);
    vec3 top = t00 * (1.f - ax) + t10 * ax;
    vec3 bot = t01 * (1.f - ax) + t11 * ax;
    return top * (1.f - ay) + bot * ay;
}

// Light::Emission for the environment: u = phi / 2PI, v = theta / PI with y up.
RT_HD vec3 env_emission(const DevScene& sc, vec3 wi) {
    float theta = fm::acosf_(wi.y);
    float phi = fm::atan2f_(wi.x, wi.z);
    if (phi < 0) phi = (float)((double)phi + 2.0 * kPiD);
    float u = (float)((double)phi / (2.0 * kPiD));
    float v = (float)((double)theta / kPiD);
    return env_bilinear(sc, u, v);
}

// HDRTexture::SamplePDF: piecewise-constant pdf in uv-space.
RT_HD float env_pdf_uv(const DevScene& sc, float u, float v) {
    u = clampf(0.0f, 0.999999f, u);
    v = clampf(0.0f, 0.999999f, v);
    int ui = (int)(u * (float)sc.env_w);
    if (ui > sc.env_w - 1) ui = sc.env_w - 1;
    int vi = (int)(v * (float)sc.env_h);
    if (vi > sc.env_h - 1) vi = sc.env_h - 1;
    Global<float>* row = &gref(sc.env_row_cdf + (size_t)vi * sc.env_w);
    Global<float>* marg = &gref(sc.env_marg_cdf);
    float p_row = vi == 0 ? marg[0] : marg[vi] - marg[vi - 1];
    float p_col = ui == 0 ? row[0] : row[ui] - row[ui - 1];
    float p = p_row * p_col;
    return (p * (float)sc.env_w) * (float)sc.env_h;
}

// Two binary searches (row, then column) over the CDFs (texture.cu:112-153).
RT_HD void env_sample_uv(const DevScene& sc, float rv, float ru, float& u, float& v) {
    int left = 0, right = sc.env_h - 1;
    while (left <= right) {
        int mid = (left + right) / 2;
        if (gref(sc.env_marg_cdf + mid) < rv) left = mid + 1;
        else right = mid - 1;
    }
    int vi = left;
    Global<float>* row = &gref(sc.env_row_cdf + (size_t)(vi < sc.env_h ? vi : sc.env_h - 1) * sc.env_w);
    left = 0;
    right = sc.env_w - 1;
    while (left <= right) {
        int mid = (left + right) / 2;
        if (row[mid] < ru) left = mid + 1;
        else right = mid - 1;
    }
    int ui = left;
    u = ((float)ui + 0.5f) / (float)sc.env_w;
    v = ((float)vi + 0.5f) / (float)sc.env_h;
}

// O(1) pick from an alias table of n entries: cell floor(u n), accepted when
// the fractional part is below the cell's threshold, else its alias.
RT_HD int alias_pick(const float2* t, int n, float u) {
    float x = u * (float)n;
    int i = (int)x;
    if (i > n - 1) i = n - 1;
    float f = x - (float)i;
    const float* e = reinterpret_cast<const float*>(t + i);  // (threshold, alias bits)
    Global<float>* g = &gref(e);
    return f < g[0] ? i : f_as_i(g[1]);
}

// Alias-table sampling of the same texel distribution as env_sample_uv (row
// from the marginal table, column from the row's table), at the texel centre.
RT_HD void env_sample_uv_alias(const DevScene& sc, float rv, float ru, float& u, float& v) {
    int vi = alias_pick(sc.env_alias_marg, sc.env_h, rv);
    int ui = alias_pick(sc.env_alias_row + (size_t)vi * sc.env_w, sc.env_w, ru);
    u = ((float)ui + 0.5f) / (float)sc.env_w;
    v = ((float)vi + 0.5f) / (float)sc.env_h;
}

// Light::SampleEnvLight: a point on a sphere of radius 2*diag around the scene
// centre, polar angle acos(1 - 2v), normal towards the centre.
RT_HD void env_sample_point(const DevScene& sc, float rv, float ru, vec3& pos, vec3& nrm) {
    float u, v;
    if (sc.env_alias) env_sample_uv_alias(sc, rv, ru, u, v);
    else env_sample_uv(sc, rv, ru, u, v);
    float theta = (float)((double)(u * 2.f) * kPiD);
    float phi = fm::acosf_(1.f - 2.f * v);
    float x = fm::sinf_(phi) * fm::sinf_(theta);
    float z = fm::sinf_(phi) * fm::cosf_(theta);
    float y = fm::cosf_(phi);
    vec3 dir = v3(x, y, z);
    vec3 center = (sc.bounds_lo + sc.bounds_hi) * 0.5f;
    float diag = length(sc.bounds_hi - sc.bounds_lo);
    pos = center + (dir * 2.f) * diag;
    nrm = -dir;
}

// Light::SampleEnvLightSolidAnglePDF: note the flipped v (light.cu:125-139).
RT_HD float env_pdf_from(const DevScene& sc, vec3 origin, vec3 lp) {
    vec3 dir = normalize(lp - origin);
    float theta = fm::acosf_(dir.y);
    float phi = fm::atan2f_(dir.x, dir.z);
    if (phi < 0) phi = (float)((double)phi + 2.0 * kPiD);
    float u = (float)((double)phi / (2.0 * kPiD));
    float v = (float)(1.0 - (double)theta / kPiD);
    return env_pdf_uv(sc, u, v);
}

}  // namespace rt
