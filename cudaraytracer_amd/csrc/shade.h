// BSDFs, light sampling and the per-bounce shading step of the wavefront path
// tracer.  Host/device code; every formula cites the reference function it
// restates and keeps the reference's FP32 / FP64 promotion rules (the double
// literal PI = 3.1415926 promotes scalar expressions to double, while
// Vec3f-by-double operations convert the scalar to float first).
#pragma once

#include "rt_math.h"
#include "scene.h"

namespace rt {

// PI of the reference (Src/Core/global.h:7): a double literal.
constexpr double kPiD = 3.1415926;
constexpr float kPiF = (float)3.1415926;
constexpr double kTwoPiD = 2.0 * 3.1415926;
constexpr double kEpsD = 1e-6;  // eps (Src/Core/global.h:6)

struct DevMaterial {
    int32_t type;
    vec3 basecolor;
    vec3 ke;
    float roughness, metallic, transmission, ior;
    vec3 scatter;
    int32_t texture, valid;
};

// The random-walk subsurface mode's coefficients, read only by that mode.
struct DevMedium {
    vec3 sigma_a, sigma_s;
    int32_t method;
};

RT_HD DevMedium load_medium(const MaterialRec* mats, int id) {
    Global<MaterialRec>& m = gref(mats + id);
    DevMedium r;
    r.sigma_a = v3(m.sigma_a[0], m.sigma_a[1], m.sigma_a[2]);
    r.sigma_s = v3(m.sigma_s[0], m.sigma_s[1], m.sigma_s[2]);
    r.method = m.sss_method;
    return r;
}

RT_HD DevMaterial load_material(const MaterialRec* mats, int id) {
    Global<MaterialRec>& m = gref(mats + id);
    DevMaterial r;
    r.type = m.type;
    r.basecolor = v3(m.basecolor[0], m.basecolor[1], m.basecolor[2]);
    r.ke = v3(m.ke[0], m.ke[1], m.ke[2]);
    r.roughness = m.roughness;
    r.metallic = m.metallic;
    r.transmission = m.transmission;
    r.ior = m.ior;
    r.scatter = v3(m.scatter[0], m.scatter[1], m.scatter[2]);
    r.texture = m.texture;
    r.valid = m.valid;
    return r;
}

// Power heuristic with beta = 2: a^2 / (a^2 + b^2) (reference: Src/Core/renderer.cuh:165-168).
RT_HD float mis_weight(float a, float b) {
    float t1 = a * a, t2 = b * b;
    return t1 / (t1 + t2);
}

// ---------------------------------------------------------------- diffuse
// Argument convention of the reference: wi = towards the viewer, wo = the
// scattered / light direction (reference: Src/Core/material.cu:68-94).
RT_HD vec3 eval_diffuse(const DevMaterial& m, vec3 wi, vec3 n, vec3 wo) {
    (void)wi;
    if (dot(wo, n) > 0.f) return m.basecolor / kPiF;
    return v3(0.f, 0.f, 0.f);
}
RT_HD float pdf_diffuse(vec3 n, vec3 wo) {
    float c = dot(wo, n);
    if (c > 0.f) return (float)((double)c / kPiD);
    return 0.f;
}
// Cosine-weighted direction: theta = acos(1 - 2 t1) / 2, phi = 2 PI t2 (double).
RT_HD vec3 sample_diffuse_dir(vec3 n, float t1, float t2) {
    float theta = 0.5f * fm::acosf_(1.f - 2.f * t1);
    float phi = (float)(kTwoPiD * (double)t2);
    float st = fm::sinf_(theta);
    float x = st * fm::cosf_(phi), y = st * fm::sinf_(phi), z = fm::cosf_(theta);
    return local_to_world(v3(x, y, z), n);
}

// ---------------------------------------------------------------- GGX PBR
// (reference: Src/Core/material.cu:98-114, 193-324)
RT_HD float ggx_d(const DevMaterial& m, vec3 wi, vec3 n, vec3 wo) {
    vec3 h = normalize(wi + wo);
    float ndoth = fmaxf(dot(n, h), 0.f);
    float alpha = m.roughness * m.roughness;
    float alpha2 = alpha * alpha;
    float ndoth2 = ndoth * ndoth;
    float denom = ndoth2 * (alpha2 - 1.f) + 1.f;
    denom = (float)((kPiD * (double)denom) * (double)denom);
    return alpha2 / denom;
}
RT_HD vec3 fresnel_schlick(vec3 wi, vec3 wo, vec3 f0) {
    vec3 h = normalize(wi + wo);
    float hdotv = fmaxf(dot(h, wi), 0.f);
    float p = fm::pow5(1.f - hdotv);
    return f0 + (v3(1.f, 1.f, 1.f) - f0) * p;
}
RT_HD float geometry_schlick(const DevMaterial& m, vec3 wi, vec3 n, vec3 wo) {
    float ndotv = fmaxf(dot(n, wi), 0.f);
    float ndotl = fmaxf(dot(n, wo), 0.f);
    float r = m.roughness + 1.f;
    float k = (r * r) / 8.f;
    float g1v = ndotv / (ndotv * (1.f - k) + k);
    float g1l = ndotl / (ndotl * (1.f - k) + k);
    return g1v * g1l;
}
RT_HD vec3 eval_pbr(const DevMaterial& m, vec3 wi, vec3 n, vec3 wo) {
    float ndotv = fmaxf(dot(n, wi), 0.0f);
    float ndotl = fmaxf(dot(n, wo), 0.0f);
    if (ndotl <= 0.0f || ndotv <= 0.0f) return v3(0.f, 0.f, 0.f);
    vec3 albedo = m.basecolor;
    vec3 f0 = lerp(v3(0.04f, 0.04f, 0.04f), albedo, m.metallic);
    float d = ggx_d(m, wi, n, wo);
    float g = geometry_schlick(m, wi, n, wo);
    vec3 f = fresnel_schlick(wi, wo, f0);
    vec3 numerator = f * (d * g);
    float denominator = (4.0f * ndotv) * ndotl + 0.0001f;
    vec3 specular = numerator / denominator;
    vec3 kd = (v3(1.f, 1.f, 1.f) - f) * (1.0f - m.metallic);
    vec3 diffuse = v3(0.f, 0.f, 0.f);
    if (dot(wi, n) > 0.f && dot(wo, n) > 0.f) diffuse = cmul(kd, albedo) / kPiF;
    return diffuse + specular;
}
RT_HD float pbr_spec_prob(const DevMaterial& m, vec3 wi, vec3 n) {
    vec3 albedo = m.basecolor;
    vec3 f0 = lerp(v3(0.04f, 0.04f, 0.04f), albedo, m.metallic);
    float ndotv = fmaxf(dot(n, wi), 0.0f);
    vec3 fa = f0 + (v3(1.f, 1.f, 1.f) - f0) * fm::pow5(1.0f - ndotv);
    float spec = luminance(fa);
    float diff = luminance(albedo) * (1.0f - m.metallic);
    float p = fmaxf(spec / (spec + diff), 1e-6f);
    return clampf(0.01f, 0.99f, p);
}
RT_HD float pdf_pbr(const DevMaterial& m, vec3 wi, vec3 n, vec3 wo) {
    float p_spec = pbr_spec_prob(m, wi, n);
    float ndotl = fmaxf(dot(n, wo), 0.0f);
    if (ndotl <= 0.0f) return 0.f;
    vec3 h = normalize(wi + wo);
    float vdoth = fmaxf(dot(wi, h), 1e-6f);
    float ndoth = fmaxf(dot(n, h), 1e-6f);
    float d = ggx_d(m, wi, n, wo);
    float pdf_spec = (d * ndoth) / (4.0f * vdoth);
    float pdf_diff = (float)((double)ndotl / kPiD);
    return p_spec * pdf_spec + (1.0f - p_spec) * pdf_diff;
}
RT_HD vec3 sample_pbr_dir(const DevMaterial& m, vec3 wi, vec3 n, float r1, float r2, float choose) {
    float p_spec = pbr_spec_prob(m, wi, n);
    float phi = (float)(kTwoPiD * (double)r1);
    if (choose < p_spec) {
        float a = m.roughness * m.roughness;
        float cos_t = sqrtf((1.0f - r2) / (1.0f + (a * a - 1.0f) * r2));
        float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
        vec3 hl = v3(sin_t * fm::cosf_(phi), sin_t * fm::sinf_(phi), cos_t);
        vec3 hw = local_to_world(hl, n);
        return normalize(reflect(-wi, hw));
    }
    float cos_t = sqrtf(r2);
    float sin_t = sqrtf(1.0f - r2);
    vec3 ll = v3(sin_t * fm::cosf_(phi), sin_t * fm::sinf_(phi), cos_t);
    return local_to_world(ll, n);
}

// Material::Evaluate / Pdf dispatch (reference: Src/Core/material.cu:14-44):
// dielectric is a delta lobe (0), everything else but PBR evaluates as diffuse.
RT_HD vec3 eval_bsdf(const DevMaterial& m, vec3 wi, vec3 n, vec3 wo) {
    if (m.type == kPbr) return eval_pbr(m, wi, n, wo);
    if (m.type == kDielectric) return v3(0.f, 0.f, 0.f);
    return eval_diffuse(m, wi, n, wo);
}
RT_HD float pdf_bsdf(const DevMaterial& m, vec3 wi, vec3 n, vec3 wo) {
    if (m.type == kPbr) return pdf_pbr(m, wi, n, wo);
    if (m.type == kDielectric) return 0.f;
    return pdf_diffuse(n, wo);
}

// ---------------------------------------------------------------- dielectric
// (reference: Src/Core/material.cu:121-191)
RT_HD float fresnel_dielectric(float cos_i, float eta_i, float eta_t) {
    cos_i = clampf(-1.0f, 1.0f, cos_i);
    bool entering = cos_i > 0.0f;
    if (!entering) {
        float t = eta_i;
        eta_i = eta_t;
        eta_t = t;
        cos_i = fabsf(cos_i);
    }
    float sin_i = sqrtf(fmaxf(0.0f, 1.0f - cos_i * cos_i));
    float sin_t = (eta_i / eta_t) * sin_i;
    if (sin_t >= 1.0f) return 1.0f;
    float cos_t = sqrtf(fmaxf(0.0f, 1.0f - sin_t * sin_t));
    float r_par = ((eta_t * cos_i) - (eta_i * cos_t)) / ((eta_t * cos_i) + (eta_i * cos_t));
    float r_perp = ((eta_i * cos_i) - (eta_t * cos_t)) / ((eta_i * cos_i) + (eta_t * cos_t));
    return 0.5f * (r_par * r_par + r_perp * r_perp);
}
RT_HD bool refract_dir(vec3 inc, vec3 n, float eta, vec3& out) {
    float cos_i = fmaxf(0.0f, -dot(inc, n));
    float sin2_t = (eta * eta) * fmaxf(0.0f, 1.0f - cos_i * cos_i);
    if (sin2_t >= 1.0f) return false;
    float cos_t = sqrtf(fmaxf(0.0f, 1.0f - sin2_t));
    out = eta * inc + (eta * cos_i - cos_t) * n;
    return true;
}
// One random number `u` chooses reflection with probability F.
RT_HD void sample_dielectric(const DevMaterial& m, vec3 wi, vec3 normal, float u, vec3& wo, float& pdf, vec3& weight) {
    vec3 inc = -normalize(wi);
    vec3 on = normalize(normal);
    float eta_i = 1.0f, eta_t = m.ior;
    if (dot(inc, on) > 0.0f) {
        on = -on;
        float t = eta_i;
        eta_i = eta_t;
        eta_t = t;
    }
    float cos_i = fmaxf(0.0f, -dot(inc, on));
    float f = fresnel_dielectric(cos_i, eta_i, eta_t);
    if (u < f) {
        wo = normalize(reflect(inc, on));
        pdf = fmaxf(f, 1e-6f);
        weight = v3(1.f, 1.f, 1.f);
        return;
    }
    vec3 r;
    float eta = eta_i / eta_t;
    if (!refract_dir(inc, on, eta, r)) {
        wo = normalize(reflect(inc, on));
        pdf = 1.0f;
        weight = v3(1.f, 1.f, 1.f);
        return;
    }
    wo = normalize(r);
    pdf = fmaxf(1.0f - f, 1e-6f);
    weight = lerp(v3(1.f, 1.f, 1.f), m.basecolor, m.transmission);
}

// ---------------------------------------------------------------- lights
// Solid-angle pdf of an area-light point seen from `p` (reference: Src/Core/light.cu:79-89).
RT_HD float area_light_pdf_sa(vec3 p, vec3 lp, vec3 ln, float area) {
    vec3 wi = normalize(p - lp);
    float d2 = norm2(p - lp);
    float c = fmaxf(dot(ln, wi), 1e-6f);
    return d2 / (area * c);
}

}  // namespace rt
