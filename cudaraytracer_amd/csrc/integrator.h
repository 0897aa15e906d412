// One shading event of a path (reference: Src/Core/renderer.cuh:170-422,
// PathTracing) split for a wavefront pipeline: everything except visibility is
// decided here; the shadow ray and its unoccluded contribution are emitted as
// a record for the shadow stage, which adds the contribution when the light
// point is visible (reference visibility rule: Src/Core/light.cu:5-27).
#pragma once

#include "device.h"
#include "envmap.h"
#include "rt_math.h"
#include "shade.h"
#include "texture.h"
#include "traverse.h"

namespace rt {

struct PathState {
    vec3 o;         // ray origin (previous hit point)
    vec3 d;         // ray direction
    vec3 w;         // throughput
    float pdf;      // pdf of the last BSDF sample (for light-hit MIS)
    uint32_t flags; // kFlagFirst | kFlagSpecular
};

struct ShadowRec {
    vec3 lp;        // light point offset along its normal
    float sdist;    // |hp - lp|
    vec3 dir;       // normalised (hp - lp)
    vec3 hp;        // shading point offset along its geometric normal
    vec3 contrib;   // added to the path's direct (first bounce) or indirect sum when visible
};

struct ShadeOut {
    bool alive;
    bool shadow;
    int sss_rays;       // rays the subsurface event traced (Burley probes or walk segments)
    bool add_direct;    // contribution target: direct (first bounce) or indirect
    vec3 emit;          // radiance added now (light hit / environment)
    bool has_emit;
    ShadowRec rec;
};

// Light sample: point, normal and emitted radiance (Light::SamplePoint and
// Light::Emission, reference: Src/Core/light.cu:29-69, 91-107, 141-155;
// Triangle::Sample, Src/Core/triangle.cu:135-143).
RT_HD bool sample_light(const DevScene& sc, Rng& rng, int& light_id, float& p_sel, vec3& lp, vec3& ln, bool& is_env) {
    if (sc.num_lights == 0) return false;
    int id = (int)(rng.next() * (float)sc.num_lights);
    if (id == sc.num_lights) id = sc.num_lights - 1;
    p_sel = 1.0f / (float)sc.num_lights;
    light_id = id;
    Global<LightRec>& L = gref(sc.lights + id);
    is_env = L.type == kEnvLight;
    if (is_env) {
        float rv = rng.next();
        float ru = rng.next();
        env_sample_point(sc, rv, ru, lp, ln);
        return true;
    }
    if (L.mesh < 0 || L.tri_count <= 0) return false;
    int t = (int)(rng.next() * (float)L.tri_count);
    if (t == L.tri_count) t = L.tri_count - 1;
    float u = rng.next();
    float v = rng.next();
    if (u + v > 1) {
        u = 1 - u;
        v = 1 - v;
    }
    const float4* tp = sc.light_tris + 4 * (size_t)(L.tri_start + t);
    float4 a = ldg4(tp), b = ldg4(tp + 1), c = ldg4(tp + 2), n = ldg4(tp + 3);
    float s0 = (1 - u) - v;
    lp = (s0 * v3(a.x, a.y, a.z) + u * v3(b.x, b.y, b.z)) + v * v3(c.x, c.y, c.z);
    ln = v3(n.x, n.y, n.z);
    return true;
}

RT_HD vec3 light_emission(const DevScene& sc, int light_id, vec3 wi) {
    Global<LightRec>& L = gref(sc.lights + light_id);
    if (L.type == kEnvLight) return env_emission(sc, wi);
    return v3(L.emission[0], L.emission[1], L.emission[2]);
}

RT_HD float light_pdf_sa(const DevScene& sc, int light_id, vec3 origin, vec3 lp, vec3 ln) {
    Global<LightRec>& L = gref(sc.lights + light_id);
    if (L.type == kEnvLight) return env_pdf_from(sc, origin, lp);
    return area_light_pdf_sa(origin, lp, ln, L.area);
}

// Shadow-ray geometry of Light::Visible: both end points pushed 1e-4 along
// their normals, ray cast from the light towards the shading point.
RT_HD void make_shadow(vec3 hit_p, vec3 hit_ng, vec3 lp, vec3 ln, ShadowRec& r) {
    vec3 hp = hit_p + 1e-4f * hit_ng;
    vec3 l = lp + 1e-4f * ln;
    r.hp = hp;
    r.lp = l;
    r.sdist = length(hp - l);
    r.dir = normalize(hp - l);
}

// Visibility decision given the closest hit of the shadow ray (found or not).
RT_HD bool shadow_visible(const ShadowRec& r, bool hit, vec3 test_p) {
    if (!hit) return true;
    float test_dist = length(test_p - r.lp);
    // (double)x > 1e-1 holds exactly when x >= 0.1f for every float x.
    if (length(r.hp - test_p) >= 0.1f && fabsf(test_dist - r.sdist) >= 0.1f) return false;
    return true;
}

// Burley diffusion profile R(r) for one channel (reference: Src/Core/material.cu:458-469).
RT_HD float burley(const DevMaterial& m, float r, int ch) {
    float d = ch == 0 ? m.scatter.x : (ch == 1 ? m.scatter.y : m.scatter.z);
    if (r < 1e-6f || d < 1e-6f) return 0.f;
    float e1 = fm::expf_(-r / d);
    float e2 = fm::expf_(-r / (3.f * d));
    return (float)((double)(e1 + e2) / (((8.0 * kPiD) * (double)d) * (double)r));
}
RT_HD float burley_pdf(const DevMaterial& m, float r) {
    float p0 = burley(m, r, 0), p1 = burley(m, r, 1), p2 = burley(m, r, 2);
    return ((p0 + p1) + p2) / 3.f;
}

// Subsurface probe: two closest-hit rays along +-n through the disc point
// (reference: Src/Core/material.cu:405-441).
__device__ __forceinline__ void sss_probe(const DevScene& sc, vec3 hit_p, vec3 normal, float dist, float angle,
                                         vec3& out_p, vec3& out_n, uint32_t* stk, int stride) {
    vec3 t, b;
    vec3 pn = normalize(normal);
    create_onb(pn, t, b);
    vec3 off = dist * (fm::cosf_(angle) * t + fm::sinf_(angle) * b);
    out_p = hit_p + off;
    float ray_off = fmaxf(dist, 1e-3f);
    float tmin = 1e-4f;
    float tmax = fmaxf(dist * 2.f, 1e-2f);
    RayInv fw = make_ray(out_p + pn * ray_off, -pn);
    RayInv bw = make_ray(out_p - pn * ray_off, pn);
    TravCount c{0, 0};
    Hit hf = trace_closest<false>(sc, fw, tmin, tmax, stk, stride, c);
    Hit hb = trace_closest<false>(sc, bw, tmin, tmax, stk, stride, c);
    bool f = hf.prim >= 0, bk = hb.prim >= 0;
    if (f && bk) {
        const Hit& h = hf.t <= hb.t ? hf : hb;
        SurfaceHit s = surface_at(sc.tri_attr, h.prim, h.b1, h.b2);
        out_p = s.p;
        out_n = s.ng;
    } else if (f) {
        SurfaceHit s = surface_at(sc.tri_attr, hf.prim, hf.b1, hf.b2);
        out_p = s.p;
        out_n = s.ng;
    } else if (bk) {
        SurfaceHit s = surface_at(sc.tri_attr, hb.prim, hb.b1, hb.b2);
        out_p = s.p;
        out_n = s.ng;
    } else {
        out_p = hit_p;
        out_n = pn;
    }
}

// Random-walk subsurface transport, the "random_walk" subsurface method (the
// reference has only the Burley profile; this follows the walk of Chiang,
// Kutz and Burley, "Practical and controllable subsurface scattering for
// production path tracing", SIGGRAPH 2016 talk, with isotropic scattering).
// The path enters the medium in a cosine-distributed direction about the
// inward normal (diffuse transmission) and walks.  Each step picks one colour
// channel uniformly, draws a free-flight distance s = -ln(u) / sigma_t of that
// channel and looks for the medium's own boundary along the segment to s: a
// closest-hit query over the medium's tree (its mesh's triangles only), both
// faces counting (the walk sees the boundary from inside); the pdf is the
// mean over the channels (one-sample MIS over wavelengths).  A surface closer
// than s ends the walk there with weight T / mean(T), T = exp(-sigma_t t) per
// channel; otherwise it scatters at s with weight sigma_s T / mean(sigma_t T),
// turns to a uniformly distributed direction and plays Russian roulette on its
// largest channel.  Absorbed and over-long walks end the path.  Step k draws
// from its own coordinate (bounce << 8 | k, kStreamWalk): channel, distance,
// then two direction numbers and the roulette number when it scatters.
constexpr int kWalkMaxSteps = 256;

struct WalkExit {
    vec3 p;     // exit point on the surface
    vec3 n;     // unit geometric normal there, facing out of the medium
    vec3 w;     // the walk's weight per channel
    bool ok;
    int steps;  // segments traced
};

// Clearance of point p in its medium: the grid cell's lower bound of the
// distance to the mesh, 0 outside the grid.
__device__ __forceinline__ float medium_clearance(const DevScene& sc, int mesh, vec3 p) {
    const float4 g0 = ldg4(sc.media_grid + 2 * mesh), g1 = ldg4(sc.media_grid + 2 * mesh + 1);
    const int n = f_as_i(g0.w);
    const int x = (int)floorf((p.x - g0.x) * g1.x), y = (int)floorf((p.y - g0.y) * g1.y),
              z = (int)floorf((p.z - g0.z) * g1.z);
    if (x < 0 || y < 0 || z < 0 || x >= n || y >= n || z >= n) return 0.f;
    return gref(sc.media_clear + f_as_i(g1.w) + ((size_t)z * n + y) * n + x);
}

__device__ __forceinline__ WalkExit sss_random_walk(const DevScene& sc, const DevMedium& md, int mesh, uint32_t root,
                                                   vec3 hit_p, vec3 n_out, float e1, float e2, uint32_t seed,
                                                   uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t* stk,
                                                   int stride) {
    WalkExit ex;
    ex.ok = false;
    ex.p = hit_p;
    ex.n = n_out;
    ex.w = v3(0.f, 0.f, 0.f);
    ex.steps = 0;
    const vec3 st = md.sigma_a + md.sigma_s;
    vec3 thr = v3(1.f, 1.f, 1.f);
    vec3 d = sample_diffuse_dir(-n_out, e1, e2);
    vec3 p = hit_p - n_out * 1e-4f;
    for (int k = 0; k < kWalkMaxSteps; ++k) {
        Rng g;
        g.init(seed, pixel, sample, (bounce << 8) | (uint32_t)k, kStreamWalk);
        int ch = clampi(0, 2, (int)(g.next() * 3.f));
        float stc = ch == 0 ? st.x : (ch == 1 ? st.y : st.z);
        float s = -fm::logf_(g.next()) / stc;
        // Only a surface nearer than the sampled distance matters, so the
        // segment query stops at s and skips every box entered beyond it; a
        // flight shorter than the point's clearance needs no query at all
        // (no triangle lies within reach: the query would find none).
        Hit h;
        h.prim = -1;
        if (!(s < medium_clearance(sc, mesh, p))) {
            RayInv r = make_ray(p, d);
            TravCount c{0, 0};
            h = trace_closest_tree<false, false>(sc.media_nodes, sc.media_isect, root, r, 0.f, fminf(s, 1e8f), stk,
                                                 stride, c);
        }
        ex.steps = k + 1;
        if (h.prim >= 0) {
            vec3 T = v3(fm::expf_(-st.x * h.t), fm::expf_(-st.y * h.t), fm::expf_(-st.z * h.t));
            float pdf = ((T.x + T.y) + T.z) / 3.f;
            if (!(pdf > 0.f)) return ex;
            SurfaceHit q = surface_at(sc.tri_attr, gref(sc.media_prim + h.prim), h.b1, h.b2);
            vec3 ng = normalize(q.ng);
            ex.p = q.p;
            ex.n = dot(ng, d) < 0.f ? -ng : ng;
            ex.w = cmul(thr, T) / pdf;
            ex.ok = true;
            return ex;
        }
        vec3 T = v3(fm::expf_(-st.x * s), fm::expf_(-st.y * s), fm::expf_(-st.z * s));
        float pdf = ((st.x * T.x + st.y * T.y) + st.z * T.z) / 3.f;
        if (!(pdf > 0.f)) return ex;
        thr = cmul(thr, cmul(md.sigma_s, T)) / pdf;
        p = p + d * s;
        float z = 1.f - 2.f * g.next();
        float phi = (float)(kTwoPiD * (double)g.next());
        float sr = sqrtf(fmaxf(0.f, 1.f - z * z));
        d = v3(sr * fm::cosf_(phi), sr * fm::sinf_(phi), z);
        float q = fmaxf(thr.x, fmaxf(thr.y, thr.z));
        if (q < 1.f) {
            if (g.next() >= q) return ex;
            thr = thr / q;
        }
    }
    return ex;
}

// One bounce of one path.  `hit` holds (t, b1, b2, prim) from the extend
// stage; CLS is the shading class of what was hit, fixed at compile time so
// each class kernel carries only its own material code.
template <int CLS>
__device__ __forceinline__ ShadeOut shade_path(const DevScene& sc, const FrameParams& fp, PathState& ps,
                                              float4 hit, uint32_t pixel, uint32_t sample, uint32_t* stk,
                                              int stride) {
    ShadeOut out;
    out.alive = false;
    out.shadow = false;
    out.has_emit = false;
    out.sss_rays = 0;
    out.add_direct = (ps.flags & kFlagFirst) != 0;
    int prim = f_as_i(hit.w);
    const bool first = (ps.flags & kFlagFirst) != 0;
    const bool specular = (ps.flags & kFlagSpecular) != 0;

    // Misses never get here: the extend stage ends them with the environment's
    // radiance (reference: Src/Core/renderer.cuh:142-159).

    // The light sample of next-event estimation takes the first random numbers
    // of the bounce and does not depend on the surface, so it is drawn before
    // the surface is fetched and its table reads overlap that fetch (the
    // random-number sequence is the reference's either way).
    Rng rng;
    rng.init(fp.seed, pixel, sample, (uint32_t)fp.bounce, kStreamPath);
    int lid = 0;
    float p_sel = 0.f;
    vec3 lp = v3(0.f, 0.f, 0.f), ln = v3(0.f, 0.f, 0.f);
    bool is_env = false, have_light = false;
    if (CLS == kClsDiffuse || CLS == kClsPbr) have_light = sample_light(sc, rng, lid, p_sel, lp, ln, is_env);

    SurfaceHit sh = surface_at(sc.tri_attr, prim, hit.y, hit.z);
    DevMaterial mat = load_material(sc.materials, sh.material);
    // A diffuse texture replaces the base colour of every lobe that reads it
    // (all but the dielectric's, material.cu:190, and lights).
    if (CLS != kClsTerminal && CLS != kClsDielectric && mat.texture >= 0)
        mat.basecolor = texture_albedo(sc, mat.texture, sh.u, sh.v);
    vec3 hit_p = sh.p;

    if (CLS == kClsTerminal) {
        // Light hit (renderer.cuh:192-216): the path ends.
        if (first) {
            out.emit = mat.ke;
            out.has_emit = true;
        } else if (specular) {
            out.emit = cmul(ps.w, mat.ke);
            out.has_emit = true;
        } else {
            // MIS against the BSDF pdf of the previous bounce; the light-selection
            // probability is not applied here (renderer.cuh:205-208).
            float mis = 1.f;
            if (sh.light >= 0 && sh.light < sc.num_lights) {
                float pdf_light = light_pdf_sa(sc, sh.light, ps.o, hit_p, sh.ng);
                mis = mis_weight(ps.pdf, pdf_light);
            }
            out.emit = cmul(ps.w, mat.ke) * mis;
            out.has_emit = true;
        }
        return out;
    }

    vec3 geom_n = normalize(sh.ng);
    vec3 hit_n = (double)norm2(sh.ns) > kEpsD ? normalize(sh.ns) : geom_n;
    if (dot(hit_n, geom_n) < 0.f) hit_n = -hit_n;
    vec3 wo = -ps.d;

    if (CLS == kClsDielectric) {
        // Delta lobe (renderer.cuh:218-240).
        vec3 wi, weight;
        float pdf = 1.0f;
        float u = rng.next();
        sample_dielectric(mat, wo, hit_n, u, wi, pdf, weight);
        ps.pdf = pdf;
        ps.o = hit_p + wi * 1e-4f;
        ps.d = wi;
        ps.w = cmul(ps.w, weight);
        if (fp.bounce >= fp.rr_start) {
            float p = rng.next();
            if (p > fp.rr) return out;
            ps.w = ps.w / fp.rr;
        }
        ps.flags = kFlagSpecular;
        out.alive = true;
        return out;
    }

    if (CLS == kClsDiffuse || CLS == kClsPbr) {
        // Next-event estimation (renderer.cuh:242-273), with the light sample drawn above.
        if (have_light) {
            vec3 dir_wi = normalize(lp - hit_p);
            float pdf_light = light_pdf_sa(sc, lid, hit_p, lp, ln) * p_sel;
            float cos_t = dot(hit_n, dir_wi);
            vec3 f = CLS == kClsPbr ? eval_pbr(mat, wo, hit_n, dir_wi) : eval_diffuse(mat, wo, hit_n, dir_wi);
            vec3 le = light_emission(sc, lid, dir_wi);
            vec3 dir_l = (cmul(le, f) * cos_t) / pdf_light;
            float pdf_b = CLS == kClsPbr ? pdf_pbr(mat, wo, hit_n, dir_wi) : pdf_diffuse(hit_n, dir_wi);
            float mis = mis_weight(pdf_light, pdf_b);
            vec3 c = mis * cmul(ps.w, dir_l);
            // A zero contribution leaves the path sums unchanged, so no ray is needed.
            if (c.x != 0.f || c.y != 0.f || c.z != 0.f) {
                make_shadow(hit_p, sh.ng, lp, ln, out.rec);
                out.rec.contrib = c;
                out.shadow = true;
            }
        }
        // BSDF sample (renderer.cuh:274-289).
        vec3 wi, f;
        float pdf;
        if (CLS == kClsPbr) {
            float r1 = rng.next();
            float r2 = rng.next();
            float ch = rng.next();
            wi = sample_pbr_dir(mat, wo, hit_n, r1, r2, ch);
            pdf = pdf_pbr(mat, wo, hit_n, wi);
            f = eval_pbr(mat, wo, hit_n, wi);
        } else {
            float t1 = rng.next();
            float t2 = rng.next();
            wi = sample_diffuse_dir(hit_n, t1, t2);
            pdf = pdf_diffuse(hit_n, wi);
            f = eval_diffuse(mat, wo, hit_n, wi);
        }
        ps.pdf = pdf;
        ps.o = hit_p;
        ps.d = wi;
        float cos_t = dot(hit_n, wi);
        if (!((double)pdf < kEpsD || cos_t < 0)) {
            ps.w = (cmul(ps.w, f) * cos_t) / pdf;
        } else {
            return out;
        }
    }

    if (CLS == kClsSubsurface) {
        // Fresnel-chosen specular reflection or diffusion-profile transmission
        // with an exit-point NEE (renderer.cuh:290-404).
        float fin = fresnel_dielectric(dot(hit_n, wo), 1.0f, mat.ior);
        vec3 f_in = v3(fin, fin, fin);
        float reflect_prob = clampf(0.02f, 0.98f, fin);
        float transmit_prob = fmaxf(1.f - reflect_prob, 1e-6f);
        float ev = rng.next();
        if (ev < reflect_prob) {
            vec3 rd;
            if (mat.roughness > 1e-3f) {
                float r1 = rng.next();
                float r2 = rng.next();
                float a = fmaxf(mat.roughness * mat.roughness, 0.02f);
                float phi = (float)(kTwoPiD * (double)r1);
                float cos_t = sqrtf((1.0f - r2) / (1.0f + (a * a - 1.0f) * r2));
                float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
                vec3 hl = v3(sin_t * fm::cosf_(phi), sin_t * fm::sinf_(phi), cos_t);
                vec3 hw = local_to_world(hl, hit_n);
                rd = normalize(reflect(-wo, hw));
                if (dot(rd, hit_n) <= 0.0f) rd = normalize(reflect(-wo, hit_n));
            } else {
                rd = normalize(reflect(-wo, hit_n));
            }
            ps.o = hit_p + rd * 1e-4f;
            ps.d = rd;
            ps.w = cmul(ps.w, f_in) / reflect_prob;
            ps.pdf = 1.0f;
            if (fp.bounce >= fp.rr_start) {
                float p = rng.next();
                if (p > fp.rr) return out;
                ps.w = ps.w / fp.rr;
            }
            ps.flags = kFlagSpecular;
            out.alive = true;
            return out;
        }
        vec3 out_p, out_n;
        const DevMedium md = load_medium(sc.materials, sh.material);
        if (md.method == kSssRandomWalk) {
            // Random walk through the medium (sss_random_walk above), in the tree of the hit's mesh.
            vec3 n_out = dot(geom_n, wo) < 0.f ? -geom_n : geom_n;
            float e1 = rng.next();
            float e2 = rng.next();
            const uint32_t root = gref(sc.media_root + sh.mesh);
            WalkExit ex = sss_random_walk(sc, md, sh.mesh, root, hit_p, n_out, e1, e2, fp.seed, pixel, sample,
                                          (uint32_t)fp.bounce, stk, stride);
            out.sss_rays = ex.steps;
            if (!ex.ok) return out;
            out_p = ex.p;
            out_n = ex.n;
            ps.w = cmul(cmul(ps.w, v3(1.f, 1.f, 1.f) - f_in), ex.w) / transmit_prob;
        } else {
            // Burley radius (material.cu:436-456), angle, probe (material.cu:415-424).
            float dist = 0.f;
            int ch = (int)(rng.next() * 3.f);
            ch = clampi(0, 2, ch);
            float dch = ch == 0 ? mat.scatter.x : (ch == 1 ? mat.scatter.y : mat.scatter.z);
            if (!(dch < 1e-6f)) {
                float r1 = rng.next();
                if (r1 < 0.5f) {
                    r1 = r1 * 2.f;
                    dist = -dch * fm::logf_(1.f - r1);
                } else {
                    r1 = (r1 - 0.5f) * 2.f;
                    dist = (-3.f * dch) * fm::logf_(1.f - r1);
                }
            }
            float angle = (float)((double)(rng.next() * 2.f) * kPiD);
            sss_probe(sc, hit_p, geom_n, dist, angle, out_p, out_n, stk, stride);
            out.sss_rays = 2;
            float sd = length(out_p - hit_p);
            float spatial_pdf = sd < 1e-6f ? 0.f : burley_pdf(mat, sd);
            if ((double)spatial_pdf < kEpsD) return out;
            vec3 profile = v3(0.f, 0.f, 0.f);
            if (!(sd < 1e-6f))
                profile = cmul(mat.basecolor, v3(burley(mat, sd, 0), burley(mat, sd, 1), burley(mat, sd, 2))) / kPiF;
            ps.w = cmul(cmul(ps.w, v3(1.f, 1.f, 1.f) - f_in), profile) / (transmit_prob * spatial_pdf);
        }

        int lid;
        float p_sel;
        vec3 lp, ln;
        bool is_env;
        if (sample_light(sc, rng, lid, p_sel, lp, ln, is_env)) {
            vec3 dir_wi = normalize(lp - out_p);
            float fnee = fresnel_dielectric(dot(out_n, dir_wi), 1.0f, mat.ior);
            float ft_out = 1 - fnee;
            float pdf_light = light_pdf_sa(sc, lid, out_p, lp, ln) * p_sel;
            float cos_t = dot(out_n, dir_wi);
            if (cos_t > 0.f && (double)pdf_light > kEpsD) {
                vec3 brdf = v3(1.f, 1.f, 1.f) / kPiF;
                vec3 le = light_emission(sc, lid, dir_wi);
                vec3 dir_l = (cmul(le, brdf) * cos_t) / pdf_light;
                float pdf_b = (float)((double)cos_t / kPiD);
                float mis = mis_weight(pdf_light, pdf_b);
                vec3 c = (mis * cmul(ps.w, dir_l)) * ft_out;
                if (c.x != 0.f || c.y != 0.f || c.z != 0.f) {
                    make_shadow(out_p, out_n, lp, ln, out.rec);
                    out.rec.contrib = c;
                    out.shadow = true;
                }
            }
        }
        float t1 = rng.next();
        float t2 = rng.next();
        vec3 od = sample_diffuse_dir(out_n, t1, t2);
        float od_pdf = pdf_diffuse(out_n, od);
        float cos_out = dot(out_n, od);
        float fout = fresnel_dielectric(dot(out_n, od), 1.0f, mat.ior);
        float ft = 1 - fout;
        if ((double)od_pdf < kEpsD || cos_out < 0.f) return out;
        ps.o = out_p + out_n * 1e-4f;
        ps.d = od;
        ps.w = ps.w * ft;
        ps.pdf = od_pdf;
    }

    // Russian roulette (renderer.cuh:407-421), optionally delayed.
    if (fp.bounce >= fp.rr_start) {
        float p = rng.next();
        if (p > fp.rr) return out;
        ps.w = ps.w / fp.rr;
    }
    ps.flags = 0u;
    out.alive = true;
    return out;
}

}  // namespace rt
