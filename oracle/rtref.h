/*
 * rtref -- CPU oracle for the wavefront path tracer.  TEST INFRASTRUCTURE
 * ONLY: it is loaded by tests/, by __graft_entry__.smoke() as the checker and
 * by bench.py's cpu_baseline leg, never by the library.
 *
 * A plain-C restatement of the reference's hot path, written in the
 * reference's own structure (recursive SAH build over 32-byte LinearBVHNode
 * records, node-by-node stack traversal, one sequential loop per path), with
 * every function citing the reference file:line it follows.  Built with
 * -ffp-contract=off so its FP32 results are comparable bit for bit with the
 * GPU kernels.  Randomness is the counter-based Philox4x32-10 stream
 * (cuRAND's XORWOW is NVIDIA's and is not reproduced).
 */
#ifndef RTREF_H
#define RTREF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Material record: 24 floats. */
#define REF_MAT_STRIDE 24
/* [0] type (0 light, 1 diffuse, 2 pbr, 3 dielectric, 4 subsurface)
   [1..3] basecolor  [4..6] emission  [7] roughness  [8] metallic
   [9] transmission  [10] ior  [11..13] scatter distance
   [14] diffuse texture index + 1 (0 = none)  [15] unused
   [16..18] sigma_a  [19..21] sigma_s  [22] subsurface method (0 Burley, the
   reference's; 1 random walk, the library's addition)  [23] unused */

typedef struct ref_scene_desc {
    int num_tris;
    const float* pos;      /* 9 per triangle, scene order */
    const float* nrm;      /* 9 per triangle */
    const float* uv;       /* 6 per triangle */
    const int32_t* mesh;   /* mesh per triangle */
    int num_meshes;
    const int32_t* mesh_material;
    const int32_t* mesh_light;
    int num_materials;
    const float* materials;
    int num_lights;
    const int32_t* light_type;  /* 0 area, 2 env */
    const int32_t* light_mesh;
    const float* light_emission; /* 3 per light */
    int env_w, env_h;
    const float* env_rgb;        /* env_w*env_h*3, row 0 = top */
    int width, height;
    float fovy;
    float direction[3];
    int num_textures;
    const int32_t* tex_info;  /* (first texel, width, height) per texture */
    const uint8_t* tex_rgba;  /* RGBA8 texels of all textures, row 0 = top */
} ref_scene_desc;

typedef struct ref_params {
    int max_bounces;
    float rr;
    int rr_start_depth;
    uint32_t seed;
    int flush_survivors;
    int env_alias; /* 1: environment sampled with alias tables instead of the CDF search */
    int shadow_mode; /* 0: the reference's shadow test; 2: the library's segment test (its fast
                        mode 2; mode 1 gives mode 0's results and is not restated) */
} ref_params;

typedef struct ref_scene ref_scene;

/* Orbit camera (camera.h:11-84): target, distance, polar angle from +y, azimuth (degrees). */
typedef struct ref_orbit {
    float target[3];
    float distance, theta_deg, phi_deg;
} ref_orbit;

/* G-buffer planes, width*height each, row 0 = bottom (Shader/gbuffer.fs:2-5). */
typedef struct ref_gbuf {
    float* pos_tri;       /* 4 floats: world position, triangle index within its mesh */
    uint16_t* nrm_mat;    /* 4: normal*0.5+0.5 in 1/65535 units, material id */
    uint16_t* bary_mesh;  /* 4: barycentrics in 1/65535 units, mesh id + 1 (0 = background) */
    float* motion_depth;  /* 4: NDC motion (current - previous), linear depth, fwidth(depth) */
} ref_gbuf;

/* DenoiseParam (Src/Utils/ui.h:14-25). */
typedef struct ref_svgf_params {
    int max_history;
    float sigma_light, sigma_normal, sigma_depth, reproj_depth;
    int atrous_iters;
} ref_svgf_params;

ref_scene* ref_scene_create(const ref_scene_desc* d);
void ref_scene_destroy(ref_scene* s);
/* New positions (9 per triangle, scene order) and optionally shading normals
   (NULL: kept) for the same triangles; node boxes refit over the same tree. */
void ref_scene_refit(ref_scene* s, const float* pos, const float* nrm);

int ref_bvh_num_nodes(const ref_scene* s);
int ref_bvh_max_depth(const ref_scene* s);
void ref_bvh_nodes(const ref_scene* s, int32_t* out8);   /* 8 words per node */
void ref_bvh_order(const ref_scene* s, int32_t* out);
void ref_camera(const ref_scene* s, float* out12);       /* position, left_down, du, dv */

/* Closest hit for rays (o.xyz, tmin, d.xyz, tmax): hits (t, b1, b2, ordered prim as int bits),
   counts (interior nodes whose children were tested, triangle tests). */
void ref_trace(const ref_scene* s, const float* rays8, int n, float* hits4, uint32_t* counts2);
/* The segment test of the library's shadow mode 2 for n rays (o.xyz, tmin, d.xyz, tmax): out[i] = 1 when a
   triangle met from its back lies strictly inside (tmin, tmax); by the BVH (counts2: node visits and
   triangle tests, may be NULL) or, brute = 1, by testing every triangle. */
void ref_occluded(const ref_scene* s, const float* rays8, int n, int brute, int32_t* out, uint32_t* counts2);

/* Accumulate samples [sample0, sample0+spp) for the listed pixels (all when pixels == NULL),
   in sample order; acc_* are npix*4 floats.  stats[0..2] += extension / shadow / subsurface
   rays (subsurface: the two Burley probe rays per event, or the random walk's segments).
   Uses OpenMP threads over pixels. */
void ref_render(const ref_scene* s, const ref_params* p, int sample0, int spp, const int32_t* pixels, int npix,
                float* acc_direct, float* acc_indirect, uint64_t* stats);

/* Same estimator in double precision (FP64 twin) for converged-image comparisons. */
void ref_render_f64(const ref_scene* s, const ref_params* p, int sample0, int spp, const int32_t* pixels, int npix,
                    double* acc_direct, double* acc_indirect);

/* Current orbit of the path-tracing camera; moving it changes the camera of ref_render. */
void ref_get_orbit(const ref_scene* s, ref_orbit* out);
void ref_set_orbit(ref_scene* s, const ref_orbit* o);
/* 4x4 row-major projection * view of the G-buffer pass for an orbit (mathutil.cu:163-194). */
void ref_view_projection(const ref_scene* s, const ref_orbit* o, float near_plane, float far_plane, float* out16);
/* Primary-visibility G-buffer of the current camera at pixel centres; motion against `prev`
   (renderer.cuh:1446-1503 with Shader/gbuffer.*).  Both faces count. */
void ref_gbuffer(const ref_scene* s, const ref_orbit* prev, const ref_gbuf* out);
/* SVGF passes over width*height pixels (renderer.cuh:442-954).  Colours are 4 floats per
   pixel (.w variance), moments 2 floats, history lengths int32. */
void ref_svgf_temporal(int w, int h, int first, const ref_svgf_params* p, float inv_spp, const float* acc_d,
                       const float* acc_i, const float* hist_prev_d, const float* hist_prev_i,
                       const float* mom_prev_d, const float* mom_prev_i, const int32_t* len_prev,
                       const ref_gbuf* cur, const ref_gbuf* prev, float* hist_d, float* hist_i, float* mom_d,
                       float* mom_i, int32_t* len);
void ref_svgf_variance(int w, int h, const ref_svgf_params* p, const float* in_d, const float* in_i,
                       const float* mom_d, const float* mom_i, const int32_t* len, const ref_gbuf* g,
                       float* out_d, float* out_i);
void ref_svgf_atrous(int w, int h, int step, const ref_svgf_params* p, const float* in_d, const float* in_i,
                     const int32_t* len, const ref_gbuf* g, float* out_d, float* out_i);

/* Scalar pieces exposed for known-answer and closed-form tests. */
void ref_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float ref_math(int op, float a, float b); /* 2 sin 3 cos 4 acos 5 exp 6 log 7 atan2 10 pow */
/* BSDF helpers: type from material; eval returns rgb into out3, pdf returned. */
float ref_bsdf_eval(const float* mat, const float* wi, const float* n, const float* wo, float* out3);
float ref_bsdf_sample(const float* mat, const float* wi, const float* n, const float* u3, float* wo_out);
float ref_fresnel_dielectric(float cos_i, float eta_i, float eta_t);
float ref_ggx_d(const float* mat, const float* wi, const float* n, const float* wo);
float ref_burley(const float* mat, float r, int ch);
/* SSS: exit distance from one uniform for scatter distance d; the two-sided probe
   from p (normal n) at (dist, angle) returning the exit point and its normal. */
float ref_sss_dist(float scatter_distance, float u);
/* Random-walk subsurface transport from surface point p (outward normal
   n_out) into mesh `mesh`, a random-walk medium, entry direction from the
   uniforms (e1, e2), step numbers from (seed, pixel, sample, bounce).  Returns
   1 and the exit point, outward normal and per-channel weight when the walk
   leaves the medium, 0 when it ends inside (or the mesh is no such medium);
   *steps receives the segments traced. */
int ref_sss_walk(const ref_scene* s, int mesh, const float* p, const float* n_out, float e1, float e2, uint32_t seed,
                 uint32_t pixel, uint32_t sample, uint32_t bounce, float* out_p, float* out_n, float* out_w,
                 int* steps);
void ref_sss_walk_batch(const ref_scene* s, int mesh, const float* p, const float* n_out, int n, const float* e,
                        uint32_t seed, int32_t* ok, float* out_p, float* out_n, float* out_w, int32_t* steps);
void ref_sss_probe(const ref_scene* s, const float* p, const float* n, float dist, float angle, float* out_p,
                   float* out_n);
void ref_env_tables(const ref_scene* s, float* row_cdf, float* marg_cdf);
/* Alias tables (threshold, alias index bits) per row entry (2*w*h floats) and per row (2*h). */
void ref_env_alias_tables(const ref_scene* s, float* row2, float* marg2);
void ref_env_sample_alias(const ref_scene* s, float rv, float ru, float* uv_out);
/* Walker/Vose table of n double weights. */
void ref_alias_build(const double* w, int n, float* prob, int32_t* alias);
void ref_env_sample(const ref_scene* s, float rv, float ru, float* uv_out);
float ref_env_pdf_uv(const ref_scene* s, float u, float v);
/* Texture::Sample (texture.cu:242-249): albedo of texture `tex` at (u, v). */
void ref_texture_albedo(const ref_scene* s, int tex, float u, float v, float* rgb);

#ifdef __cplusplus
}
#endif

#endif
