// Headless G-buffer and SVGF (temporal reprojection, variance estimate,
// edge-aware a-trous) for gfx950.  The filter semantics are the reference's
// (Src/Core/renderer.cuh:442-954, parameters ui.h:14-25); its OpenGL G-buffer
// pass (renderer.cuh:1446-1503, Shader/gbuffer.fs) is replaced by a
// primary-visibility ray cast that writes the same four channels.
#pragma once

#include "device.h"
#include "rt_math.h"
#include "traverse.h"

namespace rt {

// DenoiseParam defaults (reference: Src/Utils/ui.h:19-24).
struct SvgfParams {
    int max_history = 24;
    float sigma_light = 8.f;
    float sigma_normal = 256.f;
    float sigma_depth = 1.f;
    float reproj_depth = 0.1f;
    int atrous_iters = 5;
};

// One G-buffer: the reference's four render targets, row 0 = bottom.
struct GBufferView {
    float4* pos_tri;      // world position, triangle index within its mesh
    ushort4* nrm_mat;     // normal * 0.5 + 0.5 in 16-bit units, material id
    ushort4* bary_mesh;   // barycentrics in 16-bit units, mesh id + 1 (0 = background)
    float4* motion_depth; // NDC motion (current - previous), linear depth, fwidth(depth)
};

struct GBufferParams {
    vec3 cam_pos, left_down, du, dv;
    float vp[16], prev_vp[16];
    float near_plane, far_plane;
    int width, height;
};

RT_HD float4 mat_mul_point(const float* m, vec3 p) {
    float4 r;
    r.x = ((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3] * 1.f;
    r.y = ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7] * 1.f;
    r.z = ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11] * 1.f;
    r.w = ((m[12] * p.x + m[13] * p.y) + m[14] * p.z) + m[15] * 1.f;
    return r;
}

// Linear depth from the clip-space point, as the G-buffer fragment shader
// derives it from gl_FragCoord.z (gbuffer.fs:30-33).
RT_HD float linear_depth(const GBufferParams& g, float4 clip) {
    float zndc = clip.z / clip.w;
    float n = g.near_plane, f = g.far_plane;
    return ((2.f * n) * f) / ((f + n) - zndc * (f - n));
}

RT_HD vec3 pixel_center_dir(const GBufferParams& g, int i, int j) {
    vec3 pp = (g.left_down + ((float)i + 0.5f) * g.du) + ((float)j + 0.5f) * g.dv;
    return normalize(pp - g.cam_pos);
}

RT_HD unsigned short enc16(float v) { return (unsigned short)(unsigned int)(v * 65535.0f); }

RT_HD vec3 decode_normal(ushort4 q) {
    return normalize(v3(((float)q.x / 65535.0f) * 2.f - 1.f, ((float)q.y / 65535.0f) * 2.f - 1.f,
                        ((float)q.z / 65535.0f) * 2.f - 1.f));
}

// Primary-visibility G-buffer at pixel centres.  fwidth(depth) is the sum of
// the absolute depth differences to the horizontal and vertical quad partner,
// measured on the plane of the same triangle (what a rasteriser's helper
// lanes extrapolate).
__device__ void gbuffer_pixel(const DevScene& sc, const GBufferParams& g, int i, int j, uint32_t* stk, int stride,
                              GBufferView out) {
    size_t idx = (size_t)j * g.width + i;
    vec3 d = pixel_center_dir(g, i, j);
    RayInv r = make_ray(g.cam_pos, d);
    TravCount c{0, 0};
    Hit h = trace_closest_wave<false, false>(sc, r, 0.f, 1e8f, stk, stride, c);
    if (h.prim < 0) {
        out.pos_tri[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
        out.nrm_mat[idx] = make_ushort4(0, 0, 0, 0);
        out.bary_mesh[idx] = make_ushort4(0, 0, 0, 0);
        out.motion_depth[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    SurfaceHit s = surface_at(sc.tri_attr, h.prim, h.b1, h.b2);
    const float4* a = sc.tri_attr + kAttrStride * (size_t)h.prim;
    float4 a0 = a[0];
    int tri_in_mesh = f_as_i(a0.w);
    out.pos_tri[idx] = make_float4(s.p.x, s.p.y, s.p.z, (float)tri_in_mesh);
    vec3 ne = (s.ns * 0.5f + v3(0.5f, 0.5f, 0.5f)) * 65535.0f;
    out.nrm_mat[idx] = make_ushort4((unsigned short)(unsigned int)ne.x, (unsigned short)(unsigned int)ne.y,
                                    (unsigned short)(unsigned int)ne.z, (unsigned short)s.material);
    float b0 = (1.f - h.b1) - h.b2;
    out.bary_mesh[idx] = make_ushort4(enc16(b0), enc16(h.b1), enc16(h.b2), (unsigned short)(s.mesh + 1));
    float4 cc = mat_mul_point(g.vp, s.p);
    float4 pc = mat_mul_point(g.prev_vp, s.p);
    float mvx = cc.x / cc.w - pc.x / pc.w;
    float mvy = cc.y / cc.w - pc.y / pc.w;
    float depth = linear_depth(g, cc);
    // Quad partners on the triangle's plane.
    vec3 v0 = v3(a0.x, a0.y, a0.z);
    float pn = dot(v0 - g.cam_pos, s.ng);
    float dz = 0.f;
    for (int axis = 0; axis < 2; ++axis) {
        int pi = axis == 0 ? (i ^ 1) : i, pj = axis == 0 ? j : (j ^ 1);
        vec3 dd = pixel_center_dir(g, pi, pj);
        float den = dot(dd, s.ng);
        if (den != 0.f) {
            vec3 q = g.cam_pos + dd * (pn / den);
            dz += fabsf(linear_depth(g, mat_mul_point(g.vp, q)) - depth);
        }
    }
    out.motion_depth[idx] = make_float4(mvx, mvy, depth, dz);
}

// ---------------------------------------------------------------- temporal
// The frame's radiance comes in as per-pixel sample sums (scaled by inv_spp);
// the blended colour goes to this frame's history buffer, which is also the
// variance pass's input (the reference blends in place and then copies the
// result to its history buffer: renderer.cuh:1212-1217).  A pixel whose
// history reaches kVarianceHistory frames keeps that colour through the
// variance pass (renderer.cuh:703, 794-803), so the temporal pass also writes it
// to the variance pass's output (var_d / var_i) and the variance pass visits
// only the shorter histories.
constexpr int kVarianceHistory = 4;

struct TemporalIO {
    const float4 *acc_d, *acc_i;        // this frame's radiance sums
    float inv_spp;
    float4 *cur_d, *cur_i;              // blended colour, .w = variance (this frame's history)
    float4 *var_d, *var_i;              // variance pass output, written for long histories
    const float4 *hist_d, *hist_i;      // previous frame's history
    const float2 *mom_prev_d, *mom_prev_i;
    const int* len_prev;
    float2 *mom_d, *mom_i;
    int* len;
    // Filter guide of this frame, one record per pixel, written here once so the
    // variance and a-trous taps need no normal decode: (decoded normal, linear
    // depth) and (mesh << 16 | material, history length).
    float4* guide_nd;
    int2* guide_kl;
    float4* lum;  // luminance guide of cur_d / cur_i (see lum_guide)
};

// Luminance guide of the temporal pass's colours, read by the variance pass:
// (lum_d, sqrt(max(lum_d, 1e-4)), lum_i, sqrt(max(lum_i, 1e-4))).  A 7x7 tap
// then reads its neighbour's luminance instead of recomputing it, and its
// sqrt(max(lum_c, lum_n, 1e-4)) becomes max(root_c, root_n): a correctly
// rounded square root is monotone, so the two are the same float (also when
// either luminance is NaN, which fmaxf drops in both forms).  The a-trous
// passes compute luminance themselves: one more 16-byte load per tap cost
// them more than the arithmetic it saves.
RT_HD float4 lum_guide(float4 d, float4 i) {
    const float ld = luminance(v3(d.x, d.y, d.z)), li = luminance(v3(i.x, i.y, i.z));
    return make_float4(ld, sqrtf(fmaxf(ld, 1e-4f)), li, sqrtf(fmaxf(li, 1e-4f)));
}

// sqrtf(n) for n = 0..18, the squared lengths of the 7x7 and 5x5 filter
// offsets: the correctly rounded float roots, so a tap reads its distance
// instead of taking a square root.
static __constant__ float kTapDist[19] = {
    0x0p+0f,       0x1p+0f,       0x1.6a09e6p+0f, 0x1.bb67aep+0f, 0x1p+1f,       0x1.1e377ap+1f, 0x1.3988e2p+1f,
    0x1.52a7fap+1f, 0x1.6a09e6p+1f, 0x1.8p+1f,     0x1.94c584p+1f, 0x1.a8872ap+1f, 0x1.bb67aep+1f, 0x1.cd82b4p+1f,
    0x1.deeea2p+1f, 0x1.efbdecp+1f, 0x1p+2f,       0x1.07e0f6p+2f, 0x1.0f876cp+2f};

RT_HD int2 guide_key_len(ushort4 nrm_mat, ushort4 bary_mesh, int len) {
    return make_int2((int)(((uint32_t)bary_mesh.w << 16) | (uint32_t)nrm_mat.w), len);
}

// IsValidTemporalReprojectionTap (reference: renderer.cuh:460-497).
RT_HD bool temporal_tap_ok(int tx, int ty, int w, int h, int mesh, int mat, int tri, vec3 n, float depth, float dz,
                           float pix_dist, const GBufferView& prev, float reproj) {
    if (tx < 0 || tx >= w || ty < 0 || ty >= h) return false;
    size_t t = (size_t)ty * w + tx;
    float4 lp = prev.pos_tri[t];
    ushort4 lb = prev.bary_mesh[t];
    ushort4 ln = prev.nrm_mat[t];
    float4 lm = prev.motion_depth[t];
    vec3 lnrm = decode_normal(ln);
    float dn = dot(n, lnrm);
    float dd = fabsf(depth - lm.z);
    float thr = fmaxf(dz, lm.w) * (pix_dist + 1.f) + reproj * fmaxf(depth, lm.z);
    return mesh == (int)lb.w && mat == (int)ln.w && tri == (int)lp.w && dn >= 0.9f && dd <= thr;
}

struct TapSum {
    vec3 d, i;
    float2 md, mi;
    float len;
};

RT_HD void temporal_add(TapSum& s, float wgt, size_t t, const TemporalIO& io) {
    float4 hd = io.hist_d[t], hi = io.hist_i[t];
    s.d = s.d + v3(hd.x, hd.y, hd.z) * wgt;
    s.i = s.i + v3(hi.x, hi.y, hi.z) * wgt;
    float2 a = io.mom_prev_d[t], b = io.mom_prev_i[t];
    s.md.x += a.x * wgt;
    s.md.y += a.y * wgt;
    s.mi.x += b.x * wgt;
    s.mi.y += b.y * wgt;
    s.len += (float)io.len_prev[t] * wgt;
}

RT_HD void temporal_scale(TapSum& s, float inv) {
    s.d = s.d * inv;
    s.i = s.i * inv;
    s.md.x *= inv;
    s.md.y *= inv;
    s.mi.x *= inv;
    s.mi.y *= inv;
    s.len *= inv;
}

// TemporalDenoising (reference: renderer.cuh:524-677).  Returns the pixel's
// history length, which decides whether the variance pass estimates or copies.
RT_HD int temporal_pixel(int i, int j, int w, int h, bool first, const SvgfParams& p, const TemporalIO& io,
                          const GBufferView& cur, const GBufferView& prev) {
    size_t idx = (size_t)j * w + i;
    float4 cd = io.acc_d[idx], ci = io.acc_i[idx];
    cd.x *= io.inv_spp; cd.y *= io.inv_spp; cd.z *= io.inv_spp;
    ci.x *= io.inv_spp; ci.y *= io.inv_spp; ci.z *= io.inv_spp;
    if (first) {
        float ld = luminance(v3(cd.x, cd.y, cd.z)), li = luminance(v3(ci.x, ci.y, ci.z));
        io.mom_d[idx] = make_float2(ld, ld * ld);
        io.mom_i[idx] = make_float2(li, li * li);
        cd.w = 1.f;
        ci.w = 1.f;
        io.cur_d[idx] = cd;
        io.cur_i[idx] = ci;
        io.lum[idx] = lum_guide(cd, ci);
        io.len[idx] = 1;
        {
            ushort4 gn0 = cur.nrm_mat[idx];
            vec3 n0 = decode_normal(gn0);
            io.guide_nd[idx] = make_float4(n0.x, n0.y, n0.z, cur.motion_depth[idx].z);
            io.guide_kl[idx] = guide_key_len(gn0, cur.bary_mesh[idx], 1);
        }
        return 1;
    }
    float cx = (float)i + 0.5f, cy = (float)j + 0.5f;
    float4 gp = cur.pos_tri[idx];
    ushort4 gb = cur.bary_mesh[idx];
    ushort4 gn = cur.nrm_mat[idx];
    float4 gm = cur.motion_depth[idx];
    int mat = gn.w, mesh = gb.w, tri = (int)gp.w;
    vec3 n = decode_normal(gn);
    float depth = gm.z, dz = gm.w;
    float mvx = (gm.x * (float)w) * 0.5f, mvy = (gm.y * (float)h) * 0.5f;
    float pix_dist = sqrtf(mvx * mvx + mvy * mvy);
    float px = cx - mvx, py = cy - mvy;
    bool valid = false;
    TapSum s;
    s.d = v3(0.f, 0.f, 0.f);
    s.i = v3(0.f, 0.f, 0.f);
    s.md = make_float2(0.f, 0.f);
    s.mi = make_float2(0.f, 0.f);
    s.len = 0.f;
    if (px >= 0.5f && px < (float)w - 0.5f && py >= 0.5f && py < (float)h - 0.5f) {
        float fx0 = px - 0.5f, fy0 = py - 0.5f;
        int bx = (int)floorf(fx0), by = (int)floorf(fy0);
        float fx = fx0 - (float)bx, fy = fy0 - (float)by;
        float total = 0.f;
        bool all4 = true;
        for (int oy = 0; oy <= 1; ++oy)
            for (int ox = 0; ox <= 1; ++ox) {
                int tx = bx + ox, ty = by + oy;
                if (!temporal_tap_ok(tx, ty, w, h, mesh, mat, tri, n, depth, dz, pix_dist, prev, p.reproj_depth)) {
                    all4 = false;
                    continue;
                }
                float wx = ox == 0 ? (1.f - fx) : fx;
                float wy = oy == 0 ? (1.f - fy) : fy;
                float wgt = wx * wy;
                temporal_add(s, wgt, (size_t)ty * w + tx, io);
                total += wgt;
            }
        if (all4 && total > 1e-6f) {
            temporal_scale(s, 1.f / total);
            valid = true;
        } else {
            // 3x3 fallback with weights (1/2, 1/4) around the reprojected pixel.
            s.d = v3(0.f, 0.f, 0.f);
            s.i = v3(0.f, 0.f, 0.f);
            s.md = make_float2(0.f, 0.f);
            s.mi = make_float2(0.f, 0.f);
            s.len = 0.f;
            total = 0.f;
            int cxi = clampi(0, w - 1, (int)floorf(px)), cyi = clampi(0, h - 1, (int)floorf(py));
            for (int oy = -1; oy <= 1; ++oy)
                for (int ox = -1; ox <= 1; ++ox) {
                    int tx = cxi + ox, ty = cyi + oy;
                    if (!temporal_tap_ok(tx, ty, w, h, mesh, mat, tri, n, depth, dz, pix_dist, prev, p.reproj_depth))
                        continue;
                    float wgt = (ox == 0 ? 0.5f : 0.25f) * (oy == 0 ? 0.5f : 0.25f);
                    temporal_add(s, wgt, (size_t)ty * w + tx, io);
                    total += wgt;
                }
            if (total > 1e-6f) {
                temporal_scale(s, 1.f / total);
                valid = true;
            }
        }
    }
    int hl = 1;
    if (valid) {
        int last = (int)floorf(s.len + 0.5f);
        if (last < 1) last = 1;
        hl = last + 1 < p.max_history ? last + 1 : p.max_history;
    }
    float alpha = valid ? 1.f / (float)hl : 1.0f;
    float ld = luminance(v3(cd.x, cd.y, cd.z)), li = luminance(v3(ci.x, ci.y, ci.z));
    cd.x = alpha * cd.x + (1 - alpha) * s.d.x;
    cd.y = alpha * cd.y + (1 - alpha) * s.d.y;
    cd.z = alpha * cd.z + (1 - alpha) * s.d.z;
    ci.x = alpha * ci.x + (1 - alpha) * s.i.x;
    ci.y = alpha * ci.y + (1 - alpha) * s.i.y;
    ci.z = alpha * ci.z + (1 - alpha) * s.i.z;
    float2 lmd = valid ? s.md : make_float2(0.f, 0.f);
    float2 lmi = valid ? s.mi : make_float2(0.f, 0.f);
    float2 md, mi;
    md.x = alpha * ld + (1 - alpha) * lmd.x;
    md.y = (alpha * ld) * ld + (1 - alpha) * lmd.y;
    mi.x = alpha * li + (1 - alpha) * lmi.x;
    mi.y = (alpha * li) * li + (1 - alpha) * lmi.y;
    io.mom_d[idx] = md;
    io.mom_i[idx] = mi;
    if (valid) {
        cd.w = md.y - md.x * md.x;
        ci.w = mi.y - mi.x * mi.x;
        io.len[idx] = hl;
    } else {
        cd.w = 1.f;
        ci.w = 1.f;
        io.len[idx] = 1;
    }
    io.cur_d[idx] = cd;
    io.cur_i[idx] = ci;
    io.lum[idx] = lum_guide(cd, ci);
    io.guide_nd[idx] = make_float4(n.x, n.y, n.z, depth);
    io.guide_kl[idx] = guide_key_len(gn, gb, valid ? hl : 1);
    if (valid && hl >= kVarianceHistory) {
        io.var_d[idx] = cd;
        io.var_i[idx] = ci;
    }
    return valid ? hl : 1;
}

// ---------------------------------------------------------------- variance
struct FilterIO {
    const float4 *in_d, *in_i;
    float4 *out_d, *out_i;
    const float2 *mom_d, *mom_i;
    const int* len;
    const float4* guide_nd;  // decoded normal, linear depth
    const int2* guide_kl;    // mesh << 16 | material, history length
    const float4* lum;       // variance pass: luminance guide of in_d / in_i
};

// ComputeVarianceKernel (reference: renderer.cuh:689-805): a 7x7 bilateral
// estimate while the history is shorter than kVarianceHistory frames, else a
// copy of the input (the temporal pass writes that copy).  This is the
// estimate for pixel (i, j).  Taps read the per-pixel guide (the same decoded
// normals and depths the reference's G-buffer fetches give); denominators that
// depend only on the centre are formed once.
RT_HD void variance_estimate(int i, int j, int w, int h, const SvgfParams& p, const FilterIO& io,
                             const GBufferView& g) {
    size_t idx = (size_t)j * w + i;
    int hl = io.len[idx];
    const float4 cl = io.lum[idx];
    const float cld = cl.x, cli = cl.z;
    const float4 cg = io.guide_nd[idx];
    const int ckey = io.guide_kl[idx].x;
    float cdepth = cg.w;
    float dz = fmaxf(g.motion_depth[idx].w, 1e-4f);
    vec3 cn = v3(cg.x, cg.y, cg.z);
    vec3 fd = v3(0.f, 0.f, 0.f), fi = v3(0.f, 0.f, 0.f);
    float fmd0 = 0.f, fmd1 = 0.f, fmi0 = 0.f, fmi1 = 0.f, twd = 0.f, twi = 0.f;
    float sigma_depth = p.sigma_depth * fmaxf(dz, 1e-4f);
    const float wind_den = p.sigma_light / sqrtf(cli);
    for (int kj = -3; kj <= 3; ++kj)
        for (int ki = -3; ki <= 3; ++ki) {
            int ni = i + ki, nj = j + kj;
            if (!(ni >= 0 && ni < w && nj >= 0 && nj < h)) continue;
            size_t n = (size_t)nj * w + ni;
            if (io.guide_kl[n].x != ckey) continue;
            float4 ng = io.guide_nd[n];
            float4 nd4 = io.in_d[n], ni4 = io.in_i[n];
            vec3 nd = v3(nd4.x, nd4.y, nd4.z), nic = v3(ni4.x, ni4.y, ni4.z);
            const float4 nl = io.lum[n];
            float nld = nl.x, nli = nl.z;
            float2 nmd = io.mom_d[n], nmi = io.mom_i[n];
            float dist = kTapDist[ki * ki + kj * kj];
            float wz = -fabsf(cdepth - ng.w) / (sigma_depth * dist + 1e-4f);
            float wn = fm::powf_(fmaxf(0.f, dot(cn, v3(ng.x, ng.y, ng.z))), p.sigma_normal);
            // sqrt(max(cld, nld, 1e-4)) as the larger of the two guide roots
            float wdir = -fabsf(cld - nld) / (p.sigma_light / fmaxf(cl.y, nl.y));
            float wind = -fabsf(cli - nli) / wind_den;
            float dw = fm::expf_(wz + wdir) * wn;
            float iw = fm::expf_(wz + wind) * wn;
            fd = fd + nd * dw;
            fi = fi + nic * iw;
            fmd0 += nmd.x * dw;
            fmd1 += nmd.y * dw;
            fmi0 += nmi.x * iw;
            fmi1 += nmi.y * iw;
            twd += dw;
            twi += iw;
        }
    twd = fmaxf(twd, 1e-6f);
    twi = fmaxf(twi, 1e-6f);
    fmd0 /= twd;
    fmd1 /= twd;
    fmi0 /= twi;
    fmi1 /= twi;
    fd = fd / twd;
    fi = fi / twi;
    float vd = fmaxf(0.f, fmd1 - fmd0 * fmd0);
    float vi = fmaxf(0.f, fmi1 - fmi0 * fmi0);
    vd *= 4.f / (float)hl;
    vi *= 4.f / (float)hl;
    io.out_d[idx] = make_float4(fd.x, fd.y, fd.z, vd);
    io.out_i[idx] = make_float4(fi.x, fi.y, fi.z, vi);
}

// ---------------------------------------------------------------- a-trous
// AtrousKernel (reference: renderer.cuh:807-954): 3x3-filtered variance,
// history bootstrap, 5x5 B3 taps at `step`, variance propagated with w^2.
// Taps read the per-pixel guide; centre-only denominators are formed once.
//
// The pixel's arithmetic is written once (atrous_core) over a tap source:
// tap (ki, kj) is the pixel at (i + ki * step, j + kj * step) and variance
// tap (vi, vj) the pixel at (i + vi, j + vj).  key() reports whether the tap
// lies inside the image and gives its guide key and history length; data()
// gives its guide normal/depth and the two input colours; var() gives the two
// input variances of a unit-offset neighbour.  AtrousGlobal reads global
// memory; the tiled kernel's source reads the LDS tile it staged.

struct AtrousGlobal {
    const FilterIO& io;
    int i, j, w, h, step;
    RT_HD bool key(int ki, int kj, int2& kl) const {
        const int ni = i + ki * step, nj = j + kj * step;
        if (!(ni >= 0 && ni < w && nj >= 0 && nj < h)) return false;
        kl = io.guide_kl[(size_t)nj * w + ni];
        return true;
    }
    RT_HD void data(int ki, int kj, float4& nd, float4& cd, float4& ci) const {
        const size_t n = (size_t)(j + kj * step) * w + (i + ki * step);
        nd = io.guide_nd[n];
        cd = io.in_d[n];
        ci = io.in_i[n];
    }
    RT_HD bool var(int vi, int vj, float& vd, float& vv) const {
        const int si = i + vi, sj = j + vj;
        if (si < 0 || si >= w || sj < 0 || sj >= h) return false;
        const size_t v = (size_t)sj * w + si;
        vd = io.in_d[v].w;
        vv = io.in_i[v].w;
        return true;
    }
};

template <class Src>
RT_HD void atrous_core(size_t idx, int step, const SvgfParams& p, const FilterIO& io, const GBufferView& g,
                       const Src& src) {
    int2 ckl;
    src.key(0, 0, ckl);
    int chl = ckl.y > 1 ? ckl.y : 1;
    float4 cg, cd4, ci4;
    src.data(0, 0, cg, cd4, ci4);
    vec3 cd = v3(cd4.x, cd4.y, cd4.z), ci = v3(ci4.x, ci4.y, ci4.z);
    float cld = luminance(cd), cli = luminance(ci);
    float cdepth = cg.w;
    float dz = fmaxf(g.motion_depth[idx].w, 1e-4f);
    vec3 cn = v3(cg.x, cg.y, cg.z);
    float fvd = 0.f, fvi = 0.f, fvw = 0.f;
    for (int vj = -1; vj <= 1; ++vj)
        for (int vi = -1; vi <= 1; ++vi) {
            float vd, vv;
            if (!src.var(vi, vj, vd, vv)) continue;
            float vw = (vi == 0 ? 0.5f : 0.25f) * (vj == 0 ? 0.5f : 0.25f);
            fvd += vw * fmaxf(0.0f, vd);
            fvi += vw * fmaxf(0.0f, vv);
            fvw += vw;
        }
    fvw = fmaxf(fvw, 1e-6f);
    fvd /= fvw;
    fvi /= fvw;
    float sdd = sqrtf(fvd) + 1e-3f, sdi = sqrtf(fvi) + 1e-3f;
    float boot = fminf(1.f, fmaxf(0.f, (4.f - (float)chl) / 3.f));
    float sigma_depth = (fmaxf(p.sigma_depth, 1e-3f) * fmaxf(dz, 1e-4f)) * (1.f + 2.5f * boot);
    float sld = p.sigma_light * (1.f + 2.0f * boot);
    float sli = p.sigma_light * (1.f + 1.5f * boot);
    float sn = fmaxf(8.f, p.sigma_normal * (1.f - 0.85f * boot));
    const float sld_sdd = sld * sdd;
    const float wind_den = (sli * sdi) / sqrtf(fmaxf(cli, 1e-4f));
    vec3 fd = v3(0.f, 0.f, 0.f), fi = v3(0.f, 0.f, 0.f);
    float twd = 0.f, twi = 0.f, vard = 0.f, vari = 0.f;
    for (int kj = -2; kj <= 2; ++kj)
        for (int ki = -2; ki <= 2; ++ki) {
            int2 nkl;
            if (!src.key(ki, kj, nkl)) continue;
            if (nkl.x != ckl.x) continue;
            int nhl = nkl.y > 1 ? nkl.y : 1;
            float4 ng, nd4, ni4;
            src.data(ki, kj, ng, nd4, ni4);
            vec3 nd = v3(nd4.x, nd4.y, nd4.z), nic = v3(ni4.x, ni4.y, ni4.z);
            float nvd = fmaxf(0.0f, nd4.w), nvi = fmaxf(0.0f, ni4.w);
            float dist = kTapDist[ki * ki + kj * kj] * (float)step;
            float wz = -fabsf(cdepth - ng.w) / (sigma_depth * dist + 1e-6f);
            float wn = fm::powf_(fmaxf(0.f, dot(cn, v3(ng.x, ng.y, ng.z))), sn);
            float ld = luminance(nd), li = luminance(nic);
            float dref = fmaxf(fmaxf(cld, ld), 1e-4f);
            float wdir = -fabsf(cld - ld) / (sld_sdd / sqrtf(dref));
            float wind = -fabsf(cli - li) / wind_den;
            float conf = boot > 0.f ? (0.5f + 0.5f * fminf((float)nhl / 4.f, 1.f)) : 1.f;
            float kw = (ki == 0 ? 0.375f : (ki == 1 || ki == -1 ? 0.25f : 0.0625f)) *
                       (kj == 0 ? 0.375f : (kj == 1 || kj == -1 ? 0.25f : 0.0625f));
            float dw = ((fm::expf_(wz + wdir) * wn) * kw) * conf;
            float iw = ((fm::expf_(wz + wind) * wn) * kw) * conf;
            fd = fd + dw * nd;
            fi = fi + iw * nic;
            twd += dw;
            twi += iw;
            vard += (dw * dw) * nvd;
            vari += (iw * iw) * nvi;
        }
    if (twd > 1e-6f) io.out_d[idx] = make_float4(fd.x / twd, fd.y / twd, fd.z / twd, vard / (twd * twd));
    else io.out_d[idx] = make_float4(cd.x, cd.y, cd.z, 1.f);
    if (twi > 1e-6f) io.out_i[idx] = make_float4(fi.x / twi, fi.y / twi, fi.z / twi, vari / (twi * twi));
    else io.out_i[idx] = make_float4(ci.x, ci.y, ci.z, 1.f);
}

RT_HD void atrous_pixel(int i, int j, int w, int h, int step, const SvgfParams& p, const FilterIO& io,
                        const GBufferView& g) {
    atrous_core((size_t)j * w + i, step, p, io, g, AtrousGlobal{io, i, j, w, h, step});
}

}  // namespace rt
