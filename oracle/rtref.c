/*
 * rtref -- CPU oracle (test infrastructure only; see rtref.h).
 *
 * Restates, in the reference's own structure:
 *   Src/Core/bounds.h:60-78        slab test with 1e-2 slack
 *   Src/Core/triangle.cu:3-10      triangle normal / area
 *   Src/Core/triangle.cu:102-133   Moller-Trumbore, back faces culled
 *   Src/Core/bvh.cu:44-206         binned SAH build + DFS flatten
 *   Src/Core/bvh.cu:253-306        stack traversal, nearest child first
 *   Src/Core/camera.h:58-84, scene.cpp:646-667   orbit pinhole camera
 *   Src/Core/renderer.cuh:96-440   ray generation, path loop, shading
 *   Src/Core/material.cu           diffuse / GGX / dielectric / Burley SSS
 *   Src/Core/light.cu, lightmanager.cu, texture.cu   lights and env map
 * Eigen evaluation order (Redux.h:99-113): a 3-term sum is a0 + (a1 + a2).
 */
#include "rtref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------ float math */
typedef struct { float x, y, z; } V3;

static V3 v3(float x, float y, float z) { V3 r = {x, y, z}; return r; }
static V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
static V3 vscale(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static V3 vdivs(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
static float vdot(V3 a, V3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
static V3 vcross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static float vlen(V3 a) { return sqrtf(vdot(a, a)); }
static V3 vnormalize(V3 a) {
    float z = vdot(a, a);
    if (z > 0.0f) return vdivs(a, sqrtf(z));
    return a;
}
static float vget(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

/* Cephes single-precision methods (S. L. Moshier): Cody-Waite reduction by
   pi/4 in three parts, minimax polynomials.  Restated here for the oracle. */
static float c_sinpoly(float r, float z) {
    return ((((-1.9515295891e-4f * z) + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z) * r + r;
}
static float c_cospoly(float z) {
    return (((((2.443315711809948e-5f * z) - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z) * z - 0.5f * z) + 1.0f;
}
static float c_reduce(float ax, int* j) {
    int q = (int)(ax * 1.27323954473516f);
    float y = (float)q;
    if (q & 1) { q += 1; y += 1.0f; }
    *j = q & 7;
    return ((ax - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
}
static float c_sin(float x) {
    float sgn = 1.0f, ax = x;
    int j;
    float r, z, y;
    if (x < 0.0f) { sgn = -1.0f; ax = -x; }
    r = c_reduce(ax, &j);
    if (j > 3) { sgn = -sgn; j -= 4; }
    z = r * r;
    y = (j == 1 || j == 2) ? c_cospoly(z) : c_sinpoly(r, z);
    return sgn < 0.0f ? -y : y;
}
static float c_cos(float x) {
    float ax = fabsf(x), sgn = 1.0f, r, z, y;
    int j;
    r = c_reduce(ax, &j);
    if (j > 3) { j -= 4; sgn = -sgn; }
    if (j > 1) sgn = -sgn;
    z = r * r;
    y = (j == 1 || j == 2) ? c_sinpoly(r, z) : c_cospoly(z);
    return sgn < 0.0f ? -y : y;
}
static float c_asin_small(float a) {
    float z = a * a;
    return (((((((4.2163199048e-2f * z) + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z + 1.6666752422e-1f) * z) * a) + a;
}
static float c_acos(float x) {
    if (x < -1.0f || x > 1.0f) return (x - x) / (x - x);
    if (x < -0.5f) return 3.14159265358979323846f - 2.0f * c_asin_small(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * c_asin_small(sqrtf(0.5f * (1.0f - x)));
    return 1.57079632679489661923f - c_asin_small(x);
}
static float c_pow2i(int n) {
    uint32_t bits = (uint32_t)(n + 127) << 23;
    float f;
    memcpy(&f, &bits, 4);
    return f;
}
static float c_exp(float x) {
    float z, r, r2, p, e;
    int n, n1, n2;
    if (x != x) return x;
    if (x > 88.72283905206835f) return HUGE_VALF;
    if (x < -103.278929903431851103f) return 0.0f;
    z = floorf(1.44269504088896341f * x + 0.5f);
    r = x - z * 0.693359375f;
    r = r - z * -2.12194440e-4f;
    n = (int)z;
    r2 = r * r;
    p = ((((((1.9875691500e-4f * r) + 1.3981999507e-3f) * r + 8.3334519073e-3f) * r + 4.1665795894e-2f) * r + 1.6666665459e-1f) * r + 5.0000001201e-1f);
    e = ((p * r2) + r) + 1.0f;
    n1 = n / 2;
    n2 = n - n1;
    return (e * c_pow2i(n1)) * c_pow2i(n2);
}
static float c_log(float x) {
    uint32_t bits;
    int e = 0;
    float m, z, y, fe, r;
    if (x != x) return x;
    if (x < 0.0f) return (x - x) / (x - x);
    if (x == 0.0f) return -HUGE_VALF;
    if (x == HUGE_VALF) return x;
    memcpy(&bits, &x, 4);
    if ((bits >> 23) == 0) {
        x = x * 33554432.0f;
        memcpy(&bits, &x, 4);
        e = -25;
    }
    e += (int)((bits >> 23) & 0xff) - 126;
    bits = (bits & 0x807fffffu) | 0x3f000000u;
    memcpy(&m, &bits, 4);
    if (m < 0.707106781186547524f) {
        e -= 1;
        m = (m + m) - 1.0f;
    } else {
        m = m - 1.0f;
    }
    z = m * m;
    y = (((((((((7.0376836292e-2f * m) - 1.1514610310e-1f) * m + 1.1676998740e-1f) * m - 1.2420140846e-1f) * m + 1.4249322787e-1f) * m - 1.6668057665e-1f) * m + 2.0000714765e-1f) * m - 2.4999993993e-1f) * m + 3.3333331174e-1f);
    y = (y * m) * z;
    fe = (float)e;
    y = y + -2.12194440e-4f * fe;
    y = y + -0.5f * z;
    r = m + y;
    r = r + 0.693359375f * fe;
    return r;
}
static float c_atan(float x) {
    float sgn = 1.0f, y, z;
    if (x < 0.0f) { sgn = -1.0f; x = -x; }
    if (x > 2.414213562373095f) {
        y = 1.57079632679489661923f;
        x = -(1.0f / x);
    } else if (x > 0.4142135623730950f) {
        y = 0.78539816339744830962f;
        x = (x - 1.0f) / (x + 1.0f);
    } else {
        y = 0.0f;
    }
    z = x * x;
    y = y + (((((((8.05374449538e-2f * z) - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z) * x) + x);
    return sgn < 0.0f ? -y : y;
}
static float c_atan2(float y, float x) {
    int code = 0;
    float w = 0.0f;
    if (x < 0.0f) code = 2;
    if (y < 0.0f) code |= 1;
    if (x == 0.0f) {
        if (code & 1) return -1.57079632679489661923f;
        if (y == 0.0f) return 0.0f;
        return 1.57079632679489661923f;
    }
    if (y == 0.0f) {
        if (code & 2) return 3.14159265358979323846f;
        return 0.0f;
    }
    if (code == 2) w = 3.14159265358979323846f;
    else if (code == 3) w = -3.14159265358979323846f;
    return w + c_atan(y / x);
}

/* x^y for x >= 0 as exp(y log x), 0^y = 0 for y > 0 (stands in for powf in the
   SVGF normal weight, renderer.cuh:755, 919). */
static float c_pow(float x, float y) {
    if (x == 0.0f) return y > 0.0f ? 0.0f : 1.0f;
    return c_exp(y * c_log(x));
}

float ref_math(int op, float a, float b) {
    switch (op) {
        case 10: return c_pow(a, b);
        case 2: return c_sin(a);
        case 3: return c_cos(a);
        case 4: return c_acos(a);
        case 5: return c_exp(a);
        case 6: return c_log(a);
        case 7: return c_atan2(a, b);
        default: return 0.0f;
    }
}

/* ------------------------------------------------------------ Philox4x32-10 */
void ref_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    int r;
    for (r = 0; r < 10; ++r) {
        uint64_t p0, p1;
        uint32_t n0, n2;
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        p0 = (uint64_t)0xD2511F53u * c0;
        p1 = (uint64_t)0xCD9E8D57u * c2;
        n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Draw d of coordinate (pixel, sample, bounce, stream): Philox block d/4,
   lane d%4, top 24 bits mapped to (0, 1] like curand_uniform's range. */
typedef struct {
    uint32_t seed, pixel, sample, bounce, stream;
    uint32_t blk[4];
    int dim;
} Rng;

static void rng_init(Rng* g, uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t stream) {
    g->seed = seed;
    g->pixel = pixel;
    g->sample = sample;
    g->bounce = bounce;
    g->stream = stream;
    g->dim = 0;
}
static float rng_next(Rng* g) {
    int lane = g->dim & 3;
    uint32_t x;
    if (lane == 0) {
        uint32_t ctr[4], key[2];
        ctr[0] = g->pixel;
        ctr[1] = g->sample;
        ctr[2] = (g->bounce << 8) | (uint32_t)(g->dim >> 2);
        ctr[3] = g->stream;
        key[0] = g->seed;
        key[1] = 0x5EEDC0DEu;
        ref_philox(ctr, key, g->blk);
    }
    x = g->blk[lane];
    g->dim++;
    return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f);
}

/* ------------------------------------------------------------ scene */
#define FLOATMAX_F ((float)3.40282e+038) /* global.h:8 */
#define PI_D 3.1415926                   /* global.h:7 */

typedef struct {
    float lo[3], hi[3];
} Box;

typedef struct {
    int32_t tri_index; /* -1 interior */
    int32_t second;    /* right child or triangle count */
    float lo[3], hi[3];
} LNode;

typedef struct {
    /* Triangle record (triangle.h:38-61): vertices, normals, uv, geometric normal, area, mesh. */
    V3 p[3];
    V3 n[3];
    float uv[3][2];
    V3 gn;
    float area;
    int mesh;
    Box b;
} Tri;

struct ref_scene {
    int nt;
    Tri* tris;       /* scene order */
    Tri* ordered;    /* BVH order */
    int32_t* order;
    LNode* nodes;
    int nn, cap, max_depth;
    int nm;
    int32_t* mesh_material;
    int32_t* mesh_light;
    float* mesh_area;
    int32_t* mesh_start;
    int32_t* mesh_count;
    int nmat;
    float* mats;
    int nl;
    int32_t* light_type;
    int32_t* light_mesh;
    float* light_em;
    int env_light;
    int env_w, env_h;
    float* env;       /* rgb */
    float* env_row;   /* row-conditional CDF */
    float* env_marg;  /* marginal CDF */
    float* env_alias_row;   /* (threshold, alias bits) per texel, row tables */
    float* env_alias_marg;  /* (threshold, alias bits) per row */
    Box bounds;
    int W, H;
    float fovy;
    V3 dir;
    V3 cam_pos, cam_ld, cam_du, cam_dv;
    V3 cam_lookat, cam_up;
    ref_orbit orbit;  /* target, distance, theta, phi (degrees) of the current camera */
    int ntex;
    int32_t* tex_info;  /* (first texel, width, height) per texture */
    uint8_t* tex_rgba;
    struct ref_scene** medium;  /* per mesh: the random-walk medium's own scene, or NULL */
};

static float smin(float a, float b) { return (b < a) ? b : a; } /* std::min */
static float smax(float a, float b) { return (a < b) ? b : a; } /* std::max */

static Box box_empty(void) {
    Box b;
    int k;
    for (k = 0; k < 3; ++k) { b.lo[k] = FLOATMAX_F; b.hi[k] = -FLOATMAX_F; }
    return b;
}
static Box box_union(Box a, Box o) { /* bounds.h:35-40 */
    int k;
    for (k = 0; k < 3; ++k) { a.lo[k] = smin(a.lo[k], o.lo[k]); a.hi[k] = smax(a.hi[k], o.hi[k]); }
    return a;
}
static Box box_point(Box a, const float* p) { /* bounds.h:28-33 */
    int k;
    for (k = 0; k < 3; ++k) { a.lo[k] = smin(a.lo[k], p[k]); a.hi[k] = smax(a.hi[k], p[k]); }
    return a;
}
static float box_area(Box b) { /* bounds.h:81-87 */
    float ex, ey, ez;
    if (b.lo[0] > b.hi[0] || b.lo[1] > b.hi[1] || b.lo[2] > b.hi[2]) return 0.f;
    ex = b.hi[0] - b.lo[0];
    ey = b.hi[1] - b.lo[1];
    ez = b.hi[2] - b.lo[2];
    return 2.f * ((ex * ey + ey * ez) + ez * ex);
}
static int box_max_extent(Box b) { /* bounds.h:104-109, first index of the maximum */
    float e0 = b.hi[0] - b.lo[0], e1 = b.hi[1] - b.lo[1], e2 = b.hi[2] - b.lo[2], m = e0;
    int a = 0;
    if (e1 > m) { m = e1; a = 1; }
    if (e2 > m) a = 2;
    return a;
}
static void tri_centroid(const Tri* t, float* c) {
    int k;
    for (k = 0; k < 3; ++k) c[k] = (t->b.lo[k] + t->b.hi[k]) * 0.5f;
}
static int bucket(Box cb, const float* c, int axis) { /* bvh.cu:109-114 */
    float o = c[axis] - cb.lo[axis];
    float ext = cb.hi[axis] - cb.lo[axis];
    int b;
    if (ext > 0) o /= ext;
    b = (int)(12.0f * o);
    if (b == 12) b = 11;
    return b;
}

static int push_node(ref_scene* s) {
    if (s->nn == s->cap) {
        s->cap = s->cap ? 2 * s->cap : 1024;
        s->nodes = (LNode*)realloc(s->nodes, (size_t)s->cap * sizeof(LNode));
    }
    s->nodes[s->nn].tri_index = -1;
    s->nodes[s->nn].second = -1;
    return s->nn++;
}

/* BVH::RecursiveBuild (bvh.cu:44-175) over the working array `w` (partitioned in place). */
static int build_rec(ref_scene* s, Tri* w, int start, int end, int depth, int* n_ordered) {
    int self = push_node(s), i, j, axis, best = -1, mid;
    Box bounds = box_empty(), cb = box_empty(), bb[12];
    int counts[12];
    float cost[11], min_cost = FLOATMAX_F, sa;
    if (depth > s->max_depth) s->max_depth = depth;
    for (i = start; i < end; ++i) bounds = box_union(bounds, w[i].b);
    memcpy(s->nodes[self].lo, bounds.lo, 12);
    memcpy(s->nodes[self].hi, bounds.hi, 12);
    if (end - start <= 4) {
        s->nodes[self].tri_index = *n_ordered;
        s->nodes[self].second = end - start;
        for (i = start; i < end; ++i) s->ordered[(*n_ordered)++] = w[i];
        return self;
    }
    for (i = start; i < end; ++i) {
        float c[3];
        tri_centroid(&w[i], c);
        cb = box_point(cb, c);
    }
    axis = box_max_extent(cb);
    for (i = 0; i < 12; ++i) { counts[i] = 0; bb[i] = box_empty(); }
    for (i = start; i < end; ++i) {
        float c[3];
        int b;
        tri_centroid(&w[i], c);
        b = bucket(cb, c, axis);
        counts[b]++;
        bb[b] = box_union(bb[b], w[i].b);
    }
    sa = box_area(bounds);
    for (i = 0; i < 11; ++i) { /* right side stops at bucket 10 (bvh.cu:128-131) */
        Box b0 = box_empty(), b1 = box_empty();
        int c0 = 0, c1 = 0;
        for (j = 0; j <= i; ++j) { b0 = box_union(bb[j], b0); c0 += counts[j]; }
        for (j = i + 1; j < 11; ++j) { b1 = box_union(bb[j], b1); c1 += counts[j]; }
        cost[i] = .125f + ((float)c0 * box_area(b0) + (float)c1 * box_area(b1)) / sa;
    }
    for (i = 1; i < 11; ++i) /* split 0 is never chosen (bvh.cu:139) */
        if (cost[i] < min_cost) { min_cost = cost[i]; best = i; }
    mid = -1;
    if (best >= 0) {
        /* libstdc++ std::partition for bidirectional iterators. */
        int first = start, last = end;
        for (;;) {
            float c[3];
            for (;;) {
                if (first == last) goto done;
                tri_centroid(&w[first], c);
                if (bucket(cb, c, axis) <= best) ++first;
                else break;
            }
            --last;
            for (;;) {
                if (first == last) goto done;
                tri_centroid(&w[last], c);
                if (!(bucket(cb, c, axis) <= best)) --last;
                else break;
            }
            {
                Tri t = w[first];
                w[first] = w[last];
                w[last] = t;
            }
            ++first;
        }
    done:
        mid = first;
    }
    if (mid <= start || mid >= end) mid = (start + end) / 2; /* reference would recurse forever */
    build_rec(s, w, start, mid, depth + 1, n_ordered);
    {
        int right = build_rec(s, w, mid, end, depth + 1, n_ordered);
        s->nodes[self].tri_index = -1;
        s->nodes[self].second = right;
    }
    return self;
}

/* ------------------------------------------------------------ traversal */
typedef struct {
    V3 o, d, inv;
    int neg[3];
} Ray;

static Ray make_ray(V3 o, V3 d) {
    Ray r;
    r.o = o;
    r.d = d;
    r.inv = v3(1.f / d.x, 1.f / d.y, 1.f / d.z); /* bvh.cu:256 */
    r.neg[0] = r.inv.x < 0;
    r.neg[1] = r.inv.y < 0;
    r.neg[2] = r.inv.z < 0;
    return r;
}

/* Bounds3D::IsIntersect with tMin/tMax (bounds.h:60-78).  minCoeff/maxCoeff
   are fminf/fmaxf on the device; the slack test is evaluated in double as in
   the reference. */
static int slab(const Ray* r, const LNode* n, float tmin, float tmax, float* t) {
    float en[3], ex[3], exit_, enter;
    int k;
    for (k = 0; k < 3; ++k) {
        float a = (n->lo[k] - vget(r->o, k)) * vget(r->inv, k);
        float b = (n->hi[k] - vget(r->o, k)) * vget(r->inv, k);
        if (r->neg[k]) { en[k] = b; ex[k] = a; }
        else { en[k] = a; ex[k] = b; }
    }
    exit_ = fminf(ex[0], fminf(ex[1], ex[2]));
    enter = fmaxf(en[0], fmaxf(en[1], en[2]));
    *t = enter;
    return (double)(exit_ - enter) + 1e-2 >= 0 && (double)exit_ + 1e-2 >= 0 && enter <= tmax && exit_ >= tmin;
}

typedef struct {
    float t, b1, b2;
    int prim; /* ordered index, -1 on miss */
} Hit;

/* Triangle::IsIntersect (triangle.cu:102-133): double det, 1.f/det rounded
   through double, own [0, 1e8] range. */
static int tri_isect(const Ray* r, const Tri* tr, float* hit_t, float* hb1, float* hb2, int cull) {
    V3 e1 = vsub(tr->p[1], tr->p[0]);
    V3 e2 = vsub(tr->p[2], tr->p[0]);
    V3 s = vsub(r->o, tr->p[0]);
    V3 s1 = vcross(r->d, e2);
    V3 s2 = vcross(s, e1);
    double det = vdot(e1, s1);
    float inv, t, b1, b2;
    if (det == 0 || (cull && det < 0)) return 0; /* cull = 0: both faces, as the rasterised G-buffer */
    inv = (float)(1.0 / det);
    t = inv * vdot(s2, e2);
    if (t >= *hit_t) return 0;
    b1 = inv * vdot(s1, s);
    if (b1 < 0 || b1 > 1) return 0;
    b2 = inv * vdot(s2, r->d);
    if (b2 < 0 || b1 + b2 > 1) return 0;
    if (t < 0.f || t > 1e8f) return 0;
    *hit_t = t;
    *hb1 = b1;
    *hb2 = b2;
    return 1;
}

/* BVH::IsIntersect (bvh.cu:253-306). */
static Hit bvh_trace_c(const ref_scene* s, const Ray* r, float tmin, float tmax, uint32_t* nv, uint32_t* nt,
                       int cull) {
    Hit h;
    int stack[64], sp = 0, cur = 0;
    h.t = tmax;
    h.b1 = h.b2 = 0.f;
    h.prim = -1;
    for (;;) {
        const LNode* n = &s->nodes[cur];
        if (n->tri_index != -1) {
            int k;
            for (k = 0; k < n->second; ++k) {
                if (nt) (*nt)++;
                if (tri_isect(r, &s->ordered[n->tri_index + k], &h.t, &h.b1, &h.b2, cull)) h.prim = n->tri_index + k;
            }
        } else {
            int first = n->second, second = cur + 1;
            float tf, ts;
            int hf, hs;
            if (nv) (*nv)++;
            hf = slab(r, &s->nodes[first], tmin, h.t, &tf);
            hs = slab(r, &s->nodes[second], tmin, h.t, &ts);
            if (hf && hs) {
                if (tf < ts) { stack[sp++] = second; stack[sp++] = first; }
                else { stack[sp++] = first; stack[sp++] = second; }
            } else if (hf) {
                stack[sp++] = first;
            } else if (hs) {
                stack[sp++] = second;
            }
        }
        if (sp == 0) break;
        cur = stack[--sp];
    }
    return h;
}

static Hit bvh_trace(const ref_scene* s, const Ray* r, float tmin, float tmax, uint32_t* nv, uint32_t* nt) {
    return bvh_trace_c(s, r, tmin, tmax, nv, nt, 1);
}

/* The library's shadow fast mode (rt_render_params.shadow_any_hit = 2; no
   reference counterpart): the segment from the shading point towards the
   light point is occluded by any triangle met from its back -- the faces the
   reference's light-to-point ray meets from the front -- strictly inside
   (tmin, tmax).  FP32 in the library's operation order, so that decisions
   at the boundaries are the library's too. */
static int seg_isect(const Ray* r, const Tri* tr, float tmin, float tmax) {
    V3 e1 = vsub(tr->p[1], tr->p[0]), e2 = vsub(tr->p[2], tr->p[0]);
    V3 s1 = vcross(r->d, e2), so, s2;
    float det = vdot(e1, s1), inv, b1, b2, t;
    if (!(det < 0.f)) return 0;
    inv = 1.f / det;
    so = vsub(r->o, tr->p[0]);
    b1 = inv * vdot(s1, so);
    s2 = vcross(so, e1);
    b2 = inv * vdot(s2, r->d);
    t = inv * vdot(s2, e2);
    return b1 >= 0.f && b2 >= 0.f && b1 + b2 <= 1.f && t > tmin && t < tmax;
}

/* The segment test over the BVH: nodes as bvh_trace_c visits them (the nearer
   child first), stopping at the first occluding triangle. */
static int bvh_occluded(const ref_scene* s, const Ray* r, float tmin, float tmax, uint32_t* nv, uint32_t* nt) {
    int stack[64], sp = 0, cur = 0;
    for (;;) {
        const LNode* n = &s->nodes[cur];
        if (n->tri_index != -1) {
            int k;
            for (k = 0; k < n->second; ++k) {
                if (nt) (*nt)++;
                if (seg_isect(r, &s->ordered[n->tri_index + k], tmin, tmax)) return 1;
            }
        } else {
            int first = n->second, second = cur + 1;
            float tf, ts;
            int hf, hs;
            if (nv) (*nv)++;
            hf = slab(r, &s->nodes[first], tmin, tmax, &tf);
            hs = slab(r, &s->nodes[second], tmin, tmax, &ts);
            if (hf && hs) {
                if (tf < ts) { stack[sp++] = second; stack[sp++] = first; }
                else { stack[sp++] = first; stack[sp++] = second; }
            } else if (hf) {
                stack[sp++] = first;
            } else if (hs) {
                stack[sp++] = second;
            }
        }
        if (sp == 0) break;
        cur = stack[--sp];
    }
    return 0;
}

void ref_occluded(const ref_scene* s, const float* rays8, int n, int brute, int32_t* out, uint32_t* counts2) {
    int i;
#pragma omp parallel for schedule(dynamic, 256)
    for (i = 0; i < n; ++i) {
        const float* q = rays8 + 8 * (size_t)i;
        Ray r = make_ray(v3(q[0], q[1], q[2]), v3(q[4], q[5], q[6]));
        uint32_t nv = 0, nt = 0;
        int hit = 0, k;
        if (brute) {
            for (k = 0; k < s->nt && !hit; ++k) hit = seg_isect(&r, &s->ordered[k], q[3], q[7]);
        } else {
            hit = bvh_occluded(s, &r, q[3], q[7], &nv, &nt);
        }
        out[i] = hit;
        if (counts2) {
            counts2[2 * (size_t)i] = nv;
            counts2[2 * (size_t)i + 1] = nt;
        }
    }
}

void ref_trace(const ref_scene* s, const float* rays8, int n, float* hits4, uint32_t* counts2) {
    int i;
#pragma omp parallel for schedule(dynamic, 256)
    for (i = 0; i < n; ++i) {
        const float* q = rays8 + 8 * (size_t)i;
        Ray r = make_ray(v3(q[0], q[1], q[2]), v3(q[4], q[5], q[6]));
        uint32_t nv = 0, nt = 0;
        Hit h = bvh_trace(s, &r, q[3], q[7], &nv, &nt);
        hits4[4 * (size_t)i] = h.t;
        hits4[4 * (size_t)i + 1] = h.b1;
        hits4[4 * (size_t)i + 2] = h.b2;
        memcpy(&hits4[4 * (size_t)i + 3], &h.prim, 4);
        if (counts2) {
            counts2[2 * (size_t)i] = nv;
            counts2[2 * (size_t)i + 1] = nt;
        }
    }
}

/* ------------------------------------------------------------ scene set-up */
static float deg2rad(float a) { return (float)((double)(a / 180.f) * PI_D); }  /* mathutil.cu:4-7 */
static float rad2deg(float r) { return (float)((double)(r * 180.f) / PI_D); }  /* mathutil.cu:9-12 */

typedef struct {
    V3 pos, lookat, up, ld, du, dv;
} CamBasis;

/* Camera::ComputeCameraParam (camera.h:58-84): orbit -> pinhole basis. */
static CamBasis orbit_basis(const ref_scene* s, const ref_orbit* o) {
    CamBasis c;
    V3 target = v3(o->target[0], o->target[1], o->target[2]);
    float dist = o->distance, theta = deg2rad(o->theta_deg), phi = deg2rad(o->phi_deg);
    float x = (dist * sinf(theta)) * sinf(phi), y = dist * cosf(theta), z = (dist * sinf(theta)) * cosf(phi);
    V3 w, u, v;
    float focal, ratio, h_, w_;
    c.pos = vadd(target, v3(x, y, z));
    c.lookat = target;
    c.up = v3(0.f, 1.f, 0.f);
    w = vneg(vnormalize(vsub(c.lookat, c.pos)));
    u = vnormalize(vcross(vnormalize(c.up), w));
    v = vnormalize(vcross(w, u));
    focal = vlen(vsub(c.lookat, c.pos));
    ratio = (1.f * (float)s->W) / (float)s->H;
    h_ = (2.f * focal) * tanf(deg2rad(s->fovy / 2.f));
    w_ = h_ * ratio;
    c.du = vdivs(vscale(u, w_), (float)s->W);
    c.dv = vdivs(vscale(v, h_), (float)s->H);
    c.ld = vsub(vsub(c.lookat, vscale(c.du, (float)(s->W / 2))), vscale(c.dv, (float)(s->H / 2)));
    return c;
}

static void apply_orbit(ref_scene* s, const ref_orbit* o) {
    CamBasis c = orbit_basis(s, o);
    s->orbit = *o;
    s->cam_pos = c.pos;
    s->cam_ld = c.ld;
    s->cam_du = c.du;
    s->cam_dv = c.dv;
    s->cam_lookat = c.lookat;
    s->cam_up = c.up;
}

static void setup_camera(ref_scene* s) { /* scene.cpp:646-667 */
    V3 lo = v3(s->bounds.lo[0], s->bounds.lo[1], s->bounds.lo[2]);
    V3 hi = v3(s->bounds.hi[0], s->bounds.hi[1], s->bounds.hi[2]);
    V3 target = vscale(vadd(lo, hi), 0.5f);
    float diag = vlen(vsub(hi, lo));
    V3 dir = vnormalize(s->dir);
    ref_orbit o;
    o.target[0] = target.x;
    o.target[1] = target.y;
    o.target[2] = target.z;
    o.distance = ((diag * 0.5f) / sinf(deg2rad(s->fovy) / 2.f)) * 0.6f;
    o.theta_deg = rad2deg(acosf(dir.y / vlen(dir)));
    o.phi_deg = rad2deg(atan2f(dir.x, dir.z));
    apply_orbit(s, &o);
}

/* Walker's alias method with Vose's two worklists: scaled weights below 1
   are paired, last in first out, with one at or above 1, which donates the
   remainder and returns to the list its new weight falls in; leftovers get
   threshold 1. */
void ref_alias_build(const double* w, int n, float* prob, int32_t* alias) {
    double total = 0.0, *q;
    int *small, *large, ns = 0, nl = 0, i;
    for (i = 0; i < n; ++i) total += w[i];
    if (!(total > 0.0)) {
        for (i = 0; i < n; ++i) {
            prob[i] = 1.f;
            alias[i] = i;
        }
        return;
    }
    q = (double*)malloc(sizeof(double) * (size_t)n);
    small = (int*)malloc(sizeof(int) * (size_t)n);
    large = (int*)malloc(sizeof(int) * (size_t)n);
    for (i = 0; i < n; ++i) {
        q[i] = (w[i] * (double)n) / total;
        if (q[i] < 1.0) small[ns++] = i;
        else large[nl++] = i;
    }
    while (ns > 0 && nl > 0) {
        int a = small[--ns], b = large[--nl];
        prob[a] = (float)q[a];
        alias[a] = b;
        q[b] = (q[b] + q[a]) - 1.0;
        if (q[b] < 1.0) small[ns++] = b;
        else large[nl++] = b;
    }
    while (nl > 0) {
        int b = large[--nl];
        prob[b] = 1.f;
        alias[b] = b;
    }
    while (ns > 0) {
        int a = small[--ns];
        prob[a] = 1.f;
        alias[a] = a;
    }
    free(q);
    free(small);
    free(large);
}

static void setup_env_alias(ref_scene* s, const float* row_totals) {
    int w = s->env_w, h = s->env_h, x, y, n = w > h ? w : h;
    double* wt = (double*)malloc(sizeof(double) * (size_t)n);
    float* pr = (float*)malloc(sizeof(float) * (size_t)n);
    int32_t* al = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    s->env_alias_row = (float*)calloc(2 * (size_t)w * h, sizeof(float));
    s->env_alias_marg = (float*)calloc(2 * (size_t)h, sizeof(float));
    for (y = 0; y < h; ++y) {
        float theta = (float)((double)(((float)y + 0.5f) / (float)h) * PI_D);
        float st = sinf(theta);
        for (x = 0; x < w; ++x) {
            const float* c = &s->env[3 * ((size_t)y * w + x)];
            float lum = (0.2126f * c[0] + 0.7152f * c[1]) + 0.0722f * c[2];
            wt[x] = (double)(lum * st);
        }
        ref_alias_build(wt, w, pr, al);
        for (x = 0; x < w; ++x) {
            s->env_alias_row[2 * ((size_t)y * w + x)] = pr[x];
            memcpy(&s->env_alias_row[2 * ((size_t)y * w + x) + 1], &al[x], 4);
        }
    }
    for (y = 0; y < h; ++y) wt[y] = (double)row_totals[y];
    ref_alias_build(wt, h, pr, al);
    for (y = 0; y < h; ++y) {
        s->env_alias_marg[2 * (size_t)y] = pr[y];
        memcpy(&s->env_alias_marg[2 * (size_t)y + 1], &al[y], 4);
    }
    free(wt);
    free(pr);
    free(al);
}

static void setup_env(ref_scene* s) { /* texture.cu:38-80 */
    int w = s->env_w, h = s->env_h, x, y;
    s->env_row = (float*)calloc((size_t)w * h, sizeof(float));
    s->env_marg = (float*)calloc((size_t)h, sizeof(float));
    for (y = 0; y < h; ++y) {
        float vsum = 0.0f;
        float theta = (float)((double)(((float)y + 0.5f) / (float)h) * PI_D);
        float st = sinf(theta);
        for (x = 0; x < w; ++x) {
            size_t idx = (size_t)y * w + x;
            const float* c = &s->env[3 * idx];
            float lum = (0.2126f * c[0] + 0.7152f * c[1]) + 0.0722f * c[2];
            lum *= st;
            vsum += lum;
            s->env_row[idx] = vsum;
        }
        s->env_marg[y] = vsum;
    }
    for (y = 0; y < h; ++y) {
        float rs = s->env_marg[y];
        for (x = 0; x < w; ++x) {
            size_t idx = (size_t)y * w + x;
            s->env_row[idx] = rs > 0.0f ? s->env_row[idx] / rs : 0.0f;
        }
    }
    setup_env_alias(s, s->env_marg); /* row totals, before the prefix sum */
    for (y = 1; y < h; ++y) s->env_marg[y] += s->env_marg[y - 1];
    {
        float total = h > 0 ? s->env_marg[h - 1] : 0.f;
        for (y = 0; y < h; ++y) s->env_marg[y] = total > 0.0f ? s->env_marg[y] / total : 0.0f;
    }
}

static void build_media(ref_scene* s);

ref_scene* ref_scene_create(const ref_scene_desc* d) {
    ref_scene* s = (ref_scene*)calloc(1, sizeof(ref_scene));
    int i, k, n_ordered = 0;
    Tri* work;
    s->nt = d->num_tris;
    s->tris = (Tri*)calloc((size_t)s->nt, sizeof(Tri));
    s->ordered = (Tri*)calloc((size_t)s->nt, sizeof(Tri));
    s->order = (int32_t*)calloc((size_t)s->nt, sizeof(int32_t));
    for (i = 0; i < s->nt; ++i) {
        Tri* t = &s->tris[i];
        V3 nn;
        for (k = 0; k < 3; ++k) {
            const float* p = d->pos + 9 * (size_t)i + 3 * k;
            const float* q = d->nrm + 9 * (size_t)i + 3 * k;
            t->p[k] = v3(p[0], p[1], p[2]);
            t->n[k] = v3(q[0], q[1], q[2]);
            t->uv[k][0] = d->uv[6 * (size_t)i + 2 * k];
            t->uv[k][1] = d->uv[6 * (size_t)i + 2 * k + 1];
        }
        /* Triangle constructor (triangle.cu:3-10). */
        nn = vcross(vsub(t->p[1], t->p[0]), vsub(t->p[2], t->p[0]));
        t->area = 0.5f * vlen(nn);
        t->gn = vnormalize(nn);
        t->mesh = i; /* original index for the permutation; mesh id kept separately */
        for (k = 0; k < 3; ++k) {
            t->b.lo[k] = smin(vget(t->p[0], k), smin(vget(t->p[1], k), vget(t->p[2], k)));
            t->b.hi[k] = smax(vget(t->p[0], k), smax(vget(t->p[1], k), vget(t->p[2], k)));
        }
    }
    s->nm = d->num_meshes;
    s->mesh_material = (int32_t*)malloc(sizeof(int32_t) * (size_t)(s->nm + 1));
    s->mesh_light = (int32_t*)malloc(sizeof(int32_t) * (size_t)(s->nm + 1));
    s->mesh_area = (float*)calloc((size_t)(s->nm + 1), sizeof(float));
    s->mesh_start = (int32_t*)calloc((size_t)(s->nm + 1), sizeof(int32_t));
    s->mesh_count = (int32_t*)calloc((size_t)(s->nm + 1), sizeof(int32_t));
    for (i = 0; i < s->nm; ++i) {
        s->mesh_material[i] = d->mesh_material[i];
        s->mesh_light[i] = d->mesh_light[i];
        s->mesh_start[i] = -1;
    }
    for (i = 0; i < s->nt; ++i) {
        int m = d->mesh[i];
        if (s->mesh_start[m] < 0) s->mesh_start[m] = i;
        s->mesh_count[m]++;
        s->mesh_area[m] += s->tris[i].area; /* scene.cpp:515-519 */
    }
    s->nmat = d->num_materials;
    s->mats = (float*)malloc(sizeof(float) * REF_MAT_STRIDE * (size_t)(s->nmat + 1));
    memcpy(s->mats, d->materials, sizeof(float) * REF_MAT_STRIDE * (size_t)s->nmat);
    s->ntex = d->num_textures > 0 && d->tex_info && d->tex_rgba ? d->num_textures : 0;
    if (s->ntex) {
        size_t texels = 0;
        s->tex_info = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)s->ntex);
        memcpy(s->tex_info, d->tex_info, sizeof(int32_t) * 3 * (size_t)s->ntex);
        for (i = 0; i < s->ntex; ++i) {
            size_t end = (size_t)s->tex_info[3 * i] + (size_t)s->tex_info[3 * i + 1] * (size_t)s->tex_info[3 * i + 2];
            if (end > texels) texels = end;
        }
        s->tex_rgba = (uint8_t*)malloc(4 * texels + 4);
        memcpy(s->tex_rgba, d->tex_rgba, 4 * texels);
    }
    s->nl = d->num_lights;
    s->light_type = (int32_t*)malloc(sizeof(int32_t) * (size_t)(s->nl + 1));
    s->light_mesh = (int32_t*)malloc(sizeof(int32_t) * (size_t)(s->nl + 1));
    s->light_em = (float*)malloc(sizeof(float) * 3 * (size_t)(s->nl + 1));
    s->env_light = -1;
    for (i = 0; i < s->nl; ++i) {
        s->light_type[i] = d->light_type[i];
        s->light_mesh[i] = d->light_mesh[i];
        memcpy(&s->light_em[3 * i], &d->light_emission[3 * i], 12);
        if (s->light_type[i] == 2) s->env_light = i;
    }
    /* Build on a working copy; the mesh field carries the original index. */
    work = (Tri*)malloc(sizeof(Tri) * (size_t)s->nt);
    memcpy(work, s->tris, sizeof(Tri) * (size_t)s->nt);
    if (s->nt > 0) build_rec(s, work, 0, s->nt, 0, &n_ordered);
    free(work);
    for (i = 0; i < s->nt; ++i) {
        s->order[i] = s->ordered[i].mesh;
        s->ordered[i].mesh = d->mesh[s->order[i]];
    }
    for (i = 0; i < s->nt; ++i) s->tris[i].mesh = d->mesh[i];
    if (s->nn > 0) {
        memcpy(s->bounds.lo, s->nodes[0].lo, 12);
        memcpy(s->bounds.hi, s->nodes[0].hi, 12);
    }
    s->W = d->width;
    s->H = d->height;
    s->fovy = d->fovy;
    s->dir = v3(d->direction[0], d->direction[1], d->direction[2]);
    setup_camera(s);
    s->env_w = d->env_w;
    s->env_h = d->env_h;
    if (s->env_light >= 0 && d->env_rgb && s->env_w > 0 && s->env_h > 0) {
        s->env = (float*)malloc(sizeof(float) * 3 * (size_t)s->env_w * s->env_h);
        memcpy(s->env, d->env_rgb, sizeof(float) * 3 * (size_t)s->env_w * s->env_h);
        setup_env(s);
    } else {
        s->env_light = -1;
    }
    build_media(s);
    return s;
}

/* Random-walk media (the library's subsurface mode, no reference
   counterpart): for each mesh of a random-walk subsurface material, a scene
   of that mesh's own triangles in scene order, whose BVH the walk's segment
   queries use.  It carries no materials, lights, textures or environment. */
static int is_walk_medium(const ref_scene* s, int m) {
    int mat = s->mesh_material[m];
    const float* r;
    if (mat < 0 || mat >= s->nmat || s->mesh_count[m] <= 0) return 0;
    r = s->mats + REF_MAT_STRIDE * (size_t)mat;
    return (int)r[0] == 4 && (int)r[22] == 1;
}

static void medium_arrays(const ref_scene* s, int m, float* pos, float* nrm, float* uv) {
    int i, k, j = 0;
    for (i = 0; i < s->nt; ++i) {
        const Tri* t = &s->tris[i];
        if (t->mesh != m) continue;
        for (k = 0; k < 3; ++k) {
            pos[9 * (size_t)j + 3 * k] = t->p[k].x;
            pos[9 * (size_t)j + 3 * k + 1] = t->p[k].y;
            pos[9 * (size_t)j + 3 * k + 2] = t->p[k].z;
            nrm[9 * (size_t)j + 3 * k] = t->n[k].x;
            nrm[9 * (size_t)j + 3 * k + 1] = t->n[k].y;
            nrm[9 * (size_t)j + 3 * k + 2] = t->n[k].z;
            if (uv) {
                uv[6 * (size_t)j + 2 * k] = t->uv[k][0];
                uv[6 * (size_t)j + 2 * k + 1] = t->uv[k][1];
            }
        }
        ++j;
    }
}

static void build_media(ref_scene* s) {
    int m;
    s->medium = (ref_scene**)calloc((size_t)(s->nm + 1), sizeof(ref_scene*));
    for (m = 0; m < s->nm; ++m) {
        ref_scene_desc d;
        int n = s->mesh_count[m];
        int32_t zero = 0, none = -1, no_light = 0;
        float no_f = 0.f;
        float *pos, *nrm, *uv;
        int32_t* mesh;
        if (!is_walk_medium(s, m)) continue;
        pos = (float*)malloc(sizeof(float) * 9 * (size_t)n);
        nrm = (float*)malloc(sizeof(float) * 9 * (size_t)n);
        uv = (float*)malloc(sizeof(float) * 6 * (size_t)n);
        mesh = (int32_t*)calloc((size_t)n, sizeof(int32_t));
        medium_arrays(s, m, pos, nrm, uv);
        memset(&d, 0, sizeof(d));
        d.num_tris = n;
        d.pos = pos;
        d.nrm = nrm;
        d.uv = uv;
        d.mesh = mesh;
        d.num_meshes = 1;
        d.mesh_material = &none;
        d.mesh_light = &none;
        d.num_materials = 0;
        d.materials = &no_f;
        d.num_lights = 0;
        d.light_type = &zero;
        d.light_mesh = &none;
        d.light_emission = &no_f;
        (void)no_light;
        d.width = 1;
        d.height = 1;
        d.fovy = 45.f;
        d.direction[2] = -1.f;
        s->medium[m] = ref_scene_create(&d);
        free(pos);
        free(nrm);
        free(uv);
        free(mesh);
    }
}

/* Refit (the reference has none; it reconstructs the Scene, scene.cpp:50-57):
   new vertex positions, and shading normals unless nrm is NULL, for the same
   triangles.  The triangle constructor's normal, area and box
   (triangle.cu:3-10, bounds.h:22-26) and the mesh areas (scene.cpp:515-519)
   are recomputed; over the unchanged tree a leaf's box is the union of its
   triangles' boxes in leaf order from the empty box, as the build forms a
   node's box (bvh.cu:61-64), and an interior node's box is its left child's
   expanded by its right child's (bounds.h:35-40).  The camera is kept. */
void ref_scene_refit(ref_scene* s, const float* pos, const float* nrm) {
    int i, k;
    for (i = 0; i < s->nt; ++i) {
        Tri* t = &s->tris[i];
        V3 nn;
        for (k = 0; k < 3; ++k) {
            const float* p = pos + 9 * (size_t)i + 3 * k;
            t->p[k] = v3(p[0], p[1], p[2]);
            if (nrm) {
                const float* q = nrm + 9 * (size_t)i + 3 * k;
                t->n[k] = v3(q[0], q[1], q[2]);
            }
        }
        nn = vcross(vsub(t->p[1], t->p[0]), vsub(t->p[2], t->p[0]));
        t->area = 0.5f * vlen(nn);
        t->gn = vnormalize(nn);
        for (k = 0; k < 3; ++k) {
            t->b.lo[k] = smin(vget(t->p[0], k), smin(vget(t->p[1], k), vget(t->p[2], k)));
            t->b.hi[k] = smax(vget(t->p[0], k), smax(vget(t->p[1], k), vget(t->p[2], k)));
        }
    }
    for (i = 0; i <= s->nm; ++i) s->mesh_area[i] = 0.f;
    for (i = 0; i < s->nt; ++i) s->mesh_area[s->tris[i].mesh] += s->tris[i].area;
    for (i = 0; i < s->nt; ++i) s->ordered[i] = s->tris[s->order[i]];
    /* Children follow their parent in DFS pre-order: one reverse sweep. */
    for (i = s->nn - 1; i >= 0; --i) {
        LNode* n = &s->nodes[i];
        Box b = box_empty();
        if (n->tri_index >= 0) {
            for (k = 0; k < n->second; ++k) b = box_union(b, s->ordered[n->tri_index + k].b);
        } else {
            Box l, r;
            memcpy(l.lo, s->nodes[i + 1].lo, 12); memcpy(l.hi, s->nodes[i + 1].hi, 12);
            memcpy(r.lo, s->nodes[n->second].lo, 12); memcpy(r.hi, s->nodes[n->second].hi, 12);
            b = box_union(box_union(b, l), r);
        }
        memcpy(n->lo, b.lo, 12);
        memcpy(n->hi, b.hi, 12);
    }
    if (s->nn > 0) {
        memcpy(s->bounds.lo, s->nodes[0].lo, 12);
        memcpy(s->bounds.hi, s->nodes[0].hi, 12);
    }
    /* Media trees: refit over their unchanged topology, as the scene's. */
    for (i = 0; i < s->nm; ++i) {
        float *mp, *mn;
        if (!s->medium || !s->medium[i]) continue;
        mp = (float*)malloc(sizeof(float) * 9 * (size_t)s->mesh_count[i]);
        mn = (float*)malloc(sizeof(float) * 9 * (size_t)s->mesh_count[i]);
        medium_arrays(s, i, mp, mn, NULL);
        ref_scene_refit(s->medium[i], mp, mn);
        free(mp);
        free(mn);
    }
}

void ref_scene_destroy(ref_scene* s) {
    int m;
    if (!s) return;
    if (s->medium) {
        for (m = 0; m < s->nm; ++m) ref_scene_destroy(s->medium[m]);
        free(s->medium);
    }
    free(s->tris); free(s->ordered); free(s->order); free(s->nodes);
    free(s->mesh_material); free(s->mesh_light); free(s->mesh_area); free(s->mesh_start); free(s->mesh_count);
    free(s->mats); free(s->light_type); free(s->light_mesh); free(s->light_em);
    free(s->env); free(s->env_row); free(s->env_marg); free(s->env_alias_row); free(s->env_alias_marg);
    free(s->tex_info); free(s->tex_rgba);
    free(s);
}

int ref_bvh_num_nodes(const ref_scene* s) { return s->nn; }
int ref_bvh_max_depth(const ref_scene* s) { return s->max_depth; }
void ref_bvh_nodes(const ref_scene* s, int32_t* out8) { memcpy(out8, s->nodes, sizeof(LNode) * (size_t)s->nn); }
void ref_bvh_order(const ref_scene* s, int32_t* out) { memcpy(out, s->order, sizeof(int32_t) * (size_t)s->nt); }
void ref_camera(const ref_scene* s, float* o) {
    V3 v[4] = {s->cam_pos, s->cam_ld, s->cam_du, s->cam_dv};
    int k;
    for (k = 0; k < 4; ++k) { o[3 * k] = v[k].x; o[3 * k + 1] = v[k].y; o[3 * k + 2] = v[k].z; }
}
void ref_env_tables(const ref_scene* s, float* row_cdf, float* marg_cdf) {
    if (s->env_light < 0) return;
    memcpy(row_cdf, s->env_row, sizeof(float) * (size_t)s->env_w * s->env_h);
    memcpy(marg_cdf, s->env_marg, sizeof(float) * (size_t)s->env_h);
}

/* ------------------------------------------------------------ path tracing */
/* FP32 instantiation: the reference's float / double promotion rules. */
#define SFX _f
/* Texture::Sample (texture.cu:242-249) with the texture object of
   Texture::Load (texture.cu:226-237): u, v clamped to [0, 0.999999] (Clamp,
   mathutil.cu:72-79), v flipped, then the linear filter of a normalised-
   coordinate texture with U wrapping and V clamping: sample point
   (u W - 1/2, v' H - 1/2) between the four nearest texel centres, texels read
   as c / 255.  Weights in FP32 (the hardware rounds them to 8 fractional bits). */
static V3 tex_texel(const ref_scene* s, int tex, int x, int y) {
    const uint8_t* t = s->tex_rgba + 4 * ((size_t)s->tex_info[3 * tex] + (size_t)y * s->tex_info[3 * tex + 1] + x);
    return v3((float)t[0] / 255.f, (float)t[1] / 255.f, (float)t[2] / 255.f);
}

static V3 tex_albedo(const ref_scene* s, int tex, float u, float v) {
    int w = s->tex_info[3 * tex + 1], h = s->tex_info[3 * tex + 2];
    float x, y, ax, ay;
    int x0, y0, x1, y1;
    V3 a, b, c, d, top, bot;
    u = u >= 0.0f ? (u > 0.999999f ? 0.999999f : u) : 0.0f; /* NaN -> 0 */
    v = v >= 0.0f ? (v > 0.999999f ? 0.999999f : v) : 0.0f;
    x = u * (float)w - 0.5f;
    y = (1.0f - v) * (float)h - 0.5f;
    x0 = (int)floorf(x);
    y0 = (int)floorf(y);
    ax = x - (float)x0;
    ay = y - (float)y0;
    x1 = (x0 + 1) % w;
    x0 = ((x0 % w) + w) % w;
    y1 = y0 + 1 > h - 1 ? h - 1 : y0 + 1;
    y0 = y0 < 0 ? 0 : (y0 > h - 1 ? h - 1 : y0);
    a = tex_texel(s, tex, x0, y0);
    b = tex_texel(s, tex, x1, y0);
    c = tex_texel(s, tex, x0, y1);
    d = tex_texel(s, tex, x1, y1);
    top = v3(a.x * (1.f - ax) + b.x * ax, a.y * (1.f - ax) + b.y * ax, a.z * (1.f - ax) + b.z * ax);
    bot = v3(c.x * (1.f - ax) + d.x * ax, c.y * (1.f - ax) + d.y * ax, c.z * (1.f - ax) + d.z * ax);
    return v3(top.x * (1.f - ay) + bot.x * ay, top.y * (1.f - ay) + bot.y * ay, top.z * (1.f - ay) + bot.z * ay);
}

void ref_texture_albedo(const ref_scene* s, int tex, float u, float v, float* rgb) {
    V3 c = tex_albedo(s, tex, u, v);
    rgb[0] = c.x;
    rgb[1] = c.y;
    rgb[2] = c.z;
}

#define REAL float
#define FP32_MODE 1
#include "rtref_path.inc"
#undef SFX
#undef REAL
#undef FP32_MODE

/* FP64 twin: same estimator, shading in double. */
#define SFX _d
#define REAL double
#define FP32_MODE 0
#include "rtref_path.inc"
#undef SFX
#undef REAL
#undef FP32_MODE

void ref_render(const ref_scene* s, const ref_params* p, int sample0, int spp, const int32_t* pixels, int npix,
                float* acc_direct, float* acc_indirect, uint64_t* stats) {
    int i;
    uint64_t ext = 0, sh = 0, sss = 0;
    if (!pixels) npix = s->W * s->H;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : ext, sh, sss)
    for (i = 0; i < npix; ++i) {
        uint32_t px = pixels ? (uint32_t)pixels[i] : (uint32_t)i;
        int k;
        for (k = 0; k < spp; ++k) {
            float d[3], in[3];
            uint64_t st[3] = {0, 0, 0};
            path_radiance_f(s, p, px, (uint32_t)(sample0 + k), d, in, st);
            acc_direct[4 * (size_t)i] += d[0];
            acc_direct[4 * (size_t)i + 1] += d[1];
            acc_direct[4 * (size_t)i + 2] += d[2];
            acc_indirect[4 * (size_t)i] += in[0];
            acc_indirect[4 * (size_t)i + 1] += in[1];
            acc_indirect[4 * (size_t)i + 2] += in[2];
            ext += st[0];
            sh += st[1];
            sss += st[2];
        }
    }
    if (stats) {
        stats[0] += ext;
        stats[1] += sh;
        stats[2] += sss;
    }
}

void ref_render_f64(const ref_scene* s, const ref_params* p, int sample0, int spp, const int32_t* pixels, int npix,
                    double* acc_direct, double* acc_indirect) {
    int i;
    if (!pixels) npix = s->W * s->H;
#pragma omp parallel for schedule(dynamic, 64)
    for (i = 0; i < npix; ++i) {
        uint32_t px = pixels ? (uint32_t)pixels[i] : (uint32_t)i;
        int k;
        for (k = 0; k < spp; ++k) {
            double d[3], in[3];
            uint64_t st[3] = {0, 0, 0};
            path_radiance_d(s, p, px, (uint32_t)(sample0 + k), d, in, st);
            acc_direct[4 * (size_t)i] += d[0];
            acc_direct[4 * (size_t)i + 1] += d[1];
            acc_direct[4 * (size_t)i + 2] += d[2];
            acc_indirect[4 * (size_t)i] += in[0];
            acc_indirect[4 * (size_t)i + 1] += in[1];
            acc_indirect[4 * (size_t)i + 2] += in[2];
        }
    }
}

/* ------------------------------------------------------------ test hooks */
static V3 v3p(const float* p) { return v3(p[0], p[1], p[2]); }

float ref_bsdf_eval(const float* mat, const float* wi, const float* n, const float* wo, float* out3) {
    Mat_f m = load_mat_f(mat);
    Vec_f a = from_f(v3p(wi)), nn = from_f(v3p(n)), b = from_f(v3p(wo));
    Vec_f f = eval_f(&m, a, nn, b);
    out3[0] = f.x;
    out3[1] = f.y;
    out3[2] = f.z;
    return pdf_f(&m, a, nn, b);
}

/* u3: three uniforms; diffuse uses (u0, u1), PBR (r1, r2, choose), dielectric u0. */
float ref_bsdf_sample(const float* mat, const float* wi, const float* n, const float* u3, float* wo_out) {
    Mat_f m = load_mat_f(mat);
    Vec_f a = from_f(v3p(wi)), nn = from_f(v3p(n)), w;
    float pdf;
    if (m.type == 2) {
        w = sample_pbr_f(&m, a, nn, u3[0], u3[1], u3[2]);
        pdf = pdf_pbr_f(&m, a, nn, w);
    } else if (m.type == 3) {
        Vec_f wt;
        sample_dielectric_f(&m, a, nn, u3[0], &w, &pdf, &wt);
    } else {
        w = sample_diffuse_f(nn, u3[0], u3[1]);
        pdf = pdf_diffuse_f(nn, w);
    }
    wo_out[0] = w.x;
    wo_out[1] = w.y;
    wo_out[2] = w.z;
    return pdf;
}

float ref_fresnel_dielectric(float cos_i, float eta_i, float eta_t) { return fresnel_f(cos_i, eta_i, eta_t); }

float ref_ggx_d(const float* mat, const float* wi, const float* n, const float* wo) {
    Mat_f m = load_mat_f(mat);
    return ggx_d_f(&m, from_f(v3p(wi)), from_f(v3p(n)), from_f(v3p(wo)));
}

float ref_burley(const float* mat, float r, int ch) {
    Mat_f m = load_mat_f(mat);
    return burley_f(&m, r, ch);
}

float ref_sss_dist(float scatter_distance, float u) { return sss_dist_f(scatter_distance, u); }

void ref_sss_probe(const ref_scene* s, const float* p, const float* n, float dist, float angle, float* out_p,
                   float* out_n) {
    Vec_f op, on;
    probe_f(s, from_f(v3p(p)), from_f(v3p(n)), dist, angle, &op, &on);
    out_p[0] = op.x;
    out_p[1] = op.y;
    out_p[2] = op.z;
    out_n[0] = on.x;
    out_n[1] = on.y;
    out_n[2] = on.z;
}

int ref_sss_walk(const ref_scene* s, int mesh, const float* p, const float* n_out, float e1, float e2, uint32_t seed,
                 uint32_t pixel, uint32_t sample, uint32_t bounce, float* out_p, float* out_n, float* out_w,
                 int* steps) {
    Mat_f m;
    Vec_f op = mk_f(0, 0, 0), on = mk_f(0, 0, 0), w = mk_f(0, 0, 0);
    int ok;
    *steps = 0;
    if (mesh < 0 || mesh >= s->nm || !s->medium[mesh]) return 0;
    m = load_mat_f(s->mats + REF_MAT_STRIDE * (size_t)s->mesh_material[mesh]);
    ok = rw_walk_f(s->medium[mesh], &m, from_f(v3p(p)), from_f(v3p(n_out)), e1, e2, seed, pixel, sample, bounce, &op,
                   &on, &w, steps);
    out_p[0] = op.x; out_p[1] = op.y; out_p[2] = op.z;
    out_n[0] = on.x; out_n[1] = on.y; out_n[2] = on.z;
    out_w[0] = w.x; out_w[1] = w.y; out_w[2] = w.z;
    return ok;
}

/* n walks from the same entry point and normal, walk i with pixel i and the
   entry uniforms e[2i], e[2i+1] (OpenMP over walks). */
void ref_sss_walk_batch(const ref_scene* s, int mesh, const float* p, const float* n_out, int n, const float* e,
                        uint32_t seed, int32_t* ok, float* out_p, float* out_n, float* out_w, int32_t* steps) {
    int i;
#pragma omp parallel for schedule(dynamic, 256)
    for (i = 0; i < n; ++i) {
        int st = 0;
        ok[i] = ref_sss_walk(s, mesh, p, n_out, e[2 * i], e[2 * i + 1], seed, (uint32_t)i, 0u, 0u, out_p + 3 * (size_t)i,
                             out_n + 3 * (size_t)i, out_w + 3 * (size_t)i, &st);
        steps[i] = st;
    }
}

void ref_env_sample(const ref_scene* s, float rv, float ru, float* uv_out) {
    float u, v;
    env_sample_uv_f(s, rv, ru, &u, &v);
    uv_out[0] = u;
    uv_out[1] = v;
}

float ref_env_pdf_uv(const ref_scene* s, float u, float v) { return env_pdf_uv_f(s, u, v); }

void ref_env_alias_tables(const ref_scene* s, float* row2, float* marg2) {
    if (s->env_light < 0) return;
    memcpy(row2, s->env_alias_row, sizeof(float) * 2 * (size_t)s->env_w * s->env_h);
    memcpy(marg2, s->env_alias_marg, sizeof(float) * 2 * (size_t)s->env_h);
}

void ref_env_sample_alias(const ref_scene* s, float rv, float ru, float* uv_out) {
    float u, v;
    env_sample_uv_alias_f(s, rv, ru, &u, &v);
    uv_out[0] = u;
    uv_out[1] = v;
}

#include "rtref_svgf.inc"
