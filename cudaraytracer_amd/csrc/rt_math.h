// Scalar and 3-vector arithmetic shared by the host loader/BVH builder and the
// gfx950 kernels.  Every operation is written with an explicit evaluation order
// so that host and device produce bit-identical FP32 results when both are
// compiled with -ffp-contract=off (no FMA contraction).
//
// Evaluation order follows the Eigen 3.4 expressions the reference uses
// (reference: ThirdParty/Eigen3/.../Core/Redux.h:99-113 — a 3-term reduction is
// evaluated as a0 + (a1 + a2); Geometry/OrthoMethods.h:44-48 — cross product;
// Core/Dot.h:124-134 — normalized() divides by sqrt(squaredNorm) when > 0).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#define RT_HD __host__ __device__ __forceinline__

namespace rt {

struct vec3 {
    float x, y, z;
};

RT_HD vec3 v3(float x, float y, float z) { vec3 r; r.x = x; r.y = y; r.z = z; return r; }
RT_HD int32_t f_as_i(float f) { int32_t i; __builtin_memcpy(&i, &f, 4); return i; }
RT_HD float i_as_f(int32_t i) { float f; __builtin_memcpy(&f, &i, 4); return f; }
RT_HD vec3 operator+(vec3 a, vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
RT_HD vec3 operator-(vec3 a, vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
RT_HD vec3 operator-(vec3 a) { return v3(-a.x, -a.y, -a.z); }
// vector * scalar and scalar * vector are the same IEEE products.
RT_HD vec3 operator*(vec3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
RT_HD vec3 operator*(float s, vec3 a) { return v3(s * a.x, s * a.y, s * a.z); }
RT_HD vec3 operator/(vec3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
RT_HD vec3 cmul(vec3 a, vec3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
RT_HD float dot(vec3 a, vec3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
RT_HD float norm2(vec3 a) { return dot(a, a); }
RT_HD float length(vec3 a) { return sqrtf(norm2(a)); }
RT_HD vec3 cross(vec3 a, vec3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
RT_HD vec3 normalize(vec3 a) {
    float z = norm2(a);
    if (z > 0.0f) return a / sqrtf(z);
    return a;
}
// Linear interpolation v1*(1-t) + v2*t (reference: Src/Utils/mathutil.cu:135-137).
RT_HD vec3 lerp(vec3 a, vec3 b, float t) { return a * (1.0f - t) + b * t; }
// I - (2 I.N) N (reference: Src/Utils/mathutil.cu:139-141).
RT_HD vec3 reflect(vec3 i, vec3 n) { return i - (2.0f * dot(i, n)) * n; }
// Rec.709 luminance (reference: Src/Utils/mathutil.cu:143-146).
RT_HD float luminance(vec3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
// Clamp(lo, hi, v) (reference: Src/Utils/mathutil.cu:72-79).
RT_HD float clampf(float lo, float hi, float v) {
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}
RT_HD int clampi(int lo, int hi, int v) {
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

// Orthonormal basis from a normal and a local direction
// (reference: Src/Utils/mathutil.cu:48-60).
RT_HD vec3 local_to_world(vec3 v, vec3 normal) {
    vec3 n = normalize(normal);
    vec3 up = fabsf(n.x) > fabsf(n.y) ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f);
    vec3 a = normalize(cross(up, n));
    vec3 b = normalize(cross(n, a));
    return (a * v.x + b * v.y) + n * v.z;
}

// Tangent frame used by the subsurface probe (reference: Src/Utils/mathutil.cu:148-155).
RT_HD void create_onb(vec3 n, vec3& t, vec3& b) {
    if (fabsf(n.x) > fabsf(n.y))
        t = normalize(v3(-n.z, 0.f, n.x));
    else
        t = normalize(v3(0.f, n.z, -n.y));
    b = normalize(cross(n, t));
}

// ---------------------------------------------------------------------------
// Transcendentals.  Both the kernels and the host use these (never the
// platform libm) so that per-sample radiance is reproducible bit-for-bit
// between the GPU and a host build.  The methods are S. L. Moshier's Cephes
// single-precision routines ("Methods and Programs for Mathematical
// Functions", 1989): Cody-Waite argument reduction followed by minimax
// polynomials.  Accuracy against libm double is checked in the test suite.
// ---------------------------------------------------------------------------
namespace fm {
constexpr float kPi = 3.14159265358979323846f;
constexpr float kPiO2 = 1.57079632679489661923f;
constexpr float kPiO4 = 0.78539816339744830962f;

RT_HD float sin_poly(float r, float z) {
    return ((((-1.9515295891e-4f * z) + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z) * r + r;
}
RT_HD float cos_poly(float z) {
    return (((((2.443315711809948e-5f * z) - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z) * z - 0.5f * z) + 1.0f;
}
// Reduce |x| to r in [-pi/4, pi/4] and octant j in [0, 7] (Cody-Waite, 3-part pi/4).
RT_HD float reduce_quarter_pi(float ax, int& j) {
    int q = (int)(ax * 1.27323954473516f);
    float y = (float)q;
    if (q & 1) { q += 1; y += 1.0f; }
    j = q & 7;
    return ((ax - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
}
RT_HD float sinf_(float x) {
    float sign = 1.0f;
    float ax = x;
    if (x < 0.0f) { sign = -1.0f; ax = -x; }
    int j;
    float r = reduce_quarter_pi(ax, j);
    if (j > 3) { sign = -sign; j -= 4; }
    float z = r * r;
    float y = (j == 1 || j == 2) ? cos_poly(z) : sin_poly(r, z);
    return sign < 0.0f ? -y : y;
}
RT_HD float cosf_(float x) {
    float ax = fabsf(x);
    int j;
    float r = reduce_quarter_pi(ax, j);
    float sign = 1.0f;
    if (j > 3) { j -= 4; sign = -sign; }
    if (j > 1) sign = -sign;
    float z = r * r;
    float y = (j == 1 || j == 2) ? sin_poly(r, z) : cos_poly(z);
    return sign < 0.0f ? -y : y;
}
// asin for |a| <= 0.5 (odd polynomial in a).
RT_HD float asin_small(float a) {
    float z = a * a;
    return (((((((4.2163199048e-2f * z) + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z + 1.6666752422e-1f) * z) * a) + a;
}
RT_HD float asinf_(float x) {
    float a = fabsf(x);
    float r;
    if (a > 1.0f) return (x - x) / (x - x);  // NaN outside the domain
    if (a > 0.5f) {
        float s = sqrtf(0.5f * (1.0f - a));
        r = kPiO2 - 2.0f * asin_small(s);
    } else {
        r = asin_small(a);
    }
    return x < 0.0f ? -r : r;
}
RT_HD float acosf_(float x) {
    if (x < -1.0f || x > 1.0f) return (x - x) / (x - x);
    if (x < -0.5f) return kPi - 2.0f * asin_small(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * asin_small(sqrtf(0.5f * (1.0f - x)));
    return kPiO2 - asin_small(x);
}
// 2^n as a float for n in [-126, 127].
RT_HD float exp2i(int n) {
    uint32_t bits = (uint32_t)(n + 127) << 23;
    float f;
    __builtin_memcpy(&f, &bits, 4);
    return f;
}
RT_HD float expf_(float x) {
    if (x != x) return x;
    if (x > 88.72283905206835f) return __builtin_huge_valf();
    if (x < -103.278929903431851103f) return 0.0f;
    float z = floorf(1.44269504088896341f * x + 0.5f);
    float r = x - z * 0.693359375f;
    r = r - z * -2.12194440e-4f;
    int n = (int)z;
    float r2 = r * r;
    float p = ((((((1.9875691500e-4f * r) + 1.3981999507e-3f) * r + 8.3334519073e-3f) * r + 4.1665795894e-2f) * r + 1.6666665459e-1f) * r + 5.0000001201e-1f);
    float e = ((p * r2) + r) + 1.0f;
    // scale by 2^n in two steps so that n outside [-126, 127] stays exact.
    int n1 = n / 2;
    int n2 = n - n1;
    return (e * exp2i(n1)) * exp2i(n2);
}
RT_HD float logf_(float x) {
    if (x != x) return x;
    if (x < 0.0f) return (x - x) / (x - x);
    if (x == 0.0f) return -__builtin_huge_valf();
    if (x == __builtin_huge_valf()) return x;
    uint32_t bits;
    __builtin_memcpy(&bits, &x, 4);
    int e = 0;
    if ((bits >> 23) == 0) {  // subnormal: scale up by 2^25
        x = x * 33554432.0f;
        __builtin_memcpy(&bits, &x, 4);
        e = -25;
    }
    e += (int)((bits >> 23) & 0xff) - 126;
    bits = (bits & 0x807fffffu) | 0x3f000000u;  // mantissa in [0.5, 1)
    float m;
    __builtin_memcpy(&m, &bits, 4);
    if (m < 0.707106781186547524f) {
        e -= 1;
        m = (m + m) - 1.0f;
    } else {
        m = m - 1.0f;
    }
    float z = m * m;
    float y = (((((((((7.0376836292e-2f * m) - 1.1514610310e-1f) * m + 1.1676998740e-1f) * m - 1.2420140846e-1f) * m + 1.4249322787e-1f) * m - 1.6668057665e-1f) * m + 2.0000714765e-1f) * m - 2.4999993993e-1f) * m + 3.3333331174e-1f);
    y = (y * m) * z;
    float fe = (float)e;
    y = y + -2.12194440e-4f * fe;
    y = y + -0.5f * z;
    float r = m + y;
    r = r + 0.693359375f * fe;
    return r;
}
RT_HD float atanf_(float x) {
    float sign = 1.0f;
    if (x < 0.0f) { sign = -1.0f; x = -x; }
    float y;
    if (x > 2.414213562373095f) {
        y = kPiO2;
        x = -(1.0f / x);
    } else if (x > 0.4142135623730950f) {
        y = kPiO4;
        x = (x - 1.0f) / (x + 1.0f);
    } else {
        y = 0.0f;
    }
    float z = x * x;
    y = y + (((((((8.05374449538e-2f * z) - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z) * x) + x);
    return sign < 0.0f ? -y : y;
}
RT_HD float atan2f_(float y, float x) {
    int code = 0;
    if (x < 0.0f) code = 2;
    if (y < 0.0f) code |= 1;
    if (x == 0.0f) {
        if (code & 1) return -kPiO2;
        if (y == 0.0f) return 0.0f;
        return kPiO2;
    }
    if (y == 0.0f) {
        if (code & 2) return kPi;
        return 0.0f;
    }
    float w = 0.0f;
    if (code == 2) w = kPi;
    else if (code == 3) w = -kPi;
    return w + atanf_(y / x);
}
// x^5 as two squarings and a product.
RT_HD float pow5(float x) {
    float x2 = x * x;
    return (x2 * x2) * x;
}
// x^y for x >= 0 as exp(y log x); 0^y = 0 for y > 0 and 1 otherwise.
RT_HD float powf_(float x, float y) {
    if (x == 0.0f) return y > 0.0f ? 0.0f : 1.0f;
    return expf_(y * logf_(x));
}
}  // namespace fm

// ---------------------------------------------------------------------------
// Counter-based RNG: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11).  A draw is addressed by
// (pixel, sample, bounce, dimension, stream), so a path's random numbers do
// not depend on queue order, batch size or launch shape.
// ---------------------------------------------------------------------------
struct philox_out {
    uint32_t v[4];
};

RT_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                         uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Uniform float in (0, 1] from the top 24 bits: the same half-open range as
// the reference's sampler (curand_uniform excludes 0 and includes 1).
RT_HD float u01(uint32_t x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }

// kStreamWalk: one coordinate per step of a random walk through a
// subsurface medium, bounce field (bounce << 8) | step.
enum RngStream : uint32_t { kStreamCamera = 0u, kStreamPath = 1u, kStreamWalk = 2u };

// Sequential draws within one (pixel, sample, bounce, stream) coordinate.
// Draw d uses Philox block d/4, lane d%4.
struct Rng {
    uint32_t k0, k1, pixel, sample, bounce, stream;
    uint32_t b0, b1, b2, b3;
    int dim;
    RT_HD void init(uint32_t seed, uint32_t pixel_, uint32_t sample_, uint32_t bounce_, uint32_t stream_) {
        k0 = seed;
        k1 = 0x5EEDC0DEu;
        pixel = pixel_;
        sample = sample_;
        bounce = bounce_;
        stream = stream_;
        dim = 0;
        b0 = b1 = b2 = b3 = 0u;
    }
    RT_HD float next() {
        int lane = dim & 3;
        if (lane == 0) {
            uint32_t o[4];
            philox4x32_10(pixel, sample, (bounce << 8) | (uint32_t)(dim >> 2), stream, k0, k1, o);
            b0 = o[0]; b1 = o[1]; b2 = o[2]; b3 = o[3];
        }
        uint32_t x = lane == 0 ? b0 : (lane == 1 ? b1 : (lane == 2 ? b2 : b3));
        ++dim;
        return u01(x);
    }
};

}  // namespace rt
