// Scene-JSON and OBJ loading (reference semantics: Src/Core/scene.cpp:129-690,
// Src/Utils/objloader.cpp:66-210), camera set-up and BVH build.
#include "scene.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <iterator>
#include <limits>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <exception>
#include <mutex>
#include <system_error>
#include <thread>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

#include "json.h"

namespace rt {
namespace {

// 4x4 float matrix, row-major m[r][c].
struct Mat4 {
    float m[4][4];
};
Mat4 identity() {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = i == j ? 1.f : 0.f;
    return r;
}
// Row r, column c = ((a[r][0] b[0][c] + a[r][1] b[1][c]) + a[r][2] b[2][c]) + a[r][3] b[3][c].
Mat4 mul(const Mat4& a, const Mat4& b) {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i][j] = ((a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j]) + a.m[i][3] * b.m[3][j];
    return r;
}
// AngleToRadian: angle / 180 * PI with PI the double 3.1415926 (reference:
// Src/Utils/mathutil.cu:4-7, Src/Core/global.h:7).
float deg2rad(float a) { return (float)((double)(a / 180.f) * 3.1415926); }
float rad2deg(float r) { return (float)((double)(r * 180.f) / 3.1415926); }

Mat4 rotate_axis(int axis, float deg) {
    float rad = deg2rad(deg);
    float c = cosf(rad), s = sinf(rad);
    Mat4 r = identity();
    if (axis == 0) { r.m[1][1] = c; r.m[1][2] = -s; r.m[2][1] = s; r.m[2][2] = c; }
    else if (axis == 1) { r.m[0][0] = c; r.m[0][2] = s; r.m[2][0] = -s; r.m[2][2] = c; }
    else { r.m[0][0] = c; r.m[0][1] = -s; r.m[1][0] = s; r.m[1][1] = c; }
    return r;
}
// Matrix4f::inverse as the reference's Eigen 3.4.0 computes it on x86-64
// (Eigen/src/LU/arch/InverseSize4.h, the SSE path: inverse = adjugate / det
// formed block-wise from the four 2x2 blocks), written out lane by lane in
// FP32 with the same products, sums and their order; bit-identical to that
// code on random and affine matrices (checked against a build of the
// reference's Eigen in the tests).  The 2x2 blocks are named by their place
// in the matrix: p = rows 0-1 x cols 0-1, q = rows 2-3 x cols 0-1,
// r = rows 0-1 x cols 2-3, s = rows 2-3 x cols 2-3, each read column by
// column (x0 = (0,0), x1 = (1,0), x2 = (0,1), x3 = (1,1) of the block).
Mat4 inverse(const Mat4& a) {
    const float p0 = a.m[0][0], p1 = a.m[1][0], p2 = a.m[0][1], p3 = a.m[1][1];
    const float q0 = a.m[2][0], q1 = a.m[3][0], q2 = a.m[2][1], q3 = a.m[3][1];
    const float r0 = a.m[0][2], r1 = a.m[1][2], r2 = a.m[0][3], r3 = a.m[1][3];
    const float s0 = a.m[2][2], s1 = a.m[3][2], s2 = a.m[2][3], s3 = a.m[3][3];
    // adj(p) * q and adj(s) * r
    const float pq[4] = {p3 * q0 - p1 * q2, p3 * q1 - p1 * q3, p0 * q2 - p2 * q0, p0 * q3 - p2 * q1};
    const float sr[4] = {s3 * r0 - s1 * r2, s3 * r1 - s1 * r3, s0 * r2 - s2 * r0, s0 * r3 - s2 * r1};
    const float dp = p3 * p0 - p1 * p2, dq = q3 * q0 - q1 * q2, dr = r3 * r0 - r1 * r2, ds = s3 * s0 - s1 * s2;
    const float tr = (sr[0] * pq[0] + sr[1] * pq[2]) + (sr[2] * pq[1] + sr[3] * pq[3]);
    const float det = (dp * ds + dq * dr) - tr;
    if (det == 0.f || !std::isfinite(det)) throw std::runtime_error("object transform is singular");
    const float rd = 1.0f / det;
    const float sg[4] = {rd, -rd, -rd, rd};
    float ip[4], iq[4], ir[4], is[4];
    ip[0] = p0 * ds - (q0 * sr[0] + q1 * sr[2]);
    ip[1] = p1 * ds - (q0 * sr[1] + q1 * sr[3]);
    ip[2] = p2 * ds - (q2 * sr[0] + q3 * sr[2]);
    ip[3] = p3 * ds - (q2 * sr[1] + q3 * sr[3]);
    is[0] = s0 * dp - (r0 * pq[0] + r1 * pq[2]);
    is[1] = s1 * dp - (r0 * pq[1] + r1 * pq[3]);
    is[2] = s2 * dp - (r2 * pq[0] + r3 * pq[2]);
    is[3] = s3 * dp - (r2 * pq[1] + r3 * pq[3]);
    iq[0] = r0 * dq - (s0 * pq[3] - s1 * pq[2]);
    iq[1] = r1 * dq - (s1 * pq[0] - s0 * pq[1]);
    iq[2] = r2 * dq - (s2 * pq[3] - s3 * pq[2]);
    iq[3] = r3 * dq - (s3 * pq[0] - s2 * pq[1]);
    ir[0] = q0 * dr - (p0 * sr[3] - p1 * sr[2]);
    ir[1] = q1 * dr - (p1 * sr[0] - p0 * sr[1]);
    ir[2] = q2 * dr - (p2 * sr[3] - p3 * sr[2]);
    ir[3] = q3 * dr - (p3 * sr[0] - p2 * sr[1]);
    for (int k = 0; k < 4; ++k) {
        ip[k] *= sg[k];
        iq[k] *= sg[k];
        ir[k] *= sg[k];
        is[k] *= sg[k];
    }
    Mat4 r;
    const float cols[4][4] = {{ip[3], ip[1], iq[3], iq[1]},
                              {ip[2], ip[0], iq[2], iq[0]},
                              {ir[3], ir[1], is[3], is[1]},
                              {ir[2], ir[0], is[2], is[0]}};
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row) r.m[row][c] = cols[c][row];
    return r;
}
vec3 xform_point(const Mat4& t, vec3 p) {
    float o[3];
    for (int r = 0; r < 3; ++r)
        o[r] = ((t.m[r][0] * p.x + t.m[r][1] * p.y) + t.m[r][2] * p.z) + t.m[r][3] * 1.f;
    return v3(o[0], o[1], o[2]);
}
// transformInv.transpose() * (n, 1), then normalized (reference: Src/Utils/objloader.cpp:36-41).
// Eigen evaluates each coefficient of a transposed (row-major) 4x4 times a
// vector as one 4-wide product summed pairwise, (x0 + x2) + (x1 + x3).
vec3 xform_normal(const Mat4& inv, vec3 n) {
    float o[3];
    for (int r = 0; r < 3; ++r)
        o[r] = (inv.m[0][r] * n.x + inv.m[2][r] * n.z) + (inv.m[1][r] * n.y + inv.m[3][r] * 1.f);
    return normalize(v3(o[0], o[1], o[2]));
}

std::string read_file(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open file: " + path);
    std::string data;
    if (std::fseek(f, 0, SEEK_END) == 0) {
        long size = std::ftell(f);
        if (size > 0 && std::fseek(f, 0, SEEK_SET) == 0) {
            data.resize((size_t)size);
            data.resize(std::fread(&data[0], 1, (size_t)size, f));
        }
    }
    // Files whose size is not known up front (pipes) are read to the end.
    char buf[1 << 16];
    for (size_t got; (got = std::fread(buf, 1, sizeof buf, f)) > 0;) data.append(buf, got);
    bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) throw std::runtime_error("cannot read file: " + path);
    return data;
}
std::string dir_of(const std::string& p) {
    size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}
// GetFilenameFromPath: basename without the last extension (reference: Src/Utils/stringutil.cpp:18-33).
std::string stem_of(const std::string& p) {
    size_t k = p.find_last_of('/');
    std::string f = k == std::string::npos ? p : p.substr(k + 1);
    size_t d = f.find_last_of('.');
    if (d != std::string::npos && d > 0) return f.substr(0, d);
    return f;
}
std::string resolve_path(const std::string& base_dir, const std::string& p) {
    if (p.empty() || p[0] == '/') return p;
    return base_dir + p;
}

// ---------------------------------------------------------------- OBJ ----
// OBJ import as the reference performs it: OBJLoader::LoadObject
// (Src/Utils/objloader.cpp:66-210) over its vendored tinyobjloader
// (ThirdParty/tinyobj/tiny_obj_loader.h: LoadObj with triangulation).  Kept:
//  - lines end at '\n', '\r' or "\r\n"; leading spaces and tabs are skipped,
//    '#' starts a comment;
//  - numbers are read as tinyobj's parseReal / tryParseDouble read them
//    (:804-953): digits gathered into a double (integer digits times ten,
//    fraction digits times 10^-k from a table of doubles up to k = 7 and pow
//    beyond), an exponent applied as ldexp(m * pow(5, e), e), then rounded to
//    float; an unreadable field keeps its default (0);
//  - face corners v, v/vt, v//vn, v/vt/vn with tinyobj's index rule (fixIndex,
//    :738-759): k > 0 is k - 1, k < 0 counts back from the elements read so far;
//    a 0 or unreadable index makes LoadObj fail, and the object then yields no
//    meshes (Scene::ParseObjects warns and skips it);
//  - polygons are triangulated by tinyobj's ear clipping (exportGroupsToShape,
//    :1330-1560) in FP32, which gives up on the rest of a polygon when it finds
//    no ear;
//  - 'g', 'o' and 'usemtl' move faces between shapes exactly as LoadObj does
//    (:2306-2420, :2575-2585): a 'usemtl' that changes the material closes the
//    current face group into the current shape, 'g' closes the shape when it
//    holds triangles, 'o' when its last group held any primitive, end of file
//    when either holds;
//  - materials are the 'newmtl' names of the first readable file of each
//    'mtllib' line (base = the OBJ's directory; first definition of a name
//    wins); a 'usemtl' name no such file defines is no material;
//  - one mesh per (shape, material) in order of first appearance, named as
//    BuildMeshName names it; positions transform * (p, 1), normals
//    (transformInv^T (n, 1)).normalized(), and the face normal of the
//    transformed positions where a triangle lacks a normal index.
// Deviations, only where the reference reads out of bounds: a corner past the
// last vertex, normal or texture coordinate of the file is an error, and the
// ear clipping sees the vertices read up to the face's own line (tinyobj: up
// to the end of its face group; the same for any file whose faces only name
// vertices defined before them).
//
// The text is parsed in line-aligned chunks on host threads; the ear clipping
// runs per face in parallel; one pass in file order replays the statements;
// the triangles are written in parallel.
struct LoadedMesh {
    std::string name;
    std::string material;
    std::vector<float> pos, nrm, uv;  // 9, 9, 6 per triangle
    int count = 0;
};

namespace {

// tinyobj's tryParseDouble (tiny_obj_loader.h:804-930).
bool obj_parse_double(const char* s, const char* e, double* out) {
    static const double kFrac[8] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001};
    if (s >= e) return false;
    double mant = 0.0;
    int expo = 0, read = 0;
    char sign = '+', esign = '+';
    const char* c = s;
    bool lead_dot = false, more;
    auto digit = [](char ch) { return ch >= '0' && ch <= '9'; };
    if (*c == '+' || *c == '-') {
        sign = *c++;
        if (c != e && *c == '.') lead_dot = true;
    } else if (digit(*c)) {
    } else if (*c == '.') {
        lead_dot = true;
    } else {
        return false;
    }
    more = c != e;
    if (!lead_dot) {
        while (more && digit(*c)) {
            mant *= 10;
            mant += (int)(*c - '0');
            ++c;
            ++read;
            more = c != e;
        }
        if (read == 0) return false;
    }
    if (!more) goto assemble;
    if (*c == '.') {
        ++c;
        read = 1;
        more = c != e;
        while (more && digit(*c)) {
            mant += (int)(*c - '0') * (read < 8 ? kFrac[read] : std::pow(10.0, -read));
            ++read;
            ++c;
            more = c != e;
        }
    } else if (*c == 'e' || *c == 'E') {
    } else {
        goto assemble;
    }
    if (!more) goto assemble;
    if (*c == 'e' || *c == 'E') {
        ++c;
        more = c != e;
        if (more && (*c == '+' || *c == '-')) esign = *c++;
        else if (!digit(*c)) return false;
        read = 0;
        more = c != e;
        while (more && digit(*c)) {
            expo *= 10;
            expo += *c - '0';
            ++c;
            ++read;
            more = c != e;
        }
        expo *= esign == '+' ? 1 : -1;
        if (read == 0) return false;
    }
assemble:
    *out = (sign == '+' ? 1 : -1) * (expo ? std::ldexp(mant * std::pow(5.0, expo), expo) : mant);
    return true;
}

// parseReal: one whitespace-delimited field, `def` when it does not parse.
float obj_real(const char*& t, double def = 0.0) {
    t += std::strspn(t, " \t");
    const char* end = t + std::strcspn(t, " \t\r");
    double v = def;
    obj_parse_double(t, end, &v);
    t = end;
    return (float)v;
}

// A field of tinyobj's parseString: skip spaces and tabs, read to the next blank.
std::string obj_string(const char*& t) {
    t += std::strspn(t, " \t");
    size_t n = std::strcspn(t, " \t\r");
    std::string s(t, n);
    t += n;
    return s;
}

inline bool obj_space(char c) { return c == ' ' || c == '\t'; }
inline bool obj_newline(char c) { return c == '\r' || c == '\n' || c == '\0'; }

constexpr int32_t kObjAbsent = INT32_MIN;  // corner field not given
struct ObjCorner {
    int32_t v, vt, vn;  // raw OBJ indices (0 = unreadable), kObjAbsent when not given
};
struct ObjFace {
    uint32_t first, count;  // corners in the chunk's corner array
    int32_t nv, nvt, nvn;   // the chunk's element counts at the face line
};
enum ObjEventKind { kObjGroup, kObjObject, kObjUsemtl, kObjMtllib, kObjPrimitive };
struct ObjEvent {
    size_t face;  // chunk faces before the statement
    int kind;
    std::string text;
};
struct ObjChunk {
    std::vector<float> V, VN, VT;
    std::vector<ObjCorner> corners;
    std::vector<ObjFace> faces;
    std::vector<ObjEvent> events;
    // ear clipping: per face its first triangle and count in `tris` (corner positions within the face)
    std::vector<uint32_t> tri_first, tri_count;
    std::vector<std::array<uint32_t, 3>> tris;
};

// parseTriple's index reading: atoi of the field, then the rest of it skipped.
int32_t obj_index(const char*& t) {
    int32_t v = (int32_t)std::atoi(t);
    t += std::strcspn(t, "/ \t\r");
    return v;
}

void parse_obj_line(const char* c, ObjChunk& ck) {
    c += std::strspn(c, " \t");
    if (c[0] == '\0' || c[0] == '#') return;
    if (c[0] == 'v' && obj_space(c[1])) {
        const char* t = c + 2;
        float x = obj_real(t), y = obj_real(t), z = obj_real(t);
        ck.V.push_back(x); ck.V.push_back(y); ck.V.push_back(z);
    } else if (c[0] == 'v' && c[1] == 'n' && obj_space(c[2])) {
        const char* t = c + 3;
        float x = obj_real(t), y = obj_real(t), z = obj_real(t);
        ck.VN.push_back(x); ck.VN.push_back(y); ck.VN.push_back(z);
    } else if (c[0] == 'v' && c[1] == 't' && obj_space(c[2])) {
        const char* t = c + 3;
        float x = obj_real(t), y = obj_real(t);
        ck.VT.push_back(x); ck.VT.push_back(y);
    } else if (c[0] == 'f' && obj_space(c[1])) {
        const char* t = c + 2;
        t += std::strspn(t, " \t");
        ObjFace f{(uint32_t)ck.corners.size(), 0, (int32_t)(ck.V.size() / 3), (int32_t)(ck.VT.size() / 2),
                  (int32_t)(ck.VN.size() / 3)};
        while (!obj_newline(t[0])) {
            ObjCorner k{obj_index(t), kObjAbsent, kObjAbsent};
            if (t[0] == '/') {
                ++t;
                if (t[0] == '/') {
                    ++t;
                    k.vn = obj_index(t);
                } else {
                    k.vt = obj_index(t);
                    if (t[0] == '/') {
                        ++t;
                        k.vn = obj_index(t);
                    }
                }
            }
            ck.corners.push_back(k);
            ++f.count;
            t += std::strspn(t, " \t\r");
        }
        ck.faces.push_back(f);
    } else if ((c[0] == 'l' || c[0] == 'p') && obj_space(c[1])) {
        ck.events.push_back(ObjEvent{ck.faces.size(), kObjPrimitive, std::string()});
    } else if (c[0] == 'g' && obj_space(c[1])) {
        const char* t = c;
        std::vector<std::string> names;
        while (!obj_newline(t[0])) {
            names.push_back(obj_string(t));
            t += std::strspn(t, " \t\r");
        }
        std::string joined;
        for (size_t i = 1; i < names.size(); ++i) joined += (i > 1 ? " " : "") + names[i];
        ck.events.push_back(ObjEvent{ck.faces.size(), kObjGroup, joined});
    } else if (c[0] == 'o' && obj_space(c[1])) {
        ck.events.push_back(ObjEvent{ck.faces.size(), kObjObject, std::string(c + 2)});
    } else if (std::strncmp(c, "usemtl", 6) == 0 && obj_space(c[6])) {
        ck.events.push_back(ObjEvent{ck.faces.size(), kObjUsemtl, std::string(c + 7)});
    } else if (std::strncmp(c, "mtllib", 6) == 0 && obj_space(c[6])) {
        ck.events.push_back(ObjEvent{ck.faces.size(), kObjMtllib, std::string(c + 7)});
    }
}

// Lines of [b, e) end at '\n', '\r' or "\r\n" (tinyobj's safeGetline).
template <typename F>
void for_each_line(const char* b, const char* e, F&& f) {
    std::string line;
    while (b < e) {
        const char* q = b;
        while (q < e && *q != '\n' && *q != '\r') ++q;
        line.assign(b, q);  // NUL-terminated copy
        f(line);
        if (q < e && *q == '\r' && q + 1 < e && q[1] == '\n') ++q;
        b = q < e ? q + 1 : e;
    }
}

// pnpoly (tiny_obj_loader.h:1316-1327): crossings of a ray from (tx, ty).
bool obj_pnpoly3(const float* vx, const float* vy, float tx, float ty) {
    bool c = false;
    for (int i = 0, j = 2; i < 3; j = i++)
        if (((vy[i] > ty) != (vy[j] > ty)) && (tx < (vx[j] - vx[i]) * (ty - vy[i]) / (vy[j] - vy[i]) + vx[i])) c = !c;
    return c;
}

// tinyobj's ear clipping of one face (exportGroupsToShape's triangulate branch,
// tiny_obj_loader.h:1358-1560), FP32.  vi: the face's resolved vertex indices;
// V / nv: vertex coordinates and the vertex count the clipping may read.
// Appends triangles as positions within the face.
void obj_ear_clip(const int32_t* vi, uint32_t np, const float* V, size_t nv,
                  std::vector<std::array<uint32_t, 3>>& out) {
    auto ok = [&](int32_t v) { return v >= 0 && (size_t)v < nv; };
    size_t ax0 = 1, ax1 = 2;
    for (uint32_t k = 0; k < np; ++k) {
        int32_t a = vi[k % np], b = vi[(k + 1) % np], c = vi[(k + 2) % np];
        if (!ok(a) || !ok(b) || !ok(c)) continue;
        const float* p0 = V + 3 * (size_t)a;
        const float* p1 = V + 3 * (size_t)b;
        const float* p2 = V + 3 * (size_t)c;
        float e0x = p1[0] - p0[0], e0y = p1[1] - p0[1], e0z = p1[2] - p0[2];
        float e1x = p2[0] - p1[0], e1y = p2[1] - p1[1], e1z = p2[2] - p1[2];
        float cx = std::fabs(e0y * e1z - e0z * e1y);
        float cy = std::fabs(e0z * e1x - e0x * e1z);
        float cz = std::fabs(e0x * e1y - e0y * e1x);
        const float eps = std::numeric_limits<float>::epsilon();
        if (cx > eps || cy > eps || cz > eps) {
            if (!(cx > cy && cx > cz)) {
                ax0 = 0;
                if (cz > cx && cz > cy) ax1 = 1;
            }
            break;
        }
    }
    float area = 0.f;
    for (uint32_t k = 0; k < np; ++k) {
        int32_t a = vi[k % np], b = vi[(k + 1) % np];
        if (!ok(a) || !ok(b)) continue;
        float x0 = V[3 * (size_t)a + ax0], y0 = V[3 * (size_t)a + ax1];
        float x1 = V[3 * (size_t)b + ax0], y1 = V[3 * (size_t)b + ax1];
        area += (x0 * y1 - y0 * x1) * 0.5f;
    }
    std::vector<uint32_t> rem(np);
    for (uint32_t k = 0; k < np; ++k) rem[k] = k;
    size_t guess = 0, iters = np, prev = np;
    while (rem.size() > 3 && iters > 0) {
        const size_t n = rem.size();
        if (guess >= n) guess -= n;
        if (prev != n) {
            prev = n;
            iters = n;
        } else {
            --iters;
        }
        uint32_t ind[3];
        float vx[3], vy[3];
        for (int k = 0; k < 3; ++k) {
            ind[k] = rem[(guess + k) % n];
            int32_t v = vi[ind[k]];
            vx[k] = ok(v) ? V[3 * (size_t)v + ax0] : 0.f;
            vy[k] = ok(v) ? V[3 * (size_t)v + ax1] : 0.f;
        }
        float e0x = vx[1] - vx[0], e0y = vy[1] - vy[0], e1x = vx[2] - vx[1], e1y = vy[2] - vy[1];
        float cross = e0x * e1y - e0y * e1x;
        if (cross * area < 0.f) {
            guess += 1;
            continue;
        }
        bool overlap = false;
        for (size_t other = 3; other < n; ++other) {
            int32_t v = vi[rem[(guess + other) % n]];
            if (!ok(v)) continue;
            if (obj_pnpoly3(vx, vy, V[3 * (size_t)v + ax0], V[3 * (size_t)v + ax1])) {
                overlap = true;
                break;
            }
        }
        if (overlap) {
            guess += 1;
            continue;
        }
        out.push_back({ind[0], ind[1], ind[2]});
        rem.erase(rem.begin() + (long)((guess + 1) % n));
    }
    if (rem.size() == 3) out.push_back({rem[0], rem[1], rem[2]});
}

// The 'newmtl' names of an MTL file in order (LoadMtl, tiny_obj_loader.h:1638-2012):
// trailing blanks trimmed, the last material flushed even without a name.
bool obj_read_mtl_names(const std::string& path, std::vector<std::string>& names) {
    std::ifstream in(path.c_str(), std::ios::binary);
    if (!in) return false;
    std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    std::string cur;
    for_each_line(text.data(), text.data() + text.size(), [&](std::string& line) {
        size_t last = line.find_last_not_of(" \t");
        line.erase(last == std::string::npos ? 0 : last + 1);
        const char* t = line.c_str();
        t += std::strspn(t, " \t");
        if (t[0] == '\0' || t[0] == '#') return;
        if (std::strncmp(t, "newmtl", 6) == 0 && obj_space(t[6])) {
            if (!cur.empty()) names.push_back(cur);
            cur = std::string(t + 7);
        }
    });
    names.push_back(cur);
    return true;
}

int host_threads() {
    unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, h ? h : 1u));
}

// Run f(i) for i in [0, n) on the calling thread plus helper threads, at most
// host_threads() helpers across all concurrent calls in the process (scenes
// may be loaded from several threads at once).
std::atomic<int> g_helpers{0};

template <typename F>
void parallel_for(size_t n, F&& f) {
    int want = (int)std::min<size_t>((size_t)host_threads(), n) - 1;
    int got = 0;
    while (want > 0) {
        int cur = g_helpers.load();
        int take = std::min(want, host_threads() - cur);
        if (take <= 0) break;
        if (g_helpers.compare_exchange_weak(cur, cur + take)) {
            got = take;
            break;
        }
    }
    std::atomic<size_t> next{0};
    std::exception_ptr err;
    std::mutex err_mu;
    auto work = [&]() {
        try {
            for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
        } catch (...) {
            std::lock_guard<std::mutex> lk(err_mu);
            if (!err) err = std::current_exception();
            next.store(n);
        }
    };
    std::vector<std::thread> pool;
    pool.reserve(got);
    try {
        for (int t = 0; t < got; ++t) pool.emplace_back(work);
    } catch (const std::system_error&) {  // fewer helpers if threads cannot be created
        g_helpers.fetch_sub(got - (int)pool.size());
        got = (int)pool.size();
    }
    work();
    for (auto& th : pool) th.join();
    g_helpers.fetch_sub(got);
    if (err) std::rethrow_exception(err);
}

}  // namespace

// Returns the meshes of `path` (empty when the file does not parse as LoadObj requires).
std::vector<LoadedMesh> load_obj(const std::string& path, const Mat4& xf, const Mat4& xf_inv) {
    const std::string text = read_file(path);
    // Line-aligned chunks of about 1 MiB.
    std::vector<size_t> cuts{0};
    const size_t target = (size_t)1 << 20;
    while (cuts.back() < text.size()) {
        size_t c = std::min(text.size(), cuts.back() + target);
        while (c < text.size() && text[c - 1] != '\n' && text[c - 1] != '\r') ++c;
        cuts.push_back(c);
    }
    const size_t nchunks = cuts.size() - 1;
    std::vector<ObjChunk> chunks(nchunks);
    parallel_for(nchunks, [&](size_t i) {
        for_each_line(text.data() + cuts[i], text.data() + cuts[i + 1],
                      [&](const std::string& line) { parse_obj_line(line.c_str(), chunks[i]); });
    });

    // Global element arrays and the chunks' offsets into them.
    std::vector<size_t> v0(nchunks + 1, 0), vt0(nchunks + 1, 0), vn0(nchunks + 1, 0), f0(nchunks + 1, 0);
    for (size_t i = 0; i < nchunks; ++i) {
        v0[i + 1] = v0[i] + chunks[i].V.size() / 3;
        vt0[i + 1] = vt0[i] + chunks[i].VT.size() / 2;
        vn0[i + 1] = vn0[i] + chunks[i].VN.size() / 3;
        f0[i + 1] = f0[i] + chunks[i].faces.size();
    }
    std::vector<float> V(3 * v0[nchunks]), VT(2 * vt0[nchunks]), VN(3 * vn0[nchunks]);
    std::atomic<bool> unreadable{false};
    parallel_for(nchunks, [&](size_t i) {
        ObjChunk& ck = chunks[i];
        std::copy(ck.V.begin(), ck.V.end(), V.begin() + 3 * v0[i]);
        std::copy(ck.VT.begin(), ck.VT.end(), VT.begin() + 2 * vt0[i]);
        std::copy(ck.VN.begin(), ck.VN.end(), VN.begin() + 3 * vn0[i]);
        // fixIndex against the counts at each face line; a 0 fails LoadObj.
        for (const ObjFace& f : ck.faces)
            for (uint32_t k = 0; k < f.count; ++k) {
                ObjCorner& c = ck.corners[f.first + k];
                auto fix = [&](int32_t& x, size_t n) {
                    if (x == kObjAbsent) { x = -1; return; }
                    if (x == 0) { unreadable = true; return; }
                    x = x > 0 ? x - 1 : (int32_t)((int64_t)n + x);
                };
                fix(c.v, v0[i] + (size_t)f.nv);
                fix(c.vt, vt0[i] + (size_t)f.nvt);
                fix(c.vn, vn0[i] + (size_t)f.nvn);
            }
    });
    if (unreadable) return {};
    // Ear clipping per face.
    parallel_for(nchunks, [&](size_t i) {
        ObjChunk& ck = chunks[i];
        ck.tri_first.resize(ck.faces.size());
        ck.tri_count.resize(ck.faces.size());
        std::vector<int32_t> vi;
        for (size_t fi = 0; fi < ck.faces.size(); ++fi) {
            const ObjFace& f = ck.faces[fi];
            ck.tri_first[fi] = (uint32_t)ck.tris.size();
            if (f.count >= 3) {
                vi.resize(f.count);
                for (uint32_t k = 0; k < f.count; ++k) vi[k] = ck.corners[f.first + k].v;
                obj_ear_clip(vi.data(), f.count, V.data(), v0[i] + (size_t)f.nv, ck.tris);
            }
            ck.tri_count[fi] = (uint32_t)ck.tris.size() - ck.tri_first[fi];
        }
    });

    // Per chunk, the triangles of faces [a, b) are tris[tri_prefix[a] .. tri_prefix[b]).
    std::vector<std::vector<uint32_t>> tri_prefix(nchunks), tri_face(nchunks);
    parallel_for(nchunks, [&](size_t i) {
        const ObjChunk& ck = chunks[i];
        tri_prefix[i].resize(ck.faces.size() + 1);
        tri_prefix[i][0] = 0;
        for (size_t fi = 0; fi < ck.faces.size(); ++fi) tri_prefix[i][fi + 1] = tri_prefix[i][fi] + ck.tri_count[fi];
        tri_face[i].resize(ck.tris.size());
        for (size_t fi = 0; fi < ck.faces.size(); ++fi)
            for (uint32_t t = 0; t < ck.tri_count[fi]; ++t) tri_face[i][ck.tri_first[fi] + t] = (uint32_t)fi;
    });

    // LoadObj's statements in file order; faces move between shapes in runs of
    // consecutive faces of one chunk.
    struct ObjRun {
        uint32_t chunk, a, b;  // faces [a, b) of the chunk
        int material;
    };
    struct ObjShape {
        std::string name;
        std::vector<ObjRun> runs;
        bool has_tris = false;
    };
    std::vector<ObjShape> shapes;
    ObjShape shape;
    std::string name;
    int material = -1;
    std::vector<ObjRun> group;  // faces since the last export
    bool group_prims = false;   // a line or point statement since the last group was cleared
    std::map<std::string, int> mat_id;
    std::vector<std::string> mat_names;
    const std::string base = dir_of(path);
    auto export_group = [&]() {
        if (group.empty() && !group_prims) return false;
        shape.name = name;
        for (ObjRun r : group) {
            r.material = material;
            shape.runs.push_back(r);
            if (tri_prefix[r.chunk][r.b] > tri_prefix[r.chunk][r.a]) shape.has_tris = true;
        }
        return true;
    };
    for (size_t i = 0; i < nchunks; ++i) {
        uint32_t f = 0;
        auto take_until = [&](uint32_t stop) {
            if (stop > f) group.push_back(ObjRun{(uint32_t)i, f, stop, -1});
            f = stop;
        };
        for (const ObjEvent& ev : chunks[i].events) {
            take_until((uint32_t)ev.face);
            switch (ev.kind) {
                case kObjPrimitive: group_prims = true; break;
                case kObjUsemtl: {
                    auto it = mat_id.find(ev.text);
                    int id = it == mat_id.end() ? -1 : it->second;
                    if (id != material) {
                        export_group();
                        group.clear();
                        material = id;
                    }
                    break;
                }
                case kObjMtllib: {
                    std::vector<std::string> files;
                    std::stringstream ss(ev.text);
                    for (std::string item; std::getline(ss, item, ' ');) files.push_back(item);
                    for (const std::string& fn : files) {
                        std::vector<std::string> names;
                        if (!obj_read_mtl_names(base + fn, names)) continue;
                        for (const std::string& n : names) {
                            mat_id.emplace(n, (int)mat_names.size());
                            mat_names.push_back(n);
                        }
                        break;
                    }
                    break;
                }
                case kObjGroup:
                    export_group();
                    if (shape.has_tris) shapes.push_back(shape);
                    shape = ObjShape();
                    group.clear();
                    group_prims = false;
                    name = ev.text;
                    break;
                case kObjObject:
                    if (export_group()) shapes.push_back(shape);
                    group.clear();
                    group_prims = false;
                    shape = ObjShape();
                    name = ev.text;
                    break;
            }
        }
        take_until((uint32_t)chunks[i].faces.size());
    }
    if (export_group() || shape.has_tris) shapes.push_back(shape);

    // Meshes per (shape, material) in order of first appearance; each run's
    // triangles occupy consecutive slots of its mesh.
    struct Segment {
        uint32_t chunk, t0, t1, mesh, slot;  // chunk triangles [t0, t1) -> mesh slots from `slot`
    };
    std::vector<Segment> segs;
    std::vector<LoadedMesh> meshes;
    std::map<std::pair<int, int>, int> lookup;
    const std::string file_stem = stem_of(path);
    for (size_t si = 0; si < shapes.size(); ++si)
        for (const ObjRun& r : shapes[si].runs) {
            const uint32_t t0 = tri_prefix[r.chunk][r.a], t1 = tri_prefix[r.chunk][r.b];
            if (t1 == t0) continue;
            auto key = std::make_pair((int)si, r.material);
            auto it = lookup.find(key);
            if (it == lookup.end()) {
                LoadedMesh m;
                const std::string mname =
                    r.material >= 0 && r.material < (int)mat_names.size() ? mat_names[r.material] : std::string();
                std::string sn = shapes[si].name.empty() ? file_stem + "_shape_" + std::to_string(si) : shapes[si].name;
                m.name = mname.empty() ? sn : sn + "::" + mname;
                m.material = mname;
                it = lookup.emplace(key, (int)meshes.size()).first;
                meshes.push_back(std::move(m));
            }
            // ck.tris of a chunk are in face order, so a run's triangles are contiguous
            const uint32_t first = chunks[r.chunk].tri_first.empty() ? 0 : chunks[r.chunk].tri_first[r.a];
            LoadedMesh& m = meshes[it->second];
            const uint32_t block = 4096;
            for (uint32_t t = 0; t < t1 - t0; t += block)
                segs.push_back(Segment{r.chunk, first + t, first + std::min(t1 - t0, t + block), (uint32_t)it->second,
                                       (uint32_t)m.count + t});
            m.count += (int)(t1 - t0);
        }
    for (LoadedMesh& m : meshes) {
        m.pos.resize(9 * (size_t)m.count);
        m.nrm.resize(9 * (size_t)m.count);
        m.uv.resize(6 * (size_t)m.count);
    }

    // The triangles, with BuildVertex's arithmetic (objloader.cpp:10-50, 186-202).
    std::atomic<bool> out_of_range{false};
    parallel_for(segs.size(), [&](size_t sgi) {
        const Segment& sg = segs[sgi];
        const ObjChunk& ck = chunks[sg.chunk];
        LoadedMesh& m = meshes[sg.mesh];
        for (uint32_t tri = sg.t0; tri < sg.t1; ++tri) {
            const ObjFace& f = ck.faces[tri_face[sg.chunk][tri]];
            vec3 p[3], n[3];
            float u[3], v[3];
            bool hn[3];
            for (int k = 0; k < 3; ++k) {
                const ObjCorner& c = ck.corners[f.first + ck.tris[tri][k]];
                if ((c.v >= 0 && 3 * (size_t)c.v + 2 >= V.size()) || (c.vn >= 0 && 3 * (size_t)c.vn + 2 >= VN.size()) ||
                    (c.vt >= 0 && 2 * (size_t)c.vt + 1 >= VT.size())) {
                    out_of_range = true;
                    return;
                }
                p[k] = c.v >= 0 ? xform_point(xf, v3(V[3 * (size_t)c.v], V[3 * (size_t)c.v + 1], V[3 * (size_t)c.v + 2]))
                                : v3(0.f, 0.f, 0.f);
                hn[k] = c.vn >= 0;
                n[k] = hn[k] ? xform_normal(xf_inv, v3(VN[3 * (size_t)c.vn], VN[3 * (size_t)c.vn + 1], VN[3 * (size_t)c.vn + 2]))
                             : v3(0.f, 0.f, 0.f);
                u[k] = c.vt >= 0 ? VT[2 * (size_t)c.vt] : 0.f;
                v[k] = c.vt >= 0 ? VT[2 * (size_t)c.vt + 1] : 0.f;
            }
            if (!(hn[0] && hn[1] && hn[2])) {
                vec3 fn = normalize(cross(p[1] - p[0], p[2] - p[0]));
                n[0] = n[1] = n[2] = fn;
            }
            const size_t o = sg.slot + (tri - sg.t0);
            for (int k = 0; k < 3; ++k) {
                m.pos[9 * o + 3 * k] = p[k].x; m.pos[9 * o + 3 * k + 1] = p[k].y; m.pos[9 * o + 3 * k + 2] = p[k].z;
                m.nrm[9 * o + 3 * k] = n[k].x; m.nrm[9 * o + 3 * k + 1] = n[k].y; m.nrm[9 * o + 3 * k + 2] = n[k].z;
                m.uv[6 * o + 2 * k] = u[k]; m.uv[6 * o + 2 * k + 1] = v[k];
            }
        }
    });
    if (out_of_range) throw std::runtime_error("OBJ face references an element past the end of " + path);
    return meshes;
}

// ---------------------------------------------------------------- env ----
// Radiance RGBE (.hdr) as the reference reads it: HDRTexture::Load calls
// stbi_loadf(path, &w, &h, &c, 4) (Src/Core/texture.cu:17-25), i.e. the
// vendored stb_image 2.30's Radiance reader (ThirdParty/stb_image/stb_image.h:
// 7083-7288).  Its rules, kept here:
//  - the file starts with "#?RADIANCE\n" or "#?RGBE\n"; header lines (at most
//    1023 characters kept, the rest of a longer line skipped) run to an empty
//    line and one of them must be "FORMAT=32-bit_rle_rgbe";
//  - the resolution line is "-Y <h>" then spaces then "+X <w>" (strtol numbers);
//    sizes above 2^24, negative sizes and w*h*16 bytes above INT_MAX are refused;
//  - widths below 8 or from 32768 are stored flat (4 bytes per pixel);
//    otherwise each scanline opens with 2, 2, w >> 8, w & 255 and holds the four
//    channels one after another as runs (count > 128: count - 128 copies of the
//    next byte) and literal dumps (count bytes); a zero count, a count past the
//    end of the row or another length is an error;
//  - a scanline whose first bytes are not 2, 2 and a length byte below 128 is
//    not run-length encoded: those four bytes are pixel 0 of ROW 0, and the rest
//    of the file is read flat from row 0, pixel 1 on, whatever row it occurred
//    in (stb's goto into its flat loop);
//  - a texel is byte * 2^(e - 136) per channel when e != 0, else 0 (exact in
//    float32: a multiple of 2^-149 below 2^128).
// Deviations, only where stb's result is unspecified or empty: data that ends
// early is an error (stb reads zeros or leaves texels unset), and so is a zero
// width or height (stb returns an empty image the reference cannot sample).
// Row 0 is the first scanline (the orientation stb_image returns).
void parse_rgbe(const std::string& data, const std::string& path, EnvMap& env) {
    const uint8_t* d = (const uint8_t*)data.data();
    const size_t size = data.size();
    size_t p = 0;
    auto fail = [&](const char* what) { throw std::runtime_error(std::string(what) + " in RGBE file " + path); };
    auto get8 = [&]() -> uint8_t {
        if (p >= size) fail("truncated data");
        return d[p++];
    };
    // stbi__hdr_gettoken: up to '\n' or the end of the file, 1023 characters kept.
    auto token = [&]() {
        std::string t;
        if (p >= size) return t;
        char c = (char)d[p++];
        while (p < size && c != '\n') {
            t.push_back(c);
            if (t.size() == 1023) {
                while (p < size && d[p++] != '\n') {
                }
                break;
            }
            c = (char)d[p++];
        }
        return t;
    };
    auto starts = [&](const char* sig) { return data.compare(0, std::strlen(sig), sig) == 0; };
    if (!starts("#?RADIANCE\n") && !starts("#?RGBE\n")) fail("missing #?RADIANCE or #?RGBE signature");
    token();
    bool format_ok = false;
    for (;;) {
        std::string t = token();
        if (t.empty()) break;
        if (t == "FORMAT=32-bit_rle_rgbe") format_ok = true;
    }
    if (!format_ok) fail("unsupported format (no FORMAT=32-bit_rle_rgbe line)");
    std::string res = token();
    if (res.compare(0, 3, "-Y ") != 0) fail("unsupported data layout");
    const char* q = res.c_str() + 3;
    char* after = nullptr;
    long h = std::strtol(q, &after, 10);
    q = after;
    while (*q == ' ') ++q;
    if (std::strncmp(q, "+X ", 3) != 0) fail("unsupported data layout");
    long w = std::strtol(q + 3, nullptr, 10);
    if (h > (1L << 24) || w > (1L << 24) || h < 0 || w < 0 || (w > 0 && h > INT32_MAX / 16 / w))
        fail("image size out of range");
    if (w == 0 || h == 0) fail("zero-sized image");
    const int W = (int)w, H = (int)h;
    env.width = W;
    env.height = H;
    env.rgb.assign((size_t)W * H * 3, 0.f);
    auto put = [&](size_t texel, const uint8_t* e4) {
        float* o = &env.rgb[texel * 3];
        if (e4[3] != 0) {
            float f = std::ldexp(1.0f, (int)e4[3] - (128 + 8));
            o[0] = (float)e4[0] * f;
            o[1] = (float)e4[1] * f;
            o[2] = (float)e4[2] * f;
        } else {
            o[0] = o[1] = o[2] = 0.f;
        }
    };
    auto flat_from = [&](size_t texel) {
        const size_t n = (size_t)W * H;
        if (size - p < (n - texel) * 4) fail("truncated data");
        for (; texel < n; ++texel, p += 4) put(texel, d + p);
    };
    if (W < 8 || W >= 32768) {
        flat_from(0);
        return;
    }
    std::vector<uint8_t> scan((size_t)W * 4);
    for (int y = 0; y < H; ++y) {
        uint8_t c1 = get8(), c2 = get8(), len = get8();
        if (c1 != 2 || c2 != 2 || (len & 0x80)) {
            uint8_t first[4] = {c1, c2, len, get8()};
            put(0, first);
            flat_from(1);
            return;
        }
        int n = (len << 8) | get8();
        if (n != W) fail("invalid decoded scanline length");
        for (int c = 0; c < 4; ++c) {
            int x = 0;
            while (x < W) {
                int count = get8();
                if (count > 128) {
                    uint8_t value = get8();
                    count -= 128;
                    if (count == 0 || count > W - x) fail("bad run-length data");
                    for (int k = 0; k < count; ++k) scan[(size_t)(x++) * 4 + c] = value;
                } else {
                    if (count == 0 || count > W - x) fail("bad run-length data");
                    for (int k = 0; k < count; ++k) scan[(size_t)(x++) * 4 + c] = get8();
                }
            }
        }
        for (int x = 0; x < W; ++x) put((size_t)y * W + x, &scan[(size_t)x * 4]);
    }
}

// Environment maps: Radiance RGBE (.hdr, parse_rgbe) or portable float map
// (.pfm, this project's lossless addition).  Stored with row 0 = top.
void load_env(const std::string& path, EnvMap& env) {
    std::string data = read_file(path);
    if (data.size() >= 2 && data[0] == 'P' && (data[1] == 'F' || data[1] == 'f')) {
        bool color = data[1] == 'F';
        size_t p = 2;
        auto next_tok = [&]() {
            while (p < data.size() && std::isspace((unsigned char)data[p])) ++p;
            size_t b = p;
            while (p < data.size() && !std::isspace((unsigned char)data[p])) ++p;
            return data.substr(b, p - b);
        };
        int w = std::atoi(next_tok().c_str());
        int h = std::atoi(next_tok().c_str());
        double scale = std::atof(next_tok().c_str());
        ++p;  // single whitespace after the scale
        bool little = scale < 0;
        int ch = color ? 3 : 1;
        size_t need = (size_t)w * h * ch * 4;
        if (w <= 0 || h <= 0 || data.size() < p + need) throw std::runtime_error("truncated PFM: " + path);
        env.width = w;
        env.height = h;
        env.rgb.assign((size_t)w * h * 3, 0.f);
        for (int y = 0; y < h; ++y) {
            int row = h - 1 - y;  // PFM stores the bottom row first
            for (int x = 0; x < w; ++x)
                for (int c = 0; c < 3; ++c) {
                    int sc = ch == 3 ? c : 0;
                    uint32_t bits;
                    std::memcpy(&bits, data.data() + p + (((size_t)y * w + x) * ch + sc) * 4, 4);
                    if (!little) bits = __builtin_bswap32(bits);
                    float f;
                    std::memcpy(&f, &bits, 4);
                    env.rgb[((size_t)row * w + x) * 3 + c] = f;
                }
        }
        return;
    }
    parse_rgbe(data, path, env);
}

vec3 vec_of(const Json& a, const char* what) {
    if (!a.is_array() || a.size() < 3) throw JsonError(std::string("expected a 3-vector for '") + what + "'");
    return v3(a[0].as_float(what), a[1].as_float(what), a[2].as_float(what));
}

}  // namespace

void finalize_triangles(HostScene& sc) {
    const size_t n = (size_t)sc.num_tris();
    sc.gnormal.resize(3 * n);
    sc.area.resize(n);
    // Per triangle independent, so blocks on several threads give the same bits.
    const size_t block = 1 << 15;
    parallel_for((n + block - 1) / block, [&](size_t bi) {
        for (size_t t = bi * block; t < std::min(n, (bi + 1) * block); ++t) {
            const float* p = &sc.pos[9 * t];
            vec3 a = v3(p[0], p[1], p[2]), b = v3(p[3], p[4], p[5]), c = v3(p[6], p[7], p[8]);
            vec3 nn = cross(b - a, c - a);
            sc.area[t] = 0.5f * length(nn);
            nn = normalize(nn);
            sc.gnormal[3 * t] = nn.x;
            sc.gnormal[3 * t + 1] = nn.y;
            sc.gnormal[3 * t + 2] = nn.z;
        }
    });
    // Mesh areas are sequential FP32 sums in scene order (reference: Src/Core/scene.cpp:515-519).
    for (auto& m : sc.meshes) {
        m.area = 0.f;
        for (int t = m.start; t < m.start + m.count; ++t) m.area += sc.area[t];
    }
}

void set_scene_geometry(HostScene& sc, const float* pos, const float* nrm) {
    const size_t n9 = 9 * (size_t)sc.num_tris();
    for (size_t i = 0; i < n9; ++i)
        if (!std::isfinite(pos[i])) throw std::invalid_argument("non-finite vertex coordinate");
    std::memcpy(sc.pos.data(), pos, n9 * sizeof(float));
    if (nrm) std::memcpy(sc.nrm.data(), nrm, n9 * sizeof(float));
    finalize_triangles(sc);
    for (auto& l : sc.lights)
        if (l.type == kAreaLight && l.mesh >= 0) l.area = sc.meshes[l.mesh].area;
}

namespace {

// Squared distance from point p to triangle (a, b, c), double precision
// (closest point by Voronoi regions: Ericson, Real-Time Collision Detection, 5.1.5).
double dist2_point_triangle(const double* p, const double* a, const double* b, const double* c) {
    auto sub = [](const double* x, const double* y, double* o) { o[0] = x[0] - y[0]; o[1] = x[1] - y[1]; o[2] = x[2] - y[2]; };
    auto dot = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    double ab[3], ac[3], ap[3], bp[3], cp[3], q[3];
    sub(b, a, ab); sub(c, a, ac); sub(p, a, ap);
    auto at = [&](const double* o, double u, const double* e1, double v, const double* e2) {
        for (int k = 0; k < 3; ++k) q[k] = o[k] + u * e1[k] + v * e2[k];
    };
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) at(a, 0, ab, 0, ac);
    else {
        sub(p, b, bp);
        double d3 = dot(ab, bp), d4 = dot(ac, bp);
        if (d3 >= 0 && d4 <= d3) at(b, 0, ab, 0, ac);
        else {
            double vc = d1 * d4 - d3 * d2;
            if (vc <= 0 && d1 >= 0 && d3 <= 0) at(a, d1 / (d1 - d3), ab, 0, ac);
            else {
                sub(p, c, cp);
                double d5 = dot(ab, cp), d6 = dot(ac, cp);
                if (d6 >= 0 && d5 <= d6) at(c, 0, ab, 0, ac);
                else {
                    double vb = d5 * d2 - d1 * d6;
                    if (vb <= 0 && d2 >= 0 && d6 <= 0) at(a, 0, ab, d2 / (d2 - d6), ac);
                    else {
                        double va = d3 * d6 - d5 * d4;
                        if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
                            double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
                            double bc[3];
                            sub(c, b, bc);
                            at(b, w, bc, 0, ac);
                        } else {
                            double den = 1.0 / (va + vb + vc);
                            at(a, vb * den, ab, vc * den, ac);
                        }
                    }
                }
            }
        }
    }
    double dq[3];
    sub(p, q, dq);
    return dot(dq, dq);
}

// Distance from p to the nearest triangle of a tree (branch and bound over its
// reference-layout nodes, nearer box first), given an upper bound of it.
double nearest_distance(const BvhBuild& b, const float* pos, const double* p, double bound) {
    double best = bound * bound;
    auto box_d2 = [&](const LinearNode& n) {
        double d2 = 0;
        for (int k = 0; k < 3; ++k) {
            double e = std::max({(double)n.lo[k] - p[k], 0.0, p[k] - (double)n.hi[k]});
            d2 += e * e;
        }
        return d2;
    };
    int stack[2 * 64 + 2];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        int i = stack[--sp];
        const LinearNode& n = b.nodes[i];
        if (box_d2(n) >= best) continue;
        if (n.tri_index >= 0) {
            for (int k = 0; k < n.second; ++k) {
                const float* t = pos + 9 * (size_t)b.order[(size_t)n.tri_index + k];
                double a[3] = {t[0], t[1], t[2]}, bb[3] = {t[3], t[4], t[5]}, c[3] = {t[6], t[7], t[8]};
                best = std::min(best, dist2_point_triangle(p, a, bb, c));
            }
        } else {
            int l = i + 1, r = n.second;
            if (box_d2(b.nodes[l]) < box_d2(b.nodes[r])) std::swap(l, r);
            stack[sp++] = l;
            stack[sp++] = r;
        }
    }
    return std::sqrt(best);
}

constexpr int kClearanceGrid = 24;

// Lower bounds of the distance to the mesh per cell: the cell centre's exact
// distance less half the cell diagonal and a rounding margin well above the
// FP32 error of a triangle test at the mesh's coordinates (a query of the
// walk can only report a hit at t when a triangle lies within about t of its
// origin).  Rounded down to float.
void clearance_grid(MediumTree& t, const float* pos) {
    const Box& bx = t.bvh.bounds;
    const int n = kClearanceGrid;
    t.grid_n = n;
    t.grid_lo = bx.lo;
    double ext[3] = {(double)bx.hi.x - bx.lo.x, (double)bx.hi.y - bx.lo.y, (double)bx.hi.z - bx.lo.z};
    double cell[3], mag = 0;
    float inv[3];
    for (int k = 0; k < 3; ++k) {
        cell[k] = std::max(ext[k], 1e-6) / n;
        inv[k] = (float)(1.0 / cell[k]);
    }
    for (float v : {bx.lo.x, bx.lo.y, bx.lo.z, bx.hi.x, bx.hi.y, bx.hi.z}) mag = std::max(mag, (double)std::fabs(v));
    t.grid_inv = v3(inv[0], inv[1], inv[2]);
    // A point may land one cell over from floor((p - lo) * inv) rounding: allow
    // the full cell diagonal, not half of it.
    const double reach = std::sqrt(cell[0] * cell[0] + cell[1] * cell[1] + cell[2] * cell[2]);
    const double margin = 1e-4 * (1.0 + mag);
    t.clearance.assign((size_t)n * n * n, 0.f);
    // Slabs of z on host threads; along x each cell starts from its
    // neighbour's distance plus one cell as the upper bound (exact minimum
    // either way, the bound only prunes).
    if (t.bvh.max_depth > 64) throw std::runtime_error("medium tree deeper than 64 levels");
    auto slab = [&](int z) {
        for (int y = 0; y < n; ++y) {
            double prev = std::numeric_limits<double>::infinity();
            for (int x = 0; x < n; ++x) {
                double c[3] = {bx.lo.x + (x + 0.5) * cell[0], bx.lo.y + (y + 0.5) * cell[1], bx.lo.z + (z + 0.5) * cell[2]};
                double dist = nearest_distance(t.bvh, pos, c, prev + cell[0]);
                prev = dist;
                double d = dist - reach - margin;
                float f = d > 0 ? (float)d : 0.f;
                if ((double)f > d) f = std::nextafter(f, 0.f);
                t.clearance[((size_t)z * n + y) * n + x] = d > 0 ? f : 0.f;
            }
        }
    };
    const int workers = std::max(1, std::min<int>(n, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int w = 1; w < workers; ++w)
        pool.emplace_back([&]() {
            for (int z; (z = next.fetch_add(1)) < n;) slab(z);
        });
    for (int z; (z = next.fetch_add(1)) < n;) slab(z);
    for (auto& th : pool) th.join();
}

}  // namespace

void build_media(HostScene& sc) {
    sc.media.clear();
    for (size_t m = 0; m < sc.meshes.size(); ++m) {
        const MeshRec& mr = sc.meshes[m];
        if (mr.count <= 0 || mr.material < 0) continue;
        const MaterialRec& mat = sc.materials[mr.material];
        if (mat.type != kSubsurface || mat.sss_method != kSssRandomWalk) continue;
        MediumTree t;
        t.mesh = (int)m;
        build_bvh(sc.pos.data() + 9 * (size_t)mr.start, mr.count, t.bvh);
        clearance_grid(t, sc.pos.data() + 9 * (size_t)mr.start);
        sc.media.push_back(std::move(t));
    }
}

void refit_media(HostScene& sc) {
    for (MediumTree& t : sc.media) {
        const MeshRec& mr = sc.meshes[t.mesh];
        refit_bvh(sc.pos.data() + 9 * (size_t)mr.start, mr.count, t.bvh);
        clearance_grid(t, sc.pos.data() + 9 * (size_t)mr.start);
    }
}

void compute_camera(CameraRec& c) {
    // Camera::ComputeCameraParam (reference: Src/Core/camera.h:58-84).
    float theta = deg2rad(c.theta), phi = deg2rad(c.phi);
    float x = (c.distance * sinf(theta)) * sinf(phi);
    float y = c.distance * cosf(theta);
    float z = (c.distance * sinf(theta)) * cosf(phi);
    c.position = c.target + v3(x, y, z);
    c.lookat = c.target;
    c.up = v3(0.f, 1.f, 0.f);
    c.w = -normalize(c.lookat - c.position);
    c.u = normalize(cross(normalize(c.up), c.w));
    c.v = normalize(cross(c.w, c.u));
    c.focal = length(c.lookat - c.position);
    float ratio = (1.f * (float)c.width) / (float)c.height;
    float h_ = (2.f * c.focal) * tanf(deg2rad(c.fovy / 2.f));
    float w_ = h_ * ratio;
    c.du = (w_ * c.u) / (float)c.width;
    c.dv = (h_ * c.v) / (float)c.height;
    c.left_down = (c.lookat - (float)(c.width / 2) * c.du) - (float)(c.height / 2) * c.dv;
}

void setup_camera(HostScene& sc) {
    CameraRec& c = sc.camera;
    Box b = sc.bvh.bounds;
    vec3 center = (b.lo + b.hi) * 0.5f;
    float diag = length(b.hi - b.lo);
    // Scene::InitCameraParam (reference: Src/Core/scene.cpp:646-667).
    c.target = center;
    c.distance = ((diag * 0.5f) / sinf(deg2rad(c.fovy) / 2.f)) * 0.6f;
    vec3 dir = normalize(c.direction);
    float th = acosf(dir.y / length(dir));
    float ph = atan2f(dir.x, dir.z);
    c.theta = rad2deg(th);
    c.phi = rad2deg(ph);
    compute_camera(c);
}

void camera_view_projection(const CameraRec& c, float near_plane, float far_plane, float vp[16]) {
    // LookAt and Perspective (reference: Src/Utils/mathutil.cu:163-194), then
    // projection * view as the G-buffer vertex shader composes them.
    vec3 f = normalize(c.lookat - c.position);
    vec3 s = normalize(cross(f, c.up));
    vec3 u = cross(s, f);
    Mat4 view = identity();
    view.m[0][0] = s.x; view.m[0][1] = s.y; view.m[0][2] = s.z;
    view.m[1][0] = u.x; view.m[1][1] = u.y; view.m[1][2] = u.z;
    view.m[2][0] = -f.x; view.m[2][1] = -f.y; view.m[2][2] = -f.z;
    view.m[0][3] = -dot(s, c.position);
    view.m[1][3] = -dot(u, c.position);
    view.m[2][3] = dot(f, c.position);
    float fy = 1.0f / tanf(deg2rad(c.fovy) / 2.0f);
    float aspect = (1.f * (float)c.width) / (float)c.height;
    Mat4 proj;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) proj.m[i][j] = 0.f;
    proj.m[0][0] = fy / aspect;
    proj.m[1][1] = fy;
    proj.m[2][2] = (far_plane + near_plane) / (near_plane - far_plane);
    proj.m[2][3] = ((2.f * far_plane) * near_plane) / (near_plane - far_plane);
    proj.m[3][2] = -1.0f;
    Mat4 m = mul(proj, view);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) vp[4 * i + j] = m.m[i][j];
}

void load_scene(const std::string& json_path, HostScene& sc, int bvh_builder) {
    sc = HostScene();
    sc.path = json_path;
    std::string base = dir_of(json_path);
    Json root = Json::parse(read_file(json_path));

    // ---- camera (reference: Src/Core/scene.cpp:187-252)
    const Json& cam = root["camera"];
    if (!cam.contains("type")) throw std::runtime_error("camera missing 'type' field");
    if (cam["type"].as_string("camera.type") != "pinhole")
        throw std::runtime_error("unsupported camera type " + cam["type"].as_string("camera.type"));
    if (cam.contains("resolution")) {
        sc.camera.width = cam["resolution"][0].as_int("resolution");
        sc.camera.height = cam["resolution"][1].as_int("resolution");
    }
    if (sc.camera.width <= 0 || sc.camera.height <= 0) throw std::runtime_error("resolution must be positive");
    if (cam.contains("fov")) sc.camera.fovy = cam["fov"].as_float("fov");
    if (cam.contains("max_bounces")) sc.max_bounces = cam["max_bounces"].as_int("max_bounces");
    if (cam.contains("rr")) sc.rr = cam["rr"].as_float("rr");
    if (cam.contains("spp")) sc.spp = cam["spp"].as_int("spp");
    if (cam.contains("direction")) sc.camera.direction = vec_of(cam["direction"], "direction");
    if (cam.contains("denoise")) sc.denoise = cam["denoise"].as_bool("denoise");
    // Additive keys (not in the reference schema).
    if (cam.contains("seed")) sc.seed = (uint32_t)cam["seed"].as_number("seed");
    if (cam.contains("rr_start_depth")) sc.rr_start_depth = cam["rr_start_depth"].as_int("rr_start_depth");

    // ---- textures (reference: Src/Core/scene.cpp:618-644, Texture::Load texture.cu:207-240):
    // every entry is decoded to RGBA8 at load; an entry without 'path' or
    // 'name' keeps its slot, empty.  A file that cannot be read or decoded is
    // an error (the reference prints a message and continues with no image).
    const Json& texj = root["textures"];
    if (!texj.is_array()) throw std::runtime_error("'textures' should be an array");
    std::map<std::string, int> tex_map;
    sc.textures.assign(texj.size(), LdrImage());
    for (size_t i = 0; i < texj.size(); ++i) {
        if (!texj[i].contains("path") || !texj[i].contains("name")) {
            sc.warnings.push_back("texture missing 'path' or 'name'");
            continue;
        }
        std::string tname = texj[i]["name"].as_string("texture.name");
        load_ldr_image(resolve_path(base, texj[i]["path"].as_string("texture.path")), sc.textures[i]);
        tex_map[tname] = (int)i;
    }

    // ---- materials (reference: Src/Core/scene.cpp:254-396)
    const Json& matj = root["materials"];
    if (!matj.is_array()) throw std::runtime_error("'materials' should be an array");
    std::unordered_map<std::string, int> mat_map;
    sc.materials.resize(matj.size());
    sc.material_names.resize(matj.size());
    for (size_t i = 0; i < matj.size(); ++i) {
        const Json& m = matj[i];
        if (!m.contains("name") || !m.contains("type")) {
            sc.warnings.push_back("material missing 'name' or 'type'");
            continue;
        }
        MaterialRec r;
        std::string name = m["name"].as_string("material.name");
        std::string type = m["type"].as_string("material.type");
        if (type == "diffuse") r.type = kDiffuse;
        else if (type == "light") r.type = kLight;
        else if (type == "pbr") r.type = kPbr;
        else if (type == "subsurface") r.type = kSubsurface;
        else if (type == "dielectric" || type == "glass") r.type = kDielectric;
        else {
            sc.warnings.push_back("unsupported material type " + type);
            continue;
        }
        auto set3 = [&](float* dst, const char* key, float d0, float d1, float d2) {
            if (m.contains(key)) {
                vec3 v = vec_of(m[key], key);
                dst[0] = v.x; dst[1] = v.y; dst[2] = v.z;
            } else {
                dst[0] = d0; dst[1] = d1; dst[2] = d2;
            }
        };
        if (r.type == kDiffuse) {
            set3(r.basecolor, "basecolor", 0.f, 0.f, 0.f);
        } else if (r.type == kLight) {
            set3(r.ke, "emission", 0.f, 0.f, 0.f);
        } else if (r.type == kPbr) {
            set3(r.basecolor, "basecolor", 0.f, 0.f, 0.f);
            r.metallic = m.contains("metallic") ? m["metallic"].as_float("metallic") : 0.0f;
            r.roughness = m.contains("roughness") ? m["roughness"].as_float("roughness") : 0.5f;
        } else if (r.type == kSubsurface) {
            set3(r.basecolor, "basecolor", 0.f, 0.f, 0.f);
            set3(r.scatter, "scatter_distance", 1.f, 1.f, 1.f);
            if (m.contains("ior")) r.ior = m["ior"].as_float("ior");
            if (m.contains("roughness")) r.roughness = m["roughness"].as_float("roughness");
            r.metallic = 0.0f;
            // Random-walk mode (no counterpart in the reference's schema).
            if (m.contains("subsurface_method")) {
                std::string sm = m["subsurface_method"].as_string("subsurface_method");
                if (sm == "random_walk") r.sss_method = kSssRandomWalk;
                else if (sm == "burley") r.sss_method = kSssBurley;
                else throw std::runtime_error("unsupported subsurface_method '" + sm + "' for material '" + name + "'");
            }
            set3(r.sigma_a, "sigma_a", 0.f, 0.f, 0.f);
            set3(r.sigma_s, "sigma_s", 1.f, 1.f, 1.f);
            for (int c = 0; c < 3; ++c)
                if (!(r.sigma_a[c] >= 0.f) || !(r.sigma_s[c] >= 0.f) || !std::isfinite(r.sigma_a[c] + r.sigma_s[c]) ||
                    (r.sss_method == kSssRandomWalk && !(r.sigma_a[c] + r.sigma_s[c] > 0.f)))
                    throw std::runtime_error("sigma_a and sigma_s of material '" + name +
                                             "' must be finite, >= 0 and sum to > 0 per channel");
        } else if (r.type == kDielectric) {
            set3(r.basecolor, "basecolor", 1.f, 1.f, 1.f);
            r.ior = m.contains("ior") ? m["ior"].as_float("ior") : 1.5f;
            r.transmission = m.contains("transmission") ? m["transmission"].as_float("transmission") : 1.0f;
            r.roughness = m.contains("roughness") ? m["roughness"].as_float("roughness") : 0.0f;
            r.metallic = 0.0f;
        }
        if (m.contains("diffuse_texture")) {
            std::string tn = m["diffuse_texture"].as_string("diffuse_texture");
            auto it = tex_map.find(tn);
            if (it != tex_map.end()) r.texture = it->second;
            else sc.warnings.push_back("texture '" + tn + "' not found for material '" + name + "'");
        }
        r.valid = 1;
        mat_map[name] = (int)i;
        sc.materials[i] = r;
        sc.material_names[i] = name;
    }

    // ---- objects (reference: Src/Core/scene.cpp:398-537)
    using Clock = std::chrono::steady_clock;
    auto ms_since = [](Clock::time_point t) {
        return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
    };
    auto t_load = Clock::now();
    const Json& objj = root["objects"];
    if (!objj.is_array()) throw std::runtime_error("'objects' should be an array");
    std::unordered_map<std::string, int> mesh_map;
    for (size_t oi = 0; oi < objj.size(); ++oi) {
        const Json& o = objj[oi];
        if (!o.contains("name") || !o.contains("path")) {
            sc.warnings.push_back("object missing 'name' or 'path'");
            continue;
        }
        std::string oname = o["name"].as_string("object.name");
        std::string opath = resolve_path(base, o["path"].as_string("object.path"));
        Mat4 tr = identity(), rot = identity(), scl = identity();
        if (o.contains("translate")) {
            vec3 t = vec_of(o["translate"], "translate");
            tr.m[0][3] = t.x; tr.m[1][3] = t.y; tr.m[2][3] = t.z;
        }
        if (o.contains("rotate")) {
            const Json& ro = o["rotate"];
            for (size_t k = 0; k < ro.size(); ++k) {
                const Json& r = ro[k];
                if (r.contains("rotateX")) rot = mul(rotate_axis(0, r["rotateX"].as_float("rotateX")), rot);
                else if (r.contains("rotateY")) rot = mul(rotate_axis(1, r["rotateY"].as_float("rotateY")), rot);
                else if (r.contains("rotateZ")) rot = mul(rotate_axis(2, r["rotateZ"].as_float("rotateZ")), rot);
            }
        }
        if (o.contains("scale")) {
            vec3 s = vec_of(o["scale"], "scale");
            scl.m[0][0] = s.x; scl.m[1][1] = s.y; scl.m[2][2] = s.z;
        }
        Mat4 xf = mul(mul(tr, rot), scl);
        Mat4 xf_inv = inverse(xf);
        std::vector<LoadedMesh> loaded = load_obj(opath, xf, xf_inv);
        if (loaded.empty()) {
            sc.warnings.push_back("object '" + oname + "' produced no meshes");
            continue;
        }
        bool expanded = loaded.size() > 1;
        size_t more = 0;
        for (const LoadedMesh& lm : loaded) more += (size_t)lm.count;
        sc.pos.reserve(sc.pos.size() + 9 * more);
        sc.nrm.reserve(sc.nrm.size() + 9 * more);
        sc.uv.reserve(sc.uv.size() + 6 * more);
        sc.mesh.reserve(sc.mesh.size() + more);
        for (LoadedMesh& lm : loaded) {
            if (lm.count == 0) continue;
            // ResolveMaterialName (reference: Src/Core/scene.cpp:12-36).
            std::string mname;
            if (o.contains("material")) mname = o["material"].as_string("material");
            else if (o.contains("material_overrides") && o["material_overrides"].is_object() && !lm.material.empty() &&
                     o["material_overrides"].contains(lm.material))
                mname = o["material_overrides"][lm.material].as_string("material_overrides");
            else if (!lm.material.empty() && mat_map.count(lm.material)) mname = lm.material;
            else if (o.contains("default_material")) mname = o["default_material"].as_string("default_material");
            if (mname.empty() || !mat_map.count(mname)) {
                sc.warnings.push_back("no valid material binding for mesh '" + lm.name + "' of object '" + oname + "'");
                continue;
            }
            MeshRec mr;
            mr.name = expanded ? oname + "::" + lm.name : oname;
            mr.material = mat_map[mname];
            mr.start = sc.num_tris();
            mr.count = lm.count;
            int mid = (int)sc.meshes.size();
            if (sc.pos.empty()) {  // the first mesh: take its arrays
                sc.pos.swap(lm.pos);
                sc.nrm.swap(lm.nrm);
                sc.uv.swap(lm.uv);
            } else {
                sc.pos.insert(sc.pos.end(), lm.pos.begin(), lm.pos.end());
                sc.nrm.insert(sc.nrm.end(), lm.nrm.begin(), lm.nrm.end());
                sc.uv.insert(sc.uv.end(), lm.uv.begin(), lm.uv.end());
            }
            sc.mesh.insert(sc.mesh.end(), (size_t)lm.count, mid);
            mesh_map[mr.name] = mid;
            sc.meshes.push_back(mr);
        }
    }
    sc.load_phase_ms[0] = ms_since(t_load);
    t_load = Clock::now();
    finalize_triangles(sc);
    sc.load_phase_ms[1] = ms_since(t_load);

    // ---- lights (reference: Src/Core/scene.cpp:539-616)
    const Json& lj = root["lights"];
    if (!lj.is_array()) throw std::runtime_error("'lights' should be an array");
    sc.lights.resize(lj.size());
    for (size_t i = 0; i < lj.size(); ++i) {
        const Json& l = lj[i];
        LightRec rec;
        if (!l.contains("type")) {
            sc.warnings.push_back("light missing 'type'");
            sc.lights[i] = rec;
            continue;
        }
        std::string type = l["type"].as_string("light.type");
        if (type == "area") {
            rec.type = kAreaLight;
            if (l.contains("ke")) {
                vec3 e = vec_of(l["ke"], "ke");
                rec.emission[0] = e.x; rec.emission[1] = e.y; rec.emission[2] = e.z;
            }
            if (l.contains("material")) {
                std::string mn = l["material"].as_string("light.material");
                auto it = mat_map.find(mn);
                if (it != mat_map.end()) {
                    const MaterialRec& mr = sc.materials[it->second];
                    rec.emission[0] = mr.ke[0]; rec.emission[1] = mr.ke[1]; rec.emission[2] = mr.ke[2];
                } else {
                    sc.warnings.push_back("material '" + mn + "' not found for light");
                }
            }
            std::string lm = l["lightMeshName"].as_string("lightMeshName");
            auto it = mesh_map.find(lm);
            if (it == mesh_map.end()) throw std::runtime_error("light mesh '" + lm + "' not found");
            rec.mesh = it->second;
            sc.meshes[rec.mesh].light = (int)i;
            rec.tri_count = sc.meshes[rec.mesh].count;
            rec.area = sc.meshes[rec.mesh].area;
        } else if (type == "env") {
            rec.type = kEnvLight;
            if (!l.contains("path")) throw std::runtime_error("env light missing 'path'");
            auto t_env = Clock::now();
            load_env(resolve_path(base, l["path"].as_string("light.path")), sc.env);
            sc.load_phase_ms[2] += ms_since(t_env);
            sc.env_light = (int)i;
        } else {
            throw std::runtime_error("unsupported light type " + type);
        }
        sc.lights[i] = rec;
    }
    // Light triangles are packed contiguously per area light.
    int cursor = 0;
    for (auto& l : sc.lights) {
        if (l.type == kAreaLight && l.mesh >= 0) {
            l.tri_start = cursor;
            cursor += l.tri_count;
        }
    }

    if (sc.num_tris() == 0) throw std::runtime_error("scene has no triangles");
    for (float v : sc.pos)
        if (!std::isfinite(v)) throw std::runtime_error("non-finite vertex coordinate");
    bool gpu = bvh_builder == kBvhGpu;
    if (bvh_builder == kBvhAuto && sc.num_tris() >= kGpuBuildMinTris) {
        int devices = 0;
        gpu = hipGetDeviceCount(&devices) == hipSuccess && devices > 0;
    }
    auto t0 = std::chrono::steady_clock::now();
    if (gpu) build_bvh_gpu(sc.pos.data(), sc.num_tris(), sc.bvh);
    else build_bvh(sc.pos.data(), sc.num_tris(), sc.bvh);
    sc.bvh_build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    sc.bvh_on_gpu = gpu;
    build_media(sc);
    setup_camera(sc);
}

}  // namespace rt
