// TEST INFRASTRUCTURE ONLY: OBJ import and object transforms computed with the
// reference's own vendored libraries, tinyobjloader (ThirdParty/tinyobj) and
// Eigen 3.4.0 (ThirdParty/Eigen3), compiled where they lie in the reference
// tree (`make -C oracle ref` -> oracle/_ref/libobjref.so).  The steps around
// the library calls restate the reference's code, which itself cannot be
// compiled here (its headers pull in the CUDA runtime):
//  - the object transform of Scene::ParseObjects (Src/Core/scene.cpp:405-448):
//    T = Translate, R = each rotation left-multiplied in order, S = Scale,
//    transform = T * R * S and transformInv = transform.inverse(), with the
//    matrix builders of Src/Utils/mathutil.cu:91-135 and AngleToRadian
//    (mathutil.cu:4-7, PI = 3.1415926 from Src/Core/global.h:7).  cos/sin of a
//    float are the float overloads (what CUDA's host headers give the
//    reference's .cu files);
//  - OBJLoader::LoadObject (Src/Utils/objloader.cpp:66-210), with the
//    reference's own path helpers (Src/Utils/stringutil.cpp): tinyobj::LoadObj
//    with triangulation, one mesh per (shape, material id), per-vertex
//    positions transform * (p, 1), normals
//    (transformInv.transpose() * (n, 1)).normalized(), and the face normal of
//    the transformed positions on faces that lack a normal index.
// Tests compare the library's importer and the oracle's Python parser with
// what this returns, bit for bit.
#define TINYOBJLOADER_IMPLEMENTATION
#include <tiny_obj_loader.h>

#include <Eigen/Dense>

#include "stringutil.h"  // the reference's Src/Utils/stringutil.{h,cpp}, compiled where they lie

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

typedef Eigen::Matrix4f Mat4f;
typedef Eigen::Vector4f Vec4f;
typedef Eigen::Vector3f Vec3f;
typedef Eigen::Vector2f Vec2f;

const double kPi = 3.1415926;  // Src/Core/global.h:7

float angle_to_radian(float angle) { return angle / 180 * kPi; }

Mat4f translate_m(const Vec3f& t) {
    Mat4f m = Mat4f::Identity();
    m(0, 3) = t(0);
    m(1, 3) = t(1);
    m(2, 3) = t(2);
    return m;
}
Mat4f scale_m(const Vec3f& s) {
    Mat4f m = Mat4f::Identity();
    m(0, 0) = s(0);
    m(1, 1) = s(1);
    m(2, 2) = s(2);
    return m;
}
Mat4f rotate_m(int axis, float angle) {
    float rad = angle_to_radian(angle);
    float c = std::cos(rad), s = std::sin(rad);
    Mat4f m = Mat4f::Identity();
    int i = axis == 0 ? 1 : 0, j = axis == 2 ? 1 : 2;
    m(i, i) = c;
    m(j, j) = c;
    if (axis == 1) {  // RotateY: (0,2) = sin, (2,0) = -sin
        m(i, j) = s;
        m(j, i) = -s;
    } else {  // RotateX / RotateZ: (i,j) = -sin, (j,i) = sin
        m(i, j) = -s;
        m(j, i) = s;
    }
    return m;
}

void build_transform(const float* t, const int* axes, const float* degs, int nrot, const float* s, Mat4f& xf,
                     Mat4f& inv) {
    Mat4f T = Mat4f::Identity(), R = Mat4f::Identity(), S = Mat4f::Identity();
    if (t) T = translate_m(Vec3f(t[0], t[1], t[2]));
    for (int k = 0; k < nrot; ++k) R = rotate_m(axes[k], degs[k]) * R;
    if (s) S = scale_m(Vec3f(s[0], s[1], s[2]));
    xf = T * R * S;
    inv = xf.inverse();
}

Vec4f to4(const Vec3f& v) { return Vec4f(v(0), v(1), v(2), 1.f); }
Vec3f to3(const Vec4f& v) { return Vec3f(v(0), v(1), v(2)); }

struct Loaded {
    std::vector<float> pos, nrm, uv;  // 9, 9, 6 per triangle
    std::vector<int32_t> mesh;        // mesh index per triangle, meshes in creation order
    std::vector<std::string> mesh_names, mesh_materials;
    std::vector<std::vector<float>> pos_m, nrm_m, uv_m;
};

}  // namespace

extern "C" {

// Row-major transform and inverse of an object (translate/scale may be NULL).
void ref_object_transform(const float* t, const int* axes, const float* degs, int nrot, const float* s,
                          float* xf16, float* inv16) {
    Mat4f xf, inv;
    build_transform(t, axes, degs, nrot, s, xf, inv);
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            xf16[4 * r + c] = xf(r, c);
            inv16[4 * r + c] = inv(r, c);
        }
}

// Eigen's Matrix4f::inverse of a row-major matrix (the SSE path on x86-64).
void ref_inverse4(const float* m16, float* inv16) {
    Mat4f m;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) m(r, c) = m16[4 * r + c];
    Mat4f inv = m.inverse();
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv16[4 * r + c] = inv(r, c);
}

// Points through transform * (p, 1) and normals through
// (transformInv.transpose() * (n, 1)).normalized(), as BuildVertex does.
void ref_xform_points(const float* xf16, const float* p, int n, float* out) {
    Mat4f m;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) m(r, c) = xf16[4 * r + c];
    for (int i = 0; i < n; ++i) {
        Vec3f q = to3(m * to4(Vec3f(p[3 * i], p[3 * i + 1], p[3 * i + 2])));
        out[3 * i] = q(0);
        out[3 * i + 1] = q(1);
        out[3 * i + 2] = q(2);
    }
}
void ref_xform_normals(const float* inv16, const float* nrm, int n, float* out) {
    Mat4f inv;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv(r, c) = inv16[4 * r + c];
    for (int i = 0; i < n; ++i) {
        Vec3f q = to3(inv.transpose() * to4(Vec3f(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]))).normalized();
        out[3 * i] = q(0);
        out[3 * i + 1] = q(1);
        out[3 * i + 2] = q(2);
    }
}

// OBJLoader::LoadObject of `path` under the object transform; returns a handle
// (NULL when tinyobj fails).  Read it with ref_obj_* below, free with ref_obj_free.
void* ref_obj_load(const char* path, const float* t, const int* axes, const float* degs, int nrot, const float* s) {
    Mat4f xf, inv;
    build_transform(t, axes, degs, nrot, s, xf, inv);
    tinyobj::attrib_t attrib;
    std::vector<tinyobj::shape_t> shapes;
    std::vector<tinyobj::material_t> materials;
    std::string warn, err, p(path);
    std::string base = GetDirectoryFromPath(p);
    if (!tinyobj::LoadObj(&attrib, &shapes, &materials, &warn, &err, path, base.c_str(), true)) return nullptr;
    Loaded* L = new Loaded;
    std::unordered_map<std::string, size_t> lookup;
    auto vertex = [&](const tinyobj::index_t& ix, Vec3f& pos, Vec3f& nrm, Vec2f& uv) {
        pos = Vec3f(0.f, 0.f, 0.f);
        nrm = Vec3f(0.f, 0.f, 0.f);
        uv = Vec2f(0.f, 0.f);
        if (ix.vertex_index >= 0) {
            pos = Vec3f(attrib.vertices[3 * ix.vertex_index], attrib.vertices[3 * ix.vertex_index + 1],
                        attrib.vertices[3 * ix.vertex_index + 2]);
            pos = to3(xf * to4(pos));
        }
        if (ix.normal_index >= 0) {
            Vec3f n(attrib.normals[3 * ix.normal_index], attrib.normals[3 * ix.normal_index + 1],
                    attrib.normals[3 * ix.normal_index + 2]);
            nrm = to3(inv.transpose() * to4(n)).normalized();
        }
        if (ix.texcoord_index >= 0)
            uv = Vec2f(attrib.texcoords[2 * ix.texcoord_index], attrib.texcoords[2 * ix.texcoord_index + 1]);
    };
    for (int si = 0; si < (int)shapes.size(); ++si) {
        const tinyobj::shape_t& sh = shapes[si];
        size_t off = 0;
        for (int f = 0; f < (int)sh.mesh.num_face_vertices.size(); ++f) {
            int nv = sh.mesh.num_face_vertices[f];
            if (nv != 3) {
                off += nv;
                continue;
            }
            int mi = f < (int)sh.mesh.material_ids.size() ? sh.mesh.material_ids[f] : -1;
            std::string mname = mi >= 0 && mi < (int)materials.size() ? materials[mi].name : std::string();
            std::string key = std::to_string(si) + ":" + std::to_string(mi);
            auto it = lookup.find(key);
            size_t m;
            if (it == lookup.end()) {
                m = L->mesh_names.size();
                lookup[key] = m;
                std::string shape_name;
                if (sh.name.empty()) {
                    shape_name = GetFilenameFromPath(p) + "_shape_" + std::to_string(si);
                } else {
                    shape_name = sh.name;
                }
                L->mesh_names.push_back(mname.empty() ? shape_name : shape_name + "::" + mname);
                L->mesh_materials.push_back(mname);
                L->pos_m.emplace_back();
                L->nrm_m.emplace_back();
                L->uv_m.emplace_back();
            } else {
                m = it->second;
            }
            Vec3f P[3], N[3];
            Vec2f U[3];
            for (int k = 0; k < 3; ++k) vertex(sh.mesh.indices[off + k], P[k], N[k], U[k]);
            bool has_n = sh.mesh.indices[off].normal_index >= 0 && sh.mesh.indices[off + 1].normal_index >= 0 &&
                         sh.mesh.indices[off + 2].normal_index >= 0;
            off += nv;
            if (!has_n) {
                Vec3f fn = (P[1] - P[0]).cross(P[2] - P[0]).normalized();
                N[0] = N[1] = N[2] = fn;
            }
            for (int k = 0; k < 3; ++k) {
                for (int c = 0; c < 3; ++c) {
                    L->pos_m[m].push_back(P[k](c));
                    L->nrm_m[m].push_back(N[k](c));
                }
                L->uv_m[m].push_back(U[k](0));
                L->uv_m[m].push_back(U[k](1));
            }
        }
    }
    // Scene order: meshes one after another (Scene::ParseObjects appends each mesh's triangles).
    for (size_t m = 0; m < L->mesh_names.size(); ++m) {
        L->pos.insert(L->pos.end(), L->pos_m[m].begin(), L->pos_m[m].end());
        L->nrm.insert(L->nrm.end(), L->nrm_m[m].begin(), L->nrm_m[m].end());
        L->uv.insert(L->uv.end(), L->uv_m[m].begin(), L->uv_m[m].end());
        L->mesh.insert(L->mesh.end(), L->pos_m[m].size() / 9, (int32_t)m);
    }
    return L;
}

int ref_obj_num_triangles(void* h) { return (int)((Loaded*)h)->mesh.size(); }
int ref_obj_num_meshes(void* h) { return (int)((Loaded*)h)->mesh_names.size(); }
void ref_obj_triangles(void* h, float* pos9, float* nrm9, float* uv6, int32_t* mesh) {
    Loaded* L = (Loaded*)h;
    std::memcpy(pos9, L->pos.data(), L->pos.size() * 4);
    std::memcpy(nrm9, L->nrm.data(), L->nrm.size() * 4);
    std::memcpy(uv6, L->uv.data(), L->uv.size() * 4);
    std::memcpy(mesh, L->mesh.data(), L->mesh.size() * 4);
}
// Name of mesh m and the material name it carries ("" when none), copied into
// buffers of `cap` bytes each.
void ref_obj_mesh_name(void* h, int m, char* name, char* material, int cap) {
    Loaded* L = (Loaded*)h;
    std::snprintf(name, cap, "%s", L->mesh_names[m].c_str());
    std::snprintf(material, cap, "%s", L->mesh_materials[m].c_str());
}
void ref_obj_free(void* h) { delete (Loaded*)h; }
}
