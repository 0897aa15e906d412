// Scene-JSON and OBJ loading (reference semantics: Src/Core/scene.cpp:129-690,
// Src/Utils/objloader.cpp:66-210), camera set-up and BVH build.
#include "scene.h"

#include <hip/hip_runtime.h>

#include <algorithm>
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
// General inverse by Gauss-Jordan elimination in double, rounded to float.
Mat4 inverse(const Mat4& a) {
    double m[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) m[i][j] = j < 4 ? (double)a.m[i][j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(m[r][c]) > std::fabs(m[piv][c])) piv = r;
        if (m[piv][c] == 0.0) throw std::runtime_error("object transform is singular");
        if (piv != c)
            for (int j = 0; j < 8; ++j) std::swap(m[c][j], m[piv][j]);
        double d = m[c][c];
        for (int j = 0; j < 8; ++j) m[c][j] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            double f = m[r][c];
            if (f == 0.0) continue;
            for (int j = 0; j < 8; ++j) m[r][j] -= f * m[c][j];
        }
    }
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = (float)m[i][j + 4];
    return r;
}
vec3 xform_point(const Mat4& t, vec3 p) {
    float o[3];
    for (int r = 0; r < 3; ++r)
        o[r] = ((t.m[r][0] * p.x + t.m[r][1] * p.y) + t.m[r][2] * p.z) + t.m[r][3] * 1.f;
    return v3(o[0], o[1], o[2]);
}
// transformInv.transpose() * (n, 1), then normalized (reference: Src/Utils/objloader.cpp:36-41).
vec3 xform_normal(const Mat4& inv, vec3 n) {
    float o[3];
    for (int r = 0; r < 3; ++r)
        o[r] = ((inv.m[0][r] * n.x + inv.m[1][r] * n.y) + inv.m[2][r] * n.z) + inv.m[3][r] * 1.f;
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
struct ObjIdx {
    int v = -1, vt = -1, vn = -1;
};
struct LoadedMesh {
    std::string name;
    std::string material;
    std::vector<float> pos, nrm, uv;  // 9, 9, 6 per triangle
    int count = 0;
};

int fix_index(long idx, size_t n) {
    if (idx > 0) return (int)(idx - 1);
    if (idx < 0) return (int)((long)n + idx);
    return -1;
}

// Triangulated OBJ import, one mesh per (shape, material) pair, face normals
// when a face lacks per-vertex normals (reference: Src/Utils/objloader.cpp:66-210).
//
// The text is parsed in line-aligned chunks on host threads: each chunk
// parses its vertex lines (strtof) and the raw corner indices of its faces
// (fan-triangulated), noting the chunk's vertex counts at every face line for
// relative indices.  A sequential pass resolves the indices, replays the
// object / group / material statements in file order and assigns every
// triangle its mesh; the triangles are then written in parallel, each into its
// place in its mesh, with the same per-vertex arithmetic.
namespace {

struct ObjCorner {
    int32_t v, vt, vn;  // raw OBJ indices (1-based or negative), 0 = absent
};
struct ObjTri {
    ObjCorner c[3];
    int32_t nv, nvt, nvn;  // the chunk's vertex counts before the face line
};
struct ObjEvent {
    size_t tri;  // chunk triangles before the statement
    int kind;    // 0 object / group, 1 material
    std::string name;
};
struct ObjChunk {
    std::vector<float> V, VN, VT;
    std::vector<ObjTri> tris;
    std::vector<ObjEvent> events;
};

std::string trim_ws(const char* b, const char* e) {
    while (b < e && (*b == ' ' || *b == '\t')) ++b;
    while (e > b && (e[-1] == ' ' || e[-1] == '\t')) --e;
    return std::string(b, e);
}

// Integer token prefix as strtol reads it (optional sign, decimal digits; 0 if none).
long parse_long(const char*& t, const char* end) {
    const char* q = t;
    bool neg = false;
    if (q < end && (*q == '+' || *q == '-')) neg = *q++ == '-';
    if (q >= end || *q < '0' || *q > '9') return 0;  // no conversion: t stays
    long v = 0;
    while (q < end && *q >= '0' && *q <= '9') v = v * 10 + (*q++ - '0');
    t = q;
    return neg ? -v : v;
}

void parse_obj_chunk(const char* b, const char* e, ObjChunk& ck) {
    std::string line;
    std::vector<ObjCorner> poly;
    while (b < e) {
        const char* nl = static_cast<const char*>(std::memchr(b, '\n', (size_t)(e - b)));
        const char* le = nl ? nl : e;
        line.assign(b, le);  // NUL-terminated copy for strtof
        b = nl ? nl + 1 : e;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        size_t s = line.find_first_not_of(" \t");
        if (s == std::string::npos || line[s] == '#') continue;
        const char* c = line.c_str() + s;
        const char* ce = line.c_str() + line.size();
        if (c[0] == 'v' && (c[1] == ' ' || c[1] == '\t')) {
            char* q;
            float x = std::strtof(c + 2, &q);
            float y = std::strtof(q, &q);
            float z = std::strtof(q, &q);
            ck.V.push_back(x); ck.V.push_back(y); ck.V.push_back(z);
        } else if (c[0] == 'v' && c[1] == 'n') {
            char* q;
            float x = std::strtof(c + 2, &q);
            float y = std::strtof(q, &q);
            float z = std::strtof(q, &q);
            ck.VN.push_back(x); ck.VN.push_back(y); ck.VN.push_back(z);
        } else if (c[0] == 'v' && c[1] == 't') {
            char* q;
            float x = std::strtof(c + 2, &q);
            float y = std::strtof(q, &q);
            ck.VT.push_back(x); ck.VT.push_back(y);
        } else if (c[0] == 'f' && (c[1] == ' ' || c[1] == '\t')) {
            poly.clear();
            const char* t = c + 2;
            while (true) {
                while (t < ce && std::isspace((unsigned char)*t)) ++t;
                if (t >= ce) break;
                const char* te = t;
                while (te < ce && !std::isspace((unsigned char)*te)) ++te;
                ObjCorner k{0, 0, 0};
                const char* q = t;
                k.v = (int32_t)parse_long(q, te);
                if (q < te && *q == '/') {
                    ++q;
                    if (q < te && *q != '/') k.vt = (int32_t)parse_long(q, te);
                    if (q < te && *q == '/') {
                        ++q;
                        k.vn = (int32_t)parse_long(q, te);
                    }
                }
                poly.push_back(k);
                t = te;
            }
            const int32_t nv = (int32_t)(ck.V.size() / 3), nvt = (int32_t)(ck.VT.size() / 2),
                          nvn = (int32_t)(ck.VN.size() / 3);
            for (size_t k = 1; k + 1 < poly.size(); ++k) ck.tris.push_back(ObjTri{{poly[0], poly[k], poly[k + 1]}, nv, nvt, nvn});
        } else if ((c[0] == 'o' || c[0] == 'g') && (c[1] == ' ' || c[1] == '\t' || c[1] == 0)) {
            ck.events.push_back(ObjEvent{ck.tris.size(), 0, c[1] ? trim_ws(c + 2, ce) : std::string()});
        } else if (std::strncmp(c, "usemtl", 6) == 0) {
            ck.events.push_back(ObjEvent{ck.tris.size(), 1, trim_ws(c + 6, ce)});
        }
    }
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

std::vector<LoadedMesh> load_obj(const std::string& path, const Mat4& xf, const Mat4& xf_inv) {
    const std::string text = read_file(path);
    // Line-aligned chunks of about 1 MiB.
    std::vector<size_t> cuts{0};
    const size_t target = (size_t)1 << 20;
    while (cuts.back() < text.size()) {
        size_t c = std::min(text.size(), cuts.back() + target);
        while (c < text.size() && text[c - 1] != '\n') ++c;
        cuts.push_back(c);
    }
    const size_t nchunks = cuts.size() - 1;
    std::vector<ObjChunk> chunks(nchunks);
    parallel_for(nchunks, [&](size_t i) { parse_obj_chunk(text.data() + cuts[i], text.data() + cuts[i + 1], chunks[i]); });

    // Global vertex arrays and the chunks' offsets into them.
    std::vector<size_t> v0(nchunks + 1, 0), vt0(nchunks + 1, 0), vn0(nchunks + 1, 0), t0(nchunks + 1, 0);
    for (size_t i = 0; i < nchunks; ++i) {
        v0[i + 1] = v0[i] + chunks[i].V.size() / 3;
        vt0[i + 1] = vt0[i] + chunks[i].VT.size() / 2;
        vn0[i + 1] = vn0[i] + chunks[i].VN.size() / 3;
        t0[i + 1] = t0[i] + chunks[i].tris.size();
    }
    std::vector<float> V(3 * v0[nchunks]), VT(2 * vt0[nchunks]), VN(3 * vn0[nchunks]);
    parallel_for(nchunks, [&](size_t i) {
        std::copy(chunks[i].V.begin(), chunks[i].V.end(), V.begin() + 3 * v0[i]);
        std::copy(chunks[i].VT.begin(), chunks[i].VT.end(), VT.begin() + 2 * vt0[i]);
        std::copy(chunks[i].VN.begin(), chunks[i].VN.end(), VN.begin() + 3 * vn0[i]);
    });

    // Statements in file order: shape names, and every triangle's (shape, material).
    const size_t ntris = t0[nchunks];
    std::vector<std::string> shape_names{""};
    std::vector<std::string> materials;
    std::map<std::string, int> material_id;
    std::vector<std::pair<int, int>> tri_key(ntris);
    int shape = 0, mtl = -1;
    bool shape_has_faces = false;
    for (size_t i = 0; i < nchunks; ++i) {
        size_t t = 0;
        auto emit_until = [&](size_t stop) {
            if (t < stop) shape_has_faces = true;
            for (; t < stop; ++t) tri_key[t0[i] + t] = {shape, mtl};
        };
        for (const ObjEvent& ev : chunks[i].events) {
            emit_until(ev.tri);
            if (ev.kind == 0) {
                if (shape_has_faces) {
                    shape_names.push_back(ev.name);
                    ++shape;
                    shape_has_faces = false;
                } else {
                    shape_names[shape] = ev.name;
                }
            } else if (ev.name.empty()) {
                mtl = -1;  // the same (empty) material as before any usemtl
            } else {
                auto it = material_id.find(ev.name);
                if (it == material_id.end()) {
                    it = material_id.emplace(ev.name, (int)materials.size()).first;
                    materials.push_back(ev.name);
                }
                mtl = it->second;
            }
        }
        emit_until(chunks[i].tris.size());
    }

    // Meshes in order of first appearance; each triangle's slot in its mesh.
    std::vector<LoadedMesh> meshes;
    std::map<std::pair<int, int>, int> lookup;
    std::vector<int32_t> tri_mesh(ntris), tri_slot(ntris);
    std::pair<int, int> last_key{-2, -2};
    int last_mesh = -1;
    for (size_t t = 0; t < ntris; ++t) {
        if (tri_key[t] != last_key) {
            auto it = lookup.find(tri_key[t]);
            if (it == lookup.end()) {
                LoadedMesh m;
                const int sh = tri_key[t].first, mt = tri_key[t].second;
                const std::string& sn = shape_names[sh];
                const std::string mname = mt >= 0 ? materials[mt] : std::string();
                std::string base = sn.empty() ? stem_of(path) + "_shape_" + std::to_string(sh) : sn;
                m.name = mname.empty() ? base : base + "::" + mname;
                m.material = mname;
                it = lookup.emplace(tri_key[t], (int)meshes.size()).first;
                meshes.push_back(std::move(m));
            }
            last_key = tri_key[t];
            last_mesh = it->second;
        }
        tri_mesh[t] = last_mesh;
        tri_slot[t] = meshes[last_mesh].count++;
    }
    for (LoadedMesh& m : meshes) {
        m.pos.resize(9 * (size_t)m.count);
        m.nrm.resize(9 * (size_t)m.count);
        m.uv.resize(6 * (size_t)m.count);
    }

    auto resolve = [](int32_t idx, size_t n) -> int { return fix_index(idx, n); };
    const size_t block = 4096;
    parallel_for((ntris + block - 1) / block, [&](size_t bi) {
        size_t ci = std::upper_bound(t0.begin(), t0.end(), bi * block) - t0.begin() - 1;
        for (size_t t = bi * block; t < std::min(ntris, (bi + 1) * block); ++t) {
            while (t >= t0[ci + 1]) ++ci;
            const ObjTri& tr = chunks[ci].tris[t - t0[ci]];
            vec3 p[3], n[3];
            float u[3], v[3];
            bool hn[3];
            for (int k = 0; k < 3; ++k) {
                const ObjCorner& c = tr.c[k];
                const int iv = resolve(c.v, v0[ci] + (size_t)tr.nv);
                const int ivt = resolve(c.vt, vt0[ci] + (size_t)tr.nvt);
                const int ivn = resolve(c.vn, vn0[ci] + (size_t)tr.nvn);
                if (iv < 0 || (size_t)(3 * (size_t)iv + 2) >= V.size())
                    throw std::runtime_error("OBJ face references a missing vertex in " + path);
                p[k] = xform_point(xf, v3(V[3 * (size_t)iv], V[3 * (size_t)iv + 1], V[3 * (size_t)iv + 2]));
                hn[k] = ivn >= 0 && (size_t)(3 * (size_t)ivn + 2) < VN.size();
                n[k] = hn[k] ? xform_normal(xf_inv, v3(VN[3 * (size_t)ivn], VN[3 * (size_t)ivn + 1], VN[3 * (size_t)ivn + 2]))
                             : v3(0.f, 0.f, 0.f);
                if (ivt >= 0 && (size_t)(2 * (size_t)ivt + 1) < VT.size()) {
                    u[k] = VT[2 * (size_t)ivt];
                    v[k] = VT[2 * (size_t)ivt + 1];
                } else {
                    u[k] = 0.f;
                    v[k] = 0.f;
                }
            }
            if (!(hn[0] && hn[1] && hn[2])) {
                vec3 fn = normalize(cross(p[1] - p[0], p[2] - p[0]));
                n[0] = n[1] = n[2] = fn;
            }
            LoadedMesh& m = meshes[tri_mesh[t]];
            const size_t o = (size_t)tri_slot[t];
            for (int k = 0; k < 3; ++k) {
                m.pos[9 * o + 3 * k] = p[k].x; m.pos[9 * o + 3 * k + 1] = p[k].y; m.pos[9 * o + 3 * k + 2] = p[k].z;
                m.nrm[9 * o + 3 * k] = n[k].x; m.nrm[9 * o + 3 * k + 1] = n[k].y; m.nrm[9 * o + 3 * k + 2] = n[k].z;
                m.uv[6 * o + 2 * k] = u[k]; m.uv[6 * o + 2 * k + 1] = v[k];
            }
        }
    });
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
    setup_camera(sc);
}

}  // namespace rt
