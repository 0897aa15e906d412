// Host-side scene: scene-JSON + OBJ loading, camera set-up and BVH build.
#pragma once

#include <stdint.h>
#include <string>
#include <vector>

#include "bvh.h"
#include "image.h"
#include "rt_math.h"

namespace rt {

// Material kinds in the reference's enum order (reference: Src/Core/material.h:6-12).
enum MaterialType : int32_t { kLight = 0, kDiffuse = 1, kPbr = 2, kDielectric = 3, kSubsurface = 4 };
enum LightType : int32_t { kAreaLight = 0, kPointLight = 1, kEnvLight = 2 };
// Subsurface transport of a "subsurface" material: the reference's Burley
// diffusion profile with two probe rays (material.cu:326-441), or a random walk
// through the medium's absorption and scattering coefficients (this project's
// addition; JSON "subsurface_method": "random_walk").
enum SssMethod : int32_t { kSssBurley = 0, kSssRandomWalk = 1 };

// Device material record (96 bytes).  Defaults are the reference's Material
// member initialisers (reference: Src/Core/material.h:57-71); sigma_a, sigma_s
// and sss_method are read by the random-walk subsurface mode only.
struct MaterialRec {
    int32_t type = kDiffuse;
    float basecolor[3] = {0.f, 0.f, 0.f};
    float ke[3] = {0.f, 0.f, 0.f};
    float roughness = 0.5f;
    float metallic = 0.0f;
    float transmission = 1.0f;
    float ior = 1.0f;
    float scatter[3] = {0.f, 0.f, 0.f};
    int32_t texture = -1;
    int32_t valid = 0;
    float sigma_a[3] = {0.f, 0.f, 0.f};  // absorption coefficient per scene unit, RGB
    float sigma_s[3] = {1.f, 1.f, 1.f};  // scattering coefficient per scene unit, RGB
    int32_t sss_method = kSssBurley;
    int32_t pad_ = 0;
};
static_assert(sizeof(MaterialRec) == 96, "material record is 96 bytes");

// Device light record (32 bytes).
struct LightRec {
    int32_t type = kAreaLight;
    int32_t mesh = -1;
    float emission[3] = {0.f, 0.f, 0.f};
    int32_t tri_start = 0;  // first triangle in the light-triangle array
    int32_t tri_count = 0;
    float area = 0.f;       // mesh area (sum of triangle areas, scene order)
};
static_assert(sizeof(LightRec) == 32, "light record is 32 bytes");

struct MeshRec {
    std::string name;
    int32_t material = -1;
    int32_t light = -1;
    int32_t start = 0;
    int32_t count = 0;
    float area = 0.f;
};

// Pinhole orbit camera (reference: Src/Core/camera.h:11-84; Src/Core/scene.cpp:646-667).
struct CameraRec {
    int width = 800, height = 600;
    float fovy = 45.f;
    vec3 direction = {0.f, 0.f, -1.f};
    vec3 target, position, lookat, up, u, v, w, du, dv, left_down;
    float distance = 0.f, theta = 0.f, phi = 0.f, focal = 0.f;
};

struct EnvMap {
    int width = 0, height = 0;
    std::vector<float> rgb;  // width*height*3, row 0 = top
};

// The random walk's own tree: a BVH over the triangles of one mesh whose
// material is a random-walk subsurface medium (built like the scene's, over
// that mesh's triangles in scene order), so the walk's segment queries see
// only the medium's own boundary.
struct MediumTree {
    int mesh = -1;
    BvhBuild bvh;
    // Clearance grid: n^3 cells over the tree's box; cell value = a lower
    // bound, safe under FP32 rounding, of the distance from any point of the
    // cell to the mesh (0 near the surface).  A walk step whose free flight is
    // shorter than its cell's value cannot meet the surface and needs no query.
    int grid_n = 0;
    vec3 grid_lo, grid_inv;     // cell index = floor((p - lo) * inv)
    std::vector<float> clearance;
};

struct HostScene {
    std::string path;
    // Scene-order triangle soup.
    std::vector<float> pos;      // 9 per triangle
    std::vector<float> nrm;      // 9 per triangle (per-vertex shading normals)
    std::vector<float> uv;       // 6 per triangle
    std::vector<int32_t> mesh;   // mesh index per triangle
    std::vector<float> gnormal;  // 3 per triangle, normalised geometric normal
    std::vector<float> area;     // per triangle
    std::vector<MeshRec> meshes;
    std::vector<MaterialRec> materials;
    std::vector<std::string> material_names;
    std::vector<LightRec> lights;
    int env_light = -1;
    EnvMap env;
    std::vector<LdrImage> textures;  // "textures" entries in order, RGBA8 (row 0 = top)
    CameraRec camera;
    int max_bounces = 5;
    float rr = 0.8f;
    int spp = 16;
    bool denoise = false;
    uint32_t seed = 1234u;      // reference: Src/Core/scene.cpp:167
    int rr_start_depth = 0;     // 0: roulette from the first bounce, as the reference
    BvhBuild bvh;
    std::vector<MediumTree> media;  // one per mesh of a random-walk subsurface material
    double bvh_build_ms = 0.0;  // wall time of the BVH build
    double load_phase_ms[3] = {0, 0, 0};  // OBJ import, triangle finalize, environment map
    bool bvh_on_gpu = false;
    std::vector<std::string> warnings;

    int num_tris() const { return (int)mesh.size(); }
};

// Where the BVH is built: automatic = on the GPU when one is present and the
// scene has at least kGpuBuildMinTris triangles.  Both give the same tree.
enum BvhBuilder : int { kBvhAuto = 0, kBvhHost = 1, kBvhGpu = 2 };
constexpr int kGpuBuildMinTris = 1 << 16;
// Parse a scene JSON (+ OBJ files) and build the BVH.  Throws std::runtime_error.
void load_scene(const std::string& json_path, HostScene& sc, int bvh_builder = kBvhAuto);
// Triangle constructor math: geometric normal and area (reference: Src/Core/triangle.cu:3-10).
void finalize_triangles(HostScene& sc);
// New vertex positions (9 floats per triangle, scene order) and, unless nrm is
// null, shading normals for the same triangles: geometric normals, areas and
// mesh and light areas follow.  The BVH and the camera are not touched (see
// refit_bvh).  Throws std::invalid_argument for a non-finite coordinate,
// before anything changes.
void set_scene_geometry(HostScene& sc, const float* pos, const float* nrm);
// Medium trees: built after the scene's BVH; refit (same topology, new boxes)
// after set_scene_geometry.
void build_media(HostScene& sc);
void refit_media(HostScene& sc);
// Camera parameters from the scene bounds (reference: Src/Core/scene.cpp:646-667, camera.h:58-84).
void setup_camera(HostScene& sc);
// Pinhole basis and pixel steps from the orbit parameters (target, distance, theta, phi).
void compute_camera(CameraRec& c);
// Row-major projection * view of the G-buffer pass (reference: renderer.cuh:1463-1474).
void camera_view_projection(const CameraRec& c, float near_plane, float far_plane, float vp[16]);

}  // namespace rt
