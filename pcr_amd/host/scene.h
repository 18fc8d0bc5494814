// scene.h -- host-side scene model and ingestion for the HIP backend.
//
// Mirrors the reference's scene model (code/Scenes/Scene.h:22-129) and its
// ingestion (code/Scenes/SceneLoader.cpp:89-516): materials with RGB albedo /
// emission, optional measured spectra (SPD files), Jakob-Hanika sigmoid fits,
// spheres, planes (walls and plane lights), triangles (BVH-permuted) and area
// lights.  The result packs into the flat pcrhip_scene_desc of include/pcrhip.h.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pcrhip.h"

namespace pcr {

struct Vec3 {
    float x = 0.f, y = 0.f, z = 0.f;
};

// Jakob & Hanika sigmoid-polynomial coefficients (RGBToSpectrum.h:98-104)
struct SigmoidFit {
    float c0 = 0.f, c1 = 0.f, c2 = 0.f, scale = 0.f;
};

struct Material {
    Vec3 albedo, emissive;
    std::array<float, 61> tabAlbedo{}, tabEmissive{};
    bool useTabAlbedo = false, useTabEmissive = false;
    bool metallic = false, transparent = false;
    float ior = 1.5f, cauchyB = 0.f;
    SigmoidFit fitAlbedo, fitEmissive;
    bool isEmissive() const { return emissive.x > 0 || emissive.y > 0 || emissive.z > 0; }
};

struct Sphere {
    Vec3 center;
    float radius = 1.f;
    int mat = 0;
};
struct Plane {
    Vec3 o, u, v;
    int mat = 0;
};
struct Triangle {
    Vec3 v0, v1, v2, n0, n1, n2;
    bool smooth = false;
    int mat = 0;
    int orig = 0;  // declaration index, kept through the BVH permutation
};
struct Light {
    int kind = 0;                 // 0 plane, 1 triangle set
    int plane = -1;               // index into Scene::lightPlanes
    std::vector<Triangle> tris;   // copies (the BVH permutes Scene::triangles)
    std::vector<float> cum;       // cumulative areas
    float area = 0.f;
};
struct BvhNode {
    float bmin[3], bmax[3];
    int32_t left, right, count;
};

struct Scene {
    std::string name, version, path;
    Vec3 cam;
    float fov = 65.f;
    std::vector<Material> materials;
    std::vector<std::string> materialNames;
    std::vector<Sphere> spheres;
    std::vector<Plane> walls;
    std::vector<Plane> lightPlanes;
    std::vector<Triangle> triangles;  // BVH leaf order after load
    std::vector<BvhNode> bvh;
    std::vector<Light> lights;
};

struct SceneError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// spectral.cpp
float ybar_integral();
Vec3 spectrum_to_xyz(const std::array<float, 61> &s);
Vec3 xyz_to_linear_srgb(Vec3 xyz);
Vec3 linear_srgb_to_xyz(Vec3 rgb);
SigmoidFit fit_sigmoid(Vec3 rgb);
void populate_spectra(Material &m);
bool load_spd_file(const std::string &path, std::array<float, 61> &out, std::string *err);

// bvh.cpp: bit-exact restatement of the reference builder (code/Bvh/Bvh.cpp:122-325)
std::vector<BvhNode> build_bvh(std::vector<Triangle> &tris);
// The same build through a content-keyed cache (in memory, and in the
// directory of bvh_cache_set_dir or $PCR_BVH_CACHE when set).
std::vector<BvhNode> build_bvh_cached(std::vector<Triangle> &tris);
struct BvhCacheStats {
    uint64_t builds, memory_hits, disk_hits;
};
void bvh_cache_set_dir(const char *dir);  // NULL or "" keeps builds in memory only
void bvh_cache_clear_memory();
BvhCacheStats bvh_cache_stats();
int bvh_depth(const std::vector<BvhNode> &nodes);

// loader.cpp
Scene load_scene(const std::string &path, const std::vector<std::string> &spdDirs = {});

// Flat view of a Scene for pcrhip_scene_upload; owns the arrays desc points to.
struct PackedScene {
    std::vector<float> matAlbedo, matEmissive, matFitA, matFitE, matTabA, matTabE, matIor, matCauchy;
    std::vector<uint32_t> matFlags;
    std::vector<float> spheres;
    std::vector<int32_t> sphereMat;
    std::vector<float> planes;
    std::vector<int32_t> planeMat;
    std::vector<float> tris, triNormals;
    std::vector<uint8_t> triSmooth;
    std::vector<int32_t> triMat, triOrig;
    std::vector<int32_t> nodes;
    std::vector<int32_t> lightKind, lightFirst, lightCount;
    std::vector<float> lightArea;
    std::vector<float> lightPlanes;
    std::vector<int32_t> lightPlaneMat;
    std::vector<float> lightTris, lightTriCum;
    std::vector<int32_t> lightTriMat;
    pcrhip_scene_desc desc{};
};
void pack_scene(const Scene &s, PackedScene &p);

}  // namespace pcr
