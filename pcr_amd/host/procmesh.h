// procmesh.h -- seeded procedural closed mesh: a displaced icosphere.
//
// Used for the traversal-stress scene (icosphere at subdivision 6 = 81,920
// triangles, radially displaced by 3-octave value noise).  The reference has
// no procedural meshes; its scenes load OBJ files (code/Scenes/MeshLoader.cpp),
// so this generator is an input format of this project.  Every operation is a
// double-precision + - * / sqrt or floor, so any conforming compiler produces
// the same float32 vertices bit for bit.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

namespace pcr {

struct IcoParams {
    int subdiv = 6;
    uint32_t seed = 42;
    double radius = 100.0;     // base radius (world units)
    double amplitude = 0.15;   // displacement amplitude as a fraction of radius
    double center[3] = {0.0, 0.0, 0.0};
    int octaves = 3;
    double base_freq = 2.0;    // lattice frequency of the first octave on the unit sphere
};

namespace detail {

inline uint32_t lattice_hash(uint32_t seed, int32_t x, int32_t y, int32_t z) {
    uint32_t h = seed * 0x9E3779B1u;
    h ^= (uint32_t)x * 0x85EBCA77u;
    h = (h << 13) | (h >> 19);
    h ^= (uint32_t)y * 0xC2B2AE3Du;
    h = (h << 17) | (h >> 15);
    h ^= (uint32_t)z * 0x27D4EB2Fu;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}
// value in [-1, 1]
inline double lattice_value(uint32_t seed, int32_t x, int32_t y, int32_t z) {
    return (double)(lattice_hash(seed, x, y, z) >> 8) * (2.0 / 16777215.0) - 1.0;
}
inline double smooth(double t) { return t * t * (3.0 - 2.0 * t); }
inline double lerp(double a, double b, double t) { return a + (b - a) * t; }

inline double value_noise(uint32_t seed, double x, double y, double z) {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    double tx = smooth(x - fx), ty = smooth(y - fy), tz = smooth(z - fz);
    double c000 = lattice_value(seed, ix, iy, iz), c100 = lattice_value(seed, ix + 1, iy, iz);
    double c010 = lattice_value(seed, ix, iy + 1, iz), c110 = lattice_value(seed, ix + 1, iy + 1, iz);
    double c001 = lattice_value(seed, ix, iy, iz + 1), c101 = lattice_value(seed, ix + 1, iy, iz + 1);
    double c011 = lattice_value(seed, ix, iy + 1, iz + 1), c111 = lattice_value(seed, ix + 1, iy + 1, iz + 1);
    double x00 = lerp(c000, c100, tx), x10 = lerp(c010, c110, tx);
    double x01 = lerp(c001, c101, tx), x11 = lerp(c011, c111, tx);
    double y0 = lerp(x00, x10, ty), y1 = lerp(x01, x11, ty);
    return lerp(y0, y1, tz);
}
// Octave sum normalised to [-1, 1]: frequency doubles, amplitude halves.
inline double fbm(const IcoParams &p, double x, double y, double z) {
    double sum = 0.0, norm = 0.0, amp = 1.0, f = p.base_freq;
    for (int o = 0; o < p.octaves; o++) {
        sum = sum + amp * value_noise(p.seed + (uint32_t)o * 1013u, x * f, y * f, z * f);
        norm = norm + amp;
        amp = amp * 0.5;
        f = f * 2.0;
    }
    return sum / norm;
}
struct D3 { double x, y, z; };
inline D3 unit(D3 a) {
    double m = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return D3{a.x / m, a.y / m, a.z / m};
}

}  // namespace detail

// Returns 9 floats per triangle (v0, v1, v2), counter-clockwise seen from
// outside, in a fixed recursion order.  20 * 4^subdiv triangles.
inline std::vector<float> make_displaced_icosphere(const IcoParams &p) {
    using namespace detail;
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<D3> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto &q : v) q = unit(q);
    std::vector<std::array<int, 3>> f = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
                                         {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                         {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
                                         {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (int s = 0; s < p.subdiv; s++) {
        std::map<std::pair<int, int>, int> mids;
        auto mid = [&](int a, int b) {
            auto key = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
            auto it = mids.find(key);
            if (it != mids.end()) return it->second;
            D3 m{(v[a].x + v[b].x) * 0.5, (v[a].y + v[b].y) * 0.5, (v[a].z + v[b].z) * 0.5};
            v.push_back(unit(m));
            int idx = (int)v.size() - 1;
            mids.emplace(key, idx);
            return idx;
        };
        std::vector<std::array<int, 3>> nf;
        nf.reserve(f.size() * 4);
        for (const auto &tri : f) {
            int a = mid(tri[0], tri[1]), b = mid(tri[1], tri[2]), c = mid(tri[2], tri[0]);
            nf.push_back({tri[0], a, c});
            nf.push_back({tri[1], b, a});
            nf.push_back({tri[2], c, b});
            nf.push_back({a, b, c});
        }
        f.swap(nf);
    }
    std::vector<float> pos(v.size() * 3);
    for (size_t i = 0; i < v.size(); i++) {
        double r = p.radius * (1.0 + p.amplitude * fbm(p, v[i].x, v[i].y, v[i].z));
        pos[3 * i] = (float)(p.center[0] + v[i].x * r);
        pos[3 * i + 1] = (float)(p.center[1] + v[i].y * r);
        pos[3 * i + 2] = (float)(p.center[2] + v[i].z * r);
    }
    std::vector<float> out;
    out.reserve(f.size() * 9);
    for (const auto &tri : f)
        for (int k = 0; k < 3; k++)
            for (int c = 0; c < 3; c++) out.push_back(pos[3 * tri[k] + c]);
    return out;
}

}  // namespace pcr
