// pcr_kernels.h -- job descriptors and kernel argument blocks shared by the
// host-side driver (pcrhip.cpp) and the kernels (wavefront.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pcr_device.h"

namespace pcr {

// One render job's view of the current pass.  Lives in a device-resident job
// table indexed by the slot id carried in every path's state, so paths of
// different jobs (different scenes, resolutions, depths) share one queue.
struct DJob {
    DScene S;
    int32_t W, H;            // full frame
    int32_t x0, y0, cw, ch;  // crop window rendered by this job
    int32_t depth, shadow, spp, hero;
    uint32_t flags;
    int32_t strata;          // round(sqrt(spp)) when stratified
    uint64_t seed;
    float scale, aspect;     // tan(fov/2) (host libm) and W/H
    int32_t guide_bounces;   // guided jobs: vertices (bounce index below it) drawn from the learned mixture
    int32_t spp_pass;        // samples per pixel in one pass
    int32_t s_begin;         // first sample index of the current pass
    float4 *radiance;        // [cw*ch*spp_pass] channel radiance of each path
    double *accum;           // [cw*ch*3] running sums of XYZ (or RGB); adaptive: sums of the open group
    float4 *welford;         // adaptive: [cw*ch*2] (mean.xyz, m2.x), (m2.y, m2.z, points, done)
    int32_t group;           // adaptive: samples per Welford point
};

// Structure-of-arrays ray queue (80 bytes per ray).
struct QueueView {
    float4 *od0;   // origin.xyz, hero wavelength
    float4 *od1;   // direction.xyz, packed state word
    float4 *tp;    // throughput per channel
    float4 *rad;   // radiance gathered so far per channel
    uint4 *rng;    // PCG32 state lo/hi, path index, spare
};

// Generation segment: queue items [begin, begin+count) are paths
// [path0, path0+count) of job slot `slot`.
struct GenSeg {
    int32_t slot;
    uint32_t begin, count, path0;
};
struct ResolveSeg {
    int32_t slot;
    uint32_t begin, count;
    uint32_t pad;
};

constexpr int kMaxSegs = 64;

// Sizes of the current pass, read by the kernels from device memory so that
// the launch arguments stay the same from pass to pass (a captured pass
// graph is replayed unchanged).
struct PassInfo {
    uint32_t n_gen;    // generating bounce: number of work items (paths)
    int32_t n_segs;    // GenSeg entries
    uint32_t n_pix;    // resolve: number of pixels
    int32_t n_rsegs;   // ResolveSeg entries
};

// A small scene copied into LDS by every block of a bounce launch, in float4
// units: spheres, sphere materials (4 per float4), plane records, triangles,
// triangle normals, inner nodes, leaves (2 per float4), materials (6 per
// material), lights (5 per light), light triangles, spectral tables (4 floats
// per float4).  Used when all jobs of a pass share one scene that fits the
// block's capacity, which keeps 3 waves per SIMD within the 160 KB of LDS next
// to the traversal stacks.
struct GeoLayout {
    int sph, sphm, pl, tri, trin, bvh, leaf, mat, light, ltri, tabs, total;
};
__host__ __device__ inline GeoLayout geo_layout(const DScene &S) {
    GeoLayout g;
    g.sph = 0;
    g.sphm = S.n_sph;
    g.pl = g.sphm + (S.n_sph + 3) / 4;
    g.tri = g.pl + 4 * S.n_pl;
    g.trin = g.tri + 3 * S.n_tri;
    g.bvh = g.trin + 4 * S.n_tri;
    g.leaf = g.bvh + 4 * S.n_inner;
    g.mat = g.leaf + (S.n_leaves + 1) / 2;
    g.light = g.mat + (int)(sizeof(DMat) / 16) * S.n_mat;
    g.ltri = g.light + (int)(sizeof(DLight) / 16) * S.n_lights;
    g.tabs = g.ltri + 4 * S.n_ltri;
    g.total = g.tabs + (S.n_tabs + 3) / 4;
    return g;
}
static_assert(sizeof(DMat) % 16 == 0 && sizeof(DLight) % 16 == 0, "LDS scene records are float4 multiples");
__host__ __device__ constexpr int geo_cap(int block) { return block >= 256 ? 1024 : block >= 128 ? 512 : 256; }

struct BounceArgs {
    const DJob *jobs;
    const GenSeg *segs;
    const PassInfo *pass;
    QueueView in, out;
    const uint32_t *count_in;     // [kMaxSegs] live rays per segment in `in`
    uint32_t *count_out;          // [kMaxSegs] rays appended per segment to `out`
    uint32_t *work;               // work-stealing counter of this launch
    uint32_t *overflow;           // set if an append would exceed its region
    unsigned long long *counters; // paths, extension rays, shadow rays, emissive hits
};

struct ResolveArgs {
    const DJob *jobs;
    const ResolveSeg *segs;
    const PassInfo *pass;
};

struct IntersectArgs {
    DScene S;
    const float *rays;
    uint32_t n;
    float tmax;
    const int32_t *tri_orig;  // leaf-order -> original triangle index (may be null)
    float *t;
    int32_t *kind, *prim, *mat;
    float *hit, *normal;
};

}  // namespace pcr

extern "C" {
// lds_scene: every job of the pass uses one scene whose geo_layout fits
// geo_cap(block) (stacks of at most 16 entries); the blocks read it from LDS.
hipError_t pcr_launch_bounce(const pcr::BounceArgs *a, int first, int stack, int grid, int block, int lds_scene,
                             hipStream_t s);
// Threads per block the bounce launcher actually uses for a stack depth.
int pcr_bounce_block(int stack, int block);
hipError_t pcr_launch_resolve(const pcr::ResolveArgs *a, int grid, hipStream_t s);
hipError_t pcr_launch_intersect(const pcr::IntersectArgs *a, int stack, hipStream_t s);
hipError_t pcr_launch_reset(uint32_t *p, int n, hipStream_t s);
}
