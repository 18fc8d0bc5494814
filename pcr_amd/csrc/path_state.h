// path_state.h -- the per-ray state carried in the SoA wavefront queues
// (80 bytes per ray) and its packed control word, shared by the bounce and
// guiding kernels.
#pragma once

#include "pcr_kernels.h"

namespace pcr {

struct PathState {
    V3 o, d;
    float lam0;
    uint32_t meta;
    float tp[4];
    float rad[4];
    uint64_t rng;
    uint32_t path;
};
// meta: bounce (bits 0-7), first diffuse vertex pending (8), channel count
// (9-11), first channel (12-13), guided direction pending (14: d holds the
// surface normal until the guiding kernel draws the direction), ended by the
// guiding kernel (15: radiance still to be written), job slot (16-23).
constexpr uint32_t ST_GUIDE = 1u << 14;
constexpr uint32_t ST_DEAD = 1u << 15;
__device__ __forceinline__ int st_bounce(uint32_t m) { return (int)(m & 0xFFu); }
__device__ __forceinline__ bool st_first(uint32_t m) { return (m >> 8) & 1u; }
__device__ __forceinline__ int st_nch(uint32_t m) { return (int)((m >> 9) & 7u); }
__device__ __forceinline__ int st_ch0(uint32_t m) { return (int)((m >> 12) & 3u); }
__device__ __forceinline__ int st_slot(uint32_t m) { return (int)((m >> 16) & 0xFFu); }
__device__ __forceinline__ uint32_t st_pack(int bounce, bool first, int nch, int ch0, int slot) {
    return (uint32_t)bounce | ((uint32_t)first << 8) | ((uint32_t)nch << 9) | ((uint32_t)ch0 << 12) |
           ((uint32_t)slot << 16);
}

__device__ __forceinline__ void st_load(const QueueView &q, uint32_t i, PathState &s) {
    float4 a = q.od0[i], b = q.od1[i], t = q.tp[i], r = q.rad[i];
    uint4 g = q.rng[i];
    s.o = v3(a.x, a.y, a.z);
    s.lam0 = a.w;
    s.d = v3(b.x, b.y, b.z);
    s.meta = __float_as_uint(b.w);
    s.tp[0] = t.x; s.tp[1] = t.y; s.tp[2] = t.z; s.tp[3] = t.w;
    s.rad[0] = r.x; s.rad[1] = r.y; s.rad[2] = r.z; s.rad[3] = r.w;
    s.rng = (uint64_t)g.x | ((uint64_t)g.y << 32);
    s.path = g.z;
}
__device__ __forceinline__ void st_store(const QueueView &q, uint32_t i, const PathState &s) {
    q.od0[i] = make_float4(s.o.x, s.o.y, s.o.z, s.lam0);
    q.od1[i] = make_float4(s.d.x, s.d.y, s.d.z, __uint_as_float(s.meta));
    q.tp[i] = make_float4(s.tp[0], s.tp[1], s.tp[2], s.tp[3]);
    q.rad[i] = make_float4(s.rad[0], s.rad[1], s.rad[2], s.rad[3]);
    q.rng[i] = make_uint4((uint32_t)s.rng, (uint32_t)(s.rng >> 32), s.path, 0u);
}

// Cosine-weighted direction about N (Ray::genRayFromIntersection, Ray.h:28-53).
__device__ __forceinline__ V3 fm_hemisphere(V3 N, float r1, float r2) {
    V3 helper = fabsf(N.x) <= fabsf(N.y) ? v3(1.f, 0.f, 0.f) : v3(0.f, 1.f, 0.f);
    V3 T = normalize(cross(N, helper));
    V3 B = cross(N, T);
    float r = sqrt_rn(r1);
    float sn, cs;
    fm_sincos_turns(r2, sn, cs);
    float x = fmul(r, cs);
    float y = fmul(r, sn);
    float z = sqrt_rn(smax(0.f, fsub(fsub(1.f, fmul(x, x)), fmul(y, y))));
    return add(add(mul(T, x), mul(B, y)), mul(N, z));
}

}  // namespace pcr
