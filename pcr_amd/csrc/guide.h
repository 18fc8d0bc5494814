// guide.h -- the learned directional proposal ("path guiding") of the
// wavefront tracer: a 3-layer MLP (BF16 weights, FP32 accumulation on MFMA)
// maps a diffuse vertex (position, normal; 8-frequency sin/cos encoding) to
// an 8-lobe von Mises-Fisher mixture; the next direction is drawn from the
// one-sample MIS mixture 0.5 * cosine + 0.5 * vMF mixture, so the estimator
// stays unbiased whatever the network outputs.  The network is trained online
// on the tracer's own radiance samples (KL divergence between the normalised
// radiance-weighted target and the mixture), with Adam.
//
// The reference renderer has no guiding; the estimator it modifies is the
// diffuse bounce of code/Renderer.cpp:526-557 (cosine sampling, Ray.h:28-53).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pcr_kernels.h"

namespace pcr {

constexpr int kGIn = 128;      // encoded input width (102 used, zero padded for 16-wide MFMA steps)
constexpr int kGInUsed = 102;  // 6 raw + 6 x 8 frequencies x (sin, cos)
constexpr int kGHid = 64;
constexpr int kGLobes = 8;
constexpr int kGOut = 5 * kGLobes;  // per lobe: mean direction (3, unnormalised), log kappa, weight logit
constexpr int kGOutPad = 48;
constexpr int kGW1 = kGIn * kGHid, kGW2 = kGHid * kGHid, kGW3 = kGHid * kGOutPad;
// FP32 parameter vector: W1 [in][hid], b1, W2 [hid][hid], b2, W3 [hid][outpad], b3
constexpr int kGOffW1 = 0, kGOffB1 = kGW1, kGOffW2 = kGOffB1 + kGHid, kGOffB2 = kGOffW2 + kGW2,
              kGOffW3 = kGOffB2 + kGHid, kGOffB3 = kGOffW3 + kGW3, kGParams = kGOffB3 + kGOutPad;
constexpr float kGLogKappaMin = -3.f, kGLogKappaMax = 7.f;
constexpr int kGMaxSteps = 16;  // training steps per pass

// Device state of one job's network.
struct GuideNet {
    float *w;          // [kGParams] FP32 master weights
    uint16_t *wb;      // [2 (kGW1 + kGW2 + kGW3)] BF16 copies of W1, W2, W3: [in][out], then [out][in]
    float *grad;       // [kGParams] gradient accumulator
    float *m, *v;      // Adam moments
    float *wsum;       // [2 * kGMaxSteps]: per training step, weight total and loss
    uint32_t *n_rec;   // records written this pass
    int32_t *step;     // Adam steps taken
    float lo[3], inv[3];  // position normalisation: (p - lo) * inv in [0, 1]
};

// One training sample, written where a guided direction is drawn.
struct GuideRecord {
    float o[3], n[3], w[3];  // vertex (ray origin), surface normal, sampled direction
    float pdf;               // mixture pdf of w
    uint32_t path;           // job-local path index of the pass
    uint32_t chan;           // first channel | channel count << 8
    float rad[4];            // radiance gathered before the vertex's continuation, per channel
    float tp[4];             // throughput after the vertex's sampling weight, per channel
    float weight;            // training weight, filled after the pass
    float pad[3];
};
static_assert(sizeof(GuideRecord) == 96, "record layout");

// Training hyper-parameters of a pass.
struct GuideTrain {
    float lr;
    float beta1, beta2, eps;
};

// ------------------------------------------------------------ BF16 -------
// Round to nearest even (finite inputs), as the host-side copies do.
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// Arguments of the direction-drawing kernel of one bounce.
struct GuideArgs {
    GuideNet net;
    QueueView q;
    const uint32_t *count;  // live rays of the job's region (one job per guided pass)
    uint32_t cap;           // rays the region holds
    GuideRecord *rec;
    uint32_t rec_cap;
    uint32_t keep;          // record a vertex if its 24-bit hash is below keep
};
}  // namespace pcr

extern "C" {
hipError_t pcr_launch_guide(const pcr::GuideArgs *a, int grid, hipStream_t s);
hipError_t pcr_launch_guide_prep(const pcr::GuideNet *g, pcr::GuideRecord *rec, uint32_t rec_cap,
                                 const float4 *radiance, int S, hipStream_t s);
hipError_t pcr_launch_guide_grad(const pcr::GuideNet *g, const pcr::GuideRecord *rec, uint32_t rec_cap, int step,
                                 int S, int grid, hipStream_t s);
hipError_t pcr_launch_guide_adam(const pcr::GuideNet *g, const pcr::GuideTrain *h, int t, hipStream_t s);
hipError_t pcr_launch_guide_sync(const pcr::GuideNet *g, hipStream_t s);
hipError_t pcr_launch_guide_forward(const pcr::GuideNet *g, const float *pn, uint32_t n, float *out, int grid,
                                    hipStream_t s);
}
