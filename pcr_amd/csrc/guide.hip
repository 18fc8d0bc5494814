// guide.hip -- gfx950 kernels of the learned directional proposal: encoding,
// the 3-layer MLP (forward and backward as 16x16x16 BF16 MFMA tiles with FP32
// accumulation, one wave per 64 rays or records), vMF-mixture sampling with
// one-sample MIS against cosine sampling, training-record preparation and Adam.
#include "guide.h"
#include "path_state.h"

namespace pcr {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------ MFMA tiles --
// One wave computes C[M x N] = A[M x K] * B[K x N] as 16x16x16 tiles of
// v_mfma_f32_16x16x16_bf16.  Operand layout of the instruction: lane l holds
// A[l % 16][4 (l / 16) + i], B[4 (l / 16) + i][l % 16] and receives
// C[4 (l / 16) + i][l % 16], i = 0..3.  a(m, k) and b(k, n) return BF16 bits;
// out(m, n, v) receives each FP32 result.  Every lane of the wave must be
// active.
template <int M, int N, int K, class FA, class FB, class FC>
__device__ __forceinline__ void wave_gemm(FA a, FB b, FC out) {
    const int l = (int)__lane_id();
    const int r = l & 15, q = l >> 4;
    for (int mt = 0; mt < M / 16; mt++)
        for (int nt = 0; nt < N / 16; nt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int kt = 0; kt < K / 16; kt++) {
                bf16x4 A, B;
                _Pragma("unroll") for (int i = 0; i < 4; i++) {
                    A[i] = (short)a(mt * 16 + r, kt * 16 + 4 * q + i);
                    B[i] = (short)b(kt * 16 + 4 * q + i, nt * 16 + r);
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, B, acc, 0, 0, 0);
            }
            _Pragma("unroll") for (int i = 0; i < 4; i++) out(mt * 16 + 4 * q + i, nt * 16 + r, acc[i]);
        }
}

// ------------------------------------------------------------- encoding --
// 6 raw inputs in [0, 1] (position normalised by the scene box, normal mapped
// from [-1, 1]) followed by sin and cos of 2^j * pi * x, j = 0..7 (in turns:
// x * 2^(j-1)), padded with zeros to kGIn.
__device__ __forceinline__ void encode_row(const GuideNet &g, V3 p, V3 n, uint16_t *row) {
    float s[6] = {fmul(fsub(p.x, g.lo[0]), g.inv[0]), fmul(fsub(p.y, g.lo[1]), g.inv[1]),
                  fmul(fsub(p.z, g.lo[2]), g.inv[2]), fmul(fadd(n.x, 1.f), 0.5f),
                  fmul(fadd(n.y, 1.f), 0.5f),       fmul(fadd(n.z, 1.f), 0.5f)};
    _Pragma("unroll") for (int d = 0; d < 6; d++) row[d] = f32_to_bf16(s[d]);
    float f = 0.5f;
    for (int j = 0; j < 8; j++) {
        _Pragma("unroll") for (int d = 0; d < 6; d++) {
            float sn, cs;
            fm_sincos_turns(fmul(s[d], f), sn, cs);
            row[6 + 12 * j + 2 * d] = f32_to_bf16(sn);
            row[6 + 12 * j + 2 * d + 1] = f32_to_bf16(cs);
        }
        f = fmul(f, 2.f);
    }
    for (int k = kGInUsed; k < kGIn; k++) row[k] = 0;
}

// LDS working set of one wave (64 rows) and the block's copy of the weights.
// Rows are padded by 4 BF16 so the 16 rows a fragment load touches fall in
// different banks; every fragment is one 8-byte load (4 consecutive k).
constexpr int kLdX = kGIn + 4, kLdH = kGHid + 4;
struct WaveMem {
    alignas(16) uint16_t x[64 * kLdX];   // encoded inputs (BF16)
    alignas(16) uint16_t h1[64 * kLdH];  // relu(layer 1) (BF16)
    alignas(16) uint16_t h2[64 * kLdH];  // relu(layer 2) (BF16)
    float o[64 * kGOutPad];              // layer 3 outputs (FP32)
};
struct WaveGrad {
    uint16_t dout[64 * kGOutPad];  // dLoss/dOutput (BF16)
    uint16_t da2[64 * kGHid];      // dLoss/d(pre-activation 2)
    uint16_t da1[64 * kGHid];      // dLoss/d(pre-activation 1)
};
struct NetLds {  // transposed weights [out][in]
    alignas(16) uint16_t w1t[kGHid * kLdX];
    alignas(16) uint16_t w2t[kGHid * kLdH];
    alignas(16) uint16_t w3t[kGOutPad * kLdH];
};

// Block-cooperative copy of the transposed BF16 weights into LDS.
__device__ __forceinline__ void load_weights(const GuideNet &g, NetLds &W) {
    const uint16_t *t1 = g.wb + (kGW1 + kGW2 + kGW3);
    const uint16_t *t2 = t1 + kGW1, *t3 = t2 + kGW2;
    for (int i = threadIdx.x; i < kGHid * kGIn / 4; i += blockDim.x) {
        const int r = (4 * i) / kGIn, c = (4 * i) % kGIn;
        *(uint2 *)(W.w1t + r * kLdX + c) = *(const uint2 *)(t1 + 4 * i);
    }
    for (int i = threadIdx.x; i < kGHid * kGHid / 4; i += blockDim.x) {
        const int r = (4 * i) / kGHid, c = (4 * i) % kGHid;
        *(uint2 *)(W.w2t + r * kLdH + c) = *(const uint2 *)(t2 + 4 * i);
    }
    for (int i = threadIdx.x; i < kGOutPad * kGHid / 4; i += blockDim.x) {
        const int r = (4 * i) / kGHid, c = (4 * i) % kGHid;
        *(uint2 *)(W.w3t + r * kLdH + c) = *(const uint2 *)(t3 + 4 * i);
    }
    __syncthreads();
}

// C[M x N] = A[M x K] * B[K x N] with A row-major in LDS (leading dimension
// lda) and B given transposed, Bt[n][k] in LDS (ldb): both fragments are
// single 8-byte LDS loads.
template <int M, int N, int K, class FC>
__device__ __forceinline__ void wave_gemm_lds(const uint16_t *A, int lda, const uint16_t *Bt, int ldb, FC out) {
    const int l = (int)__lane_id();
    const int r = l & 15, q = l >> 4;
    for (int mt = 0; mt < M / 16; mt++)
        for (int nt = 0; nt < N / 16; nt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            _Pragma("unroll") for (int kt = 0; kt < K / 16; kt++) {
                const bf16x4 a = *(const bf16x4 *)(A + (mt * 16 + r) * lda + kt * 16 + 4 * q);
                const bf16x4 b = *(const bf16x4 *)(Bt + (nt * 16 + r) * ldb + kt * 16 + 4 * q);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
            }
            _Pragma("unroll") for (int i = 0; i < 4; i++) out(mt * 16 + 4 * q + i, nt * 16 + r, acc[i]);
        }
}

// Forward pass of the wave's 64 rows (x already in LDS).
__device__ __forceinline__ void net_forward(const GuideNet &g, const NetLds &W, WaveMem &M) {
    const float *b1 = g.w + kGOffB1, *b2 = g.w + kGOffB2, *b3 = g.w + kGOffB3;
    wave_gemm_lds<64, kGHid, kGIn>(M.x, kLdX, W.w1t, kLdX, [&](int m, int n, float v) {
        M.h1[m * kLdH + n] = f32_to_bf16(smax(0.f, fadd(v, b1[n])));
    });
    __syncthreads();
    wave_gemm_lds<64, kGHid, kGHid>(M.h1, kLdH, W.w2t, kLdH, [&](int m, int n, float v) {
        M.h2[m * kLdH + n] = f32_to_bf16(smax(0.f, fadd(v, b2[n])));
    });
    __syncthreads();
    wave_gemm_lds<64, kGOutPad, kGHid>(M.h2, kLdH, W.w3t, kLdH, [&](int m, int n, float v) {
        M.o[m * kGOutPad + n] = fadd(v, b3[n]);
    });
    __syncthreads();
}

// ------------------------------------------------------------ mixture ----
struct Lobe {
    V3 mu;
    float kappa, e2, len;  // e2 = exp(-2 kappa); len = |unnormalised mean|
};
__device__ __forceinline__ Lobe lobe_of(const float *o, int k) {
    Lobe L;
    V3 m = v3(o[5 * k], o[5 * k + 1], o[5 * k + 2]);
    L.len = sqrtf(dot(m, m));
    L.mu = L.len > 1e-12f ? mul(m, 1.f / L.len) : v3(0.f, 0.f, 1.f);
    L.kappa = expf(fminf(fmaxf(o[5 * k + 3], kGLogKappaMin), kGLogKappaMax));
    L.e2 = expf(-2.f * L.kappa);
    return L;
}
// log of the vMF density on the sphere: kappa / (2 pi (1 - e^-2k)) e^{k (mu.w - 1)}
__device__ __forceinline__ float vmf_log_pdf(const Lobe &L, V3 w) {
    return logf(L.kappa) - 1.8378770664093453f - logf(1.f - L.e2) + L.kappa * (dot(L.mu, w) - 1.f);
}
// weights: softmax of the logits
__device__ __forceinline__ void mix_weights(const float *o, float *pi) {
    float amax = o[4];
    for (int k = 1; k < kGLobes; k++) amax = fmaxf(amax, o[5 * k + 4]);
    float s = 0.f;
    for (int k = 0; k < kGLobes; k++) {
        pi[k] = expf(o[5 * k + 4] - amax);
        s += pi[k];
    }
    for (int k = 0; k < kGLobes; k++) pi[k] /= s;
}
__device__ __forceinline__ float mix_pdf(const float *o, const float *pi, V3 w) {
    float p = 0.f;
    for (int k = 0; k < kGLobes; k++) p += pi[k] * expf(vmf_log_pdf(lobe_of(o, k), w));
    return p;
}
// vMF sample about mu (Wood 1994): cos(theta) = 1 + log(u1 + (1 - u1) e^-2k) / k
__device__ __forceinline__ V3 vmf_sample(const Lobe &L, float u1, float u2) {
    float c = 1.f + logf(u1 + (1.f - u1) * L.e2) / L.kappa;
    c = fminf(1.f, fmaxf(-1.f, c));
    float s = sqrtf(fmaxf(0.f, 1.f - c * c));
    V3 helper = fabsf(L.mu.x) <= fabsf(L.mu.y) ? v3(1.f, 0.f, 0.f) : v3(0.f, 1.f, 0.f);
    V3 T = normalize(cross(L.mu, helper));
    V3 B = cross(L.mu, T);
    float sn, cs;
    fm_sincos_turns(u2, sn, cs);
    return add(add(mul(T, s * cs), mul(B, s * sn)), mul(L.mu, c));
}

__device__ __forceinline__ uint32_t record_hash(uint32_t path, uint32_t bounce) {
    uint32_t h = path * 0x9E3779B1u ^ (bounce + 1u) * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return h >> 8;  // 24 bits
}

// --------------------------------------------------------------- k_guide --
// Draws the pending directions of the queue: the MLP runs on 64 rays per wave
// (rays without a pending direction ride along and are left untouched).
// Draw order from the path's stream: select, lobe, u1, u2.
__global__ __launch_bounds__(64) void k_guide(GuideArgs a) {
    __shared__ WaveMem M;
    __shared__ NetLds W;
    const uint32_t n = min(*a.count, a.cap);
    if (blockIdx.x * 64u >= n) return;
    load_weights(a.net, W);
    const int lane = (int)__lane_id();
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t i = base + (uint32_t)lane;
        PathState ps;
        bool pend = false;
        if (i < n) {
            st_load(a.q, i, ps);
            pend = (ps.meta & ST_GUIDE) != 0;
        }
        if (!__ballot(pend)) continue;  // uniform: the whole wave skips
        encode_row(a.net, pend ? ps.o : v3(0.f, 0.f, 0.f), pend ? ps.d : v3(0.f, 0.f, 1.f), M.x + lane * kLdX);
        __syncthreads();
        net_forward(a.net, W, M);
        if (pend) {
            const float *o = M.o + lane * kGOutPad;
            bool net_ok = true;
            for (int k = 0; k < kGOut; k++) net_ok = net_ok && isfinite(o[k]);
            float pi[kGLobes];
            mix_weights(o, pi);
            const V3 N = ps.d;
            const float u_sel = rng_next(ps.rng), u_lobe = rng_next(ps.rng);
            const float u1 = rng_next(ps.rng), u2 = rng_next(ps.rng);
            V3 w;
            float pdf;
            if (!net_ok) {  // fallback: plain cosine sampling
                w = fm_hemisphere(N, u1, u2);
                pdf = fmaxf(dot(N, w), 0.f) * 0.31830988618379067f;
            } else {
                if (u_sel < 0.5f) {
                    w = fm_hemisphere(N, u1, u2);
                } else {
                    int k = 0;
                    float cum = pi[0];
                    while (k < kGLobes - 1 && u_lobe >= cum) cum += pi[++k];
                    w = vmf_sample(lobe_of(o, k), u1, u2);
                }
                pdf = 0.5f * fmaxf(dot(N, w), 0.f) * 0.31830988618379067f + 0.5f * mix_pdf(o, pi, w);
            }
            const float cosw = dot(N, w);
            const int nch = st_nch(ps.meta), ch0 = st_ch0(ps.meta);
            ps.meta &= ~ST_GUIDE;
            if (cosw <= 0.f || !(pdf > 0.f)) {
                _Pragma("unroll") for (int c = 0; c < 4; c++) ps.tp[c] = 0.f;
                ps.meta |= ST_DEAD;
            } else {
                // cosine BSDF sample weight (the albedo is already applied):
                // (cos / pi) / pdf, which is 1 for pure cosine sampling
                const float r = cosw * 0.31830988618379067f / pdf;
                _Pragma("unroll") for (int c = 0; c < 4; c++) if (c < nch) ps.tp[c] *= r;
                if (record_hash(ps.path, (uint32_t)st_bounce(ps.meta)) < a.keep) {
                    uint32_t slot = atomicAdd(a.net.n_rec, 1u);
                    if (slot < a.rec_cap) {
                        GuideRecord &R = a.rec[slot];
                        R.o[0] = ps.o.x; R.o[1] = ps.o.y; R.o[2] = ps.o.z;
                        R.n[0] = N.x; R.n[1] = N.y; R.n[2] = N.z;
                        R.w[0] = w.x; R.w[1] = w.y; R.w[2] = w.z;
                        R.pdf = pdf;
                        R.path = ps.path;
                        R.chan = (uint32_t)ch0 | ((uint32_t)nch << 8);
                        _Pragma("unroll") for (int c = 0; c < 4; c++) {
                            R.rad[c] = ps.rad[c];
                            R.tp[c] = ps.tp[c];
                        }
                        R.weight = 0.f;
                    }
                }
            }
            ps.d = w;
            st_store(a.q, i, ps);
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------- training ----
// Training weight of each record once its path has finished: the incident
// radiance along w, (final - gathered before) / throughput averaged over the
// record's channels, times cos / pdf: a sample of the target density
// proportional to L_i(w) cos(w) under the sampling density.  Records are
// split into S equal steps; wsum[s] collects each step's weight total.
__global__ void k_guide_prep(GuideNet g, GuideRecord *rec, uint32_t rec_cap, const float4 *radiance, int S) {
    const uint32_t n = min(*g.n_rec, rec_cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    // whole waves iterate together so the per-step sums reduce across the wave
    for (uint32_t r0 = blockIdx.x * blockDim.x; r0 < n; r0 += stride) {
        const uint32_t r = r0 + threadIdx.x;
        float wt = 0.f;
        int s = 0;
        if (r < n) {
        GuideRecord &R = rec[r];
        const float4 f4 = radiance[R.path];
        const float fin[4] = {f4.x, f4.y, f4.z, f4.w};
        const int ch0 = (int)(R.chan & 0xFFu), nch = (int)(R.chan >> 8);
        float L = 0.f;
        int cnt = 0;
        for (int c = 0; c < nch && ch0 + c < 4; c++)
            if (R.tp[c] > 0.f) {
                L += (fin[ch0 + c] - R.rad[c]) / R.tp[c];
                cnt++;
            }
        L = cnt ? fmaxf(L / (float)cnt, 0.f) : 0.f;
        const float cosw = R.n[0] * R.w[0] + R.n[1] * R.w[1] + R.n[2] * R.w[2];
        wt = L * fmaxf(cosw, 0.f) / R.pdf;
        wt = isfinite(wt) ? wt : 0.f;
        R.weight = wt;
        s = (int)((uint64_t)r * (uint64_t)S / n);
        }
        const int s0 = __shfl(s, 0, 64);
        if (__ballot(s != s0 && r < n) == 0) {  // one step in the wave: reduce, one atomic
            float t = wt;
            for (int dd = 32; dd >= 1; dd >>= 1) t += __shfl_xor(t, dd, 64);
            if (__lane_id() == 0 && t != 0.f) atomicAdd(g.wsum + 2 * s0, t);
        } else if (r < n && wt != 0.f) {
            atomicAdd(g.wsum + 2 * s, wt);
        }
    }
}

// dLoss/dOutput of one record, loss = -scale * log q(w) with q the learned vMF
// mixture; returns the record's loss.
__device__ __forceinline__ float output_grad(const float *o, V3 w, float scale, float *d) {
    float pi[kGLobes], t[kGLobes];
    mix_weights(o, pi);
    Lobe L[kGLobes];
    float tmax = -3.0e38f;
    for (int k = 0; k < kGLobes; k++) {
        L[k] = lobe_of(o, k);
        t[k] = logf(fmaxf(pi[k], 1e-30f)) + vmf_log_pdf(L[k], w);
        tmax = fmaxf(tmax, t[k]);
    }
    float s = 0.f;
    for (int k = 0; k < kGLobes; k++) s += expf(t[k] - tmax);
    const float logq = tmax + logf(s);
    for (int k = 0; k < kGLobes; k++) {
        const float gam = expf(t[k] - logq);
        const float c = dot(L[k].mu, w);
        const float rk = o[5 * k + 3];
        const bool live = rk > kGLogKappaMin && rk < kGLogKappaMax;
        const float kap = L[k].kappa;
        const float dr = live ? gam * kap * (1.f / kap - 2.f * L[k].e2 / (1.f - L[k].e2) + (c - 1.f)) : 0.f;
        const float dm = L[k].len > 1e-12f ? gam * kap / L[k].len : 0.f;
        d[5 * k + 0] = -scale * dm * (w.x - c * L[k].mu.x);
        d[5 * k + 1] = -scale * dm * (w.y - c * L[k].mu.y);
        d[5 * k + 2] = -scale * dm * (w.z - c * L[k].mu.z);
        d[5 * k + 3] = -scale * dr;
        d[5 * k + 4] = -scale * (gam - pi[k]);
    }
    for (int k = kGOut; k < kGOutPad; k++) d[k] = 0.f;
    return -scale * logq;
}

// One training step's gradient: records [s n / S, (s + 1) n / S), each
// weighted by weight / wsum[s].  Forward and backward as MFMA tiles; weight
// gradients are added to g.grad.
__global__ __launch_bounds__(64) void k_guide_grad(GuideNet g, const GuideRecord *rec, uint32_t rec_cap, int s,
                                                   int S) {
    __shared__ WaveMem M;
    __shared__ WaveGrad D;
    __shared__ NetLds Wl;
    const uint32_t n = min(*g.n_rec, rec_cap);
    const uint32_t r0 = (uint32_t)((uint64_t)s * n / S), r1 = (uint32_t)((uint64_t)(s + 1) * n / S);
    const float W = g.wsum[2 * s];
    if (!(W > 0.f) || r0 + blockIdx.x * 64u >= r1) return;
    load_weights(g, Wl);
    const int lane = (int)__lane_id();
    const uint16_t *w2 = g.wb + kGW1, *w3 = w2 + kGW2;  // [in][out] layouts
    float *gw1 = g.grad + kGOffW1, *gb1 = g.grad + kGOffB1, *gw2 = g.grad + kGOffW2, *gb2 = g.grad + kGOffB2;
    float *gw3 = g.grad + kGOffW3, *gb3 = g.grad + kGOffB3;
    for (uint32_t base = r0 + blockIdx.x * 64u; base < r1; base += gridDim.x * 64u) {
        const uint32_t r = base + (uint32_t)lane;
        const bool ok = r < r1 && rec[r].weight > 0.f;
        GuideRecord R = rec[ok ? r : r0];
        encode_row(g, v3(R.o[0], R.o[1], R.o[2]), v3(R.n[0], R.n[1], R.n[2]), M.x + lane * kLdX);
        __syncthreads();
        net_forward(g, Wl, M);
        float d[kGOutPad];
        if (ok) {
            const float loss = output_grad(M.o + lane * kGOutPad, v3(R.w[0], R.w[1], R.w[2]), R.weight / W, d);
            atomicAdd(g.wsum + 2 * s + 1, loss);
        } else {
            for (int k = 0; k < kGOutPad; k++) d[k] = 0.f;
        }
        for (int k = 0; k < kGOutPad; k++) D.dout[lane * kGOutPad + k] = f32_to_bf16(d[k]);
        __syncthreads();
        // dA2 = (dOut W3^T) masked by relu(h2); W3 is [hid][outpad]
        wave_gemm<64, kGHid, kGOutPad>([&](int m, int k) { return D.dout[m * kGOutPad + k]; },
                                       [&](int k, int nn) { return w3[nn * kGOutPad + k]; },
                                       [&](int m, int nn, float v) {
                                           D.da2[m * kGHid + nn] = M.h2[m * kLdH + nn] ? f32_to_bf16(v) : (uint16_t)0;
                                       });
        // dW3 = h2^T dOut
        wave_gemm<kGHid, kGOutPad, 64>([&](int m, int k) { return M.h2[k * kLdH + m]; },
                                       [&](int k, int nn) { return D.dout[k * kGOutPad + nn]; },
                                       [&](int m, int nn, float v) { if (v != 0.f) atomicAdd(gw3 + m * kGOutPad + nn, v); });
        __syncthreads();
        // dA1 = (dA2 W2^T) masked by relu(h1); W2 is [hid][hid]
        wave_gemm<64, kGHid, kGHid>([&](int m, int k) { return D.da2[m * kGHid + k]; },
                                    [&](int k, int nn) { return w2[nn * kGHid + k]; },
                                    [&](int m, int nn, float v) {
                                        D.da1[m * kGHid + nn] = M.h1[m * kLdH + nn] ? f32_to_bf16(v) : (uint16_t)0;
                                    });
        // dW2 = h1^T dA2
        wave_gemm<kGHid, kGHid, 64>([&](int m, int k) { return M.h1[k * kLdH + m]; },
                                    [&](int k, int nn) { return D.da2[k * kGHid + nn]; },
                                    [&](int m, int nn, float v) { if (v != 0.f) atomicAdd(gw2 + m * kGHid + nn, v); });
        __syncthreads();
        // dW1 = x^T dA1
        wave_gemm<kGIn, kGHid, 64>([&](int m, int k) { return M.x[k * kLdX + m]; },
                                   [&](int k, int nn) { return D.da1[k * kGHid + nn]; },
                                   [&](int m, int nn, float v) { if (v != 0.f) atomicAdd(gw1 + m * kGHid + nn, v); });
        // biases: column sums of the (BF16) output and pre-activation gradients
        {
            float b3 = 0.f, b2 = 0.f, b1 = 0.f;
            for (int m = 0; m < 64; m++) {
                if (lane < kGOutPad) b3 += bf16_to_f32(D.dout[m * kGOutPad + lane]);
                b2 += bf16_to_f32(D.da2[m * kGHid + lane]);
                b1 += bf16_to_f32(D.da1[m * kGHid + lane]);
            }
            if (lane < kGOutPad && b3 != 0.f) atomicAdd(gb3 + lane, b3);
            if (b2 != 0.f) atomicAdd(gb2 + lane, b2);
            if (b1 != 0.f) atomicAdd(gb1 + lane, b1);
        }
        __syncthreads();
    }
}

// BF16 copies of the weight matrices, [in][out] and transposed [out][in].
__device__ __forceinline__ void write_bf16(const GuideNet &g, int p, float v) {
    const uint16_t b = f32_to_bf16(v);
    uint16_t *wt = g.wb + (kGW1 + kGW2 + kGW3);
    if (p < kGOffB1) {
        const int i = p / kGHid, o = p % kGHid;
        g.wb[p] = b;
        wt[o * kGIn + i] = b;
    } else if (p >= kGOffW2 && p < kGOffB2) {
        const int q = p - kGOffW2, i = q / kGHid, o = q % kGHid;
        g.wb[kGW1 + q] = b;
        wt[kGW1 + o * kGHid + i] = b;
    } else if (p >= kGOffW3 && p < kGOffB3) {
        const int q = p - kGOffW3, i = q / kGOutPad, o = q % kGOutPad;
        g.wb[kGW1 + kGW2 + q] = b;
        wt[kGW1 + kGW2 + o * kGHid + i] = b;
    }
}

// Adam step t (1-based) on every parameter; clears the gradient.  The padded
// output columns of W3/b3 stay zero.
__global__ void k_guide_adam(GuideNet g, GuideTrain h, int t) {
    const float c1 = 1.f - powf(h.beta1, (float)t), c2 = 1.f - powf(h.beta2, (float)t);
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < kGParams; p += gridDim.x * blockDim.x) {
        bool pad = (p >= kGOffW3 && p < kGOffB3 && (p - kGOffW3) % kGOutPad >= kGOut) ||
                   (p >= kGOffB3 && p - kGOffB3 >= kGOut);
        const float gr = pad ? 0.f : g.grad[p];
        const float m = h.beta1 * g.m[p] + (1.f - h.beta1) * gr;
        const float v = h.beta2 * g.v[p] + (1.f - h.beta2) * gr * gr;
        g.m[p] = m;
        g.v[p] = v;
        const float w = g.w[p] - h.lr * (m / c1) / (sqrtf(v / c2) + h.eps);
        g.w[p] = w;
        g.grad[p] = 0.f;
        write_bf16(g, p, w);
    }
}

// BF16 copies from the FP32 weights (after an upload).
__global__ void k_guide_sync(GuideNet g) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < kGParams; p += gridDim.x * blockDim.x)
        write_bf16(g, p, g.w[p]);
}

// Forward pass on given vertices (verification entry): out[n][kGOut].
__global__ __launch_bounds__(64) void k_guide_forward(GuideNet g, const float *pn, uint32_t n, float *out) {
    __shared__ WaveMem M;
    __shared__ NetLds W;
    if (blockIdx.x * 64u >= n) return;
    load_weights(g, W);
    const int lane = (int)__lane_id();
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t i = base + (uint32_t)lane;
        const uint32_t j = i < n ? i : 0u;
        encode_row(g, v3(pn[6 * j], pn[6 * j + 1], pn[6 * j + 2]), v3(pn[6 * j + 3], pn[6 * j + 4], pn[6 * j + 5]),
                   M.x + lane * kLdX);
        __syncthreads();
        net_forward(g, W, M);
        if (i < n)
            for (int k = 0; k < kGOut; k++) out[(size_t)i * kGOut + k] = M.o[lane * kGOutPad + k];
        __syncthreads();
    }
}

}  // namespace pcr

// ------------------------------------------------------------- launchers --
using namespace pcr;
extern "C" hipError_t pcr_launch_guide(const GuideArgs *a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_guide, dim3(grid), dim3(64), 0, s, *a);
    return hipGetLastError();
}
extern "C" hipError_t pcr_launch_guide_prep(const GuideNet *g, GuideRecord *rec, uint32_t rec_cap,
                                            const float4 *radiance, int S, hipStream_t s) {
    hipLaunchKernelGGL(k_guide_prep, dim3(256), dim3(256), 0, s, *g, rec, rec_cap, radiance, S);
    return hipGetLastError();
}
extern "C" hipError_t pcr_launch_guide_grad(const GuideNet *g, const GuideRecord *rec, uint32_t rec_cap, int step,
                                            int S, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_guide_grad, dim3(grid), dim3(64), 0, s, *g, rec, rec_cap, step, S);
    return hipGetLastError();
}
extern "C" hipError_t pcr_launch_guide_adam(const GuideNet *g, const GuideTrain *h, int t, hipStream_t s) {
    hipLaunchKernelGGL(k_guide_adam, dim3((kGParams + 255) / 256), dim3(256), 0, s, *g, *h, t);
    return hipGetLastError();
}
extern "C" hipError_t pcr_launch_guide_sync(const GuideNet *g, hipStream_t s) {
    hipLaunchKernelGGL(k_guide_sync, dim3((kGParams + 255) / 256), dim3(256), 0, s, *g);
    return hipGetLastError();
}
extern "C" hipError_t pcr_launch_guide_forward(const GuideNet *g, const float *pn, uint32_t n, float *out, int grid,
                                               hipStream_t s) {
    hipLaunchKernelGGL(k_guide_forward, dim3(grid), dim3(64), 0, s, *g, pn, n, out);
    return hipGetLastError();
}
