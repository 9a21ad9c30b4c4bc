// guide.hpp - learned directional proposal: MLP layout, bf16/MFMA helpers and the von Mises-Fisher
// mixture the network predicts.
//
// Network: 9 inputs (position in scene bounds, outgoing direction, shading normal, all in [0,1])
// -> frequency encoding (raw inputs + sin/cos of 2*pi*frac(x*f_k), f_k = 2^(k/2), k < 16, half-octave
// steps from 1 to 181 cycles over the input range) = 297 features padded
// to 304 -> 4 hidden layers of 64 (ReLU) -> 48 outputs, of which 40 parameterise 8 lobes
// (logit, unnormalised mean direction, log concentration).  Weights are bf16, accumulation is fp32
// on MFMA (v_mfma_f32_16x16x16_bf16), biases stay fp32.  The renderer has no counterpart in
// kennyalive/YAR; the estimator it plugs into is trace_path (src/ref/path_tracing.cpp:86-104).
#pragma once

#include "device.hpp"
#include "../../../include/yar_gpu.h"

namespace yg {

constexpr int kGIn = 9, kGFreq = 16, kGEnc = 297, kGK1 = 304, kGH = 64;
constexpr int kGLobes = 8, kGOut = 40, kGOutPad = 48;
constexpr float kGSMin = -3.f, kGSMax = 7.f;       // log-concentration clamp

// Canonical parameter order (fp32 master copy): W1 [304][64], b1, W2 [64][64], b2, W3, b3, W4, b4,
// W5 [64][48], b5.  W1 rows >= 297 and W5/b5 columns >= 40 are padding and stay zero.
constexpr int kOffW[5] = {0, 19520, 23680, 27840, 32000};
constexpr int kOffB[5] = {19456, 23616, 27776, 31936, 35072};
constexpr int kLayerK[5] = {kGK1, kGH, kGH, kGH, kGH};
constexpr int kLayerN[5] = {kGH, kGH, kGH, kGH, kGOutPad};
constexpr int kGParams = 35120;
// bf16 copies: forward layout (per layer [N][K], K contiguous) and canonical layout ([K][N])
constexpr int kOffT[5] = {0, 19456, 23552, 27648, 31744};
constexpr int kGWeights = 34816;

struct GuideNet {
    const uint16_t* wt;      // forward layout, kGWeights bf16
    const uint16_t* wr;      // canonical layout, kGWeights bf16 (offsets kOffT as well)
    const float* master;     // fp32 master parameters (biases are read from here)
};

// Training activations of one batch, blocked by four samples so an MFMA lane reads the four
// consecutive samples of one feature with one 8-byte load: element (m, k) of a [n][K] matrix sits
// at ((m / 4) * K + k) * 4 + m % 4.
struct GuideActs {
    uint16_t* h[5];          // h0 (K 304) .. h4 (K 64): layer inputs
    uint16_t* d[5];          // d1 .. d5: loss gradients w.r.t. pre-activations (N 64, d5 N 48)
    float* o;                // network outputs [n][48]
};

// One training record (64 bytes): inputs, loss weight, sampled direction.
struct GuideRecord {
    float x[9];
    float wt;
    float w[3];
    float pad[3];
};

YG_INLINE uint16_t f2bf(float f) {                 // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
YG_INLINE float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

typedef short mfma_ab __attribute__((ext_vector_type(4)));
typedef float mfma_cd __attribute__((ext_vector_type(4)));

// D = A (16x16) * B (16x16) + C.  Lane l holds A[l%16][4*(l/16)+i], B[4*(l/16)+i][l%16] and
// C/D[4*(l/16)+i][l%16] for i = 0..3.
YG_INLINE mfma_cd mfma16(mfma_ab a, mfma_ab b, mfma_cd c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
YG_INLINE mfma_ab ld_ab(const uint16_t* p) { return *reinterpret_cast<const mfma_ab*>(p); }

// ---- von Mises-Fisher mixture ------------------------------------------------------------
struct VmfLobe { V3 mu; float kappa, pi, norm; };   // norm = kappa / (2 pi (1 - exp(-2 kappa)))

YG_INLINE void mixture_from_outputs(const float* o, VmfLobe* L) {
    float mx = o[0];
    for (int j = 1; j < kGLobes; j++) mx = fmaxf(mx, o[5 * j]);
    float sum = 0.f;
    for (int j = 0; j < kGLobes; j++) {
        const float e = expf(o[5 * j] - mx);
        L[j].pi = e;
        sum += e;
        const V3 r = v3(o[5 * j + 1], o[5 * j + 2], o[5 * j + 3]);
        const float rn = fmaxf(length(r), 1e-12f);
        L[j].mu = vdiv(r, rn);
        const float s = fminf(fmaxf(o[5 * j + 4], kGSMin), kGSMax);
        const float k = expf(s);
        L[j].kappa = k;
        L[j].norm = k / (2.f * kPi * -expm1f(-2.f * k));
    }
    const float inv = 1.f / sum;
    for (int j = 0; j < kGLobes; j++) L[j].pi *= inv;
}

YG_INLINE float mixture_pdf(const VmfLobe* L, V3 w) {
    float p = 0.f;
    for (int j = 0; j < kGLobes; j++) p += L[j].pi * L[j].norm * expf(L[j].kappa * (dot(L[j].mu, w) - 1.f));
    return p;
}

// Orthonormal basis around unit n (Duff et al. 2017, "Building an Orthonormal Basis, Revisited").
YG_INLINE void onb(V3 n, V3& t, V3& b) {
    const float s = n.z >= 0.f ? 1.f : -1.f;
    const float a = -1.f / (s + n.z);
    const float bb = n.x * n.y * a;
    t = v3(1.f + s * n.x * n.x * a, s * bb, -s * n.x);
    b = v3(bb, s + n.y * n.y * a, -n.y);
}

// Picks a lobe with u_lobe and inverts the vMF CDF with (u1, u2).  The lobe loop is unrolled with
// selects so the mixture stays in registers.
YG_INLINE V3 mixture_sample(const VmfLobe* L, float u_lobe, float u1, float u2) {
    V3 mu = L[kGLobes - 1].mu;
    float k = L[kGLobes - 1].kappa;
    float c = 0.f;
    bool chosen = false;
#pragma unroll
    for (int j = 0; j < kGLobes - 1; j++) {
        c += L[j].pi;
        const bool take = !chosen && u_lobe < c;
        mu = take ? L[j].mu : mu;
        k = take ? L[j].kappa : k;
        chosen = chosen || take;
    }
    float wz = 1.f + log1pf(u1 * expm1f(-2.f * k)) / k;
    wz = fminf(fmaxf(wz, -1.f), 1.f);
    const float r = sqrtf(fmaxf(0.f, 1.f - wz * wz));
    const float phi = 2.f * kPi * u2;
    V3 t, b;
    onb(mu, t, b);
    return wz * mu + (r * cosf(phi)) * t + (r * sinf(phi)) * b;
}

// Network inputs at a shading point.
YG_INLINE void guide_inputs(V3 p, V3 lo, V3 inv_ext, V3 wo, V3 n, float* x) {
    const float v[9] = {(p.x - lo.x) * inv_ext.x, (p.y - lo.y) * inv_ext.y, (p.z - lo.z) * inv_ext.z,
                        0.5f * (wo.x + 1.f), 0.5f * (wo.y + 1.f), 0.5f * (wo.z + 1.f),
                        0.5f * (n.x + 1.f), 0.5f * (n.y + 1.f), 0.5f * (n.z + 1.f)};
    for (int i = 0; i < 9; i++) x[i] = fminf(fmaxf(v[i], 0.f), 1.f);
}

// ---- launches (guide.hip) ---------------------------------------------------------------
// Forward pass for n queries read as records of `in_stride` floats (inputs at offset 0); writes
// out[m * out_stride + j] for j < n_write.  With acts != nullptr the activations are kept for
// a training step.
void guide_forward(const GuideNet& net, const float* in, int in_stride, long long n, float* out, int out_stride,
                   int n_write, const GuideActs* acts, int grid, hipStream_t s);
// Same with the query count read from device memory (graph-capturable, persistent grid).
void guide_forward_dev(const GuideNet& net, const float* in, int in_stride, const uint32_t* n_dev, float* out,
                       int out_stride, int grid, hipStream_t s);
void guide_prepare_device();
size_t guide_forward_lds();
void guide_backward(const GuideNet& net, const GuideRecord* rec, long long n, float inv_batch, const GuideActs& acts,
                    float* loss_sum, hipStream_t s);
void guide_weight_grads(const GuideActs& acts, long long n, int slice, float* partial, hipStream_t s);
void guide_apply(float* master, float* m, float* v, const float* partial, int n_slices, float lr, float b1, float b2,
                 float eps, float bc1, float bc2, uint16_t* wt, uint16_t* wr, float* grad_out, hipStream_t s);
void guide_sync_weights(const float* master, uint16_t* wt, uint16_t* wr, hipStream_t s);
// Render-context access to a yar_guide (guide.hip): its network, configuration and device, and
// `steps` Adam steps on device-resident records, launched on the caller's stream.
GuideNet guide_net(yar_guide* g);
const yar_guide_config& guide_config(const yar_guide* g);
int guide_device(const yar_guide* g);
// Returns the last batch's mean log-likelihood gain over a uniform density (nats per unit weight).
float guide_train_on(yar_guide* g, const GuideRecord* rec, long long count, int steps, hipStream_t s);
void guide_gather(const GuideRecord* all, long long count, long long batch, uint32_t seed, uint32_t step,
                  GuideRecord* out, hipStream_t s);

}  // namespace yg
