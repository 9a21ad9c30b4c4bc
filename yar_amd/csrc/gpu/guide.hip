// guide.hip - MFMA kernels of the learned directional proposal (forward, loss gradient, backward,
// weight gradients, Adam) and the yar_guide_* C ABI.
//
// Layouts: weights live in HBM as bf16 in two layouts (forward: [N][K] per layer so an MFMA lane
// reads four consecutive K of one output column; canonical [K][N] for the backward product) next
// to the fp32 master copy that Adam updates.  A block stages all weights it needs in LDS (forward
// 73 KB, backward 34 KB) and runs 4 waves in lockstep, each wave on a 16-row tile.
#include "guide.hpp"

#include <new>
#include "../../../include/yar_gpu.h"
#include "../host/common.hpp"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace yg {

// ---------------------------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------------------------
constexpr int kFwdStride[5] = {kGK1 + 8, kGH + 8, kGH + 8, kGH + 8, kGH + 8};
constexpr int kFwdRows[5] = {kGH, kGH, kGH, kGH, kGOutPad};
constexpr int kFwdOff[5] = {0, 19968, 24576, 29184, 33792};     // bf16 elements
constexpr int kFwdW = 37248;
constexpr int kEStride = kGK1 + 8, kHStride = kGH + 8;
constexpr int kWaveBuf = 16 * kEStride + 2 * 16 * kHStride;      // bf16 elements per wave
constexpr size_t kFwdLds = (size_t)kFwdW * 2 + 304 * 4 + (size_t)4 * kWaveBuf * 2;

size_t guide_forward_lds() { return kFwdLds; }

YG_INLINE void copy_rows(uint16_t* dst, int dst_stride, const uint16_t* src, int rows, int cols) {
    const int chunks = rows * (cols / 8);
    for (int c = threadIdx.x; c < chunks; c += blockDim.x) {
        const int row = c / (cols / 8), c8 = c % (cols / 8);
        *reinterpret_cast<uint4*>(dst + row * dst_stride + c8 * 8) =
            *reinterpret_cast<const uint4*>(src + (size_t)row * cols + c8 * 8);
    }
}

template <int KSTEPS, int NT>
YG_INLINE void tile_product(const uint16_t* A, int a_stride, const uint16_t* W, int w_stride, int r, int g,
                            mfma_cd* acc) {
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = mfma_cd{0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
    for (int ks = 0; ks < KSTEPS; ks++) {
        const mfma_ab a = ld_ab(A + r * a_stride + 16 * ks + 4 * g);
#pragma unroll
        for (int t = 0; t < NT; t++) acc[t] = mfma16(a, ld_ab(W + (16 * t + r) * w_stride + 16 * ks + 4 * g), acc[t]);
    }
}

// Bias + ReLU + bf16 of a hidden layer's 16x64 tile into the next layer's LDS input (and the
// training copy of that activation).
YG_INLINE void hidden_epilogue(const mfma_cd* acc, const float* bias, uint16_t* H, int r, int g, long long m0,
                               uint16_t* act) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int col = 16 * t + r;
        uint16_t hv[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            hv[v] = f2bf(fmaxf(acc[t][v] + bias[col], 0.f));
            H[(4 * g + v) * kHStride + col] = hv[v];
        }
        if (act) {
            const size_t idx = ((size_t)(m0 / 4 + g) * kGH + col) * 4;
            *reinterpret_cast<uint2*>(act + idx) =
                make_uint2((uint32_t)hv[0] | ((uint32_t)hv[1] << 16), (uint32_t)hv[2] | ((uint32_t)hv[3] << 16));
        }
    }
}

__global__ __launch_bounds__(256) void k_guide_forward(GuideNet net, const float* __restrict__ in, int in_stride,
                                                       long long n, const uint32_t* n_dev, float* __restrict__ out,
                                                       int out_stride, int n_write, GuideActs acts, int store) {
    extern __shared__ __align__(16) uint16_t lds[];
    if (n_dev) n = *n_dev;
    uint16_t* W = lds;
    float* bias = reinterpret_cast<float*>(lds + kFwdW);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    uint16_t* E = lds + kFwdW + 608 + wave * kWaveBuf;          // 304 fp32 of biases = 608 bf16 slots
    uint16_t* Ha = E + 16 * kEStride;
    uint16_t* Hb = Ha + 16 * kHStride;
    for (int l = 0; l < 5; l++) copy_rows(W + kFwdOff[l], kFwdStride[l], net.wt + kOffT[l], kFwdRows[l], kLayerK[l]);
    for (int i = threadIdx.x; i < 4 * kGH + kGOutPad; i += blockDim.x) {
        const int l = i / kGH, c = i % kGH;
        bias[i] = net.master[kOffB[l] + c];
    }
    __syncthreads();
    const long long n_btiles = (n + 63) / 64;
    for (long long bt = blockIdx.x; bt < n_btiles; bt += gridDim.x) {
        const long long m0 = bt * 64 + wave * 16;
        // frequency encoding of query m0 + r: lane group g fills features g, g+4, ... of its row
        {
            const long long m = m0 + r;
            float x[kGIn];
#pragma unroll
            for (int i = 0; i < kGIn; i++) x[i] = m < n ? fminf(fmaxf(in[m * in_stride + i], 0.f), 1.f) : 0.f;
            uint16_t* row = E + r * kEStride;
            uint16_t* act = store ? acts.h[0] + ((size_t)(m / 4) * kGK1) * 4 + (m & 3) : nullptr;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int j = g + 4 * t;
                const int f = j < kGIn ? j : kGEnc + (j - kGIn);
                float xv = 0.f;
#pragma unroll
                for (int i = 0; i < kGIn; i++) xv = (j == i) ? x[i] : xv;
                const uint16_t b = f2bf(xv);
                row[f] = b;
                if (act) act[(size_t)f * 4] = b;
            }
            for (int t = 0; t < 36; t++) {
                const int p = g + 4 * t, i = p >> 4, k = p & 15;
                float xi = x[0];
#pragma unroll
                for (int q = 1; q < kGIn; q++) xi = (i == q) ? x[q] : xi;
                const float fk = ldexpf((k & 1) ? 1.41421356f : 1.f, k >> 1);   // 2^(k/2)
                const float tt = xi * fk;
                const float fr = tt - floorf(tt);                    // revolutions in [0, 1)
                const uint16_t bs = f2bf(__builtin_amdgcn_sinf(fr)), bc = f2bf(__builtin_amdgcn_cosf(fr));
                row[kGIn + 2 * p] = bs;
                row[kGIn + 2 * p + 1] = bc;
                if (act) {
                    act[(size_t)(kGIn + 2 * p) * 4] = bs;
                    act[(size_t)(kGIn + 2 * p + 1) * 4] = bc;
                }
            }
        }
        __syncthreads();
        mfma_cd acc[4];
        tile_product<kGK1 / 16, 4>(E, kEStride, W + kFwdOff[0], kFwdStride[0], r, g, acc);
        hidden_epilogue(acc, bias, Ha, r, g, m0, store ? acts.h[1] : nullptr);
        __syncthreads();
        tile_product<4, 4>(Ha, kHStride, W + kFwdOff[1], kFwdStride[1], r, g, acc);
        hidden_epilogue(acc, bias + 64, Hb, r, g, m0, store ? acts.h[2] : nullptr);
        __syncthreads();
        tile_product<4, 4>(Hb, kHStride, W + kFwdOff[2], kFwdStride[2], r, g, acc);
        hidden_epilogue(acc, bias + 128, Ha, r, g, m0, store ? acts.h[3] : nullptr);
        __syncthreads();
        tile_product<4, 4>(Ha, kHStride, W + kFwdOff[3], kFwdStride[3], r, g, acc);
        hidden_epilogue(acc, bias + 192, Hb, r, g, m0, store ? acts.h[4] : nullptr);
        __syncthreads();
        tile_product<4, 3>(Hb, kHStride, W + kFwdOff[4], kFwdStride[4], r, g, acc);
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const int col = 16 * t + r;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const long long m = m0 + 4 * g + v;
                const float o = acc[t][v] + bias[256 + col];
                if (m < n && col < n_write) out[m * out_stride + col] = o;
                if (store) acts.o[m * kGOutPad + col] = o;
            }
        }
        __syncthreads();
    }
}

void guide_forward(const GuideNet& net, const float* in, int in_stride, long long n, float* out, int out_stride,
                   int n_write, const GuideActs* acts, int grid, hipStream_t s) {
    if (n <= 0) return;
    GuideActs a{};
    if (acts) a = *acts;
    const long long tiles = (n + 63) / 64;
    const int blocks = (int)std::min<long long>(tiles, grid);
    hipLaunchKernelGGL(k_guide_forward, dim3(blocks), dim3(256), kFwdLds, s, net, in, in_stride, n,
                       (const uint32_t*)nullptr, out, out_stride, n_write, a, acts ? 1 : 0);
}

void guide_forward_dev(const GuideNet& net, const float* in, int in_stride, const uint32_t* n_dev, float* out,
                       int out_stride, int grid, hipStream_t s) {
    GuideActs a{};
    hipLaunchKernelGGL(k_guide_forward, dim3(grid), dim3(256), kFwdLds, s, net, in, in_stride, 0ll, n_dev, out,
                       out_stride, out_stride, a, 0);
}

void guide_prepare_device() {
    (void)hipFuncSetAttribute((const void*)k_guide_forward, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFwdLds);
}

// ---------------------------------------------------------------------------------------------
// loss gradient and backward pass
// ---------------------------------------------------------------------------------------------
// dL/do for one sample of L = -wt * log q(w) * inv_batch; returns the sample's loss term.
YG_INLINE float loss_grad(const float* o, V3 w, float wt, float inv_batch, float* d) {
    float mx = o[0];
    for (int j = 1; j < kGLobes; j++) mx = fmaxf(mx, o[5 * j]);
    float pi[kGLobes], a[kGLobes], c[kGLobes], k[kGLobes], rn[kGLobes];
    V3 mu[kGLobes];
    float se = 0.f;
    for (int j = 0; j < kGLobes; j++) { pi[j] = expf(o[5 * j] - mx); se += pi[j]; }
    float amax = -f_inf();
    for (int j = 0; j < kGLobes; j++) {
        pi[j] /= se;
        const V3 r = v3(o[5 * j + 1], o[5 * j + 2], o[5 * j + 3]);
        rn[j] = fmaxf(length(r), 1e-12f);
        mu[j] = vdiv(r, rn[j]);
        const float s = fminf(fmaxf(o[5 * j + 4], kGSMin), kGSMax);
        k[j] = expf(s);
        c[j] = dot(mu[j], w);
        a[j] = logf(pi[j]) + logf(k[j]) - logf(2.f * kPi) - logf(-expm1f(-2.f * k[j])) + k[j] * (c[j] - 1.f);
        amax = fmaxf(amax, a[j]);
    }
    float sa = 0.f;
    for (int j = 0; j < kGLobes; j++) sa += expf(a[j] - amax);
    const float logq = amax + logf(sa);
    const float gs = -wt * inv_batch;
    for (int j = 0; j < kGLobes; j++) {
        const float gam = expf(a[j] - logq);
        d[5 * j] = gs * (gam - pi[j]);
        const float f = gs * gam * k[j] / rn[j];
        d[5 * j + 1] = f * (w.x - c[j] * mu[j].x);
        d[5 * j + 2] = f * (w.y - c[j] * mu[j].y);
        d[5 * j + 3] = f * (w.z - c[j] * mu[j].z);
        const float s = o[5 * j + 4];
        const float dlk = 1.f - 2.f * k[j] / expm1f(2.f * k[j]) + k[j] * (c[j] - 1.f);
        d[5 * j + 4] = (s > kGSMin && s < kGSMax) ? gs * gam * dlk : 0.f;
    }
    return -wt * logq * inv_batch;
}

constexpr int kBwdStride[5] = {0, kGH + 8, kGH + 8, kGH + 8, kGOutPad + 8};
constexpr int kBwdOff[5] = {0, 0, 4608, 9216, 13824};
constexpr int kBwdW = 17408;
constexpr int kDStride = kGH + 8;
constexpr size_t kBwdLds = (size_t)kBwdW * 2 + (size_t)4 * 2 * 16 * kDStride * 2;

// Backward step l -> l-1: dh = D * W_l^T (W_l canonical [64][N] in LDS), masked by h_l > 0.
template <int KSTEPS>
YG_INLINE void backward_layer(const uint16_t* D, const uint16_t* Wl, int w_stride, const uint16_t* hl,
                              uint16_t* Dn, uint16_t* dl, int r, int g, long long m0) {
    mfma_cd acc[4];
    tile_product<KSTEPS, 4>(D, kDStride, Wl, w_stride, r, g, acc);
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int col = 16 * t + r;
        const size_t idx = ((size_t)(m0 / 4 + g) * kGH + col) * 4;
        const uint2 hp = *reinterpret_cast<const uint2*>(hl + idx);
        const uint16_t hv[4] = {(uint16_t)(hp.x & 0xffff), (uint16_t)(hp.x >> 16), (uint16_t)(hp.y & 0xffff),
                                (uint16_t)(hp.y >> 16)};
        uint16_t dv[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            dv[v] = f2bf(bf2f(hv[v]) > 0.f ? acc[t][v] : 0.f);
            Dn[(4 * g + v) * kDStride + col] = dv[v];
        }
        *reinterpret_cast<uint2*>(dl + idx) =
            make_uint2((uint32_t)dv[0] | ((uint32_t)dv[1] << 16), (uint32_t)dv[2] | ((uint32_t)dv[3] << 16));
    }
}

__global__ __launch_bounds__(256) void k_guide_backward(GuideNet net, const GuideRecord* __restrict__ rec, long long n,
                                                        float inv_batch, GuideActs acts, float* loss_sum) {
    extern __shared__ __align__(16) uint16_t lds[];
    uint16_t* W = lds;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    uint16_t* Da = lds + kBwdW + wave * 2 * 16 * kDStride;
    uint16_t* Db = Da + 16 * kDStride;
    for (int l = 1; l < 5; l++) copy_rows(W + kBwdOff[l], kBwdStride[l], net.wr + kOffT[l], kGH, kLayerN[l]);
    __syncthreads();
    const long long n_btiles = (n + 63) / 64;
    for (long long bt = blockIdx.x; bt < n_btiles; bt += gridDim.x) {
        const long long m0 = bt * 64 + wave * 16;
        float loss = 0.f, wsum = 0.f;
        if (g == 0) {
            const long long m = m0 + r;
            float d[kGOut];
            for (int j = 0; j < kGOut; j++) d[j] = 0.f;
            if (m < n) {
                const GuideRecord R = rec[m];
                float o[kGOut];
                for (int j = 0; j < kGOut; j++) o[j] = acts.o[m * kGOutPad + j];
                if (R.wt != 0.f) loss = loss_grad(o, v3(R.w[0], R.w[1], R.w[2]), R.wt, inv_batch, d);
                wsum = R.wt * inv_batch;
            }
            uint16_t* d5 = acts.d[4] + ((size_t)(m / 4) * kGOutPad) * 4 + (m & 3);
            for (int j = 0; j < kGOutPad; j++) {
                const uint16_t b = f2bf(j < kGOut ? d[j] : 0.f);
                Da[r * kDStride + j] = b;
                d5[(size_t)j * 4] = b;
            }
        }
        if (loss_sum) {                          // [0] batch loss, [1] batch mean weight
            for (int off = 32; off > 0; off >>= 1) {
                loss += __shfl_xor(loss, off);
                wsum += __shfl_xor(wsum, off);
            }
            if (lane == 0 && loss != 0.f) atomicAdd(&loss_sum[0], loss);
            if (lane == 0 && wsum != 0.f) atomicAdd(&loss_sum[1], wsum);
        }
        __syncthreads();
        backward_layer<3>(Da, W + kBwdOff[4], kBwdStride[4], acts.h[4], Db, acts.d[3], r, g, m0);
        __syncthreads();
        backward_layer<4>(Db, W + kBwdOff[3], kBwdStride[3], acts.h[3], Da, acts.d[2], r, g, m0);
        __syncthreads();
        backward_layer<4>(Da, W + kBwdOff[2], kBwdStride[2], acts.h[2], Db, acts.d[1], r, g, m0);
        __syncthreads();
        backward_layer<4>(Db, W + kBwdOff[1], kBwdStride[1], acts.h[1], Da, acts.d[0], r, g, m0);
        __syncthreads();
    }
}

void guide_backward(const GuideNet& net, const GuideRecord* rec, long long n, float inv_batch, const GuideActs& acts,
                    float* loss_sum, hipStream_t s) {
    if (n <= 0) return;
    const long long tiles = (n + 63) / 64;
    hipLaunchKernelGGL(k_guide_backward, dim3((unsigned)std::min<long long>(tiles, 1024)), dim3(256), kBwdLds, s, net,
                       rec, n, inv_batch, acts, loss_sum);
}

// ---------------------------------------------------------------------------------------------
// weight gradients: dW_l = sum_m h_l[m]^T d_l[m] over one slice of the batch, per 16x16 tile
// ---------------------------------------------------------------------------------------------
constexpr int kTileStart[6] = {0, 76, 92, 108, 124, 136};

__global__ __launch_bounds__(256) void k_guide_wgrad(GuideActs acts, long long n_pad, int slice, float* partial) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= kTileStart[5]) return;
    int l = 0;
    while (tile >= kTileStart[l + 1]) l++;
    const int K = kLayerK[l], N = kLayerN[l];
    const int nts = N / 16;
    const int kt = (tile - kTileStart[l]) / nts, nt = (tile - kTileStart[l]) % nts;
    const uint16_t* X = acts.h[l];
    const uint16_t* D = acts.d[l];
    const long long s0 = (long long)blockIdx.y * slice;
    const long long s1 = std::min<long long>(s0 + slice, n_pad);
    mfma_cd acc = {0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    for (long long m0 = s0; m0 < s1; m0 += 16) {
        const mfma_ab a = ld_ab(X + ((size_t)(m0 / 4 + g) * K + 16 * kt + r) * 4);
        const mfma_ab b = ld_ab(D + ((size_t)(m0 / 4 + g) * N + 16 * nt + r) * 4);
        acc = mfma16(a, b, acc);
        if (kt == 0) {
#pragma unroll
            for (int v = 0; v < 4; v++) bsum += bf2f((uint16_t)b[v]);
        }
    }
    float* P = partial + (size_t)blockIdx.y * kGParams;
#pragma unroll
    for (int v = 0; v < 4; v++) P[kOffW[l] + (16 * kt + 4 * g + v) * N + 16 * nt + r] = acc[v];
    if (kt == 0) {
        bsum += __shfl_xor(bsum, 16);
        bsum += __shfl_xor(bsum, 32);
        if (g == 0) P[kOffB[l] + 16 * nt + r] = bsum;
    }
}

void guide_weight_grads(const GuideActs& acts, long long n_pad, int slice, float* partial, hipStream_t s) {
    const int n_slices = (int)((n_pad + slice - 1) / slice);
    hipLaunchKernelGGL(k_guide_wgrad, dim3((kTileStart[5] + 3) / 4, n_slices), dim3(256), 0, s, acts, n_pad, slice,
                       partial);
}

// ---------------------------------------------------------------------------------------------
// reduce + Adam, weight copies
// ---------------------------------------------------------------------------------------------
struct ParamPos { int layer, k, n; bool bias, pad; };

YG_INLINE ParamPos param_pos(int i) {
    ParamPos p{0, 0, 0, false, false};
    int l = 4;
    while (l > 0 && i < kOffW[l]) l--;
    p.layer = l;
    const int N = kLayerN[l];
    if (i >= kOffB[l]) {
        p.bias = true;
        p.n = i - kOffB[l];
        p.pad = l == 4 && p.n >= kGOut;
    } else {
        p.k = (i - kOffW[l]) / N;
        p.n = (i - kOffW[l]) % N;
        p.pad = (l == 0 && p.k >= kGEnc) || (l == 4 && p.n >= kGOut);
    }
    return p;
}

YG_INLINE void write_copies(const ParamPos& p, float w, uint16_t* wt, uint16_t* wr) {
    if (p.bias) return;
    const uint16_t b = f2bf(w);
    const int K = kLayerK[p.layer], N = kLayerN[p.layer];
    wr[kOffT[p.layer] + p.k * N + p.n] = b;
    wt[kOffT[p.layer] + p.n * K + p.k] = b;
}

__global__ void k_guide_apply(float* master, float* m, float* v, const float* partial, int n_slices, float lr, float b1,
                              float b2, float eps, float bc1, float bc2, uint16_t* wt, uint16_t* wr, float* grad_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kGParams) return;
    const ParamPos p = param_pos(i);
    float gsum = 0.f;
    for (int s = 0; s < n_slices; s++) gsum += partial[(size_t)s * kGParams + i];
    if (p.pad) gsum = 0.f;
    if (grad_out) { grad_out[i] = gsum; return; }
    if (p.pad) return;
    const float mi = b1 * m[i] + (1.f - b1) * gsum;
    const float vi = b2 * v[i] + (1.f - b2) * gsum * gsum;
    m[i] = mi;
    v[i] = vi;
    const float w = master[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    master[i] = w;
    write_copies(p, w, wt, wr);
}

void guide_apply(float* master, float* m, float* v, const float* partial, int n_slices, float lr, float b1, float b2,
                 float eps, float bc1, float bc2, uint16_t* wt, uint16_t* wr, float* grad_out, hipStream_t s) {
    hipLaunchKernelGGL(k_guide_apply, dim3((kGParams + 255) / 256), dim3(256), 0, s, master, m, v, partial, n_slices, lr,
                       b1, b2, eps, bc1, bc2, wt, wr, grad_out);
}

__global__ void k_guide_sync(const float* master, uint16_t* wt, uint16_t* wr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kGParams) return;
    const ParamPos p = param_pos(i);
    write_copies(p, p.pad ? 0.f : master[i], wt, wr);
}

void guide_sync_weights(const float* master, uint16_t* wt, uint16_t* wr, hipStream_t s) {
    hipLaunchKernelGGL(k_guide_sync, dim3((kGParams + 255) / 256), dim3(256), 0, s, master, wt, wr);
}

YG_INLINE uint32_t mix32(uint32_t h) {         // murmur3 finaliser
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    return h;
}

__global__ void k_guide_gather(const GuideRecord* all, long long count, long long batch, uint32_t seed, uint32_t step,
                               GuideRecord* out) {
    const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const uint32_t h = mix32((uint32_t)b * 0x9e3779b1u ^ mix32(step * 0x85ebca77u + seed));
    const long long idx = (long long)(((uint64_t)h * (uint64_t)count) >> 32);
    out[b] = all[idx];
}

void guide_gather(const GuideRecord* all, long long count, long long batch, uint32_t seed, uint32_t step,
                  GuideRecord* out, hipStream_t s) {
    if (batch <= 0 || count <= 0) return;
    hipLaunchKernelGGL(k_guide_gather, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, s, all, count, batch, seed,
                       step, out);
}

// mixture pdf / sampling for yar_guide_pdf and yar_guide_sample
__global__ void k_guide_mixture(const float* __restrict__ o, const float* __restrict__ dirs, const float* __restrict__ u,
                                long long n, float* out_dirs, float* out_pdf) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    VmfLobe L[kGLobes];
    mixture_from_outputs(o + i * kGOut, L);
    V3 w;
    if (u) {
        w = mixture_sample(L, u[3 * i], u[3 * i + 1], u[3 * i + 2]);
        out_dirs[3 * i] = w.x; out_dirs[3 * i + 1] = w.y; out_dirs[3 * i + 2] = w.z;
    } else {
        w = v3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
    }
    out_pdf[i] = mixture_pdf(L, w);
}

}  // namespace yg

// =============================================================================================
// C ABI
// =============================================================================================
namespace {

using namespace yg;

struct GuideErr { std::string msg; };

#define GHIP(expr)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) throw GuideErr{std::string(#expr) + ": " + hipGetErrorString(e_)};     \
    } while (0)

template <class T>
struct GBuf {
    T* p = nullptr;
    size_t n = 0;
    void reserve(size_t count) {
        if (count <= n) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        GHIP(hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T)));
        n = count;
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
};

struct Pcg {                                   // pcg32 (O'Neill), parameter initialisation only
    uint64_t s, inc;
    Pcg(uint64_t seed, uint64_t seq) : s(0), inc((seq << 1) | 1) { next(); s += seed; next(); }
    uint32_t next() {
        uint64_t o = s;
        s = o * 6364136223846793005ull + inc;
        uint32_t x = (uint32_t)(((o >> 18) ^ o) >> 27), rot = (uint32_t)(o >> 59);
        return (x >> rot) | (x << ((32 - rot) & 31));
    }
    float uniform() { return (next() >> 8) * (1.f / 16777216.f); }
};

}  // namespace

struct yar_guide {
    std::string error;
    int device = 0;
    int cu_count = 1;
    hipStream_t stream = nullptr;
    yar_guide_config cfg{};
    int step = 0;                              // Adam steps taken
    GBuf<float> master, adam_m, adam_v, partial, loss;
    GBuf<uint16_t> wt, wr;
    // training batch
    GBuf<GuideRecord> batch, samples;
    GBuf<uint16_t> h0, h, d;
    GBuf<float> o;
    long long batch_cap = 0;
    // query scratch
    GBuf<float> q_in, q_out, q_dirs, q_u, q_pdf;

    ~yar_guide() {
        (void)hipSetDevice(device);
        master.release(); adam_m.release(); adam_v.release(); partial.release(); loss.release();
        wt.release(); wr.release(); batch.release(); samples.release(); h0.release(); h.release(); d.release();
        o.release(); q_in.release(); q_out.release(); q_dirs.release(); q_u.release(); q_pdf.release();
        if (stream) (void)hipStreamDestroy(stream);
    }

    GuideNet net() const { return GuideNet{wt.p, wr.p, master.p}; }

    // activations of a batch of up to `cap` samples (rounded to 64)
    GuideActs acts(long long cap) {
        const long long np = (cap + 63) / 64 * 64;
        if (np > batch_cap) {
            h0.reserve((size_t)np * kGK1);
            h.reserve((size_t)np * kGH * 4);
            d.reserve((size_t)np * (kGH * 4 + kGOutPad));
            o.reserve((size_t)np * kGOutPad);
            batch.reserve((size_t)np);
            partial.reserve((size_t)((np + 1023) / 1024) * kGParams);
            batch_cap = np;
        }
        GuideActs a;
        a.h[0] = h0.p;
        for (int l = 1; l < 5; l++) a.h[l] = h.p + (size_t)(l - 1) * np * kGH;
        for (int l = 0; l < 4; l++) a.d[l] = d.p + (size_t)l * np * kGH;
        a.d[4] = d.p + (size_t)4 * np * kGH;
        a.o = o.p;
        return a;
    }

    // forward + backward + weight gradients of records rec[0..n) (n <= batch capacity)
    void gradients(const GuideRecord* rec, long long n, float inv_batch, bool want_loss, hipStream_t s) {
        GuideActs a = acts(n);
        const long long np = (n + 63) / 64 * 64;
        guide_forward(net(), reinterpret_cast<const float*>(rec), 16, n, o.p, kGOutPad, 0, &a, cu_count, s);
        if (want_loss) GHIP(hipMemsetAsync(loss.p, 0, 2 * sizeof(float), s));
        guide_backward(net(), rec, n, inv_batch, a, want_loss ? loss.p : nullptr, s);
        guide_weight_grads(a, np, 1024, partial.p, s);
    }

    int slices(long long n) const { return (int)(((n + 63) / 64 * 64 + 1023) / 1024); }

    void adam_step(long long n, hipStream_t s) {
        step++;
        const float bc1 = 1.f - std::pow(cfg.beta1, (float)step), bc2 = 1.f - std::pow(cfg.beta2, (float)step);
        guide_apply(master.p, adam_m.p, adam_v.p, partial.p, slices(n), cfg.learning_rate, cfg.beta1, cfg.beta2,
                    cfg.epsilon, bc1, bc2, wt.p, wr.p, nullptr, s);
    }

    // `steps` Adam steps on batches drawn from device records rec[0..count), on stream s
    void train_records(const GuideRecord* rec, long long count, int steps, hipStream_t s, bool want_loss) {
        const long long B = cfg.batch_size;
        acts(B);
        for (int k = 0; k < steps; k++) {
            guide_gather(rec, count, B, cfg.seed, (uint32_t)step, batch.p, s);
            gradients(batch.p, B, 1.f / (float)B, want_loss && k == steps - 1, s);
            adam_step(B, s);
        }
    }
};

namespace {

int gfail(yar_guide* g, int code, const std::string& msg) {
    if (g) g->error = msg;
    yar::set_last_error(msg);
    return code;
}

void fill_defaults(yar_guide_config& c) {
    if (c.batch_size <= 0) c.batch_size = 16384;
    c.batch_size = (c.batch_size + 63) / 64 * 64;
    if (c.learning_rate <= 0.f) c.learning_rate = 5e-3f;
    if (c.beta1 <= 0.f) c.beta1 = 0.9f;
    if (c.beta2 <= 0.f) c.beta2 = 0.99f;
    if (c.epsilon <= 0.f) c.epsilon = 1e-8f;
    if (c.steps_per_pass <= 0) c.steps_per_pass = 32;
    if (c.guide_fraction <= 0.f || c.guide_fraction > 1.f) c.guide_fraction = 0.5f;
    if (c.first_pass_spp <= 0) c.first_pass_spp = 8;
    if (c.record_capacity <= 0) c.record_capacity = 1 << 22;
    if (c.pass_count <= 0) c.pass_count = 2;
}

std::vector<GuideRecord> pack_records(const float* x, const float* dirs, const float* wt, int64_t n) {
    std::vector<GuideRecord> r((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        GuideRecord& R = r[(size_t)i];
        std::memset(&R, 0, sizeof(R));
        for (int k = 0; k < 9; k++) R.x[k] = x[9 * i + k];
        if (dirs) for (int k = 0; k < 3; k++) R.w[k] = dirs[3 * i + k];
        R.wt = wt ? wt[i] : 0.f;
    }
    return r;
}

}  // namespace

extern "C" {

int64_t yar_guide_param_count(void) { return kGParams; }

int yar_guide_create(int32_t device, const yar_guide_config* config, yar_guide** out) {
    if (!out) return gfail(nullptr, YAR_ERR_INVALID_ARGUMENT, "yar_guide_create: null argument");
    *out = nullptr;
    auto* g = new (std::nothrow) yar_guide;
    if (!g) return gfail(nullptr, YAR_ERR_OUT_OF_MEMORY, "yar_guide_create: out of host memory");
    g->device = device;
    if (config) g->cfg = *config;
    fill_defaults(g->cfg);
    try {
        int ndev = 0;
        GHIP(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw GuideErr{"no such HIP device"};
        GHIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        GHIP(hipGetDeviceProperties(&prop, device));
        g->cu_count = std::max(1, prop.multiProcessorCount);
        if (prop.sharedMemPerBlock < kFwdLds)
            throw GuideErr{"device offers " + std::to_string(prop.sharedMemPerBlock) + " bytes of LDS per block, the MLP needs " +
                           std::to_string(kFwdLds)};
        GHIP(hipFuncSetAttribute((const void*)k_guide_forward, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFwdLds));
        GHIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
        g->master.reserve(kGParams);
        g->adam_m.reserve(kGParams);
        g->adam_v.reserve(kGParams);
        g->wt.reserve(kGWeights);
        g->wr.reserve(kGWeights);
        g->loss.reserve(2);
        // He-uniform initialisation for the ReLU layers, a 0.1x smaller output layer so the initial
        // mixture is close to 8 broad lobes of equal weight; biases zero.
        std::vector<float> p(kGParams, 0.f);
        Pcg rng(g->cfg.seed, 0x9d1de);
        for (int l = 0; l < 5; l++) {
            const int fan_in = l == 0 ? kGEnc : kGH;
            const float a = std::sqrt(6.f / fan_in) * (l == 4 ? 0.1f : 1.f);
            for (int k = 0; k < kLayerK[l]; k++)
                for (int n = 0; n < kLayerN[l]; n++) {
                    const float u = rng.uniform();
                    const bool pad = (l == 0 && k >= kGEnc) || (l == 4 && n >= kGOut);
                    p[kOffW[l] + k * kLayerN[l] + n] = pad ? 0.f : (2.f * u - 1.f) * a;
                }
        }
        GHIP(hipMemcpyAsync(g->master.p, p.data(), kGParams * sizeof(float), hipMemcpyHostToDevice, g->stream));
        GHIP(hipMemsetAsync(g->adam_m.p, 0, kGParams * sizeof(float), g->stream));
        GHIP(hipMemsetAsync(g->adam_v.p, 0, kGParams * sizeof(float), g->stream));
        guide_sync_weights(g->master.p, g->wt.p, g->wr.p, g->stream);
        GHIP(hipGetLastError());
        GHIP(hipStreamSynchronize(g->stream));
    } catch (const GuideErr& e) {
        delete g;
        return gfail(nullptr, YAR_ERR_HIP, "yar_guide_create: " + e.msg);
    } catch (const std::bad_alloc&) {
        delete g;
        return gfail(nullptr, YAR_ERR_OUT_OF_MEMORY, "yar_guide_create: " + std::string("out of host memory"));
    } catch (const std::exception& e) {
        delete g;
        return gfail(nullptr, YAR_ERR_INTERNAL, "yar_guide_create: " + std::string(e.what()));
    } catch (...) {
        delete g;
        return gfail(nullptr, YAR_ERR_INTERNAL, "yar_guide_create: " + std::string("unknown exception"));
    }
    *out = g;
    return YAR_OK;
}

void yar_guide_destroy(yar_guide* g) { delete g; }

const char* yar_guide_last_error(const yar_guide* g) { return g ? g->error.c_str() : yar::get_last_error(); }

int yar_guide_get_params(yar_guide* g, float* out) {
    if (!g || !out) return gfail(g, YAR_ERR_INVALID_ARGUMENT, "yar_guide_get_params: null argument");
    try {
        GHIP(hipSetDevice(g->device));
        GHIP(hipMemcpyAsync(out, g->master.p, kGParams * sizeof(float), hipMemcpyDeviceToHost, g->stream));
        GHIP(hipStreamSynchronize(g->stream));
    } catch (const GuideErr& e) {
        return gfail(g, YAR_ERR_HIP, "yar_guide_get_params: " + e.msg);
    } catch (const std::bad_alloc&) {
        return gfail(g, YAR_ERR_OUT_OF_MEMORY, "yar_guide_get_params: " + std::string("out of host memory"));
    } catch (const std::exception& e) {
        return gfail(g, YAR_ERR_INTERNAL, "yar_guide_get_params: " + std::string(e.what()));
    } catch (...) {
        return gfail(g, YAR_ERR_INTERNAL, "yar_guide_get_params: " + std::string("unknown exception"));
    }
    return YAR_OK;
}

int yar_guide_set_params(yar_guide* g, const float* params) {
    if (!g || !params) return gfail(g, YAR_ERR_INVALID_ARGUMENT, "yar_guide_set_params: null argument");
    try {
        GHIP(hipSetDevice(g->device));
        std::vector<float> p(params, params + kGParams);
        for (int l = 0; l < 5; l++)
            for (int k = 0; k < kLayerK[l]; k++)
                for (int n = 0; n < kLayerN[l]; n++)
                    if ((l == 0 && k >= kGEnc) || (l == 4 && n >= kGOut)) p[kOffW[l] + k * kLayerN[l] + n] = 0.f;
        for (int n = kGOut; n < kGOutPad; n++) p[kOffB[4] + n] = 0.f;
        GHIP(hipMemcpyAsync(g->master.p, p.data(), kGParams * sizeof(float), hipMemcpyHostToDevice, g->stream));
        GHIP(hipMemsetAsync(g->adam_m.p, 0, kGParams * sizeof(float), g->stream));
        GHIP(hipMemsetAsync(g->adam_v.p, 0, kGParams * sizeof(float), g->stream));
        guide_sync_weights(g->master.p, g->wt.p, g->wr.p, g->stream);
        GHIP(hipGetLastError());
        GHIP(hipStreamSynchronize(g->stream));
        g->step = 0;
    } catch (const GuideErr& e) {
        return gfail(g, YAR_ERR_HIP, "yar_guide_set_params: " + e.msg);
    } catch (const std::bad_alloc&) {
        return gfail(g, YAR_ERR_OUT_OF_MEMORY, "yar_guide_set_params: " + std::string("out of host memory"));
    } catch (const std::exception& e) {
        return gfail(g, YAR_ERR_INTERNAL, "yar_guide_set_params: " + std::string(e.what()));
    } catch (...) {
        return gfail(g, YAR_ERR_INTERNAL, "yar_guide_set_params: " + std::string("unknown exception"));
    }
    return YAR_OK;
}

int yar_guide_forward(yar_guide* g, const float* x, int64_t n, float* out) {
    if (!g || n < 0 || (n > 0 && (!x || !out))) return gfail(g, YAR_ERR_INVALID_ARGUMENT, "yar_guide_forward: invalid argument");
    if (n == 0) return YAR_OK;
    try {
        GHIP(hipSetDevice(g->device));
        g->q_in.reserve((size_t)n * 9);
        g->q_out.reserve((size_t)n * kGOut);
        GHIP(hipMemcpyAsync(g->q_in.p, x, (size_t)n * 9 * sizeof(float), hipMemcpyHostToDevice, g->stream));
        guide_forward(g->net(), g->q_in.p, 9, n, g->q_out.p, kGOut, kGOut, nullptr, g->cu_count, g->stream);
        GHIP(hipGetLastError());
        GHIP(hipMemcpyAsync(out, g->q_out.p, (size_t)n * kGOut * sizeof(float), hipMemcpyDeviceToHost, g->stream));
        GHIP(hipStreamSynchronize(g->stream));
    } catch (const GuideErr& e) {
        return gfail(g, YAR_ERR_HIP, "yar_guide_forward: " + e.msg);
    } catch (const std::bad_alloc&) {
        return gfail(g, YAR_ERR_OUT_OF_MEMORY, "yar_guide_forward: " + std::string("out of host memory"));
    } catch (const std::exception& e) {
        return gfail(g, YAR_ERR_INTERNAL, "yar_guide_forward: " + std::string(e.what()));
    } catch (...) {
        return gfail(g, YAR_ERR_INTERNAL, "yar_guide_forward: " + std::string("unknown exception"));
    }
    return YAR_OK;
}

int yar_guide_gradients(yar_guide* g, const float* x, const float* dirs, const float* wt, int64_t n, float* out_grad,
                        float* out_loss) {
    if (!g || n <= 0 || !x || !dirs || !wt || !out_grad)
        return gfail(g, YAR_ERR_INVALID_ARGUMENT, "yar_guide_gradients: invalid argument");
    if (n > (int64_t)1 << 24) return gfail(g, YAR_ERR_INVALID_ARGUMENT, "yar_guide_gradients: batch larger than 2^24");
    try {
        GHIP(hipSetDevice(g->device));
        std::vector<GuideRecord> r = pack_records(x, dirs, wt, n);
        g->acts(n);
        GHIP(hipMemcpyAsync(g->batch.p, r.data(), r.size() * sizeof(GuideRecord), hipMemcpyHostToDevice, g->stream));
        g->gradients(g->batch.p, n, 1.f / (float)n, true, g->stream);
        g->q_out.reserve(kGParams);
        guide_apply(g->master.p, g->adam_m.p, g->adam_v.p, g->partial.p, g->slices(n), 0.f, 0.f, 0.f, 0.f, 1.f, 1.f,
                    g->wt.p, g->wr.p, g->q_out.p, g->stream);
        GHIP(hipGetLastError());
        GHIP(hipMemcpyAsync(out_grad, g->q_out.p, kGParams * sizeof(float), hipMemcpyDeviceToHost, g->stream));
        float l = 0.f;
        GHIP(hipMemcpyAsync(&l, g->loss.p, sizeof(float), hipMemcpyDeviceToHost, g->stream));
        GHIP(hipStreamSynchronize(g->stream));
        if (out_loss) *out_loss = l;
    } catch (const GuideErr& e) {
        return gfail(g, YAR_ERR_HIP, "yar_guide_gradients: " + e.msg);
    } catch (const std::bad_alloc&) {
        return gfail(g, YAR_ERR_OUT_OF_MEMORY, "yar_guide_gradients: " + std::string("out of host memory"));
    } catch (const std::exception& e) {
        return gfail(g, YAR_ERR_INTERNAL, "yar_guide_gradients: " + std::string(e.what()));
    } catch (...) {
        return gfail(g, YAR_ERR_INTERNAL, "yar_guide_gradients: " + std::string("unknown exception"));
    }
    return YAR_OK;
}

int yar_guide_train(yar_guide* g, const float* x, const float* dirs, const float* wt, int64_t n, int32_t steps,
                    float* out_loss) {
    if (!g || n <= 0 || !x || !dirs || !wt || steps < 0)
        return gfail(g, YAR_ERR_INVALID_ARGUMENT, "yar_guide_train: invalid argument");
    try {
        GHIP(hipSetDevice(g->device));
        std::vector<GuideRecord> r = pack_records(x, dirs, wt, n);
        g->samples.reserve((size_t)n);
        GHIP(hipMemcpyAsync(g->samples.p, r.data(), r.size() * sizeof(GuideRecord), hipMemcpyHostToDevice, g->stream));
        float l = 0.f;
        g->train_records(g->samples.p, n, steps, g->stream, true);
        GHIP(hipGetLastError());
        if (steps > 0) GHIP(hipMemcpyAsync(&l, g->loss.p, sizeof(float), hipMemcpyDeviceToHost, g->stream));
        GHIP(hipStreamSynchronize(g->stream));
        if (out_loss) *out_loss = l;
    } catch (const GuideErr& e) {
        return gfail(g, YAR_ERR_HIP, "yar_guide_train: " + e.msg);
    } catch (const std::bad_alloc&) {
        return gfail(g, YAR_ERR_OUT_OF_MEMORY, "yar_guide_train: " + std::string("out of host memory"));
    } catch (const std::exception& e) {
        return gfail(g, YAR_ERR_INTERNAL, "yar_guide_train: " + std::string(e.what()));
    } catch (...) {
        return gfail(g, YAR_ERR_INTERNAL, "yar_guide_train: " + std::string("unknown exception"));
    }
    return YAR_OK;
}

static int guide_mixture_call(yar_guide* g, const float* x, const float* dirs, const float* u, int64_t n, float* out_dirs,
                              float* out_pdf, const char* name) {
    if (!g || n < 0 || (n > 0 && (!x || !out_pdf || (!dirs && !u) || (u && !out_dirs))))
        return gfail(g, YAR_ERR_INVALID_ARGUMENT, std::string(name) + ": invalid argument");
    if (n == 0) return YAR_OK;
    try {
        GHIP(hipSetDevice(g->device));
        g->q_in.reserve((size_t)n * 9);
        g->q_out.reserve((size_t)n * kGOut);
        g->q_dirs.reserve((size_t)n * 3);
        g->q_u.reserve((size_t)n * 3);
        g->q_pdf.reserve((size_t)n);
        GHIP(hipMemcpyAsync(g->q_in.p, x, (size_t)n * 9 * sizeof(float), hipMemcpyHostToDevice, g->stream));
        if (u) GHIP(hipMemcpyAsync(g->q_u.p, u, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, g->stream));
        else GHIP(hipMemcpyAsync(g->q_dirs.p, dirs, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, g->stream));
        guide_forward(g->net(), g->q_in.p, 9, n, g->q_out.p, kGOut, kGOut, nullptr, g->cu_count, g->stream);
        hipLaunchKernelGGL(k_guide_mixture, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g->stream, g->q_out.p,
                           g->q_dirs.p, u ? g->q_u.p : nullptr, (long long)n, g->q_dirs.p, g->q_pdf.p);
        GHIP(hipGetLastError());
        if (u) GHIP(hipMemcpyAsync(out_dirs, g->q_dirs.p, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, g->stream));
        GHIP(hipMemcpyAsync(out_pdf, g->q_pdf.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, g->stream));
        GHIP(hipStreamSynchronize(g->stream));
    } catch (const GuideErr& e) {
        return gfail(g, YAR_ERR_HIP, std::string(name) + ": " + e.msg);
    } catch (const std::bad_alloc&) {
        return gfail(g, YAR_ERR_OUT_OF_MEMORY, std::string(name) + ": " + std::string("out of host memory"));
    } catch (const std::exception& e) {
        return gfail(g, YAR_ERR_INTERNAL, std::string(name) + ": " + std::string(e.what()));
    } catch (...) {
        return gfail(g, YAR_ERR_INTERNAL, std::string(name) + ": " + std::string("unknown exception"));
    }
    return YAR_OK;
}

int yar_guide_pdf(yar_guide* g, const float* x, const float* dirs, int64_t n, float* out_pdf) {
    return guide_mixture_call(g, x, dirs, nullptr, n, nullptr, out_pdf, "yar_guide_pdf");
}

int yar_guide_sample(yar_guide* g, const float* x, const float* u, int64_t n, float* out_dirs, float* out_pdf) {
    return guide_mixture_call(g, x, nullptr, u, n, out_dirs, out_pdf, "yar_guide_sample");
}

}  // extern "C"

namespace yg {

GuideNet guide_net(yar_guide* g) { return g->net(); }
const yar_guide_config& guide_config(const yar_guide* g) { return g->cfg; }
int guide_device(const yar_guide* g) { return g->device; }
float guide_train_on(yar_guide* g, const GuideRecord* rec, long long count, int steps, hipStream_t s) {
    g->train_records(rec, count, steps, s, true);
    float l[2] = {0.f, 0.f};
    GHIP(hipMemcpyAsync(l, g->loss.p, sizeof(l), hipMemcpyDeviceToHost, s));
    GHIP(hipStreamSynchronize(s));
    // mean log-likelihood gain over the uniform sphere density on the last batch, in nats per
    // unit weight: (w log 4 pi - loss) / w
    return l[1] > 0.f ? (l[1] * std::log(4.f * kPi) - l[0]) / l[1] : 0.f;
}

}  // namespace yg
