// mlp.hip -- the candidate-proposer network: a 6 -> 128 -> 128 -> 4 MLP
// with BF16 weights and activations on the MFMA units, FP32 accumulation,
// trained on the GPU by hand-written forward, backward and Adam kernels.
//
// Layout: parameters live as one FP32 master vector (W1 [128][6], b1, W2
// [128][128], b2, W3 [4][128], b3, rows = output units); every update also
// writes the BF16 operand copies the kernels read: W1 [128][16] (inputs padded
// to 16), W2 [128][128], W3 [16][128] (outputs padded to 16) for the forward
// pass, and W2^T, W3^T for the backward pass.
//
// One workgroup (4 waves) handles a tile of 32 samples.  Every product is a
// sequence of 16x16x32 BF16 MFMA tiles (16x16x16 for a last K step of 16)
// whose operands are both K-contiguous
// (A row-major M x K, B given as B^T row-major N x K); activations are kept in
// LDS in both orientations so that forward and backward products read
// K-contiguous rows:
//   forward   z1 = x W1^T + b1, h1 = bf16(relu z1); z2 = h1 W2^T + b2,
//             h2 = bf16(relu z2); o = h2 W3^T + b3
//   loss      (o_0..2 - t_0..2)^2 * wc  +  pinball_tau(t_3 - o_3)
//   backward  g = dL/do (bf16), dW3 = g^T h2, dh2 = g W3, dz2 = dh2 [h2 > 0],
//             dW2 = dz2^T h1, dh1 = dz2 W2, dz1 = dh1 [h1 > 0], dW1 = dz1^T x
// Per-workgroup gradient partials are summed in workgroup order by the Adam
// kernel, so training is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "axis_hip.hpp"
#include "mlp.hpp"

namespace axis_hip {

namespace mlp {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;

// BF16 operand copies inside the packed weight buffer
constexpr int bW1 = 0;                   // [kH][kInP]
constexpr int bW2 = bW1 + kH * kInP;     // [kH][kH]
constexpr int bW3 = bW2 + kH * kH;       // [kOutP][kH]
constexpr int bW2T = bW3 + kOutP * kH;   // [kH][kH]
constexpr int bW3T = bW2T + kH * kH;     // [kH][kOutP]
constexpr int kPacked = bW3T + kH * kOutP;

__device__ inline std::uint16_t f2bf(float f) { return round_bf16(f); }
__device__ inline float bf2f(std::uint16_t h) { return __uint_as_float(static_cast<unsigned>(h) << 16); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// C tile (16 x 16) += A[16 x K] * B[K x 16]; A row-major (lda), B given as
// B^T row-major (ldb), K a multiple of 16, lda and ldb multiples of 8.  Steps
// of 32 go through gfx950's 16x16x32 BF16 MFMA (lane l feeds A[l % 16][8
// (l / 16) ..+7] and B[8 (l / 16) ..+7][l % 16]), a remaining 16 through the
// 16x16x16 one (4 (l / 16) ..+3); lane l receives C[4 (l / 16) + i][l % 16],
// i < 4, from both.
__device__ inline f32x4 mma_tile(const std::uint16_t *A, int lda, const std::uint16_t *BT, int ldb, int K, f32x4 acc) {
    const int lane = static_cast<int>(threadIdx.x) & 63, r = lane & 15;
    int k0 = 0;
    for (; k0 + 32 <= K; k0 += 32) {
        const int kq = (lane >> 4) * 8;
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(A + r * lda + k0 + kq);
        const bf16x8 b = *reinterpret_cast<const bf16x8 *>(BT + r * ldb + k0 + kq);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    for (; k0 < K; k0 += 16) {
        const int kq = (lane >> 4) * 4;
        const bf16x4 a = *reinterpret_cast<const bf16x4 *>(A + r * lda + k0 + kq);
        const bf16x4 b = *reinterpret_cast<const bf16x4 *>(BT + r * ldb + k0 + kq);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    return acc;
}

struct Smem {
    alignas(16) std::uint16_t xb[kTile * kInP], xbT[kInP * kTile];
    alignas(16) std::uint16_t h1[kTile * kH], h1T[kH * kTile];
    alignas(16) std::uint16_t h2[kTile * kH], h2T[kH * kTile];
    alignas(16) std::uint16_t g3[kTile * kOutP], g3T[kOutP * kTile];
    float o[kTile * kOutP];
    float dz[kTile * kH];
    float red[kThreads / 64];
};

// forward pass of one tile; leaves xb, h1, h2 (both orientations) and o in LDS
__device__ void forward_tile(Smem &S, const std::uint16_t *wb, const float *p, const float *X, index_t s0, index_t n) {
    const int tid = static_cast<int>(threadIdx.x), wave = tid >> 6, lane = tid & 63;
    for (int idx = tid; idx < kTile * kInP; idx += kThreads) {
        const int b = idx / kInP, i = idx % kInP;
        const index_t s = s0 + b;
        const float v = (i < kIn && s < n) ? X[s * kIn + i] : 0.0f;
        const std::uint16_t h = f2bf(v);
        S.xb[b * kInP + i] = h;
        S.xbT[i * kTile + b] = h;
    }
    __syncthreads();
    // z1, z2: 2 x 8 output tiles, 4 per wave
    for (int layer = 0; layer < 2; ++layer) {
        const std::uint16_t *A = layer == 0 ? S.xb : S.h1;
        const int lda = layer == 0 ? kInP : kH, K = lda;
        const std::uint16_t *BT = wb + (layer == 0 ? bW1 : bW2);
        const float *bias = p + (layer == 0 ? oB1 : oB2);
        std::uint16_t *H = layer == 0 ? S.h1 : S.h2, *HT = layer == 0 ? S.h1T : S.h2T;
        for (int t = wave * 4; t < wave * 4 + 4; ++t) {
            const int m0 = (t / 8) * 16, n0 = (t % 8) * 16;
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mma_tile(A + m0 * lda, lda, BT + n0 * K, K, K, acc);
            const int col = n0 + (lane & 15);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int b = m0 + (lane >> 4) * 4 + i;
                const float z = acc[i] + bias[col];
                const std::uint16_t h = f2bf(z > 0.0f ? z : 0.0f);
                H[b * kH + col] = h;
                HT[col * kTile + b] = h;
            }
        }
        __syncthreads();
    }
    // o: 2 tiles (rows 0-15, 16-31), waves 0 and 1
    if (wave < 2) {
        const int m0 = wave * 16;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma_tile(S.h2 + m0 * kH, kH, wb + bW3, kH, kH, acc);
        const int q = lane & 15;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int b = m0 + (lane >> 4) * 4 + i;
            S.o[b * kOutP + q] = q < kOut ? acc[i] + p[oB3 + q] : 0.0f;
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kThreads) forward_kernel(const std::uint16_t *wb, const float *p, const float *X, index_t n, float *out) {
    __shared__ Smem S;
    const index_t s0 = static_cast<index_t>(blockIdx.x) * kTile;
    forward_tile(S, wb, p, X, s0, n);
    for (int idx = threadIdx.x; idx < kTile * kOut; idx += kThreads) {
        const int b = idx / kOut, q = idx % kOut;
        if (s0 + b < n) out[(s0 + b) * kOut + q] = S.o[b * kOutP + q];
    }
}

// block sum of one float per thread (result valid in thread 0)
__device__ float block_sum(Smem &S, float v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0) S.red[threadIdx.x >> 6] = v;
    __syncthreads();
    float t = 0.0f;
    if (threadIdx.x == 0)
        for (int w = 0; w < kThreads / 64; ++w) t += S.red[w];
    __syncthreads();
    return t;
}

// Forward + backward of one tile; writes this workgroup's gradient partial
// part[blockIdx.x][kParams] (already scaled by `scale` = 1 / batch size) and
// its loss partial.
__global__ void __launch_bounds__(kThreads) grad_kernel(const std::uint16_t *wb, const float *p, const float *X, const float *T, index_t s_begin,
                                                        index_t s_end, float scale, float wc, float tau, float *part, float *loss_part) {
    __shared__ Smem S;
    const int tid = static_cast<int>(threadIdx.x), wave = tid >> 6, lane = tid & 63;
    const index_t s0 = s_begin + static_cast<index_t>(blockIdx.x) * kTile;
    forward_tile(S, wb, p, X, s0, s_end);
    float *G = part + static_cast<index_t>(blockIdx.x) * kParams;
    // loss and dL/do
    float lsum = 0.0f;
    for (int idx = tid; idx < kTile * kOutP; idx += kThreads) {
        const int b = idx / kOutP, q = idx % kOutP;
        const index_t s = s0 + b;
        float g = 0.0f;
        if (q < kOut && s < s_end) {
            const float o = S.o[idx], t = T[s * kOut + q];
            if (q < 3) {
                const float d = o - t;
                lsum += wc * d * d;
                g = 2.0f * wc * d;
            } else {
                const float u = t - o;
                lsum += u > 0.0f ? tau * u : (tau - 1.0f) * u;
                g = u > 0.0f ? -tau : 1.0f - tau;
            }
            g *= scale;
        }
        S.o[idx] = g;  // FP32 dL/do, for the bias gradient
        const std::uint16_t h = f2bf(g);
        S.g3[b * kOutP + q] = h;
        S.g3T[q * kTile + b] = h;
    }
    const float tot = block_sum(S, lsum);  // contains barriers
    if (tid == 0) loss_part[blockIdx.x] = tot * scale;
    // dW3 = g^T h2: [16 x 128], 8 tiles, 2 per wave; db3
    for (int t = wave * 2; t < wave * 2 + 2; ++t) {
        const int n0 = t * 16;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma_tile(S.g3T, kTile, S.h2T + n0 * kTile, kTile, kTile, acc);
        const int i = n0 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int q = (lane >> 4) * 4 + r;
            if (q < kOut) G[oW3 + q * kH + i] = acc[r];
        }
    }
    if (tid < kOut) {
        float sb = 0.0f;
        for (int b = 0; b < kTile; ++b) sb += S.o[b * kOutP + tid];
        G[oB3 + tid] = sb;
    }
    __syncthreads();  // h2T is overwritten below
    // dh2 = g W3 (B^T = W3^T [128][16]); dz2 = dh2 [h2 > 0]
    for (int t = wave * 4; t < wave * 4 + 4; ++t) {
        const int m0 = (t / 8) * 16, n0 = (t % 8) * 16;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma_tile(S.g3 + m0 * kOutP, kOutP, wb + bW3T + n0 * kOutP, kOutP, kOutP, acc);
        const int i = n0 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int b = m0 + (lane >> 4) * 4 + r;
            const float dz = bf2f(S.h2[b * kH + i]) > 0.0f ? acc[r] : 0.0f;
            S.dz[b * kH + i] = dz;
            const std::uint16_t h = f2bf(dz);
            S.h2[b * kH + i] = h;  // h2 -> dz2 (the mask was read by this lane)
            S.h2T[i * kTile + b] = h;
        }
    }
    __syncthreads();
    // db2; dW2 = dz2^T h1: [128 x 128], 64 tiles, 16 per wave
    if (tid < kH) {
        float sb = 0.0f;
        for (int b = 0; b < kTile; ++b) sb += S.dz[b * kH + tid];
        G[oB2 + tid] = sb;
    }
    for (int t = wave * 16; t < wave * 16 + 16; ++t) {
        const int m0 = (t / 8) * 16, n0 = (t % 8) * 16;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma_tile(S.h2T + m0 * kTile, kTile, S.h1T + n0 * kTile, kTile, kTile, acc);
        const int i = n0 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) G[oW2 + (m0 + (lane >> 4) * 4 + r) * kH + i] = acc[r];
    }
    __syncthreads();  // dz and h1T are overwritten below
    // dh1 = dz2 W2 (B^T = W2^T); dz1 = dh1 [h1 > 0], stored transposed in h1T
    for (int t = wave * 4; t < wave * 4 + 4; ++t) {
        const int m0 = (t / 8) * 16, n0 = (t % 8) * 16;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma_tile(S.h2 + m0 * kH, kH, wb + bW2T + n0 * kH, kH, kH, acc);
        const int i = n0 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int b = m0 + (lane >> 4) * 4 + r;
            const float dz = bf2f(S.h1[b * kH + i]) > 0.0f ? acc[r] : 0.0f;
            S.dz[b * kH + i] = dz;
            S.h1T[i * kTile + b] = f2bf(dz);
        }
    }
    __syncthreads();
    // db1; dW1 = dz1^T x: [128 x 16], 8 tiles, 2 per wave
    if (tid < kH) {
        float sb = 0.0f;
        for (int b = 0; b < kTile; ++b) sb += S.dz[b * kH + tid];
        G[oB1 + tid] = sb;
    }
    for (int t = wave * 2; t < wave * 2 + 2; ++t) {
        const int m0 = t * 16;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma_tile(S.h1T + m0 * kTile, kTile, S.xbT, kTile, kTile, acc);
        const int i = lane & 15;
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (i < kIn) G[oW1 + (m0 + (lane >> 4) * 4 + r) * kIn + i] = acc[r];
    }
}

// BF16 operand copies of parameter j
__device__ inline void pack_param(int j, float v, std::uint16_t *wb) {
    const std::uint16_t h = f2bf(v);
    if (j < oB1) {
        const int o = j / kIn, i = j % kIn;
        wb[bW1 + o * kInP + i] = h;
    } else if (j >= oW2 && j < oB2) {
        const int o = (j - oW2) / kH, i = (j - oW2) % kH;
        wb[bW2 + o * kH + i] = h;
        wb[bW2T + i * kH + o] = h;
    } else if (j >= oW3 && j < oB3) {
        const int q = (j - oW3) / kH, i = (j - oW3) % kH;
        wb[bW3 + q * kH + i] = h;
        wb[bW3T + i * kOutP + q] = h;
    }
}

__global__ void pack_kernel(const float *p, std::uint16_t *wb) {
    const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (j < kParams) pack_param(j, p[j], wb);
}

// sum of the partial gradients (workgroup order) -> grad
__global__ void reduce_kernel(const float *part, int n_part, float *grad) {
    const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= kParams) return;
    float g = 0.0f;
    for (int w = 0; w < n_part; ++w) g += part[static_cast<index_t>(w) * kParams + j];
    grad[j] = g;
}

// Adam on the FP32 masters (bias-corrected), then the BF16 copies
__global__ void adam_kernel(const float *part, int n_part, float *p, float *m, float *v, float lr, float b1, float b2, float eps, float c1,
                            float c2, std::uint16_t *wb) {
    const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= kParams) return;
    float g = 0.0f;
    for (int w = 0; w < n_part; ++w) g += part[static_cast<index_t>(w) * kParams + j];
    const float mj = b1 * m[j] + (1.0f - b1) * g;
    const float vj = b2 * v[j] + (1.0f - b2) * g * g;
    m[j] = mj;
    v[j] = vj;
    const float pj = p[j] - lr * (mj / c1) / (sqrtf(vj / c2) + eps);
    p[j] = pj;
    pack_param(j, pj, wb);
}

__global__ void sum_kernel(const float *x, int n, float *out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float s = 0.0f;
        for (int i = 0; i < n; ++i) s += x[i];
        *out = s;
    }
}

}  // namespace

void init_params(std::uint64_t seed, float *out) {
    // SplitMix64 stream -> U(-a, a): He-uniform for the ReLU layers, a tenth
    // of it for the output layer; zero biases
    std::uint64_t k = 0;
    auto uniform = [&]() {
        std::uint64_t z = seed + (++k) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
    };
    std::fill(out, out + kParams, 0.0f);
    const double a1 = std::sqrt(6.0 / kIn), a2 = std::sqrt(6.0 / kH), a3 = 0.1 * std::sqrt(6.0 / kH);
    for (int j = oW1; j < oB1; ++j) out[j] = static_cast<float>((2.0 * uniform() - 1.0) * a1);
    for (int j = oW2; j < oB2; ++j) out[j] = static_cast<float>((2.0 * uniform() - 1.0) * a2);
    for (int j = oW3; j < oB3; ++j) out[j] = static_cast<float>((2.0 * uniform() - 1.0) * a3);
}

Model::Model(const float *params_host, float feature_scale) : feature_scale_(feature_scale) {
    if (!(feature_scale > 0.0f)) throw InvalidArgument("Model: feature scale must be positive");
    p_.resize(kParams);
    m_.resize(kParams);
    v_.resize(kParams);
    wb_.resize(kPacked);
    AXIS_HIP_CHECK(hipMemset(wb_.get(), 0, sizeof(std::uint16_t) * kPacked));
    set_params(params_host, nullptr);
}

void Model::set_params(const float *params_host, hipStream_t s) {
    p_.upload(params_host, kParams, s);
    AXIS_HIP_CHECK(hipMemsetAsync(m_.get(), 0, sizeof(float) * kParams, s));
    AXIS_HIP_CHECK(hipMemsetAsync(v_.get(), 0, sizeof(float) * kParams, s));
    step_ = 0;
    pack_kernel<<<grid_for(kParams, 256), 256, 0, s>>>(p_.get(), wb_.get());
    AXIS_LAUNCH_CHECK();
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
}

void Model::get_params(float *params_host, hipStream_t s) const {
    p_.download(params_host, kParams, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
}

void Model::forward(const float *x, index_t n, float *out, hipStream_t s) const {
    if (n <= 0) return;
    forward_kernel<<<grid_for(static_cast<std::size_t>(n), kTile), kThreads, 0, s>>>(wb_.get(), p_.get(), x, n, out);
    AXIS_LAUNCH_CHECK();
}

float Model::gradient(const float *x, const float *t, index_t n, float wc, float tau, float *grad, hipStream_t s) {
    if (n <= 0) throw InvalidArgument("Model::gradient: no samples");
    const unsigned nt = grid_for(static_cast<std::size_t>(n), kTile);
    part_.resize(static_cast<std::size_t>(nt) * kParams);
    loss_.resize(nt + 1);
    grad_kernel<<<nt, kThreads, 0, s>>>(wb_.get(), p_.get(), x, t, 0, n, 1.0f / static_cast<float>(n), wc, tau, part_.get(), loss_.get());
    AXIS_LAUNCH_CHECK();
    reduce_kernel<<<grid_for(kParams, 256), 256, 0, s>>>(part_.get(), static_cast<int>(nt), grad);
    AXIS_LAUNCH_CHECK();
    sum_kernel<<<1, 64, 0, s>>>(loss_.get(), static_cast<int>(nt), loss_.get() + nt);
    AXIS_LAUNCH_CHECK();
    float l = 0.0f;
    AXIS_HIP_CHECK(hipMemcpyAsync(&l, loss_.get() + nt, sizeof(float), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    return l;
}

float Model::train(const float *x, const float *t, index_t n, index_t batch, float lr, float wc, float tau, hipStream_t s) {
    if (n <= 0 || batch <= 0) throw InvalidArgument("Model::train: no samples");
    batch = (batch + kTile - 1) / kTile * kTile;
    const unsigned nt = grid_for(static_cast<std::size_t>(batch), kTile);
    part_.resize(static_cast<std::size_t>(nt) * kParams);
    const index_t n_steps = (n + batch - 1) / batch;
    loss_.resize(static_cast<std::size_t>(nt) * n_steps + 1);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (index_t k = 0; k < n_steps; ++k) {
        const index_t s0 = k * batch, s1 = std::min(n, s0 + batch);
        const unsigned ntk = grid_for(static_cast<std::size_t>(s1 - s0), kTile);
        grad_kernel<<<ntk, kThreads, 0, s>>>(wb_.get(), p_.get(), x, t, s0, s1, 1.0f / static_cast<float>(s1 - s0), wc, tau, part_.get(),
                                             loss_.get() + static_cast<std::size_t>(k) * nt);
        AXIS_LAUNCH_CHECK();
        ++step_;
        const float c1 = 1.0f - std::pow(b1, static_cast<float>(step_)), c2 = 1.0f - std::pow(b2, static_cast<float>(step_));
        adam_kernel<<<grid_for(kParams, 256), 256, 0, s>>>(part_.get(), static_cast<int>(ntk), p_.get(), m_.get(), v_.get(), lr, b1, b2, eps, c1,
                                                           c2, wb_.get());
        AXIS_LAUNCH_CHECK();
        if (ntk < nt) AXIS_HIP_CHECK(hipMemsetAsync(loss_.get() + static_cast<std::size_t>(k) * nt + ntk, 0, sizeof(float) * (nt - ntk), s));
    }
    // mean of the per-step losses
    std::vector<float> h(static_cast<std::size_t>(nt) * n_steps);
    loss_.download(h.data(), h.size(), s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    double tot = 0.0;
    for (float v : h) tot += v;
    return static_cast<float>(tot / static_cast<double>(n_steps));
}

}  // namespace mlp
}  // namespace axis_hip
