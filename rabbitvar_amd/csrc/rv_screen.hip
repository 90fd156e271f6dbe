// Realignment-screen kernels for gfx950: candidate features (one thread per candidate) and the MLP's
// inference and training steps on MFMA (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 sums).
//
// MLP work layout: a workgroup is MLP_WAVES wavefronts; a wavefront takes 16 samples at a time (the
// M dimension of one MFMA tile).  The weights sit in LDS as BF16; each wavefront keeps its tile's
// activations and gradients in its own LDS slice.  Weight gradients over a wavefront's samples are
// MFMA products with the samples as the K dimension, so they accumulate in registers across the
// tiles a wavefront visits; the workgroup adds them up in LDS and then once into the global
// gradient.
#include <hip/hip_runtime.h>

#include "rv_screen.h"
#include "rv_screen_kernels.h"

namespace rv {

namespace {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int MLP_WAVES = 4;
constexpr int MLP_THREADS = 64 * MLP_WAVES;
constexpr int MLP_TILE = 16;
constexpr int MLP_TILES_PER_WAVE = 4;  // training: 4 waves x 4 tiles x 16 = 256 samples per workgroup
constexpr int MLP_WG_SAMPLES = MLP_WAVES * MLP_TILES_PER_WAVE * MLP_TILE;

struct MlpWeights {
    uint16_t w1[SCR_F * SCR_H];  // W1[f][h]
    uint16_t w2[SCR_H * SCR_H];  // W2[k][n]
    float b1[SCR_H], b2[SCR_H];
    float w3[SCR_H];  // BF16-rounded values
    float b3;
};
struct MlpTile {
    uint16_t x[MLP_TILE * SCR_F];
    uint16_t h1[MLP_TILE * SCR_H];
    uint16_t h2[MLP_TILE * SCR_H];
    uint16_t g2[MLP_TILE * SCR_H];
    uint16_t g1[MLP_TILE * SCR_H];
    float dy[MLP_TILE];
};

__device__ inline bf4 load4(const uint16_t* p, int stride) {
    bf4 v;
    v[0] = (short)p[0];
    v[1] = (short)p[stride];
    v[2] = (short)p[2 * stride];
    v[3] = (short)p[3 * stride];
    return v;
}

__device__ void load_weights(MlpWeights& W, const float* P, int tid) {
    for (int i = tid; i < SCR_F * SCR_H; i += MLP_THREADS) W.w1[i] = f2bf(P[SCR_W1 + i]);
    for (int i = tid; i < SCR_H * SCR_H; i += MLP_THREADS) W.w2[i] = f2bf(P[SCR_W2 + i]);
    for (int i = tid; i < SCR_H; i += MLP_THREADS) {
        W.b1[i] = P[SCR_B1 + i];
        W.b2[i] = P[SCR_B2 + i];
        W.w3[i] = bf2f(f2bf(P[SCR_W3 + i]));
    }
    if (tid == 0) W.b3 = P[SCR_B3];
}

// Forward pass of one 16-sample tile (rows base .. base+15 of X; rows past n are zero).  Returns,
// in every lane, y of sample lane % 16.  Needs the block-wide barrier after the X stage, so every
// wavefront of the workgroup calls it the same number of times.
__device__ float forward_tile(const MlpWeights& W, MlpTile& T, const float* X, const int32_t* idx, int64_t base,
                              int64_t n, int lane) {
    for (int e = lane; e < MLP_TILE * SCR_F; e += 64) {
        const int i = e / SCR_F, f = e % SCR_F;
        const int64_t s = base + i;
        float v = 0.f;
        if (s < n) v = X[(size_t)(idx ? idx[s] : s) * SCR_F + f];
        T.x[e] = f2bf(v);
    }
    __syncthreads();
    const int r = lane & 15, g = lane >> 4;
    // layer 1: [16 x 32] x [32 x 64]
    for (int nt = 0; nt < SCR_H / 16; nt++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < SCR_F; k0 += 16) {
            const bf4 a = load4(&T.x[r * SCR_F + k0 + 4 * g], 1);
            const bf4 b = load4(&W.w1[(k0 + 4 * g) * SCR_H + nt * 16 + r], SCR_H);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
        const int nn = nt * 16 + r;
        for (int j = 0; j < 4; j++) {
            const float z = acc[j] + W.b1[nn];
            T.h1[(4 * g + j) * SCR_H + nn] = f2bf(z > 0.f ? z : 0.f);
        }
    }
    __syncthreads();
    // layer 2: [16 x 64] x [64 x 64]
    for (int nt = 0; nt < SCR_H / 16; nt++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < SCR_H; k0 += 16) {
            const bf4 a = load4(&T.h1[r * SCR_H + k0 + 4 * g], 1);
            const bf4 b = load4(&W.w2[(k0 + 4 * g) * SCR_H + nt * 16 + r], SCR_H);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
        const int nn = nt * 16 + r;
        for (int j = 0; j < 4; j++) {
            const float z = acc[j] + W.b2[nn];
            T.h2[(4 * g + j) * SCR_H + nn] = f2bf(z > 0.f ? z : 0.f);
        }
    }
    __syncthreads();
    // layer 3: lane (r, g) sums h2[r][16g .. 16g+15] * w3, then the four groups are added
    float s = 0.f;
    for (int q = 0; q < 16; q++) s += bf2f(T.h2[r * SCR_H + 16 * g + q]) * W.w3[16 * g + q];
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    return s + W.b3;
}

}  // namespace

__global__ void __launch_bounds__(256) k_screen_features(ScreenIn S) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < S.n) screen_features(S, i);
}

// Scores: out[s] = sigmoid(y) for s < n.  Grid-strided over 16-sample tiles.
__global__ void __launch_bounds__(MLP_THREADS) k_mlp_forward(const float* P, const float* X, int64_t n, float* out) {
    __shared__ MlpWeights W;
    __shared__ MlpTile T[MLP_WAVES];
    const int tid = (int)threadIdx.x, wave = tid / 64, lane = tid % 64;
    load_weights(W, P, tid);
    __syncthreads();
    const int64_t tiles = (n + MLP_TILE - 1) / MLP_TILE;
    const int64_t rounds = (tiles + (int64_t)gridDim.x * MLP_WAVES - 1) / ((int64_t)gridDim.x * MLP_WAVES);
    for (int64_t rd = 0; rd < rounds; rd++) {
        const int64_t tile = (rd * gridDim.x + blockIdx.x) * MLP_WAVES + wave;
        const int64_t base = tile * MLP_TILE;
        const float y = forward_tile(W, T[wave], X, nullptr, base, n, lane);
        if (lane < 16 && base + lane < n) out[base + lane] = 1.f / (1.f + expf(-y));
        __syncthreads();  // the tile buffers are reused next round
    }
}

// One training step's gradient over the batch idx[0 .. nb): grad += d(loss)/d(params), loss += the
// batch's summed weighted BCE times inv_batch.  The caller zeroes grad and loss.
__global__ void __launch_bounds__(MLP_THREADS) k_mlp_grad(const float* P, const float* X, const uint8_t* labels,
                                                         const int32_t* idx, int64_t nb, float inv_batch,
                                                         float pos_weight, float* grad, float* loss) {
    __shared__ MlpWeights W;
    __shared__ MlpTile T[MLP_WAVES];
    __shared__ float G[SCR_NPARAM];
    const int tid = (int)threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int r = lane & 15, g = lane >> 4;
    load_weights(W, P, tid);
    for (int i = tid; i < SCR_NPARAM; i += MLP_THREADS) G[i] = 0.f;
    __syncthreads();
    MlpTile& Tw = T[wave];
    f4 dw2[4][4], dw1[2][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) dw2[a][b] = f4{0.f, 0.f, 0.f, 0.f};
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 4; b++) dw1[a][b] = f4{0.f, 0.f, 0.f, 0.f};
    float lsum = 0.f, db3 = 0.f, dw3 = 0.f, db2 = 0.f, db1 = 0.f;
    for (int t = 0; t < MLP_TILES_PER_WAVE; t++) {
        const int64_t base = (int64_t)blockIdx.x * MLP_WG_SAMPLES + ((int64_t)t * MLP_WAVES + wave) * MLP_TILE;
        const float y = forward_tile(W, Tw, X, idx, base, nb, lane);
        // dL/dy and the loss, one lane per sample
        if (lane < 16) {
            float dy = 0.f;
            const int64_t s = base + lane;
            if (s < nb) {
                const int tl = labels[idx[s]] ? 1 : 0;
                const float wt = tl ? pos_weight : 1.f;
                const float p = 1.f / (1.f + expf(-y));
                dy = wt * (p - (float)tl) * inv_batch;
                // BCE as softplus(-y) for a positive, softplus(y) for a negative
                const float z = tl ? -y : y;
                lsum += wt * ((z > 0.f ? z : 0.f) + log1pf(expf(-fabsf(z))));
                db3 += dy;
            }
            Tw.dy[lane] = dy;
        }
        __syncthreads();
        // output layer and the gradient reaching h2 (lane = hidden unit n)
        {
            const int nn = lane;
            float a = 0.f, b2s = 0.f;
            for (int i = 0; i < MLP_TILE; i++) {
                const float h = bf2f(Tw.h2[i * SCR_H + nn]);
                a += Tw.dy[i] * h;
                const float gv = h > 0.f ? Tw.dy[i] * W.w3[nn] : 0.f;
                const uint16_t gb = f2bf(gv);
                Tw.g2[i * SCR_H + nn] = gb;
                b2s += bf2f(gb);
            }
            dw3 += a;
            db2 += b2s;
        }
        __syncthreads();
        // dW2 += h1^T g2 (samples as K); g1 = (g2 W2^T) masked by h1 > 0
        for (int mt = 0; mt < 4; mt++)
            for (int nt = 0; nt < 4; nt++) {
                const bf4 a = load4(&Tw.h1[(4 * g) * SCR_H + mt * 16 + r], SCR_H);
                const bf4 b = load4(&Tw.g2[(4 * g) * SCR_H + nt * 16 + r], SCR_H);
                dw2[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, dw2[mt][nt], 0, 0, 0);
            }
        for (int nt = 0; nt < 4; nt++) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < SCR_H; k0 += 16) {
                const bf4 a = load4(&Tw.g2[r * SCR_H + k0 + 4 * g], 1);
                const bf4 b = load4(&W.w2[(nt * 16 + r) * SCR_H + k0 + 4 * g], 1);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
            }
            const int k = nt * 16 + r;
            for (int j = 0; j < 4; j++) {
                const int i = 4 * g + j;
                Tw.g1[i * SCR_H + k] = bf2f(Tw.h1[i * SCR_H + k]) > 0.f ? f2bf(acc[j]) : (uint16_t)0;
            }
        }
        __syncthreads();
        {
            float s1 = 0.f;
            for (int i = 0; i < MLP_TILE; i++) s1 += bf2f(Tw.g1[i * SCR_H + lane]);
            db1 += s1;
        }
        // dW1 += X^T g1
        for (int mt = 0; mt < 2; mt++)
            for (int nt = 0; nt < 4; nt++) {
                const bf4 a = load4(&Tw.x[(4 * g) * SCR_F + mt * 16 + r], SCR_F);
                const bf4 b = load4(&Tw.g1[(4 * g) * SCR_H + nt * 16 + r], SCR_H);
                dw1[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, dw1[mt][nt], 0, 0, 0);
            }
        __syncthreads();  // the tile buffers are reused next round
    }
    // the wavefront's sums into the workgroup's
    for (int mt = 0; mt < 4; mt++)
        for (int nt = 0; nt < 4; nt++)
            for (int j = 0; j < 4; j++)
                atomicAdd(&G[SCR_W2 + (mt * 16 + 4 * g + j) * SCR_H + nt * 16 + r], dw2[mt][nt][j]);
    for (int mt = 0; mt < 2; mt++)
        for (int nt = 0; nt < 4; nt++)
            for (int j = 0; j < 4; j++)
                atomicAdd(&G[SCR_W1 + (mt * 16 + 4 * g + j) * SCR_H + nt * 16 + r], dw1[mt][nt][j]);
    atomicAdd(&G[SCR_W3 + lane], dw3);
    atomicAdd(&G[SCR_B2 + lane], db2);
    atomicAdd(&G[SCR_B1 + lane], db1);
    if (lane < 16) {
        atomicAdd(&G[SCR_B3], db3);
        atomicAdd(&loss[0], lsum * inv_batch);
    }
    __syncthreads();
    for (int i = tid; i < SCR_NPARAM; i += MLP_THREADS) atomicAdd(&grad[i], G[i]);
}

// Adam on the FP32 master parameters; clears the gradient.  c1, c2: bias corrections 1/(1-beta^t).
__global__ void __launch_bounds__(256) k_adam(float* P, float* grad, float* m, float* v, int n, float lr, float beta1,
                                              float beta2, float eps, float c1, float c2) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const float gi = grad[i];
    const float mi = beta1 * m[i] + (1.f - beta1) * gi;
    const float vi = beta2 * v[i] + (1.f - beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    grad[i] = 0.f;
}

// ---------------------------------------------------------------- launch wrappers
hipError_t launch_screen_features(const ScreenIn& S, hipStream_t s) {
    if (S.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_screen_features, dim3((unsigned)((S.n + 255) / 256)), dim3(256), 0, s, S);
    return hipGetLastError();
}
hipError_t launch_mlp_forward(const float* P, const float* X, int64_t n, float* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t tiles = (n + MLP_TILE - 1) / MLP_TILE;
    int64_t blocks = (tiles + MLP_WAVES - 1) / MLP_WAVES;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)blocks), dim3(MLP_THREADS), 0, s, P, X, n, out);
    return hipGetLastError();
}
hipError_t launch_mlp_grad(const float* P, const float* X, const uint8_t* labels, const int32_t* idx, int64_t nb,
                           float inv_batch, float pos_weight, float* grad, float* loss, hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    const int64_t blocks = (nb + MLP_WG_SAMPLES - 1) / MLP_WG_SAMPLES;
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)blocks), dim3(MLP_THREADS), 0, s, P, X, labels, idx, nb, inv_batch,
                       pos_weight, grad, loss);
    return hipGetLastError();
}
hipError_t launch_adam(float* P, float* grad, float* m, float* v, int n, float lr, float beta1, float beta2, float eps,
                       float c1, float c2, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P, grad, m, v, n, lr, beta1, beta2,
                       eps, c1, c2);
    return hipGetLastError();
}

}  // namespace rv
