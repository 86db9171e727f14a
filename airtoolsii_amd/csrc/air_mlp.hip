// air_mlp.hip -- the learned relaxation network.
//
// A 16-64-64-2 perceptron with ReLU hidden layers maps per-iteration residual
// statistics of a SIRT run to (o0, o1): lambda_k * rho = MLP_LAMBDA_MAX sigmoid(o0) and a
// stop logit o1.  Training runs here with hand-written kernels: forward and
// backward on MFMA (BF16 operands, FP32 accumulation) over tiles of 16
// samples per wave, per-block weight gradients reduced in a fixed order, and
// Adam on FP32 master weights.  Parameter layout (FP32, 5378 values):
//   W1[64][16], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]   (row = output)
#include "air_internal.h"
#include "air_device.h"

namespace air {

namespace {

constexpr int MI = MLP_IN, MH = MLP_HID, MO = MLP_OUT;
constexpr int P_W1 = 0, P_B1 = P_W1 + MH * MI, P_W2 = P_B1 + MH, P_B2 = P_W2 + MH * MH, P_W3 = P_B2 + MH,
              P_B3 = P_W3 + MO * MH, NPARAM = P_B3 + MO;
static_assert(NPARAM == MLP_PARAMS, "parameter count");
constexpr int WAVES = 4, NT = 64 * WAVES, TILE = 16;

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

__device__ __forceinline__ u16 bf16_bits(float f) {   // round to nearest even
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (u16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
__device__ __forceinline__ float bf16_val(u16 b) { return __uint_as_float((unsigned)b << 16); }
// ReLU derivative from a stored BF16 activation (+0 and -0 are inactive)
__device__ __forceinline__ bool active(u16 b) { return (b & 0x7fffu) != 0; }
// LDS writes of this wave become visible to its other lanes
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// D += A * B for 16x16x16 tiles.  Lane l holds A[l%16][4(l/16)+i],
// B[4(l/16)+i][l%16] and D[4(l/16)+i][l%16], i = 0..3.
__device__ __forceinline__ f4v mfma(bf4 a, bf4 b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// 4 consecutive BF16 of row r, from column k (the A / B operand of a lane)
__device__ __forceinline__ bf4 row4(const u16* m, int ld, int r, int k) {
    const u16* p = m + r * ld + k;
    return bf4{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
}
// 4 BF16 of column c, rows k..k+3 (an operand read across rows)
__device__ __forceinline__ bf4 col4(const u16* m, int ld, int k, int c) {
    return bf4{(short)m[k * ld + c], (short)m[(k + 1) * ld + c], (short)m[(k + 2) * ld + c],
               (short)m[(k + 3) * ld + c]};
}

struct WeightsLds {
    u16 w1[MH * MI];      // [out][in]
    u16 w2[MH * MH];      // [out][in]
    u16 w2t[MH * MH];     // [in][out]
    u16 w3[TILE * MH];    // [out][in], rows 2..15 zero
    u16 w3t[MH * TILE];   // [in][out], columns 2..15 zero
    float b1[MH], b2[MH], b3[MO];
};
struct WaveLds {
    u16 x[TILE * MI];     // [sample][feature]
    u16 h1[TILE * MH];    // post-ReLU activations
    u16 h2[TILE * MH];
    u16 d2[TILE * MH];    // dL/dz2
    u16 d1[TILE * MH];    // dL/dz1
    u16 dout[TILE * TILE];
};

__device__ void stage_weights(const float* __restrict__ W, WeightsLds& L) {
    for (int i = threadIdx.x; i < MH * MI; i += blockDim.x) L.w1[i] = bf16_bits(W[P_W1 + i]);
    for (int i = threadIdx.x; i < MH * MH; i += blockDim.x) {
        const u16 v = bf16_bits(W[P_W2 + i]);
        L.w2[i] = v;
        L.w2t[(i % MH) * MH + i / MH] = v;
    }
    for (int i = threadIdx.x; i < TILE * MH; i += blockDim.x) {
        const int o = i / MH, j = i % MH;
        const u16 v = o < MO ? bf16_bits(W[P_W3 + o * MH + j]) : (u16)0;
        L.w3[i] = v;
        L.w3t[j * TILE + o] = v;
    }
    for (int i = threadIdx.x; i < MH; i += blockDim.x) {
        L.b1[i] = W[P_B1 + i];
        L.b2[i] = W[P_B2 + i];
    }
    if (threadIdx.x < MO) L.b3[threadIdx.x] = W[P_B3 + threadIdx.x];
}

// Forward pass of one 16-sample tile in one wave; returns the raw outputs of
// this lane: D[4(l/16)+q][l%16] + b3, valid for l%16 < 2.
__device__ f4v forward_tile(const WeightsLds& L, WaveLds& S, int lane) {
    const int r = lane % 16, k4 = 4 * (lane / 16);
    for (int nb = 0; nb < MH / 16; ++nb) {                  // layer 1
        f4v acc = {0, 0, 0, 0};
        acc = mfma(row4(S.x, MI, r, k4), row4(L.w1, MI, nb * 16 + r, k4), acc);
        const int o = nb * 16 + r;
#pragma unroll
        for (int q = 0; q < 4; ++q) S.h1[(k4 + q) * MH + o] = bf16_bits(fmaxf(acc[q] + L.b1[o], 0.f));
    }
    wave_sync();
    for (int nb = 0; nb < MH / 16; ++nb) {                  // layer 2
        f4v acc = {0, 0, 0, 0};
        for (int kb = 0; kb < MH / 16; ++kb)
            acc = mfma(row4(S.h1, MH, r, kb * 16 + k4), row4(L.w2, MH, nb * 16 + r, kb * 16 + k4), acc);
        const int o = nb * 16 + r;
#pragma unroll
        for (int q = 0; q < 4; ++q) S.h2[(k4 + q) * MH + o] = bf16_bits(fmaxf(acc[q] + L.b2[o], 0.f));
    }
    wave_sync();
    f4v acc = {0, 0, 0, 0};                                  // layer 3
    for (int kb = 0; kb < MH / 16; ++kb)
        acc = mfma(row4(S.h2, MH, r, kb * 16 + k4), row4(L.w3, MH, r, kb * 16 + k4), acc);
    const float bias = r < MO ? L.b3[r] : 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] += bias;
    return acc;
}

__device__ void load_tile(const float* __restrict__ X, const int* __restrict__ idx, int64_t base, int64_t n,
                          WaveLds& S, int lane) {
    for (int i = lane; i < TILE * MI; i += 64) {
        const int s = i / MI, f = i % MI;
        const int64_t j = base + s;
        float v = 0.f;
        if (j < n) v = X[(int64_t)(idx ? idx[j] : j) * MI + f];
        S.x[i] = bf16_bits(v);
    }
    wave_sync();
}

__device__ __forceinline__ float sigmoidf_(float z) { return 1.f / (1.f + __expf(-z)); }

}  // namespace

// Raw outputs (o0, o1) for n samples.
__global__ __launch_bounds__(NT) void k_mlp_forward(const float* __restrict__ W, const float* __restrict__ X,
                                                    int64_t n, float* __restrict__ out) {
    __shared__ WeightsLds L;
    __shared__ WaveLds S[WAVES];
    stage_weights(W, L);
    __syncthreads();
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    for (int64_t base = ((int64_t)blockIdx.x * WAVES + w) * TILE; base < n; base += (int64_t)gridDim.x * WAVES * TILE) {
        load_tile(X, nullptr, base, n, S[w], lane);
        const f4v o = forward_tile(L, S[w], lane);
        const int r = lane % 16, k4 = 4 * (lane / 16);
        if (r < MO)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (base + k4 + q < n) out[(base + k4 + q) * MO + r] = o[q];
        wave_sync();
    }
}

// Loss and gradient of one minibatch (samples idx[0..nb)): per-block partial
// gradients part[block][NPARAM] and loss partials.  Loss per sample:
//   (MLP_LAMBDA_MAX sigmoid(o0) - t0)^2 + stop_w * BCE(o1, t1),   averaged over nb.
// A target t0 < 0 drops the lambda term of that sample.
__global__ __launch_bounds__(NT) void k_mlp_grad(const float* __restrict__ W, const float* __restrict__ X,
                                                 const float* __restrict__ T, const int* __restrict__ idx,
                                                 int64_t nb, float stop_w, float* __restrict__ part,
                                                 double* __restrict__ loss_part) {
    __shared__ WeightsLds L;
    __shared__ WaveLds S[WAVES];
    __shared__ double lsh[WAVES];
    static_assert(sizeof(S) >= sizeof(float) * NPARAM, "gradient staging fits in the tile buffers");
    stage_weights(W, L);
    __syncthreads();
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int r = lane % 16, k4 = 4 * (lane / 16);
    const float inv_n = 1.f / (float)nb;
    f4v g1[4], g2[16], g3[4];
    float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb3 = 0.f;
    for (int i = 0; i < 4; ++i) g1[i] = g3[i] = f4v{0, 0, 0, 0};
    for (int i = 0; i < 16; ++i) g2[i] = f4v{0, 0, 0, 0};
    double loss = 0;
    WaveLds& Sw = S[w];
    for (int64_t base = ((int64_t)blockIdx.x * WAVES + w) * TILE; base < nb; base += (int64_t)gridDim.x * WAVES * TILE) {
        load_tile(X, idx, base, nb, Sw, lane);
        const f4v o = forward_tile(L, Sw, lane);
        // output gradients (lane owns output r of samples k4+q)
        f4v dout = {0, 0, 0, 0};
        if (r < MO) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t j = base + k4 + q;
                if (j >= nb) continue;
                const int64_t sidx = idx ? idx[j] : j;
                const float t = T[sidx * MO + r];
                const float sg = sigmoidf_(o[q]);
                if (r == 0) {
                    if (t >= 0.f) {
                        const float e = (float)MLP_LAMBDA_MAX * sg - t;
                        loss += (double)(e * e);
                        dout[q] = 2.f * e * (float)MLP_LAMBDA_MAX * sg * (1.f - sg) * inv_n;
                    }
                } else {
                    const float z = o[q];
                    // BCE with logits: max(z,0) - z t + log(1 + exp(-|z|))
                    loss += (double)(stop_w * (fmaxf(z, 0.f) - z * t + log1pf(__expf(-fabsf(z)))));
                    dout[q] = stop_w * (sg - t) * inv_n;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) Sw.dout[(k4 + q) * TILE + r] = bf16_bits(r < MO ? dout[q] : 0.f);
        if (r < MO) gb3 += dout[0] + dout[1] + dout[2] + dout[3];
        wave_sync();
        // dz2 = (dout W3) * [h2 > 0];  dW3 += dout^T h2
        for (int nb2 = 0; nb2 < MH / 16; ++nb2) {
            f4v acc = {0, 0, 0, 0};
            acc = mfma(row4(Sw.dout, TILE, r, k4), row4(L.w3t, TILE, nb2 * 16 + r, k4), acc);
            const int jcol = nb2 * 16 + r;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float dz = active(Sw.h2[(k4 + q) * MH + jcol]) ? acc[q] : 0.f;
                Sw.d2[(k4 + q) * MH + jcol] = bf16_bits(dz);
                gb2[nb2] += dz;
            }
            g3[nb2] = mfma(col4(Sw.dout, TILE, k4, r), col4(Sw.h2, MH, k4, nb2 * 16 + r), g3[nb2]);
        }
        wave_sync();
        // dz1 = (dz2 W2) * [h1 > 0];  dW2 += dz2^T h1
        for (int nb1 = 0; nb1 < MH / 16; ++nb1) {
            f4v acc = {0, 0, 0, 0};
            for (int kb = 0; kb < MH / 16; ++kb)
                acc = mfma(row4(Sw.d2, MH, r, kb * 16 + k4), row4(L.w2t, MH, nb1 * 16 + r, kb * 16 + k4), acc);
            const int icol = nb1 * 16 + r;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float dz = active(Sw.h1[(k4 + q) * MH + icol]) ? acc[q] : 0.f;
                Sw.d1[(k4 + q) * MH + icol] = bf16_bits(dz);
                gb1[nb1] += dz;
            }
        }
        for (int jb = 0; jb < MH / 16; ++jb)
            for (int ib = 0; ib < MH / 16; ++ib)
                g2[jb * 4 + ib] = mfma(col4(Sw.d2, MH, k4, jb * 16 + r), col4(Sw.h1, MH, k4, ib * 16 + r),
                                       g2[jb * 4 + ib]);
        wave_sync();
        // dW1 += dz1^T x
        for (int ib = 0; ib < MH / 16; ++ib)
            g1[ib] = mfma(col4(Sw.d1, MH, k4, ib * 16 + r), col4(Sw.x, MI, k4, r), g1[ib]);
        wave_sync();
    }
    // combine the four waves in a fixed order (wave 0 stores, waves 1..3 add)
    // in LDS that the tiles no longer use
    __syncthreads();
    float* red = reinterpret_cast<float*>(S);
    for (int ww = 0; ww < WAVES; ++ww) {
        if (w == ww) {
            auto put = [&](int p, float v) {
                if (ww == 0) red[p] = v; else red[p] += v;
            };
            for (int ib = 0; ib < 4; ++ib)
#pragma unroll
                for (int q = 0; q < 4; ++q) put(P_W1 + (ib * 16 + k4 + q) * MI + r, g1[ib][q]);
            for (int jb = 0; jb < 4; ++jb)
                for (int ib = 0; ib < 4; ++ib)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        put(P_W2 + (jb * 16 + k4 + q) * MH + ib * 16 + r, g2[jb * 4 + ib][q]);
            for (int jb = 0; jb < 4; ++jb)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (k4 + q < MO) put(P_W3 + (k4 + q) * MH + jb * 16 + r, g3[jb][q]);
            // biases: lanes r, r+16, r+32, r+48 hold partial sums of the same column
            for (int c = 0; c < 4; ++c) {
                float s1 = gb1[c], s2 = gb2[c];
                s1 += __shfl_xor(s1, 16); s1 += __shfl_xor(s1, 32);
                s2 += __shfl_xor(s2, 16); s2 += __shfl_xor(s2, 32);
                if (lane < 16) {
                    put(P_B1 + c * 16 + r, s1);
                    put(P_B2 + c * 16 + r, s2);
                }
            }
            float s3 = gb3;
            s3 += __shfl_xor(s3, 16); s3 += __shfl_xor(s3, 32);
            if (lane < MO) put(P_B3 + lane, s3);
        }
        __syncthreads();
    }
    for (int p = threadIdx.x; p < NPARAM; p += NT) part[(int64_t)blockIdx.x * NPARAM + p] = red[p];
    double lw = loss;
    for (int d = 32; d > 0; d >>= 1) lw += __shfl_xor(lw, d);
    if (lane == 0) lsh[w] = lw;
    __syncthreads();
    if (threadIdx.x == 0) loss_part[blockIdx.x] = (lsh[0] + lsh[1]) + (lsh[2] + lsh[3]);
}

// grad[p] = sum over blocks (fixed order); optional Adam step on FP32 master weights.
__global__ void k_mlp_reduce(const float* __restrict__ part, int nblk, float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    float g = 0.f;
    for (int b = 0; b < nblk; ++b) g += part[(int64_t)b * NPARAM + p];
    grad[p] = g;
}

__global__ void k_mlp_adam(float* __restrict__ W, float* __restrict__ m1, float* __restrict__ m2,
                           const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1,
                           float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    const float g = grad[p];
    const float a = b1 * m1[p] + (1.f - b1) * g;
    const float v = b2 * m2[p] + (1.f - b2) * g * g;
    m1[p] = a;
    m2[p] = v;
    W[p] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// ---------------------------------------------------------------- host
void mlp_forward(const float* W, const float* X, int64_t n, float* out, hipStream_t s) {
    if (n <= 0) return;
    const unsigned nblk = (unsigned)std::min<int64_t>((n + WAVES * TILE - 1) / (WAVES * TILE), 1024);
    k_mlp_forward<<<nblk, NT, 0, s>>>(W, X, n, out);
    AIR_HIP(hipGetLastError());
}

unsigned mlp_grad_blocks(int64_t nb) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((nb + WAVES * TILE - 1) / (WAVES * TILE), 256));
}

void mlp_grad(const float* W, const float* X, const float* T, const int* idx, int64_t nb, float stop_w,
              float* part, double* loss_part, float* grad, hipStream_t s) {
    const unsigned nblk = mlp_grad_blocks(nb);
    k_mlp_grad<<<nblk, NT, 0, s>>>(W, X, T, idx, nb, stop_w, part, loss_part);
    AIR_HIP(hipGetLastError());
    k_mlp_reduce<<<(NPARAM + 255) / 256, 256, 0, s>>>(part, (int)nblk, grad);
    AIR_HIP(hipGetLastError());
}

void mlp_adam(float* W, float* m1, float* m2, const float* grad, float lr, int step, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
    k_mlp_adam<<<(NPARAM + 255) / 256, 256, 0, s>>>(W, m1, m2, grad, lr, b1, b2, 1e-8f, c1, c2);
    AIR_HIP(hipGetLastError());
}

}  // namespace air
