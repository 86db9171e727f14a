// air_multi.hip -- SIRT with many right-hand sides on one operator (SpMM).
//
// Up to 64 right-hand sides share A: the vectors are stored RHS-minor
// (X[j][0..63], R[i][0..63]), one wave walks one row (A X) or one column
// (A' M R) and lane l carries right-hand side l.  Every matrix entry is read
// once per pass for all right-hand sides; the gathered operand of an entry is
// one contiguous 256-byte row of X or M R.  The residual epilogue (r = b - A x,
// M r, per-RHS FP64 norms) and the update with projection are fused into the
// two passes.
#include "air_internal.h"
#include "air_device.h"

namespace air {

constexpr int MR_LANES = 64;          // right-hand sides per wave

namespace {

template <typename V>
__device__ __forceinline__ float entry_val(const V* val, int64_t q) { return to_f32(val[q]); }

// acc[lane] = sum over the entries [s, e) of val * Xg[col][lane]
template <typename V>
__device__ __forceinline__ float gather_dot(const int* __restrict__ idx, const V* __restrict__ val, int64_t s,
                                            int64_t e, const float* __restrict__ Xg, int lane) {
    float acc = 0.f;
    for (int64_t e0 = s; e0 < e; e0 += 64) {
        const int ne = (int)min<int64_t>(64, e - e0);
        const int c = lane < ne ? idx[e0 + lane] : 0;
        const float v = lane < ne ? entry_val(val, e0 + lane) : 0.f;
        int j = 0;
        for (; j + 4 <= ne; j += 4) {
            const int c0 = __builtin_amdgcn_readlane(c, j), c1 = __builtin_amdgcn_readlane(c, j + 1);
            const int c2 = __builtin_amdgcn_readlane(c, j + 2), c3 = __builtin_amdgcn_readlane(c, j + 3);
            const float x0 = Xg[(int64_t)c0 * MR_LANES + lane], x1 = Xg[(int64_t)c1 * MR_LANES + lane];
            const float x2 = Xg[(int64_t)c2 * MR_LANES + lane], x3 = Xg[(int64_t)c3 * MR_LANES + lane];
            acc += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j)) * x0;
            acc += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j + 1)) * x1;
            acc += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j + 2)) * x2;
            acc += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j + 3)) * x3;
        }
        for (; j < ne; ++j) {
            const int cj = __builtin_amdgcn_readlane(c, j);
            acc += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j)) * Xg[(int64_t)cj * MR_LANES + lane];
        }
    }
    return acc;
}

}  // namespace

// r = b - A x, M r and per-RHS partial norms (sum r^2, block totals per lane)
template <typename V>
__global__ __launch_bounds__(256) void k_multi_resid(const int64_t* __restrict__ ptr, const int* __restrict__ idx,
                                                     const V* __restrict__ val, int64_t m, int nrhs,
                                                     const float* __restrict__ X, const float* __restrict__ Bm,
                                                     const float* __restrict__ M, float* __restrict__ R,
                                                     float* __restrict__ MRv, double* __restrict__ part) {
    __shared__ double sh[4][MR_LANES];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double nrm = 0;
    for (int64_t i = (int64_t)blockIdx.x * 4 + w; i < m; i += (int64_t)gridDim.x * 4) {
        const float ax = gather_dot(idx, val, ptr[i], ptr[i + 1], X, lane);
        const int64_t o = i * MR_LANES + lane;
        const float r = lane < nrhs ? Bm[o] - ax : 0.f;
        R[o] = r;
        MRv[o] = M ? M[i] * r : r;
        nrm += (double)r * (double)r;
    }
    sh[w][lane] = nrm;
    __syncthreads();
    if (w == 0) part[(int64_t)blockIdx.x * MR_LANES + lane] = (sh[0][lane] + sh[1][lane]) + (sh[2][lane] + sh[3][lane]);
}

// x = P(x + lam D (A' M r)) column by column (CSC of A)
template <typename V>
__global__ __launch_bounds__(256) void k_multi_update(const int64_t* __restrict__ ptr, const int* __restrict__ idx,
                                                      const V* __restrict__ val, int64_t n, int nrhs,
                                                      const float* __restrict__ MRv, const float* __restrict__ D,
                                                      float* __restrict__ X, double lam, int lb_kind, float lb,
                                                      int ub_kind, float ub) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int64_t j = (int64_t)blockIdx.x * 4 + w; j < n; j += (int64_t)gridDim.x * 4) {
        const float g = gather_dot(idx, val, ptr[j], ptr[j + 1], MRv, lane);
        if (lane >= nrhs) continue;
        const int64_t o = j * MR_LANES + lane;
        const double dg = D ? (double)D[j] * (double)g : (double)g;
        X[o] = project((double)X[o] + lam * dg, (int)j, lb_kind, lb, nullptr, ub_kind, ub, nullptr);
    }
}

__global__ void k_multi_norms(const double* __restrict__ part, int nblk, double* __restrict__ out) {
    const int lane = threadIdx.x;
    double s = 0;
    for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * MR_LANES + lane];
    out[lane] = s;
}

unsigned multi_grid(int64_t rows, int num_cu) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((rows + 3) / 4, (int64_t)num_cu * 16));
}

void multi_resid(const DevSparse& A, int nrhs, const float* X, const float* B, const float* M, float* R, float* MRv,
                 double* part, unsigned grid, hipStream_t s) {
    dispatch_dtype(A.dtype, [&](auto v) {
        using V = decltype(v);
        k_multi_resid<V><<<grid, 256, 0, s>>>(A.ptr.as<int64_t>(), A.idx.as<int>(), A.val.as<V>(), A.nrows, nrhs, X,
                                              B, M, R, MRv, part);
        return 0;
    });
    AIR_HIP(hipGetLastError());
}

void multi_update(const DevSparse& AT, int nrhs, const float* MRv, const float* D, float* X, double lam, int lb_kind,
                  float lb, int ub_kind, float ub, unsigned grid, hipStream_t s) {
    dispatch_dtype(AT.dtype, [&](auto v) {
        using V = decltype(v);
        k_multi_update<V><<<grid, 256, 0, s>>>(AT.ptr.as<int64_t>(), AT.idx.as<int>(), AT.val.as<V>(), AT.nrows,
                                               nrhs, MRv, D, X, lam, lb_kind, lb, ub_kind, ub);
        return 0;
    });
    AIR_HIP(hipGetLastError());
}

void multi_norms(const double* part, int nblk, double* out, hipStream_t s) {
    k_multi_norms<<<1, MR_LANES, 0, s>>>(part, nblk, out);
    AIR_HIP(hipGetLastError());
}

}  // namespace air
