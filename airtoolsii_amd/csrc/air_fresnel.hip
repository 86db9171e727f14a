// air_fresnel.hip -- seismicwavetomo: Fresnel-zone sensitivity matrix on gfx950.
//
// Reference algorithm: testprobs/seismicwavetomo.m:114-288.  Row (i, j) is the
// sensitivity kernel of the wave from source S_i = (N/2, y0_i) to receiver
// R_j = (xp_j, yp_j) at every pixel centre X (x = -N/2+0.5+c, y = N/2-0.5-r):
//   delta = |X - S| + |X - R| - |S - R|
//   K(X)  = cos(2 pi delta w) exp(-(alpha delta w)^2),  w = omega/N, alpha = 10
//   A(row, X) = |S - R| K(X) / sum_X K(X),  kept where it is >= 1e-6
// (simple_fat_kernel, :273-288; threshold :117,210), pixels in column-major
// order (find, :212-228), so each row's columns come out ascending.
//
// GPU form: one 256-lane workgroup per row.  The first kernel evaluates K over
// the image, reduces its sum in FP64 and counts the kept pixels; the second
// re-evaluates K and writes the kept pixels in order with a ballot + LDS
// prefix per 256-pixel round.  All arithmetic is FP64 in the reference's
// operation order (-ffp-contract=off); cos and exp are the device library's,
// so values agree with the CPU restatement to a few ulp, not bit for bit.
#include <cmath>
#include <vector>

#include "air_internal.h"

namespace air {

namespace {

constexpr int FZ_BLOCK = 256;
constexpr double FZ_TOL = 1e-6;      // seismicwavetomo.m:117
constexpr double FZ_ALPHA = 10.0;    // seismicwavetomo.m:277

struct FresnelRow {
    double sx, sy, rx, ry, dist;
};

__device__ __forceinline__ double fresnel_k(const FresnelRow& g, int N, int64_t q, double w) {
    const double half = (double)N / 2.0;
    const int64_t c = q / N, r = q - c * N;
    const double x = (-half + 0.5) + (double)c;
    const double y = (half - 0.5) - (double)r;
    const double dxs = x - g.sx, dys = y - g.sy, dxr = x - g.rx, dyr = y - g.ry;
    const double tsx = sqrt(dxs * dxs + dys * dys);
    const double trx = sqrt(dxr * dxr + dyr * dyr);
    const double dt = (tsx + trx) - g.dist;
    const double a = (FZ_ALPHA * dt) * w;
    return cos(((2.0 * M_PI) * dt) * w) * exp(-(a * a));
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0;
    for (int k = 0; k < FZ_BLOCK / 64; ++k) t += sh[k];
    return t;
}

__global__ __launch_bounds__(FZ_BLOCK) void k_fresnel_count(const FresnelRow* __restrict__ rows_g, int64_t rows,
                                                            int N, double w, double* __restrict__ rowsum,
                                                            int64_t* __restrict__ len) {
    __shared__ double sh[FZ_BLOCK / 64];
    const int64_t n2 = (int64_t)N * N;
    for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const FresnelRow g = rows_g[row];
        double part = 0;
        for (int64_t q = threadIdx.x; q < n2; q += FZ_BLOCK) part += fresnel_k(g, N, q, w);
        const double sum = block_sum(part, sh);
        double kept = 0;
        for (int64_t q = threadIdx.x; q < n2; q += FZ_BLOCK) {
            const double v = (g.dist * fresnel_k(g, N, q, w)) / sum;
            kept += (v < FZ_TOL) ? 0.0 : 1.0;
        }
        const double cnt = block_sum(kept, sh);
        if (threadIdx.x == 0) {
            rowsum[row] = sum;
            len[row] = (int64_t)cnt;
        }
    }
}

template <typename V>
__global__ __launch_bounds__(FZ_BLOCK) void k_fresnel_fill(const FresnelRow* __restrict__ rows_g, int64_t rows, int N,
                                                           double w, const double* __restrict__ rowsum,
                                                           const int64_t* __restrict__ rowptr, int* __restrict__ col,
                                                           V* __restrict__ val) {
    __shared__ int wsum[FZ_BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t n2 = (int64_t)N * N;
    for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const FresnelRow g = rows_g[row];
        const double sum = rowsum[row];
        int64_t out = rowptr[row];
        for (int64_t q0 = 0; q0 < n2; q0 += FZ_BLOCK) {
            const int64_t q = q0 + threadIdx.x;
            double v = 0;
            bool keep = false;
            if (q < n2) {
                v = (g.dist * fresnel_k(g, N, q, w)) / sum;
                keep = !(v < FZ_TOL);
            }
            const uint64_t m = __ballot(keep);
            __syncthreads();                      // wsum of the previous round consumed
            if (lane == 0) wsum[wid] = __popcll(m);
            __syncthreads();
            int before = 0, total = 0;
            for (int k = 0; k < FZ_BLOCK / 64; ++k) {
                before += k < wid ? wsum[k] : 0;
                total += wsum[k];
            }
            if (keep) {
                const int64_t idx = out + before + __popcll(m & ((1ull << lane) - 1ull));
                col[idx] = (int)q;
                val[idx] = from_f64<V>(v);
            }
            out += total;
        }
    }
}

}  // namespace

void fresnel_to_csr(int N, int s, int p, double omega, int dtype, DevSparse& A, hipStream_t st) {
    AIR_CHECK(N > 0 && s > 0 && p > 0, AIR_E_ARG, "invalid seismicwavetomo arguments");
    const int64_t n = (int64_t)N * N;
    AIR_CHECK(n < (int64_t)INT32_MAX, AIR_E_ARG, "N too large for int32 column indices");
    const int64_t m = (int64_t)s * p;
    A.nrows = m;
    A.ncols = n;
    A.dtype = dtype;
    std::vector<double> y0, xp, yp;
    seismic_positions(N, s, p, y0, xp, yp);
    const double N2 = (double)N / 2.0;
    std::vector<FresnelRow> g((size_t)std::max<int64_t>(m, 1));
    for (int i = 0; i < s; ++i)
        for (int j = 0; j < p; ++j) {
            FresnelRow& r = g[(size_t)i * p + j];
            r.sx = N2;
            r.sy = y0[i];
            r.rx = xp[j];
            r.ry = yp[j];
            const double ex = r.sx - r.rx, ey = r.sy - r.ry;   // sqrt(sum((S-R).^2)), :281
            r.dist = std::sqrt(ex * ex + ey * ey);
        }
    const double w = omega / N;                          // simple_fat_kernel(R,S,omega/N,...), :209
    DevBuf dg, rowsum, len;
    dg.alloc(sizeof(FresnelRow) * g.size());
    AIR_HIP(hipMemcpyAsync(dg.p, g.data(), sizeof(FresnelRow) * g.size(), hipMemcpyHostToDevice, st));
    rowsum.alloc(8 * (size_t)std::max<int64_t>(m, 1));
    len.alloc(8 * (size_t)(m + 1));
    AIR_HIP(hipMemsetAsync(len.p, 0, 8 * (size_t)(m + 1), st));
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(m, 1 << 16));
    if (m > 0) {
        k_fresnel_count<<<grid, FZ_BLOCK, 0, st>>>(dg.as<FresnelRow>(), m, N, w, rowsum.as<double>(),
                                                   len.as<int64_t>());
        AIR_HIP(hipGetLastError());
    }
    A.ptr.alloc(8 * (size_t)(m + 1));
    scan_exclusive_i64(len.as<int64_t>(), A.ptr.as<int64_t>(), m, st);
    int64_t nnz = 0;
    AIR_HIP(hipMemcpyAsync(&nnz, A.ptr.as<int64_t>() + m, 8, hipMemcpyDeviceToHost, st));
    AIR_HIP(hipStreamSynchronize(st));
    A.nnz = nnz;
    alloc_sparse_arrays(A, nnz, dtype, st);
    if (m > 0) {
        dispatch_dtype(dtype, [&](auto v) {
            using V = decltype(v);
            k_fresnel_fill<V><<<grid, FZ_BLOCK, 0, st>>>(dg.as<FresnelRow>(), m, N, w, rowsum.as<double>(),
                                                         A.ptr.as<int64_t>(), A.idx.as<int>(), A.val.as<V>());
            return 0;
        });
        AIR_HIP(hipGetLastError());
    }
    AIR_HIP(hipStreamSynchronize(st));
}

}  // namespace air
