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

// ---------------------------------------------------------------- matrix-free form
// seismicwavetomo.m:94-99,257-264: the function handle re-evaluates the
// kernels on every product.  The row sums of K are computed once; every
// product then forms A(row, X) = |S - R| K(X) / sum exactly as the matrix
// builder does (same function, same sum), so the entries are the stored
// matrix's, narrowed to FP32 as a stored FP32 matrix would be.

// y[row] = sum_X A(row, X) u[X]: one work-group per row
__global__ __launch_bounds__(FZ_BLOCK) void k_fresnel_mf_fwd(const FresnelRow* __restrict__ rows_g, int64_t rows,
                                                             int N, double w, const double* __restrict__ rowsum,
                                                             const float* __restrict__ u, float* __restrict__ y) {
    __shared__ double sh[FZ_BLOCK / 64];
    const int64_t n2 = (int64_t)N * N;
    for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const FresnelRow g = rows_g[row];
        const double sum = rowsum[row];
        float acc = 0.f;
        for (int64_t q = threadIdx.x; q < n2; q += FZ_BLOCK) {
            const double v = (g.dist * fresnel_k(g, N, q, w)) / sum;
            if (!(v < FZ_TOL)) acc += (float)v * u[q];
        }
        const double t = block_sum((double)acc, sh);
        if (threadIdx.x == 0) y[row] = (float)t;
    }
}

// g[X] = sum_row A(row, X) v[row]: one thread per pixel, rows in order (no atomics)
__global__ __launch_bounds__(FZ_BLOCK) void k_fresnel_mf_adj(const FresnelRow* __restrict__ rows_g, int64_t rows,
                                                             int N, double w, const double* __restrict__ rowsum,
                                                             const float* __restrict__ v, double* __restrict__ g) {
    const int64_t n2 = (int64_t)N * N;
    for (int64_t q = (int64_t)blockIdx.x * FZ_BLOCK + threadIdx.x; q < n2; q += (int64_t)gridDim.x * FZ_BLOCK) {
        double acc = 0;
        for (int64_t row = 0; row < rows; ++row) {
            const float vi = v[row];
            if (vi == 0.f) continue;
            const FresnelRow gr = rows_g[row];
            const double a = (gr.dist * fresnel_k(gr, N, q, w)) / rowsum[row];
            if (!(a < FZ_TOL)) acc += (double)vi * (double)(float)a;
        }
        g[q] = acc;
    }
}

// per row: sum a^2, sum |a|, sum a^2 colw[X] over the kept entries (FP32 values)
__global__ __launch_bounds__(FZ_BLOCK) void k_fresnel_mf_rowstats(const FresnelRow* __restrict__ rows_g, int64_t rows,
                                                                  int N, double w, const double* __restrict__ rowsum,
                                                                  const double* __restrict__ colw,
                                                                  double* __restrict__ sq, double* __restrict__ ab,
                                                                  double* __restrict__ sqw) {
    __shared__ double sh[FZ_BLOCK / 64];
    const int64_t n2 = (int64_t)N * N;
    for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const FresnelRow g = rows_g[row];
        const double sum = rowsum[row];
        double s2 = 0, s1 = 0, s3 = 0;
        for (int64_t q = threadIdx.x; q < n2; q += FZ_BLOCK) {
            const double v = (g.dist * fresnel_k(g, N, q, w)) / sum;
            if (v < FZ_TOL) continue;
            const double a = (double)(float)v;
            s2 += a * a;
            s1 += fabs(a);
            if (colw) s3 += (a * a) * colw[q];
        }
        s2 = block_sum(s2, sh);
        s1 = block_sum(s1, sh);
        s3 = block_sum(s3, sh);
        if (threadIdx.x == 0) {
            if (sq) sq[row] = s2;
            if (ab) ab[row] = s1;
            if (sqw) sqw[row] = s3;
        }
    }
}

// per pixel: sum |a| and the number of kept entries
__global__ __launch_bounds__(FZ_BLOCK) void k_fresnel_mf_colstats(const FresnelRow* __restrict__ rows_g, int64_t rows,
                                                                  int N, double w, const double* __restrict__ rowsum,
                                                                  double* __restrict__ colabs, int* __restrict__ colcnt) {
    const int64_t n2 = (int64_t)N * N;
    for (int64_t q = (int64_t)blockIdx.x * FZ_BLOCK + threadIdx.x; q < n2; q += (int64_t)gridDim.x * FZ_BLOCK) {
        double s1 = 0;
        int c = 0;
        for (int64_t row = 0; row < rows; ++row) {
            const FresnelRow gr = rows_g[row];
            const double v = (gr.dist * fresnel_k(gr, N, q, w)) / rowsum[row];
            if (v < FZ_TOL) continue;
            const double a = (double)(float)v;
            s1 += fabs(a);
            c += a != 0.0;
        }
        if (colabs) colabs[q] = s1;
        if (colcnt) colcnt[q] = c;
    }
}

// the rows' geometry: source (N/2, y0_i), receiver (xp_j, yp_j), |S - R|
std::vector<FresnelRow> fresnel_rows(int N, int s, int p) {
    std::vector<double> y0, xp, yp;
    seismic_positions(N, s, p, y0, xp, yp);
    const double N2 = (double)N / 2.0;
    std::vector<FresnelRow> g((size_t)std::max<int64_t>((int64_t)s * p, 1));
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
    return g;
}

unsigned rows_grid(int64_t m) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(m, 1 << 16)); }
unsigned pix_grid(int64_t n2) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n2 + FZ_BLOCK - 1) / FZ_BLOCK, 1 << 16)); }

}  // namespace

void fresnel_mf_prepare(Geometry& G, int N, int s, int p, double omega, hipStream_t st) {
    AIR_CHECK(N > 0 && s > 0 && p > 0, AIR_E_ARG, "invalid seismicwavetomo arguments");
    AIR_CHECK((int64_t)N * N < (int64_t)INT32_MAX, AIR_E_ARG, "N too large for int32 column indices");
    const int64_t m = (int64_t)s * p;
    const std::vector<FresnelRow> g = fresnel_rows(N, s, p);
    G.model = 1;
    G.N = N;
    G.nrays = m;
    G.fw = omega / N;                                   // simple_fat_kernel(R,S,omega/N,...), :209
    G.frows.alloc(sizeof(FresnelRow) * g.size());
    AIR_HIP(hipMemcpyAsync(G.frows.p, g.data(), sizeof(FresnelRow) * g.size(), hipMemcpyHostToDevice, st));
    G.frowsum.alloc(8 * (size_t)std::max<int64_t>(m, 1));
    DevBuf len;
    len.alloc(8 * (size_t)(m + 1));
    if (m > 0) {
        k_fresnel_count<<<rows_grid(m), FZ_BLOCK, 0, st>>>(G.frows.as<FresnelRow>(), m, N, G.fw,
                                                           G.frowsum.as<double>(), len.as<int64_t>());
        AIR_HIP(hipGetLastError());
    }
    AIR_HIP(hipStreamSynchronize(st));
}

void fresnel_mf_forward(const Geometry& G, const float* u, float* y, hipStream_t st) {
    if (G.nrays == 0) return;
    k_fresnel_mf_fwd<<<rows_grid(G.nrays), FZ_BLOCK, 0, st>>>(G.frows.as<FresnelRow>(), G.nrays, G.N, G.fw,
                                                              G.frowsum.as<double>(), u, y);
    AIR_HIP(hipGetLastError());
}

void fresnel_mf_adjoint(const Geometry& G, const float* v, double* g, hipStream_t st) {
    const int64_t n2 = (int64_t)G.N * G.N;
    k_fresnel_mf_adj<<<pix_grid(n2), FZ_BLOCK, 0, st>>>(G.frows.as<FresnelRow>(), G.nrays, G.N, G.fw,
                                                        G.frowsum.as<double>(), v, g);
    AIR_HIP(hipGetLastError());
}

void fresnel_mf_row_stats(const Geometry& G, const double* colw, double* sq, double* ab, double* sqw, hipStream_t st) {
    if (G.nrays == 0) return;
    k_fresnel_mf_rowstats<<<rows_grid(G.nrays), FZ_BLOCK, 0, st>>>(G.frows.as<FresnelRow>(), G.nrays, G.N, G.fw,
                                                                   G.frowsum.as<double>(), colw, sq, ab, sqw);
    AIR_HIP(hipGetLastError());
}

void fresnel_mf_col_stats(const Geometry& G, double* colabs, int* colcnt, hipStream_t st) {
    const int64_t n2 = (int64_t)G.N * G.N;
    k_fresnel_mf_colstats<<<pix_grid(n2), FZ_BLOCK, 0, st>>>(G.frows.as<FresnelRow>(), G.nrays, G.N, G.fw,
                                                             G.frowsum.as<double>(), colabs, colcnt);
    AIR_HIP(hipGetLastError());
}

void fresnel_to_csr(int N, int s, int p, double omega, int dtype, DevSparse& A, hipStream_t st) {
    AIR_CHECK(N > 0 && s > 0 && p > 0, AIR_E_ARG, "invalid seismicwavetomo arguments");
    const int64_t n = (int64_t)N * N;
    AIR_CHECK(n < (int64_t)INT32_MAX, AIR_E_ARG, "N too large for int32 column indices");
    const int64_t m = (int64_t)s * p;
    A.nrows = m;
    A.ncols = n;
    A.dtype = dtype;
    const std::vector<FresnelRow> g = fresnel_rows(N, s, p);
    const double w = omega / N;                          // simple_fat_kernel(R,S,omega/N,...), :209
    DevBuf dg, rowsum, len;
    dg.alloc(sizeof(FresnelRow) * g.size());
    AIR_HIP(hipMemcpyAsync(dg.p, g.data(), sizeof(FresnelRow) * g.size(), hipMemcpyHostToDevice, st));
    rowsum.alloc(8 * (size_t)std::max<int64_t>(m, 1));
    len.alloc(8 * (size_t)(m + 1));
    AIR_HIP(hipMemsetAsync(len.p, 0, 8 * (size_t)(m + 1), st));
    const unsigned grid = rows_grid(m);
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
