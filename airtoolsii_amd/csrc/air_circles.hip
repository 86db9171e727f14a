// air_circles.hip -- sphericaltomo: circle-quadrature system matrix on gfx950.
//
// Reference algorithm: testprobs/sphericaltomo.m:108-270.  Row (m, n) samples
// circle n (radius r_n = 2n/nC) about centre xi_m = (cosd a_m, sind a_m) at
// nPhi_n + 1 angles k_t = t dPhi_n, t = 0..nPhi_n (both ends of the period,
// :203), rounds every sample to a pixel (MATLAB round: half away from zero),
// drops samples outside the image, and stores (2 pi r_n / nPhi_n) times the
// number of samples per pixel (unique + histc, :227-234).
//
// GPU form: one lane per sample writes the sample's 0-based column (or a
// sentinel past the last column when it is outside the image) into a key
// array laid out row by row; a segmented radix sort orders every row's keys;
// one wave per row then turns each run of equal keys into one CSR entry,
// counting runs across 64-key rounds with ballots.  cos(k_t) and sin(k_t)
// depend on the circle only, so they are tabulated once on the host with the
// C library and the per-sample arithmetic is FP64 in the reference's order
// (this file is compiled with -ffp-contract=off): every entry is
// bit-identical to the CPU restatement.
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

#include "air_internal.h"

namespace air {

namespace {

// linspace(0,2,nC+1)(2:end) as MATLAB evaluates it: d1 + (k*(d2-d1))/(n-1), last point exact.
std::vector<double> circle_radii(int nC) {
    std::vector<double> r(nC);
    const double n1 = (double)nC;
    for (int k = 1; k <= nC; ++k) r[k - 1] = 0.0 + ((double)k * (2.0 - 0.0)) / n1;
    if (nC >= 1) r[nC - 1] = 2.0;
    return r;
}

// MATLAB cosd/sind at the circle centres: exact at multiples of 90 degrees.
void cosd_sind(double deg, double* c, double* s) {
    if (!std::isfinite(deg)) { *c = NAN; *s = NAN; return; }
    double r = std::fmod(deg, 360.0);
    double q = std::nearbyint(r / 90.0);
    double z = r - 90.0 * q;
    int iq = ((int)q % 4 + 4) % 4;
    double zr = z * (M_PI / 180.0);
    double sz = std::sin(zr), cz = std::cos(zr);
    if (z == 0.0) { sz = 0.0; cz = 1.0; }
    double ss, cc;
    switch (iq) {
    case 0: ss = sz; cc = cz; break;
    case 1: ss = cz; cc = -sz; break;
    case 2: ss = -sz; cc = -cz; break;
    default: ss = -cz; cc = sz; break;
    }
    if (ss == 0.0) ss = 0.0;
    if (cc == 0.0) cc = 0.0;
    *s = ss;
    *c = cc;
}

// Column (0-based) of sample t of a circle of radius r about centre (xix, xiy),
// or -1 outside the image: sphericaltomo.m:207-222 in the same FP64 order.
__device__ __forceinline__ int64_t circle_sample_col(double xix, double xiy, double r, double c, double sn, int N,
                                                     double dx, double cimg) {
    const double xx = (xix + r * c) / dx + cimg;
    const double yy = (xiy + r * sn) / dx + cimg;
    const double col = round(xx), row = round(yy);
    const bool in = col > 0.0 && col <= (double)N && row > 0.0 && row <= (double)N;
    return in ? (int64_t)(((double)N - row + 1.0) + (col - 1.0) * (double)N - 1.0) : -1;
}

// circle of sample t of an angle's range: last n with cum[n] <= t
__device__ __forceinline__ int circle_of(const int64_t* __restrict__ cum, int nC, int64_t t) {
    int lo = 0, hi = nC - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cum[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// key of every sample of angles [m0, m0 + mcount): S samples per angle, circle
// n occupying [cum[n], cum[n+1]) of an angle's range.
__global__ __launch_bounds__(256) void k_circle_keys(const double* __restrict__ cosk, const double* __restrict__ sink,
                                                     const int64_t* __restrict__ cum, const double* __restrict__ radii,
                                                     int nC, const double* __restrict__ xi, int m0, int64_t total,
                                                     int64_t S, int N, double dx, double cimg, uint32_t none,
                                                     uint32_t* __restrict__ keys) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t mm = q / S;
        const int64_t t = q - mm * S;
        const int n = circle_of(cum, nC, t);
        const int64_t col = circle_sample_col(xi[2 * (m0 + mm)], xi[2 * (m0 + mm) + 1], radii[n], cosk[t], sink[t], N,
                                              dx, cimg);
        keys[q] = col >= 0 ? (uint32_t)col : none;
    }
}

// Offsets of the rows of one chunk of angles (row r = mm*nC + n), relative to
// the chunk's first key; row r ends where row r + 1 starts.
__global__ void k_circle_offsets(const int64_t* __restrict__ cum, int nC, int64_t rows, int64_t S,
                                 int* __restrict__ off) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= rows; r += (int64_t)gridDim.x * blockDim.x)
        off[r] = (int)((r / nC) * S + cum[r % nC]);
}

__device__ __forceinline__ int popc64(uint64_t v) { return __popcll(v); }

// Runs of equal keys per row (the row's CSR length).
__global__ __launch_bounds__(256) void k_circle_count(const uint32_t* __restrict__ keys, const int64_t* __restrict__ cum,
                                                      int nC, int64_t rows, int64_t S, uint32_t none,
                                                      int64_t* __restrict__ len) {
    const int lane = threadIdx.x & 63;
    for (int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); row < rows; row += (int64_t)gridDim.x * 4) {
        const int n = (int)(row % nC);
        const int64_t b = (row / nC) * S + cum[n], e = (row / nC) * S + cum[n + 1];
        int64_t runs = 0;
        for (int64_t p0 = b; p0 < e; p0 += 64) {
            const int64_t pos = p0 + lane;
            const uint32_t key = pos < e ? keys[pos] : none;
            const uint32_t prev = (pos < e && pos > b) ? keys[pos - 1] : none;
            const bool st = key != none && (pos == b || key != prev);
            runs += popc64(__ballot(st));
        }
        if (lane == 0) len[row] = runs;
    }
}

// CSR entries of each row: column = key, value = base_n * run length.  The
// run that is still open at the end of a 64-key round is finished by the first
// start of a later round or, at the end, by the end of the valid keys.
template <typename V>
__global__ __launch_bounds__(256) void k_circle_fill(const uint32_t* __restrict__ keys, const int64_t* __restrict__ cum,
                                                     const double* __restrict__ base, int nC, int64_t rows, int64_t S,
                                                     uint32_t none, const int64_t* __restrict__ rowptr,
                                                     int* __restrict__ col, V* __restrict__ val) {
    const int lane = threadIdx.x & 63;
    for (int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); row < rows; row += (int64_t)gridDim.x * 4) {
        const int n = (int)(row % nC);
        const int64_t b = (row / nC) * S + cum[n], e = (row / nC) * S + cum[n + 1];
        const double w = base[n];
        const int64_t out = rowptr[row];
        int64_t nrun = 0, pend_idx = -1, pend_pos = 0, vend = b;
        for (int64_t p0 = b; p0 < e; p0 += 64) {
            const int64_t pos = p0 + lane;
            const uint32_t key = pos < e ? keys[pos] : none;
            const uint32_t prev = (pos < e && pos > b) ? keys[pos - 1] : none;
            const bool valid = key != none;
            const bool st = valid && (pos == b || key != prev);
            const uint64_t vmask = __ballot(valid);
            const uint64_t smask = __ballot(st);
            if (vmask) vend = p0 + 64 - __clzll((long long)vmask);
            if (!smask) continue;
            const int first = __ffsll((unsigned long long)smask) - 1;
            if (pend_idx >= 0 && lane == 0) val[out + pend_idx] = from_f64<V>(w * (double)(p0 + first - pend_pos));
            if (st) {
                const int rank = popc64(smask & ((1ull << lane) - 1ull));
                const int64_t idx = out + nrun + rank;
                col[idx] = (int)key;
                const uint64_t later = lane < 63 ? (smask >> (lane + 1)) : 0ull;
                if (later) val[idx] = from_f64<V>(w * (double)(__ffsll((unsigned long long)later)));
            }
            const int last = 63 - __clzll((long long)smask);
            const int cnt = popc64(smask);
            pend_idx = nrun + cnt - 1;
            pend_pos = p0 + last;
            nrun += cnt;
        }
        if (pend_idx >= 0 && lane == 0) val[out + pend_idx] = from_f64<V>(w * (double)(vend - pend_pos));
    }
}

unsigned blocks_for(int64_t n, int64_t per_block) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + per_block - 1) / per_block, 1 << 20));
}

// Host tables of sphericaltomo.m:114-124: radii, samples per circle (both
// period ends), quadrature weight 2 pi r / nPhi, the samples' cos/sin with
// the C library, and the circle centres (cosd, sind of the angles).
struct CircleTabs {
    std::vector<double> radii, base, ck, sk, xi;
    std::vector<int64_t> cum;
    int64_t S = 0;
    double dx = 0, cimg = 0;
};

CircleTabs circle_tables(int N, const double* angles, int nA, int nC) {
    CircleTabs T;
    T.radii = circle_radii(nC);
    T.cimg = std::ceil(N / 2.0);
    T.dx = std::sqrt(2.0) / N;
    T.cum.assign(nC + 1, 0);
    T.base.assign(nC, 0.0);
    for (int c = 0; c < nC; ++c) {
        const double nphi = std::ceil((4 * M_PI / T.dx) * T.radii[c]);
        AIR_CHECK(nphi >= 1 && nphi < 1e9, AIR_E_ARG, "sphericaltomo: circle sample count out of range");
        T.cum[c + 1] = T.cum[c] + (int64_t)nphi + 1;
        T.base[c] = 2 * M_PI * T.radii[c] / nphi;
    }
    T.S = T.cum[nC];
    AIR_CHECK(T.S < (int64_t)1 << 30, AIR_E_UNSUPPORTED, "sphericaltomo: too many samples per angle");
    T.ck.assign((size_t)T.S, 0.0);
    T.sk.assign((size_t)T.S, 0.0);
    for (int c = 0; c < nC; ++c) {
        const int64_t nphi = T.cum[c + 1] - T.cum[c] - 1;
        const double dphi = 2 * M_PI / (double)nphi;
        for (int64_t t = 0; t <= nphi; ++t) {
            const double k = (double)t * dphi;
            T.ck[T.cum[c] + t] = std::cos(k);
            T.sk[T.cum[c] + t] = std::sin(k);
        }
    }
    T.xi.assign(2 * (size_t)std::max(nA, 1), 0.0);
    for (int a = 0; a < nA; ++a) cosd_sind(angles[a], &T.xi[2 * a], &T.xi[2 * a + 1]);
    return T;
}

}  // namespace

void circles_to_csr(int N, const double* angles, int nA, int nC, int dtype, DevSparse& A, hipStream_t s) {
    AIR_CHECK(N > 0 && nA >= 0 && nC > 0, AIR_E_ARG, "invalid sphericaltomo arguments");
    const int64_t n = (int64_t)N * N;
    AIR_CHECK(n < (int64_t)INT32_MAX, AIR_E_ARG, "N too large for int32 column indices");
    const int64_t m = (int64_t)nA * nC;
    A.nrows = m;
    A.ncols = n;
    A.dtype = dtype;
    const CircleTabs CT = circle_tables(N, angles, nA, nC);
    const std::vector<double>& radii = CT.radii;
    const std::vector<int64_t>& cum = CT.cum;
    const std::vector<double>& base = CT.base;
    const std::vector<double>& ck = CT.ck;
    const std::vector<double>& sk = CT.sk;
    const std::vector<double>& xi = CT.xi;
    const int64_t S = CT.S;
    const double dx = CT.dx, cimg = CT.cimg;

    DevBuf dck, dsk, dcum, drad, dbase, dxi;
    dck.alloc(8 * S); dsk.alloc(8 * S); dcum.alloc(8 * (nC + 1)); drad.alloc(8 * nC); dbase.alloc(8 * nC);
    dxi.alloc(8 * xi.size());
    AIR_HIP(hipMemcpyAsync(dck.p, ck.data(), 8 * S, hipMemcpyHostToDevice, s));
    AIR_HIP(hipMemcpyAsync(dsk.p, sk.data(), 8 * S, hipMemcpyHostToDevice, s));
    AIR_HIP(hipMemcpyAsync(dcum.p, cum.data(), 8 * (nC + 1), hipMemcpyHostToDevice, s));
    AIR_HIP(hipMemcpyAsync(drad.p, radii.data(), 8 * nC, hipMemcpyHostToDevice, s));
    AIR_HIP(hipMemcpyAsync(dbase.p, base.data(), 8 * nC, hipMemcpyHostToDevice, s));
    AIR_HIP(hipMemcpyAsync(dxi.p, xi.data(), 8 * xi.size(), hipMemcpyHostToDevice, s));

    // keys of all rows, sorted chunk by chunk of angles (int32 item counts)
    const uint32_t none = (uint32_t)n;                 // past the last column: sorts last
    int end_bit = 1;
    while (end_bit < 32 && ((uint64_t)1 << end_bit) <= (uint64_t)none) ++end_bit;
    const int64_t total = (int64_t)nA * S;
    const int mchunk = (int)std::max<int64_t>(1, std::min<int64_t>(nA, ((int64_t)1 << 30) / S));
    DevBuf keys_in, keys, off, temp;
    keys.alloc(4 * (size_t)std::max<int64_t>(total, 1));
    keys_in.alloc(4 * (size_t)std::max<int64_t>((int64_t)mchunk * S, 1));
    off.alloc(4 * ((size_t)mchunk * nC + 1));
    k_circle_offsets<<<blocks_for((int64_t)mchunk * nC + 1, 256), 256, 0, s>>>(dcum.as<int64_t>(), nC,
                                                                               (int64_t)mchunk * nC, S, off.as<int>());
    AIR_HIP(hipGetLastError());
    size_t temp_bytes = 0;
    AIR_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, temp_bytes, keys_in.as<uint32_t>(),
                                                       keys.as<uint32_t>(), (int)((int64_t)mchunk * S),
                                                       mchunk * nC, off.as<int>(), off.as<int>() + 1, 0, end_bit, s));
    temp.alloc(std::max<size_t>(temp_bytes, 256));
    for (int m0 = 0; m0 < nA; m0 += mchunk) {
        const int mc = std::min(mchunk, nA - m0);
        const int64_t items = (int64_t)mc * S;
        k_circle_keys<<<blocks_for(items, 256), 256, 0, s>>>(dck.as<double>(), dsk.as<double>(), dcum.as<int64_t>(),
                                                             drad.as<double>(), nC, dxi.as<double>(), m0, items, S, N,
                                                             dx, cimg, none, keys_in.as<uint32_t>());
        AIR_HIP(hipGetLastError());
        size_t tb = temp_bytes;
        AIR_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(temp.p, tb, keys_in.as<uint32_t>(),
                                                           keys.as<uint32_t>() + (int64_t)m0 * S, (int)items,
                                                           mc * nC, off.as<int>(), off.as<int>() + 1, 0, end_bit, s));
    }
    keys_in.release();
    temp.release();

    DevBuf len;
    len.alloc(8 * (size_t)(m + 1));
    AIR_HIP(hipMemsetAsync(len.p, 0, 8 * (size_t)(m + 1), s));
    if (m > 0) {
        k_circle_count<<<blocks_for(m, 4), 256, 0, s>>>(keys.as<uint32_t>(), dcum.as<int64_t>(), nC, m, S, none,
                                                       len.as<int64_t>());
        AIR_HIP(hipGetLastError());
    }
    A.ptr.alloc(8 * (size_t)(m + 1));
    scan_exclusive_i64(len.as<int64_t>(), A.ptr.as<int64_t>(), m, s);
    int64_t nnz = 0;
    AIR_HIP(hipMemcpyAsync(&nnz, A.ptr.as<int64_t>() + m, 8, hipMemcpyDeviceToHost, s));
    AIR_HIP(hipStreamSynchronize(s));
    A.nnz = nnz;
    alloc_sparse_arrays(A, nnz, dtype, s);
    if (m > 0) {
        dispatch_dtype(dtype, [&](auto v) {
            using V = decltype(v);
            k_circle_fill<V><<<blocks_for(m, 4), 256, 0, s>>>(keys.as<uint32_t>(), dcum.as<int64_t>(),
                                                              dbase.as<double>(), nC, m, S, none, A.ptr.as<int64_t>(),
                                                              A.idx.as<int>(), A.val.as<V>());
            return 0;
        });
        AIR_HIP(hipGetLastError());
    }
    AIR_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------- matrix-free form
// sphericaltomo.m:90-93,243-254: the function handle re-forms every row's
// binned samples per product (A(i) = aa' * u(jj); A(jj) += u(i) * aa, with
// aa = (2 pi r / nPhi) * counts).  Here the products run sample by sample in
// FP64 with the tracer's per-sample arithmetic: sum_j (c w_j) u_j = c sum_t
// u_{j(t)}, so they equal the reference's FP64 products up to FP64 rounding.
// The SIRT weights need the binned counts (sums of squares); they are taken
// from a temporary CSR built by circles_to_csr, so they equal the stored
// FP32 matrix's.

namespace {

// y[row] = c_n * sum over the row's samples of u[col]: one wave per row
__global__ __launch_bounds__(256) void k_circle_mf_fwd(const double* __restrict__ cosk, const double* __restrict__ sink,
                                                       const int64_t* __restrict__ cum,
                                                       const double* __restrict__ radii,
                                                       const double* __restrict__ base, int nC,
                                                       const double* __restrict__ xi, int64_t rows, int N, double dx,
                                                       double cimg, const float* __restrict__ u,
                                                       float* __restrict__ y) {
    const int lane = threadIdx.x & 63;
    for (int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); row < rows; row += (int64_t)gridDim.x * 4) {
        const int64_t mm = row / nC;
        const int n = (int)(row - mm * nC);
        const double xix = xi[2 * mm], xiy = xi[2 * mm + 1], r = radii[n];
        double acc = 0.0;
        for (int64_t t = cum[n] + lane; t < cum[n + 1]; t += 64) {
            const int64_t col = circle_sample_col(xix, xiy, r, cosk[t], sink[t], N, dx, cimg);
            if (col >= 0) acc += (double)u[col];
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (lane == 0) y[row] = (float)(base[n] * acc);
    }
}

// g[col] += c_n v[row] for every sample (FP64 atomics; g cleared first)
__global__ __launch_bounds__(256) void k_circle_mf_adj(const double* __restrict__ cosk, const double* __restrict__ sink,
                                                       const int64_t* __restrict__ cum,
                                                       const double* __restrict__ radii,
                                                       const double* __restrict__ base, int nC,
                                                       const double* __restrict__ xi, int64_t total, int64_t S, int N,
                                                       double dx, double cimg, const float* __restrict__ v,
                                                       double* __restrict__ g) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t mm = q / S;
        const int64_t t = q - mm * S;
        const int n = circle_of(cum, nC, t);
        const float vi = v[mm * nC + n];
        if (vi == 0.f) continue;
        const int64_t col = circle_sample_col(xi[2 * mm], xi[2 * mm + 1], radii[n], cosk[t], sink[t], N, dx, cimg);
        if (col >= 0) atomicAdd(&g[col], base[n] * (double)vi);
    }
}

__global__ void k_count_to_int(const double* __restrict__ in, int64_t n, int* __restrict__ out) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
        out[j] = (int)in[j];
}

}  // namespace

void circles_mf_prepare(Geometry& g, int N, const double* angles, int nA, int nC, hipStream_t s) {
    AIR_CHECK(N > 0 && nA >= 0 && nC > 0, AIR_E_ARG, "invalid sphericaltomo arguments");
    AIR_CHECK((int64_t)N * N < (int64_t)INT32_MAX, AIR_E_ARG, "N too large for int32 column indices");
    const CircleTabs CT = circle_tables(N, angles, nA, nC);
    g = Geometry();
    g.model = 2;
    g.N = N;
    g.nrays = (int64_t)nA * nC;
    g.cnA = nA;
    g.cnC = nC;
    g.cS = CT.S;
    g.cdx = CT.dx;
    g.cimg = CT.cimg;
    g.cangles.assign(angles, angles + nA);
    g.ccos.alloc(8 * CT.S);
    g.csin.alloc(8 * CT.S);
    g.ccum.alloc(8 * (nC + 1));
    g.crad.alloc(8 * nC);
    g.cbase.alloc(8 * nC);
    g.cxi.alloc(8 * CT.xi.size());
    AIR_HIP(hipMemcpyAsync(g.ccos.p, CT.ck.data(), 8 * CT.S, hipMemcpyHostToDevice, s));
    AIR_HIP(hipMemcpyAsync(g.csin.p, CT.sk.data(), 8 * CT.S, hipMemcpyHostToDevice, s));
    AIR_HIP(hipMemcpyAsync(g.ccum.p, CT.cum.data(), 8 * (nC + 1), hipMemcpyHostToDevice, s));
    AIR_HIP(hipMemcpyAsync(g.crad.p, CT.radii.data(), 8 * nC, hipMemcpyHostToDevice, s));
    AIR_HIP(hipMemcpyAsync(g.cbase.p, CT.base.data(), 8 * nC, hipMemcpyHostToDevice, s));
    AIR_HIP(hipMemcpyAsync(g.cxi.p, CT.xi.data(), 8 * CT.xi.size(), hipMemcpyHostToDevice, s));
    AIR_HIP(hipStreamSynchronize(s));
}

void circles_mf_forward(const Geometry& g, const float* u, float* y, hipStream_t s) {
    if (g.nrays == 0) return;
    k_circle_mf_fwd<<<blocks_for(g.nrays, 4), 256, 0, s>>>(g.ccos.as<double>(), g.csin.as<double>(),
                                                          g.ccum.as<int64_t>(), g.crad.as<double>(),
                                                          g.cbase.as<double>(), g.cnC, g.cxi.as<double>(), g.nrays,
                                                          g.N, g.cdx, g.cimg, u, y);
    AIR_HIP(hipGetLastError());
}

void circles_mf_adjoint(const Geometry& g, const float* v, double* gacc, hipStream_t s) {
    const int64_t n = (int64_t)g.N * g.N;
    zero_f64(gacc, n, s);     // a fill kernel: the product may be captured into a HIP graph
    const int64_t total = (int64_t)g.cnA * g.cS;
    if (total == 0) return;
    k_circle_mf_adj<<<blocks_for(total, 256), 256, 0, s>>>(g.ccos.as<double>(), g.csin.as<double>(),
                                                          g.ccum.as<int64_t>(), g.crad.as<double>(),
                                                          g.cbase.as<double>(), g.cnC, g.cxi.as<double>(), total,
                                                          g.cS, g.N, g.cdx, g.cimg, v, gacc);
    AIR_HIP(hipGetLastError());
}

void circles_mf_row_stats(const Geometry& g, const double* colw, double* sq, double* ab, double* sqw, hipStream_t s) {
    DevSparse A;
    circles_to_csr(g.N, g.cangles.data(), g.cnA, g.cnC, AIR_F32, A, s);
    row_stats(A, colw, sq, ab, sqw, nullptr, s);
    AIR_HIP(hipStreamSynchronize(s));   // A is freed on return
}

void circles_mf_col_stats(const Geometry& g, double* colabs, int* colcnt, hipStream_t s) {
    const int64_t n = (int64_t)g.N * g.N;
    DevSparse A, AT;
    circles_to_csr(g.N, g.cangles.data(), g.cnA, g.cnC, AIR_F32, A, s);
    csr_to_csc(A, AT, s);
    A = DevSparse();
    DevBuf nz;
    nz.alloc(8 * (size_t)std::max<int64_t>(n, 1));
    row_stats(AT, nullptr, nullptr, colabs, nullptr, nz.as<double>(), s);
    if (colcnt && n > 0) {
        k_count_to_int<<<blocks_for(n, 256), 256, 0, s>>>(nz.as<double>(), n, colcnt);
        AIR_HIP(hipGetLastError());
    }
    AIR_HIP(hipStreamSynchronize(s));   // AT and nz are freed on return
}

}  // namespace air
