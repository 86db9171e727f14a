// air_sirt.hip -- K2/K3/K4: the SIRT iteration on compressed rows, for gfx950.
//
// One SIRT step of sirt/sirt.m:213-243 is two bandwidth-bound passes over A:
//   update (over the CSC of A, i.e. rows of A'):
//       g_j = sum_i a_ij (M r)_i ;  x_j = P_C(x_j + lam * D_j g_j)        (:214-237)
//   resid  (over the CSR of A):
//       r_i = b_i - sum_j a_ij x_j ;  (M r)_i = M_i r_i ; FP64 ||r||^2 ... (:240-243)
// Each compressed row is handled by a group of G lanes of a 64-lane wave
// (G = 64 for the long rows of tomography matrices): values and indices are
// read with coalesced non-temporal loads so the streamed matrix does not evict
// the gathered vector (x or M r, ~1 MB at N=512) from L2.  The grid is
// persistent (a few blocks per CU) and strides over rows, which keeps the
// number of FP64 partial sums small; the last block to finish reduces them in
// a fixed order (deterministic) and evaluates the stop rule of
// auxil/check_stoprules.m on the device, so iterations are enqueued without a
// host round trip.  Every kernel returns at once after the device stop flag
// is set, which reproduces the reference's early exit of the while loop.
#include <cmath>

#include "air_device.h"

namespace air {

// ---------------------------------------------------------------- y = S x
template <typename V, int G>
__global__ __launch_bounds__(BLOCK) void k_spmv(const int64_t* __restrict__ ptr, const int* __restrict__ idx,
                                                const V* __restrict__ val, int64_t nrows,
                                                const float* __restrict__ x, float* __restrict__ y) {
    const int gl = threadIdx.x % G;
    for_each_row<G>(ptr, nrows, [&](int64_t row, int64_t s, int64_t e) {
        auto acc = row_dot<V, G>(idx, val, x, s, e, gl);
        acc = group_sum<G>(acc);
        if (gl == 0) y[row] = (float)acc;
    });
}

// ---------------------------------------------------------------- FP64 row statistics
template <typename V>
__global__ __launch_bounds__(BLOCK) void k_row_stats(const int64_t* __restrict__ ptr, const int* __restrict__ idx,
                                                     const V* __restrict__ val, int64_t nrows,
                                                     const double* __restrict__ w, double* __restrict__ sq,
                                                     double* __restrict__ ab, double* __restrict__ sqw,
                                                     double* __restrict__ nz) {
    const int gl = threadIdx.x & 63;
    int64_t row = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) >> 6;
    if (row >= nrows) return;
    double s2 = 0, s1 = 0, s3 = 0, c = 0;
    for (int64_t q = ptr[row] + gl; q < ptr[row + 1]; q += 64) {
        double a = to_f64(val[q]);
        s2 += a * a;
        s1 += fabs(a);
        if (w) s3 += (a * a) * w[idx[q]];
        c += (a != 0.0) ? 1.0 : 0.0;
    }
    s2 = group_sum<64>(s2);
    s1 = group_sum<64>(s1);
    s3 = group_sum<64>(s3);
    c = group_sum<64>(c);
    if (gl == 0) {
        if (sq) sq[row] = s2;
        if (ab) ab[row] = s1;
        if (sqw) sqw[row] = s3;
        if (nz) nz[row] = c;
    }
}

// ---------------------------------------------------------------- K2: residual pass (CSR)
template <typename V, int G, bool ME>
__global__ __launch_bounds__(BLOCK) void k_resid(const int64_t* __restrict__ ptr, const int* __restrict__ idx,
                                                 const V* __restrict__ val, int64_t m, const float* __restrict__ x,
                                                 const float* __restrict__ b, const float* __restrict__ M,
                                                 float* __restrict__ r, float* __restrict__ Mr, double* partial,
                                                 double* hist, Ctrl* ctrl, StopParams sp, int k) {
    __shared__ double sh[4][BLOCK / 64];
    if (load_stop(ctrl)) return;
    const int gl = threadIdx.x % G;
    double v[4] = {0, 0, 0, 0};
    for_each_row<G>(ptr, m, [&](int64_t row, int64_t s, int64_t e) {
        auto acc = row_dot<V, G>(idx, val, x, s, e, gl);
        acc = group_sum<G>(acc);
        if (gl == 0) {
            float ri = b[row] - (float)acc;
            float mri = M ? M[row] * ri : ri;
            if (ME) {
                double ro = (double)r[row];
                v[2] += ro * (ro + (double)ri);
                v[3] += ro * ro;
            }
            r[row] = ri;
            Mr[row] = mri;
            v[0] += (double)ri * (double)ri;
            v[1] += (double)ri * (double)mri;
        }
    });
    block_reduce<4>(v, sh);
    store_block_partials<4>(v, partial);
}

// ---------------------------------------------------------------- K3+K4: update pass (CSC)
template <typename V, int G>
__global__ __launch_bounds__(BLOCK) void k_update(const int64_t* __restrict__ ptr, const int* __restrict__ idx,
                                                  const V* __restrict__ val, int64_t n, const float* __restrict__ Mr,
                                                  const float* __restrict__ D, float* __restrict__ x, double lam,
                                                  int lb_kind, float lb, const float* __restrict__ lbv, int ub_kind,
                                                  float ub, const float* __restrict__ ubv, Ctrl* ctrl) {
    if (load_stop(ctrl)) return;
    if (isnan(lam)) lam = ctrl->lam_cur;     // NaN: chosen on the device this iteration
    const int gl = threadIdx.x % G;
    for_each_row<G>(ptr, n, [&](int64_t j, int64_t s, int64_t e) {
        auto acc = row_dot<V, G>(idx, val, Mr, s, e, gl);
        acc = group_sum<G>(acc);
        if (gl == 0) {
            double dg = D ? (double)D[j] * (double)(float)acc : (double)(float)acc;
            double xn = (double)x[j] + lam * dg;
            x[j] = project(xn, (int)j, lb_kind, lb, lbv, ub_kind, ub, ubv);
        }
    });
    if (threadIdx.x == 0 && blockIdx.x == 0) ctrl->lam_cur = lam;
}

// ---------------------------------------------------------------- line search pieces (sirt.m:217-222)
// g = A' Mr (FP64 store) and sum_j D_j g_j^2; last block sets lam = r'Mr / sum.
template <typename V, int G>
__global__ __launch_bounds__(BLOCK) void k_gcols(const int64_t* __restrict__ ptr, const int* __restrict__ idx,
                                                 const V* __restrict__ val, int64_t n, const float* __restrict__ Mr,
                                                 const float* __restrict__ D, double* __restrict__ g, double* partial,
                                                 Ctrl* ctrl) {
    __shared__ double sh[1][BLOCK / 64];
    if (load_stop(ctrl)) return;
    const int gl = threadIdx.x % G;
    double v[1] = {0};
    for_each_row<G>(ptr, n, [&](int64_t j, int64_t s, int64_t e) {
        auto acc = row_dot<V, G>(idx, val, Mr, s, e, gl);
        acc = group_sum<G>(acc);
        if (gl == 0) {
            double gj = (double)(float)acc;
            g[j] = gj;
            v[0] += D ? (double)D[j] * (gj * gj) : gj * gj;
        }
    });
    block_reduce<1>(v, sh);
    store_block_partials<1>(v, partial);
}

// x = P_C(x + lam D g) with lam from the control block (line) or the argument.
__global__ __launch_bounds__(BLOCK) void k_apply(int64_t n, const double* __restrict__ g, const float* __restrict__ D,
                                                 float* __restrict__ x, double lam, int use_ctrl_lam, int lb_kind,
                                                 float lb, const float* __restrict__ lbv, int ub_kind, float ub,
                                                 const float* __restrict__ ubv, Ctrl* ctrl) {
    if (load_stop(ctrl)) return;
    double l = (use_ctrl_lam || isnan(lam)) ? ctrl->lam_cur : lam;
    for (int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x; j < n; j += (int64_t)gridDim.x * BLOCK) {
        double gj = (double)(float)g[j];
        double dg = D ? (double)D[j] * gj : gj;
        x[j] = project((double)x[j] + l * dg, (int)j, lb_kind, lb, lbv, ub_kind, ub, ubv);
    }
    if (!use_ctrl_lam && !isnan(lam) && threadIdx.x == 0 && blockIdx.x == 0) ctrl->lam_cur = lam;
}

// matrix-free line search: sum_j D_j g_j^2 over the FP64 accumulators
__global__ __launch_bounds__(BLOCK) void k_gsum(int64_t n, const double* __restrict__ g, const float* __restrict__ D,
                                                double* partial, Ctrl* ctrl) {
    __shared__ double sh[1][BLOCK / 64];
    if (load_stop(ctrl)) return;
    double v[1] = {0};
    for (int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x; j < n; j += (int64_t)gridDim.x * BLOCK) {
        double gj = (double)(float)g[j];
        v[0] += D ? (double)D[j] * (gj * gj) : gj * gj;
    }
    block_reduce<1>(v, sh);
    store_block_partials<1>(v, partial);
}

// residual from a precomputed A x (matrix-free forward projection)
template <bool ME>
__global__ __launch_bounds__(BLOCK) void k_resid_vec(int64_t m, const float* __restrict__ Ax,
                                                     const float* __restrict__ b, const float* __restrict__ M,
                                                     float* __restrict__ r, float* __restrict__ Mr, double* partial,
                                                     double* hist, Ctrl* ctrl, StopParams sp, int k) {
    __shared__ double sh[4][BLOCK / 64];
    if (load_stop(ctrl)) return;
    double v[4] = {0, 0, 0, 0};
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < m; i += (int64_t)gridDim.x * BLOCK) {
        float ri = b[i] - Ax[i];
        float mri = M ? M[i] * ri : ri;
        if (ME) {
            double ro = (double)r[i];
            v[2] += ro * (ro + (double)ri);
            v[3] += ro * ro;
        }
        r[i] = ri;
        Mr[i] = mri;
        v[0] += (double)ri * (double)ri;
        v[1] += (double)ri * (double)mri;
    }
    block_reduce<4>(v, sh);
    store_block_partials<4>(v, partial);
}

// Reduce the row pass's block partials in a fixed order and evaluate the stop
// rule of iteration k (auxil/check_stoprules.m) on the device.
// k < 0 (graph mode): the iteration number comes from ctrl->kcur, and the
// next iteration's lambda is staged in ctrl->lam_cur from lamv.
__global__ __launch_bounds__(BLOCK) void k_finalize_rows(const double* __restrict__ partial, int nblocks, double* hist,
                                                         Ctrl* ctrl, StopParams sp, int k, const double* lamv) {
    __shared__ double sh[4][BLOCK / 64];
    if (load_stop(ctrl)) return;
    if (k < 0) k = ctrl->kcur + 1;
    double v[4] = {0, 0, 0, 0};
    typedef double d4 __attribute__((ext_vector_type(4)));
    const d4* p4 = (const d4*)partial;
    int b = threadIdx.x;
    for (; b + 3 * BLOCK < nblocks; b += 4 * BLOCK) {
        d4 a0 = p4[b], a1 = p4[b + BLOCK], a2 = p4[b + 2 * BLOCK], a3 = p4[b + 3 * BLOCK];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] += ((a0[q] + a1[q]) + (a2[q] + a3[q]));
    }
    for (; b < nblocks; b += BLOCK) {
        d4 a0 = p4[b];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] += a0[q];
    }
    block_reduce<4>(v, sh);
    double c[1] = {0};
    if (sp.ncp_cols > 0) {
        __shared__ double shc[1][BLOCK / 64];
        for (int j = threadIdx.x; j < sp.ncp_cols; j += BLOCK) c[0] += sp.ncp_stat[j];
        block_reduce<1>(c, shc);
    }
    if (threadIdx.x == 0) {
        evaluate_stop(ctrl, sp, k, v[0], v[1], v[2], v[3], c[0], hist);
        ctrl->kcur = k;
        if (lamv && k < sp.kmax) ctrl->lam_cur = lamv[k];
    }
}

// NCP statistic of one residual column (check_stoprules.m:107-133): from the
// half spectrum S[0..len/2] of the column, pk = |S[f]|^2 for f = 1..q
// (q = floor(len/2)), c = cumsum(pk) / sum(pk), and the column's value is
// the 2-norm of c - (1:q)/q.  One block per column; each thread owns a
// contiguous run of frequencies, so the cumulative sum is a block scan.
__global__ __launch_bounds__(BLOCK) void k_ncp_stats(const double2* __restrict__ spec, int len,
                                                     double* __restrict__ colstat, const Ctrl* ctrl) {
    __shared__ double sc[BLOCK];
    __shared__ double sh[1][BLOCK / 64];
    if (load_stop(ctrl)) return;
    const int q = len / 2;
    const double2* S = spec + (int64_t)blockIdx.x * (len / 2 + 1);
    const int chunk = (q + BLOCK - 1) / BLOCK;
    const int f0 = 1 + (int)threadIdx.x * chunk, f1 = min(q + 1, f0 + chunk);
    double part = 0;
    for (int f = f0; f < f1; ++f) part += S[f].x * S[f].x + S[f].y * S[f].y;
    sc[threadIdx.x] = part;
    __syncthreads();
    for (int d = 1; d < BLOCK; d <<= 1) {       // inclusive scan of the thread sums
        const double add = threadIdx.x >= (unsigned)d ? sc[threadIdx.x - d] : 0.0;
        __syncthreads();
        sc[threadIdx.x] += add;
        __syncthreads();
    }
    const double total = sc[BLOCK - 1];
    double cum = sc[threadIdx.x] - part, ss[1] = {0};
    for (int f = f0; f < f1; ++f) {
        cum += S[f].x * S[f].x + S[f].y * S[f].y;
        const double dv = cum / total - (double)f / (double)q;
        ss[0] += dv * dv;
    }
    block_reduce<1>(ss, sh);
    if (threadIdx.x == 0) colstat[blockIdx.x] = sqrt(ss[0]);
}

__global__ void k_fill_f64(double* p, int n, double v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// line search: lam = r'Mr / sum_j D_j g_j^2 from the column pass's block partials
__global__ __launch_bounds__(BLOCK) void k_finalize_line(const double* __restrict__ partial, int nblocks, Ctrl* ctrl) {
    __shared__ double sh[1][BLOCK / 64];
    if (load_stop(ctrl)) return;
    double v[1] = {0};
    for (int b = threadIdx.x; b < nblocks; b += BLOCK) v[0] += partial[b];
    block_reduce<1>(v, sh);
    if (threadIdx.x == 0) ctrl->lam_cur = ctrl->rMr / v[0];
}

__global__ void k_zero_f64(double* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}
__global__ void k_f64_to_f32(const double* in, float* out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}
__global__ void k_f32_to_f64(const float* in, double* out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (double)in[i];
}
__global__ void k_mul_f32(const float* a, float* y, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = a[i] * y[i];
}

// ---------------------------------------------------------------- host launchers
static unsigned vec_grid(int64_t n) {
    int64_t b = (n + BLOCK - 1) / BLOCK;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 4096));
}

Shape default_shape(const DevSparse& S, int num_cu) {
    Shape sh;
    double L = S.avg_len();
    int ve = (int)(16 / dtype_size(S.dtype));          // values per 16-byte load
    double per_lane = L / (2.0 * ve);                   // two loads per lane and round
    sh.G = per_lane >= 48 ? 64 : (per_lane >= 24 ? 32 : (per_lane >= 12 ? 16 : 8));
    int64_t groups = (S.nrows + (BLOCK / sh.G) - 1) / (BLOCK / sh.G);
    sh.blocks = (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(groups, (int64_t)num_cu * 8), 2048));
    return sh;
}

template <typename F>
static void dispatch_G(int G, F&& f) {
    switch (G) {
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    case 64: f(std::integral_constant<int, 64>{}); break;
    default: throw Err{AIR_E_ARG, "lanes per row must be 8, 16, 32 or 64"};
    }
}

void spmv(const DevSparse& S, Shape sh, const float* x, float* y, hipStream_t s) {
    if (S.nrows == 0) return;
    dispatch_dtype(S.dtype, [&](auto v) {
        using V = decltype(v);
        dispatch_G(sh.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            k_spmv<V, G><<<sh.blocks, BLOCK, 0, s>>>(S.ptr.as<int64_t>(), S.idx.as<int>(), S.val.as<V>(), S.nrows,
                                                     x, y);
        });
        return 0;
    });
    AIR_HIP(hipGetLastError());
}

void row_stats(const DevSparse& S, const double* w, double* sq, double* ab, double* sqw, double* nz, hipStream_t s) {
    if (S.nrows == 0) return;
    dispatch_dtype(S.dtype, [&](auto v) {
        using V = decltype(v);
        unsigned grid = (unsigned)((S.nrows * 64 + BLOCK - 1) / BLOCK);
        k_row_stats<V><<<grid, BLOCK, 0, s>>>(S.ptr.as<int64_t>(), S.idx.as<int>(), S.val.as<V>(), S.nrows, w, sq, ab,
                                              sqw, nz);
        return 0;
    });
    AIR_HIP(hipGetLastError());
}

void launch_resid(const DevSparse& A, Shape sh, const SirtBufs& B, StopParams sp, int k, hipStream_t s) {
    dispatch_dtype(A.dtype, [&](auto v) {
        using V = decltype(v);
        dispatch_G(sh.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            if (sp.type == AIR_STOP_ME)
                k_resid<V, G, true><<<sh.blocks, BLOCK, 0, s>>>(A.ptr.as<int64_t>(), A.idx.as<int>(), A.val.as<V>(),
                                                                A.nrows, B.x, B.b, B.M, B.r, B.Mr, B.partial, B.hist,
                                                                B.ctrl, sp, k);
            else
                k_resid<V, G, false><<<sh.blocks, BLOCK, 0, s>>>(A.ptr.as<int64_t>(), A.idx.as<int>(), A.val.as<V>(),
                                                                 A.nrows, B.x, B.b, B.M, B.r, B.Mr, B.partial, B.hist,
                                                                 B.ctrl, sp, k);
        });
        return 0;
    });
    AIR_HIP(hipGetLastError());
    launch_finalize_rows(B, sh.blocks, sp, k, s);
}

void ncp_prepare(NcpState& N, int64_t m, int len, int ncols, int smooth, hipStream_t s) {
    if (!N.planned || N.len != len || N.ncols != ncols) {
        N.release();
        int n[1] = {len}, inembed[1] = {len}, onembed[1] = {len / 2 + 1};
        AIR_CHECK(hipfftPlanMany(&N.plan, 1, n, inembed, 1, len, onembed, 1, len / 2 + 1, HIPFFT_D2Z, ncols) ==
                      HIPFFT_SUCCESS, AIR_E_HIP, "hipfftPlanMany failed for the NCP residual transform");
        N.planned = true;
        N.len = len;
        N.ncols = ncols;
    }
    AIR_CHECK(hipfftSetStream(N.plan, s) == HIPFFT_SUCCESS, AIR_E_HIP, "hipfftSetStream failed");
    N.smooth = smooth;
    N.m = m;
    N.rd.alloc(8 * (size_t)std::max<int64_t>(m, 1));
    N.spec.alloc(16 * (size_t)ncols * (len / 2 + 1));
    N.colstat.alloc(8 * (size_t)ncols);
    N.dk.alloc(8 * (size_t)smooth);
    k_fill_f64<<<1, 64, 0, s>>>(N.dk.as<double>(), smooth, INFINITY);   // check_inputs.m:209
    AIR_HIP(hipGetLastError());
}

void ncp_enqueue(NcpState& N, const float* r, int64_t m, const Ctrl* ctrl, hipStream_t s) {
    (void)ctrl;
    k_f32_to_f64<<<(unsigned)std::min<int64_t>((m + BLOCK - 1) / BLOCK, 4096), BLOCK, 0, s>>>(r, N.rd.as<double>(), m);
    AIR_HIP(hipGetLastError());
    AIR_CHECK(hipfftExecD2Z(N.plan, N.rd.as<double>(), (hipfftDoubleComplex*)N.spec.p) == HIPFFT_SUCCESS, AIR_E_HIP,
              "hipfftExecD2Z failed");
    k_ncp_stats<<<(unsigned)N.ncols, BLOCK, 0, s>>>((const double2*)N.spec.p, N.len, N.colstat.as<double>(), ctrl);
    AIR_HIP(hipGetLastError());
}

void launch_finalize_rows(const SirtBufs& B, int nblocks, StopParams sp, int k, hipStream_t s) {
    if (sp.type == AIR_STOP_NCP && B.ncp) ncp_enqueue(*B.ncp, B.r, B.ncp->m, B.ctrl, s);
    k_finalize_rows<<<1, BLOCK, 0, s>>>(B.partial, nblocks, B.hist, B.ctrl, sp, k, B.lamv);
    AIR_HIP(hipGetLastError());
}

// greedy target: partial sums of <d, x* - x>, <d, d> and ||x - x*||^2 with d = D g (the step direction)
__global__ __launch_bounds__(BLOCK) void k_greedy_dots(int64_t n, const double* __restrict__ g,
                                                       const float* __restrict__ D, const float* __restrict__ x,
                                                       const float* __restrict__ xs, double* __restrict__ part,
                                                       const Ctrl* ctrl) {
    __shared__ double sh[3][BLOCK / 64];
    if (load_stop(ctrl)) return;
    double v[3] = {0, 0, 0};
    for (int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x; j < n; j += (int64_t)gridDim.x * BLOCK) {
        const double gj = g ? (double)(float)g[j] : 0.0;
        const double d = D ? (double)D[j] * gj : gj;
        const double e = (double)xs[j] - (double)x[j];
        v[0] += d * e;
        v[1] += d * d;
        v[2] += e * e;
    }
    block_reduce<3>(v, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) part[(int64_t)blockIdx.x * 3 + q] = v[q];
}

// Relaxation of iteration k (one wave): features of the state x_{k-1}, then
// lambda_k from the greedy target or the network, and the network's stop.
// The features use only residual statistics, so a learned lambda_k is known
// before the column pass and the update stays fused into it:
//  0 log(|r|/|r0|)   1 log(|r_prev|/|r0|)   2 log(|r|/|r_prev|)   3 log(r'Mr / r0'Mr0)
//  4 log(r'Mr / r_prev'Mr_prev)   5 lam_prev rho   6 log(|r_prev|/|r_prev2|)   7 k/100
//  8 log(1+k)/5   9 log(|r|/|e|)   10 log(|r_prev|/|e|)   11 [|r| > |e|]
//  (9-11 zero when |e| is unknown)   12 lam_prev2 rho   13 feature 4 of the last iteration
//  14 log(r'Mr / |r|^2) - log(r0'Mr0 / |r0|^2)   15 [k == 1]
__global__ __launch_bounds__(256) void k_select(const double* hist, Ctrl* ctrl, LearnBufs L, int k) {
    __shared__ float f[LEARN_NF];
    __shared__ float h1[MLP_HID], h2[MLP_HID];
    __shared__ float Ws[MLP_PARAMS];
    __shared__ double sums[4];
    if (load_stop(ctrl)) return;
    const int t = threadIdx.x;
    // the network's parameters, staged with coalesced loads while wave 0 works
    if (L.mlp)
        for (int i = t; i < MLP_PARAMS; i += 256) Ws[i] = L.mlp[i];
    // wave 0 reduces the greedy sums and computes the features
    double b0 = 0, b1 = 0, b2 = 0;
    if (L.xstar && t < 64)
        for (int i = t; i < L.gblocks; i += 64) {
            b0 += L.gpart[3 * i];
            b1 += L.gpart[3 * i + 1];
            b2 += L.gpart[3 * i + 2];
        }
    for (int d = 32; d > 0; d >>= 1) {
        b0 += __shfl_xor(b0, d);
        b1 += __shfl_xor(b1, d);
        b2 += __shfl_xor(b2, d);
    }
    if (t == 0) {
        // state: 0 |r0|  1 r0'Mr0  2 r_prev'Mr_prev  3 unused  4 f4 prev  5 lam_prev  6 lam_prev2
        double* S = L.state;
        const double nr = hist[k - 1], nr1 = k >= 2 ? hist[k - 2] : nr, nr2 = k >= 3 ? hist[k - 3] : nr1;
        const double rMr = ctrl->rMr;
        if (k == 1) {
            for (int q = 0; q < LEARN_STATE; ++q) S[q] = 0;
            S[0] = nr;
            S[1] = rMr;
            S[2] = rMr;
        }
        auto lg = [](double v) { return v > 0 ? log(v) : -30.0; };
        const double rho = (double)L.rho;
        double fv[LEARN_NF];
        fv[0] = lg(nr / S[0]);
        fv[1] = lg(nr1 / S[0]);
        fv[2] = lg(nr / nr1);
        fv[3] = lg(rMr / S[1]);
        fv[4] = lg(rMr / S[2]);
        fv[5] = S[5] * rho;
        fv[6] = lg(nr1 / nr2);
        fv[7] = (k - 1) / 100.0;
        fv[8] = log1p((double)(k - 1)) / 5.0;
        const bool hasn = L.noise > 0;
        fv[9] = hasn ? lg(nr / L.noise) : 0.0;
        fv[10] = hasn ? lg(nr1 / L.noise) : 0.0;
        fv[11] = hasn ? (nr > L.noise ? 1.0 : 0.0) : 0.0;
        fv[12] = S[6] * rho;
        fv[13] = S[4];
        fv[14] = lg(rMr / (nr * nr)) - lg(S[1] / (S[0] * S[0]));
        fv[15] = k == 1 ? 1.0 : 0.0;
        S[2] = rMr;
        S[4] = fv[4];
        for (int q = 0; q < LEARN_NF; ++q) {
            f[q] = (float)fv[q];
            if (L.feat) L.feat[(int64_t)(k - 1) * LEARN_NF + q] = (float)fv[q];
        }
        sums[0] = b0;
        sums[1] = b1;
        sums[2] = b2;
        if (L.xstar && L.err) L.err[k - 1] = sqrt(b2);
    }
    __syncthreads();                  // features and staged weights visible to the block
    // greedy target: lambda* minimises ||x + lam d - x*|| before projection, in the network's range
    double lr_star = 1.0;
    if (L.mode == AIR_RELAX_GREEDY && t == 0) {
        lr_star = sums[1] > 0 ? sums[0] / sums[1] * (double)L.rho : 1.0;
        lr_star = fmin(fmax(lr_star, 0.02), (double)MLP_LAMBDA_MAX - 0.02);
        if (L.target) L.target[k - 1] = (float)lr_star;
    }
    if (!L.mlp) {
        if (t == 0) {
            const double lam = lr_star / (double)L.rho;
            ctrl->lam_cur = lam;
            L.state[6] = L.state[5];
            L.state[5] = lam;
        }
        return;                       // uniform across the block (L.mlp is a kernel argument)
    }
    // the network: 16 -> 64 -> 64 -> 2, FP32, one hidden unit per lane.  In a
    // greedy run with parameters the network's lambda drives the iteration and
    // the greedy target is only recorded (data along the network's own path).
    const float* W = Ws;
    if (t < MLP_HID) {
        float z = W[MLP_HID * MLP_IN + t];
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) z += W[t * MLP_IN + i] * f[i];
        h1[t] = fmaxf(z, 0.f);
    }
    __syncthreads();
    if (t < MLP_HID) {
        const float* W2 = W + MLP_HID * MLP_IN + MLP_HID;
        float z = W2[MLP_HID * MLP_HID + t];
#pragma unroll 16
        for (int i = 0; i < MLP_HID; ++i) z += W2[t * MLP_HID + i] * h1[i];
        h2[t] = fmaxf(z, 0.f);
    }
    __syncthreads();
    if (t >= 64) return;
    const float* W3 = W + MLP_HID * MLP_IN + MLP_HID + MLP_HID * MLP_HID + MLP_HID;
    float o0 = W3[t] * h2[t], o1 = W3[MLP_HID + t] * h2[t];
    for (int d = 32; d > 0; d >>= 1) {
        o0 += __shfl_xor(o0, d);
        o1 += __shfl_xor(o1, d);
    }
    if (t == 0) {
        o0 += W3[2 * MLP_HID];
        o1 += W3[2 * MLP_HID + 1];
        if (L.out) {
            L.out[2 * (k - 1)] = o0;
            L.out[2 * (k - 1) + 1] = o1;
        }
        const double lam = (double)MLP_LAMBDA_MAX / (1.0 + exp(-(double)o0)) / (double)L.rho;
        ctrl->lam_cur = lam;
        L.state[6] = L.state[5];
        L.state[5] = lam;
        if (L.mode == AIR_RELAX_LEARNED && o1 > 0.f) {   // stop at the state before this update
            ctrl->finaliter = k - 1;
            ctrl->stoprule = AIR_STOP_LEARNED;
            __hip_atomic_store(&ctrl->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// learned relaxation: lambda_k (and the stop) before the fused column pass of iteration k
void launch_learned_select(const SirtBufs& B, int k, hipStream_t s) {
    const LearnBufs& L = *B.learn;
    if (L.xstar) {             // error record only (no step direction yet)
        k_greedy_dots<<<L.gblocks, BLOCK, 0, s>>>(L.nx, nullptr, nullptr, B.x, L.xstar, L.gpart, B.ctrl);
        AIR_HIP(hipGetLastError());
    }
    k_select<<<1, 256, 0, s>>>(B.hist, B.ctrl, L, k);
    AIR_HIP(hipGetLastError());
}

void launch_finalize_line(const SirtBufs& B, int nblocks, int k, hipStream_t s) {
    if (B.learn) {
        const LearnBufs& L = *B.learn;
        if (L.xstar) {
            k_greedy_dots<<<L.gblocks, BLOCK, 0, s>>>(L.nx, B.g, B.D, B.x, L.xstar, L.gpart, B.ctrl);
            AIR_HIP(hipGetLastError());
        }
        (void)nblocks;
        k_select<<<1, 256, 0, s>>>(B.hist, B.ctrl, L, k);
        AIR_HIP(hipGetLastError());
        return;
    }
    k_finalize_line<<<1, BLOCK, 0, s>>>(B.partial_cols, nblocks, B.ctrl);
    AIR_HIP(hipGetLastError());
}

void launch_update(const DevSparse& AT, Shape sh, const SirtBufs& B, double lam, int k, hipStream_t s) {
    (void)k;
    dispatch_dtype(AT.dtype, [&](auto v) {
        using V = decltype(v);
        dispatch_G(sh.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            k_update<V, G><<<sh.blocks, BLOCK, 0, s>>>(AT.ptr.as<int64_t>(), AT.idx.as<int>(), AT.val.as<V>(),
                                                       AT.nrows, B.Mr, B.D, B.x, lam, B.lb_kind, B.lb, B.lbv,
                                                       B.ub_kind, B.ub, B.ubv, B.ctrl);
        });
        return 0;
    });
    AIR_HIP(hipGetLastError());
}

void launch_line_update(const DevSparse& AT, Shape sh, const SirtBufs& B, int k, hipStream_t s) {
    (void)k;
    dispatch_dtype(AT.dtype, [&](auto v) {
        using V = decltype(v);
        dispatch_G(sh.G, [&](auto g) {
            constexpr int G = decltype(g)::value;
            k_gcols<V, G><<<sh.blocks, BLOCK, 0, s>>>(AT.ptr.as<int64_t>(), AT.idx.as<int>(), AT.val.as<V>(),
                                                      AT.nrows, B.Mr, B.D, B.g, B.partial_cols, B.ctrl);
        });
        return 0;
    });
    AIR_HIP(hipGetLastError());
    launch_finalize_line(B, sh.blocks, k, s);
    k_apply<<<vec_grid(AT.nrows), BLOCK, 0, s>>>(AT.nrows, B.g, B.D, B.x, 0.0, 1, B.lb_kind, B.lb, B.lbv, B.ub_kind,
                                                 B.ub, B.ubv, B.ctrl);
    AIR_HIP(hipGetLastError());
}

void launch_apply_line(int64_t n, const SirtBufs& B, hipStream_t s) {
    k_apply<<<vec_grid(n), BLOCK, 0, s>>>(n, B.g, B.D, B.x, 0.0, 1, B.lb_kind, B.lb, B.lbv, B.ub_kind, B.ub, B.ubv,
                                          B.ctrl);
    AIR_HIP(hipGetLastError());
}

void launch_mf_resid(const Geometry& g, const SirtBufs& B, StopParams sp, int k, float* Ax, hipStream_t s) {
    mf_forward(g, B.x, Ax, s);
    const int64_t m = g.nrays;
    unsigned grid = std::min<unsigned>(vec_grid(m), 2048);
    if (sp.type == AIR_STOP_ME)
        k_resid_vec<true><<<grid, BLOCK, 0, s>>>(m, Ax, B.b, B.M, B.r, B.Mr, B.partial, B.hist, B.ctrl, sp, k);
    else
        k_resid_vec<false><<<grid, BLOCK, 0, s>>>(m, Ax, B.b, B.M, B.r, B.Mr, B.partial, B.hist, B.ctrl, sp, k);
    AIR_HIP(hipGetLastError());
    launch_finalize_rows(B, (int)grid, sp, k, s);
}

void launch_mf_update(const Geometry& g, int64_t n, const SirtBufs& B, double lam, int line, int k, hipStream_t s) {
    (void)k;
    k_zero_f64<<<vec_grid(n), BLOCK, 0, s>>>(B.g, n);
    mf_adjoint(g, B.Mr, B.g, s);
    if (line) {
        unsigned grid = std::min<unsigned>(vec_grid(n), 2048);
        k_gsum<<<grid, BLOCK, 0, s>>>(n, B.g, B.D, B.partial_cols, B.ctrl);
        launch_finalize_line(B, (int)grid, k, s);
    }
    k_apply<<<vec_grid(n), BLOCK, 0, s>>>(n, B.g, B.D, B.x, lam, line, B.lb_kind, B.lb, B.lbv, B.ub_kind, B.ub,
                                          B.ubv, B.ctrl);
    AIR_HIP(hipGetLastError());
}

void vec_f64_to_f32(const double* in, float* out, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    k_f64_to_f32<<<vec_grid(n), BLOCK, 0, s>>>(in, out, n);
    AIR_HIP(hipGetLastError());
}
void vec_f32_to_f64(const float* in, double* out, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    k_f32_to_f64<<<vec_grid(n), BLOCK, 0, s>>>(in, out, n);
    AIR_HIP(hipGetLastError());
}
void vec_mul_f32(const float* a, float* y, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    k_mul_f32<<<vec_grid(n), BLOCK, 0, s>>>(a, y, n);
    AIR_HIP(hipGetLastError());
}
void zero_f64(double* p, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    k_zero_f64<<<vec_grid(n), BLOCK, 0, s>>>(p, n);
    AIR_HIP(hipGetLastError());
}

}  // namespace air
