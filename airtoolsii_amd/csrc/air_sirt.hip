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
// inputs of the learned selection for iteration k: |r_{k-1}|, |r_{k-2}|,
// |r_{k-3}|, r'Mr of x_{k-1}, the carried state and the stop flag
struct SelIn {
    double nr, nr1, nr2, rMr;
    double S[LEARN_STATE];
    int stop;
};
__device__ __forceinline__ void select_block(const SelIn& in, Ctrl* ctrl, const LearnBufs& L, int k, float* Ws,
                                             bool staged);
__device__ __forceinline__ void stage_weights(const float* __restrict__ mlp, float* Ws);
constexpr int NW_LANE_F = (MLP_PARAMS + 255) / 256;
__device__ __forceinline__ void load_weights(const float* __restrict__ mlp, float (&w)[NW_LANE_F]);
__device__ __forceinline__ void store_weights(const float (&w)[NW_LANE_F], float* Ws);

// k < 0 (graph mode): the iteration number comes from ctrl->kcur, and the
// next iteration's lambda is staged in ctrl->lam_cur from lamv.  sel: the
// learned relaxation's lambda_{k+1} and stop are computed here, at the end of
// iteration k, instead of in a kernel of their own before the next update.
__global__ __launch_bounds__(BLOCK) void k_finalize_rows(const double* __restrict__ partial, int nblocks, double* hist,
                                                         Ctrl* ctrl, StopParams sp, int k, const double* lamv,
                                                         LearnBufs Lsel, int sel) {
    __shared__ double sh[4][BLOCK / 64];
    __shared__ float Ws[MLP_PARAMS];
    __shared__ SelIn in;
    __shared__ double hprev[2];
    if (load_stop(ctrl)) return;
    if (k < 0) k = ctrl->kcur + 1;
    const bool do_sel = sel && k < sp.kmax;
    float wreg[NW_LANE_F];
    double spre = 0.0;
    if (do_sel) {                                    // issued now, stored after the reduction below
        load_weights(Lsel.mlp, wreg);
        if (threadIdx.x < LEARN_STATE) spre = Lsel.state[threadIdx.x];
        else if (threadIdx.x == 64) spre = k >= 1 ? hist[k - 1] : 0.0;
        else if (threadIdx.x == 65) spre = k >= 2 ? hist[k - 2] : 0.0;
    }
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
    if (do_sel) {
        store_weights(wreg, Ws);
        if (threadIdx.x < LEARN_STATE) in.S[threadIdx.x] = spre;
        else if (threadIdx.x == 64) hprev[0] = spre;
        else if (threadIdx.x == 65) hprev[1] = spre;
    }
    block_reduce<4>(v, sh);
    double c[1] = {0};
    if (sp.ncp_cols > 0) {
        __shared__ double shc[1][BLOCK / 64];
        for (int j = threadIdx.x; j < sp.ncp_cols; j += BLOCK) c[0] += sp.ncp_stat[j];
        block_reduce<1>(c, shc);
    }
    if (threadIdx.x == 0) {
        const int st = evaluate_stop(ctrl, sp, k, v[0], v[1], v[2], v[3], c[0], hist);
        ctrl->kcur = k;
        if (lamv && k < sp.kmax) ctrl->lam_cur = lamv[k];
        if (do_sel) {                                // the selection for iteration k + 1
            in.stop = st;
            in.nr = sqrt(v[0]);                      // hist[k], as evaluate_stop stores it
            in.rMr = v[1];
            in.nr1 = k >= 1 ? hprev[0] : in.nr;
            in.nr2 = k >= 2 ? hprev[1] : in.nr1;
        }
    }
    if (do_sel) {                                    // uniform: kernel arguments and k
        __syncthreads();                             // in, Ws and hprev visible to the block
        select_block(in, ctrl, Lsel, k + 1, Ws, true);
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

// transform of the residual copy in N.rd and the per-column statistics
static void ncp_stats_of_rd(NcpState& N, const Ctrl* ctrl, hipStream_t s) {
    AIR_CHECK(hipfftExecD2Z(N.plan, N.rd.as<double>(), (hipfftDoubleComplex*)N.spec.p) == HIPFFT_SUCCESS, AIR_E_HIP,
              "hipfftExecD2Z failed");
    k_ncp_stats<<<(unsigned)N.ncols, BLOCK, 0, s>>>((const double2*)N.spec.p, N.len, N.colstat.as<double>(), ctrl);
    AIR_HIP(hipGetLastError());
}

void ncp_enqueue(NcpState& N, const float* r, int64_t m, const Ctrl* ctrl, hipStream_t s) {
    k_f32_to_f64<<<(unsigned)std::min<int64_t>((m + BLOCK - 1) / BLOCK, 4096), BLOCK, 0, s>>>(r, N.rd.as<double>(), m);
    AIR_HIP(hipGetLastError());
    ncp_stats_of_rd(N, ctrl, s);
}

void ncp_enqueue_f64(NcpState& N, const double* r, int64_t m, const Ctrl* ctrl, hipStream_t s) {
    AIR_HIP(hipMemcpyAsync(N.rd.p, r, 8 * (size_t)m, hipMemcpyDeviceToDevice, s));
    ncp_stats_of_rd(N, ctrl, s);
}

// A learned run without the error record selects lambda_{k+1} inside the finalize of iteration k.
bool fused_select(const SirtBufs& B) {
    return B.learn && B.learn->mode == AIR_RELAX_LEARNED && B.learn->mlp && !B.learn->xstar;
}

// Deferred norms: block j of the launch reduces the row-pass partials of
// iteration k0 + j (ring slot (k0 + j) % slots, in the block order of
// k_finalize_rows, so hist matches the direct form bit for bit); the last
// block also runs the stop check of k1 (the rule is 'none': k1 >= kmax stops).
__global__ __launch_bounds__(BLOCK) void k_finalize_deferred(const double* __restrict__ ring, int nblocks, int slots,
                                                             int k0, int k1, double* hist, Ctrl* ctrl, StopParams sp) {
    __shared__ double sh[4][BLOCK / 64];
    if (load_stop(ctrl)) return;
    const int k = k0 + (int)blockIdx.x;
    const double* partial = ring + (size_t)(k % slots) * (size_t)nblocks * 4;
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
    if (threadIdx.x == 0) {
        if (k < k1) {
            hist[k] = sqrt(v[0]);
        } else {
            evaluate_stop(ctrl, sp, k, v[0], v[1], v[2], v[3], 0.0, hist);
            ctrl->kcur = k;
        }
    }
}

void launch_finalize_deferred(const SirtBufs& B, int nblocks, StopParams sp, int k0, int k1, hipStream_t s) {
    if (k1 < k0) return;
    k_finalize_deferred<<<(unsigned)(k1 - k0 + 1), BLOCK, 0, s>>>(B.ring, nblocks, B.ring_slots, k0, k1, B.hist,
                                                                   B.ctrl, sp);
    AIR_HIP(hipGetLastError());
}

void launch_finalize_rows(const SirtBufs& B, int nblocks, StopParams sp, int k, hipStream_t s) {
    if (sp.type == AIR_STOP_NCP && B.ncp) ncp_enqueue(*B.ncp, B.r, B.ncp->m, B.ctrl, s);
    const bool sel = fused_select(B);
    k_finalize_rows<<<1, BLOCK, 0, s>>>(B.partial, nblocks, B.hist, B.ctrl, sp, k, B.lamv,
                                        sel ? *B.learn : LearnBufs{}, sel ? 1 : 0);
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
// The features of the state x_{k-1}, the network and lambda_k (and the stop)
// for one block of 256 threads; shared arrays come from the caller.
constexpr int P_W2 = MLP_HID * MLP_IN + MLP_HID;

// The network's parameters staged input-major (W1 and W2 transposed), so that
// the 64 lanes computing one layer's outputs read 64 consecutive LDS words.
// A lane's share of the parameters: all loads are issued before the first
// LDS store, so the block waits for one memory round trip instead of one per
// step; the finalize kernel issues them before its reduction and stores them
// after it.
__device__ __forceinline__ void load_weights(const float* __restrict__ mlp, float (&w)[NW_LANE_F]) {
#pragma unroll
    for (int j = 0; j < NW_LANE_F; ++j) {
        const int q = (int)threadIdx.x + 256 * j;
        w[j] = q < MLP_PARAMS ? mlp[q] : 0.f;
    }
}

__device__ __forceinline__ void store_weights(const float (&w)[NW_LANE_F], float* Ws) {
#pragma unroll
    for (int j = 0; j < NW_LANE_F; ++j) {
        const int q = (int)threadIdx.x + 256 * j;
        if (q >= MLP_PARAMS) continue;
        int d = q;
        if (q < MLP_HID * MLP_IN) d = (q % MLP_IN) * MLP_HID + q / MLP_IN;
        else if (q >= P_W2 && q < P_W2 + MLP_HID * MLP_HID) {
            const int r = q - P_W2;
            d = P_W2 + (r % MLP_HID) * MLP_HID + r / MLP_HID;
        }
        Ws[d] = w[j];
    }
}

__device__ __forceinline__ void stage_weights(const float* __restrict__ mlp, float* Ws) {
    float w[NW_LANE_F];
    load_weights(mlp, w);
    store_weights(w, Ws);
}

__device__ __forceinline__ void select_block(const SelIn& in, Ctrl* ctrl, const LearnBufs& L, int k, float* Ws,
                                             bool staged) {
    if (in.stop) return;
    if (L.mlp && !staged) {           // uniform (kernel argument): the whole block stages
        stage_weights(L.mlp, Ws);
        __syncthreads();
    }
    // from here on wave 0 alone: the features, the greedy sums and the network
    // are wave-synchronous (lane broadcasts, no block barriers)
    if (threadIdx.x >= 64) return;
    const int t = threadIdx.x;
    double b0 = 0, b1 = 0, b2 = 0;
    if (L.xstar)
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
    // state: 0 |r0|  1 r0'Mr0  2 r_prev'Mr_prev  3 unused  4 f4 prev  5 lam_prev  6 lam_prev2
    const double nr = in.nr, nr1 = in.nr1, nr2 = in.nr2;
    const double rMr = in.rMr;
    const bool first = k == 1;
    const double* S = in.S;
    const double S0 = first ? nr : S[0], S1 = first ? rMr : S[1], S2 = first ? rMr : S[2];
    // the features are FP32 network inputs, so FP32 logs suffice; ratios are
    // clamped to [1e-30, 1e30] (non-positive ones give -30)
    auto lg = [](double v) { return v > 1e-30 ? (double)__logf((float)fmin(v, 1e30)) : -30.0; };
    // lanes pick their own arguments, then all of them make the same two log
    // calls (a switch around the calls would run 16 logs one after another)
    const double rho = (double)L.rho;
    const bool hasn = L.noise > 0;
    double an = 1.0, ad = 1.0, cn = 1.0, cd = 1.0, lin = 0.0;
    int kind = 1;                     // 0: lg(an/ad) - lg(cn/cd), 1: lin, 2: lg(lin) / 5
    switch (t) {
    case 0: kind = 0; an = nr; ad = S0; break;
    case 1: kind = 0; an = nr1; ad = S0; break;
    case 2: kind = 0; an = nr; ad = nr1; break;
    case 3: kind = 0; an = rMr; ad = S1; break;
    case 4: kind = 0; an = rMr; ad = S2; break;
    case 5: lin = (first ? 0.0 : S[5]) * rho; break;
    case 6: kind = 0; an = nr1; ad = nr2; break;
    case 7: lin = (k - 1) / 100.0; break;
    case 8: kind = 2; lin = (double)k; break;                          // log1p(k - 1) / 5
    case 9: if (hasn) { kind = 0; an = nr; ad = L.noise; } break;
    case 10: if (hasn) { kind = 0; an = nr1; ad = L.noise; } break;
    case 11: lin = hasn ? (nr > L.noise ? 1.0 : 0.0) : 0.0; break;
    case 12: lin = (first ? 0.0 : S[6]) * rho; break;
    case 13: lin = first ? 0.0 : S[4]; break;
    case 14: kind = 0; an = rMr; ad = nr * nr; cn = S1; cd = S0 * S0; break;
    case 15: lin = first ? 1.0 : 0.0; break;
    default: break;                   // lanes 16..63 carry no feature
    }
    const double a = an / ad, c = cn / cd;
    const double la = lg(kind == 2 ? lin : a), lc = lg(c);
    const double v = kind == 1 ? lin : (kind == 2 ? la / 5.0 : la - lc);
    const float fv = (float)v;
    if (t < LEARN_NF && L.feat) L.feat[(int64_t)(k - 1) * LEARN_NF + t] = fv;
    const double f4 = __shfl(v, 4, 64);
    if (t == 0) {
        if (L.xstar && L.err) L.err[k - 1] = sqrt(b2);
        double* Sw = L.state;         // the state of the next call (every lane has read in.S)
        if (first) {
            for (int q = 0; q < LEARN_STATE; ++q) Sw[q] = 0;
            Sw[0] = nr;
            Sw[1] = rMr;
        }
        Sw[4] = f4;
        Sw[2] = rMr;
    }
    // greedy target: lambda* minimises ||x + lam d - x*|| before projection, in the network's range
    double lr_star = 1.0;
    if (L.mode == AIR_RELAX_GREEDY && t == 0) {
        lr_star = b1 > 0 ? b0 / b1 * (double)L.rho : 1.0;
        lr_star = fmin(fmax(lr_star, 0.02), (double)MLP_LAMBDA_MAX - 0.02);
        if (L.target) L.target[k - 1] = (float)lr_star;
    }
    if (!L.mlp) {
        if (t == 0) {
            const double lam = lr_star / (double)L.rho;
            ctrl->lam_cur = lam;
            L.state[6] = first ? 0.0 : in.S[5];
            L.state[5] = lam;
        }
        return;
    }
    // the network: 16 -> 64 -> 64 -> 2, FP32, lane o computes hidden unit o of
    // each layer; the previous layer's values are broadcast lane by lane.  In
    // a greedy run with parameters the network's lambda drives the iteration
    // and the greedy target is only recorded (data along the network's path).
    const float* W = Ws;
    float z = W[MLP_HID * MLP_IN + t];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i)
        z += W[i * MLP_HID + t] * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fv), i));
    const float hv1 = fmaxf(z, 0.f);
    const float* W2 = W + P_W2;
    z = W2[MLP_HID * MLP_HID + t];
#pragma unroll 16
    for (int i = 0; i < MLP_HID; ++i)
        z += W2[i * MLP_HID + t] * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(hv1), i));
    const float hv2 = fmaxf(z, 0.f);
    const float* W3 = W2 + MLP_HID * MLP_HID + MLP_HID;
    float o0 = W3[t] * hv2, o1 = W3[MLP_HID + t] * hv2;
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
        const double lam = (double)MLP_LAMBDA_MAX / (1.0 + (double)expf(-o0)) / (double)L.rho;
        ctrl->lam_cur = lam;
        L.state[6] = first ? 0.0 : in.S[5];
        L.state[5] = lam;
        if (L.mode == AIR_RELAX_LEARNED && o1 > 0.f) {   // stop at the state before this update
            ctrl->finaliter = k - 1;
            ctrl->stoprule = AIR_STOP_LEARNED;
            __hip_atomic_store(&ctrl->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__global__ __launch_bounds__(256) void k_select(const double* hist, Ctrl* ctrl, LearnBufs L, int k) {
    __shared__ float Ws[MLP_PARAMS];
    __shared__ SelIn in;
    if (threadIdx.x == 0) {
        in.stop = load_stop(ctrl);
        in.nr = hist[k - 1];
        in.nr1 = k >= 2 ? hist[k - 2] : in.nr;
        in.nr2 = k >= 3 ? hist[k - 3] : in.nr1;
        in.rMr = ctrl->rMr;
        for (int q = 0; q < LEARN_STATE; ++q) in.S[q] = L.state[q];
    }
    __syncthreads();
    select_block(in, ctrl, L, k, Ws, false);
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
    launch_vec_resid(g.nrays, Ax, B, sp, k, s);
}

// r = b - Ax, M r, the FP64 norms and the stop rule from a computed A x (m values)
void launch_vec_resid(int64_t m, const float* Ax, const SirtBufs& B, StopParams sp, int k, hipStream_t s) {
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
    // every projector form writes each pixel of g; only an empty one leaves it
    if (g.nrays == 0) k_zero_f64<<<vec_grid(n), BLOCK, 0, s>>>(B.g, n);
    mf_adjoint(g, B.Mr, B.g, s);
    launch_g_update(n, B, lam, line, k, s);
}

// x <- P(x + lam D g) from a computed g = A' M r (FP64, n values); line: the
// line-search lambda from g first
void launch_g_update(int64_t n, const SirtBufs& B, double lam, int line, int k, hipStream_t s) {
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
