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

#include "air_internal.h"

namespace air {

static constexpr int BLOCK = 256;

template <typename T>
__device__ __forceinline__ T ldnt(const T* p) {
    return __builtin_nontemporal_load(p);
}
__device__ __forceinline__ float ldval(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ double ldval(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ bf16_t ldval(const bf16_t* p) {
    return bf16_t{__builtin_nontemporal_load((const uint16_t*)p)};
}
__device__ __forceinline__ f16_t ldval(const f16_t* p) {
    return f16_t{__builtin_nontemporal_load((const uint16_t*)p)};
}

template <typename V> struct AccOf { using T = float; };
template <> struct AccOf<double> { using T = double; };
__device__ __forceinline__ float widen(float v, float) { return v; }
__device__ __forceinline__ double widen(double v, double) { return v; }
template <typename V> __device__ __forceinline__ float widen(V v, float) { return to_f32(v); }

template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sum over one compressed row: sum_q val[q] * vec[idx[q]]
template <typename V, int G>
__device__ __forceinline__ typename AccOf<V>::T row_dot(const int* __restrict__ idx, const V* __restrict__ val,
                                                       const float* __restrict__ vec, int64_t s, int64_t e, int gl) {
    using A = typename AccOf<V>::T;
    A a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int64_t q = s + gl;
    for (; q + 3 * G < e; q += 4 * G) {
        V v0 = ldval(val + q), v1 = ldval(val + q + G), v2 = ldval(val + q + 2 * G), v3 = ldval(val + q + 3 * G);
        int c0 = ldnt(idx + q), c1 = ldnt(idx + q + G), c2 = ldnt(idx + q + 2 * G), c3 = ldnt(idx + q + 3 * G);
        a0 += widen(v0, A{}) * (A)vec[c0];
        a1 += widen(v1, A{}) * (A)vec[c1];
        a2 += widen(v2, A{}) * (A)vec[c2];
        a3 += widen(v3, A{}) * (A)vec[c3];
    }
    for (; q < e; q += G) a0 += widen(ldval(val + q), A{}) * (A)vec[ldnt(idx + q)];
    return (a0 + a1) + (a2 + a3);
}

__device__ __forceinline__ int load_stop(const Ctrl* c) {
    return __hip_atomic_load(&c->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- y = S x
template <typename V, int G>
__global__ __launch_bounds__(BLOCK) void k_spmv(const int64_t* __restrict__ ptr, const int* __restrict__ idx,
                                                const V* __restrict__ val, int64_t nrows,
                                                const float* __restrict__ x, float* __restrict__ y) {
    constexpr int GPB = BLOCK / G;
    const int gl = threadIdx.x % G;
    const int64_t stride = (int64_t)gridDim.x * GPB;
    for (int64_t row = (int64_t)blockIdx.x * GPB + threadIdx.x / G; row < nrows; row += stride) {
        auto acc = row_dot<V, G>(idx, val, x, ptr[row], ptr[row + 1], gl);
        acc = group_sum<G>(acc);
        if (gl == 0) y[row] = (float)acc;
    }
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

// ---------------------------------------------------------------- block reduction + last-block finalize
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double (*sh)[BLOCK / 64]) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double t = group_sum<64>(v[q]);
        if (lane == 0) sh[q][wid] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double t = 0;
            for (int w = 0; w < BLOCK / 64; ++w) t += sh[q][w];
            v[q] = t;
        }
    }
}

// Stop rules of auxil/check_stoprules.m on the device (thread 0 of the last block).
__device__ void evaluate_stop(Ctrl* ctrl, const StopParams sp, int k, double rr, double rMr, double me_num,
                              double rkm1sq, double* hist) {
    double nrk = sqrt(rr);
    hist[k] = nrk;
    ctrl->rMr = rMr;
    ctrl->last_norm = nrk;
    int stop = 0, code = 0;
    switch (sp.type) {
    case AIR_STOP_DP:   // :56-72
        if (nrk <= sp.taudelta || k >= sp.kmax) { stop = 1; code = (k != sp.kmax) ? 2 : 0; }
        break;
    case AIR_STOP_ME: { // :74-94 (rkm1 = NaN at k = 0)
        double dme = me_num / 2.0, nr1 = sqrt(rkm1sq);
        if (k > 0 && dme / nr1 <= sp.taudelta) { stop = 1; code = 3; }
        else if (k >= sp.kmax) { stop = 1; code = 0; }
        break;
    }
    default:            // NONE :153-161 (NCP is evaluated on the host)
        if (k >= sp.kmax) { stop = 1; code = 0; }
        break;
    }
    if (stop) {
        ctrl->finaliter = k;
        ctrl->stoprule = code;
        __hip_atomic_store(&ctrl->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int NV>
__device__ __forceinline__ bool publish_partials(double (&v)[NV], double* partial, unsigned int* counter) {
    __shared__ bool last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) partial[(int64_t)blockIdx.x * NV + q] = v[q];
        __threadfence();
        unsigned int prev = atomicAdd(counter, 1u);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    return last;
}

template <int NV>
__device__ __forceinline__ void reduce_partials(const double* partial, double (&out)[NV], double (*sh)[BLOCK / 64]) {
    __threadfence();
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += BLOCK) {
#pragma unroll
        for (int q = 0; q < NV; ++q)
            v[q] += __hip_atomic_load(&partial[(int64_t)b * NV + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    block_reduce<NV>(v, sh);
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = v[q];
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
    constexpr int GPB = BLOCK / G;
    const int gl = threadIdx.x % G;
    const int64_t stride = (int64_t)gridDim.x * GPB;
    double v[4] = {0, 0, 0, 0};
    for (int64_t row = (int64_t)blockIdx.x * GPB + threadIdx.x / G; row < m; row += stride) {
        auto acc = row_dot<V, G>(idx, val, x, ptr[row], ptr[row + 1], gl);
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
    }
    block_reduce<4>(v, sh);
    if (publish_partials<4>(v, partial, &ctrl->cnt_rows)) {
        double t[4];
        reduce_partials<4>(partial, t, sh);
        if (threadIdx.x == 0) {
            evaluate_stop(ctrl, sp, k, t[0], t[1], t[2], t[3], hist);
            ctrl->cnt_rows = 0;
        }
    }
}

__device__ __forceinline__ float project(double xn, int j, int lb_kind, float lb, const float* lbv, int ub_kind,
                                         float ub, const float* ubv) {
    // sirt.m:232-237: lower bound first, then upper bound (NaN-ignoring max/min)
    if (lb_kind == AIR_BOUND_SCALAR) xn = fmax(xn, (double)lb);
    else if (lb_kind == AIR_BOUND_VECTOR) xn = fmax(xn, (double)lbv[j]);
    if (ub_kind == AIR_BOUND_SCALAR) xn = fmin(xn, (double)ub);
    else if (ub_kind == AIR_BOUND_VECTOR) xn = fmin(xn, (double)ubv[j]);
    return (float)xn;
}

// ---------------------------------------------------------------- K3+K4: update pass (CSC)
template <typename V, int G>
__global__ __launch_bounds__(BLOCK) void k_update(const int64_t* __restrict__ ptr, const int* __restrict__ idx,
                                                  const V* __restrict__ val, int64_t n, const float* __restrict__ Mr,
                                                  const float* __restrict__ D, float* __restrict__ x, double lam,
                                                  int lb_kind, float lb, const float* __restrict__ lbv, int ub_kind,
                                                  float ub, const float* __restrict__ ubv, Ctrl* ctrl) {
    if (load_stop(ctrl)) return;
    constexpr int GPB = BLOCK / G;
    const int gl = threadIdx.x % G;
    const int64_t stride = (int64_t)gridDim.x * GPB;
    for (int64_t j = (int64_t)blockIdx.x * GPB + threadIdx.x / G; j < n; j += stride) {
        auto acc = row_dot<V, G>(idx, val, Mr, ptr[j], ptr[j + 1], gl);
        acc = group_sum<G>(acc);
        if (gl == 0) {
            double dg = D ? (double)D[j] * (double)(float)acc : (double)(float)acc;
            double xn = (double)x[j] + lam * dg;
            x[j] = project(xn, (int)j, lb_kind, lb, lbv, ub_kind, ub, ubv);
        }
    }
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
    constexpr int GPB = BLOCK / G;
    const int gl = threadIdx.x % G;
    const int64_t stride = (int64_t)gridDim.x * GPB;
    double v[1] = {0};
    for (int64_t j = (int64_t)blockIdx.x * GPB + threadIdx.x / G; j < n; j += stride) {
        auto acc = row_dot<V, G>(idx, val, Mr, ptr[j], ptr[j + 1], gl);
        acc = group_sum<G>(acc);
        if (gl == 0) {
            double gj = (double)(float)acc;
            g[j] = gj;
            v[0] += D ? (double)D[j] * (gj * gj) : gj * gj;
        }
    }
    block_reduce<1>(v, sh);
    if (publish_partials<1>(v, partial, &ctrl->cnt_cols)) {
        double t[1];
        reduce_partials<1>(partial, t, sh);
        if (threadIdx.x == 0) {
            ctrl->lam_cur = ctrl->rMr / t[0];
            ctrl->cnt_cols = 0;
        }
    }
}

// x = P_C(x + lam D g) with lam from the control block (line) or the argument.
__global__ __launch_bounds__(BLOCK) void k_apply(int64_t n, const double* __restrict__ g, const float* __restrict__ D,
                                                 float* __restrict__ x, double lam, int use_ctrl_lam, int lb_kind,
                                                 float lb, const float* __restrict__ lbv, int ub_kind, float ub,
                                                 const float* __restrict__ ubv, Ctrl* ctrl) {
    if (load_stop(ctrl)) return;
    double l = use_ctrl_lam ? ctrl->lam_cur : lam;
    for (int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x; j < n; j += (int64_t)gridDim.x * BLOCK) {
        double gj = (double)(float)g[j];
        double dg = D ? (double)D[j] * gj : gj;
        x[j] = project((double)x[j] + l * dg, (int)j, lb_kind, lb, lbv, ub_kind, ub, ubv);
    }
    if (!use_ctrl_lam && threadIdx.x == 0 && blockIdx.x == 0) ctrl->lam_cur = lam;
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
    if (publish_partials<1>(v, partial, &ctrl->cnt_cols)) {
        double t[1];
        reduce_partials<1>(partial, t, sh);
        if (threadIdx.x == 0) {
            ctrl->lam_cur = ctrl->rMr / t[0];
            ctrl->cnt_cols = 0;
        }
    }
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
    if (publish_partials<4>(v, partial, &ctrl->cnt_rows)) {
        double t[4];
        reduce_partials<4>(partial, t, sh);
        if (threadIdx.x == 0) {
            evaluate_stop(ctrl, sp, k, t[0], t[1], t[2], t[3], hist);
            ctrl->cnt_rows = 0;
        }
    }
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
    sh.G = L >= 96 ? 64 : (L >= 40 ? 32 : (L >= 20 ? 16 : 8));
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
                                                      AT.nrows, B.Mr, B.D, B.g, B.partial + 8192, B.ctrl);
        });
        return 0;
    });
    AIR_HIP(hipGetLastError());
    k_apply<<<vec_grid(AT.nrows), BLOCK, 0, s>>>(AT.nrows, B.g, B.D, B.x, 0.0, 1, B.lb_kind, B.lb, B.lbv, B.ub_kind,
                                                 B.ub, B.ubv, B.ctrl);
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
}

void launch_mf_update(const Geometry& g, int64_t n, const SirtBufs& B, double lam, int line, int k, hipStream_t s) {
    (void)k;
    k_zero_f64<<<vec_grid(n), BLOCK, 0, s>>>(B.g, n);
    mf_adjoint(g, B.Mr, B.g, s);
    if (line) {
        unsigned grid = std::min<unsigned>(vec_grid(n), 2048);
        k_gsum<<<grid, BLOCK, 0, s>>>(n, B.g, B.D, B.partial + 8192, B.ctrl);
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
