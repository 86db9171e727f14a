// air_device.h -- device helpers shared by the SIRT kernels (compressed rows and tiled).
#pragma once
#include "air_internal.h"

namespace air {

static constexpr int BLOCK = 256;

template <typename V> struct AccOf { using T = float; };
template <> struct AccOf<double> { using T = double; };

// 16-byte vector loads of values: VE values per lane per load.
template <typename V> struct VecOf;
template <> struct VecOf<float> {
    static constexpr int VE = 4;
    typedef float vt __attribute__((ext_vector_type(4)));
};
template <> struct VecOf<double> {
    static constexpr int VE = 2;
    typedef double vt __attribute__((ext_vector_type(2)));
};
template <> struct VecOf<bf16_t> {
    static constexpr int VE = 8;
    typedef unsigned short vt __attribute__((ext_vector_type(8)));
};
template <> struct VecOf<f16_t> {
    static constexpr int VE = 8;
    typedef unsigned short vt __attribute__((ext_vector_type(8)));
};
typedef int int4v __attribute__((ext_vector_type(4)));
typedef int int2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float elem(float v, float) { return v; }
__device__ __forceinline__ double elem(double v, double) { return v; }
__device__ __forceinline__ float elem_bits(unsigned short u, bf16_t) { return to_f32(bf16_t{u}); }
__device__ __forceinline__ float elem_bits(unsigned short u, f16_t) { return to_f32(f16_t{u}); }

template <typename V, typename A>
__device__ __forceinline__ void load_vals(const V* p, A (&out)[VecOf<V>::VE]) {
    typedef typename VecOf<V>::vt vt;
    vt w = __builtin_nontemporal_load((const vt*)p);
#pragma unroll
    for (int i = 0; i < VecOf<V>::VE; ++i) {
        if constexpr (sizeof(V) == 2) out[i] = elem_bits(w[i], V{});
        else out[i] = (A)w[i];
    }
}

template <int VE>
__device__ __forceinline__ void load_idx(const int* p, int (&out)[VE]) {
    if constexpr (VE == 2) {
        int2v w = __builtin_nontemporal_load((const int2v*)p);
        out[0] = w[0]; out[1] = w[1];
    } else {
#pragma unroll
        for (int h = 0; h < VE / 4; ++h) {
            int4v w = __builtin_nontemporal_load((const int4v*)p + h);
            out[4 * h + 0] = w[0]; out[4 * h + 1] = w[1]; out[4 * h + 2] = w[2]; out[4 * h + 3] = w[3];
        }
    }
}

template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Elements per round of one row group: G lanes x VE values x U loads.
template <typename V, int G> struct RowGeom {
    static constexpr int VE = VecOf<V>::VE;
    static constexpr int U = 2;
    static constexpr int STEP = G * VE;
};

// sum over one compressed row: sum_q val[q] * vec[idx[q]].  The row is read
// from its 16-byte aligned start with vector loads (entries outside [s,e) are
// masked); all U*VE gathers of a round are issued before any is consumed.
// The arrays carry AIR_SPARSE_PAD zeroed entries past nnz for the over-read.
template <typename V, int G>
__device__ __forceinline__ typename AccOf<V>::T row_dot(const int* __restrict__ idx, const V* __restrict__ val,
                                                       const float* __restrict__ vec, int64_t s, int64_t e, int gl) {
    using A = typename AccOf<V>::T;
    using RG = RowGeom<V, G>;
    constexpr int VE = RG::VE, U = RG::U, STEP = RG::STEP;
    A acc = 0;
    for (int64_t base = s & ~(int64_t)(VE - 1); base < e; base += (int64_t)STEP * U) {
        A v[U][VE];
        int c[U][VE];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t q0 = base + (int64_t)u * STEP + (int64_t)gl * VE;
            load_vals<V, A>(val + q0, v[u]);
            load_idx<VE>(idx + q0, c[u]);
        }
        A x[U][VE];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int i = 0; i < VE; ++i) {
                const int64_t q = base + (int64_t)u * STEP + (int64_t)gl * VE + i;
                x[u][i] = (q >= s && q < e) ? (A)vec[c[u][i]] : (A)0;
            }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int i = 0; i < VE; ++i) acc += v[u][i] * x[u][i];
    }
    return acc;
}

// Visit the rows of this lane group (grid stride), loading the next row's
// bounds before the current row is processed.
template <int G, typename F>
__device__ __forceinline__ void for_each_row(const int64_t* __restrict__ ptr, int64_t nrows, F&& f) {
    constexpr int GPB = BLOCK / G;
    const int64_t stride = (int64_t)gridDim.x * GPB;
    int64_t row = (int64_t)blockIdx.x * GPB + threadIdx.x / G;
    int64_t s = 0, e = 0;
    if (row < nrows) { s = ptr[row]; e = ptr[row + 1]; }
    while (row < nrows) {
        const int64_t nrow = row + stride;
        int64_t ns = 0, ne = 0;
        if (nrow < nrows) { ns = ptr[nrow]; ne = ptr[nrow + 1]; }
        f(row, s, e);
        row = nrow;
        s = ns;
        e = ne;
    }
}

__device__ __forceinline__ int load_stop(const Ctrl* c) {
    return __hip_atomic_load(&c->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- block reduction + last-block finalize
template <int NV, int NT = BLOCK>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double (*sh)[NT / 64]) {
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
            for (int w = 0; w < NT / 64; ++w) t += sh[q][w];
            v[q] = t;
        }
    }
}

// Stop rules of auxil/check_stoprules.m on the device (thread 0 of the finalize
// block).  ncp_sum is the sum over residual columns of the NCP statistic.
__device__ int evaluate_stop(Ctrl* ctrl, const StopParams sp, int k, double rr, double rMr, double me_num,
                             double rkm1sq, double ncp_sum, double* hist) {
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
    case AIR_STOP_NCP: { // :96-151; dk starts as Inf (check_inputs.m:207-209), max ignores NaN
        const double ncc = ncp_sum / (double)sp.ncp_cols;
        double mx_old = -INFINITY, mx_new = ncc;
        for (int q = 0; q < sp.ncp_smooth; ++q) mx_old = fmax(mx_old, sp.ncp_dk[q]);
        for (int q = 1; q < sp.ncp_smooth; ++q) mx_new = fmax(mx_new, sp.ncp_dk[q]);
        if (mx_old < mx_new || k >= sp.kmax) {
            stop = 1;
            code = (k != sp.kmax) ? 1 : 0;
        } else {
            for (int q = 0; q + 1 < sp.ncp_smooth; ++q) sp.ncp_dk[q] = sp.ncp_dk[q + 1];
            sp.ncp_dk[sp.ncp_smooth - 1] = ncc;
        }
        break;
    }
    default:            // NONE :153-161
        if (k >= sp.kmax) { stop = 1; code = 0; }
        break;
    }
    if (stop) {
        ctrl->finaliter = k;
        ctrl->stoprule = code;
        __hip_atomic_store(&ctrl->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return stop;
}

// Block totals (valid in thread 0 after block_reduce) -> partial[blockIdx.x*NV + q].
// A separate single-block kernel reduces them after the kernel boundary, which
// makes the stores visible without device-scope fences or counters.
template <int NV>
__device__ __forceinline__ void store_block_partials(const double (&v)[NV], double* partial) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) partial[(int64_t)blockIdx.x * NV + q] = v[q];
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

}  // namespace air
