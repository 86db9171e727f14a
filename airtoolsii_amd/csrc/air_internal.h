// air_internal.h -- shared declarations of the MI355X SIRT engine (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/airhip.h"

namespace air {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
struct Err {
    int code;
    std::string msg;
};
#define AIR_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw ::air::Err{AIR_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)
#define AIR_CHECK(cond, code, msg)                                                     \
    do {                                                                               \
        if (!(cond)) throw ::air::Err{(code), (msg)};                                  \
    } while (0)

// ---------------------------------------------------------------- value storage
// Values of A are stored as FP32, BF16 (raw bits), FP16 (raw bits) or FP64.
struct bf16_t { uint16_t u; };
struct f16_t { uint16_t u; };

__host__ __device__ inline float to_f32(float v) { return v; }
__host__ __device__ inline float to_f32(double v) { return (float)v; }
__host__ __device__ inline float to_f32(bf16_t v) { return __builtin_bit_cast(float, (uint32_t)v.u << 16); }
__device__ inline float to_f32(f16_t v) { return (float)__builtin_bit_cast(_Float16, v.u); }
__device__ inline double to_f64(float v) { return (double)v; }
__device__ inline double to_f64(double v) { return v; }
__device__ inline double to_f64(bf16_t v) { return (double)to_f32(v); }
__device__ inline double to_f64(f16_t v) { return (double)to_f32(v); }

// Narrowing from the tracer's FP64 value: FP32 is one RNE rounding; BF16 and
// FP16 round the FP32 value again (RNE).
template <typename V> __device__ inline V from_f64(double d);
template <> __device__ inline float from_f64<float>(double d) { return (float)d; }
template <> __device__ inline double from_f64<double>(double d) { return d; }
template <> __device__ inline bf16_t from_f64<bf16_t>(double d) {
    float f = (float)d;
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7f800000u) == 0x7f800000u) return bf16_t{(uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0))};
    u += 0x7fffu + ((u >> 16) & 1u);
    return bf16_t{(uint16_t)(u >> 16)};
}
template <> __device__ inline f16_t from_f64<f16_t>(double d) {
    _Float16 h = (_Float16)(float)d;
    return f16_t{__builtin_bit_cast(uint16_t, h)};
}

inline size_t dtype_size(int dt) { return dt == AIR_F64 ? 8 : (dt == AIR_F32 ? 4 : 2); }

// Dispatch a functor templated on the value type.
template <typename F>
auto dispatch_dtype(int dt, F&& f) {
    switch (dt) {
    case AIR_F32: return f(float{});
    case AIR_BF16: return f(bf16_t{});
    case AIR_F16: return f(f16_t{});
    case AIR_F64: return f(double{});
    }
    throw Err{AIR_E_ARG, "unknown value dtype"};
}

// ---------------------------------------------------------------- device buffers
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t b) {
        if (b <= bytes && p) return;
        release();
        if (b == 0) b = 256;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            p = nullptr;
            throw Err{AIR_E_NOMEM, "hipMalloc of " + std::to_string(b) + " bytes failed"};
        }
        bytes = b;
    }
    template <typename T> T* as() const { return (T*)p; }
};

// Sparse matrix on the device: compressed rows of A (CSR) or of A' (CSC of A).
struct DevSparse {
    int64_t nrows = 0, ncols = 0, nnz = 0;
    int dtype = AIR_F32;
    DevBuf ptr;   // int64 [nrows+1]
    DevBuf idx;   // int32 [nnz]
    DevBuf val;   // dtype [nnz]
    double avg_len() const { return nrows ? (double)nnz / (double)nrows : 0.0; }
};

// Ray of a line-model tracer: start point (x0,y0) and direction (a,b).
struct RayParam {
    double x0, y0, a, b;
};

struct Geometry {
    int N = 0;
    int edge_check = 0;   // paralleltomo.m:255-256 excludes rays on the top/right edge
    int64_t nrays = 0;
    DevBuf rays;          // RayParam [nrays]
    std::vector<RayParam> host_rays;
};

// ---------------------------------------------------------------- tracer (air_tracer.hip)
void host_rays_paralleltomo(int N, const double* theta, int nA, int p, double d, std::vector<RayParam>& out);
void host_rays_fanlineartomo(int N, const double* theta, int nA, int p, double R, double dw, double sd,
                             std::vector<RayParam>& out);
// Trace all rays into CSR (values narrowed to dtype).  Returns nnz.
void trace_to_csr(const Geometry& g, int dtype, DevSparse& A, hipStream_t s);
// Matrix-free products: y = A u (m) or g = A' v (n, FP64 accumulators).
void mf_forward(const Geometry& g, const float* u, float* y, hipStream_t s);
void mf_adjoint(const Geometry& g, const float* v, double* gacc, hipStream_t s);
// Matrix-free row stats (sum a^2, sum |a|, sum a^2 * w[col]) and column stats.
void mf_row_stats(const Geometry& g, const double* colw, double* sq, double* ab, double* sqw, hipStream_t s);
void mf_col_stats(const Geometry& g, double* colabs, int* colcnt, hipStream_t s);

// ---------------------------------------------------------------- sort/scan (air_sort.hip)
// out[i] = sum_{j<i} in[j] for i in [0, n]; in has n entries (out has n+1).
void scan_exclusive_i64(const int64_t* in, int64_t* out, int64_t n, hipStream_t s);
// CSR -> CSC on the device (rows ascending within each column).
void csr_to_csc(const DevSparse& A, DevSparse& AT, hipStream_t s);

// ---------------------------------------------------------------- kernels (air_sirt.hip)
// Per-launch shape of the compressed-row kernels: G lanes per row, grid blocks.
struct Shape {
    int G = 64;
    int blocks = 2048;
};
Shape default_shape(const DevSparse& S, int num_cu);

void spmv(const DevSparse& S, Shape sh, const float* x, float* y, hipStream_t s);
// row statistics in FP64: sq = sum a^2, ab = sum |a|, sqw = sum a^2 * w[idx] (w may be null), nz = #nonzeros
void row_stats(const DevSparse& S, const double* w, double* sq, double* ab, double* sqw, double* nz, hipStream_t s);

// Control block of one SIRT run on the device.
struct Ctrl {
    int stop;
    int finaliter;
    int stoprule;
    int pad0;
    double lam_cur;      // relaxpar of the last executed iteration
    double rMr;          // r'Mr of the current residual (line search numerator)
    double last_norm;
    unsigned int cnt_rows;
    unsigned int cnt_cols;
};

struct StopParams {
    int type;            // AIR_STOP_*
    int kmax;
    double taudelta;
};

struct SirtBufs {
    const float* b;
    const float* M;      // null = identity
    const float* D;      // null = identity
    float* x;
    float* r;
    float* Mr;
    double* g;           // line search / matrix-free: A' M r accumulators (FP64)
    const float* lbv;    // vector bounds or null
    const float* ubv;
    int lb_kind, ub_kind;
    float lb, ub;
    double* partial;     // per-block partial sums
    double* hist;        // ||r_k||, k = 0..kmax
    Ctrl* ctrl;
};

// r = b - A x; Mr = M r; FP64 reductions and the stop check of iteration k.
void launch_resid(const DevSparse& A, Shape sh, const SirtBufs& B, StopParams sp, int k, hipStream_t s);
// x = P_C(x + lam D (A' Mr)) over the CSC of A.
void launch_update(const DevSparse& AT, Shape sh, const SirtBufs& B, double lam, int k, hipStream_t s);
// line search: g = A' Mr (FP64) with sum D g^2; then lam = r'Mr / sum D g^2 and the update.
void launch_line_update(const DevSparse& AT, Shape sh, const SirtBufs& B, int k, hipStream_t s);
// matrix-free iteration pieces
void launch_mf_resid(const Geometry& g, const SirtBufs& B, StopParams sp, int k, float* Ax, hipStream_t s);
void launch_mf_update(const Geometry& g, int64_t n, const SirtBufs& B, double lam, int line, int k, hipStream_t s);
// small helpers
void vec_f64_to_f32(const double* in, float* out, int64_t n, hipStream_t s);
void vec_f32_to_f64(const float* in, double* out, int64_t n, hipStream_t s);
void vec_mul_f32(const float* a, float* y, int64_t n, hipStream_t s);   // y *= a
void zero_f64(double* p, int64_t n, hipStream_t s);

}  // namespace air
