// air_engine.hip -- the C ABI of libairhip.so: engine handle, operators,
// weights, relaxation parameter and the device-resident SIRT driver.
//
// Host logic mirrors the reference MATLAB drivers:
//   sirt/sirt.m:132-281          driver (stop check at k=0, save list K, info)
//   auxil/check_inputs.m:59-260  validation errors
//   auxil/get_mfun_dfun.m:47-250 M/D diagonals (matrix branch formulas)
//   auxil/calc_relaxpar.m:89-244 rho, default/scalar/line/psi strategies
//   auxil/afun_matrix.m:44-53    the afun product convention
// The heavy work (tracing, products, reductions) runs in HIP kernels on the
// engine's stream; the host only sequences launches.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "air_internal.h"

namespace air {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

struct SirtState {
    bool prepared = false;
    int kmax = 0, k = 0;
    int relax_kind = AIR_RELAX_DEFAULT;
    std::vector<double> lam;       // per-iteration relaxpar (scalar strategies)
    double rho = NAN;
    StopParams sp{};
    int64_t m = 0, n = 0;
    bool hasM = false, hasD = false;
    DevBuf b, x, r, Mr, M, D, g, lbv, ubv, partial, hist, ctrl, Ax;
    std::unique_ptr<NcpState> ncp;   // NCP stop rule (FFT plan and buffers), kept across runs
    LearnBufs learn;                 // learned / greedy relaxation
    DevBuf l_mlp, l_xstar, l_state, l_gpart, l_feat, l_target, l_out, l_err;
    DevBuf lamdev;                   // graph mode: the lambda schedule on the device
    DevBuf ring;                     // deferred norms: block partials of the last iterations
    int defer_first = 1;             // first iteration whose deferred norms are not reduced yet
    SirtBufs bufs{};
    double setup_s = 0, iter_s = 0;
};

}  // namespace air

struct air_engine {
    int device = 0;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    int kind = 0;   // 0 none, 1 stored matrix, 2 matrix-free geometry, 3 a caller's function
    air_afun_fn afun = nullptr;   // kind 3: the caller's products on host FP64 vectors
    void* afun_ctx = nullptr;
    std::vector<double> ext_in, ext_out;
    int64_t m = 0, n = 0;
    air::DevSparse A, AT;
    air::Shape shA, shAT;
    air::Geometry geo;
    air::TiledOp T;            // LDS-tiled layout used by the SIRT loop when ready
    air::SpmmOp S;             // layout of the many-right-hand-side products (built on first use)
    // grow-only device staging of x-like transfers (the FP32 user-order copy):
    // no allocation and no device-synchronising free per call
    air::DevBuf stg32;
    hipEvent_t it_ev[2] = {nullptr, nullptr};   // timing of air_sirt_iterate (created on first use)
    // pinned host staging of FP32 vectors (hipHostMalloc, grow-only) and the
    // event after its last host-to-device copy (the buffer is rewritten only
    // once that copy is done)
    void* pin = nullptr;
    size_t pin_bytes = 0;
    hipEvent_t pin_ev = nullptr;
    bool pin_pending = false;
    bool no_tune = false;      // skip the kernel-shape autotuner (parts of a batch)
    // batch of independent problems (block-diagonal operator without CSR)
    std::vector<int64_t> b_row_off, b_col_off;
    std::vector<double> b_rho, b_M, b_D;
    air::SirtState st;
    // weights and spectral radius of the built-in methods, kept until the operator changes
    std::vector<double> wM[AIR_SART + 1], wD[AIR_SART + 1];
    bool wvalid[AIR_SART + 1] = {};
    air::DevBuf wMdev[AIR_SART + 1], wDdev[AIR_SART + 1];   // FP32 copies in the loop's order
    bool wdev[AIR_SART + 1] = {};
    double wrho[AIR_SART + 1] = {NAN, NAN, NAN, NAN, NAN};
    // graph mode: iterations per captured HIP graph (0 = direct launches) and
    // the instantiated graph of the current prepared run
    int graph_iters = 0;
    hipGraphExec_t gexec = nullptr;
    int64_t graph_launches = 0;   // graphs launched since the engine was created
    // per-kernel timing: events bracket the kernels of every timing-th
    // iteration (0 = off); ev_now says whether the current one is bracketed
    int timing = 0;
    int64_t ev_tick = 0;
    bool ev_now = false;
    std::vector<hipEvent_t> ev;
    int ev_used = 0;
    double t_rows_ms = 0, t_cols_ms = 0, t_red_ms = 0;
    int64_t launches = 0;
};

using namespace air;

namespace {

// Scratch memory of air_apply_device, the entry point that allocates nothing:
// one fixed array per GPU, held until the process ends.  In host code the
// name `scratch` is only a stand-in for the symbol; its device address on each
// GPU comes from hipGetSymbolAddress (taken once per device in air_create, with
// that device current) and only that address is passed to kernels.
static __device__ __align__(256) unsigned char scratch[AIR_APPLY_SCRATCH_BYTES];
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;   // per device; null until the device's first air_create

void init_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    AIR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    AIR_CHECK(p != nullptr, AIR_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = p;
}

void* scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

template <typename F>
int guard(F&& f) {
    try {
        f();
        return AIR_OK;
    } catch (const Err& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return AIR_E_NOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return AIR_E_ARG;
    }
}

void use_device(air_engine* e) { AIR_HIP(hipSetDevice(e->device)); }

void need_operator(air_engine* e) {
    AIR_CHECK(e != nullptr, AIR_E_ARG, "null engine");
    AIR_CHECK(e->kind != 0, AIR_E_STATE, "no operator set: call air_set_matrix_* or air_build_* first");
}

template <typename T>
void upload(DevBuf& d, const T* h, int64_t count, hipStream_t s) {
    d.alloc(sizeof(T) * std::max<int64_t>(count, 1));
    if (count > 0) AIR_HIP(hipMemcpyAsync(d.p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
}

// The engine's pinned staging buffer, at least `bytes` long, free to rewrite.
float* pinned(air_engine* e, size_t bytes) {
    if (e->pin_pending) {
        AIR_HIP(hipEventSynchronize(e->pin_ev));
        e->pin_pending = false;
    }
    if (bytes > e->pin_bytes || !e->pin) {
        if (e->pin) (void)hipHostFree(e->pin);
        e->pin = nullptr;
        e->pin_bytes = 0;
        AIR_HIP(hipHostMalloc(&e->pin, std::max<size_t>(bytes, 4096), hipHostMallocDefault));
        e->pin_bytes = std::max<size_t>(bytes, 4096);
    }
    if (!e->pin_ev) AIR_HIP(hipEventCreateWithFlags(&e->pin_ev, hipEventDisableTiming));
    return (float*)e->pin;
}

// upload FP64 host values as FP32: narrowed on the host (round to nearest
// even, as on the device) into pinned memory and copied asynchronously
// (stream-ordered; the pinned buffer is reused only after the copy)
void upload_f32(air_engine* e, DevBuf& d, const double* h, int64_t count, hipStream_t s) {
    d.alloc(sizeof(float) * std::max<int64_t>(count, 1));
    if (count <= 0) return;
    float* f = pinned(e, sizeof(float) * (size_t)count);
    for (int64_t i = 0; i < count; ++i) f[i] = (float)h[i];
    AIR_HIP(hipMemcpyAsync(d.p, f, sizeof(float) * count, hipMemcpyHostToDevice, s));
    AIR_HIP(hipEventRecord(e->pin_ev, s));
    e->pin_pending = true;
}

// FP32 device values to FP64 host values (copied in FP32 through pinned
// memory, widened on the host)
void download_f64(air_engine* e, const float* d, double* h, int64_t count, hipStream_t s) {
    if (count <= 0) return;
    float* f = pinned(e, sizeof(float) * (size_t)count);
    AIR_HIP(hipMemcpyAsync(f, d, sizeof(float) * count, hipMemcpyDeviceToHost, s));
    AIR_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < count; ++i) h[i] = (double)f[i];
}

template <typename V>
__global__ void k_narrow(const double* in, V* out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = from_f64<V>(in[i]);
}
template <typename V>
__global__ void k_widen(const V* in, double* out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = to_f64(in[i]);
}

void reset_operator(air_engine* e) {
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    e->gexec = nullptr;
    e->A = DevSparse();
    e->AT = DevSparse();
    e->T = TiledOp();
    e->S = SpmmOp();
    e->geo = Geometry();
    e->afun = nullptr;
    e->afun_ctx = nullptr;
    e->kind = 0;
    e->m = e->n = 0;
    e->st.prepared = false;
    e->b_row_off.clear();
    e->b_col_off.clear();
    e->b_rho.clear();
    e->b_M.clear();
    e->b_D.clear();
    for (int q = 0; q <= AIR_SART; ++q) {
        e->wvalid[q] = false;
        e->wdev[q] = false;
        e->wMdev[q].release();
        e->wDdev[q].release();
        e->wrho[q] = NAN;
        e->wM[q].clear();
        e->wD[q].clear();
    }
}

bool tiled_disabled() {
    const char* v = std::getenv("AIR_DISABLE_TILED");
    return v && v[0] && v[0] != '0';
}

// AIR_SPMM=0 runs many right-hand sides on the compressed rows and columns
// (wave per row or column) instead of the LDS-staged chunk kernels.
bool spmm_disabled() {
    const char* v = std::getenv("AIR_SPMM");
    return v && v[0] == '0';
}

// CSC by transpose, then the LDS-tiled layout of the SIRT loop (image side N
// for traced geometries, 0 for a general matrix).
void finish_matrix(air_engine* e, int N) {
    csr_to_csc(e->A, e->AT, e->stream);
    e->shA = default_shape(e->A, e->num_cu);
    e->shAT = default_shape(e->AT, e->num_cu);
    e->kind = 1;
    e->m = e->A.nrows;
    e->n = e->A.ncols;
    e->T = TiledOp();
    e->S = SpmmOp();
    if (!tiled_disabled()) build_tiled(e->A, N, e->T, e->stream, !e->no_tune);
}

bool use_tiled(const air_engine* e) { return e->kind == 1 && e->T.ready; }

// A caller's operator (kind 3): the device vector goes to host FP64, the
// caller's function runs, and its result comes back.
void ext_call(air_engine* e, int trans, const double* in, double* out) {
    const int rc = e->afun(e->afun_ctx, trans, in, out);
    AIR_CHECK(rc == 0, AIR_E_CALLBACK, "the operator function failed (returned " + std::to_string(rc) + ")");
}
void ext_product(air_engine* e, int trans, const float* in) {
    const int64_t nin = trans ? e->m : e->n, nout = trans ? e->n : e->m;
    e->ext_in.resize(std::max<int64_t>(nin, 1));
    e->ext_out.assign(std::max<int64_t>(nout, 1), 0.0);
    download_f64(e, in, e->ext_in.data(), nin, e->stream);
    ext_call(e, trans, e->ext_in.data(), e->ext_out.data());
}
void ext_apply(air_engine* e, int trans, const float* in, float* out) {
    hipStream_t s = e->stream;
    ext_product(e, trans, in);
    const int64_t nout = trans ? e->n : e->m;
    DevBuf tmp;
    upload(tmp, e->ext_out.data(), nout, s);
    vec_f64_to_f32(tmp.as<double>(), out, nout, s);
    AIR_HIP(hipStreamSynchronize(s));
}
void ext_apply_f64(air_engine* e, int trans, const float* in, double* out) {
    hipStream_t s = e->stream;
    ext_product(e, trans, in);
    const int64_t nout = trans ? e->n : e->m;
    if (nout > 0) AIR_HIP(hipMemcpyAsync(out, e->ext_out.data(), 8 * nout, hipMemcpyHostToDevice, s));
    AIR_HIP(hipStreamSynchronize(s));
}

// y = A x (trans=0) or y = A' x (trans=1), device FP32 in / out.
void apply_dev(air_engine* e, int trans, const float* in, float* out) {
    hipStream_t s = e->stream;
    if (e->kind == 3) {
        ext_apply(e, trans, in, out);
        return;
    }
    if (e->kind == 1 && e->A.nnz == 0 && e->T.ready) {     // a batch: only the tiled layout exists
        DevBuf xi;
        xi.alloc(4 * (size_t)std::max<int64_t>(e->T.n_int, 1));
        if (trans) {
            tiled_backward_plain(e->T, in, xi.as<float>(), s);
            permute_to_user(e->T, xi.as<float>(), out, s);
        } else {
            permute_to_internal(e->T, in, xi.as<float>(), s);
            tiled_forward_plain(e->T, xi.as<float>(), out, s);
        }
        AIR_HIP(hipStreamSynchronize(s));
        return;
    }
    if (e->kind == 1) {
        if (trans) spmv(e->AT, e->shAT, in, out, s);
        else spmv(e->A, e->shA, in, out, s);
    } else {
        if (trans) {
            DevBuf acc;
            acc.alloc(sizeof(double) * std::max<int64_t>(e->n, 1));
            zero_f64(acc.as<double>(), e->n, s);
            mf_adjoint(e->geo, in, acc.as<double>(), s);
            vec_f64_to_f32(acc.as<double>(), out, e->n, s);
            AIR_HIP(hipStreamSynchronize(s));
        } else {
            mf_forward(e->geo, in, out, s);
        }
    }
}

// FP64 row / column statistics of the operator.
struct Stats {
    std::vector<double> rsq, rab, rsqw, cab, cnz;
};

// get_mfun_dfun.m's function-handle branches, with the caller's products on
// unit vectors (host FP64, as the handle is called):
//   Cimmino, DROP (:68-73, :150-155): rsq_i = norm(A' e_i)^2
//   CAV (:108-121): s_j = 1 / nnz(A e_j), rsqw_i = s' (A' e_i).^2
//   DROP (:179-184): cnz_j = nnz(A e_j)
//   SART (:208, :229): rab = |A 1|, cab = |A' 1|
void ext_stats(air_engine* e, int method, Stats& S) {
    const int64_t m = e->m, n = e->n;
    std::vector<double> u, v;
    if (method == AIR_CIMMINO || method == AIR_DROP || method == AIR_CAV) {
        std::vector<double> sj(n, 0.0);
        if (method != AIR_CIMMINO) {
            u.assign(n, 0.0);
            v.assign(std::max<int64_t>(m, 1), 0.0);
            for (int64_t j = 0; j < n; ++j) {
                u[j] = 1.0;
                ext_call(e, 0, u.data(), v.data());
                u[j] = 0.0;
                int64_t c = 0;
                for (int64_t i = 0; i < m; ++i) c += v[i] != 0.0;
                S.cnz[j] = (double)c;
                sj[j] = 1.0 / (double)c;
            }
        }
        {
            u.assign(m, 0.0);
            v.assign(std::max<int64_t>(n, 1), 0.0);
            for (int64_t i = 0; i < m; ++i) {
                u[i] = 1.0;
                ext_call(e, 1, u.data(), v.data());
                u[i] = 0.0;
                if (method == AIR_CAV) {
                    double a = 0.0;
                    for (int64_t j = 0; j < n; ++j) a += sj[j] * (v[j] * v[j]);
                    S.rsqw[i] = a;
                } else {
                    double a = 0.0;
                    for (int64_t j = 0; j < n; ++j) a += v[j] * v[j];
                    const double nv = std::sqrt(a);
                    S.rsq[i] = nv * nv;
                }
            }
        }
    } else if (method == AIR_SART) {
        u.assign(n, 1.0);
        v.assign(std::max<int64_t>(m, 1), 0.0);
        ext_call(e, 0, u.data(), v.data());
        for (int64_t i = 0; i < m; ++i) S.rab[i] = std::fabs(v[i]);
        u.assign(m, 1.0);
        v.assign(std::max<int64_t>(n, 1), 0.0);
        ext_call(e, 1, u.data(), v.data());
        for (int64_t j = 0; j < n; ++j) S.cab[j] = std::fabs(v[j]);
    }
}

void operator_stats(air_engine* e, int method, Stats& S) {
    hipStream_t s = e->stream;
    const int64_t m = e->m, n = e->n;
    DevBuf rsq, rab, rsqw, cab, cnz, w;
    rsq.alloc(8 * std::max<int64_t>(m, 1));
    rab.alloc(8 * std::max<int64_t>(m, 1));
    rsqw.alloc(8 * std::max<int64_t>(m, 1));
    cab.alloc(8 * std::max<int64_t>(n, 1));
    cnz.alloc(8 * std::max<int64_t>(n, 1));
    S.rsq.assign(m, 0); S.rab.assign(m, 0); S.rsqw.assign(m, 0); S.cab.assign(n, 0); S.cnz.assign(n, 0);
    bool need_cols = method == AIR_CAV || method == AIR_DROP || method == AIR_SART;
    if (e->kind == 3) {
        ext_stats(e, method, S);
        return;
    }
    if (e->kind == 1) {
        if (need_cols) row_stats(e->AT, nullptr, nullptr, cab.as<double>(), nullptr, cnz.as<double>(), s);
        if (method == AIR_CAV) {
            row_stats(e->A, cnz.as<double>(), rsq.as<double>(), rab.as<double>(), rsqw.as<double>(), nullptr, s);
        } else {
            row_stats(e->A, nullptr, rsq.as<double>(), rab.as<double>(), nullptr, nullptr, s);
        }
    } else {
        if (need_cols) {
            DevBuf cnt;
            cnt.alloc(4 * std::max<int64_t>(n, 1));
            AIR_HIP(hipMemsetAsync(cnt.p, 0, 4 * n, s));
            zero_f64(cab.as<double>(), n, s);
            mf_col_stats(e->geo, cab.as<double>(), cnt.as<int>(), s);
            std::vector<int> hc(n);
            AIR_HIP(hipMemcpyAsync(hc.data(), cnt.p, 4 * n, hipMemcpyDeviceToHost, s));
            AIR_HIP(hipStreamSynchronize(s));
            for (int64_t j = 0; j < n; ++j) S.cnz[j] = hc[j];
            upload(cnz, S.cnz.data(), n, s);
        }
        mf_row_stats(e->geo, method == AIR_CAV ? cnz.as<double>() : nullptr, rsq.as<double>(), rab.as<double>(),
                     method == AIR_CAV ? rsqw.as<double>() : nullptr, s);
    }
    AIR_HIP(hipMemcpyAsync(S.rsq.data(), rsq.p, 8 * m, hipMemcpyDeviceToHost, s));
    AIR_HIP(hipMemcpyAsync(S.rab.data(), rab.p, 8 * m, hipMemcpyDeviceToHost, s));
    AIR_HIP(hipMemcpyAsync(S.rsqw.data(), rsqw.p, 8 * m, hipMemcpyDeviceToHost, s));
    AIR_HIP(hipMemcpyAsync(S.cab.data(), cab.p, 8 * n, hipMemcpyDeviceToHost, s));
    if (e->kind == 1) AIR_HIP(hipMemcpyAsync(S.cnz.data(), cnz.p, 8 * n, hipMemcpyDeviceToHost, s));
    AIR_HIP(hipStreamSynchronize(s));
}

inline double inf_to_zero(double v) { return std::isinf(v) ? 0.0 : v; }

// get_mfun_dfun.m matrix-branch formulas; identity is returned as all ones.
void compute_weights_uncached(air_engine* e, int method, std::vector<double>& M, std::vector<double>& D);

// get_mfun_dfun.m weights depend on A alone: computed once per method and operator
void compute_weights(air_engine* e, int method, std::vector<double>& M, std::vector<double>& D) {
    AIR_CHECK(method >= AIR_LANDWEBER && method <= AIR_SART, AIR_E_ARG, "SIRT method not defined.");
    if (!e->wvalid[method]) {
        compute_weights_uncached(e, method, e->wM[method], e->wD[method]);
        e->wvalid[method] = true;
    }
    M = e->wM[method];
    D = e->wD[method];
}

void compute_weights_uncached(air_engine* e, int method, std::vector<double>& M, std::vector<double>& D) {
    const int64_t m = e->m, n = e->n;
    M.assign(m, 1.0);
    D.assign(n, 1.0);
    if (method == AIR_LANDWEBER) return;                       // :50-54
    AIR_CHECK(method >= AIR_CIMMINO && method <= AIR_SART, AIR_E_ARG, "SIRT method not defined.");
    Stats S;
    operator_stats(e, method, S);
    switch (method) {
    case AIR_CIMMINO:                                          // :56-88
        for (int64_t i = 0; i < m; ++i) M[i] = inf_to_zero((1.0 / (double)m) * (1.0 / S.rsq[i]));
        break;
    case AIR_CAV:                                              // :90-138
        for (int64_t i = 0; i < m; ++i) M[i] = inf_to_zero(1.0 / S.rsqw[i]);
        break;
    case AIR_DROP:                                             // :140-198
        for (int64_t i = 0; i < m; ++i) M[i] = inf_to_zero(1.0 / S.rsq[i]);
        for (int64_t j = 0; j < n; ++j) D[j] = inf_to_zero(1.0 / S.cnz[j]);
        break;
    case AIR_SART:                                             // :200-241
        for (int64_t i = 0; i < m; ++i) M[i] = inf_to_zero(1.0 / S.rab[i]);
        for (int64_t j = 0; j < n; ++j) D[j] = inf_to_zero(1.0 / S.cab[j]);
        break;
    }
}

// Largest eigenvalue of the symmetric T^(1/2) A' M A T^(1/2) (= rho(T A' M A)),
// Lanczos with a Sturm-bisection solve of the tridiagonal matrix.
double lanczos_rho(air_engine* e, const std::vector<double>& M, const std::vector<double>& D) {
    hipStream_t s = e->stream;
    const int64_t m = e->m, n = e->n;
    DevBuf dM, dv, dy, dg;
    upload_f32(e, dM, M.data(), m, s);
    dv.alloc(4 * std::max<int64_t>(n, 1));
    dy.alloc(4 * std::max<int64_t>(m, 1));
    dg.alloc(4 * std::max<int64_t>(n, 1));
    std::vector<double> sD(n), v(n, 1.0 / std::sqrt((double)n)), vprev(n, 0.0), w(n);
    std::vector<float> hf(n);
    for (int64_t j = 0; j < n; ++j) sD[j] = std::sqrt(D[j]);
    auto op = [&](const std::vector<double>& in, std::vector<double>& out) {
        for (int64_t j = 0; j < n; ++j) hf[j] = (float)(sD[j] * in[j]);
        AIR_HIP(hipMemcpyAsync(dv.p, hf.data(), 4 * n, hipMemcpyHostToDevice, s));
        apply_dev(e, 0, dv.as<float>(), dy.as<float>());
        vec_mul_f32(dM.as<float>(), dy.as<float>(), m, s);
        apply_dev(e, 1, dy.as<float>(), dg.as<float>());
        AIR_HIP(hipMemcpyAsync(hf.data(), dg.p, 4 * n, hipMemcpyDeviceToHost, s));
        AIR_HIP(hipStreamSynchronize(s));
        for (int64_t j = 0; j < n; ++j) out[j] = sD[j] * (double)hf[j];
    };
    std::vector<double> al, be;
    double beta = 0, theta_prev = 0, theta = 0;
    int stable = 0;
    auto top_eig = [&](int k) {
        double lo = 1e300, hi = -1e300;
        for (int i = 0; i < k; ++i) {
            double r = (i > 0 ? std::fabs(be[i - 1]) : 0) + (i + 1 < k ? std::fabs(be[i]) : 0);
            lo = std::min(lo, al[i] - r);
            hi = std::max(hi, al[i] + r);
        }
        auto count_below = [&](double x) {
            int c = 0;
            double q = al[0] - x;
            if (q < 0) c++;
            for (int i = 1; i < k; ++i) {
                if (q == 0) q = 1e-300;
                q = al[i] - x - be[i - 1] * be[i - 1] / q;
                if (q < 0) c++;
            }
            return c;
        };
        for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(1.0, std::fabs(hi)); ++it) {
            double mid = 0.5 * (lo + hi);
            if (count_below(mid) >= k) hi = mid; else lo = mid;
        }
        return 0.5 * (lo + hi);
    };
    for (int it = 0; it < 300; ++it) {
        op(v, w);
        double a = 0;
        for (int64_t j = 0; j < n; ++j) { w[j] -= beta * vprev[j]; }
        for (int64_t j = 0; j < n; ++j) a += w[j] * v[j];
        for (int64_t j = 0; j < n; ++j) w[j] -= a * v[j];
        double bnew = 0;
        for (int64_t j = 0; j < n; ++j) bnew += w[j] * w[j];
        bnew = std::sqrt(bnew);
        al.push_back(a);
        theta = top_eig((int)al.size());
        if (it > 5 && std::fabs(theta - theta_prev) <= 1e-10 * std::fabs(theta)) {
            if (++stable >= 3) break;
        } else stable = 0;
        theta_prev = theta;
        if (bnew <= 1e-14 * std::fabs(theta) || bnew == 0) break;
        be.push_back(bnew);
        vprev = v;
        for (int64_t j = 0; j < n; ++j) v[j] = w[j] / bnew;
        beta = bnew;
    }
    return theta;
}

double calczeta_one(int k) {
    // calc_relaxpar.m:217-244: Newton from 1, 6 steps; Horner on C = [0 2k-1 -ones(1,k-1)]
    auto fung = [k](double x, double& f, double& df) {
        f = 0.0;
        df = 0.0;
        for (int i = 1; i <= k; ++i) {
            double c = (i == 1) ? (double)(2 * k - 1) : -1.0;
            df = df * x + f;
            f = f * x + c;
        }
    };
    double x = 1.0, f, df;
    fung(x, f, df);
    double h = f / df;
    for (int it = 0; it < 6; ++it) {
        x = x - h;
        fung(x, f, df);
        h = f / df;
    }
    return x;
}

// calc_relaxpar.m:105-171: per-iteration relaxpar for scalar strategies.
void relax_schedule(int kind, double relaxpar, const double* vec, double rho, int kmax, std::vector<double>& lam) {
    lam.assign(std::max(kmax, 0), 0.0);
    switch (kind) {
    case AIR_RELAX_DEFAULT:
        for (auto& v : lam) v = 1.9 / rho;
        break;
    case AIR_RELAX_SCALAR:
        for (auto& v : lam) v = relaxpar;
        break;
    case AIR_RELAX_VECTOR:
        AIR_CHECK(vec != nullptr || kmax == 0, AIR_E_ARG, "relax_vec must hold kmax values");
        for (int i = 0; i < kmax; ++i) lam[i] = vec[i];
        break;
    case AIR_RELAX_PSI1: case AIR_RELAX_PSI1MOD: case AIR_RELAX_PSI2: case AIR_RELAX_PSI2MOD: {
        AIR_CHECK(kmax > 2, AIR_E_ARG, "psi relaxation strategies need kmax > 2 (calczeta of 2:kmax-1)");
        bool psi2 = kind == AIR_RELAX_PSI2 || kind == AIR_RELAX_PSI2MOD;
        double nu = kind == AIR_RELAX_PSI1MOD ? 2.0 : (kind == AIR_RELAX_PSI2MOD ? 1.5 : 1.0);
        lam[0] = std::sqrt(2.0) / rho;
        lam[1] = std::sqrt(2.0) / rho;
        for (int k = 2; k < kmax; ++k) {
            double z = calczeta_one(k);
            double v = psi2 ? (nu * 2) * (1 - z) / ((1 - std::pow(z, (double)k)) * (1 - std::pow(z, (double)k)))
                            : (nu * 2) * (1 - z);
            if (!psi2 && nu == 1.0) v = 2 * (1 - z);
            lam[k] = v / rho;
        }
        break;
    }
    default:
        break;
    }
}

// relaxation chosen on the device each iteration (line search, network, greedy target)
bool device_lambda(int kind) { return kind == AIR_RELAX_LINE || kind == AIR_RELAX_GREEDY; }

static constexpr int EV_PER_ITER = 4;

void timing_begin(air_engine* e, int niter) {
    e->ev_used = 0;
    if (!e->timing) return;
    size_t need = EV_PER_ITER * (size_t)std::max(niter, 1);
    while (e->ev.size() < need) {
        hipEvent_t v;
        AIR_HIP(hipEventCreate(&v));
        e->ev.push_back(v);
    }
}

void record(air_engine* e, int slot) {
    if (!e->ev_now) return;
    AIR_HIP(hipEventRecord(e->ev[(size_t)e->ev_used * EV_PER_ITER + slot], e->stream));
}

hipEvent_t mid_event(air_engine* e) {
    return e->ev_now ? e->ev[(size_t)e->ev_used * EV_PER_ITER + 2] : nullptr;
}

// one SIRT iteration k (sirt.m:213-243): the A'(M r) + update pass, then the
// residual pass; events 0..3 bracket them when kernel timing is on.
// Graph mode applies to the tiled layout with a host lambda schedule and the
// DP/ME/none stop rules: every kernel of such an iteration takes its
// iteration number and lambda from the device (ctrl->kcur, ctrl->lam_cur), so
// one captured run of iterations can be replayed for any k.
bool graph_mode(const air_engine* e) {
    const SirtState& st = e->st;
    return e->graph_iters > 0 && !e->timing && use_tiled(e) && !device_lambda(st.relax_kind) &&
           st.relax_kind != AIR_RELAX_LEARNED && st.relax_kind != AIR_RELAX_GREEDY && st.sp.type != AIR_STOP_NCP &&
           (!st.prepared || st.bufs.lamv != nullptr);
}

void drop_graph(air_engine* e) {
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    e->gexec = nullptr;
}

// deferred norms: reduce the pending iterations' block partials into hist and
// run the stop check of the last one
void flush_deferred(air_engine* e) {
    SirtState& st = e->st;
    if (!st.bufs.ring || st.defer_first > st.k) return;
    launch_finalize_deferred(st.bufs, (int)tiled_reduce_grid(e->T.m), st.sp, st.defer_first, st.k, e->stream);
    st.defer_first = st.k + 1;
}

void enqueue_iteration(air_engine* e, int k, bool devk = false) {
    SirtState& st = e->st;
    hipStream_t s = e->stream;
    if (devk) {                     // k and lambda from the device (graph mode)
        launch_tiled_update(e->T, st.bufs, NAN, 0, -1, s);
        launch_tiled_resid(e->T, st.bufs, st.sp, -1, s, nullptr);
        return;
    }
    const bool line = device_lambda(st.relax_kind);
    double lam = line ? 0.0 : st.lam[k - 1];
    e->ev_now = e->timing > 0 && e->ev_tick++ % e->timing == 0;
    record(e, 0);
    if (st.relax_kind == AIR_RELAX_LEARNED) {
        if (!fused_select(st.bufs)) launch_learned_select(st.bufs, k, s);
        lam = NAN;                  // the update kernels read the network's lambda_k
    }
    if (use_tiled(e)) {
        launch_tiled_update(e->T, st.bufs, lam, line ? 1 : 0, k, s);
        record(e, 1);
        launch_tiled_resid(e->T, st.bufs, st.sp, k, s, mid_event(e));
        record(e, 3);
        if (e->ev_now) e->ev_used++;
        if (st.bufs.ring && k - st.defer_first + 1 >= st.bufs.ring_slots) flush_deferred(e);
        return;
    } else if (e->kind == 1) {
        if (line) launch_line_update(e->AT, e->shAT, st.bufs, k, s);
        else launch_update(e->AT, e->shAT, st.bufs, lam, k, s);
        record(e, 1);
        launch_resid(e->A, e->shA, st.bufs, st.sp, k, s);
        record(e, 2);
    } else if (e->kind == 3) {
        ext_apply_f64(e, 1, st.bufs.Mr, st.bufs.g);
        launch_g_update(e->n, st.bufs, lam, line ? 1 : 0, k, s);
        record(e, 1);
        ext_apply(e, 0, st.bufs.x, st.Ax.as<float>());
        launch_vec_resid(e->m, st.Ax.as<float>(), st.bufs, st.sp, k, s);
        record(e, 2);
    } else {
        launch_mf_update(e->geo, e->n, st.bufs, lam, line ? 1 : 0, k, s);
        record(e, 1);
        launch_mf_resid(e->geo, st.bufs, st.sp, k, st.Ax.as<float>(), s);
        record(e, 2);
    }
    record(e, 3);
    if (e->ev_now) e->ev_used++;
}

// Enqueue iterations st.k+1 .. st.k+count.  In graph mode whole runs of
// graph_iters iterations are one hipGraphLaunch of a graph captured once per
// prepared run; the rest are launched directly in the same device-k form.
void enqueue_iterations(air_engine* e, int count) {
    SirtState& st = e->st;
    hipStream_t s = e->stream;
    if (!graph_mode(e)) {
        for (int i = 0; i < count; ++i) enqueue_iteration(e, ++st.k);
        flush_deferred(e);
        return;
    }
    const int G = e->graph_iters;
    AIR_CHECK(st.bufs.lamv != nullptr, AIR_E_STATE, "graph mode changed after air_sirt_prepare");
    while (count > 0) {
        if (count >= G) {
            if (!e->gexec) {
                hipGraph_t g = nullptr;
                AIR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                for (int i = 0; i < G; ++i) enqueue_iteration(e, 0, true);
                AIR_HIP(hipStreamEndCapture(s, &g));
                const hipError_t ie = hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                AIR_HIP(ie);
            }
            AIR_HIP(hipGraphLaunch(e->gexec, s));
            e->graph_launches++;
            st.k += G;
            count -= G;
        } else {
            enqueue_iteration(e, 0, true);
            ++st.k;
            --count;
        }
    }
}

// after the stream has drained: sum the per-kernel event intervals
void timing_end(air_engine* e) {
    if (!e->timing) return;
    for (int i = 0; i < e->ev_used; ++i) {
        float a = 0, b = 0, c = 0;
        hipEvent_t* v = &e->ev[(size_t)EV_PER_ITER * i];
        AIR_HIP(hipEventElapsedTime(&a, v[0], v[1]));
        AIR_HIP(hipEventElapsedTime(&b, v[1], v[2]));
        AIR_HIP(hipEventElapsedTime(&c, v[2], v[3]));
        e->t_cols_ms += a;
        e->t_rows_ms += b;
        e->t_red_ms += c;
        e->launches += 1;
    }
    e->ev_used = 0;
}

// length of the x-like vectors (x, D, bounds) of the SIRT loop
int64_t xlen(const air_engine* e) { return use_tiled(e) ? e->T.n_int : e->n; }

// user-order FP64 n-vector -> FP32 vector of the loop's column order
void upload_xlike(air_engine* e, DevBuf& d, const double* h, hipStream_t s) {
    if (!use_tiled(e)) {
        upload_f32(e, d, h, e->n, s);
        return;
    }
    upload_f32(e, e->stg32, h, e->n, s);
    d.alloc(4 * (size_t)e->T.n_int);
    permute_to_internal(e->T, e->stg32.as<float>(), d.as<float>(), s);
}

// FP32 vector of the loop's column order -> user-order FP64 host vector
void download_xlike(air_engine* e, const float* d, double* h, hipStream_t s) {
    if (!use_tiled(e)) {
        download_f64(e, d, h, e->n, s);
        return;
    }
    e->stg32.alloc(4 * (size_t)std::max<int64_t>(e->n, 1));
    permute_to_user(e->T, d, e->stg32.as<float>(), s);
    download_f64(e, e->stg32.as<float>(), h, e->n, s);
}

void sirt_prepare(air_engine* e, const air_sirt_options* o, const double* b, const double* x0, int kmax) {
    need_operator(e);
    AIR_CHECK(o != nullptr && b != nullptr, AIR_E_ARG, "Too few input arguments.");
    AIR_CHECK(kmax >= 0, AIR_E_ARG, "Max no. of iterations must ALWAYS be specified.");
    AIR_CHECK(o->stoprule >= 0 && o->stoprule <= 3, AIR_E_ARG, "The chosen stopping rule is not valid.");
    AIR_CHECK(o->relax_kind >= AIR_RELAX_DEFAULT && o->relax_kind <= AIR_RELAX_GREEDY, AIR_E_ARG,
              "The chosen relaxation strategy is not valid.");
    AIR_CHECK(o->relax_kind != AIR_RELAX_LEARNED || o->mlp_params, AIR_E_ARG,
              "the learned relaxation needs the network parameters (mlp_params)");
    AIR_CHECK(o->relax_kind != AIR_RELAX_GREEDY || o->x_true, AIR_E_ARG,
              "the greedy relaxation needs the exact solution (x_true)");
    if (o->stoprule == AIR_STOP_NCP) {
        // check_stoprules.m:99-120: res_dims = [len, ncols] (2-D) or [m, 0] (1-D)
        const int64_t len = o->res_dims[0], cols = o->res_dims[1] > 0 ? o->res_dims[1] : 1;
        AIR_CHECK(len > 0 && len * cols == e->m, AIR_E_ARG,
                  "options.stoprule.res_dims must give the dimensions of the residual (prod(res_dims) == m).");
        AIR_CHECK(o->ncp_smooth >= 1 && o->ncp_smooth <= 4096, AIR_E_ARG, "ncp_smooth must be a positive integer.");
    }
    if (o->stoprule == AIR_STOP_DP || o->stoprule == AIR_STOP_ME)
        AIR_CHECK(!std::isnan(o->taudelta), AIR_E_ARG,
                  "For stopping rules DP and ME a scalar taudelta must be given in options.");
    auto t0 = std::chrono::steady_clock::now();
    use_device(e);
    hipStream_t s = e->stream;
    SirtState& st = e->st;
    st.prepared = false;
    st.m = e->m;
    st.n = e->n;
    const int64_t m = st.m, n = st.n;
    // M, D (get_mfun_dfun.m); a built-in method's weights are the engine's
    // cached ones, used in place
    std::vector<double> Mc, Dc;
    const std::vector<double>* pM = &Mc;
    const std::vector<double>* pD = &Dc;
    if (o->method == AIR_CUSTOM) {
        Mc.assign(m, 1.0);
        Dc.assign(n, 1.0);
        if (o->M) std::copy(o->M, o->M + m, Mc.begin());
        if (o->D) std::copy(o->D, o->D + n, Dc.begin());
        st.hasM = o->M != nullptr;
        st.hasD = o->D != nullptr;
    } else {
        AIR_CHECK(o->method >= AIR_LANDWEBER && o->method <= AIR_SART, AIR_E_ARG, "SIRT method not defined.");
        if (!e->wvalid[o->method]) {
            compute_weights_uncached(e, o->method, e->wM[o->method], e->wD[o->method]);
            e->wvalid[o->method] = true;
        }
        pM = &e->wM[o->method];
        pD = &e->wD[o->method];
        st.hasM = o->method != AIR_LANDWEBER;
        st.hasD = o->method == AIR_DROP || o->method == AIR_SART;
    }
    const std::vector<double>& M = *pM;
    const std::vector<double>& D = *pD;
    // rho and relaxpar (sirt.m:164-169, calc_relaxpar.m:89-171)
    double rho = o->rho;
    if (o->method == AIR_SART) rho = 1.0;
    // calc_relaxpar.m:91-103 computes rho whenever it is not given
    if (std::isnan(rho)) {
        if (o->method != AIR_CUSTOM && !std::isnan(e->wrho[o->method])) {
            rho = e->wrho[o->method];                 // same operator and weights as an earlier call
        } else {
            rho = lanczos_rho(e, M, D);
            if (o->method != AIR_CUSTOM) e->wrho[o->method] = rho;
        }
    }
    st.rho = rho;
    st.relax_kind = o->relax_kind;
    relax_schedule(o->relax_kind, o->relaxpar, o->relax_vec, rho, kmax, st.lam);
    // device vectors
    upload_f32(e, st.b, b, m, s);
    if (x0) upload_xlike(e, st.x, x0, s);
    else {                                            // x0 = 0 (check_inputs.m:93)
        st.x.alloc(4 * (size_t)std::max<int64_t>(xlen(e), 1));
        AIR_HIP(hipMemsetAsync(st.x.p, 0, 4 * (size_t)std::max<int64_t>(xlen(e), 1), s));
    }
    const float *dMp = nullptr, *dDp = nullptr;
    if (o->method == AIR_CUSTOM) {
        if (st.hasM) upload_f32(e, st.M, M.data(), m, s);
        if (st.hasD) upload_xlike(e, st.D, D.data(), s);
        dMp = st.hasM ? st.M.as<float>() : nullptr;
        dDp = st.hasD ? st.D.as<float>() : nullptr;
    } else {
        // device copies of a built-in method's weights are kept with the operator
        const int q = o->method;
        if (!e->wdev[q]) {
            if (st.hasM) upload_f32(e, e->wMdev[q], M.data(), m, s);
            if (st.hasD) upload_xlike(e, e->wDdev[q], D.data(), s);
            e->wdev[q] = true;
        }
        dMp = st.hasM ? e->wMdev[q].as<float>() : nullptr;
        dDp = st.hasD ? e->wDdev[q].as<float>() : nullptr;
    }
    const int64_t nx = xlen(e);
    st.r.alloc(4 * std::max<int64_t>(m, 1));
    st.Mr.alloc(4 * std::max<int64_t>(m, 1));
    st.g.alloc(8 * std::max<int64_t>(nx, 1));
    st.Ax.alloc(4 * std::max<int64_t>(m, 1));
    // Deferred norms (the tiled loop with the stop rule 'none' and a host
    // lambda schedule, outside graph mode): no iteration reads its own norms,
    // so the row pass leaves them for a later reduction instead of a finalize
    // kernel per iteration.
    const bool defer = use_tiled(e) && o->stoprule == AIR_STOP_NONE && !device_lambda(o->relax_kind) &&
                       o->relax_kind != AIR_RELAX_LEARNED && o->relax_kind != AIR_RELAX_GREEDY &&
                       e->graph_iters == 0;
    constexpr int RING_SLOTS = 32;
    // row-pass reductions: 4 doubles per block; column-pass ones: 1 per block
    const int64_t row_blocks = use_tiled(e) ? tiled_reduce_grid(m) : 2048;
    const int64_t row_slots = 4 * std::max<int64_t>(2048, row_blocks);
    const int64_t col_slots = std::max<int64_t>(2048, e->T.ntiles_b);
    st.partial.alloc(8 * (size_t)(row_slots + col_slots));
    st.hist.alloc(8 * (size_t)(kmax + 1));
    st.ctrl.alloc(sizeof(Ctrl));
    AIR_HIP(hipMemsetAsync(st.ctrl.p, 0, sizeof(Ctrl), s));
    AIR_HIP(hipMemsetAsync(st.r.p, 0, 4 * std::max<int64_t>(m, 1), s));
    SirtBufs& B = st.bufs;
    B = SirtBufs{};
    B.b = st.b.as<float>();
    B.M = dMp;
    B.D = dDp;
    B.x = st.x.as<float>();
    B.r = st.r.as<float>();
    B.Mr = st.Mr.as<float>();
    B.g = st.g.as<double>();
    B.lb_kind = o->lb_kind;
    B.ub_kind = o->ub_kind;
    B.lb = (float)o->lb;
    B.ub = (float)o->ub;
    if (o->lb_kind == AIR_BOUND_VECTOR) {
        AIR_CHECK(o->lb_vec, AIR_E_ARG, "lbound vector missing");
        upload_xlike(e, st.lbv, o->lb_vec, s);
        B.lbv = st.lbv.as<float>();
    }
    if (o->ub_kind == AIR_BOUND_VECTOR) {
        AIR_CHECK(o->ub_vec, AIR_E_ARG, "ubound vector missing");
        upload_xlike(e, st.ubv, o->ub_vec, s);
        B.ubv = st.ubv.as<float>();
    }
    B.partial = st.partial.as<double>();
    B.partial_cols = st.partial.as<double>() + row_slots;
    B.hist = st.hist.as<double>();
    B.ctrl = st.ctrl.as<Ctrl>();
    if (defer) {
        st.ring.alloc(8 * 4 * (size_t)row_blocks * RING_SLOTS);
        B.ring = st.ring.as<double>();
        B.ring_slots = RING_SLOTS;
    }
    st.defer_first = 1;
    if (o->relax_kind == AIR_RELAX_LEARNED || o->relax_kind == AIR_RELAX_GREEDY) {
        LearnBufs& L = st.learn;
        L = LearnBufs{};
        L.mode = o->relax_kind;
        L.kmax = kmax;
        L.nx = nx;
        L.rho = (float)rho;
        L.noise = std::isnan(o->noise_norm) ? 0.0 : o->noise_norm;
        const size_t kk = (size_t)std::max(kmax, 1);
        st.l_state.alloc(8 * LEARN_STATE);
        st.l_feat.alloc(4 * kk * LEARN_NF);
        st.l_target.alloc(4 * kk);
        st.l_out.alloc(4 * kk * 2);
        st.l_err.alloc(8 * kk);
        AIR_HIP(hipMemsetAsync(st.l_state.p, 0, 8 * LEARN_STATE, s));
        L.state = st.l_state.as<double>();
        L.feat = st.l_feat.as<float>();
        L.target = st.l_target.as<float>();
        L.out = st.l_out.as<float>();
        L.err = st.l_err.as<double>();
        if (o->mlp_params) {       // learned run, or a greedy run along the network's path
            st.l_mlp.alloc(4 * (size_t)MLP_PARAMS);
            AIR_HIP(hipMemcpyAsync(st.l_mlp.p, o->mlp_params, 4 * (size_t)MLP_PARAMS, hipMemcpyHostToDevice, s));
            L.mlp = st.l_mlp.as<float>();
        }
        if (o->x_true) {   // greedy target, or the error record of a learned run
            upload_xlike(e, st.l_xstar, o->x_true, s);
            L.xstar = st.l_xstar.as<float>();
            L.gblocks = (int)std::min<int64_t>((nx + 255) / 256, 1024);   // 256-thread blocks
            st.l_gpart.alloc(8 * 3 * (size_t)std::max(L.gblocks, 1));
            L.gpart = st.l_gpart.as<double>();
        }
        B.learn = &st.learn;
    }
    st.sp = StopParams{};
    st.sp.type = o->stoprule;
    st.sp.kmax = kmax;
    st.sp.taudelta = o->taudelta;
    if (o->stoprule == AIR_STOP_NCP) {
        if (!st.ncp) st.ncp.reset(new NcpState());
        const int len = o->res_dims[0], cols = o->res_dims[1] > 0 ? o->res_dims[1] : 1;
        ncp_prepare(*st.ncp, m, len, cols, o->ncp_smooth, s);
        st.sp.ncp_stat = st.ncp->colstat.as<double>();
        st.sp.ncp_dk = st.ncp->dk.as<double>();
        st.sp.ncp_cols = cols;
        st.sp.ncp_smooth = o->ncp_smooth;
        B.ncp = st.ncp.get();
    }
    st.kmax = kmax;
    st.k = 0;
    drop_graph(e);
    if (graph_mode(e) && kmax > 0) {
        st.lamdev.alloc(8 * (size_t)kmax);
        AIR_HIP(hipMemcpyAsync(st.lamdev.p, st.lam.data(), 8 * (size_t)kmax, hipMemcpyHostToDevice, s));
        B.lamv = st.lamdev.as<double>();
    }
    // r0 = b - A x0 and the k = 0 check (sirt.m:153-162)
    if (use_tiled(e)) launch_tiled_resid(e->T, B, st.sp, 0, s, nullptr);
    else if (e->kind == 1) launch_resid(e->A, e->shA, B, st.sp, 0, s);
    else if (e->kind == 3) {
        ext_apply(e, 0, B.x, st.Ax.as<float>());
        launch_vec_resid(e->m, st.Ax.as<float>(), B, st.sp, 0, s);
    } else launch_mf_resid(e->geo, B, st.sp, 0, st.Ax.as<float>(), s);
    // no synchronisation here: the first iterations queue behind r0 on the
    // stream, host staging is guarded by pin_ev, and pageable host data
    // (the lambda schedule, network parameters) is consumed by the copy call
    st.prepared = true;
    st.setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    st.iter_s = 0;
}

// Iterations until the residual norm reaches target, from its decay between
// two checks (h0 -> h1 over dk iterations), plus one; 4..maxc.
int dp_chunk(double h0, double h1, int dk, double target, int maxc) {
    if (!(h1 > target)) return 4;
    if (!(h0 > h1) || dk <= 0 || !(target > 0)) return maxc;
    const double rate = std::log(h1 / h0) / dk;          // < 0
    const double need = std::log(target / h1) / rate;   // > 0
    if (!(need < (double)maxc)) return maxc;
    return std::min(maxc, std::max(4, (int)std::ceil(need) + 1));
}

int read_stop(air_engine* e) {
    Ctrl c;
    AIR_HIP(hipMemcpyAsync(&c, e->st.ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, e->stream));
    AIR_HIP(hipStreamSynchronize(e->stream));
    return c.stop;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* air_last_error(void) { return g_last_error.c_str(); }

int air_create(int device, air_engine** out) {
    return guard([&] {
        AIR_CHECK(out != nullptr, AIR_E_ARG, "null output handle");
        int count = 0;
        AIR_HIP(hipGetDeviceCount(&count));
        AIR_CHECK(device >= 0 && device < count, AIR_E_ARG, "device index out of range");
        std::unique_ptr<air_engine> e(new air_engine());
        e->device = device;
        AIR_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        AIR_HIP(hipGetDeviceProperties(&prop, device));
        e->num_cu = prop.multiProcessorCount;
        AIR_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
        init_scratch(device);
        *out = e.release();
    });
}

void air_destroy(air_engine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    for (auto& v : e->ev) (void)hipEventDestroy(v);
    for (auto& v : e->it_ev)
        if (v) (void)hipEventDestroy(v);
    if (e->pin_ev) (void)hipEventDestroy(e->pin_ev);
    if (e->pin) (void)hipHostFree(e->pin);
    reset_operator(e);
    e->st = SirtState();
    (void)hipStreamDestroy(e->stream);
    delete e;
}

int air_synchronize(air_engine* e) {
    return guard([&] {
        AIR_CHECK(e, AIR_E_ARG, "null engine");
        AIR_HIP(hipStreamSynchronize(e->stream));
    });
}

// Upload a host CSR into e->A (values narrowed on the device); no transpose.
static void upload_csr(air_engine* e, int64_t m, int64_t n, const int64_t* rowptr, const int32_t* colind,
                       const double* val, int dtype) {
    AIR_CHECK(e && rowptr && m >= 0 && n >= 0, AIR_E_ARG, "invalid CSR arguments");
    AIR_CHECK(n <= INT32_MAX, AIR_E_ARG, "n must fit int32 column indices");
    use_device(e);
    reset_operator(e);
    hipStream_t s = e->stream;
    int64_t nnz = rowptr[m];
    AIR_CHECK(rowptr[0] == 0 && nnz >= 0, AIR_E_ARG, "rowptr must start at 0");
    for (int64_t i = 0; i < m; ++i) AIR_CHECK(rowptr[i + 1] >= rowptr[i], AIR_E_ARG, "rowptr must be nondecreasing");
    for (int64_t q = 0; q < nnz; ++q) AIR_CHECK(colind[q] >= 0 && colind[q] < n, AIR_E_ARG, "column index out of range");
    DevSparse& A = e->A;
    A.nrows = m;
    A.ncols = n;
    A.nnz = nnz;
    A.dtype = dtype;
    (void)dtype_size(dtype);
    upload(A.ptr, rowptr, m + 1, s);
    alloc_sparse_arrays(A, nnz, dtype, s);
    if (nnz > 0) AIR_HIP(hipMemcpyAsync(A.idx.p, colind, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s));
    DevBuf v64;
    upload(v64, val, nnz, s);
    dispatch_dtype(dtype, [&](auto v) {
        using V = decltype(v);
        if (nnz > 0) k_narrow<V><<<4096, 256, 0, s>>>(v64.as<double>(), A.val.as<V>(), nnz);
        return 0;
    });
    AIR_HIP(hipGetLastError());
    AIR_HIP(hipStreamSynchronize(s));
}

// A caller's matrix with a square number of columns is tiled as an N x N
// image (the column-major pixel order of the reference's test problems), so a
// traced matrix handed over as plain CSR/CSC gets the image tiles and, where
// its rows are straight rays, the step-coded forward; any other matrix is
// just as valid in that column order.
static int image_side(int64_t n) {
    int64_t r = (int64_t)std::llround(std::sqrt((double)n));
    while (r * r > n) --r;
    while ((r + 1) * (r + 1) <= n) ++r;
    return (r * r == n && r > 0 && r < (1 << 20)) ? (int)r : 0;
}

int air_set_afun(air_engine* e, int64_t m, int64_t n, air_afun_fn fn, void* ctx) {
    return guard([&] {
        AIR_CHECK(e != nullptr && fn != nullptr, AIR_E_ARG, "null engine or operator function");
        AIR_CHECK(m >= 0 && n >= 0, AIR_E_ARG, "invalid operator size");
        use_device(e);
        reset_operator(e);
        e->afun = fn;
        e->afun_ctx = ctx;
        e->m = m;
        e->n = n;
        e->kind = 3;
    });
}

int air_set_matrix_csr(air_engine* e, int64_t m, int64_t n, const int64_t* rowptr, const int32_t* colind,
                       const double* val, int dtype) {
    return guard([&] {
        upload_csr(e, m, n, rowptr, colind, val, dtype);
        finish_matrix(e, image_side(n));
    });
}

int air_set_matrix_csc(air_engine* e, int64_t m, int64_t n, const int64_t* colptr, const int64_t* rowind,
                       const double* val, int dtype) {
    return guard([&] {
        AIR_CHECK(e && colptr && m >= 0 && n >= 0, AIR_E_ARG, "invalid CSC arguments");
        AIR_CHECK(m <= INT32_MAX, AIR_E_ARG, "m must fit int32 row indices");
        int64_t nnz = colptr[n];
        std::vector<int32_t> ri(nnz);
        for (int64_t q = 0; q < nnz; ++q) {
            AIR_CHECK(rowind[q] >= 0 && rowind[q] < m, AIR_E_ARG, "row index out of range");
            ri[q] = (int32_t)rowind[q];
        }
        // The CSC of A is the CSR of A': upload it as such, transpose it into
        // the CSR of A, and swap the two roles.
        upload_csr(e, n, m, colptr, ri.data(), val, dtype);
        csr_to_csc(e->A, e->AT, e->stream);
        std::swap(e->A, e->AT);
        e->shA = default_shape(e->A, e->num_cu);
        e->shAT = default_shape(e->AT, e->num_cu);
        e->kind = 1;
        e->m = e->A.nrows;
        e->n = e->A.ncols;
        e->T = TiledOp();
        e->S = SpmmOp();
        if (!tiled_disabled()) build_tiled(e->A, image_side(e->n), e->T, e->stream, !e->no_tune);
    });
}

static int build_geometry(air_engine* e, int N, int64_t nrays, int edge, int matrix_free, int dtype) {
    hipStream_t s = e->stream;
    e->geo.N = N;
    e->geo.edge_check = edge;
    e->geo.nrays = nrays;
    upload(e->geo.rays, e->geo.host_rays.data(), nrays, s);
    AIR_HIP(hipStreamSynchronize(s));
    if (matrix_free) {
        mf_build_tiles(e->geo, s);
        e->kind = 2;
        e->m = nrays;
        e->n = (int64_t)N * N;
    } else {
        trace_to_csr(e->geo, dtype, e->A, s);
        finish_matrix(e, N);
    }
    return AIR_OK;
}

int air_build_paralleltomo(air_engine* e, int32_t N, const double* theta, int32_t nA, int32_t p, double d,
                           int matrix_free, int dtype) {
    return guard([&] {
        AIR_CHECK(e && theta && N > 0 && nA >= 0 && p > 0, AIR_E_ARG, "invalid paralleltomo arguments");
        AIR_CHECK((int64_t)N * N <= INT32_MAX, AIR_E_ARG, "N too large for int32 column indices");
        use_device(e);
        reset_operator(e);
        (void)dtype_size(dtype);
        host_rays_paralleltomo(N, theta, nA, p, d, e->geo.host_rays);
        build_geometry(e, N, (int64_t)nA * p, 1, matrix_free, dtype);
        if (matrix_free) mf_prepare_parallel(e->geo, nA, p, d, e->stream);
    });
}

int air_build_fanlineartomo(air_engine* e, int32_t N, const double* theta, int32_t nA, int32_t p, double R,
                            double dw, double sd, int matrix_free, int dtype) {
    return guard([&] {
        AIR_CHECK(e && theta && N > 0 && nA >= 0 && p > 0, AIR_E_ARG, "invalid fanlineartomo arguments");
        AIR_CHECK((int64_t)N * N <= INT32_MAX, AIR_E_ARG, "N too large for int32 column indices");
        // fanlineartomo.m:75-90,112-114
        double sdN = sd * N, dwN = dw * N, RN = R * N;
        AIR_CHECK(!(RN < std::sqrt(2.0) / 2 * N), AIR_E_ARG, "R must be greater than half squareroot 2");
        use_device(e);
        reset_operator(e);
        (void)dtype_size(dtype);
        host_rays_fanlineartomo(N, theta, nA, p, RN, dwN, sdN, e->geo.host_rays);
        build_geometry(e, N, (int64_t)nA * p, 0, matrix_free, dtype);
    });
}

int air_build_fancurvedtomo(air_engine* e, int32_t N, const double* theta, int32_t nA, int32_t p, double R,
                            double d, int matrix_free, int dtype) {
    return guard([&] {
        AIR_CHECK(e && theta && N > 0 && nA >= 0 && p > 0, AIR_E_ARG, "invalid fancurvedtomo arguments");
        AIR_CHECK((int64_t)N * N <= INT32_MAX, AIR_E_ARG, "N too large for int32 column indices");
        // fancurvedtomo.m:104-113
        AIR_CHECK(!(R < std::sqrt(2.0) / 2), AIR_E_ARG, "R must be greater than half squareroot 2");
        AIR_CHECK(!(d < 0 || d > 180), AIR_E_ARG, "The angle of the source must be in the interval [0 180]");
        use_device(e);
        reset_operator(e);
        (void)dtype_size(dtype);
        host_rays_fancurvedtomo(N, theta, nA, p, R * N, d, e->geo.host_rays);
        build_geometry(e, N, (int64_t)nA * p, 0, matrix_free, dtype);
    });
}

int air_build_seismictomo(air_engine* e, int32_t N, int32_t s, int32_t p, int matrix_free, int dtype) {
    return guard([&] {
        AIR_CHECK(e && N > 0 && s > 0 && p > 0, AIR_E_ARG, "invalid seismictomo arguments");
        AIR_CHECK((int64_t)N * N <= INT32_MAX, AIR_E_ARG, "N too large for int32 column indices");
        use_device(e);
        reset_operator(e);
        (void)dtype_size(dtype);
        host_rays_seismictomo(N, s, p, e->geo.host_rays);
        build_geometry(e, N, (int64_t)s * p, 0, matrix_free, dtype);
    });
}

int air_build_sphericaltomo(air_engine* e, int32_t N, const double* angles, int32_t nA, int32_t numCircles,
                            int matrix_free, int dtype) {
    return guard([&] {
        AIR_CHECK(e && angles && N > 0 && nA >= 0 && numCircles > 0, AIR_E_ARG, "invalid sphericaltomo arguments");
        use_device(e);
        reset_operator(e);
        (void)dtype_size(dtype);
        if (matrix_free) {     // sphericaltomo.m:90-93: rows re-formed on every product
            circles_mf_prepare(e->geo, N, angles, nA, numCircles, e->stream);
            e->kind = 2;
            e->m = (int64_t)nA * numCircles;
            e->n = (int64_t)N * N;
            return;
        }
        circles_to_csr(N, angles, nA, numCircles, dtype, e->A, e->stream);
        finish_matrix(e, N);
    });
}

int air_build_seismicwavetomo(air_engine* e, int32_t N, int32_t s, int32_t p, double omega, int matrix_free,
                              int dtype) {
    return guard([&] {
        AIR_CHECK(e && N > 0 && s > 0 && p > 0, AIR_E_ARG, "invalid seismicwavetomo arguments");
        use_device(e);
        reset_operator(e);
        (void)dtype_size(dtype);
        if (matrix_free) {                             // seismicwavetomo.m:94-99: kernels re-evaluated per product
            fresnel_mf_prepare(e->geo, N, s, p, omega, e->stream);
            e->kind = 2;
            e->m = (int64_t)s * p;
            e->n = (int64_t)N * N;
            return;
        }
        fresnel_to_csr(N, s, p, omega, dtype, e->A, e->stream);
        finish_matrix(e, N);
    });
}

int air_size(air_engine* e, int64_t* m, int64_t* n, int64_t* nnz) {
    return guard([&] {
        need_operator(e);
        if (m) *m = e->m;
        if (n) *n = e->n;
        if (nnz) *nnz = e->kind == 1 ? (e->A.nnz ? e->A.nnz : e->T.nnz) : -1;
    });
}

int air_get_csr(air_engine* e, int64_t* rowptr, int32_t* colind, double* val) {
    return guard([&] {
        need_operator(e);
        AIR_CHECK(e->kind == 1, AIR_E_UNSUPPORTED, "matrix-free operator has no stored CSR");
        use_device(e);
        hipStream_t s = e->stream;
        const DevSparse& A = e->A;
        if (rowptr) AIR_HIP(hipMemcpyAsync(rowptr, A.ptr.p, 8 * (A.nrows + 1), hipMemcpyDeviceToHost, s));
        if (colind && A.nnz) AIR_HIP(hipMemcpyAsync(colind, A.idx.p, 4 * A.nnz, hipMemcpyDeviceToHost, s));
        if (val && A.nnz) {
            DevBuf tmp;
            tmp.alloc(8 * A.nnz);
            dispatch_dtype(A.dtype, [&](auto v) {
                using V = decltype(v);
                k_widen<V><<<4096, 256, 0, s>>>(A.val.as<V>(), tmp.as<double>(), A.nnz);
                return 0;
            });
            AIR_HIP(hipGetLastError());
            AIR_HIP(hipMemcpyAsync(val, tmp.p, 8 * A.nnz, hipMemcpyDeviceToHost, s));
            AIR_HIP(hipStreamSynchronize(s));
        }
        AIR_HIP(hipStreamSynchronize(s));
    });
}

int air_apply(air_engine* e, int transp, const double* in, double* out) {
    return guard([&] {
        need_operator(e);
        AIR_CHECK(in && out, AIR_E_ARG, "null vector");
        AIR_CHECK(transp == 0 || transp == 1, AIR_E_ARG, "transp_flag must be 'size', 'notransp' or 'transp'");
        use_device(e);
        hipStream_t s = e->stream;
        int64_t nin = transp ? e->m : e->n, nout = transp ? e->n : e->m;
        DevBuf din, dout;
        upload_f32(e, din, in, nin, s);
        dout.alloc(4 * std::max<int64_t>(nout, 1));
        apply_dev(e, transp, din.as<float>(), dout.as<float>());
        download_f64(e, dout.as<float>(), out, nout, s);
    });
}

int air_apply_device(air_engine* e, int transp, const float* in, float* out) {
    return guard([&] {
        need_operator(e);
        AIR_CHECK(in && out, AIR_E_ARG, "null vector");
        AIR_CHECK(transp == 0 || transp == 1, AIR_E_ARG, "transp_flag must be 'size', 'notransp' or 'transp'");
        use_device(e);
        hipStream_t s = e->stream;
        AIR_CHECK(e->kind != 3, AIR_E_UNSUPPORTED,
                  "air_apply_device: a caller's operator function works on host vectors");
        void* sc = scratch_ptr(e->device);
        AIR_CHECK(sc != nullptr, AIR_E_STATE, "scratch memory of this device not initialised");
        const size_t cap = AIR_APPLY_SCRATCH_BYTES;
        if (e->kind == 1 && e->A.nnz == 0 && e->T.ready) {
            // a batch (tiled layout only): the internal-order vector lives in scratch
            const size_t need = 4 * (size_t)std::max<int64_t>(e->T.n_int, 1);
            AIR_CHECK(need <= cap, AIR_E_ARG,
                      "air_apply_device: the operator needs " + std::to_string(need) +
                          " bytes of scratch, more than the fixed " + std::to_string(cap));
            float* xi = (float*)sc;
            if (transp) {
                tiled_backward_plain(e->T, in, xi, s);
                permute_to_user(e->T, xi, out, s);
            } else {
                permute_to_internal(e->T, in, xi, s);
                tiled_forward_plain(e->T, xi, out, s);
            }
        } else if (e->kind == 1) {
            if (transp) spmv(e->AT, e->shAT, in, out, s);
            else spmv(e->A, e->shA, in, out, s);
        } else if (transp) {
            // matrix-free adjoint: FP64 accumulators in scratch, cleared by a fill kernel
            const size_t need = 8 * (size_t)std::max<int64_t>(e->n, 1);
            AIR_CHECK(need <= cap, AIR_E_ARG,
                      "air_apply_device: the operator needs " + std::to_string(need) +
                          " bytes of scratch, more than the fixed " + std::to_string(cap));
            double* acc = (double*)sc;
            zero_f64(acc, e->n, s);
            mf_adjoint(e->geo, in, acc, s);
            vec_f64_to_f32(acc, out, e->n, s);
        } else {
            mf_forward(e->geo, in, out, s);
        }
        AIR_HIP(hipGetLastError());
    });
}

int air_fbp(air_engine* e, const double* b, int32_t na, int filter, double* x) {
    return guard([&] {
        need_operator(e);
        AIR_CHECK(b && x && na > 0, AIR_E_ARG, "Input A, b, theta must me specified");
        AIR_CHECK(filter >= AIR_FBP_NONE && filter <= AIR_FBP_HANN, AIR_E_ARG, "Iinvalid filter");
        use_device(e);
        hipStream_t s = e->stream;
        DevBuf din, dout;
        dout.alloc(4 * std::max<int64_t>(e->n, 1));
        if (filter == AIR_FBP_NONE) {                     // fbp.m:37
            upload_f32(e, din, b, e->m, s);
            apply_dev(e, 1, din.as<float>(), dout.as<float>());
            download_f64(e, dout.as<float>(), x, e->n, s);
            return;
        }
        // fbp.m:40-43: reshape(b, nr, na) needs m to be a multiple of na
        AIR_CHECK(e->m % na == 0, AIR_E_ARG, "fbp: size(A,1) is not a multiple of length(theta)");
        const int64_t nr = e->m / na;
        din.alloc(4 * std::max<int64_t>(e->m, 1));
        fbp_filter_sinogram(b, nr, na, filter, din.as<float>(), s);
        apply_dev(e, 1, din.as<float>(), dout.as<float>());
        download_f64(e, dout.as<float>(), x, e->n, s);
        for (int64_t j = 0; j < e->n; ++j) x[j] = x[j] * M_PI / (2.0 * na);   // fbp.m:80
    });
}

int air_weights(air_engine* e, int method, double* M_out, double* D_out) {
    return guard([&] {
        need_operator(e);
        use_device(e);
        std::vector<double> M, D;
        compute_weights(e, method, M, D);
        if (M_out) std::copy(M.begin(), M.end(), M_out);
        if (D_out) std::copy(D.begin(), D.end(), D_out);
    });
}

int air_spectral_radius(air_engine* e, int method, const double* Mc, const double* Dc, double* rho_out) {
    return guard([&] {
        need_operator(e);
        AIR_CHECK(rho_out, AIR_E_ARG, "null output");
        use_device(e);
        std::vector<double> M, D;
        if (method == AIR_CUSTOM) {
            M.assign(e->m, 1.0);
            D.assign(e->n, 1.0);
            if (Mc) std::copy(Mc, Mc + e->m, M.begin());
            if (Dc) std::copy(Dc, Dc + e->n, D.begin());
        } else {
            compute_weights(e, method, M, D);
        }
        *rho_out = lanczos_rho(e, M, D);
    });
}

int air_calczeta(int32_t kmax, double* z_out) {
    return guard([&] {
        AIR_CHECK(kmax > 2, AIR_E_ARG, "The polynomial is not defined for k < 2");
        AIR_CHECK(z_out, AIR_E_ARG, "null output");
        for (int k = 2; k < kmax; ++k) z_out[k - 2] = calczeta_one(k);
    });
}

int air_sirt_prepare(air_engine* e, const air_sirt_options* opts, const double* b, const double* x0, int32_t kmax) {
    return guard([&] { sirt_prepare(e, opts, b, x0, kmax); });
}

int air_sirt_iterate(air_engine* e, int32_t niter, double* gpu_ms) {
    return guard([&] {
        need_operator(e);
        AIR_CHECK(e->st.prepared, AIR_E_STATE, "air_sirt_prepare must be called first");
        use_device(e);
        SirtState& st = e->st;
        if (!e->it_ev[0]) {
            AIR_HIP(hipEventCreate(&e->it_ev[0]));
            AIR_HIP(hipEventCreate(&e->it_ev[1]));
        }
        hipEvent_t a = e->it_ev[0], b = e->it_ev[1];
        auto t0 = std::chrono::steady_clock::now();
        timing_begin(e, niter);
        AIR_HIP(hipEventRecord(a, e->stream));
        // with a rule that can stop early, check the stop flag every 32 iterations
        // instead of enqueueing iterations that would only return.  Under the
        // discrepancy principle the next check comes where the residual's decay
        // over the last chunk reaches tau*delta (4 to 32 iterations on), after a
        // first chunk of 16, so few iterations are launched past the stop; a
        // prediction within 4 of the iterations left takes them all.
        const bool early = st.sp.type != AIR_STOP_NONE || st.relax_kind == AIR_RELAX_LEARNED;
        const int chunk = graph_mode(e) ? e->graph_iters : (e->kind == 3 ? 1 : 32);
        const bool dp = early && st.sp.type == AIR_STOP_DP && !graph_mode(e) && e->kind != 3;
        int left = std::min(niter, st.kmax - st.k);
        int k_ref = st.k;                    // hist[k_ref] is written once the queued work is done
        int c = early ? std::min(left, dp ? 16 : chunk) : left;
        while (left > 0) {
            enqueue_iterations(e, c);
            left -= c;
            if (!early || left <= 0) break;
            Ctrl cc;
            double href = 0.0;
            AIR_HIP(hipMemcpyAsync(&cc, st.ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, e->stream));
            if (dp) AIR_HIP(hipMemcpyAsync(&href, st.hist.as<double>() + k_ref, 8, hipMemcpyDeviceToHost, e->stream));
            AIR_HIP(hipStreamSynchronize(e->stream));
            if (cc.stop) break;
            int next = chunk;
            if (dp) next = dp_chunk(href, cc.last_norm, st.k - k_ref, st.sp.taudelta, chunk);
            k_ref = st.k;
            c = left - next < 4 ? left : next;
        }
        AIR_HIP(hipEventRecord(b, e->stream));
        AIR_HIP(hipEventSynchronize(b));
        timing_end(e);
        float ms = 0;
        AIR_HIP(hipEventElapsedTime(&ms, a, b));
        st.iter_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (gpu_ms) *gpu_ms = ms;
    });
}

int air_sirt_fetch(air_engine* e, double* x_out, double* resnorm, air_sirt_info* info) {
    return guard([&] {
        need_operator(e);
        AIR_CHECK(e->st.prepared, AIR_E_STATE, "air_sirt_prepare must be called first");
        use_device(e);
        SirtState& st = e->st;
        hipStream_t s = e->stream;
        // x (whatever the stop), the control block and the norm history are
        // queued together and waited for once
        Ctrl c;
        AIR_HIP(hipMemcpyAsync(&c, st.ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
        std::vector<double> hist;
        if (resnorm) {
            hist.resize((size_t)st.kmax + 1);
            AIR_HIP(hipMemcpyAsync(hist.data(), st.hist.p, 8 * hist.size(), hipMemcpyDeviceToHost, s));
        }
        if (x_out) download_xlike(e, st.x.as<float>(), x_out, s);
        AIR_HIP(hipStreamSynchronize(s));
        int fin = c.stop ? c.finaliter : st.k;
        if (resnorm) std::copy(hist.begin(), hist.begin() + (fin + 1), resnorm);
        if (info) {
            info->stoprule = c.stop ? c.stoprule : 0;
            info->finaliter = fin;
            info->relaxpar = (device_lambda(st.relax_kind) || st.relax_kind == AIR_RELAX_LEARNED)
                                 ? c.lam_cur : (fin > 0 ? st.lam[fin - 1] : NAN);
            info->rho = st.rho;
            info->nsaved = 0;
            info->setup_time = st.setup_s;
            info->iter_time = st.iter_s;
            info->timetaken = st.setup_s + st.iter_s;
        }
    });
}

int air_sirt_records(air_engine* e, float* feat, float* target, float* out, double* err) {
    return guard([&] {
        AIR_CHECK(e != nullptr, AIR_E_ARG, "null engine");
        SirtState& st = e->st;
        AIR_CHECK(st.prepared && st.bufs.learn, AIR_E_STATE,
                  "no learned or greedy relaxation run to read records from");
        use_device(e);
        hipStream_t s = e->stream;
        Ctrl c;
        AIR_HIP(hipMemcpyAsync(&c, st.ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
        AIR_HIP(hipStreamSynchronize(s));
        const size_t rows = (size_t)(c.stop ? c.finaliter : st.k);
        // a run stopped by the network at state k-1 has computed row k-1 as well
        const size_t have = std::min<size_t>(rows + (c.stop && c.stoprule == AIR_STOP_LEARNED ? 1 : 0),
                                             (size_t)st.learn.kmax);
        if (have) {
            if (feat) AIR_HIP(hipMemcpyAsync(feat, st.l_feat.p, 4 * have * LEARN_NF, hipMemcpyDeviceToHost, s));
            if (target) AIR_HIP(hipMemcpyAsync(target, st.l_target.p, 4 * have, hipMemcpyDeviceToHost, s));
            if (out) AIR_HIP(hipMemcpyAsync(out, st.l_out.p, 8 * have, hipMemcpyDeviceToHost, s));
            if (err) AIR_HIP(hipMemcpyAsync(err, st.l_err.p, 8 * have, hipMemcpyDeviceToHost, s));
        }
        AIR_HIP(hipStreamSynchronize(s));
    });
}

int air_sirt(air_engine* e, const air_sirt_options* opts, const double* b, const double* x0, const int32_t* K,
             int32_t nK, double* X_out, int32_t* itersaved, int32_t* n_itersaved, double* resnorm,
             air_sirt_info* info) {
    return guard([&] {
        need_operator(e);
        AIR_CHECK(K != nullptr && nK > 0, AIR_E_ARG, "Max no. of iterations must ALWAYS be specified.");
        for (int i = 1; i < nK; ++i) AIR_CHECK(K[i] > K[i - 1], AIR_E_ARG, "Elements of K must be strictly increasing.");
        const int kmax = K[nK - 1];
        auto t0 = std::chrono::steady_clock::now();
        sirt_prepare(e, opts, b, x0, kmax);
        SirtState& st = e->st;
        hipStream_t s = e->stream;
        const int64_t n = st.n;
        // snapshot slots for K(l) < kmax (sirt.m:247-250)
        int nsnap = 0;
        while (nsnap < nK && K[nsnap] < kmax) nsnap++;
        const int64_t nx = xlen(e);
        DevBuf snaps;
        if (nsnap) snaps.alloc(4 * (size_t)nx * nsnap);
        int l = 0;
        const int chunk = graph_mode(e) ? e->graph_iters : (e->kind == 3 ? 1 : 32);
        bool stopped = read_stop(e) != 0;
        auto ti = std::chrono::steady_clock::now();
        while (!stopped && st.k < kmax) {
            int upto = std::min(kmax, st.k + chunk);
            timing_begin(e, upto - st.k);
            while (st.k < upto) {
                const int seg = (l < nsnap && K[l] <= upto ? K[l] : upto) - st.k;   // up to the next saved k
                enqueue_iterations(e, seg);
                if (l < nsnap && st.k == K[l]) {
                    AIR_HIP(hipMemcpyAsync(snaps.as<float>() + (size_t)l * nx, st.x.p, 4 * nx, hipMemcpyDeviceToDevice, s));
                    l++;
                }
            }
            stopped = read_stop(e) != 0;
            timing_end(e);
        }
        st.iter_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - ti).count();
        air_sirt_info inf{};
        int rc = air_sirt_fetch(e, nullptr, resnorm, &inf);
        if (rc != AIR_OK) throw Err{rc, g_last_error};
        const int fin = inf.finaliter;
        int ns = 0;
        for (int q = 0; q < nK; ++q)
            if (K[q] < fin) {
                if (itersaved) itersaved[ns] = K[q];
                ns++;
            }
        int nsaved = fin > 0 ? ns + 1 : 0;
        if (itersaved) itersaved[ns] = fin;
        if (n_itersaved) *n_itersaved = ns + 1;
        if (X_out) {
            for (int q = 0; q < ns; ++q) download_xlike(e, snaps.as<float>() + (size_t)q * nx, X_out + (size_t)q * n, s);
            if (fin > 0) download_xlike(e, st.x.as<float>(), X_out + (size_t)ns * n, s);
        }
        inf.nsaved = nsaved;
        inf.timetaken = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (info) *info = inf;
    });
}

static void need_stored(air_engine* e, const char* what) {
    AIR_CHECK(e->kind == 1 && e->A.nnz >= 0 && e->A.ptr.p && e->AT.ptr.p, AIR_E_UNSUPPORTED,
              std::string(what) + " needs a stored matrix (CSR and CSC); matrix-free and batched operators are "
                                  "not supported");
}

int air_art(air_engine* e, int method, const int32_t* order, int64_t norder, const air_row_options* opts,
            const double* b, const double* x0, const int32_t* K, int32_t nK, double* X_out, int32_t* itersaved,
            int32_t* n_itersaved, double* resnorm, air_sirt_info* info) {
    return guard([&] {
        need_operator(e);
        need_stored(e, "art");
        use_device(e);
        row_action_run(e->A, e->AT, false, method, order, norder, opts, b, x0, K, nK, X_out, itersaved, n_itersaved,
                       resnorm, info, e->stream);
    });
}

int air_cart(air_engine* e, const int32_t* order, int64_t norder, const air_row_options* opts, const double* b,
             const double* x0, const int32_t* K, int32_t nK, double* X_out, int32_t* itersaved,
             int32_t* n_itersaved, double* resnorm, air_sirt_info* info) {
    return guard([&] {
        need_operator(e);
        need_stored(e, "cart");
        use_device(e);
        row_action_run(e->A, e->AT, true, 0, order, norder, opts, b, x0, K, nK, X_out, itersaved, n_itersaved,
                       resnorm, info, e->stream);
    });
}

int air_set_kernel_timing(air_engine* e, int on) {
    return guard([&] {
        AIR_CHECK(e, AIR_E_ARG, "null engine");
        AIR_CHECK(on >= 0, AIR_E_ARG, "kernel timing takes 0 (off) or a sampling stride >= 1");
        e->timing = on;
        e->ev_tick = 0;
        e->ev_now = false;
        e->t_rows_ms = e->t_cols_ms = e->t_red_ms = 0;
        e->launches = 0;
    });
}

int air_set_graph(air_engine* e, int32_t iters_per_graph) {
    return guard([&] {
        AIR_CHECK(e, AIR_E_ARG, "null engine");
        AIR_CHECK(iters_per_graph >= 0 && iters_per_graph <= 4096, AIR_E_ARG, "iterations per graph must be 0..4096");
        use_device(e);
        drop_graph(e);
        e->graph_iters = iters_per_graph;
        e->st.prepared = false;       // the next run is prepared for the new launch mode
    });
}

int air_graph_launches(air_engine* e, int64_t* n) {
    return guard([&] {
        AIR_CHECK(e && n, AIR_E_ARG, "null argument");
        *n = e->graph_launches;
    });
}

int air_get_kernel_timing(air_engine* e, double* cols_ms, double* rows_ms, double* reduce_ms, int64_t* launches) {
    return guard([&] {
        AIR_CHECK(e, AIR_E_ARG, "null engine");
        if (cols_ms) *cols_ms = e->t_cols_ms;
        if (rows_ms) *rows_ms = e->t_rows_ms;
        if (reduce_ms) *reduce_ms = e->t_red_ms;
        if (launches) *launches = e->launches;
    });
}

int air_layout_info(air_engine* e, air_layout* out) {
    return guard([&] {
        need_operator(e);
        AIR_CHECK(out, AIR_E_ARG, "null output");
        air_layout L{};
        L.tiled = use_tiled(e) ? 1 : 0;
        L.m = e->m;
        L.n = e->n;
        L.nnz = e->kind == 1 ? (e->A.nnz ? e->A.nnz : e->T.nnz) : -1;
        L.value_bytes = e->kind == 1 ? (int)dtype_size(e->A.nnz ? e->A.dtype : e->T.dtype) : 4;
        if (L.tiled) {
            const TiledOp& T = e->T;
            L.n_internal = T.n_int;
            L.pieces = T.npieces;
            L.runs = T.nruns;
            L.tiles_fwd = T.ntiles_f;
            L.tiles_bwd = T.ntiles_b;
            L.staged_rows = T.staged_rows;
            L.lds_bytes = (int64_t)T.lds_b;
            tiled_shapes(T, &L.fwd_lanes, &L.fwd_loads, &L.bwd_lanes, &L.bwd_loads);
            L.fwd_stream = tiled_fwd_stream(T) ? 1 : 0;
            L.fwd_pieces_per_block = L.fwd_stream ? T.wave_target : T.ppb;
            L.fwd_entries = T.nfe;
            L.bwd_entries = T.nbe;
            for (int q = 0; q < 64; ++q) L.tune_fwd_us[q] = 1000.f * T.tune_fwd_ms[q];
            for (int q = 0; q < 16; ++q) L.tune_bwd_us[q] = 1000.f * T.tune_bwd_ms[q];
            L.fwd_step = tiled_fwd_step(T) ? 1 : 0;
            L.fwd_step_ok = T.fstep_ok ? 1 : 0;
            L.bwd_step = tiled_bwd_step(T) ? 1 : 0;
            L.bwd_step_ok = T.bstep_ok ? 1 : 0;
            L.bwd_step_entries = T.nbse;
            L.fwd_waves_per_block = T.fwd_wpb;
            L.bwd_step_base_bits = T.bstep_ok ? T.bstep_bb : 0;
            L.bwd_step_lds_bytes = T.bstep_ok ? (int64_t)T.lds_bs : 0;
        }
        *out = L;
    });
}


// ---- learned relaxation network
int air_mlp_forward(air_engine* e, const float* params, const float* X, int64_t n, float* out) {
    return guard([&] {
        AIR_CHECK(e && params && X && out && n >= 0, AIR_E_ARG, "invalid MLP arguments");
        use_device(e);
        hipStream_t s = e->stream;
        DevBuf W, Xd, O;
        W.alloc(4 * (size_t)MLP_PARAMS);
        Xd.alloc(4 * (size_t)std::max<int64_t>(n, 1) * MLP_IN);
        O.alloc(4 * (size_t)std::max<int64_t>(n, 1) * MLP_OUT);
        AIR_HIP(hipMemcpyAsync(W.p, params, 4 * (size_t)MLP_PARAMS, hipMemcpyHostToDevice, s));
        AIR_HIP(hipMemcpyAsync(Xd.p, X, 4 * (size_t)n * MLP_IN, hipMemcpyHostToDevice, s));
        mlp_forward(W.as<float>(), Xd.as<float>(), n, O.as<float>(), s);
        AIR_HIP(hipMemcpyAsync(out, O.p, 4 * (size_t)n * MLP_OUT, hipMemcpyDeviceToHost, s));
        AIR_HIP(hipStreamSynchronize(s));
    });
}

int air_mlp_grad(air_engine* e, const float* params, const float* X, const float* T, int64_t n,
                 double stop_weight, float* grad, double* loss) {
    return guard([&] {
        AIR_CHECK(e && params && X && T && grad && n > 0 && n < (int64_t)1 << 31, AIR_E_ARG,
                  "invalid MLP arguments");
        use_device(e);
        hipStream_t s = e->stream;
        const unsigned nblk = mlp_grad_blocks(n);
        DevBuf W, Xd, Td, part, lp, G;
        W.alloc(4 * (size_t)MLP_PARAMS);
        Xd.alloc(4 * (size_t)n * MLP_IN);
        Td.alloc(4 * (size_t)n * MLP_OUT);
        part.alloc(4 * (size_t)nblk * MLP_PARAMS);
        lp.alloc(8 * (size_t)nblk);
        G.alloc(4 * (size_t)MLP_PARAMS);
        AIR_HIP(hipMemcpyAsync(W.p, params, 4 * (size_t)MLP_PARAMS, hipMemcpyHostToDevice, s));
        AIR_HIP(hipMemcpyAsync(Xd.p, X, 4 * (size_t)n * MLP_IN, hipMemcpyHostToDevice, s));
        AIR_HIP(hipMemcpyAsync(Td.p, T, 4 * (size_t)n * MLP_OUT, hipMemcpyHostToDevice, s));
        mlp_grad(W.as<float>(), Xd.as<float>(), Td.as<float>(), nullptr, n, (float)stop_weight, part.as<float>(),
                 lp.as<double>(), G.as<float>(), s);
        std::vector<double> l(nblk);
        AIR_HIP(hipMemcpyAsync(grad, G.p, 4 * (size_t)MLP_PARAMS, hipMemcpyDeviceToHost, s));
        AIR_HIP(hipMemcpyAsync(l.data(), lp.p, 8 * (size_t)nblk, hipMemcpyDeviceToHost, s));
        AIR_HIP(hipStreamSynchronize(s));
        if (loss) {
            double t = 0;
            for (double v : l) t += v;
            *loss = t / (double)n;
        }
    });
}

int air_mlp_train(air_engine* e, float* params, const float* X, const float* T, int64_t n, int epochs,
                  int batch, double lr, double stop_weight, uint64_t seed, double* epoch_loss) {
    return guard([&] {
        AIR_CHECK(e && params && X && T && n > 0 && n < (int64_t)1 << 31 && epochs >= 0 && batch > 0, AIR_E_ARG,
                  "invalid MLP training arguments");
        use_device(e);
        hipStream_t s = e->stream;
        const int64_t nbat = (n + batch - 1) / batch;
        const unsigned maxblk = mlp_grad_blocks(batch);
        DevBuf W, M1, M2, Xd, Td, idx, part, lp, G;
        W.alloc(4 * (size_t)MLP_PARAMS);
        M1.alloc(4 * (size_t)MLP_PARAMS);
        M2.alloc(4 * (size_t)MLP_PARAMS);
        Xd.alloc(4 * (size_t)n * MLP_IN);
        Td.alloc(4 * (size_t)n * MLP_OUT);
        idx.alloc(4 * (size_t)n);
        part.alloc(4 * (size_t)maxblk * MLP_PARAMS);
        lp.alloc(8 * (size_t)nbat * maxblk);
        G.alloc(4 * (size_t)MLP_PARAMS);
        AIR_HIP(hipMemcpyAsync(W.p, params, 4 * (size_t)MLP_PARAMS, hipMemcpyHostToDevice, s));
        AIR_HIP(hipMemsetAsync(M1.p, 0, 4 * (size_t)MLP_PARAMS, s));
        AIR_HIP(hipMemsetAsync(M2.p, 0, 4 * (size_t)MLP_PARAMS, s));
        AIR_HIP(hipMemcpyAsync(Xd.p, X, 4 * (size_t)n * MLP_IN, hipMemcpyHostToDevice, s));
        AIR_HIP(hipMemcpyAsync(Td.p, T, 4 * (size_t)n * MLP_OUT, hipMemcpyHostToDevice, s));
        std::vector<int> perm(n);
        std::vector<double> l((size_t)nbat * maxblk);
        std::mt19937_64 rng(seed);
        int step = 0;
        for (int ep = 0; ep < epochs; ++ep) {
            for (int64_t i = 0; i < n; ++i) perm[i] = (int)i;
            std::shuffle(perm.begin(), perm.end(), rng);
            AIR_HIP(hipMemcpyAsync(idx.p, perm.data(), 4 * (size_t)n, hipMemcpyHostToDevice, s));
            AIR_HIP(hipMemsetAsync(lp.p, 0, 8 * (size_t)nbat * maxblk, s));
            for (int64_t bi = 0; bi < nbat; ++bi) {
                const int64_t off = bi * batch, nb = std::min<int64_t>(batch, n - off);
                mlp_grad(W.as<float>(), Xd.as<float>(), Td.as<float>(), idx.as<int>() + off, nb, (float)stop_weight,
                         part.as<float>(), lp.as<double>() + bi * maxblk, G.as<float>(), s);
                mlp_adam(W.as<float>(), M1.as<float>(), M2.as<float>(), G.as<float>(), (float)lr, ++step, s);
            }
            AIR_HIP(hipMemcpyAsync(l.data(), lp.p, 8 * l.size(), hipMemcpyDeviceToHost, s));
            AIR_HIP(hipStreamSynchronize(s));
            if (epoch_loss) {
                // each minibatch's loss partials sum to nb * (its mean loss)
                double t = 0;
                for (double v : l) t += v;
                epoch_loss[ep] = t / (double)n;
            }
        }
        AIR_HIP(hipMemcpyAsync(params, W.p, 4 * (size_t)MLP_PARAMS, hipMemcpyDeviceToHost, s));
        AIR_HIP(hipStreamSynchronize(s));
    });
}

// ---- batches of independent problems
int air_build_paralleltomo_batch(air_engine* e, int32_t nprob, const int32_t* N, const int32_t* nA,
                                 const double* theta, const int32_t* p, int method, int dtype) {
    return guard([&] {
        AIR_CHECK(e && nprob > 0 && N && nA && theta && p, AIR_E_ARG, "invalid batch arguments");
        AIR_CHECK(method >= AIR_LANDWEBER && method <= AIR_SART, AIR_E_ARG, "SIRT method not defined.");
        (void)dtype_size(dtype);
        use_device(e);
        reset_operator(e);
        std::vector<TiledOp> parts;
        parts.reserve(nprob);
        e->b_row_off.assign(1, 0);
        e->b_col_off.assign(1, 0);
        int64_t toff = 0;
        for (int q = 0; q < nprob; ++q) {
            AIR_CHECK(N[q] > 0 && nA[q] > 0 && p[q] > 0 && (int64_t)N[q] * N[q] <= INT32_MAX, AIR_E_ARG,
                      "invalid batch problem");
            air_engine t;                   // this problem alone: CSR, CSC, tiled layout, weights, rho
            t.device = e->device;
            t.num_cu = e->num_cu;
            t.stream = e->stream;
            t.no_tune = true;
            host_rays_paralleltomo(N[q], theta + toff, nA[q], p[q], (double)(p[q] - 1), t.geo.host_rays);
            build_geometry(&t, N[q], (int64_t)nA[q] * p[q], 1, 0, dtype);
            AIR_CHECK(t.T.ready, AIR_E_UNSUPPORTED, "a batch problem does not fit the tiled layout");
            std::vector<double> M, D;
            compute_weights(&t, method, M, D);
            const double rho = method == AIR_SART ? 1.0 : lanczos_rho(&t, M, D);   // calc_relaxpar.m:91-110
            const double lam = 1.9 / rho;
            for (auto& v : M) v *= lam;    // the problem's default lambda, folded into its rows of M
            e->b_M.insert(e->b_M.end(), M.begin(), M.end());
            e->b_D.insert(e->b_D.end(), D.begin(), D.end());
            e->b_rho.push_back(rho);
            e->b_row_off.push_back(e->b_row_off.back() + t.m);
            e->b_col_off.push_back(e->b_col_off.back() + t.n);
            parts.push_back(std::move(t.T));
            toff += nA[q];
            t.stream = nullptr;             // owned by e
        }
        concat_tiled(parts, e->T, e->stream);
        e->kind = 1;
        e->m = e->b_row_off.back();
        e->n = e->b_col_off.back();
    });
}

int air_batch_info(air_engine* e, int64_t* row_off, int64_t* col_off, double* rho, double* M, double* D) {
    return guard([&] {
        AIR_CHECK(e && !e->b_rho.empty(), AIR_E_STATE, "the engine holds no batch");
        const size_t P = e->b_rho.size();
        if (row_off) std::copy(e->b_row_off.begin(), e->b_row_off.end(), row_off);
        if (col_off) std::copy(e->b_col_off.begin(), e->b_col_off.end(), col_off);
        if (rho) std::copy(e->b_rho.begin(), e->b_rho.end(), rho);
        if (M) std::copy(e->b_M.begin(), e->b_M.end(), M);
        if (D) std::copy(e->b_D.begin(), e->b_D.end(), D);
        (void)P;
    });
}

// ---- many right-hand sides on one operator
int air_sirt_multi(air_engine* e, const air_sirt_options* o, int32_t nrhs, const double* B, const double* X0,
                   int32_t K, double* X_out, double* resnorm, double* gpu_ms) {
    return guard([&] {
        need_operator(e);
        AIR_CHECK(e->kind == 1 && e->A.nnz > 0, AIR_E_UNSUPPORTED,
                  "several right-hand sides need a stored matrix (CSR and CSC)");
        AIR_CHECK(o && B && X_out && nrhs >= 1 && nrhs <= 64 && K >= 0, AIR_E_ARG,
                  "invalid arguments (1 <= nrhs <= 64)");
        AIR_CHECK(o->stoprule == AIR_STOP_NONE, AIR_E_UNSUPPORTED, "several right-hand sides run K iterations");
        AIR_CHECK(o->relax_kind == AIR_RELAX_DEFAULT || o->relax_kind == AIR_RELAX_SCALAR, AIR_E_UNSUPPORTED,
                  "several right-hand sides take the default or a scalar relaxation parameter");
        AIR_CHECK(o->lb_kind != AIR_BOUND_VECTOR && o->ub_kind != AIR_BOUND_VECTOR, AIR_E_UNSUPPORTED,
                  "several right-hand sides take scalar bounds");
        use_device(e);
        hipStream_t s = e->stream;
        const int64_t m = e->m, n = e->n;
        std::vector<double> M, D;
        bool hasM, hasD;
        if (o->method == AIR_CUSTOM) {
            M.assign(m, 1.0);
            D.assign(n, 1.0);
            if (o->M) std::copy(o->M, o->M + m, M.begin());
            if (o->D) std::copy(o->D, o->D + n, D.begin());
            hasM = o->M != nullptr;
            hasD = o->D != nullptr;
        } else {
            compute_weights(e, o->method, M, D);
            hasM = o->method != AIR_LANDWEBER;
            hasD = o->method == AIR_DROP || o->method == AIR_SART;
        }
        double lam = o->relaxpar;
        if (o->relax_kind == AIR_RELAX_DEFAULT) {
            double rho = o->method == AIR_SART ? 1.0 : o->rho;
            if (std::isnan(rho)) {
                if (o->method != AIR_CUSTOM && !std::isnan(e->wrho[o->method])) rho = e->wrho[o->method];
                else {
                    rho = lanczos_rho(e, M, D);
                    if (o->method != AIR_CUSTOM) e->wrho[o->method] = rho;
                }
            }
            lam = 1.9 / rho;
        }
        // RHS-minor FP32 blocks: element (i, j) at i * 64 + j
        const size_t L = 64;
        if (use_tiled(e) && !spmm_disabled() && !e->S.ready) build_spmm(e->A, e->T, e->S, s);
        if (use_tiled(e) && !spmm_disabled() && e->S.ready) {
            // LDS-staged products on the tiled layout's internal column order
            const SpmmOp& S = e->S;
            const int64_t ni = S.n_int;
            const std::vector<int>& perm = S.perm_h;
            std::vector<float> hB((size_t)m * L, 0.f), hX((size_t)ni * L, 0.f);
            for (int32_t j = 0; j < nrhs; ++j)
                for (int64_t i = 0; i < m; ++i) hB[(size_t)i * L + j] = (float)B[(size_t)j * m + i];
            if (X0)
                for (int32_t j = 0; j < nrhs; ++j)
                    for (int64_t c = 0; c < n; ++c) hX[(size_t)perm[c] * L + j] = (float)X0[(size_t)j * n + c];
            DevBuf dB, dX, dMR, dM, dD, part, nrm;
            dB.alloc(4 * (size_t)m * L);
            dX.alloc(4 * (size_t)ni * L);
            dMR.alloc(4 * (size_t)m * L);
            AIR_HIP(hipMemcpyAsync(dB.p, hB.data(), 4 * hB.size(), hipMemcpyHostToDevice, s));
            AIR_HIP(hipMemcpyAsync(dX.p, hX.data(), 4 * hX.size(), hipMemcpyHostToDevice, s));
            if (hasM) upload_f32(e, dM, M.data(), m, s);
            if (hasD) {
                std::vector<double> Di((size_t)ni, 0.0);
                for (int64_t c = 0; c < n; ++c) Di[perm[c]] = D[c];
                upload_f32(e, dD, Di.data(), ni, s);
            }
            const unsigned gr = spmm_rows_grid(m, e->num_cu);
            part.alloc(8 * (size_t)gr * L);
            nrm.alloc(8 * L);
            const float* Mp = hasM ? dM.as<float>() : nullptr;
            const float* Dp = hasD ? dD.as<float>() : nullptr;
            spmm_forward(S, dX.as<float>(), s);
            spmm_rows(S, e->T, dB.as<float>(), Mp, dMR.as<float>(), part.as<double>(), gr, s);
            hipEvent_t a, b;
            AIR_HIP(hipEventCreate(&a));
            AIR_HIP(hipEventCreate(&b));
            AIR_HIP(hipEventRecord(a, s));
            for (int k = 1; k <= K; ++k) {
                spmm_backward_update(S, dMR.as<float>(), dX.as<float>(), Dp, lam, o->lb_kind, (float)o->lb,
                                     o->ub_kind, (float)o->ub, s);
                spmm_forward(S, dX.as<float>(), s);
                spmm_rows(S, e->T, dB.as<float>(), Mp, dMR.as<float>(), part.as<double>(), gr, s);
            }
            AIR_HIP(hipEventRecord(b, s));
            multi_norms(part.as<double>(), (int)gr, nrm.as<double>(), s);
            AIR_HIP(hipEventSynchronize(b));
            float ms = 0;
            AIR_HIP(hipEventElapsedTime(&ms, a, b));
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            if (gpu_ms) *gpu_ms = ms;
            AIR_HIP(hipMemcpyAsync(hX.data(), dX.p, 4 * hX.size(), hipMemcpyDeviceToHost, s));
            std::vector<double> hn(L);
            AIR_HIP(hipMemcpyAsync(hn.data(), nrm.p, 8 * L, hipMemcpyDeviceToHost, s));
            AIR_HIP(hipStreamSynchronize(s));
            for (int32_t j = 0; j < nrhs; ++j)
                for (int64_t c = 0; c < n; ++c) X_out[(size_t)j * n + c] = (double)hX[(size_t)perm[c] * L + j];
            if (resnorm)
                for (int32_t j = 0; j < nrhs; ++j) resnorm[j] = std::sqrt(hn[j]);
            return;
        }
        std::vector<float> hB((size_t)m * L, 0.f), hX((size_t)n * L, 0.f);
        for (int32_t j = 0; j < nrhs; ++j)
            for (int64_t i = 0; i < m; ++i) hB[(size_t)i * L + j] = (float)B[(size_t)j * m + i];
        if (X0)
            for (int32_t j = 0; j < nrhs; ++j)
                for (int64_t i = 0; i < n; ++i) hX[(size_t)i * L + j] = (float)X0[(size_t)j * n + i];
        DevBuf dB, dX, dR, dMR, dM, dD, part, nrm;
        dB.alloc(4 * (size_t)m * L);
        dX.alloc(4 * (size_t)n * L);
        dR.alloc(4 * (size_t)m * L);
        dMR.alloc(4 * (size_t)m * L);
        AIR_HIP(hipMemcpyAsync(dB.p, hB.data(), 4 * hB.size(), hipMemcpyHostToDevice, s));
        AIR_HIP(hipMemcpyAsync(dX.p, hX.data(), 4 * hX.size(), hipMemcpyHostToDevice, s));
        if (hasM) upload_f32(e, dM, M.data(), m, s);
        if (hasD) upload_f32(e, dD, D.data(), n, s);
        const unsigned gr = multi_grid(m, e->num_cu), gc = multi_grid(n, e->num_cu);
        part.alloc(8 * (size_t)gr * L);
        nrm.alloc(8 * L);
        const float* Mp = hasM ? dM.as<float>() : nullptr;
        const float* Dp = hasD ? dD.as<float>() : nullptr;
        multi_resid(e->A, nrhs, dX.as<float>(), dB.as<float>(), Mp, dR.as<float>(), dMR.as<float>(),
                    part.as<double>(), gr, s);
        hipEvent_t a, b;
        AIR_HIP(hipEventCreate(&a));
        AIR_HIP(hipEventCreate(&b));
        AIR_HIP(hipEventRecord(a, s));
        for (int k = 1; k <= K; ++k) {
            multi_update(e->AT, nrhs, dMR.as<float>(), Dp, dX.as<float>(), lam, o->lb_kind, (float)o->lb, o->ub_kind,
                         (float)o->ub, gc, s);
            multi_resid(e->A, nrhs, dX.as<float>(), dB.as<float>(), Mp, dR.as<float>(), dMR.as<float>(),
                        part.as<double>(), gr, s);
        }
        AIR_HIP(hipEventRecord(b, s));
        multi_norms(part.as<double>(), (int)gr, nrm.as<double>(), s);
        AIR_HIP(hipEventSynchronize(b));
        float ms = 0;
        AIR_HIP(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (gpu_ms) *gpu_ms = ms;
        AIR_HIP(hipMemcpyAsync(hX.data(), dX.p, 4 * hX.size(), hipMemcpyDeviceToHost, s));
        std::vector<double> hn(L);
        AIR_HIP(hipMemcpyAsync(hn.data(), nrm.p, 8 * L, hipMemcpyDeviceToHost, s));
        AIR_HIP(hipStreamSynchronize(s));
        for (int32_t j = 0; j < nrhs; ++j)
            for (int64_t i = 0; i < n; ++i) X_out[(size_t)j * n + i] = (double)hX[(size_t)i * L + j];
        if (resnorm)
            for (int32_t j = 0; j < nrhs; ++j) resnorm[j] = std::sqrt(hn[j]);
    });
}
}  // extern "C"
