// air_internal.h -- shared declarations of the MI355X SIRT engine (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
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

// Zeroed entries kept after nnz in idx/val: the row kernels read rows with
// 16-byte vector loads from an aligned start and may run past the last entry.
constexpr int64_t AIR_SPARSE_PAD = 2048;

inline void alloc_sparse_arrays(DevSparse& S, int64_t nnz, int dtype, hipStream_t s) {
    const size_t vs = dtype_size(dtype);
    S.idx.alloc(sizeof(int) * (size_t)(nnz + AIR_SPARSE_PAD));
    S.val.alloc(vs * (size_t)(nnz + AIR_SPARSE_PAD));
    AIR_HIP(hipMemsetAsync(S.idx.as<int>() + nnz, 0, sizeof(int) * AIR_SPARSE_PAD, s));
    AIR_HIP(hipMemsetAsync((char*)S.val.p + vs * nnz, 0, vs * AIR_SPARSE_PAD, s));
}

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
    // matrix-free projector: rays binned by T x T image tile (tile-major item
    // lists), and per ray the positions of its items for the forward sums
    int T = 0, ntiles_side = 0;
    int64_t nitems = 0;
    DevBuf tile_ptr;      // int64 [tiles + 1]
    DevBuf tile_rays;     // int32 [nitems]: ray of each item, ascending within a tile
    DevBuf ray_ptr;       // int64 [nrays + 1]
    DevBuf ray_items;     // int32 [nitems]: item positions of each ray
    DevBuf partial;       // float [nitems]: forward partial sums
    DevBuf raux;          // per-ray 1/a, 1/b, |d|, max(|a|,|b|) for the DDA walk
    bool mf_exact = false; // crossing walk (entries equal to the tracer's bit for bit) instead of the DDA
    int splits = 1;       // work-groups per tile
    DevBuf gpart;         // double [splits * N * N]: adjoint slices when splits > 1
    // parallel beams: pixel-driven adjoint (pnA > 0 when prepared)
    int pnA = 0, pp = 0, papc = 0, pchunks = 0;
    DevBuf pang;          // per angle: direction, reciprocals, |d|, max(|a|,|b|)
    DevBuf pdet;          // double [p]: detector positions
    DevBuf ppart;         // double [pchunks * N * N]: per-chunk sums
    // model 1: Fresnel-zone rows of seismicwavetomo (nrays = rows)
    int model = 0;        // 0: line rays, 1: Fresnel zones, 2: circles of sphericaltomo
    DevBuf frows;         // per-row source/receiver geometry
    DevBuf frowsum;       // double [rows]: sum of the kernel over the image
    double fw = 0;        // omega / N
    // model 2: sphericaltomo circles (nrays = rows = angles x circles); per
    // angle S samples, circle c at [ccum[c], ccum[c+1]) with its cos/sin table
    int cnA = 0, cnC = 0;
    int64_t cS = 0;
    double cdx = 0, cimg = 0;
    DevBuf ccos, csin, ccum, crad, cbase, cxi;
    std::vector<double> cangles;
};

// ---------------------------------------------------------------- tracer (air_tracer.hip)
void parallel_detector(int p, double d, std::vector<double>& x0);
void host_rays_paralleltomo(int N, const double* theta, int nA, int p, double d, std::vector<RayParam>& out);
void mf_prepare_parallel(Geometry& g, int nA, int p, double d, hipStream_t s);
void host_rays_fanlineartomo(int N, const double* theta, int nA, int p, double R, double dw, double sd,
                             std::vector<RayParam>& out);
void host_rays_fancurvedtomo(int N, const double* theta, int nA, int p, double R, double d,
                             std::vector<RayParam>& out);
// seismictomo.m:98-116 / seismicwavetomo.m:119-132: source heights y0 (sources
// at x = N/2), receiver positions (xp, yp) on the left and top edges.
void seismic_positions(int N, int s, int p, std::vector<double>& y0, std::vector<double>& xp,
                       std::vector<double>& yp);
void host_rays_seismictomo(int N, int s, int p, std::vector<RayParam>& out);
// fbp.m:45-77 (air_fbp.hip): frequency response (half spectrum) and the
// filtered sinogram (nr x na, FP32 on the device).
std::vector<double> fbp_filter_response(int64_t nr, int filter, int64_t* order_out);
void fbp_filter_sinogram(const double* b_host, int64_t nr, int64_t na, int filter, float* out, hipStream_t s);
// ART / CART sweeps (air_art.hip): A is the CSR, AT the CSC of the operator.
void row_action_run(const DevSparse& A, const DevSparse& AT, bool cart, int art_method, const int32_t* order,
                    int64_t norder, const air_row_options* o, const double* b_h, const double* x0_h,
                    const int32_t* K_in, int32_t nK, double* X_out, int32_t* itersaved, int32_t* n_itersaved,
                    double* resnorm, air_sirt_info* info, hipStream_t s);
// seismicwavetomo: CSR of the Fresnel-zone sensitivity matrix (air_fresnel.hip).
void fresnel_to_csr(int N, int s, int p, double omega, int dtype, DevSparse& A, hipStream_t st);
// sphericaltomo: CSR of the circle-quadrature matrix (air_circles.hip).
void circles_to_csr(int N, const double* angles, int nA, int nC, int dtype, DevSparse& A, hipStream_t s);
void circles_mf_prepare(Geometry& g, int N, const double* angles, int nA, int nC, hipStream_t s);
void circles_mf_forward(const Geometry& g, const float* u, float* y, hipStream_t s);
void circles_mf_adjoint(const Geometry& g, const float* v, double* gacc, hipStream_t s);
void circles_mf_row_stats(const Geometry& g, const double* colw, double* sq, double* ab, double* sqw, hipStream_t s);
void circles_mf_col_stats(const Geometry& g, double* colabs, int* colcnt, hipStream_t s);
// Trace all rays into CSR (values narrowed to dtype).  Returns nnz.
void trace_to_csr(const Geometry& g, int dtype, DevSparse& A, hipStream_t s);
// Matrix-free products: y = A u (m) or g = A' v (n, FP64 accumulators).
void mf_build_tiles(Geometry& g, hipStream_t s);
void fresnel_mf_prepare(Geometry& g, int N, int s, int p, double omega, hipStream_t st);
void fresnel_mf_forward(const Geometry& g, const float* u, float* y, hipStream_t st);
void fresnel_mf_adjoint(const Geometry& g, const float* v, double* gout, hipStream_t st);
void fresnel_mf_row_stats(const Geometry& g, const double* colw, double* sq, double* ab, double* sqw, hipStream_t st);
void fresnel_mf_col_stats(const Geometry& g, double* colabs, int* colcnt, hipStream_t st);
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

// ---------------------------------------------------------------- tiled operator (air_tiled.hip)
// The SIRT hot path's layout of A: columns (pixels) are renumbered tile-major
// (64x64 tiles made of 16x16 sub-tiles, column-major inside) and A is stored
// twice with 16-bit local indices:
//  forward (A x): per 64x64 tile, the row segments ("pieces") of A inside the
//    tile, each entry (value, local column); the tile of x is staged in LDS and
//    every piece yields one partial sum; a second kernel adds a row's partials.
//  backward (A' y): per 16x16 sub-tile, the columns of A with (value, local
//    row); the sub-tile's distinct rows of y are staged in LDS (as runs of
//    consecutive rows) and each column is summed with LDS gathers.
// dynamic LDS a backward tile may stage (two 512-thread blocks per CU)
constexpr size_t AIR_TILE_LDS_MAX = 60 * 1024;
// dynamic LDS of the step-coded backward's strided staging: two 512-thread
// blocks per CU with the kernel's 2.6 KB of static LDS, or one 1024-thread
// block per CU for larger stagings
constexpr size_t AIR_TILE_LDS_STEP_MAX = 76 * 1024;
constexpr size_t AIR_TILE_LDS_STEP16_MAX = 152 * 1024;

struct TiledOp {
    bool ready = false;
    int dtype = AIR_F32;
    int N = 0;                  // image side for 2-D geometries (0: generic column blocks)
    int64_t m = 0, n = 0, n_int = 0, nnz = 0;
    int64_t nfe = 0, nbe = 0;   // entries of the forward / backward streams (segments padded to 8)
    int64_t ntiles_f = 0, ntiles_b = 0, npieces = 0, nruns = 0;
    int bt_size = 256;          // columns per backward tile
    int nblk_f = 0;             // forward blocks
    int maxruns = 0, maxrows = 0;
    int64_t staged_rows = 0;    // sum over backward tiles of their distinct rows
    size_t lds_b = 0;           // dynamic LDS bytes of the backward kernel
    DevBuf perm;                // int32 [n]: user column -> internal column
    DevBuf fval, fcode, piece_ptr, blk_tile, blk_p0, rp_ptr, rp_idx, partial;
    DevBuf bval, blr, cell_ptr, run_ptr, run_row, run_off, run_len;
    int fwd_cfg = 0, bwd_cfg = 0;          // kernel shapes picked by the autotuner
    int ppb = 256;                         // pieces per forward block
    std::vector<int> piece_tile_h;         // forward tile of every piece (host copy)
    std::vector<int64_t> piece_ptr_h;      // first entry of every piece (host copy)
    // stream form of the forward product: waves own runs of whole pieces
    DevBuf wblk_tile, wave_p0, wave_e;     // wave_e: first stream entry of every wave (wave_p0's pieces)
    int nblk_w = 0, nwaves = 0, wave_target = 4096;
    int fwd_wpb = 4;                       // waves per block of the step-coded forward (4 or 16; other forms 4)
    // backward codes carry a column-start flag in bit 15 when every tile has
    // fewer than 32768 distinct rows (required by the stream form)
    bool bwd_flags = false;
    // step-coded forward lanes (one word per 8 entries); fstep_ok when every
    // piece of the matrix is a unit-step staircase
    DevBuf fstep;
    bool fstep_ok = false;
    // step-coded backward lanes: values with bridging zeros, their column
    // pointers, one word per 8 entries, the slot stride per tile, and the LDS
    // of the strided staging; bstep_ok when every column could be encoded
    DevBuf bsval, bstep_ptr, bsword, bstep_R;
    DevBuf bstep_we, bstep_wk;  // per (tile, wave of 8): entry range and first column ordinal
    int64_t nbse = 0;
    size_t lds_bs = 0;
    bool bstep_ok = false;
    // bits of a lane word's base slot: 14 (offset of the first entry in its
    // group up to 7) while every tile stages fewer than 16384 slots, else 15
    // (offset up to 3, slots up to 32767)
    int bstep_bb = 14;
    // waves per tile of the step-coded backward: 8 (two blocks per CU) while
    // the staging fits AIR_TILE_LDS_STEP_MAX, else 16 (one block per CU)
    int bstep_nw = 8;
    float tune_fwd_ms[64] = {}, tune_bwd_ms[16] = {};
};
// Many right-hand sides (air_spmm.hip): per product, the launch order of the
// 1024-thread work-groups, their table, every wave's first stream block and
// block count, the operand chunks and their per-(wave, round) step counts,
// and the record stream (1 KB blocks of 12 records).
struct SpmmPass {
    DevBuf lperm, wgt, wblk, wnb, chd, ctab, stream;
    int nwg_launch = 0;
    int64_t nblocks = 0;
    int64_t step_slots = 0;     // 4 x the steps of all rounds (entries + the zeros padding them)
    double wave_balance = 1.0;  // sum over chunks of the mean over waves of records / of the max
};
struct SpmmOp {
    bool ready = false;
    int64_t m = 0, n = 0, n_int = 0, npieces = 0;
    std::vector<int> perm_h;    // user column -> internal column (the tiled layout's)
    SpmmPass fwd, bwd;
    DevBuf rowlist;             // distinct rows of every 32 x 32 block, in order
    DevBuf partial;             // float [pieces][64]
};
bool build_spmm(const DevSparse& A, const struct TiledOp& T, SpmmOp& S, hipStream_t s);
void spmm_forward(const SpmmOp& S, const float* X, hipStream_t s);
void spmm_backward_update(const SpmmOp& S, const float* MR, float* X, const float* D, double lam, int lb_kind,
                          float lb, int ub_kind, float ub, hipStream_t s);
unsigned spmm_rows_grid(int64_t m, int num_cu);
void spmm_rows(const SpmmOp& S, const struct TiledOp& T, const float* B, const float* M, float* MR, double* part,
               unsigned grid, hipStream_t s);

// Build the tiled operator from the CSR of A (N > 0: N x N image columns).
// Returns false (and leaves T not ready) when a tile needs more LDS than allowed.
bool build_tiled(const DevSparse& A, int N, TiledOp& T, hipStream_t s, bool tune = true);
// Concatenate independent problems' tiled layouts into one block-diagonal operator (air_batch.hip).
bool concat_tiled(std::vector<TiledOp>& parts, TiledOp& B, hipStream_t s);
void tiled_forward_plain(const TiledOp& T, const float* x_int, float* y, hipStream_t s);
void autotune_tiled(TiledOp& T, hipStream_t s);
void set_fwd_blocks(TiledOp& T, int ppb, hipStream_t s);
void set_fwd_waves(TiledOp& T, int target_entries, hipStream_t s);
void set_bstep_waves(TiledOp& T, hipStream_t s);
bool tiled_fwd_stream(const TiledOp& T);
bool tiled_fwd_step(const TiledOp& T);
bool tiled_bwd_step(const TiledOp& T);
void tiled_shapes(const TiledOp& T, int* fwd_lanes, int* fwd_loads, int* bwd_lanes, int* bwd_loads);
void tiled_backward_plain(const TiledOp& T, const float* y, float* g_int, hipStream_t s);
void permute_to_internal(const TiledOp& T, const float* x_user, float* x_int, hipStream_t s);
void permute_to_user(const TiledOp& T, const float* x_int, float* x_user, hipStream_t s);

// ---------------------------------------------------------------- many right-hand sides (air_multi.hip)
unsigned multi_grid(int64_t rows, int num_cu);
void multi_resid(const DevSparse& A, int nrhs, const float* X, const float* B, const float* M, float* R, float* MRv,
                 double* part, unsigned grid, hipStream_t s);
void multi_update(const DevSparse& AT, int nrhs, const float* MRv, const float* D, float* X, double lam, int lb_kind,
                  float lb, int ub_kind, float ub, unsigned grid, hipStream_t s);
void multi_norms(const double* part, int nblk, double* out, hipStream_t s);

// ---------------------------------------------------------------- learned relaxation (air_mlp.hip)
constexpr int MLP_IN = AIR_MLP_IN, MLP_HID = AIR_MLP_HID, MLP_OUT = AIR_MLP_OUT, MLP_PARAMS = AIR_MLP_PARAMS;
constexpr int MLP_LAMBDA_MAX = AIR_MLP_LAMBDA_MAX;
void mlp_forward(const float* W, const float* X, int64_t n, float* out, hipStream_t s);
unsigned mlp_grad_blocks(int64_t nb);
// part: mlp_grad_blocks(nb) x MLP_PARAMS floats; loss_part: mlp_grad_blocks(nb) doubles
void mlp_grad(const float* W, const float* X, const float* T, const int* idx, int64_t nb, float stop_w,
              float* part, double* loss_part, float* grad, hipStream_t s);
void mlp_adam(float* W, float* m1, float* m2, const float* grad, float lr, int step, hipStream_t s);

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
    int kcur;            // last iteration whose residual and stop check are done
    double lam_cur;      // relaxpar of the last executed iteration (graph mode: of the next one)
    double rMr;          // r'Mr of the current residual (line search numerator)
    double last_norm;
};

struct StopParams {
    int type;            // AIR_STOP_*
    int kmax;
    double taudelta;
    // NCP (check_stoprules.m:96-151): per-column statistics and the filter window
    const double* ncp_stat;
    double* ncp_dk;
    int ncp_cols;
    int ncp_smooth;
};

// NCP state of one SIRT run: a batched real FFT of the residual columns.
struct NcpState {
    int64_t m = 0;
    int len = 0, ncols = 0, smooth = 2;
    hipfftHandle plan = 0;
    bool planned = false;
    DevBuf rd, spec, colstat, dk;
    NcpState() = default;
    NcpState(const NcpState&) = delete;
    NcpState& operator=(const NcpState&) = delete;
    ~NcpState() { release(); }
    void release() {
        if (planned) (void)hipfftDestroy(plan);
        planned = false;
    }
};

// Device state of the learned / greedy relaxation (AIR_RELAX_LEARNED, AIR_RELAX_GREEDY).
constexpr int LEARN_NF = 16;          // features per iteration
constexpr int LEARN_STATE = 8;        // doubles of state carried between iterations
struct LearnBufs {
    int mode = 0;
    int kmax = 0;
    int64_t nx = 0;
    float rho = 1.f;
    double noise = 0.0;
    const float* mlp = nullptr;       // learned: FP32 parameters
    const float* xstar = nullptr;     // greedy: the exact solution (loop column order)
    double* state = nullptr;
    double* gpart = nullptr;          // greedy: 3 partial sums per block
    int gblocks = 0;
    float* feat = nullptr;            // kmax x LEARN_NF
    float* target = nullptr;          // kmax
    float* out = nullptr;             // kmax x 2
    double* err = nullptr;            // kmax
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
    double* partial;     // per-block partial sums of the row pass (4 per block)
    double* partial_cols;  // per-block partial sums of the column pass (1 per block)
    double* hist;        // ||r_k||, k = 0..kmax
    Ctrl* ctrl;
    NcpState* ncp = nullptr;  // NCP stop rule state (null otherwise)
    const double* lamv = nullptr;  // graph mode: lambda_1..lambda_kmax on the device (null otherwise)
    const LearnBufs* learn = nullptr;  // learned / greedy relaxation (null otherwise)
    // deferred norms (tiled loop, stop rule none, lambda from the host): the row
    // pass of iteration k >= 1 leaves its block partials in slot k % ring_slots
    // and launch_finalize_deferred turns a run of slots into hist[k] and the
    // stop check afterwards, so no finalize kernel sits between the two passes
    double* ring = nullptr;
    int ring_slots = 0;
};

// r = b - A x; Mr = M r; FP64 reductions and the stop check of iteration k.
void launch_resid(const DevSparse& A, Shape sh, const SirtBufs& B, StopParams sp, int k, hipStream_t s);
// x = P_C(x + lam D (A' Mr)) over the CSC of A.
void launch_update(const DevSparse& AT, Shape sh, const SirtBufs& B, double lam, int k, hipStream_t s);
// line search: g = A' Mr (FP64) with sum D g^2; then lam = r'Mr / sum D g^2 and the update.
void launch_line_update(const DevSparse& AT, Shape sh, const SirtBufs& B, int k, hipStream_t s);
// the same two passes on the tiled operator (x, D, bounds in the internal column order)
void launch_tiled_resid(const TiledOp& T, const SirtBufs& B, StopParams sp, int k, hipStream_t s, hipEvent_t mid);
void launch_tiled_update(const TiledOp& T, const SirtBufs& B, double lam, int line, int k, hipStream_t s);
void launch_apply_line(int64_t n, const SirtBufs& B, hipStream_t s);
// single-block reductions of the per-block partials (stop rule / line-search lambda)
void launch_finalize_rows(const SirtBufs& B, int nblocks, StopParams sp, int k, hipStream_t s);
// NCP: FFT of the residual columns and their statistics (before the finalize)
void ncp_prepare(NcpState& N, int64_t m, int len, int ncols, int smooth, hipStream_t s);
void ncp_enqueue(NcpState& N, const float* r, int64_t m, const Ctrl* ctrl, hipStream_t s);
// the same on an FP64 residual (the row-action methods keep theirs in FP64)
void ncp_enqueue_f64(NcpState& N, const double* r, int64_t m, const Ctrl* ctrl, hipStream_t s);
// lambda of a line-search style iteration k: r'Mr / sum D g^2, or the learned / greedy choice
void launch_finalize_line(const SirtBufs& B, int nblocks, int k, hipStream_t s);
// learned relaxation: the network's lambda_k (read by the update kernels when
// they are given lam = NaN) and stop, from the residual statistics of x_{k-1}
void launch_learned_select(const SirtBufs& B, int k, hipStream_t s);
// true when the finalize kernel of the row pass computes the learned lambda (no separate kernel)
bool fused_select(const SirtBufs& B);
// blocks of the tiled row pass
unsigned tiled_reduce_grid(int64_t m);
// allow the step-coded backward kernels the tile's staging when it exceeds 64 KB
void tiled_bstep_lds_attr(const TiledOp& T);
// hist[k], k = k0..k1, from the deferred block partials; the stop check of k1
void launch_finalize_deferred(const SirtBufs& B, int nblocks, StopParams sp, int k0, int k1, hipStream_t s);
// matrix-free iteration pieces
void launch_mf_resid(const Geometry& g, const SirtBufs& B, StopParams sp, int k, float* Ax, hipStream_t s);
void launch_mf_update(const Geometry& g, int64_t n, const SirtBufs& B, double lam, int line, int k, hipStream_t s);
void launch_vec_resid(int64_t m, const float* Ax, const SirtBufs& B, StopParams sp, int k, hipStream_t s);
void launch_g_update(int64_t n, const SirtBufs& B, double lam, int line, int k, hipStream_t s);
// small helpers
void vec_f64_to_f32(const double* in, float* out, int64_t n, hipStream_t s);
void vec_f32_to_f64(const float* in, double* out, int64_t n, hipStream_t s);
void vec_mul_f32(const float* a, float* y, int64_t n, hipStream_t s);   // y *= a
void zero_f64(double* p, int64_t n, hipStream_t s);

}  // namespace air
