// air_art.hip -- ART (art/art.m) and CART (cart/cart.m) row and column
// action sweeps on gfx950.
//
// Both methods are sequential by definition: every row (ART) or column
// (CART) step reads what the previous step wrote, and in tomography
// consecutive steps share pixels (rays of one angle overlap their neighbours,
// rays of different angles cross), so there is no reordering that keeps the
// reference's result.  The GPU form therefore runs each sweep as ONE wave that
// walks the steps in the reference's order: the 64 lanes gather the step's
// entries, a butterfly sum gives the inner product in every lane, the lanes
// scatter the update, and a work-group fence orders the stores before the next
// step's loads.  x (ART) and the residual (CART) are FP64 as in MATLAB;
// matrix values are read in their storage type.  Everything around the sweep
// (row norms, residuals, stop rules, saved iterates) is wide device work.
//
// Random choices (randkaczmarz row orders, CART unflagging) cannot follow
// MATLAB's rand streams, which are not reproduced; they come from a counter-
// based generator (SplitMix64 of the seed, a stream number, the sweep and the
// step) that the host, the device and the CPU restatement evaluate alike.
#include <algorithm>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <vector>

#include "air_internal.h"

namespace air {

namespace {

constexpr uint64_t RNG_PERM = 1, RNG_SAMPLE = 2, RNG_UNFLAG = 3;

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// uniform in [0, 1) for (seed, stream, sweep k, step t)
__host__ __device__ __forceinline__ double rng_u01(uint64_t seed, uint64_t stream, uint64_t k, uint64_t t) {
    const uint64_t z = splitmix64(splitmix64(seed ^ stream) + (k << 32) + t);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double clamp_bounds(double v, int64_t j, int lb_kind, double lb, const double* lbv,
                                               int ub_kind, double ub, const double* ubv) {
    // art.m:279-284: max with the lower bound, then min with the upper bound
    if (lb_kind == AIR_BOUND_SCALAR) v = fmax(v, lb);
    else if (lb_kind == AIR_BOUND_VECTOR) v = fmax(v, lbv[j]);
    if (ub_kind == AIR_BOUND_SCALAR) v = fmin(v, ub);
    else if (ub_kind == AIR_BOUND_VECTOR) v = fmin(v, ubv[j]);
    return v;
}

__device__ __forceinline__ void order_stores() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

struct Bounds {
    int lb_kind = AIR_BOUND_NONE, ub_kind = AIR_BOUND_NONE;
    double lb = 0, ub = 0;
    const double* lbv = nullptr;
    const double* ubv = nullptr;
};

// What one ART step needs besides the row's entries, gathered for a whole
// sweep by a parallel kernel so that the sweep loads it one step ahead.
struct ArtStep {
    int64_t s, e;       // the row's entries [s, e)
    double b, nrm, lam; // b(ri), normAi(ri) (damped), relaxpar of the step
};

__global__ void k_art_steps(const int* __restrict__ rows, int64_t L, const int64_t* __restrict__ ptr,
                            const double* __restrict__ b, const double* __restrict__ nrm,
                            const double* __restrict__ lamv, double lam, ArtStep* __restrict__ out) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < L; t += (int64_t)gridDim.x * blockDim.x) {
        const int r = rows[t];
        ArtStep a;
        a.s = ptr[r];
        a.e = ptr[r + 1];
        a.b = b[r];
        a.nrm = nrm[r];
        a.lam = lamv ? lamv[t] : lam;
        out[t] = a;
    }
}

// One ART sweep (art.m:249-285) over the steps st[0..L): one wave.  The next
// step's record and its first PC*64 entries are loaded while the current
// step runs; x lives in LDS (copied in and out) when it fits, else in global
// memory with a work-group fence between steps.
template <typename V, bool LDS, int PC>
__global__ __launch_bounds__(64) void k_art_sweep(const ArtStep* __restrict__ st, int64_t L,
                                                  const int* __restrict__ idx, const V* __restrict__ val,
                                                  double* xg, Bounds bd, int64_t n, int project_all_first) {
    extern __shared__ double xlds[];
    double* x = LDS ? xlds : xg;
    const int lane = threadIdx.x;
    const bool bounded = bd.lb_kind != AIR_BOUND_NONE || bd.ub_kind != AIR_BOUND_NONE;
    if (LDS) {
        for (int64_t j = lane; j < n; j += 64) x[j] = xg[j];
        __syncthreads();
    }
    if (L <= 0) return;
    ArtStep cur = st[0];
    int cc[PC];
    double aa[PC];
#pragma unroll
    for (int u = 0; u < PC; ++u) {
        const int64_t q = cur.s + 64 * u + lane;
        cc[u] = q < cur.e ? idx[q] : 0;
        aa[u] = q < cur.e ? to_f64(val[q]) : 0.0;
    }
    for (int64_t t = 0; t < L; ++t) {
        ArtStep nxt = cur;
        int nc[PC];
        double na[PC];
        if (t + 1 < L) nxt = st[t + 1];
#pragma unroll
        for (int u = 0; u < PC; ++u) {
            const int64_t q = nxt.s + 64 * u + lane;
            const bool ok = t + 1 < L && q < nxt.e;
            nc[u] = ok ? idx[q] : 0;
            na[u] = ok ? to_f64(val[q]) : 0.0;
        }
        double dot = 0;
#pragma unroll
        for (int u = 0; u < PC; ++u)
            if (cur.s + 64 * u + lane < cur.e) dot += aa[u] * x[cc[u]];
        for (int64_t q = cur.s + 64 * PC + lane; q < cur.e; q += 64) dot += to_f64(val[q]) * x[idx[q]];
        dot = wave_sum(dot);
        const double coef = (cur.lam * (cur.b - dot)) / cur.nrm;
#pragma unroll
        for (int u = 0; u < PC; ++u)
            if (cur.s + 64 * u + lane < cur.e) {
                const int j = cc[u];
                double v = x[j] + coef * aa[u];
                if (bounded) v = clamp_bounds(v, j, bd.lb_kind, bd.lb, bd.lbv, bd.ub_kind, bd.ub, bd.ubv);
                x[j] = v;
            }
        for (int64_t q = cur.s + 64 * PC + lane; q < cur.e; q += 64) {
            const int j = idx[q];
            double v = x[j] + coef * to_f64(val[q]);
            if (bounded) v = clamp_bounds(v, j, bd.lb_kind, bd.lb, bd.lbv, bd.ub_kind, bd.ub, bd.ubv);
            x[j] = v;
        }
        if (LDS) __syncthreads();
        else order_stores();
        if (t == 0 && project_all_first && bounded) {
            // the reference projects the whole vector after every row; entries
            // off the row keep their projected value afterwards, so after the
            // first row of the run only the row's own entries can change
            for (int64_t j = lane; j < n; j += 64)
                x[j] = clamp_bounds(x[j], j, bd.lb_kind, bd.lb, bd.lbv, bd.ub_kind, bd.ub, bd.ubv);
            if (LDS) __syncthreads();
            else order_stores();
        }
        cur = nxt;
#pragma unroll
        for (int u = 0; u < PC; ++u) {
            cc[u] = nc[u];
            aa[u] = na[u];
        }
    }
    if (LDS)
        for (int64_t j = lane; j < n; j += 64) xg[j] = x[j];
}

// max |x| over the whole vector, NaN ignored (MATLAB max)
__device__ double wave_absmax(const double* x, int64_t n, int64_t* arg) {
    const int lane = threadIdx.x;
    double best = -1.0;
    int64_t bi = -1;
    for (int64_t j = lane; j < n; j += 64) {
        const double a = fabs(x[j]);
        if (a > best) { best = a; bi = j; }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const int64_t oi = __shfl_xor(bi, o, 64);
        if (ob > best || (ob == best && oi >= 0 && (bi < 0 || oi < bi))) { best = ob; bi = oi; }
    }
    *arg = bi;
    return bi < 0 ? NAN : best;
}

struct CartParams {
    double lam, thr;
    int k, kbegin, nunflag;
    uint64_t seed;
};

// One CART sweep (cart.m:226-287) over cols[0..L): one wave.  CSC of A.
template <typename V>
__global__ __launch_bounds__(64) void k_cart_sweep(const int64_t* __restrict__ cptr, const int* __restrict__ ridx,
                                                   const V* __restrict__ val, const int* __restrict__ cols,
                                                   int64_t L, const double* __restrict__ nrm, double* x, double* r,
                                                   unsigned char* flag, int* nflag, Bounds bd, int64_t n,
                                                   CartParams cp) {
    const int lane = threadIdx.x;
    const bool track = cp.k >= cp.kbegin;
    int64_t arg = -1;
    double mm = track ? wave_absmax(x, n, &arg) : NAN;
    for (int64_t t = 0; t < L; ++t) {
        const int j = cols[t];
        if (flag[j]) {
            const int64_t s = cptr[j], e = cptr[j + 1];
            double dot = 0;
            for (int64_t q = s + lane; q < e; q += 64) dot += to_f64(val[q]) * r[ridx[q]];
            dot = wave_sum(dot);
            const double delta = dot / nrm[j];
            double od = cp.lam * delta;
            const double xkj = x[j];
            if (bd.lb_kind != AIR_BOUND_NONE) {                   // cart.m:251-259
                const double lbj = bd.lb_kind == AIR_BOUND_SCALAR ? bd.lb : bd.lbv[j];
                if (od < lbj - xkj) od = lbj - xkj;
            }
            if (bd.ub_kind != AIR_BOUND_NONE) {                   // :261-267
                const double ubj = bd.ub_kind == AIR_BOUND_SCALAR ? bd.ub : bd.ubv[j];
                if (od > ubj - xkj) od = ubj - xkj;
            }
            const double xn = xkj + od;
            if (lane == 0) {
                x[j] = xn;
                if (track && fabs(od) < cp.thr * mm) {            // :273-276
                    flag[j] = 0;
                    nflag[j] = 1;
                }
            }
            for (int64_t q = s + lane; q < e; q += 64) {          // :279
                const int i = ridx[q];
                r[i] = r[i] - od * to_f64(val[q]);
            }
            order_stores();
            if (track) {                                          // keep mm = max(abs(xk)) for the next step
                const double an = fabs(xn);
                if (an >= mm || isnan(mm)) { if (!isnan(an)) { mm = an; arg = j; } }
                else if (j == arg) { mm = wave_absmax(x, n, &arg); }
            }
        } else {                                                  // :281-285
            if (lane == 0) {
                const int draw =
                    1 + (int)floor(rng_u01(cp.seed, RNG_UNFLAG, (uint64_t)cp.k, (uint64_t)j) * cp.nunflag);
                if (nflag[j] < draw) nflag[j] = nflag[j] + 1;
                else flag[j] = 1;
            }
            order_stores();                                       // a custom order may visit j again
        }
    }
}

template <typename V, bool LDS, int PC>
void art_sweep_launch(const DevSparse& A, const ArtStep* st, int64_t L, double* x, Bounds bd, int64_t n, int paf,
                      hipStream_t s) {
    const size_t lds = LDS ? 8 * (size_t)n : 0;
    if (LDS && lds > 64 * 1024)
        AIR_HIP(hipFuncSetAttribute((const void*)k_art_sweep<V, LDS, PC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    k_art_sweep<V, LDS, PC><<<1, 64, lds, s>>>(st, L, A.idx.as<int>(), A.val.as<V>(), x, bd, n, paf);
    AIR_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- chunked sweep
// The same sweep with the rows staged through LDS: the steps are cut into
// chunks (at most ART_CMAX steps and E entries each); while wave 0 runs the
// steps of chunk c from LDS, waves 1..15 copy the records and entries of
// chunk c+1 into the other half of a double buffer, and the whole work-group
// meets at one barrier per chunk.  x stays in LDS throughout.  The inner
// product is reduced with DPP row shifts and row broadcasts in FP64 (lane 63
// ends with the total), and the next step's record and first entries are read
// while the current one finishes.
constexpr int ART_CMAX = 256;
constexpr int ART_BLOCK = 1024;

struct ArtRec {
    double b, nrm, lam;
    int off, len;
};

template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_d(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(u & 0xffffffffll), CTRL, ROWMASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, ROWMASK, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// sum over the 64 lanes, returned in every lane
__device__ __forceinline__ double wave_sum_dpp(double v) {
    v += dpp_d<0x111, 0xf>(v);     // row_shr:1
    v += dpp_d<0x112, 0xf>(v);     // row_shr:2
    v += dpp_d<0x114, 0xf>(v);     // row_shr:4
    v += dpp_d<0x118, 0xf>(v);     // row_shr:8: lane 15 of each row holds the row
    v += dpp_d<0x142, 0xa>(v);     // row_bcast:15 -> rows 1, 3
    v += dpp_d<0x143, 0xc>(v);     // row_bcast:31 -> rows 2, 3: lane 63 holds the total
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), 63);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

struct ArtBuf {
    ArtRec* rec;
    double* val;
    int* idx;
};

__device__ __forceinline__ ArtBuf art_buf(char* base, int E, int which) {
    char* p = base + (size_t)which * ((size_t)ART_CMAX * sizeof(ArtRec) + (size_t)E * 12);
    ArtBuf b;
    b.rec = (ArtRec*)p;
    b.val = (double*)(p + (size_t)ART_CMAX * sizeof(ArtRec));
    b.idx = (int*)(p + (size_t)ART_CMAX * sizeof(ArtRec) + (size_t)E * 8);
    return b;
}

// copy chunk c's records and entries into buffer B.  Producer wave w takes
// steps w, w + nwaves, ...: lane l first loads the record of the wave's l-th
// step (all records in one round trip), then the wave copies the entries of
// two steps at a time.
template <typename V>
__device__ __forceinline__ void art_stage(const ArtStep* __restrict__ st, const int* __restrict__ cstart,
                                          const int* __restrict__ soff, const int* __restrict__ idx,
                                          const V* __restrict__ val, int c, ArtBuf B, int wave, int nwaves) {
    const int lane = threadIdx.x & 63;
    const int t0 = cstart[c], cnt = cstart[c + 1] - t0;
    const int mine = cnt > wave ? (cnt - wave + nwaves - 1) / nwaves : 0;   // steps of this wave (<= 64)
    int64_t s0 = 0;
    int len = 0, off = 0;
    if (lane < mine) {
        const int q = wave + nwaves * lane;
        const ArtStep a = st[t0 + q];
        ArtRec r;
        r.b = a.b;
        r.nrm = a.nrm;
        r.lam = a.lam;
        r.off = soff[t0 + q];
        r.len = (int)(a.e - a.s);
        B.rec[q] = r;
        s0 = a.s;
        len = r.len;
        off = r.off;
    }
    for (int u = 0; u < mine; u += 2) {
        const int64_t sa = __shfl(s0, u), sb = __shfl(s0, u + 1 < mine ? u + 1 : u);
        const int la = __shfl(len, u), lb = u + 1 < mine ? __shfl(len, u + 1) : 0;
        const int oa = __shfl(off, u), ob = __shfl(off, u + 1 < mine ? u + 1 : u);
        const int lmax = la > lb ? la : lb;
        for (int i = lane; i < lmax; i += 64) {
            int ia = 0, ib = 0;
            double va = 0, vb = 0;
            if (i < la) { ia = idx[sa + i]; va = to_f64(val[sa + i]); }
            if (i < lb) { ib = idx[sb + i]; vb = to_f64(val[sb + i]); }
            if (i < la) { B.idx[oa + i] = ia; B.val[oa + i] = va; }
            if (i < lb) { B.idx[ob + i] = ib; B.val[ob + i] = vb; }
        }
    }
}

template <typename V>
__global__ __launch_bounds__(ART_BLOCK) void k_art_chunked(const ArtStep* __restrict__ st, const int* __restrict__ cstart,
                                                           int nchunks, const int* __restrict__ soff,
                                                           const int* __restrict__ idx, const V* __restrict__ val,
                                                           double* xg, Bounds bd, int64_t n, int project_all_first,
                                                           int E) {
    extern __shared__ double smem[];
    double* x = smem;
    char* base = (char*)(smem + n);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = ART_BLOCK / 64;
    const bool bounded = bd.lb_kind != AIR_BOUND_NONE || bd.ub_kind != AIR_BOUND_NONE;
    for (int64_t j = threadIdx.x; j < n; j += ART_BLOCK) x[j] = xg[j];
    if (nchunks > 0) art_stage<V>(st, cstart, soff, idx, val, 0, art_buf(base, E, 0), wave, nw);
    __syncthreads();
    for (int c = 0; c < nchunks; ++c) {
        const ArtBuf B = art_buf(base, E, c & 1);
        if (wave == 0) {
            const int cnt = cstart[c + 1] - cstart[c];
            ArtRec r = B.rec[0];
            double a0 = 0, a1 = 0;
            int j0 = 0, j1 = 0;
            if (lane < r.len) { a0 = B.val[r.off + lane]; j0 = B.idx[r.off + lane]; }
            if (lane + 64 < r.len) { a1 = B.val[r.off + lane + 64]; j1 = B.idx[r.off + lane + 64]; }
            for (int q = 0; q < cnt; ++q) {
                // the next step's record and first 128 entries (they do not depend on x)
                ArtRec rn = r;
                double b0 = 0, b1 = 0;
                int k0 = 0, k1 = 0;
                if (q + 1 < cnt) {
                    rn = B.rec[q + 1];
                    if (lane < rn.len) { b0 = B.val[rn.off + lane]; k0 = B.idx[rn.off + lane]; }
                    if (lane + 64 < rn.len) { b1 = B.val[rn.off + lane + 64]; k1 = B.idx[rn.off + lane + 64]; }
                }
                // both gathers issued together (lanes past the row read x[0]
                // and contribute nothing); the values are reused by the update,
                // which is exact because a row's columns are distinct
                const bool on0 = lane < r.len, on1 = lane + 64 < r.len;
                const double x0v = x[j0], x1v = x[j1];
                double dot = on0 ? a0 * x0v : 0.0;
                dot += on1 ? a1 * x1v : 0.0;
                for (int i = lane + 128; i < r.len; i += 64) dot += B.val[r.off + i] * x[B.idx[r.off + i]];
                dot = wave_sum_dpp(dot);
                const double coef = (r.lam * (r.b - dot)) / r.nrm;
                if (on0) {
                    double v = x0v + coef * a0;
                    if (bounded) v = clamp_bounds(v, j0, bd.lb_kind, bd.lb, bd.lbv, bd.ub_kind, bd.ub, bd.ubv);
                    x[j0] = v;
                }
                if (on1) {
                    double v = x1v + coef * a1;
                    if (bounded) v = clamp_bounds(v, j1, bd.lb_kind, bd.lb, bd.lbv, bd.ub_kind, bd.ub, bd.ubv);
                    x[j1] = v;
                }
                for (int i = lane + 128; i < r.len; i += 64) {
                    const int jj = B.idx[r.off + i];
                    double v = x[jj] + coef * B.val[r.off + i];
                    if (bounded) v = clamp_bounds(v, jj, bd.lb_kind, bd.lb, bd.lbv, bd.ub_kind, bd.ub, bd.ubv);
                    x[jj] = v;
                }
                if (c == 0 && q == 0 && project_all_first && bounded) {
                    // the whole vector is projected after every row; after the
                    // first row only the row's own entries can change
                    for (int64_t j = lane; j < n; j += 64)
                        x[j] = clamp_bounds(x[j], j, bd.lb_kind, bd.lb, bd.lbv, bd.ub_kind, bd.ub, bd.ubv);
                }
                r = rn;
                a0 = b0; a1 = b1; j0 = k0; j1 = k1;
            }
        } else if (c + 1 < nchunks) {
            art_stage<V>(st, cstart, soff, idx, val, c + 1, art_buf(base, E, (c + 1) & 1), wave - 1, nw - 1);
        }
        __syncthreads();
    }
    for (int64_t j = threadIdx.x; j < n; j += ART_BLOCK) xg[j] = x[j];
}

// Chunk plan of one sweep: chunk boundaries (step indices) and each step's
// entry offset inside its chunk's buffer.
void art_plan_chunks(const std::vector<int>& steps, const std::vector<int64_t>& hp, int E, std::vector<int>& cstart,
                     std::vector<int>& soff) {
    cstart.assign(1, 0);
    soff.resize(steps.size());
    int used = 0, cnt = 0;
    for (size_t t = 0; t < steps.size(); ++t) {
        const int len = (int)(hp[(size_t)steps[t] + 1] - hp[(size_t)steps[t]]);
        if (cnt == ART_CMAX || used + len > E) {
            cstart.push_back((int)t);
            used = 0;
            cnt = 0;
        }
        soff[t] = used;
        used += len;
        cnt++;
    }
    if (!steps.empty()) cstart.push_back((int)steps.size());
}

// LDS entries per buffer for the chunked sweep, or 0 when x and two buffers
// holding the longest row do not fit
int art_chunk_entries(int64_t n, int64_t maxlen, int lim) {
    const int64_t room = (int64_t)lim - 8 * n - 2 * (int64_t)ART_CMAX * (int64_t)sizeof(ArtRec);
    int64_t E = std::min<int64_t>(4096, room / 24) & ~(int64_t)63;
    return (E >= 512 && E >= maxlen) ? (int)E : 0;
}

// ---------------------------------------------------------------- chunked CART sweep
// The column-action sweep with the residual r in LDS: wave 0 runs the
// column steps of one chunk while waves 1..15 stage the next chunk's column
// records (column, norm, the current x_j, bounds, flags) and entries.  Used
// when no column repeats in the order, so a staged x_j or flag is never
// changed by a step of the chunk being run; x, flag and nflag are written
// back to global memory by lane 0 as in the one-wave form.
constexpr int CART_CMAX = 64;

struct CartRec {
    double nrm, xj, lb, ub;
    int col, off, len, flag, nflag, pad;
};

struct CartBuf {
    CartRec* rec;
    double* val;
    int* idx;
};

__device__ __forceinline__ CartBuf cart_buf(char* base, int E, int which) {
    char* p = base + (size_t)which * ((size_t)CART_CMAX * sizeof(CartRec) + (size_t)E * 12);
    CartBuf b;
    b.rec = (CartRec*)p;
    b.val = (double*)(p + (size_t)CART_CMAX * sizeof(CartRec));
    b.idx = (int*)(p + (size_t)CART_CMAX * sizeof(CartRec) + (size_t)E * 8);
    return b;
}

template <typename V>
__device__ __forceinline__ void cart_stage(const int* __restrict__ cols, const int* __restrict__ cstart,
                                           const int* __restrict__ soff, const int64_t* __restrict__ cptr,
                                           const int* __restrict__ ridx, const V* __restrict__ val,
                                           const double* __restrict__ nrm, const double* xg,
                                           const unsigned char* flag, const int* nflag, const Bounds& bd, int c,
                                           CartBuf B, int wave, int nwaves) {
    const int lane = threadIdx.x & 63;
    const int t0 = cstart[c], cnt = cstart[c + 1] - t0;
    const int mine = cnt > wave ? (cnt - wave + nwaves - 1) / nwaves : 0;   // <= 64
    int64_t s0 = 0;
    int len = 0, off = 0;
    if (lane < mine) {
        const int q = wave + nwaves * lane;
        const int j = cols[t0 + q];
        CartRec r;
        r.nrm = nrm[j];
        r.xj = xg[j];
        r.lb = bd.lb_kind == AIR_BOUND_SCALAR ? bd.lb : (bd.lb_kind == AIR_BOUND_VECTOR ? bd.lbv[j] : 0.0);
        r.ub = bd.ub_kind == AIR_BOUND_SCALAR ? bd.ub : (bd.ub_kind == AIR_BOUND_VECTOR ? bd.ubv[j] : 0.0);
        r.col = j;
        r.off = soff[t0 + q];
        r.flag = flag[j];
        r.nflag = nflag[j];
        r.len = r.flag ? (int)(cptr[j + 1] - cptr[j]) : 0;     // entries only for columns that step
        r.pad = 0;
        B.rec[q] = r;
        s0 = cptr[j];
        len = r.len;
        off = r.off;
    }
    for (int u = 0; u < mine; ++u) {
        const int64_t sa = __shfl(s0, u);
        const int la = __shfl(len, u), oa = __shfl(off, u);
        for (int i = lane; i < la; i += 64) {
            B.idx[oa + i] = ridx[sa + i];
            B.val[oa + i] = to_f64(val[sa + i]);
        }
    }
}

template <typename V>
__global__ __launch_bounds__(ART_BLOCK) void k_cart_chunked(const int* __restrict__ cols, const int* __restrict__ cstart,
                                                            int nchunks, const int* __restrict__ soff,
                                                            const int64_t* __restrict__ cptr,
                                                            const int* __restrict__ ridx, const V* __restrict__ val,
                                                            const double* __restrict__ nrm, double* xg, double* rg,
                                                            unsigned char* flag, int* nflag, Bounds bd, int64_t n,
                                                            int64_t m, CartParams cp, int E) {
    extern __shared__ double smem[];
    double* r = smem;
    char* base = (char*)(smem + m);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = ART_BLOCK / 64;
    const bool track = cp.k >= cp.kbegin;
    for (int64_t i = threadIdx.x; i < m; i += ART_BLOCK) r[i] = rg[i];
    int64_t arg = -1;
    double mm = NAN;
    if (wave == 0 && track) mm = wave_absmax(xg, n, &arg);
    if (nchunks > 0)
        cart_stage<V>(cols, cstart, soff, cptr, ridx, val, nrm, xg, flag, nflag, bd, 0, cart_buf(base, E, 0), wave,
                      nw);
    __syncthreads();
    for (int c = 0; c < nchunks; ++c) {
        const CartBuf B = cart_buf(base, E, c & 1);
        if (wave == 0) {
            const int cnt = cstart[c + 1] - cstart[c];
            for (int q = 0; q < cnt; ++q) {
                const CartRec rc = B.rec[q];
                const int j = rc.col;
                if (rc.flag) {
                    double av[4], rv[4];
                    int iv[4];
                    bool on[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int e = lane + 64 * u;
                        on[u] = e < rc.len;
                        av[u] = on[u] ? B.val[rc.off + e] : 0.0;
                        iv[u] = on[u] ? B.idx[rc.off + e] : 0;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) rv[u] = r[iv[u]];
                    double dot = 0;
#pragma unroll
                    for (int u = 0; u < 4; ++u) dot += on[u] ? av[u] * rv[u] : 0.0;
                    for (int e = lane + 256; e < rc.len; e += 64) dot += B.val[rc.off + e] * r[B.idx[rc.off + e]];
                    dot = wave_sum_dpp(dot);
                    const double delta = dot / rc.nrm;
                    double od = cp.lam * delta;
                    const double xkj = rc.xj;
                    if (bd.lb_kind != AIR_BOUND_NONE && od < rc.lb - xkj) od = rc.lb - xkj;   // cart.m:251-259
                    if (bd.ub_kind != AIR_BOUND_NONE && od > rc.ub - xkj) od = rc.ub - xkj;   // :261-267
                    const double xn = xkj + od;
                    if (lane == 0) {
                        xg[j] = xn;
                        if (track && fabs(od) < cp.thr * mm) {                                // :273-276
                            flag[j] = 0;
                            nflag[j] = 1;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (on[u]) r[iv[u]] = rv[u] - od * av[u];                           // :279
                    for (int e = lane + 256; e < rc.len; e += 64) {
                        const int i = B.idx[rc.off + e];
                        r[i] = r[i] - od * B.val[rc.off + e];
                    }
                    if (track) {                                                          // mm = max(abs(xk))
                        const double an = fabs(xn);
                        if (an >= mm || isnan(mm)) { if (!isnan(an)) { mm = an; arg = j; } }
                        else if (j == arg) {
                            order_stores();                                               // x[j] written above
                            mm = wave_absmax(xg, n, &arg);
                        }
                    }
                } else if (lane == 0) {                                                   // :281-285
                    const int draw =
                        1 + (int)floor(rng_u01(cp.seed, RNG_UNFLAG, (uint64_t)cp.k, (uint64_t)j) * cp.nunflag);
                    if (rc.nflag < draw) nflag[j] = rc.nflag + 1;
                    else flag[j] = 1;
                }
            }
        } else if (c + 1 < nchunks) {
            cart_stage<V>(cols, cstart, soff, cptr, ridx, val, nrm, xg, flag, nflag, bd, c + 1,
                          cart_buf(base, E, (c + 1) & 1), wave - 1, nw - 1);
        }
        __syncthreads();
    }
    for (int64_t i = threadIdx.x; i < m; i += ART_BLOCK) rg[i] = r[i];
}

// LDS entries per buffer for the chunked CART sweep, or 0 when r and two
// buffers holding the longest column do not fit
int cart_chunk_entries(int64_t m, int64_t maxlen, int lim) {
    const int64_t room = (int64_t)lim - 8 * m - 2 * (int64_t)CART_CMAX * (int64_t)sizeof(CartRec);
    int64_t E = std::min<int64_t>(4096, room / 24) & ~(int64_t)63;
    return (E >= 256 && E >= maxlen) ? (int)E : 0;
}

void cart_plan_chunks(const std::vector<int>& cols, const std::vector<int64_t>& cp, int E, std::vector<int>& cstart,
                      std::vector<int>& soff) {
    cstart.assign(1, 0);
    soff.resize(cols.size());
    int used = 0, cnt = 0;
    for (size_t t = 0; t < cols.size(); ++t) {
        const int len = (int)(cp[(size_t)cols[t] + 1] - cp[(size_t)cols[t]]);
        if (cnt == CART_CMAX || used + len > E) {
            cstart.push_back((int)t);
            used = 0;
            cnt = 0;
        }
        soff[t] = used;
        used += len;
        cnt++;
    }
    if (!cols.empty()) cstart.push_back((int)cols.size());
}

void launch_art_chunked(const DevSparse& A, const ArtStep* st, const int* cstart, int nchunks, const int* soff,
                        double* x, Bounds bd, int64_t n, int paf, int E, hipStream_t s) {
    const size_t lds = 8 * (size_t)n + 2 * ((size_t)ART_CMAX * sizeof(ArtRec) + (size_t)E * 12);
    dispatch_dtype(A.dtype, [&](auto v) {
        using V = decltype(v);
        if (lds > 64 * 1024)
            AIR_HIP(hipFuncSetAttribute((const void*)k_art_chunked<V>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lds));
        k_art_chunked<V><<<1, ART_BLOCK, lds, s>>>(st, cstart, nchunks, soff, A.idx.as<int>(), A.val.as<V>(), x, bd,
                                                   n, paf, E);
        return 0;
    });
    AIR_HIP(hipGetLastError());
}

// pc: entries prefetched per step in units of 64 (2, 4 or 8)
void launch_art_sweep(const DevSparse& A, const ArtStep* st, int64_t L, double* x, Bounds bd, int64_t n, int paf,
                      bool lds, int pc, hipStream_t s) {
    dispatch_dtype(A.dtype, [&](auto v) {
        using V = decltype(v);
        if (lds) {
            if (pc <= 2) art_sweep_launch<V, true, 2>(A, st, L, x, bd, n, paf, s);
            else if (pc <= 4) art_sweep_launch<V, true, 4>(A, st, L, x, bd, n, paf, s);
            else art_sweep_launch<V, true, 8>(A, st, L, x, bd, n, paf, s);
        } else {
            if (pc <= 2) art_sweep_launch<V, false, 2>(A, st, L, x, bd, n, paf, s);
            else if (pc <= 4) art_sweep_launch<V, false, 4>(A, st, L, x, bd, n, paf, s);
            else art_sweep_launch<V, false, 8>(A, st, L, x, bd, n, paf, s);
        }
        return 0;
    });
}

// r = b - A x (FP64), one wave per row
template <typename V>
__global__ __launch_bounds__(256) void k_resid64(const int64_t* __restrict__ ptr, const int* __restrict__ idx,
                                                 const V* __restrict__ val, int64_t m, const double* __restrict__ x,
                                                 const double* __restrict__ b, double* __restrict__ r) {
    const int lane = threadIdx.x & 63;
    for (int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); i < m; i += (int64_t)gridDim.x * 4) {
        double acc = 0;
        for (int64_t q = ptr[i] + lane; q < ptr[i + 1]; q += 64) acc += to_f64(val[q]) * x[idx[q]];
        acc = wave_sum(acc);
        if (lane == 0) r[i] = b[i] - acc;
    }
}

// squared 2-norms of the compressed rows (rows of A from CSR, columns from CSC)
template <typename V>
__global__ __launch_bounds__(256) void k_sqnorm(const int64_t* __restrict__ ptr, const V* __restrict__ val,
                                                int64_t rows, double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    for (int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); i < rows; i += (int64_t)gridDim.x * 4) {
        double acc = 0;
        for (int64_t q = ptr[i] + lane; q < ptr[i + 1]; q += 64) {
            const double v = to_f64(val[q]);
            acc += v * v;
        }
        acc = wave_sum(acc);
        if (lane == 0) out[i] = acc;
    }
}

// out[0] = r'r, out[1] = p'(p + r) (the monotone-error numerator with p = r_{k-1})
__global__ __launch_bounds__(1024) void k_stop_dots(const double* __restrict__ r, const double* __restrict__ p,
                                                    int64_t m, double* __restrict__ out) {
    __shared__ double sh[2][16];
    double a = 0, c = 0;
    for (int64_t i = threadIdx.x; i < m; i += 1024) {
        a += r[i] * r[i];
        if (p) c += p[i] * (p[i] + r[i]);
    }
    a = wave_sum(a);
    c = wave_sum(c);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { sh[0][w] = a; sh[1][w] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0, sc = 0;
        for (int q = 0; q < 16; ++q) { sa += sh[0][q]; sc += sh[1][q]; }
        out[0] = sa;
        out[1] = sc;
    }
}

unsigned rows_grid(int64_t rows) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((rows + 3) / 4, 65536));
}

// ---------------------------------------------------------------- host side
struct StopState {
    int type = AIR_STOP_NONE;
    double taudelta = NAN;
    int kmax = 0;
    bool have_prev = false;     // ME: r_{k-1} exists (it is NaN at k = 0)
};

// check_stoprules.m:53-72,153-161 (DP and NONE); returns the info code or -1
int stop_code(const StopState& S, int k, double rr) {
    const double nrk = std::sqrt(rr);
    if (S.type == AIR_STOP_DP) {
        if (nrk <= S.taudelta || k >= S.kmax) return k != S.kmax ? 2 : 0;
        return -1;
    }
    return k >= S.kmax ? 0 : -1;
}

void check_order(const int32_t* order, int64_t norder, int64_t limit, const char* what) {
    for (int64_t q = 0; q < norder; ++q)
        AIR_CHECK(order[q] >= 1 && order[q] <= limit, AIR_E_ARG,
                  std::string("Index exceeds matrix dimensions: custom ") + what + " order entry out of range");
}

Bounds make_bounds(const air_row_options* o, int64_t n, DevBuf& lbd, DevBuf& ubd, hipStream_t s) {
    Bounds bd;
    bd.lb_kind = o->lb_kind;
    bd.ub_kind = o->ub_kind;
    bd.lb = o->lb;
    bd.ub = o->ub;
    if (o->lb_kind == AIR_BOUND_VECTOR) {
        AIR_CHECK(o->lb_vec, AIR_E_ARG, "lbound vector missing");
        lbd.alloc(8 * (size_t)n);
        AIR_HIP(hipMemcpyAsync(lbd.p, o->lb_vec, 8 * (size_t)n, hipMemcpyHostToDevice, s));
        bd.lbv = lbd.as<double>();
    }
    if (o->ub_kind == AIR_BOUND_VECTOR) {
        AIR_CHECK(o->ub_vec, AIR_E_ARG, "ubound vector missing");
        ubd.alloc(8 * (size_t)n);
        AIR_HIP(hipMemcpyAsync(ubd.p, o->ub_vec, 8 * (size_t)n, hipMemcpyHostToDevice, s));
        bd.ubv = ubd.as<double>();
    }
    return bd;
}

}  // namespace

// Shared driver of both methods: sweeps until a stop rule fires, saving the
// iterates of K (art.m:227-330, cart.m:211-323).
void row_action_run(const DevSparse& A, const DevSparse& AT, bool cart, int art_method, const int32_t* order,
                    int64_t norder, const air_row_options* o, const double* b_h, const double* x0_h,
                    const int32_t* K_in, int32_t nK, double* X_out, int32_t* itersaved, int32_t* n_itersaved,
                    double* resnorm, air_sirt_info* info, hipStream_t s) {
    auto t0 = std::chrono::steady_clock::now();
    AIR_CHECK(o && b_h, AIR_E_ARG, "null options or right-hand side");
    AIR_CHECK(K_in && nK > 0, AIR_E_ARG, "Max no. of iterations must ALWAYS be specified.");
    for (int q = 1; q < nK; ++q) AIR_CHECK(K_in[q] > K_in[q - 1], AIR_E_ARG, "Elements of K must be strictly increasing.");
    AIR_CHECK(o->damp >= 0, AIR_E_ARG, "Damping must be positive");
    AIR_CHECK(o->stoprule >= AIR_STOP_NONE && o->stoprule <= AIR_STOP_ME, AIR_E_ARG,
              "The chosen stopping rule is not valid.");
    const int64_t m = A.nrows, n = A.ncols;
    if (o->stoprule == AIR_STOP_NCP) {
        // check_stoprules.m:99-120: res_dims = [len, ncols] (2-D) or [m, 0] (1-D)
        const int64_t len = o->res_dims[0], cols = o->res_dims[1] > 0 ? o->res_dims[1] : 1;
        AIR_CHECK(len > 0 && len * cols == m, AIR_E_ARG,
                  "options.stoprule.res_dims must give the dimensions of the residual (prod(res_dims) == m).");
        AIR_CHECK(o->ncp_smooth >= 1 && o->ncp_smooth <= 4096, AIR_E_ARG, "ncp_smooth must be a positive integer.");
    }
    if (o->stoprule == AIR_STOP_DP || o->stoprule == AIR_STOP_ME)
        AIR_CHECK(!std::isnan(o->taudelta), AIR_E_ARG,
                  "For stopping rules DP and ME a scalar taudelta must be given in options.");
    std::vector<int32_t> K(K_in, K_in + nK);
    const bool sym = !cart && art_method == AIR_ART_SYMKACZMARZ;
    if (sym) {                                                      // art.m:123-135
        for (int q = 0; q < nK; ++q)
            AIR_CHECK(K[q] % 2 == 0, AIR_E_ARG, "For symkaczmarz only even iteration numbers can be requested.");
        for (auto& k : K) k /= 2;
    }
    const int kmax = K[nK - 1];

    // norms of the rows (ART) or columns (CART), art.m:161-181 / cart.m:154-174
    const DevSparse& S = cart ? AT : A;
    const int64_t L0 = S.nrows;
    DevBuf dnrm;
    dnrm.alloc(8 * (size_t)std::max<int64_t>(L0, 1));
    dispatch_dtype(S.dtype, [&](auto v) {
        using V = decltype(v);
        if (L0) k_sqnorm<V><<<rows_grid(L0), 256, 0, s>>>(S.ptr.as<int64_t>(), S.val.as<V>(), L0, dnrm.as<double>());
        return 0;
    });
    AIR_HIP(hipGetLastError());
    std::vector<double> nrm((size_t)L0);
    AIR_HIP(hipMemcpyAsync(nrm.data(), dnrm.p, 8 * (size_t)L0, hipMemcpyDeviceToHost, s));
    AIR_HIP(hipStreamSynchronize(s));

    // the step order: nonzero rows/columns, or the caller's order (1-based)
    std::vector<int> I;
    const bool custom = cart ? (order != nullptr) : (art_method == AIR_ART_CUSTOM);
    if (custom) {
        AIR_CHECK(order && norder >= 0, AIR_E_ARG, "custom order missing");
        check_order(order, norder, L0, cart ? "column" : "row");
        for (int64_t q = 0; q < norder; ++q)
            if (nrm[order[q] - 1] > 0) I.push_back(order[q] - 1);
    } else {
        AIR_CHECK(cart || art_method == AIR_ART_KACZMARZ || art_method == AIR_ART_SYMKACZMARZ ||
                      art_method == AIR_ART_RANDKACZMARZ,
                  AIR_E_ARG, "Unknown ART method specified");
        for (int64_t q = 0; q < L0; ++q)
            if (nrm[q] > 0) I.push_back((int)q);
    }
    if (sym) {
        const size_t h = I.size();
        for (size_t q = 0; q < h; ++q) I.push_back(I[h - 1 - q]);
    }
    // damping, art.m:206 / cart.m:193
    double mx = 0;
    for (double v : nrm) mx = std::max(mx, v);
    for (double& v : nrm) v = v + o->damp * mx;
    AIR_HIP(hipMemcpyAsync(dnrm.p, nrm.data(), 8 * (size_t)L0, hipMemcpyHostToDevice, s));
    const bool rnd = !cart && art_method == AIR_ART_RANDKACZMARZ;
    std::vector<double> cumul;
    bool fast = true;
    if (rnd) {                                                      // art.m:209-219
        double tot = 0;
        for (double v : nrm) tot += v;
        cumul.resize((size_t)L0);
        double c = 0, dev = 0;
        for (int64_t q = 0; q < L0; ++q) {
            c += nrm[q] / tot;
            cumul[q] = c;
            dev = std::max(dev, std::fabs(c - (double)(q + 1) / (double)L0));
        }
        fast = dev < 0.05;
    }

    // relaxation, calc_relaxpar.m:54-87
    double lam = cart ? 0.25 : 1.0;
    if (o->relax_kind == AIR_RELAX_SCALAR) lam = o->relaxpar;
    else if (o->relax_kind == AIR_RELAX_VECTOR) {
        AIR_CHECK(!cart, AIR_E_ARG, "For CART methods the relaxpar must be a scalar.");
        AIR_CHECK(o->relax_vec, AIR_E_ARG, "relaxation values missing");
    } else
        AIR_CHECK(o->relax_kind == AIR_RELAX_DEFAULT, AIR_E_ARG,
                  "For ART methods the relaxpar must be a scalar or function handle.");

    // device state
    DevBuf db, dx, dr, dp, dlbv, dubv, dsteps, dlam, ddots, dflag, dnflag;
    db.alloc(8 * (size_t)std::max<int64_t>(m, 1));
    dx.alloc(8 * (size_t)std::max<int64_t>(n, 1));
    dr.alloc(8 * (size_t)std::max<int64_t>(m, 1));
    ddots.alloc(16);
    AIR_HIP(hipMemcpyAsync(db.p, b_h, 8 * (size_t)m, hipMemcpyHostToDevice, s));
    if (x0_h) AIR_HIP(hipMemcpyAsync(dx.p, x0_h, 8 * (size_t)n, hipMemcpyHostToDevice, s));
    else AIR_HIP(hipMemsetAsync(dx.p, 0, 8 * (size_t)n, s));
    const Bounds bd = make_bounds(o, n, dlbv, dubv, s);
    const size_t L = I.size();
    dsteps.alloc(4 * std::max<size_t>(L, 1));
    if (!rnd) AIR_HIP(hipMemcpyAsync(dsteps.p, I.data(), 4 * L, hipMemcpyHostToDevice, s));
    const bool lam_vec = !cart && o->relax_kind == AIR_RELAX_VECTOR;
    if (lam_vec) dlam.alloc(8 * std::max<size_t>(L, 1));
    // ART: per-step records, x in LDS when it fits one work-group's LDS
    // (AIR_ART_LDS=0 keeps it in global memory), entries prefetched per step
    DevBuf dinfo, dcst, dsoff;
    bool use_lds = false;
    int pc = 2, chunkE = 0;
    std::vector<int64_t> hrowptr;
    std::vector<int> cst, soff;
    auto upload_plan = [&](const std::vector<int>& order_of_steps) {
        art_plan_chunks(order_of_steps, hrowptr, chunkE, cst, soff);
        dcst.alloc(4 * cst.size());
        dsoff.alloc(4 * std::max<size_t>(soff.size(), 1));
        AIR_HIP(hipMemcpyAsync(dcst.p, cst.data(), 4 * cst.size(), hipMemcpyHostToDevice, s));
        if (!soff.empty()) AIR_HIP(hipMemcpyAsync(dsoff.p, soff.data(), 4 * soff.size(), hipMemcpyHostToDevice, s));
    };
    if (!cart) {
        dinfo.alloc(sizeof(ArtStep) * std::max<size_t>(L, 1));
        int dev = 0, lim = 0;
        AIR_HIP(hipGetDevice(&dev));
        AIR_HIP(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const char* env = std::getenv("AIR_ART_LDS");
        use_lds = 8 * n <= (int64_t)lim && !(env && env[0] == '0');
        std::vector<int64_t> hp((size_t)m + 1);
        AIR_HIP(hipMemcpyAsync(hp.data(), A.ptr.p, 8 * ((size_t)m + 1), hipMemcpyDeviceToHost, s));
        AIR_HIP(hipStreamSynchronize(s));
        int64_t maxlen = 0;
        for (int64_t q = 0; q < m; ++q) maxlen = std::max(maxlen, hp[q + 1] - hp[q]);
        pc = maxlen <= 128 ? 2 : (maxlen <= 256 ? 4 : 8);
        // rows staged through LDS by the rest of a 1024-thread work-group
        // (AIR_ART_CHUNK=0 keeps the one-wave form)
        const char* envc = std::getenv("AIR_ART_CHUNK");
        chunkE = use_lds && !(envc && envc[0] == '0') ? art_chunk_entries(n, maxlen, lim) : 0;
        hrowptr.swap(hp);
    }
    int cartE = 0;
    if (cart) {
        dflag.alloc((size_t)std::max<int64_t>(n, 1));
        dnflag.alloc(4 * (size_t)std::max<int64_t>(n, 1));
        AIR_HIP(hipMemsetAsync(dflag.p, 1, (size_t)n, s));
        AIR_HIP(hipMemsetAsync(dnflag.p, 0, 4 * (size_t)n, s));
        // r in LDS with the columns staged by the rest of a 1024-thread work-group,
        // when r and two column buffers fit and no column repeats in the order
        // (AIR_ART_CHUNK=0 keeps the one-wave form)
        const char* envc = std::getenv("AIR_ART_CHUNK");
        std::vector<char> seen((size_t)std::max<int64_t>(n, 1), 0);
        bool repeats = false;
        for (int j : I) {
            repeats = repeats || seen[(size_t)j];
            seen[(size_t)j] = 1;
        }
        if (!(envc && envc[0] == '0') && !repeats && L) {
            int dev = 0, lim = 0;
            AIR_HIP(hipGetDevice(&dev));
            AIR_HIP(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
            hrowptr.resize((size_t)n + 1);
            AIR_HIP(hipMemcpyAsync(hrowptr.data(), AT.ptr.p, 8 * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
            AIR_HIP(hipStreamSynchronize(s));
            int64_t maxlen = 0;
            for (int64_t q = 0; q < n; ++q) maxlen = std::max(maxlen, hrowptr[q + 1] - hrowptr[q]);
            cartE = cart_chunk_entries(m, maxlen, lim);
            if (cartE > 0) {
                cart_plan_chunks(I, hrowptr, cartE, cst, soff);
                dcst.alloc(4 * cst.size());
                dsoff.alloc(4 * std::max<size_t>(soff.size(), 1));
                AIR_HIP(hipMemcpyAsync(dcst.p, cst.data(), 4 * cst.size(), hipMemcpyHostToDevice, s));
                AIR_HIP(hipMemcpyAsync(dsoff.p, soff.data(), 4 * soff.size(), hipMemcpyHostToDevice, s));
                AIR_HIP(hipStreamSynchronize(s));
            }
        }
    }
    auto residual = [&]() {
        dispatch_dtype(A.dtype, [&](auto v) {
            using V = decltype(v);
            if (m) k_resid64<V><<<rows_grid(m), 256, 0, s>>>(A.ptr.as<int64_t>(), A.idx.as<int>(), A.val.as<V>(), m,
                                                          dx.as<double>(), db.as<double>(), dr.as<double>());
            return 0;
        });
        AIR_HIP(hipGetLastError());
    };
    StopState ss;
    ss.type = o->stoprule;
    ss.taudelta = o->taudelta;
    ss.kmax = kmax;
    if (ss.type == AIR_STOP_ME) dp.alloc(8 * (size_t)std::max<int64_t>(m, 1));
    // NCP: the per-column statistics come from the SIRT engine's transform
    // (FP64 residual copy, batched D2Z, one block per column); the filter
    // window dk lives on the host, check_stoprules.m:96-151
    NcpState ncp;
    DevBuf dctrl;
    std::vector<double> dk, colstat;
    if (ss.type == AIR_STOP_NCP) {
        const int len = o->res_dims[0], cols = o->res_dims[1] > 0 ? o->res_dims[1] : 1;
        ncp_prepare(ncp, m, len, cols, o->ncp_smooth, s);
        dctrl.alloc(sizeof(Ctrl));
        AIR_HIP(hipMemsetAsync(dctrl.p, 0, sizeof(Ctrl), s));        // stop = 0: the statistics kernel runs
        dk.assign((size_t)o->ncp_smooth, INFINITY);                  // check_inputs.m:207-209
        colstat.resize((size_t)cols);
    }
    // evaluates the stop rule at sweep k on the residual in dr; returns the code or -1
    auto check = [&](int k, double* nrk_out) {
        k_stop_dots<<<1, 1024, 0, s>>>(dr.as<double>(), ss.have_prev ? dp.as<double>() : nullptr, m,
                                         ddots.as<double>());
        AIR_HIP(hipGetLastError());
        double d[2];
        double pp = 0;
        AIR_HIP(hipMemcpyAsync(d, ddots.p, 16, hipMemcpyDeviceToHost, s));
        if (ss.type == AIR_STOP_ME && ss.have_prev) {
            k_stop_dots<<<1, 1024, 0, s>>>(dp.as<double>(), nullptr, m, ddots.as<double>());
            AIR_HIP(hipGetLastError());
            AIR_HIP(hipMemcpyAsync(&pp, ddots.p, 8, hipMemcpyDeviceToHost, s));
        }
        if (ss.type == AIR_STOP_NCP) {
            ncp_enqueue_f64(ncp, dr.as<double>(), m, (const Ctrl*)dctrl.p, s);
            AIR_HIP(hipMemcpyAsync(colstat.data(), ncp.colstat.p, 8 * colstat.size(), hipMemcpyDeviceToHost, s));
        }
        AIR_HIP(hipStreamSynchronize(s));
        *nrk_out = std::sqrt(d[0]);
        int code = -1;
        if (ss.type == AIR_STOP_NCP) {                             // check_stoprules.m:129-151
            double ncc = 0;                                        // mean over the columns
            for (double v : colstat) ncc += v;
            ncc /= (double)colstat.size();
            double mx_old = -INFINITY, mx_new = ncc;               // max ignores NaN, as fmax does
            for (size_t q = 0; q < dk.size(); ++q) mx_old = std::fmax(mx_old, dk[q]);
            for (size_t q = 1; q < dk.size(); ++q) mx_new = std::fmax(mx_new, dk[q]);
            if (mx_old < mx_new || k >= kmax) code = k != kmax ? 1 : 0;
            else {
                dk.erase(dk.begin());
                dk.push_back(ncc);
            }
        } else if (ss.type == AIR_STOP_ME) {                       // check_stoprules.m:74-94
            const double dme = ss.have_prev ? d[1] / 2 : NAN;
            const double nr1 = ss.have_prev ? std::sqrt(pp) : NAN;
            if (dme / nr1 <= ss.taudelta) code = 3;
            else if (k >= kmax) code = 0;
            else {
                AIR_HIP(hipMemcpyAsync(dp.p, dr.p, 8 * (size_t)m, hipMemcpyDeviceToDevice, s));
                ss.have_prev = true;
            }
        } else {
            code = stop_code(ss, k, d[0]);
        }
        return code;
    };

    residual();                                                     // art.m:147 / cart.m:139
    double nrk = 0;
    int code = check(0, &nrk);
    if (resnorm) resnorm[0] = nrk;
    int k = 0, l = 0, fin = 0;
    double lam_last = lam;
    std::vector<int> steps(L);
    std::vector<double> lams(lam_vec ? L : 0);
    if (chunkE > 0 && !rnd && L) upload_plan(I);
    const int nunflag = o->Nunflag >= 1 ? o->Nunflag : std::max(1, (int)std::lround(kmax / 4.0));
    auto ti = std::chrono::steady_clock::now();
    while (code < 0) {
        ++k;
        if (rnd) {                                                  // art.m:242-262
            std::vector<int> perm(I);
            for (int64_t t = (int64_t)L - 1; t >= 1; --t) {
                const int64_t j = (int64_t)std::floor(rng_u01(o->seed, RNG_PERM, (uint64_t)k, (uint64_t)t) * (t + 1));
                std::swap(perm[t], perm[j]);
            }
            for (size_t t = 0; t < L; ++t) {
                if (fast) steps[t] = perm[t];
                else {
                    const double u = rng_u01(o->seed, RNG_SAMPLE, (uint64_t)k, (uint64_t)t);
                    const int64_t ri = std::lower_bound(cumul.begin(), cumul.end(), u) - cumul.begin();
                    steps[t] = (int)std::min<int64_t>(ri, L0 - 1);   // sum(cumul<rand)+1, kept in range
                }
                if (lam_vec) lams[t] = o->relax_vec[(size_t)(k - 1) * m + perm[t]];
            }
            AIR_HIP(hipMemcpyAsync(dsteps.p, steps.data(), 4 * L, hipMemcpyHostToDevice, s));
        } else if (lam_vec) {
            for (size_t t = 0; t < L; ++t) lams[t] = o->relax_vec[(size_t)(k - 1) * m + I[t]];   // relaxpar((k-1)*m+i)
        }
        if (lam_vec && L) {
            AIR_HIP(hipMemcpyAsync(dlam.p, lams.data(), 8 * L, hipMemcpyHostToDevice, s));
            lam_last = lams[L - 1];
        }
        if (cart) {
            CartParams cp{lam, o->THR, k, o->Kbegin, nunflag, o->seed};
            dispatch_dtype(AT.dtype, [&](auto v) {
                using V = decltype(v);
                if (L && cartE > 0) {
                    const size_t lds = 8 * (size_t)m + 2 * ((size_t)CART_CMAX * sizeof(CartRec) + (size_t)cartE * 12);
                    if (lds > 64 * 1024)
                        AIR_HIP(hipFuncSetAttribute((const void*)k_cart_chunked<V>,
                                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                    k_cart_chunked<V><<<1, ART_BLOCK, lds, s>>>(dsteps.as<int>(), dcst.as<int>(), (int)cst.size() - 1,
                                                                dsoff.as<int>(), AT.ptr.as<int64_t>(), AT.idx.as<int>(),
                                                                AT.val.as<V>(), dnrm.as<double>(), dx.as<double>(),
                                                                dr.as<double>(), dflag.as<unsigned char>(),
                                                                dnflag.as<int>(), bd, n, m, cp, cartE);
                } else if (L) {
                    k_cart_sweep<V><<<1, 64, 0, s>>>(AT.ptr.as<int64_t>(), AT.idx.as<int>(), AT.val.as<V>(),
                                                     dsteps.as<int>(), (int64_t)L, dnrm.as<double>(), dx.as<double>(),
                                                     dr.as<double>(), dflag.as<unsigned char>(), dnflag.as<int>(), bd,
                                                     n, cp);
                }
                return 0;
            });
        } else {
            if (L) {
                k_art_steps<<<(unsigned)std::min<int64_t>(((int64_t)L + 255) / 256, 4096), 256, 0, s>>>(
                    dsteps.as<int>(), (int64_t)L, A.ptr.as<int64_t>(), db.as<double>(), dnrm.as<double>(),
                    lam_vec ? dlam.as<double>() : nullptr, lam, dinfo.as<ArtStep>());
                AIR_HIP(hipGetLastError());
                if (chunkE > 0) {
                    if (rnd) upload_plan(steps);
                    launch_art_chunked(A, dinfo.as<ArtStep>(), dcst.as<int>(), (int)cst.size() - 1, dsoff.as<int>(),
                                       dx.as<double>(), bd, n, k == 1 ? 1 : 0, chunkE, s);
                } else {
                    launch_art_sweep(A, dinfo.as<ArtStep>(), (int64_t)L, dx.as<double>(), bd, n, k == 1 ? 1 : 0,
                                     use_lds, pc, s);
                }
            }
            residual();                                             // art.m:288
        }
        AIR_HIP(hipGetLastError());
        code = check(k, &nrk);
        if (resnorm) resnorm[k] = nrk;
        if ((l < nK && k == K[l]) || code >= 0) {                 // art.m:295-298
            if (X_out) AIR_HIP(hipMemcpyAsync(X_out + (size_t)l * n, dx.p, 8 * (size_t)n, hipMemcpyDeviceToHost, s));
            l++;
        }
    }
    AIR_HIP(hipStreamSynchronize(s));
    fin = k;
    std::vector<int32_t> Ks(K);
    if (sym) {                                                      // art.m:321-324
        fin *= 2;
        for (auto& q : Ks) q *= 2;
    }
    int ns = 0;                                                     // art.m:327
    for (int q = 0; q < nK; ++q)
        if (Ks[q] < fin) {
            if (itersaved) itersaved[ns] = Ks[q];
            ns++;
        }
    if (itersaved) itersaved[ns] = fin;
    if (n_itersaved) *n_itersaved = ns + 1;
    if (info) {
        air_sirt_info inf{};
        inf.stoprule = code;
        inf.finaliter = fin;
        inf.relaxpar = lam_vec ? lam_last : lam;
        inf.rho = NAN;
        inf.nsaved = l;
        inf.iter_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - ti).count();
        inf.timetaken = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        inf.setup_time = inf.timetaken - inf.iter_time;
        *info = inf;
    }
}

}  // namespace air
