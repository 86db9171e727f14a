// air_spmm.hip -- SIRT with up to 64 right-hand sides on one stored operand
// (sirt/sirt.m:199-261 run for many b at once), both products on operand rows
// staged in LDS.
//
// Vectors are RHS-minor: row i of X, B or M R is 64 FP32 values (256 bytes),
// one per right-hand side.  Both products use one kernel shape, a 512-thread
// work-group (8 waves) that owns 512 accumulation targets and walks a list of
// operand chunks of up to 256 rows (64 KB of LDS); with the 16 KB of record
// rings that is 80 KB, so two work-groups share a CU:
//   - a wave holds 64 targets in 16 rounds of 4 (target i = 32 t + 4 w + g:
//     round t, wave w, group g, so that every wave has targets spread over the
//     whole work-group); its lanes form four groups of 16, group g works on
//     target g of a round and lane q of a group holds right-hand sides
//     4q..4q+3, so every entry is one 16-byte LDS read per lane, the group
//     reading one whole 256-byte operand row (no bank conflicts), and four FMAs;
//   - the 16 rounds of accumulators (64 VGPRs) stay in registers over all
//     chunks of the work-group (about 125 VGPRs in all, the budget of 4 waves
//     per SIMD);
//   - a round's entries in one chunk are stored in records of 4 steps x 4
//     groups (80 bytes: 16 FP32 values, then 16 one-byte operand slots); the
//     round's step count is the largest of its four targets' entry counts and
//     the shorter targets read zero entries;
//   - every wave reads its own records, in chunk and round order, from one
//     contiguous stream of 1 KB blocks (12 records each) through a two-block
//     ring in LDS, the next block waiting in four VGPRs; waves meet only at
//     chunk changes.  The operand chunk has one buffer: at a chunk change
//     the waves meet at a barrier, the next chunk is loaded by DMA
//     (global_load_lds) and waited for, and the CU's other work-group computes
//     meanwhile.
//
//   A X    targets are pieces (one row's entries inside one 64 x 64 column tile
//          of the tiled layout, in the tiled layout's piece order); a
//          work-group takes 512 consecutive pieces of one tile and its chunks
//          are the tile's sixteen 16 x 16 sub-tiles of X.  It writes one
//          256-byte partial per piece; the row pass sums a row's partials in
//          tile order and forms r = b - A x, M r and the FP64 norms per RHS.
//   A'(M R) targets are a block of 512 consecutive internal columns; its chunks are
//          consecutive runs of 256 of the block's distinct rows of M R; the
//          epilogue is the update x <- P(x + lam D g) with scalar bounds.
//
// The layouts are built on the GPU (rocPRIM radix sorts of the entries by
// target and chunk); the per-wave streams and the work-group tables are
// planned on the host from the per-cell counts and validated there against
// every bound the kernel relies on.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "air_internal.h"
#include "air_device.h"

namespace air {

namespace {

constexpr int SP_THREADS = 512;
constexpr int SP_WAVES = 8;
constexpr int SP_ROUNDS = 16;
constexpr int SP_TARGETS = SP_WAVES * SP_ROUNDS * 4;     // 512
constexpr int SP_CELLS = SP_WAVES * SP_ROUNDS;           // 128 (wave, round) cells per chunk
constexpr int SP_BB = 9;                                 // log2 of the columns of a backward block
static_assert((1 << SP_BB) == SP_TARGETS, "a backward block has one column per target");
constexpr int SP_ROWS = 256;                             // operand rows per chunk
constexpr int SP_RHS = 64;
constexpr int SP_REC = 80;                               // bytes per record
constexpr int SP_BLK = 1024;                             // bytes per stream block
constexpr int SP_BLK_RECS = 12;                          // records per stream block
constexpr int SP_OP_BYTES = SP_ROWS * SP_RHS * 4;        // 65536
constexpr int SP_FT = 4096;                              // columns per 64 x 64 tile of the tiled layout
constexpr int NXCD = 8;

static_assert(SP_BLK_RECS * SP_REC <= SP_BLK, "a block holds its records");
static_assert(SP_OP_BYTES / 16 == 8 * SP_THREADS, "the operand is eight 16-byte DMAs per thread");

struct SpEpi {
    float* out;          // forward: partial [pieces][64]; backward: X [n_int][64]
    const float* D;      // backward: per internal column, nullptr for 1
    double lam;
    int lb_kind, ub_kind;
    float lb, ub;
};

__device__ __forceinline__ void fma4(float4& a, float v, const float4& x) {
    a.x = fmaf(v, x.x, a.x);
    a.y = fmaf(v, x.y, a.y);
    a.z = fmaf(v, x.z, a.z);
    a.w = fmaf(v, x.w, a.w);
}

// LDS byte offset of lane q's 16 bytes in the operand row named by byte J of
// the slot word: one v_perm_b32 builds [qoff.b0, slot J, 0, 0] (selector
// bytes 0-3 pick bytes of the second source, 4-7 of the first, 0x0c gives
// zero); qoff holds 16 q.
template <int J>
__device__ __forceinline__ uint32_t sp_addr(uint32_t sl, uint32_t qoff) {
    return __builtin_amdgcn_perm(sl, qoff, 0x0c0c0000u | ((4u + J) << 8));
}

// NS steps of one record: all operand rows are read before the FMAs.
template <int NS>
__device__ __forceinline__ void sp_steps(float4& a, const float4& v, uint32_t sl, uint32_t qoff,
                                         const unsigned char* op) {
    float4 x[NS];
    x[0] = *(const float4*)(op + sp_addr<0>(sl, qoff));
    if constexpr (NS > 1) x[1] = *(const float4*)(op + sp_addr<1>(sl, qoff));
    if constexpr (NS > 2) x[2] = *(const float4*)(op + sp_addr<2>(sl, qoff));
    if constexpr (NS > 3) x[3] = *(const float4*)(op + sp_addr<3>(sl, qoff));
    fma4(a, v.x, x[0]);
    if constexpr (NS > 1) fma4(a, v.y, x[1]);
    if constexpr (NS > 2) fma4(a, v.z, x[2]);
    if constexpr (NS > 3) fma4(a, v.w, x[3]);
}

typedef __attribute__((address_space(3))) void lds_void;

// LDS of the chunk kernel, 80 KB so that two work-groups share a CU (one
// computes while the other waits at a barrier or for its operand): the
// record rings of the 8 waves (2 KB each), then the operand chunk.
__shared__ __align__(64) unsigned char sp_ring[SP_WAVES * 2 * SP_BLK];
__shared__ __align__(16) unsigned char sp_op[SP_OP_BYTES];

// 16 bytes per lane from global memory straight into LDS: lane l of the wave
// writes dst + 16 l (dst is the same for the whole wave).
__device__ __forceinline__ void dma16(const void* src, unsigned char* dst) {
    __builtin_amdgcn_global_load_lds(src, (lds_void*)dst, 16, 0, 0);
}

// operand chunk {first row, rows}: wave w, DMA k covers rows 32 w + 4 k .. + 3
// (row 32 w + 4 k + lane / 16, columns 4 (lane % 16) ..)
template <int MODE>
__device__ __forceinline__ void sp_dma_operand(const int4& ch, const int* __restrict__ rowlist,
                                               const float* __restrict__ Y, int w, int lane) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int r = w * 32 + k * 4 + (lane >> 4);
        if (r < ch.y) {
            const int64_t row = MODE == 0 ? (int64_t)ch.x + r : (int64_t)rowlist[(int64_t)ch.x + r];
            dma16(Y + row * SP_RHS + (lane & 15) * 4, sp_op + (w * 32 + k * 4) * 256);
        }
    }
}

// One wave's position in its record stream: block blk (ring slot blk & 1),
// record `within` of it; `st` holds block blk + 2 once loaded.
struct SpStream {
    const uint4* src;      // the wave's first block
    int nb;                // blocks of the wave
    int blk, within;
    uint4 st;
    bool sw;               // a block change happened in the current chunk
    uint32_t off;          // ring offset of the current record
    uint32_t sl;           // its slot word for group g (read one record ahead)
};

// Advance to the next record and return its byte offset in the wave's ring.
// On a block change the staged block goes into the ring slot just left and
// the block after it is requested.
__device__ __forceinline__ uint32_t sp_next(SpStream& S, unsigned char* ring, int lane) {
    if (S.within == SP_BLK_RECS) {
        S.within = 0;
        ++S.blk;
        if (S.blk + 1 < S.nb) *(uint4*)(ring + ((S.blk + 1) & 1) * SP_BLK + lane * 16) = S.st;
        if (S.blk + 2 < S.nb) S.st = S.src[(int64_t)(S.blk + 2) * (SP_BLK / 16) + lane];
        S.sw = true;
    }
    const uint32_t off = (uint32_t)(S.blk & 1) * SP_BLK + (uint32_t)S.within * SP_REC;
    ++S.within;
    return off;
}

// The current record: its slot word (read one record ahead) and its values
// (read now); then move to the next record and read its slot word, so that
// the operand reads of a record depend on nothing read in the same step.
// Past the wave's last record the read touches stale ring bytes, never used.
__device__ __forceinline__ void sp_take(SpStream& S, unsigned char* ring, int lane, float4& v, uint32_t& sl) {
    const int g = lane >> 4;
    sl = S.sl;
    v = *(const float4*)(ring + S.off + g * 16);
    S.off = sp_next(S, ring, lane);
    S.sl = *(const uint32_t*)(ring + S.off + 64 + g * 4);
}

// One chunk: this wave's 16 rounds on the staged operand, then, if another
// chunk follows, a barrier (every wave is done with the operand), the DMA of
// the next operand, a wait for it and a second barrier.
template <int MODE>
__device__ __forceinline__ void sp_chunk(float4 (&acc)[SP_ROUNDS], SpStream& S, unsigned char* ring,
                                         const uint32_t* __restrict__ cw, bool has_next, int4 next,
                                         const int* rowlist, const float* Y, int w, int lane) {
    const int q = lane & 15;
    uint32_t cnt[SP_ROUNDS / 2];
#pragma unroll
    for (int i = 0; i < SP_ROUNDS / 2; ++i) cnt[i] = __builtin_amdgcn_readfirstlane(cw[i]);
    const uint32_t qoff = (uint32_t)q << 4;
#pragma unroll
    for (int t = 0; t < SP_ROUNDS; ++t) {
        const uint32_t C = (cnt[t >> 1] >> (16 * (t & 1))) & 0xffffu;
        if (C == 0) continue;
        const uint32_t nb = C >> 2, rem = C & 3u;
        float4 a = acc[t];
        for (uint32_t b = 0; b < nb; ++b) {
            float4 v;
            uint32_t sl;
            sp_take(S, ring, lane, v, sl);
            sp_steps<4>(a, v, sl, qoff, sp_op);
        }
        if (rem) {
            float4 v;
            uint32_t sl;
            sp_take(S, ring, lane, v, sl);
            if (rem == 1) sp_steps<1>(a, v, sl, qoff, sp_op);
            else if (rem == 2) sp_steps<2>(a, v, sl, qoff, sp_op);
            else sp_steps<3>(a, v, sl, qoff, sp_op);
        }
        acc[t] = a;
    }
    if (has_next) {
        __syncthreads();
        sp_dma_operand<MODE>(next, rowlist, Y, w, lane);
        __builtin_amdgcn_s_waitcnt(0x0F70);    // vmcnt(0): this wave's DMA (and its ring block load)
        __syncthreads();
    }
}

// A chunk descriptor in scalar registers (zeros when absent).
__device__ __forceinline__ int4 sp_chunk_desc(const int4* __restrict__ chd, int c, bool present) {
    int4 v = make_int4(0, 0, 0, 0);
    if (present) {
        const int4 r = chd[c];
        v = make_int4(__builtin_amdgcn_readfirstlane(r.x), __builtin_amdgcn_readfirstlane(r.y), 0, 0);
    }
    return v;
}

// Launch entry b runs work-group lperm[b] (-1: idle).  Work-group table
// {first chunk, chunks, A, B}: forward A = first piece, B = pieces (<= 512);
// backward A = the block of 512 internal columns, B unused.
// Per work-group and wave: first stream block and block count.  Per chunk:
// {first operand row (forward: internal column; backward: offset into the
// row list), rows}; per (chunk, wave): 16 round step counts (uint16).
template <int MODE>
__global__ __launch_bounds__(SP_THREADS, 4) void k_spmm(const int* __restrict__ lperm, const int4* __restrict__ wgt,
                                                     const int64_t* __restrict__ wblk, const int* __restrict__ wnb,
                                                     const int4* __restrict__ chd, const uint32_t* __restrict__ ctab,
                                                     const uint4* __restrict__ stream, const int* __restrict__ rowlist,
                                                     const float* __restrict__ Y, SpEpi ep) {
    const int wg = lperm[blockIdx.x];
    if (wg < 0) return;
    const int4 d = wgt[wg];
    const int chunk0 = d.x, nch = d.y;
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63, g = lane >> 4, q = lane & 15;
    unsigned char* ring = sp_ring + w * (2 * SP_BLK);

    float4 acc[SP_ROUNDS];
#pragma unroll
    for (int t = 0; t < SP_ROUNDS; ++t) acc[t] = make_float4(0.f, 0.f, 0.f, 0.f);

    if (nch > 0) {
        SpStream S;
        S.src = stream + wblk[wg * SP_WAVES + w] * (SP_BLK / 16);
        S.nb = wnb[wg * SP_WAVES + w];
        S.blk = 0;
        S.within = 0;
        S.sw = false;
        S.st = make_uint4(0u, 0u, 0u, 0u);
        if (S.nb > 0) *(uint4*)(ring + lane * 16) = S.src[lane];
        if (S.nb > 1) *(uint4*)(ring + SP_BLK + lane * 16) = S.src[SP_BLK / 16 + lane];
        if (S.nb > 2) S.st = S.src[2 * (SP_BLK / 16) + lane];
        sp_dma_operand<MODE>(sp_chunk_desc(chd, chunk0, true), rowlist, Y, w, lane);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        S.off = sp_next(S, ring, lane);     // the first record: its slot word ahead
        S.sl = *(const uint32_t*)(ring + S.off + 64 + g * 4);
        for (int c = 0; c < nch; ++c) {
            const int4 n1 = sp_chunk_desc(chd, chunk0 + c + 1, c + 1 < nch);
            sp_chunk<MODE>(acc, S, ring, ctab + ((int64_t)(chunk0 + c) * SP_WAVES + w) * (SP_ROUNDS / 2), c + 1 < nch,
                           n1, rowlist, Y, w, lane);
        }
    }

    if constexpr (MODE == 0) {
        const uint32_t p0 = (uint32_t)d.z, np = (uint32_t)d.w;
#pragma unroll
        for (int t = 0; t < SP_ROUNDS; ++t) {
            const uint32_t i = (uint32_t)(t * (4 * SP_WAVES) + w * 4 + g);
            if (i < np) *(float4*)(ep.out + (int64_t)(p0 + i) * SP_RHS + q * 4) = acc[t];
        }
    } else {
        const int64_t G = d.z;
#pragma unroll
        for (int t = 0; t < SP_ROUNDS; ++t) {
            const uint32_t i = (uint32_t)(t * (4 * SP_WAVES) + w * 4 + g);
            const int64_t px = G * SP_TARGETS + i;
            float4* xp = (float4*)(ep.out + px * SP_RHS + q * 4);
            const float4 xo = *xp;
            const double dd = ep.D ? (double)ep.D[px] : 1.0;
            float4 xn;
            xn.x = project((double)xo.x + ep.lam * (dd * (double)acc[t].x), 0, ep.lb_kind, ep.lb, nullptr, ep.ub_kind,
                           ep.ub, nullptr);
            xn.y = project((double)xo.y + ep.lam * (dd * (double)acc[t].y), 0, ep.lb_kind, ep.lb, nullptr, ep.ub_kind,
                           ep.ub, nullptr);
            xn.z = project((double)xo.z + ep.lam * (dd * (double)acc[t].z), 0, ep.lb_kind, ep.lb, nullptr, ep.ub_kind,
                           ep.ub, nullptr);
            xn.w = project((double)xo.w + ep.lam * (dd * (double)acc[t].w), 0, ep.lb_kind, ep.lb, nullptr, ep.ub_kind,
                           ep.ub, nullptr);
            *xp = xn;
        }
    }
}

// Row pass: r = b - (sum of the row's piece partials, in tile order), M r and
// FP64 sum r^2 per right-hand side (block partials).  16 lanes per row, lane q
// holds right-hand sides 4q..4q+3.
__global__ __launch_bounds__(256) void k_spmm_rows(const int64_t* __restrict__ rp_ptr,
                                                   const uint32_t* __restrict__ rp_idx,
                                                   const float* __restrict__ partial, const float* __restrict__ B,
                                                   const float* __restrict__ M, int64_t m, float* __restrict__ MR,
                                                   double* __restrict__ part) {
    __shared__ double sh[16][SP_RHS];
    const int tid = threadIdx.x, rg = tid >> 4, q = tid & 15;
    double n0 = 0, n1 = 0, n2 = 0, n3 = 0;
    for (int64_t i = (int64_t)blockIdx.x * 16 + rg; i < m; i += (int64_t)gridDim.x * 16) {
        const int64_t s = rp_ptr[i], e = rp_ptr[i + 1];
        float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
        int64_t k = s;
        for (; k + 4 <= e; k += 4) {
            const uint32_t p0 = rp_idx[k], p1 = rp_idx[k + 1], p2 = rp_idx[k + 2], p3 = rp_idx[k + 3];
            const float4 v0 = *(const float4*)(partial + (int64_t)p0 * SP_RHS + q * 4);
            const float4 v1 = *(const float4*)(partial + (int64_t)p1 * SP_RHS + q * 4);
            const float4 v2 = *(const float4*)(partial + (int64_t)p2 * SP_RHS + q * 4);
            const float4 v3 = *(const float4*)(partial + (int64_t)p3 * SP_RHS + q * 4);
            sum.x += v0.x; sum.y += v0.y; sum.z += v0.z; sum.w += v0.w;
            sum.x += v1.x; sum.y += v1.y; sum.z += v1.z; sum.w += v1.w;
            sum.x += v2.x; sum.y += v2.y; sum.z += v2.z; sum.w += v2.w;
            sum.x += v3.x; sum.y += v3.y; sum.z += v3.z; sum.w += v3.w;
        }
        for (; k < e; ++k) {
            const float4 v = *(const float4*)(partial + (int64_t)rp_idx[k] * SP_RHS + q * 4);
            sum.x += v.x; sum.y += v.y; sum.z += v.z; sum.w += v.w;
        }
        const float4 b = *(const float4*)(B + i * SP_RHS + q * 4);
        const float4 r = make_float4(b.x - sum.x, b.y - sum.y, b.z - sum.z, b.w - sum.w);
        const float mi = M ? M[i] : 1.f;
        *(float4*)(MR + i * SP_RHS + q * 4) = make_float4(mi * r.x, mi * r.y, mi * r.z, mi * r.w);
        n0 += (double)r.x * (double)r.x;
        n1 += (double)r.y * (double)r.y;
        n2 += (double)r.z * (double)r.z;
        n3 += (double)r.w * (double)r.w;
    }
    sh[rg][q * 4 + 0] = n0;
    sh[rg][q * 4 + 1] = n1;
    sh[rg][q * 4 + 2] = n2;
    sh[rg][q * 4 + 3] = n3;
    __syncthreads();
    if (tid < SP_RHS) {
        double s = 0;
        for (int k = 0; k < 16; ++k) s += sh[k][tid];
        part[(int64_t)blockIdx.x * SP_RHS + tid] = s;
    }
}

// ---------------------------------------------------------------- layout build (GPU)
unsigned sp_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536)); }

int sp_bits(uint64_t v) {
    int b = 1;
    while (b < 64 && ((uint64_t)1 << b) <= v) b++;
    return b;
}

template <typename K, typename Vv>
void sp_sort(const K* kin, K* kout, const Vv* vin, Vv* vout, int64_t n, int end_bit, hipStream_t s) {
    size_t tb = 0;
    AIR_HIP(rocprim::radix_sort_pairs(nullptr, tb, kin, kout, vin, vout, (size_t)n, 0, end_bit, s));
    DevBuf tmp;
    tmp.alloc(tb);
    AIR_HIP(rocprim::radix_sort_pairs(tmp.p, tb, kin, kout, vin, vout, (size_t)n, 0, end_bit, s));
    AIR_HIP(hipStreamSynchronize(s));
}

void sp_scan(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s) {
    size_t tb = 0;
    AIR_HIP(rocprim::exclusive_scan(nullptr, tb, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s));
    DevBuf tmp;
    tmp.alloc(tb);
    AIR_HIP(rocprim::exclusive_scan(tmp.p, tb, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s));
}

__global__ void k_sp_rowof(const int64_t* __restrict__ ptr, int64_t m, int* __restrict__ rowof) {
    const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (row >= m) return;
    for (int64_t e = ptr[row] + lane; e < ptr[row + 1]; e += 64) rowof[e] = (int)row;
}

// forward key: tile (20 bits) | row (32) | sub-tile (4) | column inside the sub-tile (8)
__global__ void k_sp_fkeys(const int* __restrict__ idx, const int* __restrict__ rowof, const int* __restrict__ perm,
                           int64_t nnz, uint64_t* __restrict__ key, uint32_t* __restrict__ iota) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t c = (uint32_t)perm[idx[e]];
        key[e] = ((c >> 12) << 44) | ((uint64_t)(uint32_t)rowof[e] << 12) | (c & 4095u);
        iota[e] = (uint32_t)e;
    }
}

// start flags of segments with equal key >> sh
__global__ void k_sp_flags(const uint64_t* __restrict__ ks, int64_t n, int sh, uint32_t* __restrict__ flag) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        flag[e] = (e == 0 || (ks[e] >> sh) != (ks[e - 1] >> sh)) ? 1u : 0u;
}
__global__ void k_sp_flags32(const uint32_t* __restrict__ ks, int64_t n, uint32_t* __restrict__ flag) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        flag[e] = (e == 0 || ks[e] != ks[e - 1]) ? 1u : 0u;
}

// first entry of every segment; count of segments per group (group = key >> gsh)
__global__ void k_sp_segs(const uint64_t* __restrict__ ks, const uint32_t* __restrict__ flag,
                          const uint32_t* __restrict__ ex, int64_t n, int gsh, uint32_t* __restrict__ segstart,
                          uint32_t* __restrict__ gcount) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        if (flag[e]) {
            if (segstart) segstart[ex[e]] = (uint32_t)e;
            if (gcount) atomicAdd(&gcount[ks[e] >> gsh], 1u);
        }
}
__global__ void k_sp_segs32(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ ex, int64_t n,
                            uint32_t* __restrict__ segstart) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        if (flag[e]) segstart[ex[e]] = (uint32_t)e;
}

// forward cells: piece -> (batch, round, wave, group), sub-tile = chunk, step k.
// Cell id = chunk * 256 + wave * 16 + round with chunk = batch * 16 + sub-tile.
__global__ void k_sp_fcells(const uint64_t* __restrict__ ks, const uint32_t* __restrict__ pflag,
                            const uint32_t* __restrict__ pex, const uint32_t* __restrict__ sflag,
                            const uint32_t* __restrict__ sex, const uint32_t* __restrict__ segstart,
                            const uint32_t* __restrict__ tfp, const uint32_t* __restrict__ bb, int64_t n,
                            uint32_t* __restrict__ cellC, uint32_t* __restrict__ ecell, uint32_t* __restrict__ ek) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = ks[e];
        const uint32_t F = (uint32_t)(key >> 44), sub = (uint32_t)(key >> 8) & 15u;
        const uint32_t p = pex[e] + pflag[e] - 1u, qp = p - tfp[F];
        const uint32_t B = bb[F] + qp / SP_TARGETS, i = qp % SP_TARGETS;
        const uint32_t t = i / (4 * SP_WAVES), w = (i >> 2) % SP_WAVES, g = i & 3u;
        const uint32_t k = (uint32_t)e - segstart[sex[e] + sflag[e] - 1u];
        const uint32_t cell = (B * 16u + sub) * SP_CELLS + w * 16u + t;
        atomicMax(&cellC[cell], k + 1u);
        ecell[e] = cell;
        ek[e] = (k << 2) | g;
    }
}

// Byte of step k of group g in the record stream: the cell's wave owns the
// blocks from wfirst[work-group * 16 + wave]; the cell starts at record r0 of
// that wave's sequence; 12 records per 1 KB block.
__device__ __forceinline__ int64_t sp_entry_byte(uint32_t cell, uint32_t k, uint32_t g,
                                                 const uint32_t* __restrict__ cell_r0,
                                                 const int* __restrict__ chunk_wg,
                                                 const int64_t* __restrict__ wfirst) {
    const uint32_t w = (cell % SP_CELLS) >> 4;
    const int64_t wb = wfirst[(int64_t)chunk_wg[cell / SP_CELLS] * SP_WAVES + w];
    const uint32_t rec = cell_r0[cell] + (k >> 2);
    return (wb + rec / SP_BLK_RECS) * SP_BLK + (rec % SP_BLK_RECS) * SP_REC + g * 16 + (k & 3u) * 4;
}

template <typename V>
__global__ void k_sp_fscatter(const uint64_t* __restrict__ ks, const uint32_t* __restrict__ ord,
                              const V* __restrict__ val, const uint32_t* __restrict__ ecell,
                              const uint32_t* __restrict__ ek, const uint32_t* __restrict__ cell_r0,
                              const int* __restrict__ chunk_wg, const int64_t* __restrict__ wfirst, int64_t n,
                              unsigned char* __restrict__ stream) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t kg = ek[e], k = kg >> 2, g = kg & 3u;
        const int64_t at = sp_entry_byte(ecell[e], k, g, cell_r0, chunk_wg, wfirst);
        *(float*)(stream + at) = to_f32(val[ord[e]]);
        // the slot byte of this step: record byte 64 + 4 g + (k & 3)
        stream[at - g * 16 - (k & 3u) * 4 + 64 + g * 4 + (k & 3u)] = (unsigned char)(ks[e] & 255u);
    }
}

// backward key: block of 512 consecutive internal columns (23 bits) | row (32)
// | column inside the block (9)
__global__ void k_sp_bkeys(const int* __restrict__ idx, const int* __restrict__ rowof, const int* __restrict__ perm,
                           int64_t nnz, uint64_t* __restrict__ key, uint32_t* __restrict__ iota) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t c = (uint32_t)perm[idx[e]];
        key[e] = ((c >> SP_BB) << (SP_BB + 32)) | ((uint64_t)(uint32_t)rowof[e] << SP_BB) | (c & (SP_TARGETS - 1));
        iota[e] = (uint32_t)e;
    }
}

// distinct rows of every block, counted per block
__global__ void k_sp_buniq(const uint64_t* __restrict__ ks, const uint32_t* __restrict__ uflag,
                           const uint32_t* __restrict__ uex, int64_t n, int* __restrict__ urow,
                           uint32_t* __restrict__ nu) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        if (uflag[e]) {
            urow[uex[e]] = (int)((ks[e] >> SP_BB) & 0xffffffffu);
            atomicAdd(&nu[ks[e] >> (SP_BB + 32)], 1u);
        }
}

// chunk of 256 distinct rows and slot inside it; key2 = chunk * 512 + column slot
__global__ void k_sp_bchunk(const uint64_t* __restrict__ ks, const uint32_t* __restrict__ uflag,
                            const uint32_t* __restrict__ uex, const uint32_t* __restrict__ tfu,
                            const uint32_t* __restrict__ cb, int64_t n, uint32_t* __restrict__ key2,
                            unsigned char* __restrict__ slot_of, uint32_t* __restrict__ iota) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = ks[e];
        const uint32_t G = (uint32_t)(key >> (SP_BB + 32));
        const uint32_t r = uex[e] + uflag[e] - 1u - tfu[G];
        key2[e] = (cb[G] + (r >> 8)) * (uint32_t)SP_TARGETS + (uint32_t)(key & (SP_TARGETS - 1));
        slot_of[e] = (unsigned char)(r & 255u);
        iota[e] = (uint32_t)e;
    }
}

__global__ void k_sp_bcells(const uint32_t* __restrict__ ks2, const uint32_t* __restrict__ flag,
                            const uint32_t* __restrict__ ex, const uint32_t* __restrict__ segstart, int64_t n,
                            uint32_t* __restrict__ cellC, uint32_t* __restrict__ ecell, uint32_t* __restrict__ ek) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t key = ks2[e], ch = key / SP_TARGETS, i = key % SP_TARGETS;
        const uint32_t t = i / (4 * SP_WAVES), w = (i >> 2) % SP_WAVES, g = i & 3u;
        const uint32_t k = (uint32_t)e - segstart[ex[e] + flag[e] - 1u];
        const uint32_t cell = ch * SP_CELLS + w * 16u + t;
        atomicMax(&cellC[cell], k + 1u);
        ecell[e] = cell;
        ek[e] = (k << 2) | g;
    }
}

template <typename V>
__global__ void k_sp_bscatter(const uint32_t* __restrict__ ord2, const uint32_t* __restrict__ ord1,
                              const unsigned char* __restrict__ slot_of, const V* __restrict__ val,
                              const uint32_t* __restrict__ ecell, const uint32_t* __restrict__ ek,
                              const uint32_t* __restrict__ cell_r0, const int* __restrict__ chunk_wg,
                              const int64_t* __restrict__ wfirst, int64_t n, unsigned char* __restrict__ stream) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t kg = ek[e], k = kg >> 2, g = kg & 3u;
        const uint32_t p1 = ord2[e];
        const int64_t at = sp_entry_byte(ecell[e], k, g, cell_r0, chunk_wg, wfirst);
        *(float*)(stream + at) = to_f32(val[ord1[p1]]);
        stream[at - g * 16 - (k & 3u) * 4 + 64 + g * 4 + (k & 3u)] = slot_of[p1];
    }
}

// Per work-group: its first global chunk and chunk count.  Chunks are
// numbered so that a work-group's chunks are consecutive; cell (chunk, wave,
// round) = chunk * 256 + wave * 16 + round.
struct WgPlan {
    int chunk0, nch;
    int4 desc;      // {chunk0, nch, A, B}
};

// Every wave's record sequence over its work-group's chunks and rounds (cell
// record offsets cell_r0), packed 12 records to a 1 KB block, the waves'
// blocks one after another; the per-chunk count table; validation of the
// bounds the kernel relies on.  chunk_op: per chunk {first row, rows}.
void plan_streams(SpmmPass& P, const std::vector<uint32_t>& cellC, const std::vector<WgPlan>& wgs,
                  const std::vector<int4>& chunk_op, int64_t op_rows_total, std::vector<uint32_t>& cell_r0,
                  std::vector<int>& chunk_wg, std::vector<int64_t>& wfirst) {
    const int64_t nchunks = (int64_t)chunk_op.size();
    AIR_CHECK((int64_t)cellC.size() == nchunks * SP_CELLS, AIR_E_STATE, "spmm layout: cell count");
    cell_r0.assign(cellC.size(), 0u);
    chunk_wg.assign(nchunks, -1);
    const int nwg = (int)wgs.size();
    wfirst.assign((size_t)nwg * SP_WAVES, 0);
    std::vector<int32_t> wnb((size_t)nwg * SP_WAVES, 0);
    std::vector<uint32_t> ctab((size_t)nchunks * SP_WAVES * (SP_ROUNDS / 2), 0u);
    int64_t blocks = 0;
    for (int wg = 0; wg < nwg; ++wg) {
        const WgPlan& W = wgs[wg];
        AIR_CHECK(W.chunk0 >= 0 && W.nch >= 0 && (int64_t)W.chunk0 + W.nch <= nchunks, AIR_E_STATE,
                  "spmm layout: chunk range");
        for (int c = 0; c < W.nch; ++c) {
            AIR_CHECK(chunk_wg[W.chunk0 + c] < 0, AIR_E_STATE, "spmm layout: a chunk belongs to two work-groups");
            chunk_wg[W.chunk0 + c] = wg;
            const int4 op = chunk_op[W.chunk0 + c];
            AIR_CHECK(op.y >= 1 && op.y <= SP_ROWS && op.x >= 0 && (int64_t)op.x + op.y <= op_rows_total,
                      AIR_E_STATE, "spmm layout: operand rows out of range");
        }
        for (int w = 0; w < SP_WAVES; ++w) {
            uint64_t r = 0;
            for (int c = 0; c < W.nch; ++c)
                for (int t = 0; t < SP_ROUNDS; ++t) {
                    const int64_t cell = (int64_t)(W.chunk0 + c) * SP_CELLS + w * 16 + t;
                    const uint32_t C = cellC[cell];
                    AIR_CHECK(C <= 0xffffu, AIR_E_STATE, "spmm layout: a target has too many entries in one chunk");
                    AIR_CHECK(r < (1ull << 31), AIR_E_UNSUPPORTED, "spmm layout: a wave's stream is too long");
                    cell_r0[cell] = (uint32_t)r;
                    r += (C + 3) / 4;
                    ctab[((size_t)(W.chunk0 + c) * SP_WAVES + w) * (SP_ROUNDS / 2) + t / 2] |= C << (16 * (t & 1));
                }
            const int64_t nb = (int64_t)((r + SP_BLK_RECS - 1) / SP_BLK_RECS);
            wfirst[(size_t)wg * SP_WAVES + w] = blocks;
            wnb[(size_t)wg * SP_WAVES + w] = (int32_t)nb;
            blocks += nb;
        }
        if (W.nch > 0)
            for (int w = 0; w < SP_WAVES; ++w)   // every wave's stream holds exactly its records
                AIR_CHECK(wnb[(size_t)wg * SP_WAVES + w] >= 0, AIR_E_STATE, "spmm layout: block count");
    }
    P.nblocks = blocks;
    {   // layout statistics: step slots filled, and how evenly a chunk's records fall on the waves
        double steps = 0, sum_mean = 0, sum_max = 0;
        for (int64_t ch = 0; ch < nchunks; ++ch) {
            double tot = 0, mx = 0;
            for (int w = 0; w < SP_WAVES; ++w) {
                double rw = 0;
                for (int t = 0; t < SP_ROUNDS; ++t) {
                    const uint32_t C = cellC[ch * SP_CELLS + w * 16 + t];
                    steps += C;
                    rw += (C + 3) / 4;
                }
                tot += rw;
                mx = std::max(mx, rw);
            }
            sum_mean += tot / SP_WAVES;
            sum_max += mx;
        }
        P.step_slots = (int64_t)(steps * 4);
        P.wave_balance = sum_max > 0 ? sum_mean / sum_max : 1.0;
    }
    const size_t nw = std::max<size_t>((size_t)nwg * SP_WAVES, 1);
    P.wblk.alloc(8 * nw);
    P.wnb.alloc(4 * nw);
    P.chd.alloc(16 * (size_t)std::max<int64_t>(nchunks, 1));
    P.ctab.alloc(4 * std::max<size_t>(ctab.size(), 1));
    std::vector<int4> wgt(std::max(nwg, 1));
    for (int wg = 0; wg < nwg; ++wg) wgt[wg] = wgs[wg].desc;
    P.wgt.alloc(16 * wgt.size());
    if (nwg > 0) {
        AIR_HIP(hipMemcpy(P.wblk.p, wfirst.data(), 8 * wfirst.size(), hipMemcpyHostToDevice));
        AIR_HIP(hipMemcpy(P.wnb.p, wnb.data(), 4 * wnb.size(), hipMemcpyHostToDevice));
        AIR_HIP(hipMemcpy(P.wgt.p, wgt.data(), 16 * wgt.size(), hipMemcpyHostToDevice));
    }
    if (nchunks > 0) {
        AIR_HIP(hipMemcpy(P.chd.p, chunk_op.data(), 16 * chunk_op.size(), hipMemcpyHostToDevice));
        AIR_HIP(hipMemcpy(P.ctab.p, ctab.data(), 4 * ctab.size(), hipMemcpyHostToDevice));
    }
}

// launch order: entry b runs on XCD b % 8 (round-robin dispatch)
void upload_lperm(SpmmPass& P, const std::vector<std::vector<int>>& per_xcd) {
    size_t len = 0;
    for (const auto& v : per_xcd) len = std::max(len, v.size());
    std::vector<int> tab(std::max<size_t>(len * NXCD, 1), -1);
    for (int x = 0; x < NXCD; ++x)
        for (size_t j = 0; j < per_xcd[x].size(); ++j) tab[j * NXCD + x] = per_xcd[x][j];
    P.nwg_launch = (int)(len * NXCD);
    P.lperm.alloc(4 * tab.size());
    AIR_HIP(hipMemcpy(P.lperm.p, tab.data(), 4 * tab.size(), hipMemcpyHostToDevice));
}

void alloc_stream(SpmmPass& P, hipStream_t s) {
    const size_t bytes = (size_t)std::max<int64_t>(P.nblocks, 1) * SP_BLK;
    P.stream.alloc(bytes);
    AIR_HIP(hipMemsetAsync(P.stream.p, 0, bytes, s));
}

// cell record offsets, owning work-groups and wave blocks on the device for the scatter
struct ScatterTabs {
    DevBuf r0, cwg, wfirst;
    void upload(const std::vector<uint32_t>& r0h, const std::vector<int>& cwgh, const std::vector<int64_t>& wfh,
                hipStream_t s) {
        r0.alloc(4 * std::max<size_t>(r0h.size(), 1));
        cwg.alloc(4 * std::max<size_t>(cwgh.size(), 1));
        wfirst.alloc(8 * std::max<size_t>(wfh.size(), 1));
        if (!r0h.empty()) AIR_HIP(hipMemcpyAsync(r0.p, r0h.data(), 4 * r0h.size(), hipMemcpyHostToDevice, s));
        if (!cwgh.empty()) AIR_HIP(hipMemcpyAsync(cwg.p, cwgh.data(), 4 * cwgh.size(), hipMemcpyHostToDevice, s));
        if (!wfh.empty()) AIR_HIP(hipMemcpyAsync(wfirst.p, wfh.data(), 8 * wfh.size(), hipMemcpyHostToDevice, s));
    }
};

}  // namespace

bool build_spmm(const DevSparse& A, const TiledOp& T, SpmmOp& S, hipStream_t s) {
    S = SpmmOp();
    const int64_t m = A.nrows, nnz = A.nnz;
    if (!T.ready || nnz == 0 || T.n_int % SP_FT != 0 || T.n_int / SP_FT >= (1 << 19)) return false;
    const int64_t nF = T.n_int / SP_FT;
    S.m = m;
    S.n = A.ncols;
    S.n_int = T.n_int;
    S.npieces = T.npieces;
    S.perm_h.resize(A.ncols);
    AIR_HIP(hipMemcpy(S.perm_h.data(), T.perm.p, 4 * (size_t)A.ncols, hipMemcpyDeviceToHost));
    DevBuf rowof, key, ks, iota, ord, f1, x1, f2, x2, seg, ecell, ek;
    rowof.alloc(4 * (size_t)nnz);
    k_sp_rowof<<<(unsigned)((m * 64 + 255) / 256), 256, 0, s>>>(A.ptr.as<int64_t>(), m, rowof.as<int>());
    AIR_HIP(hipGetLastError());
    key.alloc(8 * (size_t)nnz);
    ks.alloc(8 * (size_t)nnz);
    iota.alloc(4 * (size_t)nnz);
    ord.alloc(4 * (size_t)nnz);
    f1.alloc(4 * (size_t)nnz);
    x1.alloc(4 * (size_t)nnz);
    f2.alloc(4 * (size_t)nnz);
    x2.alloc(4 * (size_t)nnz);
    seg.alloc(4 * (size_t)nnz);
    ecell.alloc(4 * (size_t)nnz);
    ek.alloc(4 * (size_t)nnz);

    // ================= forward: pieces of each 64 x 64 tile in batches of 512
    {
        k_sp_fkeys<<<sp_grid(nnz), 256, 0, s>>>(A.idx.as<int>(), rowof.as<int>(), T.perm.as<int>(), nnz,
                                                key.as<uint64_t>(), iota.as<uint32_t>());
        AIR_HIP(hipGetLastError());
        sp_sort(key.as<uint64_t>(), ks.as<uint64_t>(), iota.as<uint32_t>(), ord.as<uint32_t>(), nnz,
                44 + sp_bits((uint64_t)nF), s);
        k_sp_flags<<<sp_grid(nnz), 256, 0, s>>>(ks.as<uint64_t>(), nnz, 12, f1.as<uint32_t>());   // pieces
        k_sp_flags<<<sp_grid(nnz), 256, 0, s>>>(ks.as<uint64_t>(), nnz, 8, f2.as<uint32_t>());    // (piece, sub-tile)
        AIR_HIP(hipGetLastError());
        sp_scan(f1.as<uint32_t>(), x1.as<uint32_t>(), nnz, s);
        sp_scan(f2.as<uint32_t>(), x2.as<uint32_t>(), nnz, s);
        DevBuf pcount;
        pcount.alloc(4 * (size_t)nF);
        AIR_HIP(hipMemsetAsync(pcount.p, 0, 4 * nF, s));
        k_sp_segs<<<sp_grid(nnz), 256, 0, s>>>(ks.as<uint64_t>(), f1.as<uint32_t>(), x1.as<uint32_t>(), nnz, 44,
                                               nullptr, pcount.as<uint32_t>());
        k_sp_segs<<<sp_grid(nnz), 256, 0, s>>>(ks.as<uint64_t>(), f2.as<uint32_t>(), x2.as<uint32_t>(), nnz, 0,
                                               seg.as<uint32_t>(), nullptr);
        AIR_HIP(hipGetLastError());
        std::vector<uint32_t> pc(nF);
        AIR_HIP(hipMemcpyAsync(pc.data(), pcount.p, 4 * nF, hipMemcpyDeviceToHost, s));
        AIR_HIP(hipStreamSynchronize(s));
        std::vector<uint32_t> tfp(nF), bb(nF);
        int64_t P = 0, nB = 0;
        for (int64_t F = 0; F < nF; ++F) {
            tfp[F] = (uint32_t)P;
            bb[F] = (uint32_t)nB;
            P += pc[F];
            nB += (pc[F] + SP_TARGETS - 1) / SP_TARGETS;
        }
        AIR_CHECK(P == T.npieces, AIR_E_STATE, "spmm layout: pieces differ from the tiled layout's");
        AIR_CHECK(nB * 16 * SP_CELLS < (int64_t)UINT32_MAX, AIR_E_UNSUPPORTED, "spmm layout: too many cells");
        DevBuf dtfp, dbb, cellC;
        dtfp.alloc(4 * (size_t)nF);
        dbb.alloc(4 * (size_t)nF);
        AIR_HIP(hipMemcpyAsync(dtfp.p, tfp.data(), 4 * nF, hipMemcpyHostToDevice, s));
        AIR_HIP(hipMemcpyAsync(dbb.p, bb.data(), 4 * nF, hipMemcpyHostToDevice, s));
        const int64_t ncell = nB * 16 * SP_CELLS;
        cellC.alloc(4 * (size_t)std::max<int64_t>(ncell, 1));
        AIR_HIP(hipMemsetAsync(cellC.p, 0, 4 * ncell, s));
        k_sp_fcells<<<sp_grid(nnz), 256, 0, s>>>(ks.as<uint64_t>(), f1.as<uint32_t>(), x1.as<uint32_t>(),
                                                 f2.as<uint32_t>(), x2.as<uint32_t>(), seg.as<uint32_t>(),
                                                 dtfp.as<uint32_t>(), dbb.as<uint32_t>(), nnz, cellC.as<uint32_t>(),
                                                 ecell.as<uint32_t>(), ek.as<uint32_t>());
        AIR_HIP(hipGetLastError());
        std::vector<uint32_t> hc(ncell);
        AIR_HIP(hipMemcpyAsync(hc.data(), cellC.p, 4 * ncell, hipMemcpyDeviceToHost, s));
        AIR_HIP(hipStreamSynchronize(s));
        // chunks: every batch walks the 16 sub-tiles of its tile; work-group = batch
        std::vector<WgPlan> wgs;
        std::vector<int4> chunk_op((size_t)nB * 16);
        std::vector<std::vector<int>> per_xcd(NXCD);
        for (int64_t F = 0; F < nF; ++F) {
            const int64_t nb = (pc[F] + SP_TARGETS - 1) / SP_TARGETS;
            for (int64_t b = 0; b < nb; ++b) {
                const int64_t B = bb[F] + b;
                AIR_CHECK(B == (int64_t)wgs.size(), AIR_E_STATE, "spmm layout: batch order");
                for (int sub = 0; sub < 16; ++sub)
                    chunk_op[B * 16 + sub] = make_int4((int)(F * SP_FT + sub * 256), SP_ROWS, 0, 0);
                const int64_t p0 = tfp[F] + b * SP_TARGETS;
                const int np = (int)std::min<int64_t>(SP_TARGETS, pc[F] - b * SP_TARGETS);
                AIR_CHECK(p0 + np <= P, AIR_E_STATE, "spmm layout: pieces of a work-group out of range");
                wgs.push_back({(int)(B * 16), 16, make_int4((int)(B * 16), 16, (int)p0, np)});
                per_xcd[F % NXCD].push_back((int)B);
            }
        }
        std::vector<uint32_t> cell_r0;
        std::vector<int> chunk_wg;
        std::vector<int64_t> wfirst;
        plan_streams(S.fwd, hc, wgs, chunk_op, T.n_int, cell_r0, chunk_wg, wfirst);
        upload_lperm(S.fwd, per_xcd);
        alloc_stream(S.fwd, s);
        ScatterTabs st;
        st.upload(cell_r0, chunk_wg, wfirst, s);
        dispatch_dtype(A.dtype, [&](auto v) {
            using V = decltype(v);
            k_sp_fscatter<V><<<sp_grid(nnz), 256, 0, s>>>(ks.as<uint64_t>(), ord.as<uint32_t>(), A.val.as<V>(),
                                                          ecell.as<uint32_t>(), ek.as<uint32_t>(),
                                                          st.r0.as<uint32_t>(), st.cwg.as<int>(),
                                                          st.wfirst.as<int64_t>(), nnz,
                                                          S.fwd.stream.as<unsigned char>());
            return 0;
        });
        AIR_HIP(hipGetLastError());
        AIR_HIP(hipStreamSynchronize(s));
    }

    // ================= backward: 32 x 32 blocks of pixels, chunks of 256 distinct rows
    {
        const int64_t nG = T.n_int / SP_TARGETS;
        k_sp_bkeys<<<sp_grid(nnz), 256, 0, s>>>(A.idx.as<int>(), rowof.as<int>(), T.perm.as<int>(), nnz,
                                                key.as<uint64_t>(), iota.as<uint32_t>());
        AIR_HIP(hipGetLastError());
        sp_sort(key.as<uint64_t>(), ks.as<uint64_t>(), iota.as<uint32_t>(), ord.as<uint32_t>(), nnz,
                SP_BB + 32 + sp_bits((uint64_t)nG), s);
        k_sp_flags<<<sp_grid(nnz), 256, 0, s>>>(ks.as<uint64_t>(), nnz, SP_BB, f1.as<uint32_t>());   // (block, row)
        AIR_HIP(hipGetLastError());
        sp_scan(f1.as<uint32_t>(), x1.as<uint32_t>(), nnz, s);
        uint32_t lastx = 0, lastf = 0;
        AIR_HIP(hipMemcpyAsync(&lastx, x1.as<uint32_t>() + nnz - 1, 4, hipMemcpyDeviceToHost, s));
        AIR_HIP(hipMemcpyAsync(&lastf, f1.as<uint32_t>() + nnz - 1, 4, hipMemcpyDeviceToHost, s));
        AIR_HIP(hipStreamSynchronize(s));
        const int64_t NU = (int64_t)lastx + lastf;
        S.rowlist.alloc(4 * (size_t)std::max<int64_t>(NU, 1));
        DevBuf nu;
        nu.alloc(4 * (size_t)nG);
        AIR_HIP(hipMemsetAsync(nu.p, 0, 4 * nG, s));
        k_sp_buniq<<<sp_grid(nnz), 256, 0, s>>>(ks.as<uint64_t>(), f1.as<uint32_t>(), x1.as<uint32_t>(), nnz,
                                                S.rowlist.as<int>(), nu.as<uint32_t>());
        AIR_HIP(hipGetLastError());
        std::vector<uint32_t> hnu(nG), tfu(nG), cb(nG);
        AIR_HIP(hipMemcpyAsync(hnu.data(), nu.p, 4 * nG, hipMemcpyDeviceToHost, s));
        AIR_HIP(hipStreamSynchronize(s));
        int64_t U = 0, nch = 0;
        for (int64_t G = 0; G < nG; ++G) {
            tfu[G] = (uint32_t)U;
            cb[G] = (uint32_t)nch;
            U += hnu[G];
            nch += (hnu[G] + SP_ROWS - 1) / SP_ROWS;
        }
        AIR_CHECK(U == NU, AIR_E_STATE, "spmm layout: distinct rows do not add up");
        AIR_CHECK(nch * SP_TARGETS < (int64_t)UINT32_MAX && nch * SP_CELLS < (int64_t)UINT32_MAX, AIR_E_UNSUPPORTED,
                  "spmm layout: too many row chunks");
        DevBuf dtfu, dcb, key2, ks2, slot_of, cellC;
        dtfu.alloc(4 * (size_t)nG);
        dcb.alloc(4 * (size_t)nG);
        AIR_HIP(hipMemcpyAsync(dtfu.p, tfu.data(), 4 * nG, hipMemcpyHostToDevice, s));
        AIR_HIP(hipMemcpyAsync(dcb.p, cb.data(), 4 * nG, hipMemcpyHostToDevice, s));
        key2.alloc(4 * (size_t)nnz);
        ks2.alloc(4 * (size_t)nnz);
        slot_of.alloc((size_t)nnz);
        DevBuf iota2, ord2;
        iota2.alloc(4 * (size_t)nnz);
        ord2.alloc(4 * (size_t)nnz);
        k_sp_bchunk<<<sp_grid(nnz), 256, 0, s>>>(ks.as<uint64_t>(), f1.as<uint32_t>(), x1.as<uint32_t>(),
                                                 dtfu.as<uint32_t>(), dcb.as<uint32_t>(), nnz, key2.as<uint32_t>(),
                                                 slot_of.as<unsigned char>(), iota2.as<uint32_t>());
        AIR_HIP(hipGetLastError());
        // stable: entries of one (chunk, pixel) keep their row order
        sp_sort(key2.as<uint32_t>(), ks2.as<uint32_t>(), iota2.as<uint32_t>(), ord2.as<uint32_t>(), nnz,
                sp_bits((uint64_t)nch * SP_TARGETS), s);
        k_sp_flags32<<<sp_grid(nnz), 256, 0, s>>>(ks2.as<uint32_t>(), nnz, f2.as<uint32_t>());
        AIR_HIP(hipGetLastError());
        sp_scan(f2.as<uint32_t>(), x2.as<uint32_t>(), nnz, s);
        k_sp_segs32<<<sp_grid(nnz), 256, 0, s>>>(f2.as<uint32_t>(), x2.as<uint32_t>(), nnz, seg.as<uint32_t>());
        const int64_t ncell = nch * SP_CELLS;
        cellC.alloc(4 * (size_t)std::max<int64_t>(ncell, 1));
        AIR_HIP(hipMemsetAsync(cellC.p, 0, 4 * ncell, s));
        k_sp_bcells<<<sp_grid(nnz), 256, 0, s>>>(ks2.as<uint32_t>(), f2.as<uint32_t>(), x2.as<uint32_t>(),
                                                 seg.as<uint32_t>(), nnz, cellC.as<uint32_t>(), ecell.as<uint32_t>(),
                                                 ek.as<uint32_t>());
        AIR_HIP(hipGetLastError());
        std::vector<uint32_t> hc(ncell);
        AIR_HIP(hipMemcpyAsync(hc.data(), cellC.p, 4 * ncell, hipMemcpyDeviceToHost, s));
        AIR_HIP(hipStreamSynchronize(s));
        std::vector<WgPlan> wgs;
        std::vector<int4> chunk_op((size_t)nch);
        for (int64_t G = 0; G < nG; ++G) {
            const int64_t nc = (hnu[G] + SP_ROWS - 1) / SP_ROWS;
            for (int64_t c = 0; c < nc; ++c)
                chunk_op[cb[G] + c] = make_int4((int)(tfu[G] + c * SP_ROWS),
                                                (int)std::min<int64_t>(SP_ROWS, hnu[G] - c * SP_ROWS), 0, 0);
            AIR_CHECK((G + 1) * SP_TARGETS <= T.n_int, AIR_E_STATE, "spmm layout: block out of range");
            wgs.push_back({(int)cb[G], (int)nc, make_int4((int)cb[G], (int)nc, (int)G, 0)});
        }
        std::vector<uint32_t> cell_r0;
        std::vector<int> chunk_wg;
        std::vector<int64_t> wfirst;
        plan_streams(S.bwd, hc, wgs, chunk_op, NU, cell_r0, chunk_wg, wfirst);
        // contiguous ranges of blocks per XCD (neighbouring blocks share rows)
        std::vector<std::vector<int>> per_xcd(NXCD);
        const int64_t per = (nG + NXCD - 1) / NXCD;
        for (int64_t G = 0; G < nG; ++G) per_xcd[G / per].push_back((int)G);
        upload_lperm(S.bwd, per_xcd);
        alloc_stream(S.bwd, s);
        ScatterTabs st;
        st.upload(cell_r0, chunk_wg, wfirst, s);
        dispatch_dtype(A.dtype, [&](auto v) {
            using V = decltype(v);
            k_sp_bscatter<V><<<sp_grid(nnz), 256, 0, s>>>(ord2.as<uint32_t>(), ord.as<uint32_t>(),
                                                          slot_of.as<unsigned char>(), A.val.as<V>(),
                                                          ecell.as<uint32_t>(), ek.as<uint32_t>(),
                                                          st.r0.as<uint32_t>(), st.cwg.as<int>(),
                                                          st.wfirst.as<int64_t>(), nnz,
                                                          S.bwd.stream.as<unsigned char>());
            return 0;
        });
        AIR_HIP(hipGetLastError());
        AIR_HIP(hipStreamSynchronize(s));
    }
    S.partial.alloc(4 * (size_t)SP_RHS * std::max<int64_t>(S.npieces, 1));
    S.ready = true;
    if (const char* v = std::getenv("AIR_SPMM_STATS"); v && v[0] == '1') {
        for (int k = 0; k < 2; ++k) {
            const SpmmPass& P = k ? S.bwd : S.fwd;
            std::fprintf(stderr, "spmm %s: %d work-groups, %lld blocks (%.1f MB), entries / step slots %.3f, wave balance %.3f\n",
                         k ? "A'y" : "Ax", P.nwg_launch, (long long)P.nblocks, P.nblocks * (double)SP_BLK / 1e6,
                         P.step_slots ? (double)nnz / (double)P.step_slots : 0.0, P.wave_balance);
        }
    }
    return true;
}

void spmm_forward(const SpmmOp& S, const float* X, hipStream_t s) {
    SpEpi ep{S.partial.as<float>(), nullptr, 0.0, 0, 0, 0.f, 0.f};
    const SpmmPass& P = S.fwd;
    k_spmm<0><<<P.nwg_launch, SP_THREADS, 0, s>>>(P.lperm.as<int>(), P.wgt.as<int4>(), P.wblk.as<int64_t>(),
                                                  P.wnb.as<int>(), P.chd.as<int4>(), P.ctab.as<uint32_t>(),
                                                  P.stream.as<uint4>(), nullptr, X, ep);
    AIR_HIP(hipGetLastError());
}

void spmm_backward_update(const SpmmOp& S, const float* MR, float* X, const float* D, double lam, int lb_kind,
                          float lb, int ub_kind, float ub, hipStream_t s) {
    SpEpi ep{X, D, lam, lb_kind, ub_kind, lb, ub};
    const SpmmPass& P = S.bwd;
    k_spmm<1><<<P.nwg_launch, SP_THREADS, 0, s>>>(P.lperm.as<int>(), P.wgt.as<int4>(), P.wblk.as<int64_t>(),
                                                  P.wnb.as<int>(), P.chd.as<int4>(), P.ctab.as<uint32_t>(),
                                                  P.stream.as<uint4>(), S.rowlist.as<int>(), MR, ep);
    AIR_HIP(hipGetLastError());
}

unsigned spmm_rows_grid(int64_t m, int num_cu) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((m + 15) / 16, (int64_t)num_cu * 8));
}

void spmm_rows(const SpmmOp& S, const TiledOp& T, const float* B, const float* M, float* MR, double* part,
               unsigned grid, hipStream_t s) {
    k_spmm_rows<<<grid, 256, 0, s>>>(T.rp_ptr.as<int64_t>(), T.rp_idx.as<uint32_t>(), S.partial.as<float>(), B, M,
                                     S.m, MR, part);
    AIR_HIP(hipGetLastError());
}

}  // namespace air
