// air_batch.hip -- many independent problems as one block-diagonal operator.
//
// Each problem's LDS-tiled layout is built on its own (its columns fill whole
// forward and backward tiles of its image), then the layouts are concatenated:
// entry arrays back to back, and every index that points into a shared array
// (entry offsets, piece numbers, rows, tiles, runs, internal columns) shifted
// by the problem's offset.  The SIRT kernels then treat the batch as one
// operator: one launch per pass covers every problem.
#include "air_internal.h"

namespace air {

namespace {

__global__ void k_shift_i64(const int64_t* __restrict__ src, int64_t n, int64_t off, int64_t* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i] + off;
}
__global__ void k_shift_i32(const int* __restrict__ src, int64_t n, int off, int* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i] + off;
}
__global__ void k_clear_start_flags(unsigned short* __restrict__ code, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        code[i] &= (unsigned short)0xfffeu;
}
unsigned grid_of(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}
void shift64(const DevBuf& src, int64_t n, int64_t off, DevBuf& dst, int64_t at, hipStream_t s) {
    if (n <= 0) return;
    k_shift_i64<<<grid_of(n), 256, 0, s>>>(src.as<int64_t>(), n, off, dst.as<int64_t>() + at);
    AIR_HIP(hipGetLastError());
}
void shift32(const DevBuf& src, int64_t n, int64_t off, DevBuf& dst, int64_t at, hipStream_t s) {
    if (n <= 0) return;
    AIR_CHECK(off < INT32_MAX, AIR_E_UNSUPPORTED, "batched operator too large for 32-bit indices");
    k_shift_i32<<<grid_of(n), 256, 0, s>>>(src.as<int>(), n, (int)off, dst.as<int>() + at);
    AIR_HIP(hipGetLastError());
}
void copy_bytes(const DevBuf& src, size_t bytes, DevBuf& dst, size_t at, hipStream_t s) {
    if (bytes) AIR_HIP(hipMemcpyAsync((char*)dst.p + at, src.p, bytes, hipMemcpyDeviceToDevice, s));
}

}  // namespace

// Concatenate the tiled layouts of independent problems (all of one storage
// type, each ready) into B; the parts are released one by one.  When some part
// cannot carry column-start flags in its backward codes, the flags of the
// others are cleared so that the whole batch uses one code format.
bool concat_tiled(std::vector<TiledOp>& parts, TiledOp& B, hipStream_t s) {
    B = TiledOp();
    AIR_CHECK(!parts.empty(), AIR_E_ARG, "empty batch");
    const int dt = parts[0].dtype;
    int64_t m = 0, n = 0, nint = 0, nnz = 0, nfe = 0, nbe = 0, P = 0, R = 0, tf = 0, tb = 0, staged = 0, rpi = 0;
    int maxruns = 0, maxrows = 0;
    bool flags = true, step = true, bstep = true;
    int64_t nbse = 0;
    size_t lds_bs = 0;
    for (auto& T : parts) {
        AIR_CHECK(T.ready && T.dtype == dt, AIR_E_STATE, "batch parts must be tiled with one storage type");
        m += T.m;
        n += T.n;
        nint += T.n_int;
        nnz += T.nnz;
        nfe += T.nfe;
        nbe += T.nbe;
        P += T.npieces;
        R += T.nruns;
        tf += T.ntiles_f;
        tb += T.ntiles_b;
        staged += T.staged_rows;
        maxruns = std::max(maxruns, T.maxruns);
        maxrows = std::max(maxrows, T.maxrows);
        flags = flags && T.bwd_flags;
        step = step && T.fstep_ok;
        bstep = bstep && T.bstep_ok && T.bstep_bb == parts[0].bstep_bb &&   // one word format
                T.bstep_nw == parts[0].bstep_nw;                            // and block shape
        nbse += T.nbse;
        if (T.bstep_ok) lds_bs = std::max(lds_bs, T.lds_bs - 8 * (size_t)T.maxruns);   // the slots
        rpi += T.npieces;
    }
    AIR_CHECK(nint < INT32_MAX && m < INT32_MAX && P < INT32_MAX && R < INT32_MAX, AIR_E_UNSUPPORTED,
              "batched operator too large for 32-bit indices");
    const size_t vb = dtype_size(dt);
    B.dtype = dt;
    B.N = 0;
    B.m = m;
    B.n = n;
    B.n_int = nint;
    B.nnz = nnz;
    B.nfe = nfe;
    B.nbe = nbe;
    B.npieces = P;
    B.nruns = R;
    B.ntiles_f = tf;
    B.ntiles_b = tb;
    B.staged_rows = staged;
    B.maxruns = maxruns;
    B.maxrows = maxrows;
    B.bt_size = parts[0].bt_size;
    B.bwd_flags = flags;
    B.lds_b = 4 * (2 * (size_t)maxruns + 1 + (size_t)maxrows);
    B.perm.alloc(4 * (size_t)std::max<int64_t>(n, 1));
    B.fval.alloc(vb * (size_t)(nfe + AIR_SPARSE_PAD));
    B.fcode.alloc(2 * (size_t)(nfe + AIR_SPARSE_PAD));
    B.bval.alloc(vb * (size_t)(nbe + AIR_SPARSE_PAD));
    B.blr.alloc(2 * (size_t)(nbe + AIR_SPARSE_PAD));
    AIR_HIP(hipMemsetAsync((char*)B.fval.p + vb * nfe, 0, vb * AIR_SPARSE_PAD, s));
    AIR_HIP(hipMemsetAsync((char*)B.fcode.p + 2 * nfe, 0, 2 * AIR_SPARSE_PAD, s));
    AIR_HIP(hipMemsetAsync((char*)B.bval.p + vb * nbe, 0, vb * AIR_SPARSE_PAD, s));
    AIR_HIP(hipMemsetAsync((char*)B.blr.p + 2 * nbe, 0, 2 * AIR_SPARSE_PAD, s));
    // step words are tile-local, so they concatenate as they are (one per 8 forward entries)
    B.fstep_ok = step;
    if (step) {
        B.fstep.alloc(4 * (size_t)(nfe / 8 + AIR_SPARSE_PAD));
        AIR_HIP(hipMemsetAsync((char*)B.fstep.p + 4 * (nfe / 8), 0, 4 * AIR_SPARSE_PAD, s));
    }
    // step-coded backward: words and strides are tile-local; column pointers are
    // shifted; the staging's run tables take the batch's longest run list
    lds_bs += 8 * (size_t)maxruns;
    bstep = bstep && lds_bs <= (parts[0].bstep_nw == 8 ? AIR_TILE_LDS_STEP_MAX : AIR_TILE_LDS_STEP16_MAX);
    B.bstep_ok = bstep;
    if (bstep) {
        B.nbse = nbse;
        B.lds_bs = lds_bs;
        B.bstep_bb = parts[0].bstep_bb;
        B.bstep_nw = parts[0].bstep_nw;
        B.bsval.alloc(vb * (size_t)(nbse + AIR_SPARSE_PAD));
        AIR_HIP(hipMemsetAsync((char*)B.bsval.p + vb * nbse, 0, vb * AIR_SPARSE_PAD, s));
        B.bsword.alloc(4 * (size_t)(nbse / 8 + AIR_SPARSE_PAD));
        AIR_HIP(hipMemsetAsync((char*)B.bsword.p + 4 * (nbse / 8), 0, 4 * AIR_SPARSE_PAD, s));
        B.bstep_ptr.alloc(8 * (size_t)(nint + 1));
        B.bstep_R.alloc(4 * (size_t)std::max<int64_t>(tb, 1));
    }
    B.piece_ptr.alloc(8 * (size_t)(P + 1));
    B.rp_ptr.alloc(8 * (size_t)(m + 1));
    B.rp_idx.alloc(4 * (size_t)std::max<int64_t>(P, 1));
    B.partial.alloc(4 * (size_t)std::max<int64_t>(P, 1));
    B.cell_ptr.alloc(8 * (size_t)(nint + 1));
    B.run_ptr.alloc(4 * (size_t)(tb + 1));
    B.run_row.alloc(4 * (size_t)std::max<int64_t>(R, 1));
    B.run_off.alloc(4 * (size_t)std::max<int64_t>(R, 1));
    B.run_len.alloc(4 * (size_t)std::max<int64_t>(R, 1));
    B.piece_tile_h.reserve((size_t)P);
    B.piece_ptr_h.reserve((size_t)P + 1);
    int64_t om = 0, on = 0, oint = 0, oef = 0, oeb = 0, oP = 0, oR = 0, otf = 0, otb = 0, obs = 0;
    for (auto& T : parts) {
        // forward
        shift32(T.perm, T.n, oint, B.perm, on, s);
        copy_bytes(T.fval, vb * (size_t)T.nfe, B.fval, vb * (size_t)oef, s);
        copy_bytes(T.fcode, 2 * (size_t)T.nfe, B.fcode, 2 * (size_t)oef, s);
        if (step) copy_bytes(T.fstep, 4 * (size_t)(T.nfe / 8), B.fstep, 4 * (size_t)(oef / 8), s);
        shift64(T.piece_ptr, T.npieces, oef, B.piece_ptr, oP, s);
        shift64(T.rp_ptr, T.m, oP, B.rp_ptr, om, s);      // rp_idx positions of a part start at its piece offset
        shift32(T.rp_idx, T.npieces, oP, B.rp_idx, oP, s);
        for (int64_t k = 0; k < T.npieces; ++k) {
            B.piece_tile_h.push_back(T.piece_tile_h[k] + (int)otf);
            B.piece_ptr_h.push_back(T.piece_ptr_h[k] + oef);
        }
        // backward
        copy_bytes(T.bval, vb * (size_t)T.nbe, B.bval, vb * (size_t)oeb, s);
        copy_bytes(T.blr, 2 * (size_t)T.nbe, B.blr, 2 * (size_t)oeb, s);
        if (!flags && T.bwd_flags && T.nbe > 0) {        // one code format for the whole batch
            k_clear_start_flags<<<grid_of(T.nbe), 256, 0, s>>>(B.blr.as<unsigned short>() + oeb, T.nbe);
            AIR_HIP(hipGetLastError());
        }
        shift64(T.cell_ptr, T.n_int, oeb, B.cell_ptr, oint, s);
        if (bstep) {
            copy_bytes(T.bsval, vb * (size_t)T.nbse, B.bsval, vb * (size_t)obs, s);
            copy_bytes(T.bsword, 4 * (size_t)(T.nbse / 8), B.bsword, 4 * (size_t)(obs / 8), s);
            shift64(T.bstep_ptr, T.n_int, obs, B.bstep_ptr, oint, s);
            copy_bytes(T.bstep_R, 4 * (size_t)T.ntiles_b, B.bstep_R, 4 * (size_t)otb, s);
            obs += T.nbse;
        }
        shift32(T.run_ptr, T.ntiles_b, oR, B.run_ptr, otb, s);
        shift32(T.run_row, T.nruns, om, B.run_row, oR, s);
        copy_bytes(T.run_off, 4 * (size_t)T.nruns, B.run_off, 4 * (size_t)oR, s);
        copy_bytes(T.run_len, 4 * (size_t)T.nruns, B.run_len, 4 * (size_t)oR, s);
        om += T.m;
        on += T.n;
        oint += T.n_int;
        oef += T.nfe;
        oeb += T.nbe;
        oP += T.npieces;
        oR += T.nruns;
        otf += T.ntiles_f;
        otb += T.ntiles_b;
        AIR_HIP(hipStreamSynchronize(s));
        T = TiledOp();                                    // release the part
    }
    // closing entries of the pointer arrays
    AIR_HIP(hipMemcpyAsync(B.piece_ptr.as<int64_t>() + P, &nfe, 8, hipMemcpyHostToDevice, s));
    AIR_HIP(hipMemcpyAsync(B.rp_ptr.as<int64_t>() + m, &P, 8, hipMemcpyHostToDevice, s));
    AIR_HIP(hipMemcpyAsync(B.cell_ptr.as<int64_t>() + nint, &nbe, 8, hipMemcpyHostToDevice, s));
    if (bstep) {
        AIR_HIP(hipMemcpyAsync(B.bstep_ptr.as<int64_t>() + nint, &nbse, 8, hipMemcpyHostToDevice, s));
        set_bstep_waves(B, s);
        tiled_bstep_lds_attr(B);
    }
    const int Ri = (int)R;
    AIR_HIP(hipMemcpyAsync(B.run_ptr.as<int>() + tb, &Ri, 4, hipMemcpyHostToDevice, s));
    B.piece_ptr_h.push_back(nfe);
    AIR_HIP(hipStreamSynchronize(s));
    set_fwd_blocks(B, 256, s);
    set_fwd_waves(B, B.wave_target, s);
    B.ready = true;
    autotune_tiled(B, s);
    return true;
}

}  // namespace air
