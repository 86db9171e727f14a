// air_sort.hip -- device scan and CSR -> CSC transpose (setup-time work).
//
// The CSC of A (the compressed rows of A') feeds the A'.y kernel.  MATLAB
// stores sparse matrices as CSC with row indices ascending per column
// (testprobs/paralleltomo.m:300); the transpose here produces the same order:
// a stable radix sort of the column keys, in CSR (row-major) order, by
// rocPRIM, then a gather of row indices and values.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "air_internal.h"

namespace air {

void scan_exclusive_i64(const int64_t* in, int64_t* out, int64_t n, hipStream_t s) {
    // out has n+1 entries; the extra input element is taken as zero by
    // scanning n+1 values of a buffer whose tail is cleared.
    DevBuf tmp_in;
    tmp_in.alloc(sizeof(int64_t) * (n + 1));
    AIR_HIP(hipMemsetAsync(tmp_in.as<int64_t>() + n, 0, sizeof(int64_t), s));
    if (n > 0) AIR_HIP(hipMemcpyAsync(tmp_in.p, in, sizeof(int64_t) * n, hipMemcpyDeviceToDevice, s));
    size_t tb = 0;
    AIR_HIP(rocprim::exclusive_scan(nullptr, tb, tmp_in.as<int64_t>(), out, (int64_t)0, (size_t)(n + 1),
                                    rocprim::plus<int64_t>(), s));
    DevBuf tmp;
    tmp.alloc(tb);
    AIR_HIP(rocprim::exclusive_scan(tmp.p, tb, tmp_in.as<int64_t>(), out, (int64_t)0, (size_t)(n + 1),
                                    rocprim::plus<int64_t>(), s));
    AIR_HIP(hipStreamSynchronize(s));
}

__global__ void k_iota(uint32_t* v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

// row index of every nonzero (wave per row)
__global__ void k_expand_rows(const int64_t* __restrict__ ptr, int64_t m, int* __restrict__ rowof) {
    int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int lane = threadIdx.x & 63;
    if (row >= m) return;
    for (int64_t q = ptr[row] + lane; q < ptr[row + 1]; q += 64) rowof[q] = (int)row;
}

__global__ void k_count_cols(const int* __restrict__ col, int64_t nnz, int64_t* __restrict__ cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nnz) atomicAdd((unsigned long long*)&cnt[col[i]], 1ull);
}

template <typename V>
__global__ void k_gather_csc(const uint32_t* __restrict__ perm, const int* __restrict__ rowof,
                             const V* __restrict__ val, int64_t nnz, int* __restrict__ rowc, V* __restrict__ valc) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nnz) return;
    uint32_t e = perm[q];
    rowc[q] = rowof[e];
    valc[q] = val[e];
}

void csr_to_csc(const DevSparse& A, DevSparse& AT, hipStream_t s) {
    const int64_t m = A.nrows, n = A.ncols, nnz = A.nnz;
    AIR_CHECK(nnz < (int64_t)UINT32_MAX, AIR_E_ARG, "transpose: nnz must be below 2^32");
    AT.nrows = n;
    AT.ncols = m;
    AT.nnz = nnz;
    AT.dtype = A.dtype;
    AT.ptr.alloc(sizeof(int64_t) * (n + 1));
    alloc_sparse_arrays(AT, nnz, A.dtype, s);
    // column counts -> colptr
    DevBuf cnt;
    cnt.alloc(sizeof(int64_t) * (n + 1));
    AIR_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int64_t) * (n + 1), s));
    if (nnz > 0) k_count_cols<<<(unsigned)((nnz + 255) / 256), 256, 0, s>>>(A.idx.as<int>(), nnz, cnt.as<int64_t>());
    AIR_HIP(hipGetLastError());
    scan_exclusive_i64(cnt.as<int64_t>(), AT.ptr.as<int64_t>(), n, s);
    if (nnz == 0) return;
    // stable sort of column keys carrying the CSR position
    DevBuf rowof, iota, keys_out, perm;
    rowof.alloc(sizeof(int) * nnz);
    iota.alloc(sizeof(uint32_t) * nnz);
    keys_out.alloc(sizeof(uint32_t) * nnz);
    perm.alloc(sizeof(uint32_t) * nnz);
    k_expand_rows<<<(unsigned)((m * 64 + 255) / 256), 256, 0, s>>>(A.ptr.as<int64_t>(), m, rowof.as<int>());
    k_iota<<<(unsigned)((nnz + 255) / 256), 256, 0, s>>>(iota.as<uint32_t>(), nnz);
    AIR_HIP(hipGetLastError());
    int end_bit = 1;
    while (end_bit < 32 && ((int64_t)1 << end_bit) < n) end_bit++;
    size_t tb = 0;
    const uint32_t* keys_in = (const uint32_t*)A.idx.p;
    AIR_HIP(rocprim::radix_sort_pairs(nullptr, tb, keys_in, keys_out.as<uint32_t>(), iota.as<uint32_t>(),
                                      perm.as<uint32_t>(), (size_t)nnz, 0, end_bit, s));
    DevBuf tmp;
    tmp.alloc(tb);
    AIR_HIP(rocprim::radix_sort_pairs(tmp.p, tb, keys_in, keys_out.as<uint32_t>(), iota.as<uint32_t>(),
                                      perm.as<uint32_t>(), (size_t)nnz, 0, end_bit, s));
    dispatch_dtype(A.dtype, [&](auto v) {
        using V = decltype(v);
        k_gather_csc<V><<<(unsigned)((nnz + 255) / 256), 256, 0, s>>>(perm.as<uint32_t>(), rowof.as<int>(),
                                                                      A.val.as<V>(), nnz, AT.idx.as<int>(),
                                                                      AT.val.as<V>());
        return 0;
    });
    AIR_HIP(hipGetLastError());
    AIR_HIP(hipStreamSynchronize(s));
}

}  // namespace air
