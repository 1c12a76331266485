// transpose.hip -- a handle that solves with the transpose of an uploaded
// triangular matrix, built on the device from that matrix's CSR.
//
// A handle stores a lower triangle M (diagonal last) and solves either M or,
// when marked reversed, P M P (P the index reversal: upper matrices).  The
// transpose of either is solved by storing S = P M^T P, which is again lower
// with the diagonal last, under the opposite reversal mark.
//
// Construction: entry p of M at (row j, column i) becomes entry (n-1-i,
// n-1-j) of S.  A stable radix sort of the entries, taken in reverse storage
// order, by their S row groups each S row's entries by descending j, that is
// by ascending S column, so the diagonal (j = i) comes last; a scan of the
// per-row counts gives S's row pointers.  Everything runs on the device; the
// host only reads back the row pointers to record the longest row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace sptrsv {

// keys[q] = S row of entry p = nnz-1-q, vals[q] = p; cnt[key] counts S row lengths.
__global__ void k_tr_keys(const int* __restrict__ ja, int n, int nnz, int* __restrict__ keys,
                          int* __restrict__ vals, int* __restrict__ cnt) {
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nnz;
         q += (long long)gridDim.x * blockDim.x) {
        const int p = nnz - 1 - (int)q;
        const int key = n - 1 - ja[p];
        keys[q] = key;
        vals[q] = p;
        atomicAdd(&cnt[key], 1);
    }
}

// S entry q takes M entry p = vals[q]; its column is n-1-(row of p), the row
// found by binary search in M's row pointers (rows are never empty).
template <bool F64>
__global__ void k_tr_fill(const int* __restrict__ ia, int n, int nnz, const int* __restrict__ vals,
                          const double* __restrict__ a, const float* __restrict__ a32, int* __restrict__ ja_s,
                          double* __restrict__ a_s, float* __restrict__ a32_s) {
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nnz;
         q += (long long)gridDim.x * blockDim.x) {
        const int p = vals[q];
        int lo = 0, hi = n - 1;  // largest j with ia[j] <= p
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ia[mid] <= p)
                lo = mid;
            else
                hi = mid - 1;
        }
        ja_s[q] = n - 1 - lo;
        if constexpr (F64) a_s[q] = a[p];
        a32_s[q] = a32[p];
    }
}

// perm[q] = src_perm[perm[q]]: the value map of the source composed in.
__global__ void k_tr_compose(int* __restrict__ perm, const int* __restrict__ src_perm, int nnz) {
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nnz;
         q += (long long)gridDim.x * blockDim.x)
        perm[q] = src_perm[perm[q]];
}

static unsigned tr_grid(long long count) {
    return (unsigned)std::max<long long>(1, std::min<long long>((count + 255) / 256, 16384));
}

}  // namespace sptrsv

using namespace sptrsv;

extern "C" {

int sptrsv_transpose(sp_mat_t* src, sp_mat_t* dst) {
    if (!src || !dst || !src->priv) return fail(SPTRSV_E_INVALID_ARG, "source matrix not allocated");
    if (src == dst) return fail(SPTRSV_E_INVALID_ARG, "source and destination must differ");
    const int n = src->nr, nnz = src->nnz;
    const bool f64 = src->a != nullptr;
    memset(dst, 0, sizeof(*dst));
    SPT_TRY(ensure_priv(dst));
    dst->nr = dst->nc = n;
    dst->nnz = nnz;
    priv_of(dst)->reversed = !priv_of(src)->reversed;
    int st = SPTRSV_OK;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && st == SPTRSV_OK) st = hip_fail(e, what, __FILE__, __LINE__);
    };
    const size_t ne = (size_t)std::max(nnz, 1);
    chk(hipMalloc(&dst->ia, sizeof(int) * (size_t)(n + 1)), "hipMalloc ia");
    chk(hipMalloc(&dst->ja, sizeof(int) * ne), "hipMalloc ja");
    chk(hipMalloc(&dst->a32, sizeof(float) * ne), "hipMalloc a32");
    if (f64) chk(hipMalloc(&dst->a, sizeof(double) * ne), "hipMalloc a");
    int *keys = nullptr, *vals = nullptr, *tk = nullptr, *tv = nullptr;
    chk(hipMalloc(&keys, sizeof(int) * ne), "hipMalloc keys");
    chk(hipMalloc(&vals, sizeof(int) * ne), "hipMalloc vals");
    chk(hipMalloc(&tk, sizeof(int) * ne), "hipMalloc sort keys");
    chk(hipMalloc(&tv, sizeof(int) * ne), "hipMalloc sort vals");
    if (st == SPTRSV_OK) chk(hipMemset(dst->ia, 0, sizeof(int) * (size_t)(n + 1)), "clear counts");
    if (st == SPTRSV_OK && n > 0) {
        hipLaunchKernelGGL(k_tr_keys, dim3(tr_grid(nnz)), dim3(256), 0, nullptr, src->ja, n, nnz, keys, vals,
                           dst->ia);
        chk(hipGetLastError(), "transpose keys");
        if (st == SPTRSV_OK) st = radix_sort_pairs(keys, vals, tk, tv, nnz, n - 1, nullptr);
        if (st == SPTRSV_OK) st = exclusive_scan(dst->ia, dst->ia, (long long)n + 1, nullptr);
        if (st == SPTRSV_OK) {
            if (f64)
                hipLaunchKernelGGL(k_tr_fill<true>, dim3(tr_grid(nnz)), dim3(256), 0, nullptr, src->ia, n, nnz,
                                   vals, src->a, src->a32, dst->ja, dst->a, dst->a32);
            else
                hipLaunchKernelGGL(k_tr_fill<false>, dim3(tr_grid(nnz)), dim3(256), 0, nullptr, src->ia, n, nnz,
                                   vals, nullptr, src->a32, dst->ja, nullptr, dst->a32);
            chk(hipGetLastError(), "transpose fill");
        }
    }
    if (st == SPTRSV_OK && n > 0) {
        std::vector<int> ia((size_t)n + 1);
        chk(hipMemcpy(ia.data(), dst->ia, sizeof(int) * ia.size(), hipMemcpyDeviceToHost), "read row pointers");
        int mr = 0;
        for (int r = 0; r < n && st == SPTRSV_OK; ++r) mr = std::max(mr, ia[r + 1] - ia[r]);
        if (st == SPTRSV_OK && ia[n] != nnz) st = fail(SPTRSV_E_INVALID_CSR, "transpose row pointers do not sum to nnz");
        priv_of(dst)->max_row = mr;
    }
    // the sorted payload is the value map of the new handle
    MatPriv* sp = priv_of(src);
    if (st == SPTRSV_OK && nnz > 0 && sp->src_perm) {
        hipLaunchKernelGGL(k_tr_compose, dim3(tr_grid(nnz)), dim3(256), 0, nullptr, vals, sp->src_perm, nnz);
        chk(hipGetLastError(), "transpose value map");
    }
    if (st == SPTRSV_OK) {
        priv_of(dst)->src_perm = vals;
        priv_of(dst)->src_count = sp->src_count;
        vals = nullptr;
    }
    for (int* p : {keys, vals, tk, tv})
        if (p) (void)hipFree(p);
    if (st == SPTRSV_OK) chk(hipDeviceSynchronize(), "transpose sync");
    if (st != SPTRSV_OK) free_mat(dst);
    return st;
}

}  // extern "C"
