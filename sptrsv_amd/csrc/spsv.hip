// spsv.hip -- a generic triangular-solve handle over a device CSR matrix, in
// the shape of the vendor sparse libraries' SpSV call sequence (create,
// analysis, solve, destroy): fill mode, diagonal type and operation are
// arguments instead of properties of the upload.
//
// The handle is built on the device from the caller's CSR:
//  * only the chosen triangle is read (entries of the other one are
//    ignored);
//  * a unit diagonal ignores any stored diagonal entry; a non-unit diagonal
//    takes the row's first stored (i, i) entry, and a row without one is
//    SPTRSV_E_INVALID_CSR;
//  * a lower triangle is stored as is with the diagonal moved last; an upper
//    one is stored as the lower triangle of the reversed index order (row
//    n-1-i, columns n-1-j), which the solves undo on b and x;
//  * the transposed operation is the device transpose of that handle.
// The caller's arrays are not referenced after creation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace sptrsv {

int convert_values_to_f32(sp_mat_t* m);

namespace {

unsigned row_grid(long long count) {
    return (unsigned)std::max<long long>(1, std::min<long long>((count + 255) / 256, 16384));
}

// Stored-row lengths: kept off-diagonals + 1 (the diagonal) at the stored
// row (i, or n-1-i for an upper triangle); dpos[i] = the diagonal entry used
// (-1 for a unit diagonal); bad bit 1 = column outside [0, n), bit 2 = a
// missing diagonal.
__global__ void k_tri_count(const int* __restrict__ ia, const int* __restrict__ ja, int n, int upper, int unit,
                            int* __restrict__ len, int* __restrict__ dpos, int* __restrict__ bad) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)t;
        int kept = 0, d = -1, flags = 0;
        for (int p = ia[i]; p < ia[i + 1]; ++p) {
            const int j = ja[p];
            if (j < 0 || j >= n) flags |= 1;
            else if (j == i) { if (d < 0) d = p; }
            else if (upper ? j > i : j < i) ++kept;
        }
        if (!unit && d < 0) flags |= 2;
        if (flags) atomicOr(bad, flags);
        len[upper ? n - 1 - i : i] = kept + 1;
        dpos[i] = unit ? -1 : d;
    }
}

// Writes stored row r of source row i: its kept entries in source order
// (columns mapped to n-1-j for an upper triangle), then the diagonal.
__global__ void k_tri_fill(const int* __restrict__ ia, const int* __restrict__ ja, const double* __restrict__ a,
                           int n, int upper, const int* __restrict__ dpos, const int* __restrict__ ia_s,
                           int* __restrict__ ja_s, double* __restrict__ a_s, int* __restrict__ perm) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)t;
        const int r = upper ? n - 1 - i : i;
        int w = ia_s[r];
        for (int p = ia[i]; p < ia[i + 1]; ++p) {
            const int j = ja[p];
            if (j < 0 || j >= n || j == i || (upper ? j < i : j > i)) continue;
            ja_s[w] = upper ? n - 1 - j : j;
            a_s[w] = a[p];
            perm[w] = p;
            ++w;
        }
        ja_s[w] = r;
        a_s[w] = dpos[i] >= 0 ? a[dpos[i]] : 1.0;
        perm[w] = dpos[i];
    }
}

template <class T>
__global__ void k_scale(T* __restrict__ x, long long count, T alpha) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < count;
         t += (long long)gridDim.x * blockDim.x)
        x[t] *= alpha;
}

// The triangle of (ia, ja, a) as a handle of the non-transposed operation.
int build_triangle(int n, int nnz, const int* ia, const int* ja, const double* a, bool upper, bool unit,
                   sp_mat_t* m) {
    memset(m, 0, sizeof(*m));
    SPT_TRY(ensure_priv(m));
    m->nr = m->nc = n;
    priv_of(m)->reversed = upper;
    int st = SPTRSV_OK;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && st == SPTRSV_OK) st = hip_fail(e, what, __FILE__, __LINE__);
    };
    int *len = nullptr, *dpos = nullptr, *bad = nullptr;
    chk(hipMalloc(&len, 4 * ((size_t)n + 1)), "hipMalloc lengths");
    chk(hipMalloc(&dpos, 4 * (size_t)std::max(n, 1)), "hipMalloc diagonal positions");
    chk(hipMalloc(&bad, 4), "hipMalloc flag");
    chk(hipMalloc(&m->ia, 4 * ((size_t)n + 1)), "hipMalloc ia");
    if (st == SPTRSV_OK) chk(hipMemset(len, 0, 4 * ((size_t)n + 1)), "clear lengths");
    if (st == SPTRSV_OK) chk(hipMemset(bad, 0, 4), "clear flag");
    if (st == SPTRSV_OK && n > 0) {
        hipLaunchKernelGGL(k_tri_count, dim3(row_grid(n)), dim3(256), 0, nullptr, ia, ja, n, upper ? 1 : 0,
                           unit ? 1 : 0, len, dpos, bad);
        chk(hipGetLastError(), "triangle count");
    }
    int h_bad = 0;
    if (st == SPTRSV_OK) chk(hipMemcpy(&h_bad, bad, 4, hipMemcpyDeviceToHost), "read flag");
    if (st == SPTRSV_OK && (h_bad & 1)) st = fail(SPTRSV_E_INVALID_CSR, "column index outside [0, n)");
    if (st == SPTRSV_OK && (h_bad & 2)) st = fail(SPTRSV_E_INVALID_CSR, "a row has no diagonal entry (non-unit diagonal)");
    if (st == SPTRSV_OK) st = exclusive_scan(len, m->ia, (long long)n + 1, nullptr);
    std::vector<int> ia_h((size_t)n + 1, 0);
    if (st == SPTRSV_OK) chk(hipMemcpy(ia_h.data(), m->ia, 4 * ia_h.size(), hipMemcpyDeviceToHost), "read ia");
    const int total = ia_h[n];
    m->nnz = total;
    if (st == SPTRSV_OK) chk(hipMalloc(&m->ja, 4 * (size_t)std::max(total, 1)), "hipMalloc ja");
    if (st == SPTRSV_OK) chk(hipMalloc(&m->a, 8 * (size_t)std::max(total, 1)), "hipMalloc a");
    if (st == SPTRSV_OK) chk(hipMalloc(&m->a32, 4 * (size_t)std::max(total, 1)), "hipMalloc a32");
    if (st == SPTRSV_OK) chk(hipMalloc(&priv_of(m)->src_perm, 4 * (size_t)std::max(total, 1)), "hipMalloc value map");
    priv_of(m)->src_count = nnz;
    if (st == SPTRSV_OK && n > 0) {
        hipLaunchKernelGGL(k_tri_fill, dim3(row_grid(n)), dim3(256), 0, nullptr, ia, ja, a, n, upper ? 1 : 0, dpos,
                           m->ia, m->ja, m->a, priv_of(m)->src_perm);
        chk(hipGetLastError(), "triangle fill");
    }
    if (st == SPTRSV_OK) {
        int mr = 0;
        for (int r = 0; r < n; ++r) mr = std::max(mr, ia_h[r + 1] - ia_h[r]);
        priv_of(m)->max_row = mr;
        chk(hipDeviceSynchronize(), "triangle sync");
    }
    if (st == SPTRSV_OK) st = convert_values_to_f32(m);
    for (int* p : {len, dpos, bad})
        if (p) (void)hipFree(p);
    if (st != SPTRSV_OK) free_mat(m);
    return st;
}

}  // namespace
}  // namespace sptrsv

using namespace sptrsv;

struct sptrsv_spsv {
    sp_mat_t mat{};
    dfr_analysis_info_t info{};
    bool analysed = false;
};

extern "C" {

int sptrsv_spsv_create(sptrsv_spsv_t** out, int n, int nnz, const int* d_ia, const int* d_ja, const double* d_a,
                       int fill, int diag, int op) {
    if (!out) return fail(SPTRSV_E_INVALID_ARG, "null handle pointer");
    *out = nullptr;
    if (n < 0 || nnz < 0 || (n > 0 && (!d_ia || (nnz > 0 && (!d_ja || !d_a)))))
        return fail(SPTRSV_E_INVALID_ARG, "bad CSR arguments");
    if ((fill != SPTRSV_FILL_LOWER && fill != SPTRSV_FILL_UPPER) ||
        (diag != SPTRSV_DIAG_NON_UNIT && diag != SPTRSV_DIAG_UNIT) ||
        (op != SPTRSV_OP_NONE && op != SPTRSV_OP_TRANSPOSE))
        return fail(SPTRSV_E_INVALID_ARG, "unknown fill mode, diagonal type or operation");
    auto* h = new sptrsv_spsv();
    sp_mat_t tri{};
    int st = build_triangle(n, nnz, d_ia, d_ja, d_a, fill == SPTRSV_FILL_UPPER, diag == SPTRSV_DIAG_UNIT, &tri);
    if (st == SPTRSV_OK && op == SPTRSV_OP_TRANSPOSE) {
        st = sptrsv_transpose(&tri, &h->mat);
        free_mat(&tri);
    } else if (st == SPTRSV_OK) {
        h->mat = tri;
    }
    if (st != SPTRSV_OK) {
        delete h;
        return st;
    }
    *out = h;
    return SPTRSV_OK;
}

int sptrsv_spsv_analysis(sptrsv_spsv_t* h, int precision, int k, int tune) {
    if (!h) return fail(SPTRSV_E_INVALID_ARG, "null handle");
    if (h->analysed) free_device_analysis(nullptr, &h->info);
    h->analysed = false;
    dfr_analysis_info_t* p = &h->info;
    SPT_TRY(multirow_analysis_base_GPU(&p, &h->mat, ALL));
    h->analysed = true;
    if (tune && h->mat.nr > 0) {
        sptrsv_config_t best{};
        float ms = 0;
        SPT_TRY(sptrsv_autotune(&h->mat, &h->info, precision, std::max(k, 1), 3, &best, &ms));
    }
    return SPTRSV_OK;
}

int sptrsv_spsv_solve(sptrsv_spsv_t* h, int precision, double alpha, const void* b, void* x, int k,
                      hipStream_t stream) {
    if (!h) return fail(SPTRSV_E_INVALID_ARG, "null handle");
    SPT_TRY(sptrsv_solve(&h->mat, h->analysed ? &h->info : nullptr, precision, nullptr, b, x, h->mat.nr, k, 1,
                         stream));
    if (alpha != 1.0 && h->mat.nr > 0) {
        const long long count = (long long)h->mat.nr * k;
        if (precision == SPTRSV_PREC_F32)
            hipLaunchKernelGGL(k_scale<float>, dim3(row_grid(count)), dim3(256), 0, stream, static_cast<float*>(x),
                               count, (float)alpha);
        else
            hipLaunchKernelGGL(k_scale<double>, dim3(row_grid(count)), dim3(256), 0, stream, static_cast<double*>(x),
                               count, alpha);
        SPT_HIP(hipGetLastError());
    }
    return SPTRSV_OK;
}

sp_mat_t* sptrsv_spsv_matrix(sptrsv_spsv_t* h) { return h ? &h->mat : nullptr; }

void sptrsv_spsv_destroy(sptrsv_spsv_t* h) {
    if (!h) return;
    free_device_analysis(&h->mat, h->analysed ? &h->info : nullptr);
    delete h;
}

}  // extern "C"
