// ingest.hip -- the reference driver's matrix preparation (main.cu:124-246)
// on the device: COO in file order -> CSR of A with symmetric mirroring ->
// lower triangle with an explicit diagonal, uploaded as a solver handle.
//
// Semantics kept from the driver (and from sptrsv_read_mtx_lower):
//  * A's CSR keeps, within each row, the order in which entries arrive: COO
//    entry e adds (r, c) to row r and, for symmetric input with r != c, its
//    mirror (c, r) to row c right after it (main.cu:160-212);
//  * the lower triangle of row i copies the row's entries in that order up
//    to the first one with column >= i, then appends the diagonal: 1.0 for
//    unit matrices, else the value of the row's first (i, i) entry, or 1.0 if
//    it has none (main.cu:216-246).
//
// Device pipeline: every emitted entry gets its emission index g (a scan of
// 1 or 2 per COO entry); a stable radix sort by row of (row, g) lists each
// row's entries in arrival order; one thread per row finds the cut and the
// diagonal; a scan of the kept lengths gives the row pointers; the kept
// entries and diagonals are then written in place.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace sptrsv {

int convert_values_to_f32(sp_mat_t* m);

namespace {

unsigned grid_for(long long count) {
    return (unsigned)std::max<long long>(1, std::min<long long>((count + 255) / 256, 16384));
}

// emit[e] = entries COO entry e adds to A (1, or 2 for a mirrored pair);
// bad is set for an index outside [0, n).
__global__ void k_ing_count(const int* __restrict__ row, const int* __restrict__ col, int nz, int n, int sym,
                            int* __restrict__ emit, int* __restrict__ bad) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nz;
         e += (long long)gridDim.x * blockDim.x) {
        const int r = row[e], c = col[e];
        if (r < 0 || r >= n || c < 0 || c >= n) {
            atomicOr(bad, 1);
            emit[e] = 0;
        } else {
            emit[e] = (sym && r != c) ? 2 : 1;
        }
    }
}

// Emission g = pos[e] (+1 for the mirror): keys[g] = its row, vals[g] = g,
// ecol[g] = its column, esrc[g] = e (the value index); cnt counts rows.
__global__ void k_ing_emit(const int* __restrict__ row, const int* __restrict__ col, const int* __restrict__ pos,
                           const int* __restrict__ emit, int nz, int* __restrict__ keys, int* __restrict__ vals,
                           int* __restrict__ ecol, int* __restrict__ esrc, int* __restrict__ cnt) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nz;
         e += (long long)gridDim.x * blockDim.x) {
        const int g = pos[e], r = row[e], c = col[e];
        keys[g] = r;
        vals[g] = g;
        ecol[g] = c;
        esrc[g] = (int)e;
        atomicAdd(&cnt[r], 1);
        if (emit[e] == 2) {
            keys[g + 1] = c;
            vals[g + 1] = g + 1;
            ecol[g + 1] = r;
            esrc[g + 1] = (int)e;
            atomicAdd(&cnt[c], 1);
        }
    }
}

// Row i of A is order[rp[i] .. rp[i+1]) (emission indices in arrival order):
// keep[i] = entries before the first column >= i, dsrc[i] = the COO entry
// giving the diagonal value (-1: the constant 1.0); len[i] = keep[i] + 1 is
// the row's length in L.
__global__ void k_ing_rows(const int* __restrict__ rp, const int* __restrict__ order, const int* __restrict__ ecol,
                           const int* __restrict__ esrc, int n, int unit, int* __restrict__ keep,
                           int* __restrict__ dsrc, int* __restrict__ len) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int s = rp[i], e = rp[i + 1];
        int cut = e;
        int d = -1;
        bool have_d = unit != 0;
        for (int t = s; t < e; ++t) {
            const int g = order[t];
            const int c = ecol[g];
            if (c >= i && cut == e) cut = t;
            if (!have_d && c == i) {
                d = esrc[g];
                have_d = true;
            }
            if (cut != e && have_d) break;
        }
        keep[i] = cut - s;
        dsrc[i] = d;
        len[i] = cut - s + 1;
    }
}

// Writes L: kept entries of row i at ia[i] + local position, the diagonal
// last; perm records each value's COO entry (-1: the constant 1.0).
__global__ void k_ing_fill(const int* __restrict__ rp, const int* __restrict__ order,
                           const int* __restrict__ rows_sorted, const int* __restrict__ ecol,
                           const int* __restrict__ esrc, const double* __restrict__ val,
                           const int* __restrict__ keep, const int* __restrict__ ia, int total, int unit,
                           int* __restrict__ ja, double* __restrict__ a, int* __restrict__ perm) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = rows_sorted[t];
        const int local = (int)t - rp[i];
        if (local < keep[i]) {
            const int g = order[t];
            ja[ia[i] + local] = ecol[g];
            a[ia[i] + local] = unit ? 1.0 : val[esrc[g]];
            perm[ia[i] + local] = unit ? -1 : esrc[g];
        }
    }
}

__global__ void k_ing_diag(const int* __restrict__ ia, const int* __restrict__ dsrc, const double* __restrict__ val,
                           int n, int* __restrict__ ja, double* __restrict__ a, int* __restrict__ perm) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int q = ia[i + 1] - 1;
        ja[q] = (int)i;
        a[q] = dsrc[i] >= 0 ? val[dsrc[i]] : 1.0;
        perm[q] = dsrc[i];
    }
}

}  // namespace
}  // namespace sptrsv

using namespace sptrsv;

extern "C" {

int sptrsv_coo_to_lower(const int* row, const int* col, const double* val, int nz, int n, int symmetric, int unit,
                        sp_mat_t* gpu_L) {
    if (!gpu_L) return fail(SPTRSV_E_INVALID_ARG, "null sp_mat_t");
    if (n < 0 || nz < 0 || (nz > 0 && (!row || !col || !val))) return fail(SPTRSV_E_INVALID_ARG, "bad COO arguments");
    if (nz > 0x3fffffff) return fail(SPTRSV_E_CAPACITY, "more than 2^30 COO entries");
    memset(gpu_L, 0, sizeof(*gpu_L));
    SPT_TRY(ensure_priv(gpu_L));
    gpu_L->nr = gpu_L->nc = n;
    int st = SPTRSV_OK;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && st == SPTRSV_OK) st = hip_fail(e, what, __FILE__, __LINE__);
    };
    std::vector<void*> tmp;
    auto dalloc = [&](void** p, size_t bytes, const char* what) {
        *p = nullptr;
        chk(hipMalloc(p, std::max<size_t>(bytes, 8)), what);
        if (*p) tmp.push_back(*p);
    };
    const size_t z = (size_t)nz;
    int *d_row, *d_col, *emit, *pos, *bad, *cnt, *keep, *len, *dsrc;
    double* d_val;
    dalloc((void**)&d_row, 4 * z, "hipMalloc row");
    dalloc((void**)&d_col, 4 * z, "hipMalloc col");
    dalloc((void**)&d_val, 8 * z, "hipMalloc val");
    dalloc((void**)&emit, 4 * (z + 1), "hipMalloc emit");
    dalloc((void**)&pos, 4 * (z + 1), "hipMalloc pos");
    dalloc((void**)&bad, 4, "hipMalloc flag");
    dalloc((void**)&cnt, 4 * ((size_t)n + 1), "hipMalloc counts");
    dalloc((void**)&keep, 4 * ((size_t)n + 1), "hipMalloc keep");
    dalloc((void**)&len, 4 * ((size_t)n + 1), "hipMalloc lengths");
    dalloc((void**)&dsrc, 4 * ((size_t)n + 1), "hipMalloc diagonal sources");
    if (st == SPTRSV_OK && nz > 0) {
        chk(hipMemcpy(d_row, row, 4 * z, hipMemcpyHostToDevice), "copy row");
        chk(hipMemcpy(d_col, col, 4 * z, hipMemcpyHostToDevice), "copy col");
        chk(hipMemcpy(d_val, val, 8 * z, hipMemcpyHostToDevice), "copy val");
    }
    if (st == SPTRSV_OK) {
        chk(hipMemset(bad, 0, 4), "clear flag");
        chk(hipMemset(emit, 0, 4 * (z + 1)), "clear emit");
        chk(hipMemset(cnt, 0, 4 * ((size_t)n + 1)), "clear counts");
        chk(hipMemset(len, 0, 4 * ((size_t)n + 1)), "clear lengths");
    }
    if (st == SPTRSV_OK && nz > 0) {
        hipLaunchKernelGGL(k_ing_count, dim3(grid_for(nz)), dim3(256), 0, nullptr, d_row, d_col, nz, n, symmetric,
                           emit, bad);
        chk(hipGetLastError(), "ingest count");
    }
    int h_bad = 0, total = 0;
    if (st == SPTRSV_OK) chk(hipMemcpy(&h_bad, bad, 4, hipMemcpyDeviceToHost), "read flag");
    if (st == SPTRSV_OK && h_bad) st = fail(SPTRSV_E_INVALID_CSR, "COO index outside [0, n)");
    if (st == SPTRSV_OK) st = exclusive_scan(emit, pos, (long long)z + 1, nullptr);
    if (st == SPTRSV_OK) chk(hipMemcpy(&total, pos + z, 4, hipMemcpyDeviceToHost), "read total");
    const size_t M = (size_t)total;
    int *keys = nullptr, *vals = nullptr, *tk = nullptr, *tv = nullptr, *ecol = nullptr, *esrc = nullptr;
    dalloc((void**)&keys, 4 * M, "hipMalloc keys");
    dalloc((void**)&vals, 4 * M, "hipMalloc order");
    dalloc((void**)&tk, 4 * M, "hipMalloc sort keys");
    dalloc((void**)&tv, 4 * M, "hipMalloc sort vals");
    dalloc((void**)&ecol, 4 * M, "hipMalloc columns");
    dalloc((void**)&esrc, 4 * M, "hipMalloc sources");
    if (st == SPTRSV_OK && nz > 0) {
        hipLaunchKernelGGL(k_ing_emit, dim3(grid_for(nz)), dim3(256), 0, nullptr, d_row, d_col, pos, emit, nz, keys,
                           vals, ecol, esrc, cnt);
        chk(hipGetLastError(), "ingest emit");
    }
    if (st == SPTRSV_OK && n > 0) st = radix_sort_pairs(keys, vals, tk, tv, total, n - 1, nullptr);
    if (st == SPTRSV_OK) st = exclusive_scan(cnt, cnt, (long long)n + 1, nullptr);  // rows of A
    if (st == SPTRSV_OK && n > 0) {
        hipLaunchKernelGGL(k_ing_rows, dim3(grid_for(n)), dim3(256), 0, nullptr, cnt, vals, ecol, esrc, n, unit, keep,
                           dsrc, len);
        chk(hipGetLastError(), "ingest rows");
    }
    // row pointers of L
    if (st == SPTRSV_OK) chk(hipMalloc(&gpu_L->ia, 4 * ((size_t)n + 1)), "hipMalloc ia");
    if (st == SPTRSV_OK) st = exclusive_scan(len, gpu_L->ia, (long long)n + 1, nullptr);
    std::vector<int> ia_h((size_t)n + 1, 0);
    if (st == SPTRSV_OK) chk(hipMemcpy(ia_h.data(), gpu_L->ia, 4 * ia_h.size(), hipMemcpyDeviceToHost), "read ia");
    const long long nnzL = ia_h[n];
    if (st == SPTRSV_OK && nnzL > 0x7fffffffLL) st = fail(SPTRSV_E_CAPACITY, "lower triangle exceeds int32 nnz");
    gpu_L->nnz = (int)nnzL;
    if (st == SPTRSV_OK) chk(hipMalloc(&gpu_L->ja, 4 * (size_t)std::max(nnzL, 1LL)), "hipMalloc ja");
    if (st == SPTRSV_OK) chk(hipMalloc(&gpu_L->a, 8 * (size_t)std::max(nnzL, 1LL)), "hipMalloc a");
    if (st == SPTRSV_OK) chk(hipMalloc(&gpu_L->a32, 4 * (size_t)std::max(nnzL, 1LL)), "hipMalloc a32");
    MatPriv* pv = priv_of(gpu_L);
    if (st == SPTRSV_OK) chk(hipMalloc(&pv->src_perm, 4 * (size_t)std::max(nnzL, 1LL)), "hipMalloc value map");
    pv->src_count = nz;
    if (st == SPTRSV_OK && total > 0) {
        hipLaunchKernelGGL(k_ing_fill, dim3(grid_for(total)), dim3(256), 0, nullptr, cnt, vals, keys, ecol, esrc,
                           d_val, keep, gpu_L->ia, total, unit, gpu_L->ja, gpu_L->a, pv->src_perm);
        chk(hipGetLastError(), "ingest fill");
    }
    if (st == SPTRSV_OK && n > 0) {
        hipLaunchKernelGGL(k_ing_diag, dim3(grid_for(n)), dim3(256), 0, nullptr, gpu_L->ia, dsrc, d_val, n,
                           gpu_L->ja, gpu_L->a, pv->src_perm);
        chk(hipGetLastError(), "ingest diagonal");
    }
    if (st == SPTRSV_OK) {
        int mr = 0;
        for (int i = 0; i < n; ++i) mr = std::max(mr, ia_h[i + 1] - ia_h[i]);
        pv->max_row = mr;
        chk(hipDeviceSynchronize(), "ingest sync");
    }
    if (st == SPTRSV_OK) st = convert_values_to_f32(gpu_L);
    for (void* p : tmp) (void)hipFree(p);
    if (st == SPTRSV_OK) chk(hipDeviceSynchronize(), "ingest sync");
    if (st != SPTRSV_OK) free_mat(gpu_L);
    return st;
}

}  // extern "C"
