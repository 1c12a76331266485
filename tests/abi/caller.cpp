// A C++ caller of libsptrsv_amd built against include/sptrsv_amd.h alone, as
// the reference's host code would be after the switch.
//
// 1. The reference's structs (dfr_syncfree.h:7-41, restated below as plain
//    layouts for VALUE_TYPE = double and float) must be a prefix of the
//    library's: same field names, same offsets, same sizes, so code written
//    against the reference's fields reads the same memory.
// 2. The reference's test flow (test.cu:4-30) written against the library:
//    upload, analyse, b = L * 1, solve, check x == 1, free.
// 3. The reference's top-level entry (test.h:4) called as main.cu:259 does.
//
// Without arguments the program only touches host-side entry points (no GPU
// needed); with "run" it solves the 64 x 64 2D Laplacian factor on the GPU.
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <vector>

#include "sptrsv_amd.h"

namespace ref {
enum MODE { SIMPLE, ORDER, MULTIROW, ALL };
template <class VALUE_TYPE>
struct sp_mat_t {
    int nr;
    int nc;
    int nnz;
#ifdef PLANT_SWAPPED_FIELDS  // a planted mistake: the layout checks must reject it
    int* ja;
    int* ia;
#else
    int* ia;
    int* ja;
#endif
    VALUE_TYPE* a;
};
struct dfr_analysis_info_t {
    MODE mode;
    int lev_ctr;
    int nlevs;
    int* lev_size;
    int* warp_lev;
    int n_warps;
    int* inv_iorder;
    int* iorder;
    int* ibase_row;
    int* ivect_size_warp;
    int* row_ctr;
};
}  // namespace ref

#define SAME_FIELD(LIB, REF, f)                                                                     \
    static_assert(offsetof(LIB, f) == offsetof(REF, f), #LIB "." #f ": offset differs");             \
    static_assert(sizeof(((LIB*)nullptr)->f) == sizeof(((REF*)nullptr)->f), #LIB "." #f ": size differs")

using RefMat64 = ref::sp_mat_t<double>;
using RefMat32 = ref::sp_mat_t<float>;
SAME_FIELD(sp_mat_t, RefMat64, nr);
SAME_FIELD(sp_mat_t, RefMat64, nc);
SAME_FIELD(sp_mat_t, RefMat64, nnz);
SAME_FIELD(sp_mat_t, RefMat64, ia);
SAME_FIELD(sp_mat_t, RefMat64, ja);
SAME_FIELD(sp_mat_t, RefMat64, a);
// VALUE_TYPE = float: the same offsets (a is a pointer either way); such a
// caller uploads with allocate_memory_f32, after which a is NULL and the
// values are in a32.
SAME_FIELD(sp_mat_t, RefMat32, a);
static_assert(sizeof(sp_mat_t) >= sizeof(RefMat64), "sp_mat_t shorter than the reference's");

SAME_FIELD(dfr_analysis_info_t, ref::dfr_analysis_info_t, mode);
SAME_FIELD(dfr_analysis_info_t, ref::dfr_analysis_info_t, lev_ctr);
SAME_FIELD(dfr_analysis_info_t, ref::dfr_analysis_info_t, nlevs);
SAME_FIELD(dfr_analysis_info_t, ref::dfr_analysis_info_t, lev_size);
SAME_FIELD(dfr_analysis_info_t, ref::dfr_analysis_info_t, warp_lev);
SAME_FIELD(dfr_analysis_info_t, ref::dfr_analysis_info_t, n_warps);
SAME_FIELD(dfr_analysis_info_t, ref::dfr_analysis_info_t, inv_iorder);
SAME_FIELD(dfr_analysis_info_t, ref::dfr_analysis_info_t, iorder);
SAME_FIELD(dfr_analysis_info_t, ref::dfr_analysis_info_t, ibase_row);
SAME_FIELD(dfr_analysis_info_t, ref::dfr_analysis_info_t, ivect_size_warp);
SAME_FIELD(dfr_analysis_info_t, ref::dfr_analysis_info_t, row_ctr);
static_assert(sizeof(dfr_analysis_info_t) >= sizeof(ref::dfr_analysis_info_t),
              "dfr_analysis_info_t shorter than the reference's");
static_assert((int)SIMPLE == (int)ref::SIMPLE && (int)ORDER == (int)ref::ORDER &&
                  (int)MULTIROW == (int)ref::MULTIROW && (int)ALL == (int)ref::ALL,
              "MODE values differ");

// The reference's test flow (test.cu:4-30) against the library.  Returns the
// number of x != 1, or a negative status.
static int reference_flow(const char* filename, int* ia, int* ja, double* a, int n) {
    sp_mat_t L;
    int st = allocate_memory(ia, ja, a, n, &L, ALL);
    if (st != SPTRSV_OK) {
        std::fprintf(stderr, "%s: %s\n", filename, sptrsv_last_error());
        return st;
    }
    dfr_analysis_info_t info = {};
    info.mode = MULTIROW;
    st = run_analysis(&info, &L);  // as interface.cu:28-38
    std::vector<double> b(n, 0.0), x(n, 0.0);
    for (int i = 0; i < n; ++i)  // b = L * ones, test.cu:13-16
        for (int j = ia[i]; j < ia[i + 1]; ++j) b[i] += a[j];
    if (st == SPTRSV_OK) st = run_solver(&L, &info, b.data(), x.data(), n);
    if (st != SPTRSV_OK) std::fprintf(stderr, "solve: %s\n", sptrsv_last_error());
    free_device_analysis(&L, &info);
    if (st != SPTRSV_OK) return st;
    int errors = 0;
    for (int i = 0; i < n; ++i) errors += (x[i] != 1.0);  // validate_x, matrix_properties.h:7-29
    return errors;
}

int main(int argc, char** argv) {
    // host-side entry points: no device needed
    if (std::strcmp(sptrsv_status_string(SPTRSV_OK), "ok") != 0) return 10;
    if (std::strcmp(sptrsv_status_string(SPTRSV_E_NO_ANALYSIS_NEEDED), "unknown status") == 0) return 11;
    int n = 0, nnz = 0, k = 0;
    int *ia = nullptr, *ja = nullptr;
    double *a = nullptr, *b = nullptr;
    if (sptrsv_generate(1, 64, 1, &n, &nnz, &ia, &ja, &a, &k, &b) != SPTRSV_OK || n != 4096) return 12;
    int rc = 0;
    if (argc > 1 && std::strcmp(argv[1], "run") == 0) {
        const int e1 = reference_flow("laplace2d_64", ia, ja, a, n);
        const int e2 = test_solve_L_analysis_multirow("laplace2d_64", ia, ja, a, n);  // as main.cu:259
        std::vector<float> a32(a, a + nnz);
        const int e3 = test_solve_L_analysis_multirow_f32("laplace2d_64 (float)", ia, ja, a32.data(), n);
        std::printf("reference flow: %d errors; test_solve_L_analysis_multirow: %d; _f32: %d\n", e1, e2, e3);
        rc = (e1 == 0 && e2 == 0 && e3 == 0) ? 0 : 1;
    } else {
        std::printf("host entry points ok: n = %d, nnz = %d\n", n, nnz);
    }
    sptrsv_free_host(ia);
    sptrsv_free_host(ja);
    sptrsv_free_host(a);
    sptrsv_free_host(b);
    return rc;
}
