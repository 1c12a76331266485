// rocsparse_baseline.cpp -- the vendor triangular solves as reported
// baselines, in place of csr_L_solve_cusparse_v2 (solve_csr_cusparse.cu:4-29,
// declared in solver.h:8): csrsv (level-scheduled direct solve) and csritsv
// (FP64 Jacobi iteration with a residual stopping test, checked on the host).
// rocSPARSE is loaded with dlopen on first use so the solver library does
// not depend on it.  Analysis is done once at creation and timed separately
// from the solves.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <string>

#include "common.hpp"

using namespace sptrsv;

namespace {

struct Api {
    void* lib = nullptr;
    decltype(&rocsparse_create_handle) create_handle = nullptr;
    decltype(&rocsparse_destroy_handle) destroy_handle = nullptr;
    decltype(&rocsparse_set_stream) set_stream = nullptr;
    decltype(&rocsparse_create_mat_descr) create_descr = nullptr;
    decltype(&rocsparse_destroy_mat_descr) destroy_descr = nullptr;
    decltype(&rocsparse_set_mat_fill_mode) set_fill = nullptr;
    decltype(&rocsparse_set_mat_diag_type) set_diag = nullptr;
    decltype(&rocsparse_create_mat_info) create_info = nullptr;
    decltype(&rocsparse_destroy_mat_info) destroy_info = nullptr;
    decltype(&rocsparse_dcsrsv_buffer_size) buffer_size = nullptr;
    decltype(&rocsparse_dcsrsv_analysis) analysis = nullptr;
    decltype(&rocsparse_dcsrsv_solve) solve = nullptr;
    decltype(&rocsparse_csrsv_clear) clear = nullptr;
    decltype(&rocsparse_dcsritsv_buffer_size) itsv_buffer_size = nullptr;
    decltype(&rocsparse_dcsritsv_analysis) itsv_analysis = nullptr;
    decltype(&rocsparse_dcsritsv_solve_ex) itsv_solve = nullptr;
    decltype(&rocsparse_csritsv_clear) itsv_clear = nullptr;
    bool have_itsv = false;
};
Api g_api;

template <class F>
bool sym(void* lib, const char* name, F& f) {
    f = reinterpret_cast<F>(dlsym(lib, name));
    return f != nullptr;
}

bool load() {
    Api& a = g_api;
    if (a.lib) return true;
    const char* names[] = {"librocsparse.so", "librocsparse.so.1", "/opt/rocm/lib/librocsparse.so"};
    for (const char* n : names) {
        a.lib = dlopen(n, RTLD_NOW | RTLD_LOCAL);
        if (a.lib) break;
    }
    if (!a.lib) return false;
    a.have_itsv = sym(a.lib, "rocsparse_dcsritsv_buffer_size", a.itsv_buffer_size) &&
                  sym(a.lib, "rocsparse_dcsritsv_analysis", a.itsv_analysis) &&
                  sym(a.lib, "rocsparse_dcsritsv_solve_ex", a.itsv_solve) &&
                  sym(a.lib, "rocsparse_csritsv_clear", a.itsv_clear);
    return sym(a.lib, "rocsparse_create_handle", a.create_handle) &&
           sym(a.lib, "rocsparse_destroy_handle", a.destroy_handle) &&
           sym(a.lib, "rocsparse_set_stream", a.set_stream) &&
           sym(a.lib, "rocsparse_create_mat_descr", a.create_descr) &&
           sym(a.lib, "rocsparse_destroy_mat_descr", a.destroy_descr) &&
           sym(a.lib, "rocsparse_set_mat_fill_mode", a.set_fill) &&
           sym(a.lib, "rocsparse_set_mat_diag_type", a.set_diag) &&
           sym(a.lib, "rocsparse_create_mat_info", a.create_info) &&
           sym(a.lib, "rocsparse_destroy_mat_info", a.destroy_info) &&
           sym(a.lib, "rocsparse_dcsrsv_buffer_size", a.buffer_size) &&
           sym(a.lib, "rocsparse_dcsrsv_analysis", a.analysis) &&
           sym(a.lib, "rocsparse_dcsrsv_solve", a.solve) &&
           sym(a.lib, "rocsparse_csrsv_clear", a.clear);
}

}  // namespace

struct sptrsv_rocsparse {
    bool itsv = false;  // csritsv (Jacobi) instead of csrsv
    sp_mat_t* mat = nullptr;
    rocsparse_handle handle = nullptr;
    rocsparse_mat_descr descr = nullptr;
    rocsparse_mat_info info = nullptr;
    void* buffer = nullptr;
};

#define RS_CHK(call)                                                                         \
    do {                                                                                     \
        rocsparse_status s_ = (call);                                                        \
        if (s_ != rocsparse_status_success)                                                  \
            return fail(SPTRSV_E_HIP, std::string("rocSPARSE ") + #call + " status " +        \
                                          std::to_string((int)s_));                          \
    } while (0)

extern "C" {

static int create_common(sp_mat_t* mat, bool itsv, sptrsv_rocsparse_t** out, float* analysis_ms) {
    if (!mat || !mat->priv || !out || !mat->a) return fail(SPTRSV_E_INVALID_ARG, "bad arguments");
    if (!load()) return fail(SPTRSV_E_UNAVAILABLE, "rocSPARSE library not loadable");
    Api& a = g_api;
    if (itsv && !a.have_itsv) return fail(SPTRSV_E_UNAVAILABLE, "rocSPARSE without csritsv");
    if (priv_of(mat)->reversed) return fail(SPTRSV_E_INVALID_ARG, "rocSPARSE baseline takes lower-triangular matrices");
    auto* h = new sptrsv_rocsparse();
    h->itsv = itsv;
    h->mat = mat;
    *out = h;
    RS_CHK(a.create_handle(&h->handle));
    RS_CHK(a.create_descr(&h->descr));
    RS_CHK(a.set_fill(h->descr, rocsparse_fill_mode_lower));
    RS_CHK(a.set_diag(h->descr, rocsparse_diag_type_non_unit));
    RS_CHK(a.create_info(&h->info));
    size_t bytes = 0;
    if (itsv)
        RS_CHK(a.itsv_buffer_size(h->handle, rocsparse_operation_none, mat->nr, mat->nnz, h->descr, mat->a,
                                  mat->ia, mat->ja, h->info, &bytes));
    else
        RS_CHK(a.buffer_size(h->handle, rocsparse_operation_none, mat->nr, mat->nnz, h->descr, mat->a,
                             mat->ia, mat->ja, h->info, &bytes));
    SPT_HIP(hipMalloc(&h->buffer, bytes > 0 ? bytes : 1));
    hipEvent_t e0, e1;
    SPT_HIP(hipEventCreate(&e0));
    SPT_HIP(hipEventCreate(&e1));
    SPT_HIP(hipEventRecord(e0, nullptr));
    if (itsv)
        RS_CHK(a.itsv_analysis(h->handle, rocsparse_operation_none, mat->nr, mat->nnz, h->descr, mat->a,
                               mat->ia, mat->ja, h->info, rocsparse_analysis_policy_force,
                               rocsparse_solve_policy_auto, h->buffer));
    else
        RS_CHK(a.analysis(h->handle, rocsparse_operation_none, mat->nr, mat->nnz, h->descr, mat->a,
                          mat->ia, mat->ja, h->info, rocsparse_analysis_policy_force,
                          rocsparse_solve_policy_auto, h->buffer));
    SPT_HIP(hipEventRecord(e1, nullptr));
    SPT_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SPT_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (analysis_ms) *analysis_ms = ms;
    return SPTRSV_OK;
}

int csr_L_rocsparse_create(sp_mat_t* mat, sptrsv_rocsparse_t** out, float* analysis_ms) {
    return create_common(mat, false, out, analysis_ms);
}

int csr_L_rocsparse_itsv_create(sp_mat_t* mat, sptrsv_rocsparse_t** out, float* analysis_ms) {
    return create_common(mat, true, out, analysis_ms);
}

int csr_L_solve_rocsparse_itsv(sptrsv_rocsparse_t* h, const double* b, double* x, double tol, int max_iter,
                               int free_iter, int* iters, hipStream_t stream) {
    if (!h || !h->buffer || !h->itsv) return fail(SPTRSV_E_INVALID_ARG, "rocSPARSE csritsv handle not created");
    if (max_iter < 1 || free_iter < 0 || !(tol > 0)) return fail(SPTRSV_E_INVALID_ARG, "bad iteration limits");
    Api& a = g_api;
    const double one = 1.0;
    rocsparse_int nit = max_iter;
    RS_CHK(a.set_stream(h->handle, stream));
    SPT_HIP(hipMemsetAsync(x, 0, sizeof(double) * (size_t)h->mat->nr, stream));  // y_0 = 0
    RS_CHK(a.itsv_solve(h->handle, &nit, free_iter, &tol, nullptr, rocsparse_operation_none, h->mat->nr,
                        h->mat->nnz, &one, h->descr, h->mat->a, h->mat->ia, h->mat->ja, h->info, b, x,
                        rocsparse_solve_policy_auto, h->buffer));
    // rocSPARSE returns the index k of the stopping test that succeeded (or
    // the last one); free_iter untested iterations run before each test and
    // after the last, so (k + 1) * free_iter + k iterations were performed.
    if (iters) *iters = (nit + 1) * free_iter + nit;
    return SPTRSV_OK;
}

int csr_L_solve_rocsparse(sptrsv_rocsparse_t* h, const double* b, double* x, hipStream_t stream) {
    if (!h || !h->buffer || h->itsv) return fail(SPTRSV_E_INVALID_ARG, "rocSPARSE csrsv handle not created");
    Api& a = g_api;
    const double one = 1.0;
    RS_CHK(a.set_stream(h->handle, stream));
    RS_CHK(a.solve(h->handle, rocsparse_operation_none, h->mat->nr, h->mat->nnz, &one, h->descr,
                   h->mat->a, h->mat->ia, h->mat->ja, h->info, b, x, rocsparse_solve_policy_auto,
                   h->buffer));
    return SPTRSV_OK;
}

void csr_L_rocsparse_destroy(sptrsv_rocsparse_t* h) {
    if (!h) return;
    Api& a = g_api;
    if (h->info && h->handle) {
        if (h->itsv)
            a.itsv_clear(h->handle, h->descr, h->info);
        else
            a.clear(h->handle, h->descr, h->info);
    }
    if (h->info) a.destroy_info(h->info);
    if (h->descr) a.destroy_descr(h->descr);
    if (h->handle) a.destroy_handle(h->handle);
    if (h->buffer) (void)hipFree(h->buffer);
    delete h;
}

}  // extern "C"
