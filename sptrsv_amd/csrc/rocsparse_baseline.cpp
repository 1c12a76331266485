// rocsparse_baseline.cpp -- the vendor triangular solve as a reported
// baseline, in place of csr_L_solve_cusparse_v2 (solve_csr_cusparse.cu:4-29,
// declared in solver.h:8).  rocSPARSE is loaded with dlopen on first use so
// the solver library does not depend on it.  Analysis (csrsv_analysis) is
// done once at creation and timed separately from the solves.
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

int csr_L_rocsparse_create(sp_mat_t* mat, sptrsv_rocsparse_t** out, float* analysis_ms) {
    if (!mat || !mat->priv || !out || !mat->a) return fail(SPTRSV_E_INVALID_ARG, "bad arguments");
    if (!load()) return fail(SPTRSV_E_UNAVAILABLE, "rocSPARSE library not loadable");
    Api& a = g_api;
    auto* h = new sptrsv_rocsparse();
    h->mat = mat;
    *out = h;
    RS_CHK(a.create_handle(&h->handle));
    RS_CHK(a.create_descr(&h->descr));
    RS_CHK(a.set_fill(h->descr, rocsparse_fill_mode_lower));
    RS_CHK(a.set_diag(h->descr, rocsparse_diag_type_non_unit));
    RS_CHK(a.create_info(&h->info));
    size_t bytes = 0;
    RS_CHK(a.buffer_size(h->handle, rocsparse_operation_none, mat->nr, mat->nnz, h->descr, mat->a,
                         mat->ia, mat->ja, h->info, &bytes));
    SPT_HIP(hipMalloc(&h->buffer, bytes > 0 ? bytes : 1));
    hipEvent_t e0, e1;
    SPT_HIP(hipEventCreate(&e0));
    SPT_HIP(hipEventCreate(&e1));
    SPT_HIP(hipEventRecord(e0, nullptr));
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

int csr_L_solve_rocsparse(sptrsv_rocsparse_t* h, const double* b, double* x, hipStream_t stream) {
    if (!h || !h->buffer) return fail(SPTRSV_E_INVALID_ARG, "rocSPARSE handle not created");
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
    if (h->info && h->handle) a.clear(h->handle, h->descr, h->info);
    if (h->info) a.destroy_info(h->info);
    if (h->descr) a.destroy_descr(h->descr);
    if (h->handle) a.destroy_handle(h->handle);
    if (h->buffer) (void)hipFree(h->buffer);
    delete h;
}

}  // extern "C"
