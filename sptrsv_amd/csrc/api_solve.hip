// api_solve.hip -- C-ABI solve entry points (solver.h:5-7 and the mixed,
// multi-RHS, autotune and bench additions), the two-pass refinement kernels
// and the workspace management.
#include <math.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "syncfree.hpp"

namespace sptrsv {

// ---- small device kernels -----------------------------------------------------

__global__ void k_cvt_f64_f32(const double* in, float* out, size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

// x = (double) x0
__global__ void k_widen(const float* in, double* out, size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = (double)in[i];
}

// FP64 update x += d
__global__ void k_update(double* x, const float* d, size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        x[i] += (double)d[i];
}

// FP64 residual r = b - L x over all entries of the row (diagonal included),
// rounded to FP32 for the FP32 correction solve.  A group of L entry lanes x
// K right-hand-side lanes per row, like the solves; no dependencies.
template <int K>
__global__ __launch_bounds__(256) void k_residual(int n, int k, const int* __restrict__ ia,
                                                  const int* __restrict__ ja,
                                                  const double* __restrict__ a,
                                                  const double* __restrict__ b,
                                                  const double* __restrict__ x,
                                                  float* __restrict__ r, int L, int lgL) {
    const int lgG = lgL + kLog2<K>();
    const int G = 1 << lgG;
    const int lane = threadIdx.x & 63;
    const int gl = lane & (G - 1);
    const int e = gl / K;
    const int rr = gl % K;
    const long long row = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> lgG;
    const bool live = row < n && rr < k;
    double s = 0.0;
    if (live) {
        const int rs = ia[row], re = ia[row + 1];
        for (int p = rs + e; p < re; p += L) s = __fma_rn(a[p], x[(size_t)ja[p] * k + rr], s);
    }
    for (int o = K; o < G; o <<= 1) s += __shfl_xor(s, o);
    if (live && e == 0) r[(size_t)row * k + rr] = (float)(b[(size_t)row * k + rr] - s);
}

__global__ void k_f64_to_f32_vals(const double* a, float* a32, int nnz) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
         i += (long long)gridDim.x * blockDim.x)
        a32[i] = (float)a[i];
}

static unsigned stream_grid(size_t count, unsigned threads = 256) {
    size_t g = (count + threads - 1) / threads;
    return (unsigned)std::max<size_t>(1, std::min<size_t>(g, 8192));
}

int jacobi_solve(sp_mat_t* m, const sptrsv_config_t& c, int k, const double* b, double* x, hipStream_t s);
int jacobi_tune(sp_mat_t* m, int k, int reps, sptrsv_config_t* best, float* best_ms);
int jacobi_margin(sp_mat_t* m, double* delta);
int reserve_jacobi(sp_mat_t* m, int k);

// ---- configuration --------------------------------------------------------------

static int ilog2(int v) {
    int l = 0;
    while ((1 << (l + 1)) <= v) ++l;
    return l;
}
static int pow2_ceil(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Lanes per row from the integer average nnz/n, as solve_csr_simple.cu:121-137
// picks its tile (0-2 -> 2, 3-5 -> 4, 6-11 -> 8, 12-23 -> 16, else 32).
static int reference_lanes(const sp_mat_t* m) {
    int avg = m->nr > 0 ? m->nnz / m->nr : 0;
    if (avg <= 2) return 2;
    if (avg <= 5) return 4;
    if (avg <= 11) return 8;
    if (avg <= 23) return 16;
    return 32;
}

static bool has_order(const dfr_analysis_info_t* info) { return info && info->iorder; }
// Long rows split into segments unless the configuration asks for whole rows.
static bool split_rows_on(const sptrsv_config_t& c) { return c.split_rows >= 0; }
static bool has_waves(const dfr_analysis_info_t* info) {
    return info && info->ibase_row64 && info->ivect_size_wave && info->iorder;
}

// Entry window: 4 when rows are long compared with the lanes walking them.
static int window_for(const MatPriv* pv, int lanes) { return (pv->max_row > 2 * lanes) ? 4 : 1; }

// Shape used for a kernel when only the kernel is named: one right-hand side
// takes the lean path (one row per wave for SIMPLE/ORDER, with lanes covering
// a typical row: the next power of two >= nnz/n, 4..64); several take the
// general engine with the reference's lanes rule.
static sptrsv_config_t shape_for(const sp_mat_t* m, int kernel, int k) {
    sptrsv_config_t c{};
    c.kernel = kernel;
    c.wg = 256;
    c.fused = 1;
    const int K = pow2_ceil(k);
    if (K == 1) {
        c.variant = 1;
        int avg = m->nr > 0 ? m->nnz / m->nr : 1;
        c.lanes = std::min(kWave, std::max(4, pow2_ceil(avg)));
    } else {
        c.variant = 0;
        int lanes = reference_lanes(m);
        while (lanes * K > kWave && lanes > 1) lanes >>= 1;
        c.lanes = lanes;
    }
    return c;
}

int default_config(const sp_mat_t* m, const dfr_analysis_info_t* info, int precision, int k,
                   sptrsv_config_t* out) {
    (void)precision;
    int kernel = SPTRSV_KERNEL_SIMPLE;
    if (info && info->mode == ORDER && has_order(info)) kernel = SPTRSV_KERNEL_ORDER;
    if (info && info->mode == MULTIROW && has_waves(info)) kernel = SPTRSV_KERNEL_MULTIROW;
    *out = shape_for(m, kernel, k);
    return SPTRSV_OK;
}

static int resolve_config(sp_mat_t* m, dfr_analysis_info_t* info, int precision, int k,
                          const sptrsv_config_t* cfg, sptrsv_config_t* out) {
    MatPriv* pv = priv_of(m);
    if (cfg && cfg->kernel != SPTRSV_KERNEL_AUTO) {
        *out = *cfg;
        if (cfg->lanes <= 0) {
            // lanes not given: those of the tuned shape for the kernel, else
            // of the default shape; so is the variant unless one is named
            const sptrsv_config_t& t = pv->tuned_cfg[precision];
            sptrsv_config_t d = (pv->tuned[precision] && pv->tuned_k[precision] == k && t.kernel == cfg->kernel)
                                    ? t : shape_for(m, cfg->kernel, k);
            out->lanes = d.lanes;
            if (cfg->variant == 0) out->variant = d.variant;
            if (cfg->wg <= 0) out->wg = d.wg;
        }
    } else if (pv->tuned[precision] && pv->tuned_k[precision] == k &&
               (pv->tuned_cfg[precision].kernel == SPTRSV_KERNEL_SIMPLE ||
                (pv->tuned_cfg[precision].kernel == SPTRSV_KERNEL_ORDER && has_order(info)) ||
                (pv->tuned_cfg[precision].kernel == SPTRSV_KERNEL_MULTIROW && has_waves(info)) ||
                (pv->tuned_cfg[precision].kernel == SPTRSV_KERNEL_JACOBI && precision == SPTRSV_PREC_MIXED))) {
        *out = pv->tuned_cfg[precision];
        if (cfg) out->fused = cfg->fused;
        if (!out->fused) out->reorder = 0;  // the copy serves the fused pass only
    } else {
        default_config(m, info, precision, k, out);
        if (cfg) out->fused = cfg->fused;
    }
    int K = pow2_ceil(k);
    if (out->lanes <= 0) {
        sptrsv_config_t d;
        default_config(m, info, precision, k, &d);
        out->lanes = d.lanes;
    }
    if (out->kernel == SPTRSV_KERNEL_JACOBI && (out->sweeps <= 0 || out->outer <= 0)) {
        // schedule not given: the tuned one, else 12 sweeps x 4 refinement steps
        const sptrsv_config_t& t = pv->tuned_cfg[precision];
        const bool tj = pv->tuned[precision] && pv->tuned_k[precision] == k && t.kernel == SPTRSV_KERNEL_JACOBI;
        if (out->sweeps <= 0) {
            out->sweeps = tj ? t.sweeps : 12;
            if (out->sweeps_next <= 0) out->sweeps_next = tj ? t.sweeps_next : 0;
            if (out->inner <= 0) out->inner = tj ? t.inner : 0;
        }
        if (out->outer <= 0) out->outer = tj ? t.outer : 4;
    }
    if (out->sweeps_next < 0) out->sweeps_next = 0;
    if (out->inner < 0 || out->inner > 64) return fail(SPTRSV_E_INVALID_ARG, "inner must be 0..64");
    if (out->wg != 64 && out->wg != 256 && out->wg != 1024) out->wg = 256;
    if (K > 1) {
        // several RHS: the general engine, ticketed (256 threads) or
        // persistent (256 or 1024 threads)
        if (out->kernel != SPTRSV_KERNEL_JACOBI && out->variant != 2 && out->variant != 3) out->variant = 0;
        if (!(out->variant == 2 && out->wg == 1024)) out->wg = 256;
    }
    if (out->cu_waves < 0) out->cu_waves = 0;
    if (out->claim_rounds < 0 || out->claim_rounds > 64) return fail(SPTRSV_E_INVALID_ARG, "claim_rounds must be 0..64");
    if (out->split_rows < -1 || out->split_rows > 1) return fail(SPTRSV_E_INVALID_ARG, "split_rows must be -1, 0 or 1");
    if (out->kernel == SPTRSV_KERNEL_JACOBI) {
        if (out->window < 0 || out->window > 64) return fail(SPTRSV_E_INVALID_ARG, "window must be 0..64");
    } else {
        if (out->window != 0 && out->window != 1 && out->window != 4 && out->window != 16)
            return fail(SPTRSV_E_INVALID_ARG, "window must be 0, 1, 4 or 16");
        if (out->window == 16 && (K > 1 || out->wg != 256)) out->window = 4;
    }
    if (out->lanes > kWave / K) out->lanes = kWave / K;
    if ((out->lanes & (out->lanes - 1)) != 0) return fail(SPTRSV_E_INVALID_ARG, "lanes must be a power of two");
    if (out->kernel == SPTRSV_KERNEL_ORDER && !has_order(info))
        return fail(SPTRSV_E_NOT_ANALYSED, "ORDER solve needs an analysis with iorder");
    if (out->kernel == SPTRSV_KERNEL_MULTIROW && !has_waves(info))
        return fail(SPTRSV_E_NOT_ANALYSED, "MULTIROW solve needs a MULTIROW/ALL analysis");
    if (out->kernel < SPTRSV_KERNEL_SIMPLE || out->kernel > SPTRSV_KERNEL_JACOBI)
        return fail(SPTRSV_E_INVALID_ARG, "unknown kernel");
    if (out->kernel == SPTRSV_KERNEL_JACOBI && precision != SPTRSV_PREC_MIXED)
        return fail(SPTRSV_E_INVALID_ARG, "JACOBI is the bulk solver of the MIXED path only");
    if (out->kernel != SPTRSV_KERNEL_JACOBI && (out->variant < 0 || out->variant > 3))
        return fail(SPTRSV_E_INVALID_ARG, "variant must be 0..3");
    if (out->variant == 3 && out->kernel != SPTRSV_KERNEL_JACOBI &&
        (K > 1 || out->kernel == SPTRSV_KERNEL_MULTIROW))
        return fail(SPTRSV_E_INVALID_ARG, "the in-degree walk (variant 3) takes SIMPLE or ORDER and one right-hand side");
    if (out->reorder < 0 || out->reorder > 1) return fail(SPTRSV_E_INVALID_ARG, "reorder must be 0 or 1");
    if (out->reorder == 1 &&
        ((out->kernel != SPTRSV_KERNEL_ORDER && out->kernel != SPTRSV_KERNEL_MULTIROW) || out->variant == 3 ||
         precision == SPTRSV_PREC_F32 || (precision == SPTRSV_PREC_MIXED && !out->fused)))
        return fail(SPTRSV_E_INVALID_ARG, "the level-ordered copy (reorder = 1) takes ORDER or MULTIROW, "
                                          "variants 0..2, FP64 or fused MIXED");
    return SPTRSV_OK;
}

// ---- workspaces ---------------------------------------------------------------------

static int reserve(sp_mat_t* m, int k, bool need_pub, bool need_two_pass) {
    MatPriv* pv = priv_of(m);
    size_t words = (size_t)m->nr * k;
    if (need_pub && pv->pub_words < words) {
        drop_graphs(pv);
        if (pv->pub) SPT_HIP(hipFree(pv->pub));
        pv->pub = nullptr;
        pv->pub_words = 0;
        SPT_HIP(hipMalloc(&pv->pub, std::max<size_t>(words, 1) * sizeof(uint64_t)));
        pv->pub_words = words;
    }
    if (need_two_pass && pv->resid_len < words) {
        drop_graphs(pv);
        if (pv->b32) SPT_HIP(hipFree(pv->b32));
        if (pv->corr) SPT_HIP(hipFree(pv->corr));
        pv->b32 = nullptr;
        pv->corr = nullptr;
        pv->resid_len = 0;
        SPT_HIP(hipMalloc(&pv->b32, std::max<size_t>(words, 1) * sizeof(float)));
        SPT_HIP(hipMalloc(&pv->corr, std::max<size_t>(words, 1) * sizeof(float)));
        pv->resid_len = words;
    }
    return SPTRSV_OK;
}

// ---- residency cap ---------------------------------------------------------------------
// Dynamic LDS that lets at most cu_waves waves of wg-thread workgroups share
// a CU (160 KB of LDS per CU on gfx950; at most 64 KB per workgroup, so the
// floor is two workgroups per CU).  The kernels do not touch this LDS.
constexpr unsigned kLdsPerCu = 160u * 1024u;
constexpr unsigned kStaticLdsSlack = 8u * 1024u + 512u;  // largest static LDS of a solve kernel

int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    return cus;
}

unsigned lds_pad_for(int cu_waves, int wg) {
    if (cu_waves <= 0) return 0;
    const int waves_per_wg = std::max(1, wg / kWave);
    const int wgs = std::max(1, cu_waves / waves_per_wg);
    long long pad = (long long)(kLdsPerCu / (unsigned)wgs) - (long long)kStaticLdsSlack;
    if (pad <= 0) return 0;
    return (unsigned)std::min<long long>(pad, 65536);
}

// ---- kernel timer ---------------------------------------------------------------------
static std::mutex g_kt_mu;
static int g_kt_on = 0;  // 0 off, 1 every timed launch, 2 sync-free passes only, 3 Jacobi sweeps only
static std::vector<hipEvent_t> g_kt_pool;   // pairs
static std::vector<int> g_kt_launches;      // kernel launches between each pair's events
static size_t g_kt_used = 0;

// Next free (start, stop) event pair when timing is enabled, else nullptr;
// the caller brackets `launches` consecutive launches of one kernel with it.
static thread_local bool g_capturing = false;  // no timer events inside a capture

// Library teardown: the kernel timer's event pool.
__attribute__((destructor)) static void timer_pool_teardown() {
    std::lock_guard<std::mutex> g(g_kt_mu);
    for (hipEvent_t e : g_kt_pool) (void)hipEventDestroy(e);
    g_kt_pool.clear();
    g_kt_used = 0;
}

KtPair kernel_timer_pair(hipStream_t, int launches, int kind) {
    KtPair p;
    if (g_capturing) return p;
    std::lock_guard<std::mutex> g(g_kt_mu);
    if (!g_kt_on || (g_kt_on == 2 && kind != kTimeSyncFree) || (g_kt_on == 3 && kind != kTimeSweeps)) return p;
    if (g_kt_used + 2 > g_kt_pool.size()) {
        for (int i = 0; i < 64; ++i) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return p;
            g_kt_pool.push_back(e);
        }
        g_kt_launches.resize(g_kt_pool.size() / 2, 1);
    }
    p.ev[0] = g_kt_pool[g_kt_used];
    p.ev[1] = g_kt_pool[g_kt_used + 1];
    g_kt_launches[g_kt_used / 2] = std::max(1, launches);
    g_kt_used += 2;
    return p;
}

// ---- one sync-free pass ----------------------------------------------------------------

// Words of segment partials a pass with configuration c needs (0: rows are
// not split): multirow walks with one right-hand side over an analysis with
// a segment table.
static size_t split_part_words(const dfr_analysis_info_t* info, const sptrsv_config_t& c, int k) {
    if (c.kernel != SPTRSV_KERNEL_MULTIROW || pow2_ceil(k) != 1 || !split_rows_on(c)) return 0;
    const SegTable* sg = segments_of(info);
    return sg ? 2 * (size_t)sg->n_segs : 0;
}

// ---- in-degree walk: the off-diagonal pattern by column ------------------------------
// Off-diagonal entry q of row i is the (q - i)-th off-diagonal entry overall
// (each earlier row has exactly one diagonal, last), so keys and rows land
// compacted without a scan.
__global__ void k_dep_keys(const int* __restrict__ ia, const int* __restrict__ ja, int n, int* __restrict__ keys,
                           int* __restrict__ rows, int* __restrict__ cnt) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int q = ia[i]; q < ia[i + 1] - 1; ++q) {
            keys[q - i] = ja[q];
            rows[q - i] = i;
            atomicAdd(&cnt[ja[q]], 1);
        }
}

// Counters and dependents of every row (a stable sort by column keeps each
// column's rows ascending).  Setup only: allocates, before any capture.
static int ensure_indeg(sp_mat_t* m) {
    MatPriv* pv = priv_of(m);
    if (pv->indeg) return SPTRSV_OK;
    const int n = m->nr;
    const long long off = (long long)m->nnz - n;  // every row holds its diagonal
    const size_t no = (size_t)std::max<long long>(off, 1);
    int *keys = nullptr, *tk = nullptr, *tv = nullptr;
    int st = SPTRSV_OK;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && st == SPTRSV_OK) st = hip_fail(e, what, __FILE__, __LINE__);
    };
    chk(hipMalloc(&pv->dep_ptr, sizeof(int) * ((size_t)n + 1)), "hipMalloc dependents");
    chk(hipMalloc(&pv->dep_row, sizeof(int) * no), "hipMalloc dependents");
    chk(hipMalloc(&keys, sizeof(int) * no), "hipMalloc keys");
    chk(hipMalloc(&tk, sizeof(int) * no), "hipMalloc keys");
    chk(hipMalloc(&tv, sizeof(int) * no), "hipMalloc keys");
    if (st == SPTRSV_OK) chk(hipMemset(pv->dep_ptr, 0, sizeof(int) * ((size_t)n + 1)), "clear counts");
    if (st == SPTRSV_OK && n > 0) {
        hipLaunchKernelGGL(k_dep_keys, dim3(stream_grid((size_t)n)), dim3(256), 0, nullptr, m->ia, m->ja, n, keys,
                           pv->dep_row, pv->dep_ptr);
        chk(hipGetLastError(), "dependents keys");
        if (st == SPTRSV_OK && off > 0) st = radix_sort_pairs(keys, pv->dep_row, tk, tv, (int)off, n - 1, nullptr);
        if (st == SPTRSV_OK) st = exclusive_scan(pv->dep_ptr, pv->dep_ptr, (long long)n + 1, nullptr);
    }
    if (st == SPTRSV_OK) chk(hipMalloc(&pv->indeg, sizeof(int) * (size_t)std::max(n, 1)), "hipMalloc counters");
    if (st == SPTRSV_OK) chk(hipDeviceSynchronize(), "dependents sync");
    for (int* q : {keys, tk, tv})
        if (q) (void)hipFree(q);
    if (st != SPTRSV_OK) {
        for (int** q : {&pv->indeg, &pv->dep_ptr, &pv->dep_row}) {
            if (*q) (void)hipFree(*q);
            *q = nullptr;
        }
    }
    return st;
}

// The matrix's partials workspace sized for `info`'s segment table (before a
// capture, where no allocation may happen).
int reserve_part(sp_mat_t* m, const dfr_analysis_info_t* info) {
    MatPriv* pv = priv_of(m);
    const SegTable* sg = segments_of(info);
    const size_t words = sg ? 2 * (size_t)sg->n_segs : 0;
    if (!pv || pv->part_words >= words) return SPTRSV_OK;
    drop_graphs(pv);
    if (pv->part) SPT_HIP(hipFree(pv->part));
    pv->part = nullptr;
    pv->part_words = 0;
    SPT_HIP(hipMalloc(&pv->part, words * sizeof(uint64_t)));
    pv->part_words = words;
    return SPTRSV_OK;
}

// ---- level-ordered copy --------------------------------------------------------------
// Row p of the copy is row iorder[p] of the matrix, its entries in their
// storage order with every column c renamed inv_iorder[c].  Every dependency
// of a row sits at an earlier level, hence earlier in iorder, so the copy is
// lower triangular with the diagonal still last; its analysis has the same
// levels, classes and waves, with iorder the identity (the level sort is
// stable), and is derived from the matrix's rather than recomputed.  The
// kernels then fold the same entries in the same order, and the rows of one
// wave are adjacent in memory.
__global__ void k_lo_len(const int* __restrict__ ia, const int* __restrict__ iorder, int n, int* __restrict__ len,
                         int* __restrict__ rows) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const int r = iorder[p];
        len[p] = ia[r + 1] - ia[r];
        rows[p] = r;
    }
}

// One wave per copied row.
__global__ void k_lo_fill(const int* __restrict__ ia, const int* __restrict__ ja, const double* __restrict__ a,
                          const float* __restrict__ a32, const int* __restrict__ iorder,
                          const int* __restrict__ inv, int n, const int* __restrict__ nia, int* __restrict__ nja,
                          double* __restrict__ na, float* __restrict__ na32) {
    const int lane = threadIdx.x & (kWave - 1);
    const long long waves = (long long)gridDim.x * (blockDim.x / kWave);
    for (long long p = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave; p < n; p += waves) {
        const int r = iorder[p];
        const int s = ia[r], len = ia[r + 1] - s, d = nia[p];
        for (int q = lane; q < len; q += kWave) {
            nja[d + q] = inv[ja[s + q]];
            if (a) na[d + q] = a[s + q];
            na32[d + q] = a32[s + q];
        }
    }
}

void drop_level_copy(MatPriv* pv) {
    if (!pv || !pv->lc_mat) return;
    drop_graphs(pv);  // captured solves hold the copy's arrays
    free_mat(pv->lc_mat);
    delete pv->lc_mat;
    pv->lc_mat = nullptr;
    if (pv->lc_info) {
        free_info_arrays(pv->lc_info);
        free(pv->lc_info);
        pv->lc_info = nullptr;
    }
    if (pv->lc_rows) (void)hipFree(pv->lc_rows);
    pv->lc_rows = nullptr;
    pv->lc_src = 0;
}

// The copy for analysis `info` (built when missing or made for another
// analysis).  Setup only: allocates, before any capture.
static int ensure_level_copy(sp_mat_t* m, dfr_analysis_info_t* info) {
    MatPriv* pv = priv_of(m);
    const unsigned long long id = analysis_id(info);
    if (!id || !info->iorder || !info->inv_iorder)
        return fail(SPTRSV_E_NOT_ANALYSED, "a level-ordered copy needs a level analysis made by this library");
    if (pv->lc_mat && pv->lc_src == id) return SPTRSV_OK;
    drop_level_copy(pv);
    const int n = m->nr;
    const size_t nz = (size_t)std::max(m->nnz, 1);
    sp_mat_t* C = new sp_mat_t{};
    int st = ensure_priv(C);
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && st == SPTRSV_OK) st = hip_fail(e, what, __FILE__, __LINE__);
    };
    if (st == SPTRSV_OK) {
        priv_of(C)->max_row = pv->max_row;
        C->nr = C->nc = n;
        C->nnz = m->nnz;
        chk(hipMalloc(&C->ia, sizeof(int) * ((size_t)n + 1)), "hipMalloc copy ia");
        chk(hipMalloc(&C->ja, sizeof(int) * nz), "hipMalloc copy ja");
        chk(hipMalloc(&C->a32, sizeof(float) * nz), "hipMalloc copy a32");
        if (m->a) chk(hipMalloc(&C->a, sizeof(double) * nz), "hipMalloc copy a");
        chk(hipMalloc(&pv->lc_rows, sizeof(int) * (size_t)std::max(n, 1)), "hipMalloc row map");
    }
    if (st == SPTRSV_OK) chk(hipMemset(C->ia + n, 0, sizeof(int)), "clear ia end");
    if (st == SPTRSV_OK && n > 0) {
        hipLaunchKernelGGL(k_lo_len, dim3(stream_grid((size_t)n)), dim3(256), 0, nullptr, m->ia, info->iorder, n,
                           C->ia, pv->lc_rows);
        chk(hipGetLastError(), "copy row lengths");
        if (st == SPTRSV_OK) st = exclusive_scan(C->ia, C->ia, (long long)n + 1, nullptr);
        if (st == SPTRSV_OK) {
            const unsigned g = (unsigned)std::min<long long>(65536, ((long long)n + 3) / 4);
            hipLaunchKernelGGL(k_lo_fill, dim3(g), dim3(256), 0, nullptr, m->ia, m->ja, m->a, m->a32, info->iorder,
                               info->inv_iorder, n, C->ia, C->ja, C->a, C->a32);
            chk(hipGetLastError(), "copy entries");
        }
    }
    if (st == SPTRSV_OK) chk(hipDeviceSynchronize(), "level-ordered copy");
    dfr_analysis_info_t* ci = static_cast<dfr_analysis_info_t*>(calloc(1, sizeof(dfr_analysis_info_t)));
    if (!ci && st == SPTRSV_OK) st = fail(SPTRSV_E_ALLOC, "copy analysis struct");
    if (st == SPTRSV_OK) st = level_copy_analysis(info, C, ci);
    if (st == SPTRSV_OK) st = reserve_part(m, ci);  // the copy's segments share the matrix's partials
    pv->lc_mat = C;
    pv->lc_info = ci;
    pv->lc_src = id;
    if (st != SPTRSV_OK) drop_level_copy(pv);
    return st;
}

struct LevelView {
    const sp_mat_t* mat;
    const int* row_orig;
};

// timed = false keeps the pass off the kernel timer (whose event pool may
// grow, which the allocation-free entry point must not do); part_ext, when
// given, holds the segment partials instead of the matrix's workspace; lv,
// when given, names a level-ordered copy whose arrays are walked (with the
// analysis `info` of the copy) while b and x stay in the caller's order.
template <class P>
static int run_pass(sp_mat_t* m, dfr_analysis_info_t* info, const sptrsv_config_t& c, int k,
                    const void* b, void* pub, double* x_out, hipStream_t s, bool timed = true,
                    uint64_t* part_ext = nullptr, const LevelView* lv = nullptr) {
    MatPriv* pv = priv_of(m);
    const sp_mat_t* sm = lv ? lv->mat : m;
    const int K = pow2_ceil(k);
    SolveArgs A{};
    A.n = sm->nr;
    A.k = k;
    A.ia = sm->ia;
    A.ja = sm->ja;
    A.a64 = sm->a;
    A.a32 = sm->a32;
    A.b = b;
    A.pub = pub;
    A.x_out = x_out;
    A.row_orig = lv ? lv->row_orig : nullptr;
    A.iorder = info ? info->iorder : nullptr;
    A.wbase = info ? info->ibase_row64 : nullptr;
    A.wsize = info ? info->ivect_size_wave : nullptr;
    A.n_waves = info ? info->n_waves : 0;
    A.L = c.lanes;
    A.lgL = ilog2(c.lanes);
    A.ticket = pv->counters;
    A.err = pv->counters + 1;
    A.wd_ticks = pv->watchdog_ticks;
    // the trace is kept in the caller's row order, which a copy does not walk
    A.trace = (!lv && pv->trace && pv->trace_len >= 3 * (size_t)m->nr * k) ? pv->trace : nullptr;
    A.backoff = pv->backoff;
    A.dyn_lds = pv->lds_pad ? pv->lds_pad : lds_pad_for(c.cu_waves, c.wg);
    A.claim_rounds = c.claim_rounds > 0 ? c.claim_rounds : 1;
    int kind = c.kernel == SPTRSV_KERNEL_ORDER ? KK_ORDER
               : c.kernel == SPTRSV_KERNEL_MULTIROW ? KK_MULTIROW : KK_SIMPLE;
    if (c.variant == 1 && K == 1) kind += KK_FAST_SIMPLE - KK_SIMPLE;
    if (c.variant == 3) {
        // in-degree walk (one right-hand side; SIMPLE or ORDER, checked when resolved)
        SPT_TRY(ensure_indeg(m));
        kind = c.kernel == SPTRSV_KERNEL_ORDER ? KK_INDEG_ORDER : KK_INDEG_SIMPLE;
        A.indeg = pv->indeg;
        A.dep_ptr = pv->dep_ptr;
        A.dep_row = pv->dep_row;
    }
    if (c.variant == 2) {
        kind += KK_PERSIST_SIMPLE - KK_SIMPLE;
        // persistent: the wave count itself is the residency cap
        A.dyn_lds = pv->lds_pad;
        const int per_cu = c.cu_waves > 0 ? c.cu_waves : 4;
        A.persist_grid = (long long)device_cus() * std::max(1, per_cu / std::max(1, c.wg / kWave));
    }
    // long rows split into segments: multirow walks, one right-hand side
    const size_t part_words = split_part_words(info, c, k);
    if (part_words) {
        const SegTable* sg = segments_of(info);
        if (!part_ext) SPT_TRY(reserve_part(m, info));
        A.items = sg->items;
        A.n_items = sg->n_items;
        A.seg_lo = sg->seg_lo;
        A.seg_hi = sg->seg_hi;
        A.row_seg = sg->row_seg;
        A.part = part_ext ? part_ext : pv->part;
    }
    size_t words = (size_t)m->nr * k;
    using word = typename P::word;
    hipLaunchKernelGGL((k_prepare<word>), dim3(stream_grid(words)), dim3(256), 0, s,
                       static_cast<word*>(pub), words, P::kUnsolved, pv->counters, A.part, part_words, sm->ia,
                       A.indeg, A.indeg ? m->nr : 0);
    SPT_HIP(hipGetLastError());
    int u = c.window > 0 ? c.window : window_for(pv, (kind == KK_MULTIROW || kind == KK_FAST_MULTIROW) ? 32 : c.lanes);
    const KtPair ev = timed ? kernel_timer_pair(s) : KtPair{};
    if (ev) SPT_HIP(hipEventRecord(ev[0], s));
    hipError_t e = launch_solve<P, 64>(kind, K, c.wg, u, A, s);
    if (e != hipSuccess) return hip_fail(e, "sync-free solve launch", __FILE__, __LINE__);
    if (ev) SPT_HIP(hipEventRecord(ev[1], s));
    return SPTRSV_OK;
}

static int residual_to_f32(sp_mat_t* m, const sptrsv_config_t& c, int k, const double* b,
                           const double* x, float* r, hipStream_t s) {
    const int K = pow2_ceil(k);
    int L = std::max(1, std::min(c.lanes, kWave / K));
    int lgL = ilog2(L);
    long long threads = ((long long)m->nr << (lgL + ilog2(K)));
    unsigned grid = ceil_div(threads, 256);
    if (grid == 0) return SPTRSV_OK;
#define SPT_RES(KK)                                                                          \
    hipLaunchKernelGGL((k_residual<KK>), dim3(grid), dim3(256), 0, s, m->nr, k, m->ia, m->ja, \
                       m->a, b, x, r, L, lgL)
    switch (K) {
        case 1: SPT_RES(1); break;
        case 2: SPT_RES(2); break;
        case 4: SPT_RES(4); break;
        case 8: SPT_RES(8); break;
        case 16: SPT_RES(16); break;
        case 32: SPT_RES(32); break;
        default: SPT_RES(64); break;
    }
#undef SPT_RES
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

// The solve proper (no argument checks).
static int solve_impl(sp_mat_t* m, dfr_analysis_info_t* info, int precision,
                      const sptrsv_config_t& c, const void* b, void* x, int k, int refine_steps,
                      hipStream_t s) {
    MatPriv* pv = priv_of(m);
    size_t words = (size_t)m->nr * k;
    const double* bd = static_cast<const double*>(b);
    double* xd = static_cast<double*>(x);
    int extra = refine_steps;
    if (c.reorder == 1) {
        // the level-ordered copy publishes in its own order (the matrix's
        // workspace) and reads b / writes x in the caller's (FP64 or fused
        // mixed, checked when resolved)
        SPT_TRY(ensure_level_copy(m, info));
        SPT_TRY(reserve(m, k, true, precision == SPTRSV_PREC_MIXED && refine_steps > 1));
        const LevelView lv{pv->lc_mat, pv->lc_rows};
        if (precision == SPTRSV_PREC_F64)
            return run_pass<PolF64>(m, pv->lc_info, c, k, b, pv->pub, xd, s, true, nullptr, &lv);
        SPT_TRY(run_pass<PolMixed>(m, pv->lc_info, c, k, b, pv->pub, xd, s, true, nullptr, &lv));
        extra = refine_steps - 1;
    } else if (precision == SPTRSV_PREC_F64) {
        return run_pass<PolF64>(m, info, c, k, b, x, nullptr, s);
    } else if (precision == SPTRSV_PREC_F32) {
        return run_pass<PolF32>(m, info, c, k, b, x, nullptr, s);
    } else if (c.kernel == SPTRSV_KERNEL_JACOBI) {
        return jacobi_solve(m, c, k, bd, xd, s);
    } else if (c.fused) {
        SPT_TRY(reserve(m, k, true, refine_steps > 1));
        SPT_TRY(run_pass<PolMixed>(m, info, c, k, b, pv->pub, xd, s));
        extra = refine_steps - 1;
    } else {
        SPT_TRY(reserve(m, k, false, true));
        hipLaunchKernelGGL(k_cvt_f64_f32, dim3(stream_grid(words)), dim3(256), 0, s, bd, pv->b32, words);
        SPT_HIP(hipGetLastError());
        SPT_TRY(run_pass<PolF32>(m, info, c, k, pv->b32, pv->corr, nullptr, s));
        hipLaunchKernelGGL(k_widen, dim3(stream_grid(words)), dim3(256), 0, s, pv->corr, xd, words);
        SPT_HIP(hipGetLastError());
    }
    for (int it = 0; it < extra; ++it) {
        SPT_TRY(residual_to_f32(m, c, k, bd, xd, pv->b32, s));
        SPT_TRY(run_pass<PolF32>(m, info, c, k, pv->b32, pv->corr, nullptr, s));
        hipLaunchKernelGGL(k_update, dim3(stream_grid(words)), dim3(256), 0, s, xd, pv->corr, words);
        SPT_HIP(hipGetLastError());
    }
    return SPTRSV_OK;
}

static int check_solve_args(sp_mat_t* m, int precision, const void* b, void* x, int n, int k,
                            int refine_steps) {
    if (!m || !m->priv) return fail(SPTRSV_E_INVALID_ARG, "matrix not allocated (allocate_memory)");
    if (n != m->nr) return fail(SPTRSV_E_INVALID_ARG, "n differs from the matrix row count");
    if (k < 1 || k > kWave) return fail(SPTRSV_E_CAPACITY, "right-hand sides must be 1..64 per call");
    if (n > 0 && (!b || !x)) return fail(SPTRSV_E_INVALID_ARG, "null b or x");
    if (b == x && n > 0) return fail(SPTRSV_E_INVALID_ARG, "b and x must not alias");
    if (precision == SPTRSV_PREC_F32) {
        if (!m->a32) return fail(SPTRSV_E_INVALID_ARG, "matrix has no FP32 values");
    } else if (precision == SPTRSV_PREC_F64 || precision == SPTRSV_PREC_MIXED) {
        if (!m->a) return fail(SPTRSV_E_INVALID_ARG, "matrix has no FP64 values (allocated as FP32)");
        if (precision == SPTRSV_PREC_MIXED && refine_steps < 1)
            return fail(SPTRSV_E_INVALID_ARG, "refine_steps must be >= 1");
    } else {
        return fail(SPTRSV_E_INVALID_ARG, "unknown precision");
    }
    return SPTRSV_OK;
}

// dst row i = src row n-1-i (k values of `T` per row).
template <class T>
__global__ void k_reverse_rows(const T* __restrict__ src, T* __restrict__ dst, long long n, int k) {
    const long long total = n * k;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long i = t / k;
        const int r = (int)(t - i * k);
        dst[t] = src[(n - 1 - i) * k + r];
    }
}

static int reverse_rows(const void* src, void* dst, int n, int k, size_t esz, hipStream_t s) {
    const unsigned g = stream_grid((size_t)n * k);
    if (esz == 8)
        hipLaunchKernelGGL(k_reverse_rows<double>, dim3(g), dim3(256), 0, s, static_cast<const double*>(src),
                           static_cast<double*>(dst), (long long)n, k);
    else
        hipLaunchKernelGGL(k_reverse_rows<float>, dim3(g), dim3(256), 0, s, static_cast<const float*>(src),
                           static_cast<float*>(dst), (long long)n, k);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

// solve_impl in the caller's index order: an upper-triangular matrix (kept as
// the reversed lower triangle) has b reversed on the way in and x on the way
// out.  The two workspaces are allocated on the first such solve.
static int solve_oriented(sp_mat_t* m, dfr_analysis_info_t* info, int precision, const sptrsv_config_t& c,
                          const void* b, void* x, int k, int refine_steps, hipStream_t s) {
    MatPriv* pv = priv_of(m);
    if (!pv->reversed) return solve_impl(m, info, precision, c, b, x, k, refine_steps, s);
    const size_t esz = precision == SPTRSV_PREC_F32 ? sizeof(float) : sizeof(double);
    const size_t bytes = std::max<size_t>((size_t)m->nr * k * esz, 8);
    if (pv->rev_bytes < bytes) {
        drop_graphs(pv);
        if (pv->rev_b) SPT_HIP(hipFree(pv->rev_b));
        if (pv->rev_x) SPT_HIP(hipFree(pv->rev_x));
        pv->rev_b = pv->rev_x = nullptr;
        pv->rev_bytes = 0;
        SPT_HIP(hipMalloc(&pv->rev_b, bytes));
        SPT_HIP(hipMalloc(&pv->rev_x, bytes));
        pv->rev_bytes = bytes;
    }
    SPT_TRY(reverse_rows(b, pv->rev_b, m->nr, k, esz, s));
    SPT_TRY(solve_impl(m, info, precision, c, pv->rev_b, pv->rev_x, k, refine_steps, s));
    return reverse_rows(pv->rev_x, x, m->nr, k, esz, s);
}

int solve_checked(sp_mat_t* m, dfr_analysis_info_t* info, int precision, const sptrsv_config_t* cfg,
                  const void* b, void* x, int n, int k, int refine_steps, hipStream_t s) {
    SPT_TRY(check_solve_args(m, precision, b, x, n, k, refine_steps));
    if (n == 0) return SPTRSV_OK;
    sptrsv_config_t c;
    SPT_TRY(resolve_config(m, info, precision, k, cfg, &c));
    return solve_oriented(m, info, precision, c, b, x, k, refine_steps, s);
}

// ---- allocation-free fused mixed solve ----------------------------------------------
// The published words of sptrsv_solve_mixed_noalloc live in this fixed array
// (one copy per GPU, held until the process ends) instead of the matrix's
// workspace.  Host code never passes the array's name to a kernel: its
// device address is looked up once per device with hipGetSymbolAddress.
static __device__ __align__(256) uint64_t scratch[SPTRSV_NOALLOC_MAX_WORDS];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint64_t* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array (looked up on the
// first call on that device, with it current).
static int scratch_on_current_device(uint64_t** out) {
    int dev = 0;
    SPT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(SPTRSV_E_CAPACITY, "device index beyond the scratch table");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        if (!p) return fail(SPTRSV_E_HIP, "scratch array has no device address");
        g_scratch_ptr[dev] = static_cast<uint64_t*>(p);
    }
    *out = g_scratch_ptr[dev];
    return SPTRSV_OK;
}

}  // namespace sptrsv

using namespace sptrsv;

extern "C" {

int sptrsv_solve(sp_mat_t* mat, dfr_analysis_info_t* info, int precision, const sptrsv_config_t* cfg,
                 const void* b, void* x, int n, int k, int refine_steps, hipStream_t stream) {
    return solve_checked(mat, info, precision, cfg, b, x, n, k, refine_steps, stream);
}

int sptrsv_solve_graph(sp_mat_t* mat, dfr_analysis_info_t* info, int precision, const sptrsv_config_t* cfg,
                       const void* b, void* x, int n, int k, int refine_steps, hipStream_t stream) {
    SPT_TRY(check_solve_args(mat, precision, b, x, n, k, refine_steps));
    if (n == 0) return SPTRSV_OK;
    sptrsv_config_t c;
    SPT_TRY(resolve_config(mat, info, precision, k, cfg, &c));
    MatPriv* pv = priv_of(mat);
    // the analysis is part of the key: a captured solve holds its arrays
    const std::array<long long, 21> key{precision, (long long)(uintptr_t)b, (long long)(uintptr_t)x, n, k,
                                        refine_steps, c.kernel, c.lanes, c.wg, c.fused, c.variant, c.cu_waves,
                                        c.sweeps, c.outer, c.window, c.sweeps_next, c.inner, c.claim_rounds,
                                        c.split_rows, c.reorder, (long long)analysis_id(info)};
    auto it = pv->graphs.find(key);
    if (it == pv->graphs.end()) {
        // one eager solve sets up every workspace and table (no allocation
        // may happen inside the capture)
        SPT_TRY(solve_oriented(mat, info, precision, c, b, x, k, refine_steps, stream));
        SPT_HIP(hipStreamSynchronize(stream));
        if (!pv->cap_stream) SPT_HIP(hipStreamCreate(&pv->cap_stream));
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        SPT_HIP(hipStreamBeginCapture(pv->cap_stream, hipStreamCaptureModeThreadLocal));
        g_capturing = true;
        const int st = solve_oriented(mat, info, precision, c, b, x, k, refine_steps, pv->cap_stream);
        g_capturing = false;
        hipError_t e = hipStreamEndCapture(pv->cap_stream, &g);
        if (st != SPTRSV_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return hip_fail(e, "end capture", __FILE__, __LINE__);
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return hip_fail(e, "graph instantiate", __FILE__, __LINE__);
        it = pv->graphs.emplace(key, ge).first;
    }
    SPT_HIP(hipGraphLaunch(it->second, stream));
    return SPTRSV_OK;
}

int sptrsv_default_config(const sp_mat_t* mat, const dfr_analysis_info_t* info, int precision, int k,
                          sptrsv_config_t* out) {
    if (!mat || !out || k < 1) return fail(SPTRSV_E_INVALID_ARG, "null argument");
    return default_config(mat, info, precision, k, out);
}

int csr_L_solve_simple(sp_mat_t* mat, const double* b, double* x, int n, int* is_solved) {
    (void)is_solved;
    sptrsv_config_t c{SPTRSV_KERNEL_SIMPLE, 0, 0, 1};
    return solve_checked(mat, nullptr, SPTRSV_PREC_F64, &c, b, x, n, 1, 1, nullptr);
}

int csr_L_solve_order(sp_mat_t* mat, dfr_analysis_info_t* info, const double* b, double* x, int n,
                      hipStream_t stream) {
    sptrsv_config_t c{SPTRSV_KERNEL_ORDER, 0, 0, 1};
    return solve_checked(mat, info, SPTRSV_PREC_F64, &c, b, x, n, 1, 1, stream);
}

int csr_L_solve_multirow(sp_mat_t* mat, dfr_analysis_info_t* info, const double* b, double* x, int n,
                         hipStream_t stream) {
    sptrsv_config_t c{SPTRSV_KERNEL_MULTIROW, 0, 0, 1};
    return solve_checked(mat, info, SPTRSV_PREC_F64, &c, b, x, n, 1, 1, stream);
}

int csr_L_solve_simple_f32(sp_mat_t* mat, const float* b, float* x, int n, int* is_solved) {
    (void)is_solved;
    sptrsv_config_t c{SPTRSV_KERNEL_SIMPLE, 0, 0, 1};
    return solve_checked(mat, nullptr, SPTRSV_PREC_F32, &c, b, x, n, 1, 1, nullptr);
}

int csr_L_solve_order_f32(sp_mat_t* mat, dfr_analysis_info_t* info, const float* b, float* x, int n,
                          hipStream_t stream) {
    sptrsv_config_t c{SPTRSV_KERNEL_ORDER, 0, 0, 1};
    return solve_checked(mat, info, SPTRSV_PREC_F32, &c, b, x, n, 1, 1, stream);
}

int csr_L_solve_multirow_f32(sp_mat_t* mat, dfr_analysis_info_t* info, const float* b, float* x,
                             int n, hipStream_t stream) {
    sptrsv_config_t c{SPTRSV_KERNEL_MULTIROW, 0, 0, 1};
    return solve_checked(mat, info, SPTRSV_PREC_F32, &c, b, x, n, 1, 1, stream);
}

int csr_L_solve_mixed(sp_mat_t* mat, dfr_analysis_info_t* info, const double* b, double* x, int n,
                      int k, int refine_steps, hipStream_t stream) {
    return solve_checked(mat, info, SPTRSV_PREC_MIXED, nullptr, b, x, n, k, refine_steps, stream);
}

int sptrsv_solve_mixed_noalloc(sp_mat_t* mat, dfr_analysis_info_t* info, const sptrsv_config_t* cfg,
                               const double* b, double* x, int n, int k, hipStream_t stream) {
    SPT_TRY(check_solve_args(mat, SPTRSV_PREC_MIXED, b, x, n, k, 1));
    if ((long long)n * k > (long long)SPTRSV_NOALLOC_MAX_WORDS)
        return fail(SPTRSV_E_CAPACITY, "n * k = " + std::to_string((long long)n * k) +
                                           " exceeds the allocation-free scratch of " +
                                           std::to_string((long long)SPTRSV_NOALLOC_MAX_WORDS) + " words");
    if (priv_of(mat)->reversed)
        return fail(SPTRSV_E_INVALID_ARG, "the allocation-free solve takes lower-triangular handles only");
    if (n == 0) return SPTRSV_OK;
    sptrsv_config_t c;
    SPT_TRY(resolve_config(mat, info, SPTRSV_PREC_MIXED, k, cfg, &c));
    const bool named = cfg && cfg->kernel != SPTRSV_KERNEL_AUTO;
    if (c.kernel == SPTRSV_KERNEL_JACOBI) {
        if (named)
            return fail(SPTRSV_E_INVALID_ARG, "the allocation-free solve runs the sync-free kernels only "
                                              "(JACOBI keeps per-matrix workspaces); name SIMPLE, ORDER or MULTIROW");
        // a tuned JACOBI choice: the default sync-free shape instead
        SPT_TRY(default_config(mat, info, SPTRSV_PREC_MIXED, k, &c));
    }
    c.fused = 1;
    if (c.variant == 3)
        return fail(SPTRSV_E_INVALID_ARG, "the allocation-free solve does not run the in-degree walk (its "
                                          "counters are a per-matrix workspace)");
    if (c.reorder == 1) {
        // a tuned choice walks the matrix as given instead; a named one is refused
        if (named && cfg->reorder == 1)
            return fail(SPTRSV_E_INVALID_ARG, "the allocation-free solve does not walk a level-ordered copy (the "
                                              "copy is a per-matrix workspace)");
        c.reorder = 0;
    }
    // the published words, then (split long rows) the segment partials from
    // the next 256-byte boundary
    const size_t pub_words = (size_t)n * k;
    const size_t part_words = split_part_words(info, c, k);
    const size_t part_at = (pub_words + 31) / 32 * 32;
    if (part_words && part_at + part_words > (size_t)SPTRSV_NOALLOC_MAX_WORDS)
        return fail(SPTRSV_E_CAPACITY, "n * k plus the split rows' " + std::to_string(part_words) +
                                           " partial words exceed the allocation-free scratch of " +
                                           std::to_string((long long)SPTRSV_NOALLOC_MAX_WORDS) + " words");
    uint64_t* pub = nullptr;
    SPT_TRY(scratch_on_current_device(&pub));
    return run_pass<PolMixed>(mat, info, c, k, b, pub, x, stream, false, part_words ? pub + part_at : nullptr);
}

int sptrsv_reserve(sp_mat_t* mat, int k) {
    if (!mat || !mat->priv || k < 1 || k > kWave) return fail(SPTRSV_E_INVALID_ARG, "bad reserve");
    SPT_TRY(reserve(mat, k, true, true));
    if (mat->a) SPT_TRY(reserve_jacobi(mat, k));
    return SPTRSV_OK;
}

int sptrsv_check(sp_mat_t* mat, hipStream_t stream) {
    if (!mat || !mat->priv) return fail(SPTRSV_E_INVALID_ARG, "matrix not allocated");
    MatPriv* pv = priv_of(mat);
    SPT_HIP(hipStreamSynchronize(stream));
    int w[4] = {0, 0, 0, 0};
    SPT_HIP(hipMemcpy(w, pv->counters, sizeof(w), hipMemcpyDeviceToHost));
    if (w[1] || w[3]) {
        int zero = 0;
        if (w[1]) SPT_HIP(hipMemcpy(pv->counters + 1, &zero, sizeof(int), hipMemcpyHostToDevice));
        if (w[3]) SPT_HIP(hipMemcpy(pv->counters + 3, &zero, sizeof(int), hipMemcpyHostToDevice));
    }
    if (w[1])
        return fail(SPTRSV_E_WATCHDOG, "a dependency wait exceeded the watchdog bound (input "
                                       "malformed or bound too small); affected rows hold NaN");
    if (w[3])
        return fail(SPTRSV_E_NOT_CONVERGED, "JACOBI refinement did not reach the certified tolerance; "
                                            "re-solve with a sync-free kernel or more refinement steps");
    return SPTRSV_OK;
}

int sptrsv_set_backoff(sp_mat_t* mat, int level) {
    if (!mat || !mat->priv || level < 0 || level > 4) return fail(SPTRSV_E_INVALID_ARG, "bad back-off");
    drop_graphs(priv_of(mat));  // captured launches hold the old setting
    priv_of(mat)->backoff = level;
    return SPTRSV_OK;
}

int sptrsv_set_lds_pad(sp_mat_t* mat, int bytes) {
    if (!mat || !mat->priv || bytes < 0 || bytes > 65536) return fail(SPTRSV_E_INVALID_ARG, "bad LDS pad");
    drop_graphs(priv_of(mat));  // captured launches hold the old setting
    priv_of(mat)->lds_pad = (unsigned)bytes;
    return SPTRSV_OK;
}

int sptrsv_trace(sp_mat_t* mat, int k, int enable) {
    if (!mat || !mat->priv || k < 1 || k > kWave) return fail(SPTRSV_E_INVALID_ARG, "bad trace request");
    MatPriv* pv = priv_of(mat);
    drop_graphs(pv);  // captured launches hold the trace buffer's address
    if (pv->trace) SPT_HIP(hipFree(pv->trace));
    pv->trace = nullptr;
    pv->trace_len = 0;
    if (enable) {
        size_t len = 3 * (size_t)mat->nr * k;
        SPT_HIP(hipMalloc(&pv->trace, std::max<size_t>(len, 1) * sizeof(long long)));
        SPT_HIP(hipMemset(pv->trace, 0, std::max<size_t>(len, 1) * sizeof(long long)));
        pv->trace_len = len;
    }
    return SPTRSV_OK;
}

int sptrsv_trace_read(sp_mat_t* mat, int64_t* host, size_t count, double* rate_khz) {
    if (!mat || !mat->priv || !host) return fail(SPTRSV_E_INVALID_ARG, "bad trace read");
    MatPriv* pv = priv_of(mat);
    if (!pv->trace || count > pv->trace_len) return fail(SPTRSV_E_INVALID_ARG, "trace not enabled or too short");
    SPT_HIP(hipDeviceSynchronize());
    SPT_HIP(hipMemcpy(host, pv->trace, count * sizeof(long long), hipMemcpyDeviceToHost));
    if (rate_khz) *rate_khz = pv->wall_rate_khz;
    return SPTRSV_OK;
}

int sptrsv_kernel_timer(int enable) {
    std::lock_guard<std::mutex> g(g_kt_mu);
    g_kt_on = (enable >= 0 && enable <= 3) ? enable : 1;
    g_kt_used = 0;
    return SPTRSV_OK;
}

int sptrsv_kernel_timer_read(float* total_ms, int* launches) {
    if (!total_ms || !launches) return fail(SPTRSV_E_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> g(g_kt_mu);
    double tot = 0;
    int count = 0;
    for (size_t i = 0; i + 1 < g_kt_used; i += 2) {
        SPT_HIP(hipEventSynchronize(g_kt_pool[i + 1]));
        float ms = 0;
        SPT_HIP(hipEventElapsedTime(&ms, g_kt_pool[i], g_kt_pool[i + 1]));
        tot += ms;
        count += g_kt_launches[i / 2];
    }
    *total_ms = (float)tot;
    *launches = count;
    g_kt_used = 0;
    return SPTRSV_OK;
}

int sptrsv_set_watchdog_ms(sp_mat_t* mat, double ms) {
    if (!mat || !mat->priv || !(ms > 0)) return fail(SPTRSV_E_INVALID_ARG, "bad watchdog bound");
    MatPriv* pv = priv_of(mat);
    drop_graphs(pv);  // captured launches hold the old bound
    pv->watchdog_ticks = (long long)(ms * pv->wall_rate_khz);
    if (pv->watchdog_ticks < 1) pv->watchdog_ticks = 1;
    return SPTRSV_OK;
}

int sptrsv_autotune(sp_mat_t* mat, dfr_analysis_info_t* info, int precision, int k, int reps,
                    sptrsv_config_t* best, float* best_ms) {
    if (!mat || !mat->priv || reps < 1 || k < 1 || k > kWave || precision < 0 || precision > 2)
        return fail(SPTRSV_E_INVALID_ARG, "bad autotune arguments");
    MatPriv* pv = priv_of(mat);
    const int n = mat->nr;
    const int K = pow2_ceil(k);
    size_t words = (size_t)n * k;
    size_t esz = precision == SPTRSV_PREC_F32 ? sizeof(float) : sizeof(double);
    void *b = nullptr, *x = nullptr;
    SPT_HIP(hipMalloc(&b, std::max<size_t>(words, 1) * esz));
    SPT_HIP(hipMalloc(&x, std::max<size_t>(words, 1) * esz));
    SPT_HIP(hipMemset(b, 0, std::max<size_t>(words, 1) * esz));  // any finite right-hand side
    if (precision == SPTRSV_PREC_MIXED) SPT_TRY(reserve(mat, k, true, true));

    // Shapes: (kernel, lanes, workgroup, variant).
    std::vector<sptrsv_config_t> shapes;
    int ref = std::min(reference_lanes(mat), kWave / K);
    int lo = (K > 1) ? 1 : std::max(1, ref / 2), hi = std::min(kWave / K, ref * 4);
    std::vector<int> wgs = (K == 1) ? std::vector<int>{256, 1024} : std::vector<int>{256};
    // general engine (several rows per wave)
    for (int L = lo; L <= hi; L <<= 1)
        for (int wg : wgs) shapes.push_back({SPTRSV_KERNEL_SIMPLE, L, wg, 1, 0, 0});
    // in level order a wave's rows are independent, so a wave of 64 rows
    // with one lane each is also tried (it wins on shallow matrices of many
    // short rows, such as many independent systems in one matrix)
    if (has_order(info))
        for (int L = (K == 1 ? 1 : lo); L <= hi; L <<= 1) shapes.push_back({SPTRSV_KERNEL_ORDER, L, 256, 1, 0, 0});
    if (has_waves(info))
        for (int wg : wgs) shapes.push_back({SPTRSV_KERNEL_MULTIROW, ref, wg, 1, 0, 0});
    // several right-hand sides: the persistent walk of the general engine
    // (no ticket per workgroup; the wave count is the residency cap)
    if (K > 1) {
        for (int wg : {256, 1024}) {
            for (int L = lo; L <= hi; L <<= 1) shapes.push_back({SPTRSV_KERNEL_SIMPLE, L, wg, 1, 2, 0});
            if (has_order(info))
                for (int L = lo; L <= hi; L <<= 1) shapes.push_back({SPTRSV_KERNEL_ORDER, L, wg, 1, 2, 0});
            if (has_waves(info)) shapes.push_back({SPTRSV_KERNEL_MULTIROW, ref, wg, 1, 2, 0});
        }
    }
    // lean path (one right-hand side): one row per wave, or single-level waves
    if (K == 1) {
        const int flo = std::max(4, ref), fhi = std::min(kWave, ref * 8);
        for (int L = flo; L <= fhi; L <<= 1)
            for (int wg : {64, 256}) shapes.push_back({SPTRSV_KERNEL_SIMPLE, L, wg, 1, 1, 0});
        if (has_order(info))
            for (int L = flo; L <= fhi; L <<= 1)
                for (int wg : {64, 256}) shapes.push_back({SPTRSV_KERNEL_ORDER, L, wg, 1, 1, 0});
        if (has_waves(info))
            for (int wg : {64, 256}) shapes.push_back({SPTRSV_KERNEL_MULTIROW, ref, wg, 1, 1, 0});
        // persistent lean path: a fixed number of resident waves walks the
        // rows; a workgroup's waves share each claimed block of rows
        for (int wg : {256, 1024}) {
            for (int L = flo; L <= fhi; L <<= 1) shapes.push_back({SPTRSV_KERNEL_SIMPLE, L, wg, 1, 2, 0});
            if (has_order(info))
                for (int L = flo; L <= fhi; L <<= 1) shapes.push_back({SPTRSV_KERNEL_ORDER, L, wg, 1, 2, 0});
            if (has_waves(info)) shapes.push_back({SPTRSV_KERNEL_MULTIROW, ref, wg, 1, 2, 0});
        }
    }

    hipEvent_t e0, e1;
    SPT_HIP(hipEventCreate(&e0));
    SPT_HIP(hipEventCreate(&e1));
    int status = SPTRSV_OK;
    auto time_cfg = [&](const sptrsv_config_t& c) -> float {
        // one warm-up, then reps timed solves
        status = solve_impl(mat, info, precision, c, b, x, k, 1, nullptr);
        if (status != SPTRSV_OK) return 1e30f;
        if (hipEventRecord(e0, nullptr) != hipSuccess) return 1e30f;
        for (int r = 0; r < reps && status == SPTRSV_OK; ++r)
            status = solve_impl(mat, info, precision, c, b, x, k, 1, nullptr);
        if (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return 1e30f;
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1e30f;
        return ms / reps;
    };
    // Stage 1: every shape without and with a tight residency cap.
    std::vector<std::pair<float, sptrsv_config_t>> ranked;
    for (auto c : shapes) {
        for (int cw : (c.variant == 2 ? std::vector<int>{2, 8} : std::vector<int>{0, 4})) {
            c.cu_waves = cw;
            float t = time_cfg(c);
            if (status != SPTRSV_OK) break;
            ranked.push_back({t, c});
        }
        if (status != SPTRSV_OK) break;
    }
    // Stage 2: the four best shapes across residency caps.
    float best_t = 1e30f;
    sptrsv_config_t bc = shapes.front();
    if (status == SPTRSV_OK) {
        std::sort(ranked.begin(), ranked.end(),
                  [](const auto& a, const auto& b) { return a.first < b.first; });
        best_t = ranked.front().first;
        bc = ranked.front().second;
        int taken = 0;
        for (size_t i = 0; i < ranked.size() && taken < 4 && status == SPTRSV_OK; ++i) {
            sptrsv_config_t c = ranked[i].second;
            bool dup = false;
            for (size_t j = 0; j < i; ++j) {
                const auto& d = ranked[j].second;
                dup |= d.kernel == c.kernel && d.lanes == c.lanes && d.wg == c.wg && d.variant == c.variant;
            }
            if (dup) continue;
            ++taken;
            for (int cw : (c.variant == 2 ? std::vector<int>{1, 3, 4, 6, 12, 16}
                                          : std::vector<int>{32, 16, 8, 2})) {
                c.cu_waves = cw;
                float t = time_cfg(c);
                if (status != SPTRSV_OK) break;
                if (t < best_t) {
                    best_t = t;
                    bc = c;
                }
            }
        }
    }
    // Stage 3 (long rows): the best shape of each kernel with 16 entries per
    // lane in flight instead of 4, which shortens the load chain of a long row.
    if (status == SPTRSV_OK && K == 1 && pv->max_row > 64) {
        for (int kern : {SPTRSV_KERNEL_SIMPLE, SPTRSV_KERNEL_ORDER, SPTRSV_KERNEL_MULTIROW}) {
            auto it = std::find_if(ranked.begin(), ranked.end(),
                                   [&](const auto& r) { return r.second.kernel == kern; });
            if (it == ranked.end()) continue;
            sptrsv_config_t c = (bc.kernel == kern) ? bc : it->second;
            c.wg = 256;
            c.window = 16;
            const float t = time_cfg(c);
            if (status != SPTRSV_OK) break;
            if (t < best_t) {
                best_t = t;
                bc = c;
            }
        }
    }
    // Stage 4 (persistent walk): other block sizes of the claimed walk.
    if (status == SPTRSV_OK && bc.variant == 2) {
        const sptrsv_config_t base = bc;
        for (int d : {2, 4}) {
            sptrsv_config_t c = base;
            c.claim_rounds = d;
            const float t = time_cfg(c);
            if (status != SPTRSV_OK) break;
            if (t < best_t) {
                best_t = t;
                bc = c;
            }
        }
    }
    // Stage 5 (split long rows): the multirow walk with whole rows instead.
    if (status == SPTRSV_OK && split_part_words(info, bc, k) > 0) {
        sptrsv_config_t c = bc;
        c.split_rows = -1;
        const float t = time_cfg(c);
        if (status == SPTRSV_OK && t < best_t) {
            best_t = t;
            bc = c;
        }
    }
    // Stage 6 (level-ordered copy): the best ORDER / MULTIROW shape walked on
    // the copy; the copy is released again when it loses.
    if (status == SPTRSV_OK && precision != SPTRSV_PREC_F32 && bc.variant != 3 &&
        (bc.kernel == SPTRSV_KERNEL_ORDER || bc.kernel == SPTRSV_KERNEL_MULTIROW)) {
        sptrsv_config_t c = bc;
        c.reorder = 1;
        const float t = time_cfg(c);
        if (status == SPTRSV_OK && t < best_t) {
            best_t = t;
            bc = c;
        }
        if (bc.reorder != 1) drop_level_copy(pv);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(b);
    (void)hipFree(x);
    if (status != SPTRSV_OK) return status;
    SPT_TRY(sptrsv_check(mat, nullptr));
    // Mixed precision on a strictly diagonally dominant matrix: also time the
    // certified Jacobi refinement and keep it when it is faster.
    double delta = 0;
    if (precision == SPTRSV_PREC_MIXED && mat->a && jacobi_margin(mat, &delta) == SPTRSV_OK && delta > 0) {
        sptrsv_config_t jc{};
        float jms = 0;
        int st = jacobi_tune(mat, k, reps, &jc, &jms);
        if (st == SPTRSV_OK && jms < best_t) {
            best_t = jms;
            bc = jc;
        } else if (st != SPTRSV_OK && st != SPTRSV_E_NOT_CONVERGED) {
            return st;
        }
    }
    pv->tuned[precision] = true;
    pv->tuned_cfg[precision] = bc;
    pv->tuned_k[precision] = k;
    if (best) *best = bc;
    if (best_ms) *best_ms = best_t;
    return SPTRSV_OK;
}

}  // extern "C"

namespace sptrsv {
int convert_values_to_f32(sp_mat_t* m) {
    if (m->nnz == 0) return SPTRSV_OK;
    hipLaunchKernelGGL(k_f64_to_f32_vals, dim3(stream_grid((size_t)m->nnz)), dim3(256), 0, nullptr,
                       m->a, m->a32, m->nnz);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}
}  // namespace sptrsv
