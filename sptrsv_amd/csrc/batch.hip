// batch.hip -- many independent lower-triangular systems solved in ONE
// launch (the reference solves one matrix per call, interface.cu:40-95).
//
// The systems are concatenated into one block-diagonal CSR with global
// column indices, so the shared row engine applies unchanged.  Work is a
// table of tasks (matrix m, first row, lanes per row) that workgroups claim
// with the ticket counter; tasks of all matrices are interleaved block by
// block, so every matrix starts its critical path immediately, and each
// matrix runs with its own lanes-per-row choice (heuristic, or timed per
// matrix by sptrsv_batch_tune).  The prepare kernel and the solve are
// captured once in a hipGraph and replayed.
#include <algorithm>
#include <map>
#include <tuple>
#include <vector>

#include "syncfree.hpp"

namespace sptrsv {

hipEvent_t* kernel_timer_pair(hipStream_t s);
int upload_concat(const int* n, const int* const* ia, const int* const* ja, const double* const* a,
                  int nmat, sp_mat_t* m, std::vector<long long>& row_off);

struct Task {
    int start;  // first row (global)
    int end;    // one past the matrix's last row (global)
    int lanes;
    int lgl;
};

// Natural-order rows of one task, LDS shortcut inside the task's rows.
template <class P, int WG>
__global__ __launch_bounds__(WG) void k_batch(SolveArgs A, const Task* __restrict__ tasks, int ntasks) {
    using word = typename P::word;
    __shared__ word s_pub[WG];
    __shared__ int s_t;
    const int tid = threadIdx.x;
    if (tid == 0) s_t = atomicAdd(A.ticket, 1);
    __syncthreads();
    const int t = s_t;
    if (t >= ntasks) return;
    const Task tk = tasks[t];
    const int rwg = WG >> tk.lgl;
    for (int i = tid; i < rwg; i += WG) s_pub[i] = P::kUnsolved;
    __syncthreads();
    const int lane = tid & (kWave - 1);
    const int pos = tk.start + (tid >> tk.lgl);
    const int row = pos < tk.end ? pos : -1;
    const int hi = min(tk.start + rwg, tk.end);
    sf_rows<P, 1, 4>(A, tk.lanes, lane, row, s_pub, tk.start, hi);
}

template <class P>
__global__ void k_batch_prepare(P* pub, size_t count, P unsolved, int* ticket) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *ticket = 0;
    for (; i < count; i += (size_t)gridDim.x * blockDim.x) pub[i] = unsolved;
}

}  // namespace sptrsv

using namespace sptrsv;

struct sptrsv_batch {
    sp_mat_t mat{};
    std::vector<long long> row_off;      // nmat+1
    std::vector<int> lanes;              // per matrix
    std::vector<int> avg;                // nnz/n per matrix
    Task* d_tasks = nullptr;
    int ntasks = 0;
    Task* d_sub = nullptr;               // tuning: tasks of one matrix
    int sub_cap = 0;
    // graph cache: (precision, b, x) -> exec
    std::map<std::tuple<int, const void*, void*>, hipGraphExec_t> graphs;
    hipStream_t cap_stream = nullptr;
};

namespace {

constexpr int kBatchWG = 256;

int lanes_rule(int avg) {
    if (avg <= 2) return 2;
    if (avg <= 5) return 4;
    if (avg <= 11) return 8;
    if (avg <= 23) return 16;
    return 32;
}

int lg(int v) {
    int l = 0;
    while ((1 << (l + 1)) <= v) ++l;
    return l;
}

void matrix_tasks(const sptrsv_batch_t* bt, int m, int lanes, std::vector<Task>& out) {
    const int rwg = kBatchWG / lanes;
    for (long long s = bt->row_off[m]; s < bt->row_off[m + 1]; s += rwg)
        out.push_back({(int)s, (int)bt->row_off[m + 1], lanes, lg(lanes)});
}

// Interleave: block 0 of every matrix, then block 1 of every matrix, ...
int build_tasks(sptrsv_batch_t* bt) {
    const int nmat = (int)bt->lanes.size();
    std::vector<std::vector<Task>> per(nmat);
    size_t total = 0, longest = 0;
    for (int m = 0; m < nmat; ++m) {
        matrix_tasks(bt, m, bt->lanes[m], per[m]);
        total += per[m].size();
        longest = std::max(longest, per[m].size());
    }
    std::vector<Task> all;
    all.reserve(total);
    for (size_t j = 0; j < longest; ++j)
        for (int m = 0; m < nmat; ++m)
            if (j < per[m].size()) all.push_back(per[m][j]);
    if (bt->d_tasks) (void)hipFree(bt->d_tasks);
    bt->d_tasks = nullptr;
    SPT_HIP(hipMalloc(&bt->d_tasks, sizeof(Task) * std::max<size_t>(all.size(), 1)));
    if (!all.empty())
        SPT_HIP(hipMemcpy(bt->d_tasks, all.data(), sizeof(Task) * all.size(), hipMemcpyHostToDevice));
    bt->ntasks = (int)all.size();
    for (auto& g : bt->graphs) (void)hipGraphExecDestroy(g.second);
    bt->graphs.clear();
    return SPTRSV_OK;
}

template <class P>
int launch_batch(sptrsv_batch_t* bt, const Task* tasks, int ntasks, const void* b, void* pub,
                 double* x_out, hipStream_t s) {
    MatPriv* pv = priv_of(&bt->mat);
    SolveArgs A{};
    A.n = bt->mat.nr;
    A.k = 1;
    A.ia = bt->mat.ia;
    A.ja = bt->mat.ja;
    A.a64 = bt->mat.a;
    A.a32 = bt->mat.a32;
    A.b = b;
    A.pub = pub;
    A.x_out = x_out;
    A.ticket = pv->counters;
    A.err = pv->counters + 1;
    A.wd_ticks = pv->watchdog_ticks;
    using word = typename P::word;
    size_t words = (size_t)bt->mat.nr;
    unsigned pg = (unsigned)std::max<size_t>(1, std::min<size_t>((words + 255) / 256, 8192));
    hipLaunchKernelGGL((k_batch_prepare<word>), dim3(pg), dim3(256), 0, s, static_cast<word*>(pub),
                       words, P::kUnsolved, pv->counters);
    SPT_HIP(hipGetLastError());
    if (ntasks > 0) {
        hipEvent_t* ev = kernel_timer_pair(s);
        if (ev) SPT_HIP(hipEventRecord(ev[0], s));
        hipLaunchKernelGGL((k_batch<P, kBatchWG>), dim3(ntasks), dim3(kBatchWG), 0, s, A, tasks, ntasks);
        SPT_HIP(hipGetLastError());
        if (ev) SPT_HIP(hipEventRecord(ev[1], s));
    }
    return SPTRSV_OK;
}

int batch_solve_tasks(sptrsv_batch_t* bt, int precision, const Task* tasks, int ntasks,
                      const double* b, double* x, hipStream_t s) {
    MatPriv* pv = priv_of(&bt->mat);
    if (precision == SPTRSV_PREC_F64) return launch_batch<PolF64>(bt, tasks, ntasks, b, x, nullptr, s);
    if (precision == SPTRSV_PREC_MIXED) {
        if (pv->pub_words < (size_t)bt->mat.nr) return fail(SPTRSV_E_INVALID_ARG, "workspace missing");
        return launch_batch<PolMixed>(bt, tasks, ntasks, b, pv->pub, x, s);
    }
    return fail(SPTRSV_E_INVALID_ARG, "batched solves take FP64 or MIXED precision");
}

}  // namespace

extern "C" {

int sptrsv_batch_create(int nmat, const int* n, const int* const* ia, const int* const* ja,
                        const double* const* a, sptrsv_batch_t** out) {
    if (nmat < 1 || !n || !ia || !ja || !a || !out) return fail(SPTRSV_E_INVALID_ARG, "bad batch arguments");
    auto* bt = new sptrsv_batch();
    int st = upload_concat(n, ia, ja, a, nmat, &bt->mat, bt->row_off);
    if (st != SPTRSV_OK) {
        delete bt;
        return st;
    }
    bt->lanes.resize(nmat);
    bt->avg.resize(nmat);
    for (int m = 0; m < nmat; ++m) {
        long long rows = bt->row_off[m + 1] - bt->row_off[m];
        long long nz = rows > 0 ? (long long)ia[m][n[m]] - ia[m][0] : 0;
        bt->avg[m] = rows > 0 ? (int)(nz / rows) : 0;
        bt->lanes[m] = lanes_rule(bt->avg[m]);
    }
    st = build_tasks(bt);
    if (st == SPTRSV_OK) st = sptrsv_reserve(&bt->mat, 1);
    if (st == SPTRSV_OK) SPT_HIP(hipStreamCreate(&bt->cap_stream));
    if (st != SPTRSV_OK) {
        sptrsv_batch_destroy(bt);
        return st;
    }
    *out = bt;
    return SPTRSV_OK;
}

int sptrsv_batch_rows(const sptrsv_batch_t* bt, int64_t* total_rows, int64_t* total_nnz) {
    if (!bt) return fail(SPTRSV_E_INVALID_ARG, "null batch");
    if (total_rows) *total_rows = bt->mat.nr;
    if (total_nnz) *total_nnz = bt->mat.nnz;
    return SPTRSV_OK;
}

int sptrsv_batch_row_offsets(const sptrsv_batch_t* bt, int64_t* row_off) {
    if (!bt || !row_off) return fail(SPTRSV_E_INVALID_ARG, "null argument");
    for (size_t i = 0; i < bt->row_off.size(); ++i) row_off[i] = bt->row_off[i];
    return SPTRSV_OK;
}

int sptrsv_batch_choice(const sptrsv_batch_t* bt, int m, sptrsv_config_t* cfg) {
    if (!bt || !cfg || m < 0 || m >= (int)bt->lanes.size()) return fail(SPTRSV_E_INVALID_ARG, "bad index");
    *cfg = sptrsv_config_t{SPTRSV_KERNEL_SIMPLE, bt->lanes[m], kBatchWG, 1};
    return SPTRSV_OK;
}

int sptrsv_batch_tune(sptrsv_batch_t* bt, int precision, int tune) {
    if (!bt) return fail(SPTRSV_E_INVALID_ARG, "null batch");
    const int nmat = (int)bt->lanes.size();
    for (int m = 0; m < nmat; ++m) bt->lanes[m] = lanes_rule(bt->avg[m]);
    if (tune) {
        double *b = nullptr, *x = nullptr;
        SPT_HIP(hipMalloc(&b, sizeof(double) * std::max(1, bt->mat.nr)));
        SPT_HIP(hipMalloc(&x, sizeof(double) * std::max(1, bt->mat.nr)));
        SPT_HIP(hipMemset(b, 0, sizeof(double) * std::max(1, bt->mat.nr)));
        hipEvent_t e0, e1;
        SPT_HIP(hipEventCreate(&e0));
        SPT_HIP(hipEventCreate(&e1));
        std::vector<Task> tl;
        int st = SPTRSV_OK;
        for (int m = 0; m < nmat && st == SPTRSV_OK; ++m) {
            int ref = lanes_rule(bt->avg[m]);
            float best = 1e30f;
            for (int L = std::max(1, ref / 2); L <= std::min(64, ref * 2) && st == SPTRSV_OK; L <<= 1) {
                tl.clear();
                matrix_tasks(bt, m, L, tl);
                if ((int)tl.size() > bt->sub_cap) {
                    if (bt->d_sub) (void)hipFree(bt->d_sub);
                    SPT_HIP(hipMalloc(&bt->d_sub, sizeof(Task) * tl.size()));
                    bt->sub_cap = (int)tl.size();
                }
                SPT_HIP(hipMemcpy(bt->d_sub, tl.data(), sizeof(Task) * tl.size(), hipMemcpyHostToDevice));
                st = batch_solve_tasks(bt, precision, bt->d_sub, (int)tl.size(), b, x, nullptr);
                SPT_HIP(hipEventRecord(e0, nullptr));
                for (int r = 0; r < 2 && st == SPTRSV_OK; ++r)
                    st = batch_solve_tasks(bt, precision, bt->d_sub, (int)tl.size(), b, x, nullptr);
                SPT_HIP(hipEventRecord(e1, nullptr));
                SPT_HIP(hipEventSynchronize(e1));
                float ms = 0;
                SPT_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    bt->lanes[m] = L;
                }
            }
        }
        hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(b);
        (void)hipFree(x);
        if (st != SPTRSV_OK) return st;
        SPT_TRY(sptrsv_check(&bt->mat, nullptr));
    }
    return build_tasks(bt);
}

int sptrsv_batch_solve(sptrsv_batch_t* bt, int precision, const double* b, double* x, hipStream_t stream) {
    if (!bt || (bt->mat.nr > 0 && (!b || !x)) || b == x) return fail(SPTRSV_E_INVALID_ARG, "bad batch solve");
    return batch_solve_tasks(bt, precision, bt->d_tasks, bt->ntasks, b, x, stream);
}

int sptrsv_batch_solve_graph(sptrsv_batch_t* bt, int precision, const double* b, double* x,
                             hipStream_t stream) {
    if (!bt || (bt->mat.nr > 0 && (!b || !x)) || b == x) return fail(SPTRSV_E_INVALID_ARG, "bad batch solve");
    auto key = std::make_tuple(precision, (const void*)b, (void*)x);
    auto it = bt->graphs.find(key);
    if (it == bt->graphs.end()) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        SPT_HIP(hipStreamBeginCapture(bt->cap_stream, hipStreamCaptureModeThreadLocal));
        int st = batch_solve_tasks(bt, precision, bt->d_tasks, bt->ntasks, b, x, bt->cap_stream);
        hipError_t e = hipStreamEndCapture(bt->cap_stream, &g);
        if (st != SPTRSV_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return hip_fail(e, "end capture", __FILE__, __LINE__);
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return hip_fail(e, "graph instantiate", __FILE__, __LINE__);
        it = bt->graphs.emplace(key, ge).first;
    }
    SPT_HIP(hipGraphLaunch(it->second, stream));
    return SPTRSV_OK;
}

void sptrsv_batch_destroy(sptrsv_batch_t* bt) {
    if (!bt) return;
    for (auto& g : bt->graphs) (void)hipGraphExecDestroy(g.second);
    if (bt->d_tasks) (void)hipFree(bt->d_tasks);
    if (bt->d_sub) (void)hipFree(bt->d_sub);
    if (bt->cap_stream) (void)hipStreamDestroy(bt->cap_stream);
    free_device_analysis(&bt->mat, nullptr);
    delete bt;
}

}  // extern "C"
