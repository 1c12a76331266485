// batch.hip -- many independent lower-triangular systems solved in ONE
// launch (the reference solves one matrix per call, interface.cu:40-95).
//
// The systems are concatenated into one block-diagonal CSR with global
// column indices, so the shared row engine applies unchanged.  Work is a
// table of tasks (matrix m, first row, lanes per row) that workgroups claim
// with the ticket counter; tasks of all matrices are interleaved block by
// block, so every matrix starts its critical path immediately, and each
// matrix runs with its own lanes-per-row choice (heuristic, or timed per
// matrix by sptrsv_batch_tune and then capped for the whole launch, which is
// bound by throughput rather than by any one member's chain).  The prepare kernel and the solve are
// captured once in a hipGraph and replayed.
//
// Mixed precision adds a second per-matrix choice: a strictly diagonally
// dominant member may instead be solved by certified Jacobi refinement
// (iterative.hip) over tiles cut inside that member, which turns a deep
// dependency chain (a banded member has one level per row) into a few tens
// of bandwidth-bound sweeps.  The Jacobi members run on a second stream
// forked from the solve stream, concurrently with the sync-free launch of
// the other members, and both halves are captured into the same graph.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <unordered_set>
#include <vector>

#include "syncfree.hpp"

namespace sptrsv {

int upload_concat(const int* n, const int* const* ia, const int* const* ja, const double* const* a,
                  int nmat, sp_mat_t* m, std::vector<long long>& row_off);
int jacobi_solve(sp_mat_t* m, const sptrsv_config_t& c, int k, const double* b, double* x, hipStream_t s);
int jacobi_tune(sp_mat_t* m, int k, int reps, sptrsv_config_t* best, float* best_ms);
int reserve_jacobi(sp_mat_t* m, int k);
int solve_checked(sp_mat_t* m, dfr_analysis_info_t* info, int precision, const sptrsv_config_t* cfg,
                  const void* b, void* x, int n, int k, int refine_steps, hipStream_t s);
int convert_values_to_f32(sp_mat_t* m);
int reserve_part(sp_mat_t* m, const dfr_analysis_info_t* info);
unsigned long long fnv1a(const void* p, size_t n, unsigned long long h);
void jacobi_host_tiles(const int* ia, int r_begin, int r_end, long long row_base, long long entry_base,
                       int window, std::vector<int>& out);
int jacobi_set_tiles(sp_mat_t* m, const std::vector<int>* tiles4, const std::vector<int>* tile_member,
                     int members, double delta);

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

__global__ void k_rows_iota(int* rows, long long count, int first) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (long long)gridDim.x * blockDim.x)
        rows[i] = first + (int)i;
}

__global__ void k_gather_rows(const double* __restrict__ src, const int* __restrict__ rows, int n,
                              double* __restrict__ dst) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[rows[i]];
}

__global__ void k_scatter_rows(const double* __restrict__ src, const int* __restrict__ rows, int n,
                               double* __restrict__ dst) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        dst[rows[i]] = src[i];
}

__global__ void k_shift(int* p, long long count, int off) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (long long)gridDim.x * blockDim.x)
        p[i] += off;
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
    std::vector<double> delta;           // per matrix: min_i |L_ii| - sum_j |L_ij|
    std::vector<std::vector<int>> mtiles[2];  // per matrix: Jacobi row tiles of 1024 / 2048 entries
    std::vector<int> choice;             // MIXED per matrix: SPTRSV_KERNEL_SIMPLE or _JACOBI
    std::vector<float> t_sync;           // MIXED per matrix: tuned sync-free time alone (ms)
    sptrsv_config_t jac_cfg{};           // schedule of the Jacobi members
    int n_jacobi = 0;
    Task* d_tasks = nullptr;             // every matrix (FP64)
    int ntasks = 0;
    Task* d_tasks_sf = nullptr;          // MIXED: the sync-free matrices only
    int ntasks_sf = 0;
    hipStream_t aux = nullptr;           // Jacobi half of a MIXED solve
    bool serial = false;                 // MIXED: Jacobi half, then the sync-free half, on one stream
    std::vector<long long> ent_off;      // nmat+1: first entry of each matrix
    std::vector<int> max_row;            // per matrix: longest row
    std::vector<double> chain;           // per matrix: fraction of rows i > 0 that depend on row i - 1
    std::vector<int> pattern;            // per matrix: first matrix with the same structure (n, ia, ja)
    int n_patterns = 0;
    // MIXED: the sync-free matrices as one matrix of their own, with its own
    // level analysis and tuned configuration; its rows are gathered from and
    // scattered back to the batch order (null: the ticketed task launch)
    sp_mat_t sub{};
    dfr_analysis_info_t* sub_info = nullptr;
    sptrsv_config_t sub_cfg{};
    int* sub_rows = nullptr;             // batch row of each row of sub
    double* sub_b = nullptr;
    double* sub_x = nullptr;
    bool use_sub = false;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    Task* d_sub = nullptr;               // tuning: tasks of one matrix
    int sub_cap = 0;
    // graph cache: (precision, b, x) -> exec
    std::map<std::tuple<int, const void*, void*>, hipGraphExec_t> graphs;
    hipStream_t cap_stream = nullptr;
};

namespace {

constexpr int kBatchWG = 256;

// Batches alive in this process, for the library's teardown.
std::mutex g_batch_mu;
std::unordered_set<sptrsv_batch*> g_batches;

// Library teardown, before the HIP runtime's own: captured graphs, streams
// and events of batches their owners never destroyed (device memory is left
// to the runtime); no HIP call when there are none.
__attribute__((destructor)) void batch_teardown() {
    std::lock_guard<std::mutex> g(g_batch_mu);
    if (!g_batches.empty()) (void)hipDeviceSynchronize();
    for (sptrsv_batch* bt : g_batches) {
        for (auto& gr : bt->graphs) (void)hipGraphExecDestroy(gr.second);
        bt->graphs.clear();
        for (hipStream_t* st : {&bt->cap_stream, &bt->aux}) {
            if (*st) (void)hipStreamDestroy(*st);
            *st = nullptr;
        }
        for (hipEvent_t* ev : {&bt->fork_ev, &bt->join_ev}) {
            if (*ev) (void)hipEventDestroy(*ev);
            *ev = nullptr;
        }
    }
}

// Destroys the batch's captured solves: they hold the device addresses of
// the task tables, the separate matrix and the Jacobi workspaces, so this
// runs before any of them is freed or rebuilt.
void drop_batch_graphs(sptrsv_batch_t* bt) {
    if (bt->graphs.empty()) return;
    (void)hipDeviceSynchronize();
    for (auto& g : bt->graphs) (void)hipGraphExecDestroy(g.second);
    bt->graphs.clear();
}

// Members with the same structure (n, row lengths, columns): pattern[m] is
// the first such member.  A structure hash finds candidates; equal arrays
// confirm them.
void find_patterns(sptrsv_batch_t* bt, const int* n, const int* const* ia, const int* const* ja) {
    const int nmat = (int)bt->lanes.size();
    bt->pattern.assign(nmat, 0);
    bt->n_patterns = 0;
    std::map<unsigned long long, std::vector<int>> by_hash;
    auto same = [&](int p, int m) {
        if (n[p] != n[m]) return false;
        const int bp = n[p] > 0 ? ia[p][0] : 0, bm = n[m] > 0 ? ia[m][0] : 0;
        for (int i = 0; i <= n[m]; ++i)
            if (ia[p][i] - bp != ia[m][i] - bm) return false;
        const int nz = n[m] > 0 ? ia[m][n[m]] - bm : 0;
        return nz == 0 || std::memcmp(ja[p] + bp, ja[m] + bm, sizeof(int) * (size_t)nz) == 0;
    };
    for (int m = 0; m < nmat; ++m) {
        unsigned long long h = fnv1a(&n[m], sizeof(int), 1469598103934665603ull);
        const int base = n[m] > 0 ? ia[m][0] : 0;
        for (int i = 1; i <= n[m]; ++i) {
            const int len = ia[m][i] - ia[m][i - 1];
            h = fnv1a(&len, sizeof(int), h);
        }
        if (n[m] > 0) h = fnv1a(ja[m] + base, sizeof(int) * (size_t)(ia[m][n[m]] - base), h);
        std::vector<int>& cand = by_hash[h];
        int rep = m;
        for (int p : cand)
            if (same(p, m)) {
                rep = p;
                break;
            }
        if (rep == m) {
            cand.push_back(m);
            ++bt->n_patterns;
        }
        bt->pattern[m] = rep;
    }
}

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
int upload_tasks(const std::vector<std::vector<Task>>& per, Task** d, int* count) {
    size_t total = 0, longest = 0;
    for (auto& v : per) {
        total += v.size();
        longest = std::max(longest, v.size());
    }
    std::vector<Task> all;
    all.reserve(total);
    for (size_t j = 0; j < longest; ++j)
        for (auto& v : per)
            if (j < v.size()) all.push_back(v[j]);
    if (*d) (void)hipFree(*d);
    *d = nullptr;
    SPT_HIP(hipMalloc(d, sizeof(Task) * std::max<size_t>(all.size(), 1)));
    if (!all.empty()) SPT_HIP(hipMemcpy(*d, all.data(), sizeof(Task) * all.size(), hipMemcpyHostToDevice));
    *count = (int)all.size();
    return SPTRSV_OK;
}

int build_tasks(sptrsv_batch_t* bt) {
    const int nmat = (int)bt->lanes.size();
    std::vector<std::vector<Task>> per(nmat), per_sf(nmat);
    for (int m = 0; m < nmat; ++m) {
        matrix_tasks(bt, m, bt->lanes[m], per[m]);
        if (bt->choice[m] != SPTRSV_KERNEL_JACOBI) per_sf[m] = per[m];
    }
    drop_batch_graphs(bt);
    SPT_TRY(upload_tasks(per, &bt->d_tasks, &bt->ntasks));
    SPT_TRY(upload_tasks(per_sf, &bt->d_tasks_sf, &bt->ntasks_sf));
    return SPTRSV_OK;
}

// Hands the tiles of the matrices whose choice is JACOBI to the Jacobi path.
int install_jacobi(sptrsv_batch_t* bt) {
    drop_batch_graphs(bt);
    const int nmat = (int)bt->choice.size();
    std::vector<int> tiles[2], member[2];
    double delta = 1e300;
    bt->n_jacobi = 0;
    for (int m = 0; m < nmat; ++m) {
        if (bt->choice[m] != SPTRSV_KERNEL_JACOBI) continue;
        ++bt->n_jacobi;
        delta = std::min(delta, bt->delta[m]);
        for (int w = 0; w < 2; ++w) {
            tiles[w].insert(tiles[w].end(), bt->mtiles[w][m].begin(), bt->mtiles[w][m].end());
            member[w].insert(member[w].end(), bt->mtiles[w][m].size() / 4, m);
        }
    }
    SPT_TRY(jacobi_set_tiles(&bt->mat, tiles, member, nmat, bt->n_jacobi ? delta : 1.0));
    if (bt->n_jacobi) SPT_TRY(reserve_jacobi(&bt->mat, 1));
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
        const KtPair ev = kernel_timer_pair(s);
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

// The sync-free matrices of a MIXED solve: the separate matrix (gather b,
// solve with its tuned configuration, scatter x) or the ticketed task launch.
int sync_free_half(sptrsv_batch_t* bt, int precision, const double* b, double* x, hipStream_t s) {
    if (bt->use_sub && bt->sub.nr > 0) {
        const int n = bt->sub.nr;
        const unsigned g = (unsigned)std::min<long long>(8192, ((long long)n + 255) / 256);
        hipLaunchKernelGGL(k_gather_rows, dim3(g), dim3(256), 0, s, b, bt->sub_rows, n, bt->sub_b);
        SPT_HIP(hipGetLastError());
        SPT_TRY(solve_checked(&bt->sub, bt->sub_info, precision, &bt->sub_cfg, bt->sub_b, bt->sub_x, n, 1, 1, s));
        hipLaunchKernelGGL(k_scatter_rows, dim3(g), dim3(256), 0, s, bt->sub_x, bt->sub_rows, n, x);
        SPT_HIP(hipGetLastError());
        return SPTRSV_OK;
    }
    if (bt->ntasks_sf > 0) return batch_solve_tasks(bt, precision, bt->d_tasks_sf, bt->ntasks_sf, b, x, s);
    return SPTRSV_OK;
}

// The whole batched solve: FP64 runs every matrix sync-free; MIXED runs the
// Jacobi matrices on the auxiliary stream while the sync-free ones run on s.
int batch_solve_all(sptrsv_batch_t* bt, int precision, const double* b, double* x, hipStream_t s) {
    if (precision != SPTRSV_PREC_MIXED) return batch_solve_tasks(bt, precision, bt->d_tasks, bt->ntasks, b, x, s);
    if (bt->n_jacobi == 0) return sync_free_half(bt, precision, b, x, s);
    if (bt->serial) {  // the two halves one after the other, each with the whole device
        SPT_TRY(jacobi_solve(&bt->mat, bt->jac_cfg, 1, b, x, s));
        return sync_free_half(bt, precision, b, x, s);
    }
    SPT_HIP(hipEventRecord(bt->fork_ev, s));
    SPT_HIP(hipStreamWaitEvent(bt->aux, bt->fork_ev, 0));
    SPT_TRY(jacobi_solve(&bt->mat, bt->jac_cfg, 1, b, x, bt->aux));
    SPT_HIP(hipEventRecord(bt->join_ev, bt->aux));
    SPT_TRY(sync_free_half(bt, precision, b, x, s));
    SPT_HIP(hipStreamWaitEvent(s, bt->join_ev, 0));
    return SPTRSV_OK;
}

void free_sub(sptrsv_batch_t* bt) {
    drop_batch_graphs(bt);
    if (bt->sub.priv) free_device_analysis(&bt->sub, bt->sub_info);
    if (bt->sub_info) free(bt->sub_info);
    bt->sub_info = nullptr;
    for (void* p : {(void*)bt->sub_rows, (void*)bt->sub_b, (void*)bt->sub_x})
        if (p) (void)hipFree(p);
    bt->sub_rows = nullptr;
    bt->sub_b = bt->sub_x = nullptr;
    memset(&bt->sub, 0, sizeof(bt->sub));
    bt->use_sub = false;
}

// The sync-free matrices (choice other than JACOBI) copied on the device
// into one matrix of their own, in batch order, with its level analysis.
int build_sub(sptrsv_batch_t* bt) {
    free_sub(bt);
    const int nmat = (int)bt->choice.size();
    long long rows = 0, nz = 0;
    int mr = 0;
    for (int m = 0; m < nmat; ++m) {
        if (bt->choice[m] == SPTRSV_KERNEL_JACOBI) continue;
        rows += bt->row_off[m + 1] - bt->row_off[m];
        nz += bt->ent_off[m + 1] - bt->ent_off[m];
        mr = std::max(mr, bt->max_row[m]);
    }
    if (rows == 0) return SPTRSV_OK;
    sp_mat_t& S = bt->sub;
    SPT_TRY(ensure_priv(&S));
    S.nr = S.nc = (int)rows;
    S.nnz = (int)nz;
    priv_of(&S)->max_row = mr;
    int st = SPTRSV_OK;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && st == SPTRSV_OK) st = hip_fail(e, what, __FILE__, __LINE__);
    };
    chk(hipMalloc(&S.ia, sizeof(int) * (size_t)(rows + 1)), "hipMalloc ia");
    chk(hipMalloc(&S.ja, sizeof(int) * (size_t)nz), "hipMalloc ja");
    chk(hipMalloc(&S.a, sizeof(double) * (size_t)nz), "hipMalloc a");
    chk(hipMalloc(&S.a32, sizeof(float) * (size_t)nz), "hipMalloc a32");
    chk(hipMalloc(&bt->sub_rows, sizeof(int) * (size_t)rows), "hipMalloc row map");
    chk(hipMalloc(&bt->sub_b, sizeof(double) * (size_t)rows), "hipMalloc b");
    chk(hipMalloc(&bt->sub_x, sizeof(double) * (size_t)rows), "hipMalloc x");
    long long r = 0, e = 0;
    for (int m = 0; m < nmat && st == SPTRSV_OK; ++m) {
        if (bt->choice[m] == SPTRSV_KERNEL_JACOBI) continue;
        const long long r0 = bt->row_off[m], nr = bt->row_off[m + 1] - r0;
        const long long e0 = bt->ent_off[m], ne = bt->ent_off[m + 1] - e0;
        if (nr == 0) continue;
        const unsigned gr = (unsigned)std::min<long long>(4096, (nr + 255) / 256);
        const unsigned ge = (unsigned)std::min<long long>(4096, (ne + 255) / 256);
        chk(hipMemcpy(S.ia + r, bt->mat.ia + r0, sizeof(int) * (size_t)nr, hipMemcpyDeviceToDevice), "copy ia");
        chk(hipMemcpy(S.ja + e, bt->mat.ja + e0, sizeof(int) * (size_t)ne, hipMemcpyDeviceToDevice), "copy ja");
        chk(hipMemcpy(S.a + e, bt->mat.a + e0, sizeof(double) * (size_t)ne, hipMemcpyDeviceToDevice), "copy a");
        if (st != SPTRSV_OK) break;
        hipLaunchKernelGGL(k_shift, dim3(gr), dim3(256), 0, nullptr, S.ia + r, nr, (int)(e - e0));
        hipLaunchKernelGGL(k_shift, dim3(ge), dim3(256), 0, nullptr, S.ja + e, ne, (int)(r - r0));
        hipLaunchKernelGGL(k_rows_iota, dim3(gr), dim3(256), 0, nullptr, bt->sub_rows + r, nr, (int)r0);
        chk(hipGetLastError(), "row block kernels");
        r += nr;
        e += ne;
    }
    const int last = (int)nz;
    if (st == SPTRSV_OK) chk(hipMemcpy(S.ia + rows, &last, sizeof(int), hipMemcpyHostToDevice), "ia end");
    if (st == SPTRSV_OK) st = convert_values_to_f32(&S);
    if (st == SPTRSV_OK) chk(hipDeviceSynchronize(), "sub build");
    if (st == SPTRSV_OK) st = multirow_analysis_base_GPU(&bt->sub_info, &S, ALL);
    if (st == SPTRSV_OK) st = reserve_part(&S, bt->sub_info);  // captured solves allocate nothing
    if (st != SPTRSV_OK) free_sub(bt);
    return st;
}

// Mean time of `reps` whole batched MIXED solves (after one warm-up).
float time_batch(sptrsv_batch_t* bt, const double* b, double* x, int reps, int* st,
                 int precision = SPTRSV_PREC_MIXED) {
    hipEvent_t e0, e1;
    *st = SPTRSV_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        *st = fail(SPTRSV_E_HIP, "event create");
        return 1e30f;
    }
    *st = batch_solve_all(bt, precision, b, x, nullptr);
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps && *st == SPTRSV_OK; ++r) *st = batch_solve_all(bt, precision, b, x, nullptr);
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / std::max(reps, 1);
}

}  // namespace

extern "C" {

int sptrsv_batch_create(int nmat, const int* n, const int* const* ia, const int* const* ja,
                        const double* const* a, sptrsv_batch_t** out) {
    if (nmat < 1 || !n || !ia || !ja || !a || !out) return fail(SPTRSV_E_INVALID_ARG, "bad batch arguments");
    auto* bt = new sptrsv_batch();
    {
        std::lock_guard<std::mutex> g(g_batch_mu);
        g_batches.insert(bt);
    }
    int st = upload_concat(n, ia, ja, a, nmat, &bt->mat, bt->row_off);
    if (st != SPTRSV_OK) {
        std::lock_guard<std::mutex> g(g_batch_mu);
        g_batches.erase(bt);
        delete bt;
        return st;
    }
    bt->lanes.resize(nmat);
    bt->avg.resize(nmat);
    bt->delta.resize(nmat);
    bt->mtiles[0].resize(nmat);
    bt->mtiles[1].resize(nmat);
    bt->choice.assign(nmat, SPTRSV_KERNEL_SIMPLE);
    bt->t_sync.assign(nmat, 0.0f);
    bt->ent_off.assign(nmat + 1, 0);
    bt->max_row.assign(nmat, 0);
    bt->chain.assign(nmat, 0.0);
    long long entry_base = 0;
    for (int m = 0; m < nmat; ++m) {
        long long rows = bt->row_off[m + 1] - bt->row_off[m];
        long long nz = rows > 0 ? (long long)ia[m][n[m]] - ia[m][0] : 0;
        bt->ent_off[m] = entry_base;
        for (int i = 0; i < n[m]; ++i) bt->max_row[m] = std::max(bt->max_row[m], ia[m][i + 1] - ia[m][i]);
        bt->avg[m] = rows > 0 ? (int)(nz / rows) : 0;
        bt->lanes[m] = lanes_rule(bt->avg[m]);
        double dm = rows > 0 ? 1e300 : 1.0;
        for (int i = 0; i < n[m]; ++i) {
            const int rs = ia[m][i] - ia[m][0], re = ia[m][i + 1] - ia[m][0];
            double off = 0.0;
            for (int p = rs; p < re - 1; ++p) off += std::fabs(a[m][ia[m][0] + p]);
            dm = std::min(dm, std::fabs(a[m][ia[m][0] + re - 1]) - off);
        }
        bt->delta[m] = dm;
        long long pred = 0;  // rows whose largest off-diagonal column is i - 1 (diagonal last)
        for (int i = 1; i < n[m]; ++i) {
            const int re = ia[m][i + 1] - ia[m][0];
            if (re - (ia[m][i] - ia[m][0]) >= 2 && ja[m][ia[m][0] + re - 2] == i - 1) ++pred;
        }
        bt->chain[m] = n[m] > 1 ? (double)pred / (n[m] - 1) : 0.0;
        if (rows > 0)
            for (int w = 0; w < 2; ++w)
                jacobi_host_tiles(ia[m], 0, n[m], bt->row_off[m], entry_base, 1024 << w, bt->mtiles[w][m]);
        entry_base += nz;
    }
    bt->ent_off[nmat] = entry_base;
    find_patterns(bt, n, ia, ja);
    st = build_tasks(bt);
    if (st == SPTRSV_OK) st = sptrsv_reserve(&bt->mat, 1);
    if (st == SPTRSV_OK) st = install_jacobi(bt);
    if (st == SPTRSV_OK) SPT_HIP(hipStreamCreate(&bt->cap_stream));
    if (st == SPTRSV_OK) SPT_HIP(hipStreamCreate(&bt->aux));
    if (st == SPTRSV_OK) SPT_HIP(hipEventCreateWithFlags(&bt->fork_ev, hipEventDisableTiming));
    if (st == SPTRSV_OK) SPT_HIP(hipEventCreateWithFlags(&bt->join_ev, hipEventDisableTiming));
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
    if (bt->choice[m] == SPTRSV_KERNEL_JACOBI) *cfg = bt->jac_cfg;
    return SPTRSV_OK;
}

int sptrsv_batch_patterns(const sptrsv_batch_t* bt, int* n_patterns, int* pattern_of) {
    if (!bt || !n_patterns) return fail(SPTRSV_E_INVALID_ARG, "null batch or count");
    *n_patterns = bt->n_patterns;
    if (pattern_of) std::copy(bt->pattern.begin(), bt->pattern.end(), pattern_of);
    return SPTRSV_OK;
}

int sptrsv_batch_schedule(const sptrsv_batch_t* bt, int* serial, int* separate) {
    if (!bt || !serial || !separate) return fail(SPTRSV_E_INVALID_ARG, "null argument");
    *serial = bt->serial ? 1 : 0;
    *separate = bt->use_sub ? 1 : 0;
    return SPTRSV_OK;
}

int sptrsv_batch_set_schedule(sptrsv_batch_t* bt, int serial, int separate) {
    if (!bt || serial < 0 || serial > 1 || separate < 0 || separate > 1)
        return fail(SPTRSV_E_INVALID_ARG, "bad batch schedule");
    if (separate && !bt->sub.priv)
        return fail(SPTRSV_E_INVALID_ARG, "no separate sync-free matrix: tune the batch (MIXED) first");
    drop_batch_graphs(bt);
    bt->serial = serial != 0;
    bt->use_sub = separate != 0;
    return SPTRSV_OK;
}

}  // extern "C"

namespace {

// Device scratch and timing events of one tuning call, released on every
// return path.
struct TuneScratch {
    double* b = nullptr;
    double* x = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~TuneScratch() {
        if (b) (void)hipFree(b);
        if (x) (void)hipFree(x);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    int init(int rows) {
        const size_t bytes = sizeof(double) * (size_t)std::max(1, rows);
        SPT_HIP(hipMalloc(&b, bytes));
        SPT_HIP(hipMalloc(&x, bytes));
        SPT_HIP(hipMemset(b, 0, bytes));
        SPT_HIP(hipEventCreate(&e0));
        SPT_HIP(hipEventCreate(&e1));
        return SPTRSV_OK;
    }
};

// Mean time of `reps` sync-free solves of matrix m alone with `lanes`
// lanes per row (after one untimed solve).
int time_member(sptrsv_batch_t* bt, TuneScratch& ts, int precision, int m, int lanes, int reps, float* ms) {
    std::vector<Task> tl;
    matrix_tasks(bt, m, lanes, tl);
    if ((int)tl.size() > bt->sub_cap) {
        if (bt->d_sub) (void)hipFree(bt->d_sub);
        bt->d_sub = nullptr;
        bt->sub_cap = 0;
        SPT_HIP(hipMalloc(&bt->d_sub, sizeof(Task) * tl.size()));
        bt->sub_cap = (int)tl.size();
    }
    SPT_HIP(hipMemcpy(bt->d_sub, tl.data(), sizeof(Task) * tl.size(), hipMemcpyHostToDevice));
    SPT_TRY(batch_solve_tasks(bt, precision, bt->d_sub, (int)tl.size(), ts.b, ts.x, nullptr));
    SPT_HIP(hipEventRecord(ts.e0, nullptr));
    for (int r = 0; r < reps; ++r)
        SPT_TRY(batch_solve_tasks(bt, precision, bt->d_sub, (int)tl.size(), ts.b, ts.x, nullptr));
    SPT_HIP(hipEventRecord(ts.e1, nullptr));
    SPT_HIP(hipEventSynchronize(ts.e1));
    SPT_HIP(hipEventElapsedTime(ms, ts.e0, ts.e1));
    *ms /= std::max(reps, 1);
    return SPTRSV_OK;
}

// Sets the matrices solved by Jacobi (the rest sync-free) and tunes the
// Jacobi schedule for that set; *jac_ms gets the Jacobi time alone.
int apply_choice(sptrsv_batch_t* bt, const std::vector<int>& jac, float* jac_ms) {
    const int nmat = (int)bt->choice.size();
    bt->choice.assign(nmat, SPTRSV_KERNEL_SIMPLE);
    for (int m : jac) bt->choice[m] = SPTRSV_KERNEL_JACOBI;
    SPT_TRY(install_jacobi(bt));
    if (!jac.empty()) {
        float jms = 0;
        SPT_TRY(jacobi_tune(&bt->mat, 1, 2, &bt->jac_cfg, &jms));
        if (jac_ms) *jac_ms = jms;
    }
    return build_tasks(bt);
}

// Lanes per row of the sync-free members launched together.  Each member's
// lanes were timed with the member alone, where its dependency chain bounds
// it; launched together the members are bound by throughput, where fewer
// lanes per row (more rows per wave, every lane busy) can win: the 128
// random members of the 256-system benchmark (about 7 entries per row) take
// 2.9 ms at their own 8 lanes and 1.7 ms at 1 or 2.  The whole batch is
// timed with the members' own lanes and with every sync-free member capped
// at 2 and at 1 lane per row, and the fastest is kept.
int tune_concurrent_lanes(sptrsv_batch_t* bt, TuneScratch& ts, int precision) {
    const std::vector<int> own = bt->lanes;
    bt->use_sub = false;  // the task launch is what the lanes shape
    SPT_TRY(build_tasks(bt));
    int st = SPTRSV_OK;
    float best = time_batch(bt, ts.b, ts.x, 3, &st, precision);
    SPT_TRY(st);
    std::vector<int> best_lanes = own;
    for (int cap : {2, 1}) {
        bool changed = false;
        for (size_t m = 0; m < own.size(); ++m) {
            const bool jacobi = precision == SPTRSV_PREC_MIXED && bt->choice[m] == SPTRSV_KERNEL_JACOBI;
            bt->lanes[m] = jacobi ? own[m] : std::min(own[m], cap);
            changed |= bt->lanes[m] != own[m];
        }
        if (!changed) continue;
        SPT_TRY(build_tasks(bt));
        const float t = time_batch(bt, ts.b, ts.x, 3, &st, precision);
        SPT_TRY(st);
        if (t < best) {
            best = t;
            best_lanes = bt->lanes;
        }
    }
    bt->lanes = best_lanes;
    return build_tasks(bt);
}

// MIXED per-matrix choice between the sync-free launch and Jacobi.  The
// Jacobi time J of all diagonally dominant matrices is measured first; a
// dominant matrix whose sync-free time alone exceeds J / 2 is "slow".  Then
// whole concurrent batches are timed -- every matrix sync-free, Jacobi for
// every dominant matrix, for the slow ones, and for those whose sync-free
// time alone exceeds 1 ms or 0.25 ms -- and the fastest is kept.  Lanes per
// row are tuned only for matrices that can stay sync-free.
int tune_mixed(sptrsv_batch_t* bt, TuneScratch& ts) {
    const int nmat = (int)bt->choice.size();
    std::vector<int> dominant;
    for (int m = 0; m < nmat; ++m)
        if (bt->delta[m] > 0) dominant.push_back(m);
    float jall = -1.0f;
    if (!dominant.empty()) {
        const int st = apply_choice(bt, dominant, &jall);
        if (st == SPTRSV_E_NOT_CONVERGED) jall = -1.0f;
        else SPT_TRY(st);
    }
    std::vector<int> slow;
    std::vector<char> lanes_tuned(nmat, 0);
    for (int m = 0; m < nmat; ++m) {
        // a member with an earlier member's structure takes that member's
        // sync-free timing and lanes (the work, hence the time, depends on
        // the structure only)
        const int p = bt->pattern[m];
        if (p != m && lanes_tuned[p]) {
            bt->lanes[m] = bt->lanes[p];
            bt->t_sync[m] = bt->t_sync[p];
            lanes_tuned[m] = 1;
            if (jall > 0 && bt->delta[m] > 0 && bt->t_sync[m] > 0.5f * jall) slow.push_back(m);
            continue;
        }
        const int ref = lanes_rule(bt->avg[m]);
        float best = 0;
        SPT_TRY(time_member(bt, ts, SPTRSV_PREC_MIXED, m, ref, 1, &best));
        bt->lanes[m] = ref;
        if (jall > 0 && bt->delta[m] > 0 && best > 0.5f * jall) {
            slow.push_back(m);  // Jacobi candidate: lanes do not move a latency-bound chain
        } else {
            for (int L : {std::max(1, ref / 2), std::min(64, ref * 2)}) {
                if (L == ref) continue;
                float ms = 0;
                SPT_TRY(time_member(bt, ts, SPTRSV_PREC_MIXED, m, L, 2, &ms));
                if (ms < best) {
                    best = ms;
                    bt->lanes[m] = L;
                }
            }
            lanes_tuned[m] = 1;
        }
        bt->t_sync[m] = best;
    }
    SPT_TRY(sptrsv_check(&bt->mat, nullptr));
    int st = SPTRSV_OK;
    SPT_TRY(apply_choice(bt, {}, nullptr));
    float best = time_batch(bt, ts.b, ts.x, 3, &st);
    SPT_TRY(st);
    std::vector<int> best_set;
    if (jall > 0) {
        // candidate Jacobi sets: every dominant matrix, and the dominant
        // matrices whose sync-free time alone exceeds a threshold (a long
        // dependency chain bounds the sync-free half however few rows it has)
        std::vector<std::vector<int>> sets{dominant, slow};
        for (float tau : {1.0f, 0.25f}) {
            std::vector<int> set;
            for (int m : dominant)
                if (bt->t_sync[m] > tau) set.push_back(m);
            sets.push_back(set);
        }
        std::sort(sets.begin(), sets.end());
        sets.erase(std::unique(sets.begin(), sets.end()), sets.end());
        for (const std::vector<int>& set : sets) {
            if (set.empty()) continue;
            SPT_TRY(apply_choice(bt, set, nullptr));
            const float t = time_batch(bt, ts.b, ts.x, 3, &st);
            SPT_TRY(st);
            if (t < best) {
                best = t;
                best_set = set;
            }
        }
    }
    SPT_TRY(apply_choice(bt, best_set, nullptr));
    SPT_TRY(tune_concurrent_lanes(bt, ts, SPTRSV_PREC_MIXED));
    // The sync-free matrices either through the ticketed task launch or as a
    // matrix of their own (level analysis, tuned kernel); the two halves
    // concurrently on two streams or one after the other (the concurrent
    // sweeps can starve latency-bound waves).  Persistent kernels claim their
    // work in order, so they progress next to the Jacobi stream too.
    bt->serial = false;
    bt->use_sub = false;
    if (bt->ntasks_sf > 0) {
        SPT_TRY(build_sub(bt));
        bool sub_ok = bt->sub.nr > 0;
        if (sub_ok) {
            float ms = 0;
            const int ast = sptrsv_autotune(&bt->sub, bt->sub_info, SPTRSV_PREC_MIXED, 1, 3, &bt->sub_cfg, &ms);
            if (ast != SPTRSV_OK) {
                free_sub(bt);
                sub_ok = false;
            }
        }
        float best_t = 1e30f;
        bool best_sub = false, best_serial = false;
        for (int use : {0, 1}) {
            if (use && !sub_ok) continue;
            for (int ser : {0, 1}) {
                if (bt->n_jacobi == 0 && ser) continue;  // nothing to overlap
                bt->use_sub = use != 0;
                bt->serial = ser != 0;
                const float t = time_batch(bt, ts.b, ts.x, 5, &st);
                SPT_TRY(st);
                if (t < best_t) {
                    best_t = t;
                    best_sub = bt->use_sub;
                    best_serial = bt->serial;
                }
            }
        }
        bt->use_sub = best_sub;
        bt->serial = best_serial;
    }
    drop_batch_graphs(bt);
    return sptrsv_check(&bt->mat, nullptr);
}

}  // namespace

extern "C" {

int sptrsv_batch_tune(sptrsv_batch_t* bt, int precision, int tune) {
    if (!bt) return fail(SPTRSV_E_INVALID_ARG, "null batch");
    if (precision != SPTRSV_PREC_F64 && precision != SPTRSV_PREC_MIXED)
        return fail(SPTRSV_E_INVALID_ARG, "batched solves take FP64 or MIXED precision");
    // tuning rebuilds task tables, Jacobi tiles and the separate matrix, and
    // may return early on any path: captured solves go first
    drop_batch_graphs(bt);
    const int nmat = (int)bt->lanes.size();
    for (int m = 0; m < nmat; ++m) bt->lanes[m] = lanes_rule(bt->avg[m]);
    if (!tune) {
        // Untuned choice: in MIXED precision a strictly dominant matrix whose
        // rows mostly depend on their predecessor (a dependency chain about as
        // long as the matrix, as in banded systems) takes block Jacobi sweeps
        // with a fixed schedule (5 sweeps x 3 refinement steps, 4 in-tile
        // iterations, 1024-product tiles); everything else stays sync-free.
        // The halves run one after the other, the sync-free one as tasks.
        bt->choice.assign(nmat, SPTRSV_KERNEL_SIMPLE);
        if (precision == SPTRSV_PREC_MIXED)
            for (int m = 0; m < nmat; ++m)
                if (bt->delta[m] > 0 && bt->chain[m] >= 0.5) bt->choice[m] = SPTRSV_KERNEL_JACOBI;
        sptrsv_config_t jc{};
        jc.kernel = SPTRSV_KERNEL_JACOBI;
        jc.lanes = 1;
        jc.wg = 256;
        jc.fused = 1;
        jc.variant = 2;
        jc.sweeps = 5;
        jc.outer = 3;
        jc.inner = 4;
        bt->jac_cfg = jc;
        bt->serial = true;
        bt->use_sub = false;
        // launched together the sync-free members are bound by throughput:
        // rows of under 12 entries take at most 2 lanes (more rows per wave)
        for (int m = 0; m < nmat; ++m)
            if (bt->choice[m] != SPTRSV_KERNEL_JACOBI && bt->avg[m] < 12) bt->lanes[m] = std::min(bt->lanes[m], 2);
        free_sub(bt);
        SPT_TRY(install_jacobi(bt));
        return build_tasks(bt);
    }
    TuneScratch ts;
    SPT_TRY(ts.init(bt->mat.nr));
    if (precision == SPTRSV_PREC_MIXED) return tune_mixed(bt, ts);
    for (int m = 0; m < nmat; ++m) {
        if (bt->pattern[m] != m) {  // timed once per structure
            bt->lanes[m] = bt->lanes[bt->pattern[m]];
            bt->t_sync[m] = bt->t_sync[bt->pattern[m]];
            continue;
        }
        const int ref = lanes_rule(bt->avg[m]);
        float best = 1e30f;
        for (int L = std::max(1, ref / 2); L <= std::min(64, ref * 2); L <<= 1) {
            float ms = 0;
            SPT_TRY(time_member(bt, ts, precision, m, L, 2, &ms));
            if (ms < best) {
                best = ms;
                bt->lanes[m] = L;
            }
        }
        bt->t_sync[m] = best;
    }
    SPT_TRY(sptrsv_check(&bt->mat, nullptr));
    return tune_concurrent_lanes(bt, ts, precision);
}

int sptrsv_batch_solve(sptrsv_batch_t* bt, int precision, const double* b, double* x, hipStream_t stream) {
    if (!bt || (bt->mat.nr > 0 && (!b || !x)) || b == x) return fail(SPTRSV_E_INVALID_ARG, "bad batch solve");
    return batch_solve_all(bt, precision, b, x, stream);
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
        int st = batch_solve_all(bt, precision, b, x, bt->cap_stream);
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

int sptrsv_batch_check(sptrsv_batch_t* bt, hipStream_t stream) {
    if (!bt) return fail(SPTRSV_E_INVALID_ARG, "null batch");
    const int st = sptrsv_check(&bt->mat, stream);
    if (st != SPTRSV_OK || !bt->use_sub || !bt->sub.priv) return st;
    return sptrsv_check(&bt->sub, stream);
}

void sptrsv_batch_destroy(sptrsv_batch_t* bt) {
    if (!bt) return;
    {
        std::lock_guard<std::mutex> g(g_batch_mu);
        g_batches.erase(bt);
    }
    drop_batch_graphs(bt);
    if (bt->d_tasks) (void)hipFree(bt->d_tasks);
    if (bt->d_sub) (void)hipFree(bt->d_sub);
    if (bt->d_tasks_sf) (void)hipFree(bt->d_tasks_sf);
    if (bt->cap_stream) (void)hipStreamDestroy(bt->cap_stream);
    if (bt->aux) (void)hipStreamDestroy(bt->aux);
    if (bt->fork_ev) (void)hipEventDestroy(bt->fork_ev);
    if (bt->join_ev) (void)hipEventDestroy(bt->join_ev);
    free_sub(bt);
    free_device_analysis(&bt->mat, nullptr);
    delete bt;
}

}  // extern "C"
