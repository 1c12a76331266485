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
#include <map>
#include <tuple>
#include <vector>

#include "syncfree.hpp"

namespace sptrsv {

hipEvent_t* kernel_timer_pair(hipStream_t s);
int upload_concat(const int* n, const int* const* ia, const int* const* ja, const double* const* a,
                  int nmat, sp_mat_t* m, std::vector<long long>& row_off);
int jacobi_solve(sp_mat_t* m, const sptrsv_config_t& c, int k, const double* b, double* x, hipStream_t s);
int jacobi_tune(sp_mat_t* m, int k, int reps, sptrsv_config_t* best, float* best_ms);
int reserve_jacobi(sp_mat_t* m, int k);
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
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
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
    SPT_TRY(upload_tasks(per, &bt->d_tasks, &bt->ntasks));
    SPT_TRY(upload_tasks(per_sf, &bt->d_tasks_sf, &bt->ntasks_sf));
    for (auto& g : bt->graphs) (void)hipGraphExecDestroy(g.second);
    bt->graphs.clear();
    return SPTRSV_OK;
}

// Hands the tiles of the matrices whose choice is JACOBI to the Jacobi path.
int install_jacobi(sptrsv_batch_t* bt) {
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

// The whole batched solve: FP64 runs every matrix sync-free; MIXED runs the
// Jacobi matrices on the auxiliary stream while the sync-free ones run on s.
int batch_solve_all(sptrsv_batch_t* bt, int precision, const double* b, double* x, hipStream_t s) {
    if (precision != SPTRSV_PREC_MIXED || bt->n_jacobi == 0)
        return batch_solve_tasks(bt, precision, bt->d_tasks, precision == SPTRSV_PREC_MIXED ? bt->ntasks_sf
                                                                                           : bt->ntasks,
                                 b, x, s);
    if (bt->serial) {  // the two halves one after the other, each with the whole device
        SPT_TRY(jacobi_solve(&bt->mat, bt->jac_cfg, 1, b, x, s));
        if (bt->ntasks_sf > 0) SPT_TRY(batch_solve_tasks(bt, precision, bt->d_tasks_sf, bt->ntasks_sf, b, x, s));
        return SPTRSV_OK;
    }
    SPT_HIP(hipEventRecord(bt->fork_ev, s));
    SPT_HIP(hipStreamWaitEvent(bt->aux, bt->fork_ev, 0));
    SPT_TRY(jacobi_solve(&bt->mat, bt->jac_cfg, 1, b, x, bt->aux));
    SPT_HIP(hipEventRecord(bt->join_ev, bt->aux));
    if (bt->ntasks_sf > 0) SPT_TRY(batch_solve_tasks(bt, precision, bt->d_tasks_sf, bt->ntasks_sf, b, x, s));
    SPT_HIP(hipStreamWaitEvent(s, bt->join_ev, 0));
    return SPTRSV_OK;
}

// Mean time of `reps` whole batched MIXED solves (after one warm-up).
float time_batch(sptrsv_batch_t* bt, const double* b, double* x, int reps, int* st) {
    hipEvent_t e0, e1;
    *st = SPTRSV_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        *st = fail(SPTRSV_E_HIP, "event create");
        return 1e30f;
    }
    *st = batch_solve_all(bt, SPTRSV_PREC_MIXED, b, x, nullptr);
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps && *st == SPTRSV_OK; ++r) *st = batch_solve_all(bt, SPTRSV_PREC_MIXED, b, x, nullptr);
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
    int st = upload_concat(n, ia, ja, a, nmat, &bt->mat, bt->row_off);
    if (st != SPTRSV_OK) {
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
    long long entry_base = 0;
    for (int m = 0; m < nmat; ++m) {
        long long rows = bt->row_off[m + 1] - bt->row_off[m];
        long long nz = rows > 0 ? (long long)ia[m][n[m]] - ia[m][0] : 0;
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
        if (rows > 0)
            for (int w = 0; w < 2; ++w)
                jacobi_host_tiles(ia[m], 0, n[m], bt->row_off[m], entry_base, 1024 << w, bt->mtiles[w][m]);
        entry_base += nz;
    }
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

int sptrsv_batch_schedule(const sptrsv_batch_t* bt, int* serial) {
    if (!bt || !serial) return fail(SPTRSV_E_INVALID_ARG, "null argument");
    *serial = bt->serial ? 1 : 0;
    return SPTRSV_OK;
}

int sptrsv_batch_set_schedule(sptrsv_batch_t* bt, int serial) {
    if (!bt || serial < 0 || serial > 1) return fail(SPTRSV_E_INVALID_ARG, "bad batch schedule");
    bt->serial = serial != 0;
    for (auto& g : bt->graphs) (void)hipGraphExecDestroy(g.second);
    bt->graphs.clear();
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

// MIXED per-matrix choice between the sync-free launch and Jacobi.  The
// Jacobi time J of all diagonally dominant matrices is measured first; a
// dominant matrix whose sync-free time alone exceeds J / 2 is "slow".  Then
// three whole concurrent batches are timed -- every matrix sync-free,
// Jacobi for every dominant matrix, Jacobi for the slow ones only -- and the
// fastest is kept.  Lanes per row are tuned only for matrices that can stay
// sync-free.
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
    for (int m = 0; m < nmat; ++m) {
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
        for (const std::vector<int>* set : {&dominant, &slow}) {
            if (set->empty() || (set == &slow && slow.size() == dominant.size())) continue;
            SPT_TRY(apply_choice(bt, *set, nullptr));
            const float t = time_batch(bt, ts.b, ts.x, 3, &st);
            SPT_TRY(st);
            if (t < best) {
                best = t;
                best_set = *set;
            }
        }
    }
    SPT_TRY(apply_choice(bt, best_set, nullptr));
    // the two halves concurrently on two streams, or one after the other
    // (the concurrent Jacobi sweeps can starve the latency-bound sync-free
    // waves)
    bt->serial = false;
    if (bt->n_jacobi > 0 && bt->ntasks_sf > 0) {
        const float tc = time_batch(bt, ts.b, ts.x, 5, &st);
        SPT_TRY(st);
        bt->serial = true;
        const float tsr = time_batch(bt, ts.b, ts.x, 5, &st);
        SPT_TRY(st);
        bt->serial = tsr < tc;
    }
    for (auto& g : bt->graphs) (void)hipGraphExecDestroy(g.second);
    bt->graphs.clear();
    return sptrsv_check(&bt->mat, nullptr);
}

}  // namespace

extern "C" {

int sptrsv_batch_tune(sptrsv_batch_t* bt, int precision, int tune) {
    if (!bt) return fail(SPTRSV_E_INVALID_ARG, "null batch");
    if (precision != SPTRSV_PREC_F64 && precision != SPTRSV_PREC_MIXED)
        return fail(SPTRSV_E_INVALID_ARG, "batched solves take FP64 or MIXED precision");
    const int nmat = (int)bt->lanes.size();
    for (int m = 0; m < nmat; ++m) bt->lanes[m] = lanes_rule(bt->avg[m]);
    if (!tune) {
        bt->choice.assign(nmat, SPTRSV_KERNEL_SIMPLE);
        SPT_TRY(install_jacobi(bt));
        return build_tasks(bt);
    }
    TuneScratch ts;
    SPT_TRY(ts.init(bt->mat.nr));
    if (precision == SPTRSV_PREC_MIXED) return tune_mixed(bt, ts);
    for (int m = 0; m < nmat; ++m) {
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
    return build_tasks(bt);
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
    return sptrsv_check(&bt->mat, stream);
}

void sptrsv_batch_destroy(sptrsv_batch_t* bt) {
    if (!bt) return;
    for (auto& g : bt->graphs) (void)hipGraphExecDestroy(g.second);
    if (bt->d_tasks) (void)hipFree(bt->d_tasks);
    if (bt->d_sub) (void)hipFree(bt->d_sub);
    if (bt->d_tasks_sf) (void)hipFree(bt->d_tasks_sf);
    if (bt->cap_stream) (void)hipStreamDestroy(bt->cap_stream);
    if (bt->aux) (void)hipStreamDestroy(bt->aux);
    if (bt->fork_ev) (void)hipEventDestroy(bt->fork_ev);
    if (bt->join_ev) (void)hipEventDestroy(bt->join_ev);
    free_device_analysis(&bt->mat, nullptr);
    delete bt;
}

}  // extern "C"
