// common.hpp -- shared host/device definitions of the gfx950 SpTRSV library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
#include <map>
#include <string>

#include "sptrsv_amd.h"

namespace sptrsv {

// ---- error plumbing --------------------------------------------------------
void set_error(const std::string& msg);
int fail(int status, const std::string& msg);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define SPT_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) return ::sptrsv::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define SPT_TRY(call)                 \
    do {                              \
        int s_ = (call);              \
        if (s_ != SPTRSV_OK) return s_; \
    } while (0)

// ---- constants -------------------------------------------------------------
constexpr int kWave = 64;
// Readiness is carried by the published word itself: the all-ones pattern (a
// NaN payload that arithmetic never produces, see canon64/canon32) means
// "row not solved yet".  Levels use 0 (levels are >= 1).
constexpr uint64_t kUnsolved64 = ~0ull;
constexpr uint32_t kUnsolved32 = ~0u;

// ---- library-private state hung off sp_mat_t::priv ------------------------
struct Workspace;  // defined in solve.hip

// Row tiles of the Jacobi passes: 4 ints per tile (first row, end row, first
// entry, end entry) and, for a batch, the member of each tile (device).
struct TileTable {
    int* tiles = nullptr;
    int* member = nullptr;   // null: one member
    int count = 0;
    int K = 0;               // right-hand-side width the tiles were cut for
    int rows_max = 0;        // most rows in one tile
    // entries of each tile sorted by column (built on request): column,
    // scaled value, position inside the tile (packed above the column's
    // cbits bits of sja when cbits > 0; else in sslot)
    int* sja = nullptr;
    float* ssv = nullptr;
    unsigned short* sslot = nullptr;
    int cbits = 0;
};

struct MatPriv {
    int* counters = nullptr;     // [0] ticket, [1] watchdog word, [2..] spare (device)
    uint64_t* pub = nullptr;     // published words, n*k 64-bit (device)
    size_t pub_words = 0;
    uint64_t* part = nullptr;    // published partial sums of long-row segments, 2 per segment (device)
    size_t part_words = 0;
    // level-ordered copy (config reorder = 1): this matrix with its rows in the
    // level order of an analysis, that copy's own analysis, and the caller's
    // row of each copied row; built on first use, for the analysis whose id
    // is lc_src (rebuilt for another), dropped when the values change
    sp_mat_t* lc_mat = nullptr;
    dfr_analysis_info_t* lc_info = nullptr;
    int* lc_rows = nullptr;
    unsigned long long lc_src = 0;
    // in-degree walk: per-row counters and the off-diagonal pattern by column
    // (built on first use; structure only, so new values keep them)
    int* indeg = nullptr;
    int* dep_ptr = nullptr;
    int* dep_row = nullptr;
    double* resid = nullptr;     // two-pass refinement buffers (device)
    float* corr = nullptr;
    float* b32 = nullptr;
    size_t resid_len = 0;
    long long watchdog_ticks = 0;  // wall-clock ticks a wait may last
    double wall_rate_khz = 100000.0;
    int device = 0;
    int max_row = 0;             // longest row (entries incl. diagonal)
    long long* trace = nullptr;  // per-row timeline (profiling), 3*n*k
    size_t trace_len = 0;
    int backoff = 0;
    unsigned lds_pad = 0;        // extra LDS per workgroup, caps resident waves
    // JACOBI refinement
    double tol = 1e-10;          // certified relative accuracy
    double margin = 0.0;         // min_i |L_ii| - sum_j |L_ij|
    bool margin_known = false;
    // upper-triangular matrix stored as the lower triangle of the reversed
    // index order: solves reverse b into rev_b and the result back from rev_x
    bool reversed = false;
    // stored entry q holds value src_perm[q] of the value array the handle
    // was created from (src_count entries; identity when src_perm is null;
    // -1 marks the constant 1.0 of a unit diagonal): sptrsv_update_values
    // gathers new values through it, keeping structure and analysis
    int* src_perm = nullptr;
    long long src_count = 0;
    void* rev_b = nullptr;
    void* rev_x = nullptr;
    size_t rev_bytes = 0;
    float* jac_r = nullptr;      // FP32 row-scaled residual D^-1 r, n*k
    float* jac_sval = nullptr;   // FP32 a_ij / a_ii (0 on the diagonal), nnz
    double* jac_diag = nullptr;  // a_ii, n (contiguous, for the first iterate)
    float* jac_a = nullptr;      // Jacobi iterates (ping-pong), n*k each
    float* jac_b = nullptr;
    size_t jac_len = 0;
    unsigned long long* jac_norms = nullptr;  // per (member, column) max |r|, max |x| (double bits), replicated
    size_t jac_norms_len = 0;    // in 64-bit words
    TileTable jac_tt[2];         // row tiles of 1024 / 2048 products (residual uses [0])
    bool jac_tiles_fixed = false;  // tiles installed by the caller (batch), never rebuilt
    int jac_members = 1;         // independent systems sharing the certificate (batch)
    // captured solves (sptrsv_solve_graph): key = precision, b, x, n, k,
    // refinement steps and the resolved configuration
    std::map<std::array<long long, 21>, hipGraphExec_t> graphs;
    hipStream_t cap_stream = nullptr;
    bool tuned[3] = {false, false, false};
    sptrsv_config_t tuned_cfg[3] = {};
    int tuned_k[3] = {0, 0, 0};
};

// Long-row segments of an analysis (its wave64 multirow layout).  A row of
// more than kSegMin off-diagonal entries keeps its first kSegOwn entries;
// the rest is cut into runs of kSegEntries, each a segment item: one wave
// folds the run in a fixed order and publishes the partial sum, and the row
// adds its partials in segment order.  The multirow walks go through
// items[] (wave w >= 0, segment g as -(g + 1)); the segments of a wave's rows
// come right before that wave, so an item only ever waits on lower items.
// The arrays live on the device.
constexpr int kSegMin = 256;
constexpr int kSegOwn = 128;
constexpr int kSegEntries = 256;
// Segments of a row with `off` off-diagonal entries.
__host__ __device__ constexpr int seg_count(int off) {
    return off > kSegMin ? (off - kSegOwn + kSegEntries - 1) / kSegEntries : 0;
}
// Published partial words: all-ones = not yet published; the poison pattern
// (NaN in every policy's view) stands in for a partial the watchdog gave up on.
constexpr uint64_t kPartPoison = 0x7ff800007fc00000ull;
struct SegTable {
    int n_items = 0;         // n_waves + n_segs
    int n_segs = 0;
    int n_split = 0;         // rows that own segments
    int* items = nullptr;    // [n_items]
    int* seg_lo = nullptr;   // first entry of each segment [n_segs]
    int* seg_hi = nullptr;   // end entry of each segment [n_segs]
    int* row_seg = nullptr;  // first segment of each row, -1 for none [n]
};
// The analysis's segment table, or null when it has none.
const SegTable* segments_of(const dfr_analysis_info_t* info);
int build_segments(dfr_analysis_info_t* info, sp_mat_t* m);
// Identifies an analysis for the captured-solve cache: unique per analysis
// built, loaded or copied (never reused); 0 for none.
unsigned long long analysis_id(const dfr_analysis_info_t* info);

MatPriv* priv_of(sp_mat_t* m);
int ensure_priv(sp_mat_t* m);
// Destroys every captured solve of the matrix.  A captured solve holds the
// device addresses of the workspaces and the launch arguments it was
// captured with, so this runs before any of them is freed, reallocated or
// changed (waits for the device first: a replay may still be running).
void drop_graphs(MatPriv* pv);
void free_mat(sp_mat_t* m);  // device arrays and private state; zeroes *m
// Releases a matrix's level-ordered copy (if any).
void drop_level_copy(MatPriv* pv);
void free_info_arrays(dfr_analysis_info_t* info);
// The analysis of `copy`, the level-ordered copy made for analysis `src`,
// derived from `src` (levels permuted, iorder the identity, the rest shared).
int level_copy_analysis(const dfr_analysis_info_t* src, sp_mat_t* copy, dfr_analysis_info_t* dst);

// out[i] = sum_{j<i} in[j] on the device (in and out may alias)
int exclusive_scan(const int* in, int* out, long long n, hipStream_t s);
// Kernel timer (api_solve.hip): the next (start, stop) event pair while
// timing is enabled and the stream is not being captured, else an empty
// pair (false); the caller records it around `launches` consecutive
// launches of one kernel.  Returned by value: the pool may grow meanwhile.
struct KtPair {
    hipEvent_t ev[2] = {nullptr, nullptr};
    explicit operator bool() const { return ev[0] != nullptr; }
    hipEvent_t operator[](int i) const { return ev[i]; }
};
constexpr int kTimeSyncFree = 1;   // sync-free solve passes (kernel timer kinds)
constexpr int kTimeSweeps = 2;     // a refinement step's Jacobi sweeps
KtPair kernel_timer_pair(hipStream_t s, int launches = 1, int kind = kTimeSyncFree);
// stable LSD radix sort of (keys, vals) by key in [0, max_key]; tk/tv scratch
int radix_sort_pairs(int* keys, int* vals, int* tk, int* tv, int n, int max_key, hipStream_t s);

// ---- launch helpers ----------------------------------------------------------
inline unsigned ceil_div(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace sptrsv
