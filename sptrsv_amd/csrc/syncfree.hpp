// syncfree.hpp -- the sync-free dependency-driven row kernel shared by the
// level analysis and every solver (simple / order / multirow; FP64, FP32 and
// fused mixed precision; one or several right-hand sides).
//
// Method (solve_csr_simple.cu:9-108, solve_csr_order.cu:83-150,
// solve_csr_multirow.cu:21-114, analysis_csr.cu:104-185): every row is
// owned by a group of lanes that walks the row's off-diagonal entries, waits
// until the row each entry depends on has been published, folds it in, then
// reduces across the group and publishes its own result.  Workgroups take a
// ticket from an atomic counter when they start, so rows are handed out in
// start order and every row a running wave waits on belongs to a wave that is
// already running (deadlock freedom without a grid barrier).
//
// gfx950 design:
//  * Readiness is carried by the published word itself (all-ones pattern =
//    unsolved, levels use 0), so publishing is ONE agent-scope relaxed store
//    and waiting is agent-scope relaxed polling loads (global_load sc1): no
//    separate flag, no release/acquire fence, no L2 writeback.
//  * Wave64 lane groups: a group is L entry-lanes x K right-hand-side lanes;
//    64/(L*K) rows per wave.  Rows of one wave may depend on each other (the
//    natural and iorder layouts allow it), so a group reduces and publishes
//    as soon as all its lanes are done, INSIDE the polling loop; the wave
//    leaves the loop only when every group has published.  (A lockstep wave
//    that reduced after the loop would deadlock on an in-wave dependency.)
//  * Rows of the same workgroup (natural order) are also published to LDS and
//    read from there.
//  * Each lane keeps a window of U entries in flight and consumes the ready
//    prefix in storage order, so long rows overlap their dependency loads
//    while the summation order stays fixed (results are bit-reproducible).
//  * Every wait is bounded by a wall-clock budget: past it the wave flags the
//    matrix's watchdog word and treats the missing value as NaN, so a
//    malformed input ends the kernel instead of hanging the GPU.
#pragma once

#include <algorithm>

#include "common.hpp"

namespace sptrsv {

struct SolveArgs {
    int n;                  // rows
    int k;                  // right-hand sides (row-major, stride k)
    const int* ia;
    const int* ja;
    const double* a64;
    const float* a32;
    const void* b;          // double* (F64, MIXED) or float* (F32); n*k
    void* pub;              // published words, n*k (x itself for F64/F32)
    double* x_out;          // MIXED: FP64 solution n*k
    const int* iorder;      // ORDER / MULTIROW
    const int* wbase;       // MULTIROW: first iorder position of each wave
    const int* wsize;       // MULTIROW: lanes per row of each wave (0 = diagonal only)
    int n_waves;            // MULTIROW
    int L;                  // SIMPLE / ORDER: entry lanes per row
    int lgL;                // log2(L)
    int* ticket;
    int* err;
    long long wd_ticks;
    long long* trace;       // optional per-row timeline: start, reduced, stored (wall clock)
    int backoff;            // 0: spin; 1..4: s_sleep between polling turns without progress
    unsigned dyn_lds;       // extra LDS bytes per workgroup (caps resident waves; unused by the kernels)
    long long persist_grid; // persistent kernels: workgroups to launch
    int claim_rounds;       // persistent kernels: items per wave in each claimed block
    // long-row segments of the multirow walks, one right-hand side (null
    // items: rows are never split)
    const int* items;       // walk order: wave w >= 0, segment g as -(g + 1)
    int n_items;
    const int* seg_lo;      // entries [seg_lo, seg_hi) of segment g
    const int* seg_hi;
    const int* row_seg;     // first segment of each row, -1 for none
    uint64_t* part;         // partial sums, 2 words per segment (all-ones = not published)
    // in-degree walk (variant 3): per-row count of unsolved dependencies and
    // the dependents of each row (the off-diagonal pattern by column)
    int* indeg;
    const int* dep_ptr;     // n + 1
    const int* dep_row;     // rows i > j with an entry in column j, ascending
    // level-ordered copy: row p of this matrix is the caller's row
    // row_orig[p], whose b is read and x written there; the published words
    // stay in this matrix's order (null: the caller's order)
    const int* row_orig;
};

// Caller's index of row `row`, right-hand side r.
__device__ __forceinline__ size_t caller_index(const SolveArgs& A, int row, int r) {
    return (size_t)(A.row_orig ? A.row_orig[row] : row) * A.k + r;
}

template <int K>
__host__ __device__ constexpr int kLog2() {
    return K <= 1 ? 0 : 1 + kLog2<K / 2>();
}

// ---- memory primitives -------------------------------------------------------
template <class T>
__device__ __forceinline__ T poll_global(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void publish_global(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T poll_lds(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ __forceinline__ void publish_lds(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ uint64_t canon64(double v) {
    uint64_t u = (uint64_t)__double_as_longlong(v);
    return u == kUnsolved64 ? 0x7ff8000000000000ull : u;
}
__device__ __forceinline__ uint32_t canon32(float v) {
    uint32_t u = __float_as_uint(v);
    return u == kUnsolved32 ? 0x7fc00000u : u;
}

// ---- cross-lane reduction steps (one lane group of G lanes, K == 1) ----------
// Step s combines lanes 2^s apart: s = 0, 1 swap within quads (DPP quad_perm),
// s = 2, 3 mirror within half-rows / rows of 16 (DPP row_half_mirror,
// row_mirror; each half already holds one value after the earlier steps, so a
// mirror pairs the two halves), s = 4, 5 exchange rows 16 / 32 lanes apart
// (v_permlane16_swap / v_permlane32_swap).  Every lane of the group ends with
// the same, bit-identical total (the combine is commutative).
__device__ __forceinline__ uint32_t dpp_partner(uint32_t v, int s) {
    switch (s) {
        case 0: return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
        case 1: return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
        case 2: return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);
        default: return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);
    }
}

template <class T, class Op>
__device__ __forceinline__ T lane_combine(T v, int s, Op op) {
    constexpr int N = sizeof(T) / 4;
    uint32_t w[N];
    __builtin_memcpy(w, &v, sizeof(T));
    uint32_t a[N], b[N];
    if (s < 4) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            a[i] = w[i];
            b[i] = dpp_partner(w[i], s);
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            auto r = (s == 4) ? __builtin_amdgcn_permlane16_swap(w[i], w[i], false, false)
                              : __builtin_amdgcn_permlane32_swap(w[i], w[i], false, false);
            a[i] = r[0];
            b[i] = r[1];
        }
    }
    T x, y;
    __builtin_memcpy(&x, a, sizeof(T));
    __builtin_memcpy(&y, b, sizeof(T));
    return op(x, y);
}

// Value of lane `src` (wave-uniform) in every lane, via v_readlane.
template <class T>
__device__ __forceinline__ T bcast(T v, int src) {
    constexpr int N = (sizeof(T) + 3) / 4;
    uint32_t w[N] = {};
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < N; ++i) w[i] = (uint32_t)__builtin_amdgcn_readlane((int)w[i], src);
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
}

struct OpAdd {
    template <class T>
    __device__ T operator()(T x, T y) const { return x + y; }
};
struct OpMax {
    __device__ int operator()(int x, int y) const { return max(x, y); }
};

// ---- precision / operation policies -----------------------------------------

// Plain FP64: x_i = (b_i - sum L_ij x_j) * (1 / L_ii), published into x.
struct PolF64 {
    using word = uint64_t;
    static constexpr word kUnsolved = kUnsolved64;
    static constexpr word kPoison = 0x7ff8000000000000ull;
    struct Acc { double s; };
    struct Val { double v; };
    struct Diag { double piv; };
    __device__ static Acc zero() { return {0.0}; }
    __device__ static Acc seed(const SolveArgs& A, size_t bi) {
        return {static_cast<const double*>(A.b)[bi]};
    }
    __device__ static Val load(const SolveArgs& A, int p) { return {A.a64[p]}; }
    __device__ static Diag diag(const SolveArgs& A, int dpos) { return {1.0 / A.a64[dpos]}; }
    __device__ static void fold(Acc& c, const Val& v, word w) {
        c.s = __fma_rn(-v.v, __longlong_as_double((long long)w), c.s);
    }
    struct Out {};
    __device__ static Acc add(const Acc& x, const Acc& y) { return {x.s + y.s}; }
    __device__ static Acc shfl_xor(const Acc& c, int o) { return {__shfl_xor(c.s, o)}; }
    __device__ static Acc reduce_step(const Acc& c, int s) { return {lane_combine(c.s, s, OpAdd{})}; }
    __device__ static word finish(const SolveArgs&, const Diag& d, const Acc& t, Out&) {
        return canon64(t.s * d.piv);
    }
    // x itself is the published array, unless the solve publishes in another
    // order (a level-ordered copy): then x receives the published word
    __device__ static void emit(const SolveArgs& A, size_t bi, const Out&, word w) {
        if (A.x_out) __builtin_nontemporal_store(__longlong_as_double((long long)w), &A.x_out[bi]);
    }
    __device__ static Out bcast_out(const Out& o, int) { return o; }
    // a segment's partial sum as published words (never all-ones)
    static constexpr int kPartWords = 1;
    // general engine: form every poll address of a turn before the first
    // poll goes out (FP64 only: measured faster there, slower for the
    // fused mixed pass)
    static constexpr bool kPollsTogether = true;
    __device__ static void part_pack(const Acc& a, uint64_t* w) { w[0] = canon64(a.s); }
    __device__ static Acc part_unpack(const uint64_t* w) { return {__longlong_as_double((long long)w[0])}; }
};

// Plain FP32 on the FP32 copy of the values, published into x (float).
struct PolF32 {
    using word = uint32_t;
    static constexpr word kUnsolved = kUnsolved32;
    static constexpr word kPoison = 0x7fc00000u;
    struct Acc { float s; };
    struct Val { float v; };
    struct Diag { float piv; };
    __device__ static Acc zero() { return {0.0f}; }
    __device__ static Acc seed(const SolveArgs& A, size_t bi) {
        return {static_cast<const float*>(A.b)[bi]};
    }
    __device__ static Val load(const SolveArgs& A, int p) { return {A.a32[p]}; }
    __device__ static Diag diag(const SolveArgs& A, int dpos) { return {1.0f / A.a32[dpos]}; }
    __device__ static void fold(Acc& c, const Val& v, word w) {
        c.s = __fmaf_rn(-v.v, __uint_as_float(w), c.s);
    }
    struct Out {};
    __device__ static Acc add(const Acc& x, const Acc& y) { return {x.s + y.s}; }
    __device__ static Acc shfl_xor(const Acc& c, int o) { return {__shfl_xor(c.s, o)}; }
    __device__ static Acc reduce_step(const Acc& c, int s) { return {lane_combine(c.s, s, OpAdd{})}; }
    __device__ static word finish(const SolveArgs&, const Diag& d, const Acc& t, Out&) {
        return canon32(t.s * d.piv);
    }
    __device__ static void emit(const SolveArgs&, size_t, const Out&, word) {}
    __device__ static Out bcast_out(const Out& o, int) { return o; }
    static constexpr int kPartWords = 1;
    static constexpr bool kPollsTogether = false;
    __device__ static void part_pack(const Acc& a, uint64_t* w) { w[0] = canon32(a.s); }
    __device__ static Acc part_unpack(const uint64_t* w) { return {__uint_as_float((uint32_t)w[0])}; }
};

// Fused mixed precision, one refinement step in the same dependency pass:
//   x0_i = (b_i - sum a32_ij x0_j) / a32_ii                      (FP32)
//   r_i  =  b_i - sum a64_ij x0_j - a64_ii x0_i                  (FP64)
//   d_i  = (r_i - sum a32_ij d_j) / a32_ii                       (FP32)
//   x_i  = x0_i + d_i                                            (FP64)
// Row i publishes the packed pair (x0_i, d_i); x0 and d have exactly the
// dependencies of x, so the correction solve rides on the bulk solve's
// critical path instead of walking it a second time.
struct PolMixed {
    using word = uint64_t;
    static constexpr word kUnsolved = kUnsolved64;
    static constexpr word kPoison = 0x7fc000007fc00000ull;
    struct Acc { float x0; float d; double r; };
    struct Val { float v32; double v64; };
    struct Diag { float piv; double a; };
    __device__ static Acc zero() { return {0.0f, 0.0f, 0.0}; }
    __device__ static Acc seed(const SolveArgs& A, size_t bi) {
        double b = static_cast<const double*>(A.b)[bi];
        return {(float)b, 0.0f, b};
    }
    __device__ static Val load(const SolveArgs& A, int p) { return {A.a32[p], A.a64[p]}; }
    __device__ static Diag diag(const SolveArgs& A, int dpos) {
        return {1.0f / A.a32[dpos], A.a64[dpos]};
    }
    __device__ static void fold(Acc& c, const Val& v, word w) {
        float x0 = __uint_as_float((uint32_t)w);
        float d = __uint_as_float((uint32_t)(w >> 32));
        c.x0 = __fmaf_rn(-v.v32, x0, c.x0);
        c.d = __fmaf_rn(-v.v32, d, c.d);
        c.r = __fma_rn(-v.v64, (double)x0, c.r);
    }
    __device__ static Acc add(const Acc& x, const Acc& y) {
        return {x.x0 + y.x0, x.d + y.d, x.r + y.r};
    }
    __device__ static Acc shfl_xor(const Acc& c, int o) {
        return {__shfl_xor(c.x0, o), __shfl_xor(c.d, o), __shfl_xor(c.r, o)};
    }
    __device__ static Acc reduce_step(const Acc& c, int s) {
        return {lane_combine(c.x0, s, OpAdd{}), lane_combine(c.d, s, OpAdd{}),
                lane_combine(c.r, s, OpAdd{})};
    }
    struct Out { double x; };
    __device__ static void emit(const SolveArgs& A, size_t bi, const Out& o, word) {
        __builtin_nontemporal_store(o.x, &A.x_out[bi]);
    }
    __device__ static Out bcast_out(const Out& o, int src) { return {bcast(o.x, src)}; }
    __device__ static word finish(const SolveArgs&, const Diag& dg, const Acc& t, Out& o) {
        const float piv = dg.piv;
        float x0 = t.x0 * piv;
        double r = __fma_rn(-dg.a, (double)x0, t.r);
        float d = ((float)r + t.d) * piv;
        o.x = (double)x0 + (double)d;
        uint32_t lo = canon32(x0);
        uint32_t hi = __float_as_uint(d);
        return ((uint64_t)hi << 32) | lo;  // never all-ones: lo is canonical
    }
    static constexpr int kPartWords = 2;
    static constexpr bool kPollsTogether = false;
    __device__ static void part_pack(const Acc& a, uint64_t* w) {
        w[0] = ((uint64_t)canon32(a.d) << 32) | canon32(a.x0);
        w[1] = canon64(a.r);
    }
    __device__ static Acc part_unpack(const uint64_t* w) {
        return {__uint_as_float((uint32_t)w[0]), __uint_as_float((uint32_t)(w[0] >> 32)),
                __longlong_as_double((long long)w[1])};
    }
};

// Level analysis: lev_i = 1 + max lev_j (analysis_csr.cu:147-181).  0 marks
// an unsolved row, so a zero-filled level array is the initial state.
struct PolLevel {
    using word = uint32_t;
    static constexpr word kUnsolved = 0u;
    static constexpr word kPoison = 1u;
    struct Acc { int m; };
    struct Val {};
    struct Diag {};
    __device__ static Acc zero() { return {0}; }
    __device__ static Acc seed(const SolveArgs&, size_t) { return {0}; }
    __device__ static Val load(const SolveArgs&, int) { return {}; }
    __device__ static Diag diag(const SolveArgs&, int) { return {}; }
    __device__ static void fold(Acc& c, const Val&, word w) { c.m = max(c.m, (int)w); }
    struct Out {};
    __device__ static Acc add(const Acc& x, const Acc& y) { return {max(x.m, y.m)}; }
    __device__ static Acc shfl_xor(const Acc& c, int o) { return {__shfl_xor(c.m, o)}; }
    __device__ static Acc reduce_step(const Acc& c, int s) { return {lane_combine(c.m, s, OpMax{})}; }
    __device__ static word finish(const SolveArgs&, const Diag&, const Acc& t, Out&) {
        return (word)(1 + t.m);
    }
    __device__ static void emit(const SolveArgs&, size_t, const Out&, word) {}
    __device__ static Out bcast_out(const Out& o, int) { return o; }
    static constexpr int kPartWords = 1;
    static constexpr bool kPollsTogether = false;
    __device__ static void part_pack(const Acc& a, uint64_t* w) { w[0] = (uint64_t)(uint32_t)a.m; }
    __device__ static Acc part_unpack(const uint64_t* w) { return {(int)(uint32_t)w[0]}; }
};

// All lanes of a group of G lanes (K right-hand-side lanes interleaved)
// receive the group's total.
template <class P, int K>
__device__ __forceinline__ typename P::Acc group_reduce(typename P::Acc t, int G) {
    if constexpr (K == 1) {
        int s = 0;
        for (int o = 1; o < G; o <<= 1, ++s) t = P::reduce_step(t, s);
    } else {
        for (int o = K; o < G; o <<= 1) t = P::add(t, P::shfl_xor(t, o));
    }
    return t;
}

// ---- the row engine -----------------------------------------------------------
// One call handles the lane group of `row` (row < 0: idle group) for the
// calling lane.  L = entry lanes per row (power of two), K = right-hand-side
// lanes (power of two, lanes with r >= A.k idle), L*K divides 64.
// s_pub/lds_lo/lds_hi: rows in [lds_lo, lds_hi) are also published to and
// read from LDS (natural order only; pass lds_hi = lds_lo to disable).
template <class P, int K, int U, bool SPLIT = false>
__device__ __forceinline__ void sf_rows(const SolveArgs& A, int L, int lane, int row,
                                        typename P::word* s_pub, int lds_lo, int lds_hi) {
    using word = typename P::word;
    using Acc = typename P::Acc;
    using Val = typename P::Val;
    word* gpub = static_cast<word*>(A.pub);

    const int G = L * K;                      // lanes per row group
    const int gl = lane % G;                  // lane within the group
    const int e = gl / K;                     // entry lane
    const int r = gl % K;                     // right-hand side
    const bool live = row >= 0 && r < A.k;

    int rs = 0, dpos = 0;
    if (live) {
        rs = A.ia[row];
        dpos = A.ia[row + 1] - 1;
    }
    // A split long row (SPLIT: multirow walks with a segment table, one
    // right-hand side) folds its first kSegOwn entries, then its segments'
    // partial sums: lane e adds partials e, e + L, ... in that order.
    int eend = dpos;
    int pg = e, pg_end = 0;  // this lane's next partial (absolute index), end of the row's partials
    if constexpr (SPLIT && K == 1) {
        if (A.items && live && dpos - rs > kSegMin) {
            const int g = A.row_seg[row];
            if (g >= 0) {
                pg = g + e;
                pg_end = g + seg_count(dpos - rs);
                eend = rs + kSegOwn;
            }
        }
    }
    const size_t xi = (size_t)(live ? row : 0) * A.k + r;           // published word
    const size_t bi = live ? caller_index(A, row, r) : 0;  // b and x (the caller's order)
    Acc acc = (live && e == 0) ? P::seed(A, bi) : P::zero();
    // The diagonal (and its reciprocal) is independent of the dependencies:
    // load it now so it is off the critical path at publish time.
    typename P::Diag dg{};
    if (live && e == 0) dg = P::diag(A, dpos);

    int p = rs + e;     // first entry of the current window
    int kc = 0;         // entries of the window already folded
    bool edone = !(live && p < eend);  // entries folded
    bool done = edone && pg >= pg_end;  // entries and partials folded
    bool published = !live;

    int col[U];
    Val val[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        int q = p + u * L;
        col[u] = 0;
        if (!edone && q < eend) {
            col[u] = A.ja[q];
            val[u] = P::load(A, q);
        }
    }
    // With 4-entry windows (rows longer than a window) the next window's
    // entries load while the current one is polled, so a long row pays the
    // entry-load latency once rather than once per window.  (16-entry
    // windows already hold most rows; doubling their registers would cost
    // more occupancy than it saves.)
    constexpr bool kPrefetch = (U == 4);
    int ncol[U];
    Val nval[U];
    if constexpr (kPrefetch) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = p + U * L + u * L;
            ncol[u] = 0;
            if (!edone && q < eend) {
                ncol[u] = A.ja[q];
                nval[u] = P::load(A, q);
            }
        }
    }

    const unsigned long long gmask =
        (G == 64) ? ~0ull : (((1ull << G) - 1ull) << ((lane / G) * G));
    const long long t0 = wall_clock64();
    if (A.trace && live && e == 0) A.trace[3 * xi] = t0;
    bool expired = false;
    int spins = 0;
    // A reduced row waits one loop turn before it is stored: on gfx9 the
    // vector-memory counter retires in order, so a poll issued after a store
    // could not be consumed before that store completed.  Issuing the turn's
    // global polls first and the stores after them keeps the store latency
    // off every poll.
    bool pending = false;
    word pend_w = 0;
    typename P::Out pend_o{};

    while (true) {
        bool moved = false;  // a dependency was folded this turn
        // (1) global polls of the window entries not folded yet
        word w[U];
        bool in_lds[U];
        if constexpr (P::kPollsTogether) {
            // every address formed before the first poll goes out (a column
            // still arriving from the previous turn's window load is waited
            // for here), so the polls leave back to back
            bool in_mem[U];
            const word* ad[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = p + u * L;
                in_lds[u] = in_mem[u] = false;
                ad[u] = gpub;
                if (!edone && u >= kc && q < eend) {
                    const int c = col[u];
                    if (c >= lds_lo && c < lds_hi) {
                        in_lds[u] = true;
                    } else {
                        in_mem[u] = true;
                        ad[u] = &gpub[(size_t)c * A.k + r];
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                w[u] = P::kUnsolved;
                if (in_mem[u]) w[u] = poll_global(ad[u]);
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = p + u * L;
                w[u] = P::kUnsolved;
                in_lds[u] = false;
                if (!edone && u >= kc && q < eend) {
                    const int c = col[u];
                    if (c >= lds_lo && c < lds_hi)
                        in_lds[u] = true;
                    else
                        w[u] = poll_global(&gpub[(size_t)c * A.k + r]);
                }
            }
        }
        // the next partial of a split row whose entries are folded
        uint64_t pw[P::kPartWords];
        const bool ppoll = SPLIT && (K == 1) && edone && !done;
#pragma unroll
        for (int i = 0; i < P::kPartWords; ++i)
            pw[i] = ppoll ? poll_global(&A.part[2 * (size_t)pg + i]) : kUnsolved64;
        asm volatile("" ::: "memory");
        // (3) LDS polls, after this wave's own LDS stores; (4) fold the ready
        // prefix of the window in storage order.  LDS results live in their
        // own registers so they never wait on the global polls.
        auto consume = [&]() {
            word wl[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                wl[u] = in_lds[u] ? poll_lds(&s_pub[col[u] - lds_lo]) : P::kUnsolved;
            if (ppoll) {
                bool ok = true;
#pragma unroll
                for (int i = 0; i < P::kPartWords; ++i) ok &= pw[i] != kUnsolved64;
                if (!ok && expired) {
#pragma unroll
                    for (int i = 0; i < P::kPartWords; ++i) pw[i] = kPartPoison;
                    ok = true;
                }
                if (ok) {
                    acc = P::add(acc, P::part_unpack(pw));
                    moved = true;
                    pg += L;
                    done = pg >= pg_end;
                }
            }
            if (!edone) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int q = p + u * L;
                    word v = in_lds[u] ? wl[u] : w[u];
                    if (expired && u >= kc && q < eend && v == P::kUnsolved) v = P::kPoison;
                    if (u == kc && q < eend && v != P::kUnsolved) {
                        P::fold(acc, val[u], v);
                        moved = true;
                        ++kc;
                    }
                }
                if (p + kc * L >= eend) {
                    edone = true;
                    done = pg >= pg_end;
                } else if (kc == U) {
                    p += U * L;
                    kc = 0;
                    if constexpr (kPrefetch) {
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            col[u] = ncol[u];
                            val[u] = nval[u];
                            const int q = p + U * L + u * L;
                            if (q < eend) {
                                ncol[u] = A.ja[q];
                                nval[u] = P::load(A, q);
                            }
                        }
                    } else {
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const int q = p + u * L;
                            if (q < eend) {
                                col[u] = A.ja[q];
                                val[u] = P::load(A, q);
                            }
                        }
                    }
                }
            }
        };
        // (2) stores of the rows reduced in the previous turn.  The branch is
        // wave-uniform and consume() is repeated in both arms, so each arm
        // waits for exactly its own outstanding memory operations.
        const unsigned long long pmask = __ballot(pending);
        if (pmask) {
            // Every lane stores: lanes without a row to publish repeat the
            // first pending lane's store (same address, same value), so the
            // store is not skipped on any path and the polls above are waited
            // for with the stores still in flight.
            const int src = __builtin_ctzll(pmask);
            const size_t sxi = pending ? xi : bcast(xi, src);
            // x's index: the published word's, or (level-ordered copy) the
            // caller's, kept from the start (a lookup here would wait for
            // the polls in flight)
            const size_t sbi = !A.row_orig ? sxi : pending ? bi : bcast(bi, src);
            const word sw = pending ? pend_w : bcast(pend_w, src);
            const typename P::Out so = pending ? pend_o : P::bcast_out(pend_o, src);
            if (pending && row >= lds_lo && row < lds_hi) publish_lds(&s_pub[row - lds_lo], pend_w);
            if (A.trace && pending) A.trace[3 * xi + 2] = wall_clock64();
            publish_global(&gpub[sxi], sw);
            P::emit(A, sbi, so, sw);  // x in the caller's order: nobody waits on it
            pending = false;
            asm volatile("" ::: "memory");
            consume();
        } else {
            consume();
        }
        // (5) reduce the groups whose lanes are all done
        const unsigned long long dm = __ballot(done);
        const bool ready = ((dm & gmask) == gmask) && !published;
        if (__ballot(ready)) {
            const Acc t = group_reduce<P, K>(acc, G);
            if (ready) {
                if (e == 0) {
                    if (A.trace) A.trace[3 * xi + 1] = wall_clock64();
                    pend_w = P::finish(A, dg, t, pend_o);
                    pending = true;
                }
                published = true;
            }
        }
        if (__ballot(!published || pending) == 0ull) break;
        if (A.backoff && !__ballot(moved || pending)) {
            switch (A.backoff) {
                case 1: __builtin_amdgcn_s_sleep(1); break;
                case 2: __builtin_amdgcn_s_sleep(2); break;
                case 3: __builtin_amdgcn_s_sleep(4); break;
                default: __builtin_amdgcn_s_sleep(8); break;
            }
        }
        if (++spins == 512) {
            spins = 0;
            if (!expired && wall_clock64() - t0 > A.wd_ticks) {
                expired = true;
                if (lane == 0) atomicOr(A.err, 1);
            }
        }
    }
}

// ---- kernels -----------------------------------------------------------------
// A.L = entry lanes per row (power of two, L*K <= 64), A.lgL = log2(A.L).

// Natural row order (simple): a workgroup owns consecutive rows and, for one
// right-hand side, also keeps their published words in LDS.
template <class P, int K, int WG, int U>
__global__ __launch_bounds__(WG) void k_simple(SolveArgs A) {
    using word = typename P::word;
    __shared__ word s_pub[(K == 1) ? WG : 1];
    __shared__ int s_base;
    const int tid = threadIdx.x;
    const int lgG = A.lgL + kLog2<K>();
    const int RWG = WG >> lgG;                 // rows per workgroup
    if (tid == 0) s_base = atomicAdd(A.ticket, 1) * RWG;
    if (K == 1)
        for (int i = tid; i < RWG; i += WG) s_pub[i] = P::kUnsolved;
    __syncthreads();
    const int base = s_base;
    const int lane = tid & (kWave - 1);
    const int pos = base + (tid >> lgG);
    const int row = pos < A.n ? pos : -1;
    const int hi = (K == 1) ? min(base + RWG, A.n) : base;
    sf_rows<P, K, U>(A, A.L, lane, row, s_pub, base, hi);
}

// Level order (order): rows at consecutive positions of iorder.
template <class P, int K, int WG, int U>
__global__ __launch_bounds__(WG) void k_order(SolveArgs A) {
    __shared__ int s_base;
    const int tid = threadIdx.x;
    const int lgG = A.lgL + kLog2<K>();
    const int RWG = WG >> lgG;
    if (tid == 0) s_base = atomicAdd(A.ticket, 1) * RWG;
    __syncthreads();
    const int lane = tid & (kWave - 1);
    const int pos = s_base + (tid >> lgG);
    const int row = pos < A.n ? A.iorder[pos] : -1;
    sf_rows<P, K, U>(A, A.L, lane, row, nullptr, 0, 0);
}

// Multirow work item t: with a segment table, items[t] (a wave, or segment g
// as -(g + 1)); else wave t.
__device__ __forceinline__ int multirow_item(const SolveArgs& A, int t) { return A.items ? A.items[t] : t; }
__device__ __forceinline__ int multirow_items(const SolveArgs& A) { return A.items ? A.n_items : A.n_waves; }

// A segment item of a split long row (defined with the lean walk below).
template <class P, int U>
__device__ __forceinline__ void fast_segment(const SolveArgs& A, int lane, int g);

// (level, class) packed waves (multirow): wave w solves the rows at iorder
// positions [wbase[w], wbase[w+1]) with wsize[w] lanes per row.  Rows of one
// wave share a level, so they never depend on each other.  Segment items
// (one right-hand side) are folded by the whole wave.
template <class P, int K, int WG, int U>
__global__ __launch_bounds__(WG) void k_multirow(SolveArgs A) {
    __shared__ int s_t;
    const int tid = threadIdx.x;
    if (tid == 0) s_t = atomicAdd(A.ticket, 1);
    __syncthreads();
    const int t = s_t * (WG / kWave) + (tid >> 6);
    if (t >= multirow_items(A)) return;
    const int w = multirow_item(A, t);
    const int lane = tid & (kWave - 1);
    if (w < 0) {
        if constexpr (K == 1) fast_segment<P, U>(A, lane, -w - 1);
        return;
    }
    const int vs = A.wsize[w];
    int L = vs == 0 ? 1 : vs;
    if (L * K > kWave) L = kWave / K;
    const int b0 = A.wbase[w];
    const int nrows = A.wbase[w + 1] - b0;
    const int cap = kWave / (L * K);           // rows one pass of the wave holds
    const int seg = lane / (L * K);
    // With several right-hand sides a wave of the 64-lane layout can hold more
    // rows than lane groups; its rows share a level, so passes are independent.
    for (int c0 = 0; c0 < nrows; c0 += cap) {
        const int row = (c0 + seg < nrows) ? A.iorder[b0 + c0 + seg] : -1;
        sf_rows<P, K, U, true>(A, L, lane, row, nullptr, 0, 0);
    }
}

// ---- fast path: no dependency inside a wave -----------------------------------
// For layouts where the rows of one wave cannot depend on each other (one row
// per wave, or a multirow wave whose rows share a level), a wave polls until
// every dependency of its current chunk of entries is present, folds the chunk
// in storage order, and reduces and publishes once at the end.  The polling
// turn is a handful of instructions (predicated loads, compares, one ballot),
// which keeps the detection delay close to the memory round trip.  Chunks are
// U entries per lane (for U <= 4 the next chunk's entries load while the
// current one is polled); folding order per lane is the same as sf_rows'.
// Folds entries [lo, hi) into acc: lane e of a group of L lanes takes
// entries lo + e, lo + e + L, ... in that order, in chunks of U whose
// dependencies are all present before the chunk is folded.
template <class P, int U>
__device__ __forceinline__ void fold_entries(const SolveArgs& A, int L, int e, int lane, bool live, int lo, int hi,
                                             typename P::Acc& acc, long long t0, bool& expired, int& spins) {
    using word = typename P::word;
    using Val = typename P::Val;
    const word* gpub = static_cast<const word*>(A.pub);
    // Entries of the next chunk are loaded before the current chunk is
    // polled (chunks of up to 4 entries per lane), so a long row pays the
    // entry-load latency once, not per chunk.
    constexpr bool kPrefetch = (U <= 4);
    int col[U], ncol[U];
    Val val[U], nval[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int q = lo + e + u * L;
        col[u] = 0;
        if (live && q < hi) {
            col[u] = A.ja[q];
            val[u] = P::load(A, q);
        }
    }
    for (int p = lo + e; __ballot(live && p < hi); p += U * L) {
        word got[U];
        bool pend[U];
        if constexpr (kPrefetch) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = p + U * L + u * L;
                ncol[u] = 0;
                if (live && q < hi) {
                    ncol[u] = A.ja[q];
                    nval[u] = P::load(A, q);
                }
            }
        } else if (p != lo + e) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = p + u * L;
                col[u] = 0;
                if (live && q < hi) {
                    col[u] = A.ja[q];
                    val[u] = P::load(A, q);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            pend[u] = live && p + u * L < hi;
            got[u] = P::kUnsolved;
        }
        while (true) {
            // every pending poll of the chunk goes out before the first is
            // examined, so a turn costs one round trip (a check right after
            // each load would make the compiler wait for that load alone)
            word w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                w[u] = P::kUnsolved;
                if (pend[u]) w[u] = poll_global(&gpub[col[u]]);
            }
            bool any = false;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (pend[u]) {
                    if (w[u] != P::kUnsolved) {
                        got[u] = w[u];
                        pend[u] = false;
                    } else if (expired) {
                        got[u] = P::kPoison;
                        pend[u] = false;
                    }
                    any |= pend[u];
                }
            }
            if (!__ballot(any)) break;
            if (A.backoff) __builtin_amdgcn_s_sleep(1);
            if (++spins == 512) {
                spins = 0;
                if (!expired && wall_clock64() - t0 > A.wd_ticks) {
                    expired = true;
                    if (lane == 0) atomicOr(A.err, 1);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (live && p + u * L < hi) P::fold(acc, val[u], got[u]);
        if constexpr (kPrefetch) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                col[u] = ncol[u];
                val[u] = nval[u];
            }
        }
    }
}

template <class P, int U, bool SPLIT = false>
__device__ __forceinline__ void fast_rows(const SolveArgs& A, int L, int lane, int row) {
    using word = typename P::word;
    using Acc = typename P::Acc;
    word* gpub = static_cast<word*>(A.pub);
    const int e = lane & (L - 1);
    const bool live = row >= 0;
    int rs = 0, dpos = 0;
    if (live) {
        rs = A.ia[row];
        dpos = A.ia[row + 1] - 1;
    }
    // a split long row (SPLIT: multirow walks with a segment table) folds its
    // first kSegOwn entries, then its segments' partial sums (lane e:
    // partials e, e + L, ...)
    int eend = dpos, seg0 = 0, nseg = 0;
    if constexpr (SPLIT) {
        if (A.items && live && dpos - rs > kSegMin) {
            const int g = A.row_seg[row];
            if (g >= 0) {
                seg0 = g;
                nseg = seg_count(dpos - rs);
                eend = rs + kSegOwn;
            }
        }
    }
    const size_t xi = live ? (size_t)row : 0;
    Acc acc = (live && e == 0) ? P::seed(A, caller_index(A, row, 0)) : P::zero();
    typename P::Diag dg{};
    if (live && e == 0) dg = P::diag(A, dpos);
    if (A.trace && live && e == 0) A.trace[3 * xi] = wall_clock64();
    const long long t0 = wall_clock64();
    bool expired = false;
    int spins = 0;
    fold_entries<P, U>(A, L, e, lane, live, rs, eend, acc, t0, expired, spins);
    for (int j = e; SPLIT && __ballot(j < nseg); j += L) {
        uint64_t pw[P::kPartWords];
        bool pend = j < nseg;
        while (true) {
            if (pend) {
                bool ok = true;
#pragma unroll
                for (int i = 0; i < P::kPartWords; ++i) {
                    pw[i] = poll_global(&A.part[2 * (size_t)(seg0 + j) + i]);
                    ok &= pw[i] != kUnsolved64;
                }
                if (!ok && expired) {
#pragma unroll
                    for (int i = 0; i < P::kPartWords; ++i) pw[i] = kPartPoison;
                    ok = true;
                }
                pend = !ok;
            }
            if (!__ballot(pend)) break;
            if (++spins == 512) {
                spins = 0;
                if (!expired && wall_clock64() - t0 > A.wd_ticks) {
                    expired = true;
                    if (lane == 0) atomicOr(A.err, 1);
                }
            }
        }
        if (j < nseg) acc = P::add(acc, P::part_unpack(pw));
    }
    const Acc t = group_reduce<P, 1>(acc, L);
    if (live && e == 0) {
        if (A.trace) A.trace[3 * xi + 1] = wall_clock64();
        typename P::Out o{};
        const word w = P::finish(A, dg, t, o);
        if (A.trace) A.trace[3 * xi + 2] = wall_clock64();
        publish_global(&gpub[xi], w);
        P::emit(A, caller_index(A, row, 0), o, w);
    }
}

// Segment item g of a split long row: the whole wave folds entries
// [seg_lo[g], seg_hi[g]) from zero and publishes the partial sum.
template <class P, int U>
__device__ __forceinline__ void fast_segment(const SolveArgs& A, int lane, int g) {
    using Acc = typename P::Acc;
    const long long t0 = wall_clock64();
    bool expired = false;
    int spins = 0;
    Acc acc = P::zero();
    fold_entries<P, U>(A, kWave, lane, lane, true, A.seg_lo[g], A.seg_hi[g], acc, t0, expired, spins);
    const Acc t = group_reduce<P, 1>(acc, kWave);
    if (lane == 0) {
        uint64_t w[P::kPartWords];
        P::part_pack(t, w);
#pragma unroll
        for (int i = 0; i < P::kPartWords; ++i) publish_global(&A.part[2 * (size_t)g + i], w[i]);
    }
}

// One row per wave, natural (ORDERED = false) or level order (true).
template <class P, int WG, int U, bool ORDERED>
__global__ __launch_bounds__(WG) void k_fast_row(SolveArgs A) {
    __shared__ int s_base;
    const int tid = threadIdx.x;
    if (tid == 0) s_base = atomicAdd(A.ticket, 1) * (WG / kWave);
    __syncthreads();
    const int pos = s_base + (tid >> 6);
    if (pos >= A.n) return;  // whole wave
    const int row = ORDERED ? A.iorder[pos] : pos;
    const int lane = tid & (kWave - 1);
    fast_rows<P, U>(A, A.L, lane, lane < A.L ? row : -1);
}

// Multirow waves (rows of one wave share a level).
template <class P, int WG, int U>
__global__ __launch_bounds__(WG) void k_fast_multirow(SolveArgs A) {
    __shared__ int s_t;
    const int tid = threadIdx.x;
    if (tid == 0) s_t = atomicAdd(A.ticket, 1);
    __syncthreads();
    const int t = s_t * (WG / kWave) + (tid >> 6);
    if (t >= multirow_items(A)) return;
    const int w = multirow_item(A, t);
    const int lane = tid & (kWave - 1);
    if (w < 0) {
        fast_segment<P, U>(A, lane, -w - 1);
        return;
    }
    const int vs = A.wsize[w];
    const int L = vs == 0 ? 1 : vs;
    const int b0 = A.wbase[w];
    const int nrows = A.wbase[w + 1] - b0;
    const int seg = lane / L;
    const int row = seg < nrows ? A.iorder[b0 + seg] : -1;
    fast_rows<P, U, true>(A, L, lane, row);
}

// ---- in-degree walk (one right-hand side) ------------------------------------------
// The column-oriented schedule of sync-free SpTRSV: every row keeps a count
// of its unsolved dependencies; a solved row decrements the count of each row
// that depends on it (its column's pattern), and a row starts folding once
// its count is zero.  A waiting row polls one word, its counter, instead of
// one per pending entry.  The values are still pulled in storage order
// through fold_entries (whose polls then succeed at once: a counter reaches
// zero only after every dependency stored its value, and a value read early
// shows the unsolved pattern and is read again), so the results are
// bit-identical to the lean walk with the same lanes and window -- no
// floating-point atomics, unlike the push-sum form of this schedule.  Rows
// are taken in ticket order, so every decrement a row waits for comes from a
// running wave.
template <class P, int U>
__device__ __forceinline__ void indeg_row(const SolveArgs& A, int L, int lane, int row) {
    using word = typename P::word;
    using Acc = typename P::Acc;
    word* gpub = static_cast<word*>(A.pub);
    const int e = lane & (L - 1);
    const bool live = lane < L;
    const int rs = A.ia[row], dpos = A.ia[row + 1] - 1;
    const size_t xi = (size_t)row;
    Acc acc = (live && e == 0) ? P::seed(A, xi) : P::zero();
    typename P::Diag dg{};
    if (live && e == 0) dg = P::diag(A, dpos);
    if (A.trace && live && e == 0) A.trace[3 * xi] = wall_clock64();
    const long long t0 = wall_clock64();
    bool expired = false;
    int spins = 0;
    // wait until every dependency has been solved (lane 0 polls for the wave)
    while (true) {
        int c = 0;
        if (lane == 0) c = poll_global(&A.indeg[row]);
        if (__shfl(c, 0) <= 0) break;
        if (++spins == 512) {
            spins = 0;
            if (wall_clock64() - t0 > A.wd_ticks) {
                expired = true;  // fold_entries then poisons what is missing
                if (lane == 0) atomicOr(A.err, 1);
                break;
            }
        }
    }
    fold_entries<P, U>(A, L, e, lane, live, rs, dpos, acc, t0, expired, spins);
    const Acc t = group_reduce<P, 1>(acc, L);
    if (live && e == 0) {
        if (A.trace) A.trace[3 * xi + 1] = wall_clock64();
        typename P::Out o{};
        const word w = P::finish(A, dg, t, o);
        if (A.trace) A.trace[3 * xi + 2] = wall_clock64();
        publish_global(&gpub[xi], w);
        P::emit(A, xi, o, w);
    }
    // hand the row to its dependents
    for (int q = A.dep_ptr[row] + lane; q < A.dep_ptr[row + 1]; q += kWave)
        __hip_atomic_fetch_add(&A.indeg[A.dep_row[q]], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One row per wave in natural (ORDERED = false) or level order.
template <class P, int WG, int U, bool ORDERED>
__global__ __launch_bounds__(WG) void k_indeg(SolveArgs A) {
    __shared__ int s_base;
    const int tid = threadIdx.x;
    if (tid == 0) s_base = atomicAdd(A.ticket, 1) * (WG / kWave);
    __syncthreads();
    const int pos = s_base + (tid >> 6);
    if (pos >= A.n) return;  // whole wave
    indeg_row<P, U>(A, A.L, tid & (kWave - 1), ORDERED ? A.iorder[pos] : pos);
}

// ---- persistent lean path --------------------------------------------------------
// A fixed number of workgroups (the residency cap) walks the work items:
// rows in natural order (MODE 0), rows in iorder (MODE 1) or multirow waves of
// the analysis (MODE 2).  A workgroup claims a block of B = waves * R
// consecutive items with one atomic on the ticket counter (blocks are handed
// out in increasing order); its waves take the block's items round-robin,
// each in increasing order, and a barrier closes the block before the next
// claim.  One claim per block rather than per item: same-address atomics
// serialise (about 11 ns each), which one claim per row would make the
// bound.
//
// Deadlock freedom, whatever the number of resident workgroups (other
// kernels may hold part of the device): every claimed block belongs to a
// workgroup that is running, and all of its waves are resident.  Let u be
// the lowest unfinished item; its block has been claimed (a workgroup claims
// only after finishing its previous block), and every item it depends on is
// lower, hence finished.  The wave that owns u in the block has finished its
// earlier items of the block (they are lower than u), so it is working on u,
// whose dependencies are all published: u completes.  Waves that are done
// with their share wait at the barrier holding no item.  So every item
// completes, and a workgroup leaves when its claim falls past the end.
template <int WG>
__device__ __forceinline__ int claim_block(int* ctr, int items_per_block, int* s_slot) {
    __syncthreads();  // the previous block is done; s_slot may be rewritten
    if (threadIdx.x == 0) *s_slot = atomicAdd(ctr, 1) * items_per_block;
    __syncthreads();
    return *s_slot;
}

template <class P, int WG, int U, int MODE>
__global__ __launch_bounds__(WG) void k_persist(SolveArgs A) {
    __shared__ int s_base;
    constexpr int NW = WG / kWave;
    const int wv = threadIdx.x >> 6;
    const int lane = threadIdx.x & (kWave - 1);
    const int count = (MODE == 2) ? multirow_items(A) : A.n;
    const int B = NW * max(1, A.claim_rounds);
    for (int b0 = claim_block<WG>(A.ticket, B, &s_base); b0 < count; b0 = claim_block<WG>(A.ticket, B, &s_base)) {
        for (int j = wv; j < B && b0 + j < count; j += NW) {
            const int t = b0 + j;
            if constexpr (MODE == 2) {
                const int w = multirow_item(A, t);
                if (w < 0) {
                    fast_segment<P, U>(A, lane, -w - 1);
                    continue;
                }
                const int vs = A.wsize[w];
                const int L = vs == 0 ? 1 : vs;
                const int wb = A.wbase[w];
                const int seg = lane / L;
                const int row = seg < A.wbase[w + 1] - wb ? A.iorder[wb + seg] : -1;
                fast_rows<P, U, true>(A, L, lane, row);
            } else {
                const int row = (MODE == 1) ? A.iorder[t] : t;
                fast_rows<P, U>(A, A.L, lane, lane < A.L ? row : -1);
            }
        }
    }
}

// Persistent general engine (any K): the same block-claimed walk for lane
// groups of L entry lanes x K right-hand-side lanes.  An item is the
// 64/(L*K) rows one wave holds, consecutive in natural order (MODE 0) or in
// iorder (MODE 1), or one multirow wave of the analysis (MODE 2, rows of one
// level, in passes of 64/(L*K) rows).  Rows of an item may depend on each
// other in MODE 0/1; sf_rows publishes inside its polling loop, so that is
// safe, and the deadlock argument is k_persist's.
template <class P, int K, int WG, int U, int MODE>
__global__ __launch_bounds__(WG) void k_persist_rows(SolveArgs A) {
    __shared__ int s_base;
    constexpr int NW = WG / kWave;
    const int wv = threadIdx.x >> 6;
    const int lane = threadIdx.x & (kWave - 1);
    const int B = NW * max(1, A.claim_rounds);
    if constexpr (MODE == 2) {
        for (int b0 = claim_block<WG>(A.ticket, B, &s_base); b0 < A.n_waves;
             b0 = claim_block<WG>(A.ticket, B, &s_base)) {
            for (int j = wv; j < B && b0 + j < A.n_waves; j += NW) {
                const int t = b0 + j;
                const int vs = A.wsize[t];
                int L = vs == 0 ? 1 : vs;
                if (L * K > kWave) L = kWave / K;
                const int wb = A.wbase[t];
                const int nrows = A.wbase[t + 1] - wb;
                const int cap = kWave / (L * K);
                const int seg = lane / (L * K);
                for (int c0 = 0; c0 < nrows; c0 += cap) {
                    const int row = (c0 + seg < nrows) ? A.iorder[wb + c0 + seg] : -1;
                    sf_rows<P, K, U>(A, L, lane, row, nullptr, 0, 0);
                }
            }
        }
    } else {
        const int lgG = A.lgL + kLog2<K>();
        const int rpw = kWave >> lgG;
        const int items = (A.n + rpw - 1) / rpw;
        for (int b0 = claim_block<WG>(A.ticket, B, &s_base); b0 < items; b0 = claim_block<WG>(A.ticket, B, &s_base)) {
            for (int j = wv; j < B && b0 + j < items; j += NW) {
                const int pos = (b0 + j) * rpw + (lane >> lgG);
                const int row = pos < A.n ? (MODE == 1 ? A.iorder[pos] : pos) : -1;
                sf_rows<P, K, U>(A, A.L, lane, row, nullptr, 0, 0);
            }
        }
    }
}

// Fill the published words with the unsolved pattern, the segment partials
// (if any) with all-ones, the in-degree counters (if any) with each row's
// off-diagonal count, and reset the ticket.
template <class T>
__global__ void k_prepare(T* pub, size_t count, T unsolved, int* ticket, uint64_t* part, size_t part_count,
                          const int* ia, int* indeg, int n_indeg) {
    const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    if (i0 == 0) *ticket = 0;
    for (size_t i = i0; i < count; i += stride) pub[i] = unsolved;
    for (size_t i = i0; i < part_count; i += stride) part[i] = kUnsolved64;
    for (size_t i = i0; i < (size_t)n_indeg; i += stride) indeg[i] = ia[i + 1] - ia[i] - 1;
}

// ---- host-side launch ----------------------------------------------------------
enum KernelKind {
    KK_SIMPLE = 1,
    KK_ORDER = 2,
    KK_MULTIROW = 3,
    KK_FAST_SIMPLE = 4,    // one row per wave, natural order (K == 1)
    KK_FAST_ORDER = 5,     // one row per wave, level order (K == 1)
    KK_FAST_MULTIROW = 6,  // multirow waves, no in-wave dependency (K == 1)
    KK_PERSIST_SIMPLE = 7,    // persistent walk over rows (lean for K == 1)
    KK_PERSIST_ORDER = 8,     // persistent walk over iorder (lean for K == 1)
    KK_PERSIST_MULTIROW = 9,  // persistent walk over multirow waves (lean for K == 1)
    KK_INDEG_SIMPLE = 10,     // in-degree walk, natural order (K == 1)
    KK_INDEG_ORDER = 11,      // in-degree walk, level order (K == 1)
};

// Launch of a persistent kernel with A.persist_grid workgroups, clamped to
// what can be resident at once (more would only start as others finish).  An
// ordinary launch: progress does not depend on all workgroups being resident.
template <class F>
hipError_t launch_persistent(F* kern, int wg, const SolveArgs& A, hipStream_t s) {
    const void* f = reinterpret_cast<const void*>(kern);
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, wg, A.dyn_lds);
    if (e != hipSuccess) return e;
    const long long cap = (long long)std::max(1, per_cu) * cus;
    const unsigned grid = (unsigned)std::max<long long>(1, std::min<long long>(A.persist_grid, cap));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(wg), A.dyn_lds, s, A);
    return hipGetLastError();
}

template <class P, int WG, int U, int MODE>
hipError_t launch_persist(const SolveArgs& A, hipStream_t s) {
    return launch_persistent(&k_persist<P, WG, U, MODE>, WG, A, s);
}

template <class P, int K, int WG, int U, int MODE>
hipError_t launch_persist_rows(const SolveArgs& A, hipStream_t s) {
    return launch_persistent(&k_persist_rows<P, K, WG, U, MODE>, WG, A, s);
}

template <class P, int K, int WG, int U>
hipError_t launch_k(int kind, const SolveArgs& A, hipStream_t s) {
    unsigned grid;
    if (kind == KK_MULTIROW || kind == KK_FAST_MULTIROW)
        grid = ceil_div(A.items ? A.n_items : A.n_waves, WG / kWave);
    else if (kind == KK_FAST_SIMPLE || kind == KK_FAST_ORDER || kind == KK_INDEG_SIMPLE || kind == KK_INDEG_ORDER)
        grid = ceil_div(A.n, WG / kWave);
    else
        grid = ceil_div(A.n, WG >> (A.lgL + kLog2<K>()));
    if (grid == 0) return hipSuccess;
    if constexpr (K == 1) {
        switch (kind) {
            case KK_FAST_SIMPLE:
                hipLaunchKernelGGL((k_fast_row<P, WG, U, false>), dim3(grid), dim3(WG), A.dyn_lds, s, A);
                return hipGetLastError();
            case KK_FAST_ORDER:
                hipLaunchKernelGGL((k_fast_row<P, WG, U, true>), dim3(grid), dim3(WG), A.dyn_lds, s, A);
                return hipGetLastError();
            case KK_FAST_MULTIROW:
                hipLaunchKernelGGL((k_fast_multirow<P, WG, U>), dim3(grid), dim3(WG), A.dyn_lds, s, A);
                return hipGetLastError();
            case KK_PERSIST_SIMPLE:
                return launch_persist<P, WG, U, 0>(A, s);
            case KK_PERSIST_ORDER:
                return launch_persist<P, WG, U, 1>(A, s);
            case KK_PERSIST_MULTIROW:
                return launch_persist<P, WG, U, 2>(A, s);
            case KK_INDEG_SIMPLE:
                hipLaunchKernelGGL((k_indeg<P, WG, U, false>), dim3(grid), dim3(WG), A.dyn_lds, s, A);
                return hipGetLastError();
            case KK_INDEG_ORDER:
                hipLaunchKernelGGL((k_indeg<P, WG, U, true>), dim3(grid), dim3(WG), A.dyn_lds, s, A);
                return hipGetLastError();
            default:
                break;
        }
    } else {
        switch (kind) {
            case KK_PERSIST_SIMPLE:
                return launch_persist_rows<P, K, WG, U, 0>(A, s);
            case KK_PERSIST_ORDER:
                return launch_persist_rows<P, K, WG, U, 1>(A, s);
            case KK_PERSIST_MULTIROW:
                return launch_persist_rows<P, K, WG, U, 2>(A, s);
            default:
                break;
        }
    }
    switch (kind) {
        case KK_SIMPLE:
            hipLaunchKernelGGL((k_simple<P, K, WG, U>), dim3(grid), dim3(WG), A.dyn_lds, s, A);
            break;
        case KK_ORDER:
            hipLaunchKernelGGL((k_order<P, K, WG, U>), dim3(grid), dim3(WG), A.dyn_lds, s, A);
            break;
        case KK_MULTIROW:
            hipLaunchKernelGGL((k_multirow<P, K, WG, U>), dim3(grid), dim3(WG), A.dyn_lds, s, A);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Persistent general engine with 1024-thread workgroups (several right-hand
// sides): 16 waves share each claimed block, 16 items per claim.
template <class P, int K, int U>
hipError_t launch_persist_wide(int kind, const SolveArgs& A, hipStream_t s) {
    switch (kind) {
        case KK_PERSIST_SIMPLE: return launch_persist_rows<P, K, 1024, U, 0>(A, s);
        case KK_PERSIST_ORDER: return launch_persist_rows<P, K, 1024, U, 1>(A, s);
        case KK_PERSIST_MULTIROW: return launch_persist_rows<P, K, 1024, U, 2>(A, s);
        default: return hipErrorInvalidValue;
    }
}

// Launch one solve pass.  K: right-hand-side lanes (1..64, power of two);
// with K == 1 the workgroup may be 64, 256 or 1024 threads and the entry
// window 1 or 4 (16 with 256 threads); with K > 1 the workgroup is 256
// threads (1024 for the persistent kinds too) and the window 1 or 4.
template <class P, int MAXK>
hipError_t launch_solve(int kind, int K, int wg, int u, const SolveArgs& A, hipStream_t s) {
    if (K == 1) {
        if (wg == 1024)
            return (u >= 4) ? launch_k<P, 1, 1024, 4>(kind, A, s) : launch_k<P, 1, 1024, 1>(kind, A, s);
        if (wg == 64)
            return (u >= 4) ? launch_k<P, 1, 64, 4>(kind, A, s) : launch_k<P, 1, 64, 1>(kind, A, s);
        if (u == 16) return launch_k<P, 1, 256, 16>(kind, A, s);
        return (u == 4) ? launch_k<P, 1, 256, 4>(kind, A, s) : launch_k<P, 1, 256, 1>(kind, A, s);
    }
    if constexpr (MAXK > 1) {
        const bool w4 = u >= 4;
        if (wg == 1024 && kind >= KK_PERSIST_SIMPLE) {
            switch (K) {
                case 2: return w4 ? launch_persist_wide<P, 2, 4>(kind, A, s) : launch_persist_wide<P, 2, 1>(kind, A, s);
                case 4: return w4 ? launch_persist_wide<P, 4, 4>(kind, A, s) : launch_persist_wide<P, 4, 1>(kind, A, s);
                case 8: return w4 ? launch_persist_wide<P, 8, 4>(kind, A, s) : launch_persist_wide<P, 8, 1>(kind, A, s);
                case 16: return w4 ? launch_persist_wide<P, 16, 4>(kind, A, s) : launch_persist_wide<P, 16, 1>(kind, A, s);
                case 32: return w4 ? launch_persist_wide<P, 32, 4>(kind, A, s) : launch_persist_wide<P, 32, 1>(kind, A, s);
                case 64: return w4 ? launch_persist_wide<P, 64, 4>(kind, A, s) : launch_persist_wide<P, 64, 1>(kind, A, s);
                default: break;
            }
        }
        switch (K) {
            case 2: return w4 ? launch_k<P, 2, 256, 4>(kind, A, s) : launch_k<P, 2, 256, 1>(kind, A, s);
            case 4: return w4 ? launch_k<P, 4, 256, 4>(kind, A, s) : launch_k<P, 4, 256, 1>(kind, A, s);
            case 8: return w4 ? launch_k<P, 8, 256, 4>(kind, A, s) : launch_k<P, 8, 256, 1>(kind, A, s);
            case 16: return w4 ? launch_k<P, 16, 256, 4>(kind, A, s) : launch_k<P, 16, 256, 1>(kind, A, s);
            case 32: return w4 ? launch_k<P, 32, 256, 4>(kind, A, s) : launch_k<P, 32, 256, 1>(kind, A, s);
            case 64: return w4 ? launch_k<P, 64, 256, 4>(kind, A, s) : launch_k<P, 64, 256, 1>(kind, A, s);
            default: break;
        }
    }
    return hipErrorInvalidValue;
}

extern template hipError_t launch_solve<PolF64, 64>(int, int, int, int, const SolveArgs&, hipStream_t);
extern template hipError_t launch_solve<PolF32, 64>(int, int, int, int, const SolveArgs&, hipStream_t);
extern template hipError_t launch_solve<PolMixed, 64>(int, int, int, int, const SolveArgs&, hipStream_t);
extern template hipError_t launch_solve<PolLevel, 1>(int, int, int, int, const SolveArgs&, hipStream_t);

}  // namespace sptrsv
