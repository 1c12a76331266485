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
};

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

// ---- precision / operation policies -----------------------------------------

// Plain FP64: x_i = (b_i - sum L_ij x_j) * (1 / L_ii), published into x.
struct PolF64 {
    using word = uint64_t;
    static constexpr word kUnsolved = kUnsolved64;
    static constexpr word kPoison = 0x7ff8000000000000ull;
    struct Acc { double s; };
    struct Val { double v; };
    __device__ static Acc zero() { return {0.0}; }
    __device__ static Acc seed(const SolveArgs& A, size_t bi) {
        return {static_cast<const double*>(A.b)[bi]};
    }
    __device__ static Val load(const SolveArgs& A, int p) { return {A.a64[p]}; }
    __device__ static void fold(Acc& c, const Val& v, word w) {
        c.s = __fma_rn(-v.v, __longlong_as_double((long long)w), c.s);
    }
    __device__ static Acc add(const Acc& x, const Acc& y) { return {x.s + y.s}; }
    __device__ static Acc shfl_xor(const Acc& c, int o) { return {__shfl_xor(c.s, o)}; }
    __device__ static word finish(const SolveArgs& A, int dpos, size_t xi, const Acc& t) {
        (void)xi;
        return canon64(t.s * (1.0 / A.a64[dpos]));
    }
};

// Plain FP32 on the FP32 copy of the values, published into x (float).
struct PolF32 {
    using word = uint32_t;
    static constexpr word kUnsolved = kUnsolved32;
    static constexpr word kPoison = 0x7fc00000u;
    struct Acc { float s; };
    struct Val { float v; };
    __device__ static Acc zero() { return {0.0f}; }
    __device__ static Acc seed(const SolveArgs& A, size_t bi) {
        return {static_cast<const float*>(A.b)[bi]};
    }
    __device__ static Val load(const SolveArgs& A, int p) { return {A.a32[p]}; }
    __device__ static void fold(Acc& c, const Val& v, word w) {
        c.s = __fmaf_rn(-v.v, __uint_as_float(w), c.s);
    }
    __device__ static Acc add(const Acc& x, const Acc& y) { return {x.s + y.s}; }
    __device__ static Acc shfl_xor(const Acc& c, int o) { return {__shfl_xor(c.s, o)}; }
    __device__ static word finish(const SolveArgs& A, int dpos, size_t xi, const Acc& t) {
        (void)xi;
        return canon32(t.s * (1.0f / A.a32[dpos]));
    }
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
    __device__ static Acc zero() { return {0.0f, 0.0f, 0.0}; }
    __device__ static Acc seed(const SolveArgs& A, size_t bi) {
        double b = static_cast<const double*>(A.b)[bi];
        return {(float)b, 0.0f, b};
    }
    __device__ static Val load(const SolveArgs& A, int p) { return {A.a32[p], A.a64[p]}; }
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
    __device__ static word finish(const SolveArgs& A, int dpos, size_t xi, const Acc& t) {
        float piv = 1.0f / A.a32[dpos];
        float x0 = t.x0 * piv;
        double r = __fma_rn(-A.a64[dpos], (double)x0, t.r);
        float d = ((float)r + t.d) * piv;
        A.x_out[xi] = (double)x0 + (double)d;
        uint32_t lo = canon32(x0);
        uint32_t hi = __float_as_uint(d);
        return ((uint64_t)hi << 32) | lo;  // never all-ones: lo is canonical
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
    __device__ static Acc zero() { return {0}; }
    __device__ static Acc seed(const SolveArgs&, size_t) { return {0}; }
    __device__ static Val load(const SolveArgs&, int) { return {}; }
    __device__ static void fold(Acc& c, const Val&, word w) { c.m = max(c.m, (int)w); }
    __device__ static Acc add(const Acc& x, const Acc& y) { return {max(x.m, y.m)}; }
    __device__ static Acc shfl_xor(const Acc& c, int o) { return {__shfl_xor(c.m, o)}; }
    __device__ static word finish(const SolveArgs&, int, size_t, const Acc& t) {
        return (word)(1 + t.m);
    }
};

// ---- the row engine -----------------------------------------------------------
// One call handles the lane group of `row` (row < 0: idle group) for the
// calling lane.  L = entry lanes per row (power of two), K = right-hand-side
// lanes (power of two, lanes with r >= A.k idle), L*K divides 64.
// s_pub/lds_lo/lds_hi: rows in [lds_lo, lds_hi) are also published to and
// read from LDS (natural order only; pass lds_hi = lds_lo to disable).
template <class P, int K, int U>
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
    const size_t xi = (size_t)(live ? row : 0) * A.k + r;
    Acc acc = (live && e == 0) ? P::seed(A, xi) : P::zero();

    int p = rs + e;     // first entry of the current window
    int kc = 0;         // entries of the window already folded
    bool done = !(live && p < dpos);
    bool published = !live;

    int col[U];
    Val val[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        int q = p + u * L;
        col[u] = 0;
        if (!done && q < dpos) {
            col[u] = A.ja[q];
            val[u] = P::load(A, q);
        }
    }

    const unsigned long long gmask =
        (G == 64) ? ~0ull : (((1ull << G) - 1ull) << ((lane / G) * G));
    const long long t0 = wall_clock64();
    bool expired = false;
    int spins = 0;

    while (true) {
        if (!done) {
            word w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int q = p + u * L;
                w[u] = P::kUnsolved;
                if (u >= kc && q < dpos) {
                    int c = col[u];
                    if (c >= lds_lo && c < lds_hi)
                        w[u] = poll_lds(&s_pub[c - lds_lo]);
                    else
                        w[u] = poll_global(&gpub[(size_t)c * A.k + r]);
                    if (expired && w[u] == P::kUnsolved) w[u] = P::kPoison;
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int q = p + u * L;
                if (u == kc && q < dpos && w[u] != P::kUnsolved) {
                    P::fold(acc, val[u], w[u]);
                    ++kc;
                }
            }
            if (p + kc * L >= dpos) {
                done = true;
            } else if (kc == U) {
                p += U * L;
                kc = 0;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    int q = p + u * L;
                    if (q < dpos) {
                        col[u] = A.ja[q];
                        val[u] = P::load(A, q);
                    }
                }
            }
        }
        const unsigned long long dm = __ballot(done);
        const bool ready = ((dm & gmask) == gmask) && !published;
        if (__ballot(ready)) {
            Acc t = acc;
            for (int o = K; o < G; o <<= 1) t = P::add(t, P::shfl_xor(t, o));
            if (ready) {
                if (e == 0) {
                    word wv = P::finish(A, dpos, xi, t);
                    if (row >= lds_lo && row < lds_hi) publish_lds(&s_pub[row - lds_lo], wv);
                    publish_global(&gpub[xi], wv);
                }
                published = true;
            }
        }
        if (__ballot(!published) == 0ull) break;
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

// (level, class) packed waves (multirow): wave w solves the rows at iorder
// positions [wbase[w], wbase[w+1]) with wsize[w] lanes per row.  Rows of one
// wave share a level, so they never depend on each other.
template <class P, int K, int WG, int U>
__global__ __launch_bounds__(WG) void k_multirow(SolveArgs A) {
    __shared__ int s_t;
    const int tid = threadIdx.x;
    if (tid == 0) s_t = atomicAdd(A.ticket, 1);
    __syncthreads();
    const int w = s_t * (WG / kWave) + (tid >> 6);
    if (w >= A.n_waves) return;
    const int lane = tid & (kWave - 1);
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
        sf_rows<P, K, U>(A, L, lane, row, nullptr, 0, 0);
    }
}

// Fill the published words with the unsolved pattern and reset the ticket.
template <class T>
__global__ void k_prepare(T* pub, size_t count, T unsolved, int* ticket) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *ticket = 0;
    for (; i < count; i += (size_t)gridDim.x * blockDim.x) pub[i] = unsolved;
}

// ---- host-side launch ----------------------------------------------------------
enum KernelKind { KK_SIMPLE = 1, KK_ORDER = 2, KK_MULTIROW = 3 };

template <class P, int K, int WG, int U>
hipError_t launch_k(int kind, const SolveArgs& A, hipStream_t s) {
    unsigned grid;
    if (kind == KK_MULTIROW)
        grid = ceil_div(A.n_waves, WG / kWave);
    else
        grid = ceil_div(A.n, WG >> (A.lgL + kLog2<K>()));
    if (grid == 0) return hipSuccess;
    switch (kind) {
        case KK_SIMPLE:
            hipLaunchKernelGGL((k_simple<P, K, WG, U>), dim3(grid), dim3(WG), 0, s, A);
            break;
        case KK_ORDER:
            hipLaunchKernelGGL((k_order<P, K, WG, U>), dim3(grid), dim3(WG), 0, s, A);
            break;
        case KK_MULTIROW:
            hipLaunchKernelGGL((k_multirow<P, K, WG, U>), dim3(grid), dim3(WG), 0, s, A);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Launch one solve pass.  K: right-hand-side lanes (1..64, power of two);
// with K == 1 the workgroup may be 256 or 1024 threads and the entry window
// 1 or 4; with K > 1 both are fixed at 256 and 1.
template <class P, int MAXK>
hipError_t launch_solve(int kind, int K, int wg, int u, const SolveArgs& A, hipStream_t s) {
    if (K == 1) {
        if (wg == 1024)
            return (u == 4) ? launch_k<P, 1, 1024, 4>(kind, A, s) : launch_k<P, 1, 1024, 1>(kind, A, s);
        return (u == 4) ? launch_k<P, 1, 256, 4>(kind, A, s) : launch_k<P, 1, 256, 1>(kind, A, s);
    }
    if constexpr (MAXK > 1) {
        switch (K) {
            case 2: return launch_k<P, 2, 256, 1>(kind, A, s);
            case 4: return launch_k<P, 4, 256, 1>(kind, A, s);
            case 8: return launch_k<P, 8, 256, 1>(kind, A, s);
            case 16: return launch_k<P, 16, 256, 1>(kind, A, s);
            case 32: return launch_k<P, 32, 256, 1>(kind, A, s);
            case 64: return launch_k<P, 64, 256, 1>(kind, A, s);
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
