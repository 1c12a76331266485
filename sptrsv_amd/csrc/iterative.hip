// iterative.hip -- FP32 Jacobi sweeps as the bulk solver of FP64 iterative
// refinement, certified on the device.
//
// Published method: iterative sparse triangular solves (Anzt, Chow, Dongarra,
// "Iterative Sparse Triangular Solves for Preconditioning", Euro-Par 2015;
// the scheme behind rocSPARSE csritsv).  With L = D + N (N strictly lower),
// the Jacobi iteration d <- D^-1 (r - N d) is exact after nlevs sweeps (N is
// nilpotent) and, for diagonally dominant L, accurate to FP32 after a few
// tens of sweeps regardless of the depth of the dependency graph.  Each sweep
// is a bandwidth-bound pass with no inter-row synchronisation, so deep
// systems trade a latency-bound walk of thousands of levels for a few tens of
// HBM-bound passes.
//
// One refinement step (x in FP64):
//   r = b - L x (FP64), rs = (float)(D^-1 r) = d_0             [k_jac_resid / k_jac_first]
//   d_s = rs - N' d_{s-1}, N' = D^-1 N in FP32, s = 1 .. m-1      [k_jac_sweep]
//   x += d_{m-1} (FP64), fused into the last sweep
// (the row scaling by D^-1 is applied once to the stored FP32 values, so a
// sweep needs no diagonal)
// Certificate: with delta = min_i(|L_ii| - sum_j |L_ij|) > 0, |L^-1|_inf <=
// 1/delta, so |x - x*|_inf <= |b - L x|_inf / delta.  A final residual pass
// takes max |r| and max |x| per independent system (one, or each member of a
// batch, and each right-hand-side column) and k_jac_certify compares the
// bound with tol * |x|_inf, leaving the verdict in a device word (no host
// round trip) and clearing the norms for the next certificate.
//
// Kernel layout (both passes): rows are cut into tiles by nnz windows -- tile
// t holds the rows whose first entry lies in [t*W, (t+1)*W), W = 1024/K
// entries for K right-hand sides -- so every tile streams about the same
// number of entries.  A 256-thread workgroup stages its tile's row pointers
// and right-hand side (and, for the residual, the diagonal) in LDS, streams
// the entries with every lane busy and all loads of a lane issued before
// their first use (phase 1: a product per (entry, rhs) into LDS), then sums
// rows out of LDS with a lane group per row sized to the tile's mean row
// length (phase 2).  A tile whose products do not fit LDS (a very long row)
// sums straight from memory with a wavefront per row.
#include <math.h>

#include <algorithm>
#include <vector>

#include "syncfree.hpp"

namespace sptrsv {

constexpr int kTileThreads = 256;
constexpr int kNormReps = 16;    // replicas of each member's norm words (spread atomics)
constexpr int kNormStride = 16;  // 64-bit words per replica (one 128-byte line)

// A tile spans an nnz window of TW / K entries for K right-hand sides (TW
// products); LDS holds up to 5/4 of that, so a tile whose last row runs past
// its window still fits unless that row is very long.
template <int TW>
__host__ __device__ constexpr int tile_cap() {
    return TW + TW / 4;
}

struct TileArgs {
    int n;
    int k;
    const int* ia;
    const int* ja;
    const float* sval;        // a_ij / a_ii in FP32 (0 on the diagonal)
    const double* a64;
    const double* diag;       // a_ii, contiguous (read by the residual and the first iterate)
    const int4* tiles;        // per tile {first row, end row, first entry, end entry}
    const int* tile_member;   // member (independent system) of each tile; null: one member
    // column-sorted tiles (sweep only): each tile's entries ordered by column,
    // with the entry's position inside its tile; null: storage order.  With
    // cbits > 0 the position is packed above the column's cbits bits of sja
    // and sslot is unused.
    const int* sja;
    const float* ssv;
    const unsigned short* sslot;
    int cbits;
    int rows_max;  // most rows in one tile of the table (sizes the sweep's LDS)
};

struct NormArgs {
    // null: no norms; else one slot per (member, right-hand-side column),
    // slot = member * k + column, laid out [slot][rep][stride]: [0] |r|, [1] |x|
    unsigned long long* norms;
    int members;
};

template <class T>
__device__ __forceinline__ T jac_group_sum(T v, int G, int K) {
    if (K == 1) {
        int s = 0;
        for (int o = 1; o < G; o <<= 1, ++s) v = lane_combine(v, s, OpAdd{});
    } else {
        for (int o = K; o < G; o <<= 1) v += __shfl_xor(v, o);
    }
    return v;
}

// Lane-group shape for phase 2: G2 lanes per row and right-hand side, the
// smallest power of two for which a lane sums at most `epl` entries of a
// row of the tile's mean length (at most 64 / K lanes).
template <int K>
__device__ __forceinline__ int group_lanes(int E, int rows, int epl) {
    const int avg = (E + rows - 1) / rows;
    int g = 1;
    while (g * epl < avg && g * K < kWave) g <<= 1;
    return g;
}

// Tile of workgroup b.  Workgroups are dispatched to the 8 XCDs round-robin
// (b mod 8), each with its own L2; this bijection hands XCD x one contiguous
// eighth of the tiles, so neighbouring tiles -- which gather the same
// stretch of the iterate -- share that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
    const int q = nblocks >> 3, r = nblocks & 7;
    const int x = b & 7, i = b >> 3;
    return x * q + min(x, r) + i;
}

// Stages one tile in LDS with every independent load issued before the
// first use: level 1 the tile record; level 2 the row pointers, the per-row
// right-hand side (V; skipped when rhs is null) and the entries; level 3 the
// gathered iterate; then all LDS stores.  Products a * v[ja] (diagonal
// included) go to prod when the tile fits.
// SORTED: 0 storage order, 1 column-sorted with positions in sslot, 2
// column-sorted with positions packed above the column's cbits bits (the
// word is kept whole until the gather needs its column).
template <int K, int TW, class V, class P, int SORTED = 0>
struct TileStage {
    static constexpr int W = TW / K;                       // entries per window = max rows per tile
    static constexpr int CAP = tile_cap<TW>();
    static constexpr int RU = (W + 1 + kTileThreads - 1) / kTileThreads;
    static constexpr int SU = (TW + kTileThreads - 1) / kTileThreads;
    static constexpr int U = CAP / kTileThreads;
    int r0, rows, p0, E;
    bool lds;

    __device__ bool run(const TileArgs& A, int tile, const V* rhs, const P* vals, const P* v, int* s_ia,
                        V* s_rhs, P* prod) {
        const int tid = threadIdx.x;
        const int k = A.k;
        const int4 tl = A.tiles[tile];
        r0 = tl.x;
        rows = tl.y - tl.x;
        p0 = tl.z;
        E = tl.w - tl.z;
        if (rows <= 0) return false;
        lds = E * K <= CAP;
        // Load order: the per-row loads (row pointers, right-hand side),
        // guarded by the tile's row count, then the entries and the
        // gathers, unconditional with their indices clamped inside the tile
        // (a slot past the tile rereads its last entry and is never stored).
        // With every guarded load issued before the entries, the compiler
        // counts the loads outstanding after each entry exactly, so a gather
        // waits only for its own entry; a guard between the entry loads and
        // the gathers made it wait conservatively and serialised the gathers.
        // level 2
        int iav[RU];
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            iav[u] = 1;
            if (u * kTileThreads <= rows) iav[u] = A.ia[r0 + min(tid + u * kTileThreads, rows)];
        }
        V rv[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            rv[u] = V(0);
            if (rhs && u * kTileThreads < rows * K) {
                const int t = tid + u * kTileThreads;
                const int lr = min(t / K, rows - 1), rr = min(t & (K - 1), k - 1);
                rv[u] = rhs[(size_t)(r0 + lr) * k + rr];
            }
        }
        int jv[U] = {}, slot[U] = {};
        P av[U] = {};
        if (lds) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = min((tid + u * kTileThreads) / K, E - 1);
                if (SORTED) {
                    jv[u] = A.sja[p0 + e];
                    av[u] = (P)A.ssv[p0 + e];
                    if (SORTED == 1) slot[u] = A.sslot[p0 + e];
                } else {
                    jv[u] = A.ja[p0 + e];
                    av[u] = vals[p0 + e];
                }
            }
        }
        // level 3
        P xv[U] = {};
        if (lds) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int rr = min((tid + u * kTileThreads) & (K - 1), k - 1);
                const int col = (SORTED == 2) ? (int)((unsigned)jv[u] & ((1u << A.cbits) - 1u)) : jv[u];
                xv[u] = v[(size_t)col * k + rr];
            }
        }
        // LDS
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const int t = tid + u * kTileThreads;
            if (t <= rows) s_ia[t] = iav[u] - p0;
        }
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const int t = tid + u * kTileThreads;
            if (rhs && t < rows * K) s_rhs[t] = (t & (K - 1)) < k ? rv[u] : V(0);
        }
        if (lds) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = tid + u * kTileThreads;
                const int sl = (SORTED == 2) ? (int)((unsigned)jv[u] >> A.cbits) : slot[u];
                if (t < E * K) prod[SORTED ? sl * K + (t & (K - 1)) : t] = av[u] * xv[u];
            }
        }
        return true;
    }
};

// One Jacobi sweep on the row-scaled system: d_new = rs - N' d_old, with
// N' = D^-1 N stored as sval and rs = D^-1 r.  mode 0 stores d_new, mode 1
// sets x = d_new (first refinement step, x was 0), mode 2 adds x += d_new.
template <int K, int TW, int SORTED>
__device__ __forceinline__ void sweep_tile(const TileArgs& A, int tile, const float* __restrict__ rs,
                                           const float* __restrict__ d_old, float* __restrict__ d_new, double* x,
                                           int mode, int epl, float* prod, int* s_ia, float* s_r) {
    using Stage = TileStage<K, TW, float, float, SORTED>;
    Stage T;
    if (!T.run(A, tile, rs, A.sval, d_old, s_ia, s_r, prod)) return;
    __syncthreads();
    const int tid = threadIdx.x;
    const int k = A.k;
    // phase 2: a group of G2*K lanes per row
    const int G2 = group_lanes<K>(T.E, T.rows, epl);
    const int GL = G2 * K;
    const int rpp = kTileThreads / GL;
    const int npass = (T.rows + rpp - 1) / rpp;
    const int gl = tid % GL, e = gl / K, rr = gl & (K - 1);
    for (int ps = 0; ps < npass; ++ps) {
        const int lr = ps * rpp + tid / GL;
        const bool live = lr < T.rows && rr < k;
        float s = 0.0f;
        if (live) {
            const int rs0 = s_ia[lr], re = s_ia[lr + 1] - 1;  // diagonal excluded
            if (T.lds) {
                for (int q = rs0 + e; q < re; q += G2) s += prod[q * K + rr];
            } else {
#pragma unroll 4
                for (int q = rs0 + e; q < re; q += G2)
                    s = __fmaf_rn(A.sval[T.p0 + q], d_old[(size_t)A.ja[T.p0 + q] * k + rr], s);
            }
        }
        s = jac_group_sum(s, GL, K);
        if (live && e == 0) {
            const size_t i = (size_t)(T.r0 + lr) * k + rr;
            const float d = s_r[lr * K + rr] - s;
            if (mode == 0)
                d_new[i] = d;
            else if (mode == 1)
                x[i] = (double)d;
            else
                x[i] += (double)d;
        }
    }
}

// One Jacobi sweep on the row-scaled system: d_new = rs - N' d_old, with
// N' = D^-1 N stored as sval and rs = D^-1 r.  mode 0 stores d_new, mode 1
// sets x = d_new (first refinement step, x was 0), mode 2 adds x += d_new.
// One workgroup per tile.  Dynamic LDS (sweep_lds_bytes): the products, then
// row pointers and right-hand sides for the table's longest tile, not for
// the worst case of TW / K one-entry rows, so typical tables fit more
// workgroups per CU.
extern __shared__ float jac_sweep_lds[];

template <int K, int TW>
__host__ __device__ constexpr int sweep_rows_cap(int rows_max) {
    return (rows_max > 0 && rows_max < TW / K) ? rows_max : TW / K;
}

template <int K, int TW>
size_t sweep_lds_bytes(int rows_max) {
    const int r = sweep_rows_cap<K, TW>(rows_max);
    return sizeof(float) * ((size_t)tile_cap<TW>() + (size_t)r + 1 + (size_t)r * K);
}

template <int K, int TW, int SORTED>
__global__ __launch_bounds__(kTileThreads) void k_jac_sweep(TileArgs A, const float* __restrict__ rs,
                                                            const float* __restrict__ d_old,
                                                            float* __restrict__ d_new, double* x, int mode,
                                                            int epl) {
    using Stage = TileStage<K, TW, float, float, SORTED>;
    const int r = sweep_rows_cap<K, TW>(A.rows_max);
    float* prod = jac_sweep_lds;
    int* s_ia = reinterpret_cast<int*>(jac_sweep_lds + Stage::CAP);
    float* s_r = jac_sweep_lds + Stage::CAP + r + 1;
    sweep_tile<K, TW, SORTED>(A, xcd_tile(blockIdx.x, gridDim.x), rs, d_old, d_new, x, mode, epl, prod, s_ia, s_r);
}

// Block sweep (one right-hand side): the coupling to earlier tiles is taken
// from the previous iterate once, and the tile's own diagonal block gets
// `inner` Jacobi iterations in LDS:
//   base_i = rs_i - sum_{j < r0} N'_ij d_old_j
//   z^0 = d_old on the tile, z^t_i = base_i - sum_{r0 <= j < i} N'_ij z^{t-1}_j
//   d_new = z^inner
// One inner iteration is the point sweep restricted to the tile; with as
// many inner iterations as the tile has rows it is the exact block solve.
// Banded and locally coupled systems need a fraction of the point sweeps,
// and an inner iteration costs LDS traffic only.  A tile too large for LDS
// takes the point update.  Dynamic LDS (bsweep_lds_bytes): entry values and
// columns for CAP entries, then row pointers, base and the two iterates for
// the table's longest tile.
template <int TW>
size_t bsweep_lds_bytes(int rows_max) {
    const int r = sweep_rows_cap<1, TW>(rows_max);
    return sizeof(float) * (2 * (size_t)tile_cap<TW>() + (size_t)r + 1 + 3 * (size_t)r);
}

template <int TW>
__global__ __launch_bounds__(kTileThreads) void k_jac_bsweep(TileArgs A, const float* __restrict__ rs,
                                                             const float* __restrict__ d_old,
                                                             float* __restrict__ d_new, double* x, int mode,
                                                             int epl, int inner) {
    constexpr int W = TW;                   // max rows of a tile (one entry per row at least)
    constexpr int CAP = tile_cap<TW>();
    constexpr int U = CAP / kTileThreads;
    constexpr int RU = (W + 1 + kTileThreads - 1) / kTileThreads;
    const int rcap = sweep_rows_cap<1, TW>(A.rows_max);
    float* const s_v = jac_sweep_lds;                                   // off-tile product, or in-tile value
    int* const s_c = reinterpret_cast<int*>(jac_sweep_lds + CAP);       // -1 off-tile, else tile-local column
    int* const s_ia = reinterpret_cast<int*>(jac_sweep_lds + 2 * CAP);  // rcap + 1
    float* const s_b = jac_sweep_lds + 2 * CAP + rcap + 1;              // base
    float* const s_z0 = s_b + rcap;                                     // inner iterates
    float* const s_z1 = s_z0 + rcap;
    const int tid = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int4 tl = A.tiles[tile];
    const int r0 = tl.x, rows = tl.y - tl.x, p0 = tl.z, E = tl.w - tl.z;
    if (rows <= 0) return;
    if (E > CAP) {  // point update straight from memory, a wavefront per row
        const int lane = tid & (kWave - 1);
        for (int lr = tid >> 6; lr < rows; lr += kTileThreads / kWave) {
            const int i = r0 + lr;
            const int q0 = A.ia[i], q1 = A.ia[i + 1] - 1;
            float sm = 0.0f;
            for (int q = q0 + lane; q < q1; q += kWave) sm = __fmaf_rn(A.sval[q], d_old[A.ja[q]], sm);
            sm = jac_group_sum(sm, kWave, 1);
            if (lane == 0) {
                const float d = rs[i] - sm;
                if (mode == 0) d_new[i] = d;
                else if (mode == 1) x[i] = (double)d;
                else x[i] += (double)d;
            }
        }
        return;
    }
    // staging: every load issued before its first use.  The per-row loads,
    // guarded by the row count, come first; the entry loads and the gathers
    // are unconditional (indices clamped inside the tile; an in-tile column
    // is gathered too and not used), so each gather waits only for its own
    // entry (the reasoning of TileStage).
    int iav[RU];
    float rv[RU], zv[RU];
#pragma unroll
    for (int u = 0; u < RU; ++u) {
        iav[u] = 0;
        rv[u] = zv[u] = 0.0f;
        if (u * kTileThreads <= rows) iav[u] = A.ia[r0 + min(tid + u * kTileThreads, rows)];
        if (u * kTileThreads < rows) {
            const int lr = min(tid + u * kTileThreads, rows - 1);
            rv[u] = rs[r0 + lr];
            zv[u] = d_old[r0 + lr];
        }
    }
    int jv[U];
    float av[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int e = min(tid + u * kTileThreads, E - 1);
        jv[u] = A.ja[p0 + e];
        av[u] = A.sval[p0 + e];
    }
    float xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = d_old[jv[u]];
#pragma unroll
    for (int u = 0; u < RU; ++u) {
        const int t = tid + u * kTileThreads;
        if (t <= rows) s_ia[t] = iav[u] - p0;
        if (t < rows) {
            s_b[t] = rv[u];
            s_z0[t] = zv[u];
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int t = tid + u * kTileThreads;
        if (t < E) {
            const bool off = jv[u] < r0;
            s_v[t] = off ? av[u] * xv[u] : av[u];
            s_c[t] = off ? -1 : jv[u] - r0;
        }
    }
    __syncthreads();
    const int G2 = group_lanes<1>(E, rows, epl);
    const int rpp = kTileThreads / G2;
    const int npass = (rows + rpp - 1) / rpp;
    const int e = tid % G2;
    // inner iterations on the diagonal block.  With one pass of rows per
    // workgroup (the usual case) each lane keeps its first EC in-tile
    // entries in registers across the iterations, so an iteration costs one
    // round of independent LDS gathers.
    constexpr int EC = 8;
    int cur = 0;
    if (npass == 1 && inner > 0) {
        // one pass over the tile gives both the base (the off-tile sums)
        // and the first inner iterate z^1 = base - N'_in z^0, each summed in
        // the order of the separate passes (so the results are the same)
        const int lr = tid / G2;
        const bool live = lr < rows;
        const int q0 = live ? s_ia[lr] + e : 0, q1 = live ? s_ia[lr + 1] - 1 : 0;
        int cc[EC];
        float cv[EC];
        float so = 0.0f, si = 0.0f;
#pragma unroll
        for (int u = 0; u < EC; ++u) {
            const int q = q0 + u * G2;
            const int c = (q < q1) ? s_c[q] : -1;
            const float v = (q < q1) ? s_v[q] : 0.0f;
            if (q < q1 && c < 0) so += v;
            cc[u] = c;
            cv[u] = (c >= 0) ? v : 0.0f;
            si = __fmaf_rn(cv[u], s_z0[max(c, 0)], si);
        }
        const int qrest = q0 + EC * G2;
        for (int q = qrest; q < q1; q += G2) {
            const int c = s_c[q];
            if (c < 0)
                so += s_v[q];
            else
                si = __fmaf_rn(s_v[q], s_z0[c], si);
        }
        so = jac_group_sum(so, G2, 1);
        si = jac_group_sum(si, G2, 1);
        if (live && e == 0) {
            const float base = s_b[lr] - so;
            s_b[lr] = base;
            s_z1[lr] = base - si;
        }
        cur = 1;
        for (int it = 1; it < inner; ++it) {
            if (it > 0) __syncthreads();
            const float* zc = cur ? s_z1 : s_z0;
            float sm = 0.0f;
#pragma unroll
            for (int u = 0; u < EC; ++u) sm = __fmaf_rn(cv[u], zc[max(cc[u], 0)], sm);
            for (int q = qrest; q < q1; q += G2) {
                const int c = s_c[q];
                if (c >= 0) sm = __fmaf_rn(s_v[q], zc[c], sm);
            }
            sm = jac_group_sum(sm, G2, 1);
            if (live && e == 0) (cur ? s_z0 : s_z1)[lr] = s_b[lr] - sm;
            cur ^= 1;
        }
    } else {
        // base: the off-tile sums
        for (int ps = 0; ps < npass; ++ps) {
            const int lr = ps * rpp + tid / G2;
            const bool live = lr < rows;
            float sm = 0.0f;
            if (live) {
                const int q1 = s_ia[lr + 1] - 1;  // diagonal excluded
                for (int q = s_ia[lr] + e; q < q1; q += G2)
                    if (s_c[q] < 0) sm += s_v[q];
            }
            sm = jac_group_sum(sm, G2, 1);
            if (live && e == 0) s_b[lr] -= sm;
        }
        for (int it = 0; it < inner; ++it) {
            __syncthreads();
            const float* zc = cur ? s_z1 : s_z0;
            float* zn = cur ? s_z0 : s_z1;
            for (int ps = 0; ps < npass; ++ps) {
                const int lr = ps * rpp + tid / G2;
                const bool live = lr < rows;
                float sm = 0.0f;
                if (live) {
                    const int q1 = s_ia[lr + 1] - 1;
                    for (int q = s_ia[lr] + e; q < q1; q += G2) {
                        const int c = s_c[q];
                        if (c >= 0) sm = __fmaf_rn(s_v[q], zc[c], sm);
                    }
                }
                sm = jac_group_sum(sm, G2, 1);
                if (live && e == 0) zn[lr] = s_b[lr] - sm;
            }
            cur ^= 1;
        }
    }
    __syncthreads();
    const float* zf = cur ? s_z1 : s_z0;
    for (int lr = tid; lr < rows; lr += kTileThreads) {
        const int i = r0 + lr;
        const float d = inner > 0 ? zf[lr] : s_b[lr];
        if (mode == 0) d_new[i] = d;
        else if (mode == 1) x[i] = (double)d;
        else x[i] += (double)d;
    }
}

__device__ __forceinline__ unsigned long long* norm_slot(const NormArgs& N, int member, int word) {
    return N.norms + ((size_t)member * kNormReps + blockIdx.x % kNormReps) * kNormStride + word;
}

// r = b - L x in FP64 (diagonal included).  Writes rs = (float)(r / a_ii)
// when rs_out is set, and with N.norms folds max |r| and max |x| per member.
// Only the products and row pointers are staged in LDS (14 KB per
// workgroup, which lets 8 workgroups share a CU); the lane that finishes a
// row loads its b, a_ii and x from memory before summing the row.
template <int K>
__global__ __launch_bounds__(kTileThreads) void k_jac_resid(TileArgs A, const double* __restrict__ b,
                                                            const double* __restrict__ x,
                                                            float* __restrict__ rs_out, NormArgs N, int epl) {
    constexpr int TW = 1024;
    using Stage = TileStage<K, TW, double, double>;
    __shared__ double prod[Stage::CAP];
    __shared__ int s_ia[Stage::W + 1];
    __shared__ unsigned long long s_nrm[2 * K];  // per column: max |r|, max |x|
    Stage T;
    if (!T.run(A, xcd_tile(blockIdx.x, gridDim.x), nullptr, A.a64, x, s_ia, nullptr, prod)) return;
    const int tid = threadIdx.x;
    const int k = A.k;
    if (N.norms && tid < 2 * K) s_nrm[tid] = 0ull;
    __syncthreads();
    const int G2 = group_lanes<K>(T.E, T.rows, epl);
    const int GL = G2 * K;
    const int rpp = kTileThreads / GL;
    const int npass = (T.rows + rpp - 1) / rpp;
    const int gl = tid % GL, e = gl / K, rr = gl & (K - 1);
    // max |r| and max |x| of this thread's rows (all in column rr), as the
    // bit patterns of non-negative doubles (ordered like the values)
    unsigned long long mr = 0ull, mx = 0ull;
    for (int ps = 0; ps < npass; ++ps) {
        const int lr = ps * rpp + tid / GL;
        const bool live = lr < T.rows && rr < k;
        const bool fin = live && e == 0;
        const size_t i = (size_t)(T.r0 + min(lr, T.rows - 1)) * k + min(rr, k - 1);
        // the finishing lane's per-row operands, in flight during the sum
        double bv = 0.0, dv = 1.0, xv = 0.0;
        if (fin) {
            bv = b[i];
            if (rs_out) dv = A.diag[T.r0 + lr];
            if (N.norms) xv = x[i];
        }
        double s = 0.0;
        if (live) {
            const int rs0 = s_ia[lr], re = s_ia[lr + 1];
            if (T.lds) {
                for (int q = rs0 + e; q < re; q += G2) s += prod[q * K + rr];
            } else {
#pragma unroll 4
                for (int q = rs0 + e; q < re; q += G2)
                    s = __fma_rn(A.a64[T.p0 + q], x[(size_t)A.ja[T.p0 + q] * k + rr], s);
            }
        }
        s = jac_group_sum(s, GL, K);
        if (fin) {
            const double r = bv - s;
            if (rs_out) rs_out[i] = (float)(r / dv);
            if (N.norms) {
                mr = max(mr, (unsigned long long)__double_as_longlong(fabs(r)));
                mx = max(mx, (unsigned long long)__double_as_longlong(fabs(xv)));
            }
        }
    }
    if (N.norms) {
        // Each column is certified on its own: a thread's column is
        // rr = tid & (K - 1), so lanes of one column are K apart.  Wave
        // maxima per column first (one LDS atomic per wave and column
        // instead of one per row; same-address atomics serialise), then one
        // global atomic per column and word.
        for (int o = kWave / 2; o >= K; o >>= 1) {
            mr = max(mr, (unsigned long long)__shfl_xor((long long)mr, o));
            mx = max(mx, (unsigned long long)__shfl_xor((long long)mx, o));
        }
        const int wl = tid & (kWave - 1);
        if (wl < K && wl < k) {  // here rr == wl
            if (mr) atomicMax(&s_nrm[2 * wl], mr);
            if (mx) atomicMax(&s_nrm[2 * wl + 1], mx);
        }
        __syncthreads();
        if (tid < 2 * K && (tid >> 1) < k && s_nrm[tid] != 0ull) {
            const int m = A.tile_member ? A.tile_member[xcd_tile(blockIdx.x, gridDim.x)] : 0;
            atomicMax(norm_slot(N, m * k + (tid >> 1), tid & 1), s_nrm[tid]);
        }
    }
}

// First refinement step (x = 0): rs = (float)(b / a_ii) on the rows of the
// tiles (one workgroup per tile; the rows of a batch's other members are
// left alone), reading the contiguous diagonal copy.
template <int K>
__global__ __launch_bounds__(256) void k_jac_first(TileArgs A, const double* __restrict__ b, float* rs) {
    const int4 tl = A.tiles[blockIdx.x];
    const int k = A.k;
    const long long begin = (long long)tl.x * k, end = (long long)tl.y * k;
    for (long long i = begin + threadIdx.x; i < end; i += blockDim.x) rs[i] = (float)(b[i] / A.diag[i / k]);
}

// sval = a_ij / a_ii in FP32 for j < i, 0 on the diagonal, and diag = a_ii
// (one row per thread).
__global__ void k_jac_scale(int n, const int* __restrict__ ia, const double* __restrict__ a, float* sval,
                            double* diag) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int rs = ia[i], re = ia[i + 1];
        const double dg = a[re - 1];
        for (int p = rs; p < re - 1; ++p) sval[p] = (float)(a[p] / dg);
        sval[re - 1] = 0.0f;
        diag[i] = dg;
    }
}

// Column-sorted copy of each tile's entries (setup): a bitonic sort of
// (column, position) pairs in LDS, one workgroup per tile.  Tiles of more
// than kSortCap entries (summed from memory by the sweep) keep storage order.
// With cbits > 0 the position is stored in the bits of sja above the
// column's cbits (the caller guarantees n <= 2^cbits and that a position the
// sweep reads, below the tile's LDS capacity, fits the 32 - cbits bits left);
// sslot is then not written.
constexpr int kSortCap = 4096;

__global__ __launch_bounds__(256) void k_jac_sort_tiles(const int4* __restrict__ tiles, const int* __restrict__ ja,
                                                        const float* __restrict__ sval, int* sja, float* ssv,
                                                        unsigned short* sslot, int cbits) {
    __shared__ unsigned long long key[kSortCap];
    const int4 tl = tiles[blockIdx.x];
    const int p0 = tl.z, E = tl.w - tl.z;
    const int tid = threadIdx.x;
    if (E <= 0) return;
    if (E > kSortCap) {
        for (int i = tid; i < E; i += blockDim.x) {
            sja[p0 + i] = ja[p0 + i];
            ssv[p0 + i] = sval[p0 + i];
            if (cbits == 0) sslot[p0 + i] = 0;  // unused: such tiles are summed from memory
        }
        return;
    }
    int N = 1;
    while (N < E) N <<= 1;
    for (int i = tid; i < N; i += blockDim.x)
        key[i] = i < E ? ((unsigned long long)(unsigned)ja[p0 + i] << 16) | (unsigned)i : ~0ull;
    for (int size = 2; size <= N; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (int i = tid; i < (N >> 1); i += blockDim.x) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool asc = (lo & size) == 0;
                const unsigned long long a = key[lo], b = key[hi];
                if ((a > b) == asc) {
                    key[lo] = b;
                    key[hi] = a;
                }
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < E; i += blockDim.x) {
        const unsigned long long kv = key[i];
        const int pos = (int)(kv & 0xffffull);
        const unsigned col = (unsigned)(kv >> 16);
        ssv[p0 + i] = sval[p0 + pos];
        if (cbits > 0) {
            // positions past the packed field belong to tiles the sweep sums
            // from memory (more entries than its LDS holds): stored as 0
            const unsigned room = 32u - (unsigned)cbits;
            const unsigned ps = ((unsigned)pos >> room) ? 0u : (unsigned)pos;
            sja[p0 + i] = (int)(col | (ps << cbits));
        } else {
            sja[p0 + i] = (int)col;
            sslot[p0 + i] = (unsigned short)pos;
        }
    }
}

// Row tiles by nnz windows: tile t = {first row whose first entry is >=
// t * window, the same for t + 1 (n for the last), their first entries}.
__device__ __forceinline__ int first_row_at(const int* ia, int n, long long v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ia[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ void k_jac_tiles(const int* ia, int n, int window, int ntiles, int4* tiles) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int r0 = first_row_at(ia, n, (long long)t * window);
    const int r1 = (t + 1 == ntiles) ? n : first_row_at(ia, n, (long long)(t + 1) * window);
    tiles[t] = make_int4(r0, r1, ia[r0], ia[r1]);
}

// status = 1 unless, for every slot (member and right-hand-side column),
// |r|_inf / delta <= tol * |x|_inf.  The norm words are cleared as they are
// read, ready for the next certificate (they are zero from allocation on).
__global__ void k_jac_certify(unsigned long long* norms, int slots, double delta, double tol, int* status) {
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < slots; m += gridDim.x * blockDim.x) {
        unsigned long long rb = 0, xb = 0;
        for (int rep = 0; rep < kNormReps; ++rep) {
            unsigned long long* w = norms + ((size_t)m * kNormReps + rep) * kNormStride;
            rb = max(rb, w[0]);
            xb = max(xb, w[1]);
            w[0] = 0ull;
            w[1] = 0ull;
        }
        const double rn = __longlong_as_double((long long)rb);
        const double xn = __longlong_as_double((long long)xb);
        const bool ok = rn == 0.0 || rn / delta <= tol * xn;
        if (!ok) *status = 1;
    }
}

// per-block minimum of |L_ii| - sum_j |L_ij| (FP64)
__global__ void k_margin(int n, const int* ia, const double* a, double* block_min) {
    __shared__ double sm[256];
    double m = 1e300;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int rs = ia[i], re = ia[i + 1];
        double off = 0.0;
        for (int p = rs; p < re - 1; ++p) off += fabs(a[p]);
        m = fmin(m, fabs(a[re - 1]) - off);
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) block_min[blockIdx.x] = sm[0];
}

// Deterministic right-hand side in [-1, 1) for tuning (hash of the index).
__global__ void k_fill_hash(double* b, size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x) {
        unsigned long long z = i * 0x9e3779b97f4a7c15ull + 0x632be59bd9b4e019ull;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        b[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
}

int jacobi_margin(sp_mat_t* m, double* delta) {
    MatPriv* pv = priv_of(m);
    if (pv->margin_known) {
        *delta = pv->margin;
        return SPTRSV_OK;
    }
    if (!m->a) return fail(SPTRSV_E_INVALID_ARG, "JACOBI needs FP64 values");
    const int blocks = (int)std::max<long long>(1, std::min<long long>(1024, ((long long)m->nr + 255) / 256));
    double* d = nullptr;
    SPT_HIP(hipMalloc(&d, sizeof(double) * blocks));
    hipLaunchKernelGGL(k_margin, dim3(blocks), dim3(256), 0, nullptr, m->nr, m->ia, m->a, d);
    std::vector<double> h(blocks);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(h.data(), d, sizeof(double) * blocks, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "dominance margin", __FILE__, __LINE__);
    double mn = 1e300;
    for (double v : h) mn = std::min(mn, v);
    pv->margin = m->nr > 0 ? mn : 1.0;
    pv->margin_known = true;
    *delta = pv->margin;
    return SPTRSV_OK;
}

static int pow2_at_least(int k) {
    int K = 1;
    while (K < k) K <<= 1;
    return K;
}

// Most rows in one tile of a table of {first row, end row, first entry, end
// entry} records.
static int tiles_rows_max(const std::vector<int>& rec) {
    int r = 0;
    for (size_t i = 0; i + 3 < rec.size(); i += 4) r = std::max(r, rec[i + 1] - rec[i]);
    return r;
}

static int free_table(TileTable& t) {
    for (void* p : {(void*)t.tiles, (void*)t.member, (void*)t.sja, (void*)t.ssv, (void*)t.sslot})
        if (p) SPT_HIP(hipFree(p));
    t = TileTable{};
    return SPTRSV_OK;
}

// Workspaces and the tile table for k right-hand sides (allocates; call
// before any capture).
static int ensure_jacobi_ws(sp_mat_t* m, int k) {
    MatPriv* pv = priv_of(m);
    const size_t words = std::max<size_t>((size_t)m->nr * k, 1);
    if (pv->jac_len < words) {
        drop_graphs(pv);
        for (float** p : {&pv->jac_r, &pv->jac_a, &pv->jac_b})
            if (*p) SPT_HIP(hipFree(*p));
        pv->jac_r = pv->jac_a = pv->jac_b = nullptr;
        pv->jac_len = 0;
        SPT_HIP(hipMalloc(&pv->jac_r, words * sizeof(float)));
        SPT_HIP(hipMalloc(&pv->jac_a, words * sizeof(float)));
        SPT_HIP(hipMalloc(&pv->jac_b, words * sizeof(float)));
        pv->jac_len = words;
    }
    // the norm words: zero between certificates (the certificate clears them)
    const size_t nw = (size_t)std::max(1, pv->jac_members) * k * kNormReps * kNormStride;
    if (pv->jac_norms_len < nw) {
        drop_graphs(pv);
        if (pv->jac_norms) SPT_HIP(hipFree(pv->jac_norms));
        pv->jac_norms = nullptr;
        pv->jac_norms_len = 0;
        SPT_HIP(hipMalloc(&pv->jac_norms, nw * sizeof(unsigned long long)));
        SPT_HIP(hipMemset(pv->jac_norms, 0, nw * sizeof(unsigned long long)));
        SPT_HIP(hipDeviceSynchronize());
        pv->jac_norms_len = nw;
    }
    if (!pv->jac_sval) {
        SPT_HIP(hipMalloc(&pv->jac_sval, sizeof(float) * (size_t)std::max(1, m->nnz)));
        SPT_HIP(hipMalloc(&pv->jac_diag, sizeof(double) * (size_t)std::max(1, m->nr)));
        const unsigned g = (unsigned)std::max<long long>(1, std::min<long long>(4096, ((long long)m->nr + 255) / 256));
        hipLaunchKernelGGL(k_jac_scale, dim3(g), dim3(256), 0, nullptr, m->nr, m->ia, m->a, pv->jac_sval,
                           pv->jac_diag);
        SPT_HIP(hipGetLastError());
        SPT_HIP(hipDeviceSynchronize());
    }
    const int K = pow2_at_least(k);
    for (int w = 0; w < 2; ++w) {
        TileTable& t = pv->jac_tt[w];
        if (t.K == K) continue;
        if (pv->jac_tiles_fixed) return fail(SPTRSV_E_CAPACITY, "batched JACOBI solves take one right-hand side");
        drop_graphs(pv);
        SPT_TRY(free_table(t));
        const int window = (1024 << w) / K;
        const int ntiles = (int)std::max<long long>(1, ((long long)m->nnz + window - 1) / window);
        SPT_HIP(hipMalloc(&t.tiles, 4 * sizeof(int) * (size_t)ntiles));
        hipLaunchKernelGGL(k_jac_tiles, dim3(ceil_div(ntiles, 256)), dim3(256), 0, nullptr, m->ia, m->nr,
                           window, ntiles, reinterpret_cast<int4*>(t.tiles));
        SPT_HIP(hipGetLastError());
        std::vector<int> rec(4 * (size_t)ntiles);
        SPT_HIP(hipMemcpy(rec.data(), t.tiles, sizeof(int) * rec.size(), hipMemcpyDeviceToHost));
        t.rows_max = tiles_rows_max(rec);
        t.count = ntiles;
        t.K = K;
    }
    return SPTRSV_OK;
}

int reserve_jacobi(sp_mat_t* m, int k) { return ensure_jacobi_ws(m, k); }

// Host-side tiles of rows [r_begin, r_end) of a CSR whose row pointers are
// ia (relative: ia[r] - ia[r_begin] is the entry offset inside the range),
// appended as {first row, end row, first entry, end entry} with rows and
// entries shifted by row_base / entry_base.  Same windows (of `window`
// entries) as k_jac_tiles, restarted at the range's first entry, so tiles
// never straddle ranges.
void jacobi_host_tiles(const int* ia, int r_begin, int r_end, long long row_base, long long entry_base,
                       int window, std::vector<int>& out) {
    const long long e0 = ia[r_begin];
    const long long nz = (long long)ia[r_end] - e0;
    int r = r_begin;
    for (long long t = 0; t * window < nz; ++t) {
        const int first = r;
        while (r < r_end && (long long)ia[r] - e0 < (t + 1) * window) ++r;
        if (r == first) continue;  // window inside one long row
        out.push_back((int)(row_base + first - r_begin));
        out.push_back((int)(row_base + r - r_begin));
        out.push_back((int)(entry_base + ia[first] - e0));
        out.push_back((int)(entry_base + ia[r] - e0));
    }
}

// Installs tile tables built on the host (one right-hand side; [0] windows of
// 1024 entries, [1] of 2048) covering the rows the JACOBI path should solve,
// with the member of each tile; members without tiles are left to other
// solvers.  delta is the dominance margin of the covered rows.
int jacobi_set_tiles(sp_mat_t* m, const std::vector<int>* tiles4, const std::vector<int>* tile_member,
                     int members, double delta) {
    MatPriv* pv = priv_of(m);
    drop_graphs(pv);
    pv->jac_tiles_fixed = true;
    pv->jac_members = std::max(1, members);
    pv->margin = delta;
    pv->margin_known = true;
    for (int w = 0; w < 2; ++w) {
        TileTable& t = pv->jac_tt[w];
        SPT_TRY(free_table(t));
        t.K = 1;
        const size_t nt = tile_member[w].size();
        if (nt == 0) continue;
        SPT_HIP(hipMalloc(&t.tiles, sizeof(int) * tiles4[w].size()));
        SPT_HIP(hipMemcpy(t.tiles, tiles4[w].data(), sizeof(int) * tiles4[w].size(), hipMemcpyHostToDevice));
        SPT_HIP(hipMalloc(&t.member, sizeof(int) * nt));
        SPT_HIP(hipMemcpy(t.member, tile_member[w].data(), sizeof(int) * nt, hipMemcpyHostToDevice));
        t.rows_max = tiles_rows_max(tiles4[w]);
        t.count = (int)nt;
    }
    return SPTRSV_OK;
}

template <int K, int TW>
static void launch_sweep(const TileArgs& AS, int ntiles, const float* rs, const float* cur, float* nxt, double* x,
                         int mode, int epl, hipStream_t s, int inner = -1) {
    if constexpr (K == 1) {
        if (inner >= 0) {
            hipLaunchKernelGGL((k_jac_bsweep<TW>), dim3(ntiles), dim3(kTileThreads), bsweep_lds_bytes<TW>(AS.rows_max),
                               s, AS, rs, cur, nxt, x, mode, epl, inner);
            return;
        }
    }
    const size_t lds = sweep_lds_bytes<K, TW>(AS.rows_max);
    if (AS.sja && AS.cbits > 0)
        hipLaunchKernelGGL((k_jac_sweep<K, TW, 2>), dim3(ntiles), dim3(kTileThreads), lds, s, AS, rs, cur, nxt, x,
                           mode, epl);
    else if (AS.sja)
        hipLaunchKernelGGL((k_jac_sweep<K, TW, 1>), dim3(ntiles), dim3(kTileThreads), lds, s, AS, rs, cur, nxt, x,
                           mode, epl);
    else
        hipLaunchKernelGGL((k_jac_sweep<K, TW, 0>), dim3(ntiles), dim3(kTileThreads), lds, s, AS, rs, cur, nxt,
                           x, mode, epl);
}

// Column bits of the packed sorted layout for tiles of tw products, or 0
// when a column and a position the sweep reads (below the tile's LDS
// capacity) do not fit one 32-bit word together.
static int packed_column_bits(int n, int tw) {
    int pbits = 0;
    while ((1 << pbits) < tw + tw / 4) ++pbits;
    const int cbits = 32 - pbits;
    return ((long long)n <= (1ll << cbits)) ? cbits : 0;
}

// Column-sorted entry copy of tile table t of tiles of tw products
// (allocates; setup only).  The position of each entry travels in the high
// bits of its column word when both fit, which saves 2 bytes per entry of
// every sweep's traffic.
static int ensure_sorted(sp_mat_t* m, TileTable& t, int tw) {
    MatPriv* pv = priv_of(m);
    if (t.sja || t.count == 0) return SPTRSV_OK;
    const size_t nz = (size_t)std::max(1, m->nnz);
    t.cbits = packed_column_bits(m->nr, tw);
    SPT_HIP(hipMalloc(&t.sja, sizeof(int) * nz));
    SPT_HIP(hipMalloc(&t.ssv, sizeof(float) * nz));
    if (t.cbits == 0) SPT_HIP(hipMalloc(&t.sslot, sizeof(unsigned short) * nz));
    hipLaunchKernelGGL(k_jac_sort_tiles, dim3(t.count), dim3(256), 0, nullptr, reinterpret_cast<const int4*>(t.tiles),
                       m->ja, pv->jac_sval, t.sja, t.ssv, t.sslot, t.cbits);
    SPT_HIP(hipGetLastError());
    SPT_HIP(hipDeviceSynchronize());
    return SPTRSV_OK;
}

template <int K>
static int jacobi_schedule(sp_mat_t* m, const sptrsv_config_t& c, int k, const double* b, double* x,
                           hipStream_t s) {
    MatPriv* pv = priv_of(m);
    const TileTable& t0 = pv->jac_tt[0];
    const int wide = (c.lanes == 2 && pv->jac_tt[1].count > 0) ? 1 : 0;  // sweep tiles: 1024 or 2048
    const TileTable& ts = pv->jac_tt[wide];
    TileArgs A{m->nr, k, m->ia, m->ja, pv->jac_sval, m->a, pv->jac_diag, reinterpret_cast<const int4*>(t0.tiles),
               t0.member, nullptr, nullptr, nullptr};
    TileArgs AS = A;
    AS.tiles = reinterpret_cast<const int4*>(ts.tiles);
    AS.tile_member = ts.member;
    AS.rows_max = ts.rows_max;
    // block sweeps: in-tile iterations per sweep (-1: point sweeps)
    const int inner = (c.variant == 2 && K == 1) ? (c.inner > 0 ? c.inner : 4) : -1;
    if (c.variant == 1 && ts.sja) {  // column-sorted tiles
        AS.sja = ts.sja;
        AS.ssv = ts.ssv;
        AS.sslot = ts.sslot;
        AS.cbits = ts.cbits;
    }

    const NormArgs none{nullptr, 1};
    const NormArgs with{pv->jac_norms, std::max(1, pv->jac_members)};
    if (t0.count == 0) return SPTRSV_OK;
    const unsigned tiles = (unsigned)t0.count;
    const int sweeps_first = std::max(2, c.sweeps);
    const int sweeps_later = c.sweeps_next > 0 ? std::max(2, c.sweeps_next) : sweeps_first;
    const int outer = std::max(1, c.outer);
    const int epl = c.window > 0 ? c.window : 8;  // entries a phase-2 lane sums
    float* rs = pv->jac_r;  // D^-1 r of the current step; also the first iterate
    for (int o = 0; o < outer; ++o) {
        if (o == 0)
            hipLaunchKernelGGL((k_jac_first<K>), dim3(tiles), dim3(256), 0, s, A, b, rs);
        else
            hipLaunchKernelGGL((k_jac_resid<K>), dim3(tiles), dim3(kTileThreads), 0, s, A, b,
                               (const double*)x, rs, none, epl);
        const float* cur = rs;
        float* nxt = pv->jac_a;
        const int sweeps = o == 0 ? sweeps_first : sweeps_later;
        // the kernel timer brackets the step's consecutive sweeps with one
        // event pair (events between launches would add their own overhead
        // to every sweep)
        const KtPair ev = kernel_timer_pair(s, sweeps - 1, kTimeSweeps);
        if (ev) SPT_HIP(hipEventRecord(ev[0], s));
        for (int sw = 1; sw < sweeps; ++sw) {
            const bool last = sw == sweeps - 1;
            const int mode = last ? (o == 0 ? 1 : 2) : 0;
            if (wide)
                launch_sweep<K, 2048>(AS, ts.count, rs, cur, nxt, x, mode, epl, s, inner);
            else
                launch_sweep<K, 1024>(AS, ts.count, rs, cur, nxt, x, mode, epl, s, inner);
            cur = nxt;
            nxt = (nxt == pv->jac_a) ? pv->jac_b : pv->jac_a;
        }
        if (ev) SPT_HIP(hipEventRecord(ev[1], s));
    }
    // certificate: final FP64 residual with max-norms per member and column
    // (the norms are zero: cleared by the previous certificate)
    const int slots = with.members * k;
    hipLaunchKernelGGL((k_jac_resid<K>), dim3(tiles), dim3(kTileThreads), 0, s, A, b, (const double*)x,
                       (float*)nullptr, with, epl);
    hipLaunchKernelGGL(k_jac_certify, dim3(ceil_div(slots, 256)), dim3(256), 0, s, pv->jac_norms, slots,
                       pv->margin, pv->tol, pv->counters + 3);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

int jacobi_solve(sp_mat_t* m, const sptrsv_config_t& c, int k, const double* b, double* x, hipStream_t s) {
    double delta = 0;
    SPT_TRY(jacobi_margin(m, &delta));
    if (!(delta > 0)) return fail(SPTRSV_E_INVALID_ARG, "JACOBI needs a strictly diagonally dominant matrix");
    SPT_TRY(ensure_jacobi_ws(m, k));
    if (c.variant == 1) {
        MatPriv* pv = priv_of(m);
        const int wide = (c.lanes == 2 && pv->jac_tt[1].count > 0) ? 1 : 0;
        SPT_TRY(ensure_sorted(m, pv->jac_tt[wide], 1024 << wide));
    }
    switch (pow2_at_least(k)) {
        case 1: return jacobi_schedule<1>(m, c, k, b, x, s);
        case 2: return jacobi_schedule<2>(m, c, k, b, x, s);
        case 4: return jacobi_schedule<4>(m, c, k, b, x, s);
        case 8: return jacobi_schedule<8>(m, c, k, b, x, s);
        case 16: return jacobi_schedule<16>(m, c, k, b, x, s);
        case 32: return jacobi_schedule<32>(m, c, k, b, x, s);
        default: return jacobi_schedule<64>(m, c, k, b, x, s);
    }
}

// New values on an unchanged structure: the row-scaled copy and the sorted
// tile copies are rebuilt in their buffers (same addresses), and the
// dominance margin is recomputed when next needed.
int jacobi_values_changed(sp_mat_t* m) {
    MatPriv* pv = priv_of(m);
    pv->margin_known = false;
    if (pv->jac_sval && m->nr > 0) {
        const unsigned g = (unsigned)std::max<long long>(1, std::min<long long>(4096, ((long long)m->nr + 255) / 256));
        hipLaunchKernelGGL(k_jac_scale, dim3(g), dim3(256), 0, nullptr, m->nr, m->ia, m->a, pv->jac_sval,
                           pv->jac_diag);
        SPT_HIP(hipGetLastError());
        for (TileTable& t : pv->jac_tt)
            if (t.sja && t.count > 0) {
                hipLaunchKernelGGL(k_jac_sort_tiles, dim3(t.count), dim3(256), 0, nullptr,
                                   reinterpret_cast<const int4*>(t.tiles), m->ja, pv->jac_sval, t.sja, t.ssv, t.sslot,
                                   t.cbits);
                SPT_HIP(hipGetLastError());
            }
        SPT_HIP(hipDeviceSynchronize());
    }
    return SPTRSV_OK;
}

// Smallest certified schedule for each sweeps-per-step candidate, timed; the
// fastest is returned.  Certification is checked with a hashed right-hand
// side and a safety factor of 8 on the tolerance.
int jacobi_tune(sp_mat_t* m, int k, int reps, sptrsv_config_t* best, float* best_ms) {
    MatPriv* pv = priv_of(m);
    double delta = 0;
    SPT_TRY(jacobi_margin(m, &delta));
    if (!(delta > 0)) return fail(SPTRSV_E_INVALID_ARG, "not strictly diagonally dominant");
    SPT_TRY(ensure_jacobi_ws(m, k));
    const size_t words = std::max<size_t>((size_t)m->nr * k, 1);
    double *b = nullptr, *x = nullptr;
    SPT_HIP(hipMalloc(&b, words * sizeof(double)));
    SPT_HIP(hipMalloc(&x, words * sizeof(double)));
    hipLaunchKernelGGL(k_fill_hash, dim3(1024), dim3(256), 0, nullptr, b, words);
    const double tol = pv->tol;
    hipEvent_t e0, e1;
    SPT_HIP(hipEventCreate(&e0));
    SPT_HIP(hipEventCreate(&e1));
    float bt = 1e30f;
    sptrsv_config_t bc{};
    int st = SPTRSV_OK;
    // whether schedule c certifies tol/8 on the hashed right-hand side
    auto certifies = [&](const sptrsv_config_t& c) -> bool {
        int zero = 0, flag = 0;
        if (hipMemcpy(pv->counters + 3, &zero, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
            st = fail(SPTRSV_E_HIP, "clear flag");
            return false;
        }
        pv->tol = tol / 8;
        st = jacobi_solve(m, c, k, b, x, nullptr);
        pv->tol = tol;
        if (st == SPTRSV_OK && hipMemcpy(&flag, pv->counters + 3, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
            st = fail(SPTRSV_E_HIP, "read flag");
        return st == SPTRSV_OK && flag == 0;
    };
    // per-solve time of schedule c: the best of three rounds of `reps`
    // solves after one warm-up solve (a single round let the choice between
    // close candidates, such as the two tile widths, flip with clock noise)
    auto timed = [&](const sptrsv_config_t& c) -> float {
        st = jacobi_solve(m, c, k, b, x, nullptr);
        float best_round = 1e30f;
        for (int round = 0; round < 3 && st == SPTRSV_OK; ++round) {
            if (hipEventRecord(e0, nullptr) != hipSuccess) {
                st = fail(SPTRSV_E_HIP, "tuning event");
                break;
            }
            for (int r = 0; r < reps && st == SPTRSV_OK; ++r) st = jacobi_solve(m, c, k, b, x, nullptr);
            float ms = 0;
            if (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                if (st == SPTRSV_OK) st = fail(SPTRSV_E_HIP, "tuning event");
                break;
            }
            best_round = std::min(best_round, ms / std::max(reps, 1));
        }
        return best_round;
    };
    // point sweeps (variant 0) and, for one right-hand side, block sweeps
    // (variant 2: exact solves of the tiles' diagonal blocks)
    const std::vector<int> point_sweeps{4, 6, 8, 10, 12, 14, 16, 20, 24};
    const std::vector<int> block_sweeps{2, 3, 4, 5, 6, 8, 10, 12, 16};
    // (block sweeps with 4 and with 8 in-tile iterations: more iterations
    // per sweep can certify with fewer refinement steps, each of which costs
    // an FP64 residual pass)
    for (int var : {0, 2, 3}) {
    if (var >= 2 && k != 1) break;
    int prev_outer = 16;
    for (int sweeps : (var >= 2 ? block_sweeps : point_sweeps)) {
        sptrsv_config_t c{SPTRSV_KERNEL_JACOBI, 1, 256, 1, std::min(var, 2), 0, sweeps, 1, 0, 0,
                          var == 2 ? 4 : var == 3 ? 8 : 0};
        // the smallest number of refinement steps that certifies tol/8
        bool found = false;
        for (int outer = 1; outer <= std::min(16, prev_outer) && st == SPTRSV_OK; ++outer) {
            c.outer = outer;
            if (certifies(c)) {
                found = true;
                break;
            }
        }
        if (st != SPTRSV_OK) break;
        if (!found) continue;
        prev_outer = c.outer;  // more sweeps per step never need more steps
        for (int width : {1, 2}) {  // sweep tiles of 1024 or 2048 products
            c.lanes = width;
            // block sweeps iterate differently with another tile width
            if (var >= 2 && width == 2 && !certifies(c)) {
                if (st != SPTRSV_OK) break;
                continue;
            }
            const float ms = timed(c);
            if (st == SPTRSV_OK && ms < bt) {
                bt = ms;
                bc = c;
            }
        }
        if (st != SPTRSV_OK) break;
    }
    if (st != SPTRSV_OK) break;
    }
    // column-sorted tiles (point sweeps) at either tile width -- the widths
    // rank differently with sorted entries, and a point sweep's iterates do
    // not depend on the tiling -- then phase-2 entries per lane (default 8),
    // each on the best schedule so far
    const bool point = bc.variant == 0;
    for (int v : {-1, -2, 2, 4, 16}) {
        if (st != SPTRSV_OK || bt >= 1e30f) break;
        sptrsv_config_t c = bc;
        if (v < 0) {
            if (!point) continue;
            c.variant = 1;  // column-sorted tiles
            c.lanes = -v;   // tiles of 1024 or 2048 products
        } else {
            c.window = v;
        }
        const float ms = timed(c);
        if (st == SPTRSV_OK && ms < bt) {
            bt = ms;
            bc = c;
        }
    }
    // block sweeps: other in-tile iteration counts on the best schedule
    for (int inn : {2, 8, 4}) {
        if (bc.variant != 2 || bt >= 1e30f || st != SPTRSV_OK) break;
        if (inn == bc.inner) continue;
        sptrsv_config_t c = bc;
        c.inner = inn;
        if (!certifies(c)) {
            if (st != SPTRSV_OK) break;
            continue;
        }
        const float ms = timed(c);
        if (st == SPTRSV_OK && ms < bt) {
            bt = ms;
            bc = c;
        }
    }
    // fewer sweeps in the steps after the first: they solve for a correction
    // that needs only the remaining relative accuracy; stop at the first
    // count that no longer certifies tol/8
    for (int s2 = bc.sweeps - 1; s2 >= 2 && bc.outer > 1 && bt < 1e30f && st == SPTRSV_OK; --s2) {
        sptrsv_config_t c = bc;
        c.sweeps_next = s2;
        if (!certifies(c)) break;
        const float ms = timed(c);
        if (st == SPTRSV_OK && ms < bt) {
            bt = ms;
            bc.sweeps_next = s2;
        }
    }
    int zero = 0;
    (void)hipMemcpy(pv->counters + 3, &zero, sizeof(int), hipMemcpyHostToDevice);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(b);
    (void)hipFree(x);
    if (st != SPTRSV_OK) return st;
    if (bt >= 1e30f) return fail(SPTRSV_E_NOT_CONVERGED, "no Jacobi schedule certified within 16 steps");
    *best = bc;
    *best_ms = bt;
    return SPTRSV_OK;
}

}  // namespace sptrsv

using namespace sptrsv;

extern "C" {

int sptrsv_set_tolerance(sp_mat_t* mat, double rel_tol) {
    if (!mat || !mat->priv || !(rel_tol > 0) || rel_tol >= 1) return fail(SPTRSV_E_INVALID_ARG, "bad tolerance");
    // captured solves certify against the tolerance they were captured with
    drop_graphs(priv_of(mat));
    priv_of(mat)->tol = rel_tol;
    return SPTRSV_OK;
}

int sptrsv_dominance_margin(sp_mat_t* mat, double* delta) {
    if (!mat || !mat->priv || !delta) return fail(SPTRSV_E_INVALID_ARG, "null argument");
    return jacobi_margin(mat, delta);
}

}  // extern "C"
