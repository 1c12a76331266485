// analysis.hip -- GPU level-set analysis (multirow_analysis_base_GPU,
// analysis_csr.cu:272-410) producing bit-exact levels, iorder and the
// (level, class) warp arrays, in both the reference's 32-lane layout and the
// 64-lane layout the wave64 multirow solver uses.
//
// Pipeline (all on the device; the host reads back nlevs and the warp counts
// only, to size arrays):
//   1. levels      sync-free pass of the shared row engine with the max-plus
//                  policy (analysis_csr.cu:104-185)
//   2. nlevs       max reduction (thrust::max_element, :305)
//   3. keys        key = 7*(lev-1) + class(off-diagonal count), bucket and
//                  level histograms (:309-348)
//   4. iorder      stable LSD radix sort of (key, row) pairs, 8-bit digits;
//                  ties keep ascending row order (thrust::stable_sort_by_key,
//                  :350-352)
//   5. buckets     exclusive scan of the bucket histogram (:374)
//   6. warps       ceil(count*size/lanes) per bucket, exclusive scan, per-warp
//                  base row and size, sentinel ibase_row[n_warps] = n
//                  (:354-385, kernel_base :243-270)
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "syncfree.hpp"

namespace sptrsv {

unsigned lds_pad_for(int cu_waves, int wg);

// Tile of the level pass from the integer average nnz/n, as
// aux_call_kernel_analysis_L_coop_groups picks it (analysis_csr.cu:215-240):
// 0-1 -> 1, 2 -> 2, 3-5 -> 4, 6-11 -> 8, 12-23 -> 16, else 32.
static void solve_config_for_analysis(const sp_mat_t* m, int* lanes, int* wg) {
    int avg = m->nr > 0 ? m->nnz / m->nr : 0;
    *lanes = avg <= 1 ? 1 : avg == 2 ? 2 : avg <= 5 ? 4 : avg <= 11 ? 8 : avg <= 23 ? 16 : 32;
    *wg = 256;
}

// ---- helpers -------------------------------------------------------------------

// vect_size_calculator (analysis_csr.cu:12-31) on the off-diagonal count.
__device__ __forceinline__ int row_class(int nz) {
    if (nz == 0) return 6;
    if (nz == 1) return 0;
    if (nz <= 2) return 1;
    if (nz <= 4) return 2;
    if (nz <= 8) return 3;
    if (nz <= 16) return 4;
    return 5;
}

__global__ void k_fill_int(int* p, size_t count, int v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_max(const int* v, int n, int* out) {
    __shared__ int s[256];
    int m = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        m = max(m, v[i]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] = max(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, s[0]);
}

// key, bucket histogram [7*nlevs], level histogram [nlevs], row ids 0..n-1.
// A level outside [1, nlevs] (impossible after a completed level pass) is
// clamped and flagged in *bad instead of indexing out of the histograms.
__global__ void k_keys(int n, const int* __restrict__ ia, const int* __restrict__ lev, int nlevs,
                       int* __restrict__ key, int* __restrict__ vals, int* bucket_cnt,
                       int* lev_size, int* bad) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int l = lev[i];
        if (l < 1 || l > nlevs) {
            atomicOr(bad, 1);
            l = 1;
        }
        int kk = 7 * (l - 1) + row_class(ia[i + 1] - ia[i] - 1);
        key[i] = kk;
        vals[i] = i;
        atomicAdd(&bucket_cnt[kk], 1);
        atomicAdd(&lev_size[l - 1], 1);
    }
}

// The same keys with the bucket histogram kept in LDS per workgroup (when
// the 7 * nlevs buckets fit), each workgroup taking one contiguous range of
// rows: a level holds up to hundreds of thousands of rows, and one global
// atomic per row on its bucket serialises on those few hot words.  Each
// workgroup adds its nonzero counts once; level sizes follow from the
// buckets (k_lev_sizes).
constexpr int kKeyLdsBins = 12288;
__global__ void k_keys_lds(int n, const int* __restrict__ ia, const int* __restrict__ lev, int nlevs,
                           int* __restrict__ key, int* __restrict__ vals, int* bucket_cnt, int* bad) {
    extern __shared__ int s_bins[];
    const int nb = 7 * nlevs;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s_bins[b] = 0;
    __syncthreads();
    const long long chunk = ((long long)n + gridDim.x - 1) / gridDim.x;
    const long long r0 = (long long)blockIdx.x * chunk;
    const long long r1 = r0 + chunk < n ? r0 + chunk : n;
    for (long long i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
        int l = lev[i];
        if (l < 1 || l > nlevs) {
            atomicOr(bad, 1);
            l = 1;
        }
        const int kk = 7 * (l - 1) + row_class(ia[i + 1] - ia[i] - 1);
        key[i] = kk;
        vals[i] = (int)i;
        atomicAdd(&s_bins[kk], 1);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        const int c = s_bins[b];
        if (c) atomicAdd(&bucket_cnt[b], c);
    }
}

__global__ void k_lev_sizes(const int* __restrict__ bucket_cnt, int nlevs, int* __restrict__ lev_size) {
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < nlevs; l += gridDim.x * blockDim.x) {
        int t = 0;
        for (int c = 0; c < 7; ++c) t += bucket_cnt[7 * l + c];
        lev_size[l] = t;
    }
}

// ---- exclusive scan (int), any length ------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ int block_excl_scan(int v, int* total) {
    __shared__ int wsum[kScanThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int off = 0, tot = 0;
    for (int i = 0; i < kScanThreads / 64; ++i) {
        if (i < w) off += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

__global__ void k_scan_reduce(const int* in, long long n, int* sums) {
    long long base = (long long)blockIdx.x * kScanTile;
    int s = 0;
    for (int j = 0; j < kScanItems; ++j) {
        long long i = base + (long long)threadIdx.x * kScanItems + j;
        if (i < n) s += in[i];
    }
    int tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void k_scan_apply(const int* in, int* out, long long n, const int* offs) {
    long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = (base + j < n) ? in[base + j] : 0;
        s += v[j];
    }
    int tot;
    int run = block_excl_scan(s, &tot) + (offs ? offs[blockIdx.x] : 0);
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
}

// out[i] = sum_{j<i} in[j]; in and out may alias.
int exclusive_scan(const int* in, int* out, long long n, hipStream_t s) {
    if (n <= 0) return SPTRSV_OK;
    long long blocks = (n + kScanTile - 1) / kScanTile;
    if (blocks == 1) {
        hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(kScanThreads), 0, s, in, out, n, nullptr);
        SPT_HIP(hipGetLastError());
        return SPTRSV_OK;
    }
    int* sums = nullptr;
    SPT_HIP(hipMalloc(&sums, blocks * sizeof(int)));
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)blocks), dim3(kScanThreads), 0, s, in, n, sums);
    int st = hipGetLastError() == hipSuccess ? exclusive_scan(sums, sums, blocks, s)
                                             : fail(SPTRSV_E_HIP, "scan reduce launch");
    if (st == SPTRSV_OK) {
        hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)blocks), dim3(kScanThreads), 0, s, in, out, n,
                           sums);
        if (hipGetLastError() != hipSuccess) st = fail(SPTRSV_E_HIP, "scan apply launch");
    }
    if (hipStreamSynchronize(s) != hipSuccess && st == SPTRSV_OK) st = fail(SPTRSV_E_HIP, "scan sync");
    (void)hipFree(sums);
    return st;
}

// ---- stable LSD radix sort of (key, value) pairs, 8-bit digits ----------------------
constexpr int kRsThreads = 256;
constexpr int kRsItems = 16;
constexpr int kRsTile = kRsThreads * kRsItems;

// hist[d * ntiles + tile] = number of digit-d keys in the tile.
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const int* keys, int n, int shift,
                                                        int* hist, int ntiles) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kRsTile;
    const int end = min(base + kRsTile, n);
    for (int i = base + threadIdx.x; i < end; i += kRsThreads)
        atomicAdd(&h[(keys[i] >> shift) & 255], 1);
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Scatter one tile in input order: rounds of 256 elements; within a round a
// lane's rank among equal digits comes from 8 ballots (earlier lanes of its
// wave) plus the counts of earlier waves; `run` carries earlier rounds.
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const int* __restrict__ kin,
                                                           const int* __restrict__ vin,
                                                           int* __restrict__ kout,
                                                           int* __restrict__ vout, int n, int shift,
                                                           const int* __restrict__ offs, int ntiles) {
    __shared__ int run[256];
    __shared__ int wcnt[kRsThreads / 64][256];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    run[tid] = offs[tid * ntiles + blockIdx.x];
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int round = 0; round < kRsItems; ++round) {
        for (int j = tid; j < (kRsThreads / 64) * 256; j += kRsThreads) (&wcnt[0][0])[j] = 0;
        __syncthreads();
        const int i = blockIdx.x * kRsTile + round * kRsThreads + tid;
        const bool valid = i < n;
        const int kv = valid ? kin[i] : 0;
        const int d = (kv >> shift) & 255;
        unsigned long long m = __ballot(valid);
        for (int bit = 0; bit < 8; ++bit) {
            unsigned long long bb = __ballot((d >> bit) & 1);
            m &= ((d >> bit) & 1) ? bb : ~bb;
        }
        const int rank = __popcll(m & lt);
        if (valid && rank == 0) wcnt[w][d] = __popcll(m);
        __syncthreads();
        if (valid) {
            int off = run[d];
            for (int q = 0; q < w; ++q) off += wcnt[q][d];
            kout[off + rank] = kv;
            vout[off + rank] = vin[i];
        }
        __syncthreads();
        int add = 0;
        for (int q = 0; q < kRsThreads / 64; ++q) add += wcnt[q][tid];
        run[tid] += add;
        __syncthreads();
    }
}

// Sorts (keys, vals) by key, stably; max_key bounds the keys.  The result is
// left in (keys, vals); tk/tv are scratch of the same length.
int radix_sort_pairs(int* keys, int* vals, int* tk, int* tv, int n, int max_key,
                            hipStream_t s) {
    if (n <= 1) return SPTRSV_OK;
    int bits = 0;
    while (bits < 31 && (max_key >> bits) != 0) ++bits;
    int passes = std::max(1, (bits + 7) / 8);
    int ntiles = (n + kRsTile - 1) / kRsTile;
    int* hist = nullptr;
    SPT_HIP(hipMalloc(&hist, sizeof(int) * 256 * (size_t)ntiles));
    int st = SPTRSV_OK;
    int *ki = keys, *vi = vals, *ko = tk, *vo = tv;
    for (int p = 0; p < passes && st == SPTRSV_OK; ++p) {
        int shift = 8 * p;
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(kRsThreads), 0, s, ki, n, shift, hist, ntiles);
        if (hipGetLastError() != hipSuccess) { st = fail(SPTRSV_E_HIP, "radix hist"); break; }
        st = exclusive_scan(hist, hist, 256LL * ntiles, s);
        if (st != SPTRSV_OK) break;
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(kRsThreads), 0, s, ki, vi, ko, vo, n, shift,
                           hist, ntiles);
        if (hipGetLastError() != hipSuccess) { st = fail(SPTRSV_E_HIP, "radix scatter"); break; }
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    if (st == SPTRSV_OK && ki != keys) {
        if (hipMemcpyAsync(keys, ki, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, s) != hipSuccess ||
            hipMemcpyAsync(vals, vi, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, s) != hipSuccess)
            st = fail(SPTRSV_E_HIP, "radix copy back");
    }
    if (hipStreamSynchronize(s) != hipSuccess && st == SPTRSV_OK) st = fail(SPTRSV_E_HIP, "radix sort sync");
    (void)hipFree(hist);
    return st;
}

// ---- warp arrays ----------------------------------------------------------------------

// get_size_from_pos + get_warps_per_bucks (analysis_csr.cu:66-101): size 2^class,
// diagonal-only class counted as 1, warps = ceil(count*size/lanes).
__global__ void k_warps_per_bucket(const int* cnt, int nb, int lanes, int* wpb) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b <= nb; b += gridDim.x * blockDim.x) {
        if (b == nb) { wpb[b] = 0; continue; }
        int cls = b % 7;
        int size = (cls == 6) ? 1 : (1 << cls);
        if (size > lanes) size = lanes;
        long long t = (long long)cnt[b] * size;
        wpb[b] = (int)((t + lanes - 1) / lanes);
    }
}

// kernel_base (analysis_csr.cu:243-270): one wave per bucket.
__global__ void k_warp_base(const int* warp_off, const int* bucket_off, int nb, int lanes,
                            int* base, int* vsize, int n) {
    const int lane = threadIdx.x & 63;
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    if (wid == 0 && lane == 0) base[warp_off[nb]] = n;
    for (int b = wid; b < nb; b += nw) {
        int off = warp_off[b], cnt = warp_off[b + 1] - off;
        int cls = b % 7;
        int size = (cls == 6) ? 0 : (1 << cls);
        if (size > lanes) size = lanes;
        int sz = size ? size : 1;
        int rows = bucket_off[b];
        for (int x = lane; x < cnt; x += 64) {
            vsize[off + x] = size;
            base[off + x] = rows + x * (lanes / sz);
        }
    }
}

__global__ void k_inverse(const int* perm, int n, int* inv) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        inv[perm[i]] = i;
}

// warp_lev[l] = first warp of level l+1: warp_off[7*l], l = 0..nlevs.
__global__ void k_warp_lev(const int* warp_off, int nlevs, int* warp_lev) {
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l <= nlevs; l += gridDim.x * blockDim.x)
        warp_lev[l] = warp_off[7 * l];
}

static unsigned grid_for(long long count) {
    long long g = (count + 255) / 256;
    return (unsigned)std::max<long long>(1, std::min<long long>(g, 8192));
}

// ---- analysis-owned state -------------------------------------------------------------
// info->priv points to an InfoPriv this library created; a caller struct that
// was never filled here (malloc'ed, uninitialised) is recognised by the
// pointer not being registered, so its garbage pointers are never freed.
struct InfoPriv {
    SegTable seg;              // long-row segments of the multirow walks (derived data)
    unsigned long long id = 0;  // unique per analysis built, loaded or copied
};
static std::mutex g_info_mu;
static std::unordered_set<void*> g_infos;
static unsigned long long g_info_ids = 0;

static InfoPriv* new_info_priv() {
    InfoPriv* p = new InfoPriv();
    std::lock_guard<std::mutex> g(g_info_mu);
    p->id = ++g_info_ids;
    g_infos.insert(p);
    return p;
}

static bool owns_info(const dfr_analysis_info_t* info) {
    std::lock_guard<std::mutex> g(g_info_mu);
    return info->priv && g_infos.count(info->priv);
}

unsigned long long analysis_id(const dfr_analysis_info_t* info) {
    if (!info || !owns_info(info)) return 0;
    return static_cast<const InfoPriv*>(info->priv)->id;
}

static void free_segments(SegTable& t) {
    for (int* p : {t.items, t.seg_lo, t.seg_hi, t.row_seg})
        if (p) (void)hipFree(p);
    t = SegTable{};
}

const SegTable* segments_of(const dfr_analysis_info_t* info) {
    if (!info || !owns_info(info)) return nullptr;
    const SegTable& t = static_cast<const InfoPriv*>(info->priv)->seg;
    return t.items ? &t : nullptr;
}

void free_info_arrays(dfr_analysis_info_t* info) {
    if (!info || !owns_info(info)) return;
    InfoPriv* ip = static_cast<InfoPriv*>(info->priv);
    {
        std::lock_guard<std::mutex> g(g_info_mu);
        g_infos.erase(ip);
    }
    free_segments(ip->seg);
    int* arrays[] = {info->lev_size, info->warp_lev, info->inv_iorder, info->iorder,
                     info->ibase_row, info->ivect_size_warp, info->row_ctr, info->level,
                     info->bucket_off, info->ibase_row64, info->ivect_size_wave};
    for (int* p : arrays)
        if (p) (void)hipFree(p);
    delete ip;
    MODE mode = info->mode;
    *info = dfr_analysis_info_t{};
    info->mode = mode;
}

// Warp arrays for one lane width.
static int build_warps(const int* bucket_cnt, const int* bucket_off, int nb, int lanes, int n,
                       int** base, int** vsize, int* count, int** warp_off_out, hipStream_t s) {
    int* wpb = nullptr;
    SPT_HIP(hipMalloc(&wpb, sizeof(int) * (size_t)(nb + 1)));
    hipLaunchKernelGGL(k_warps_per_bucket, dim3(grid_for(nb + 1)), dim3(256), 0, s, bucket_cnt, nb,
                       lanes, wpb);
    SPT_HIP(hipGetLastError());
    SPT_TRY(exclusive_scan(wpb, wpb, nb + 1, s));
    int total = 0;
    SPT_HIP(hipMemcpyAsync(&total, wpb + nb, sizeof(int), hipMemcpyDeviceToHost, s));
    SPT_HIP(hipStreamSynchronize(s));
    SPT_HIP(hipMalloc(base, sizeof(int) * (size_t)(total + 1)));
    SPT_HIP(hipMalloc(vsize, sizeof(int) * (size_t)std::max(total, 1)));
    hipLaunchKernelGGL(k_warp_base, dim3(grid_for((long long)nb * 64)), dim3(256), 0, s, wpb,
                       bucket_off, nb, lanes, *base, *vsize, n);
    SPT_HIP(hipGetLastError());
    *count = total;
    if (warp_off_out)
        *warp_off_out = wpb;
    else
        SPT_HIP(hipFree(wpb));
    return SPTRSV_OK;
}

// ---- long-row segments -------------------------------------------------------------
// Rows of the 32-lane class (the longest) with more than kSegMin off-diagonal
// entries are split: the row keeps its first kSegOwn entries and each further
// run of kSegEntries becomes a segment item.  The segment items of a wave's
// rows are placed right before that wave in the walk order.

// Segments each wave needs: the sum of seg_count over its rows.
__global__ void k_seg_count(const int* __restrict__ ia, const int* __restrict__ iorder,
                            const int* __restrict__ wbase, const int* __restrict__ wsize, int n_waves, int* cnt) {
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w <= n_waves; w += gridDim.x * blockDim.x) {
        if (w == n_waves) {
            cnt[w] = 0;
            continue;
        }
        int c = 0;
        if (wsize[w] == 32)
            for (int q = wbase[w]; q < wbase[w + 1]; ++q) {
                const int r = iorder[q];
                c += seg_count(ia[r + 1] - ia[r] - 1);
            }
        cnt[w] = c;
    }
}

// Walk order (item t: wave w >= 0, or segment g as -(g + 1)) and the segment
// records; seg_off = exclusive scan of k_seg_count, so wave w sits at item
// w + seg_off[w + 1] with its seg_off[w + 1] - seg_off[w] segments before it.
__global__ void k_seg_fill(const int* __restrict__ ia, const int* __restrict__ iorder,
                           const int* __restrict__ wbase, const int* __restrict__ wsize, int n_waves,
                           const int* __restrict__ seg_off, int* items, int* seg_lo, int* seg_hi, int* row_seg) {
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < n_waves; w += gridDim.x * blockDim.x) {
        const int g0 = seg_off[w];
        const int segs = seg_off[w + 1] - g0;
        int g = g0;
        if (segs > 0)
            for (int q = wbase[w]; q < wbase[w + 1]; ++q) {
                const int r = iorder[q];
                const int rs = ia[r], dpos = ia[r + 1] - 1;
                if (seg_count(dpos - rs) == 0) continue;
                row_seg[r] = g;
                for (int lo = rs + kSegOwn; lo < dpos; lo += kSegEntries, ++g) {
                    items[w + g] = -(g + 1);
                    seg_lo[g] = lo;
                    seg_hi[g] = min(lo + kSegEntries, dpos);
                }
            }
        items[w + g0 + segs] = w;
    }
}

__global__ void k_count_split(const int* row_seg, int n, int* count) {
    int c = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) c += row_seg[i] >= 0;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
}

__global__ void k_fill_i32(int* p, long long n, int v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// Builds (or rebuilds) the segment table of an analysis with the wave64
// layout; nothing when no row needs segments.
int build_segments(dfr_analysis_info_t* info, sp_mat_t* m) {
    if (!owns_info(info)) return SPTRSV_OK;
    SegTable& t = static_cast<InfoPriv*>(info->priv)->seg;
    free_segments(t);
    if (!info->ibase_row64 || info->n_waves == 0 || m->nr == 0) return SPTRSV_OK;
    const MatPriv* pv = priv_of(m);
    if (pv && pv->max_row <= kSegMin + 1) return SPTRSV_OK;  // no row long enough to split
    const int nw = info->n_waves, n = m->nr;
    hipStream_t s = nullptr;
    int* cnt = nullptr;
    SPT_HIP(hipMalloc(&cnt, sizeof(int) * (size_t)(nw + 1)));
    hipLaunchKernelGGL(k_seg_count, dim3(grid_for(nw + 1)), dim3(256), 0, s, m->ia, info->iorder, info->ibase_row64,
                       info->ivect_size_wave, nw, cnt);
    int st = hipGetLastError() == hipSuccess ? SPTRSV_OK : fail(SPTRSV_E_HIP, "segment count");
    if (st == SPTRSV_OK) st = exclusive_scan(cnt, cnt, nw + 1, s);
    int total = 0;
    if (st == SPTRSV_OK && hipMemcpy(&total, cnt + nw, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        st = fail(SPTRSV_E_HIP, "segment total");
    if (st == SPTRSV_OK && total > 0) {
        const long long items = (long long)nw + total;
        if (items > 0x7ffffffeLL) st = fail(SPTRSV_E_CAPACITY, "too many segments");
        auto alloc = [&](int** p, long long count) {
            if (st == SPTRSV_OK && hipMalloc(p, sizeof(int) * (size_t)count) != hipSuccess)
                st = fail(SPTRSV_E_ALLOC, "segment table");
        };
        alloc(&t.items, items);
        alloc(&t.seg_lo, total);
        alloc(&t.seg_hi, total);
        alloc(&t.row_seg, n);
        if (st == SPTRSV_OK) {
            hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(256), 0, s, t.row_seg, (long long)n, -1);
            hipLaunchKernelGGL(k_seg_fill, dim3(grid_for(nw)), dim3(256), 0, s, m->ia, info->iorder,
                               info->ibase_row64, info->ivect_size_wave, nw, cnt, t.items, t.seg_lo, t.seg_hi,
                               t.row_seg);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
                st = fail(SPTRSV_E_HIP, "segment fill");
        }
        t.n_items = (int)items;
        t.n_segs = total;
        if (st == SPTRSV_OK) {
            // rows that own segments
            int* flag = nullptr;
            if (hipMalloc(&flag, sizeof(int)) != hipSuccess || hipMemset(flag, 0, sizeof(int)) != hipSuccess)
                st = fail(SPTRSV_E_ALLOC, "segment count");
            if (st == SPTRSV_OK) {
                hipLaunchKernelGGL(k_count_split, dim3(grid_for(n)), dim3(256), 0, s, t.row_seg, n, flag);
                if (hipMemcpy(&t.n_split, flag, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
                    st = fail(SPTRSV_E_HIP, "segment count");
            }
            if (flag) (void)hipFree(flag);
        }
        if (st != SPTRSV_OK) free_segments(t);
    }
    (void)hipFree(cnt);
    return st;
}

static int analyse(dfr_analysis_info_t* info, sp_mat_t* m, MODE mode) {
    hipStream_t s = nullptr;
    MatPriv* pv = priv_of(m);
    const int n = m->nr;
    free_info_arrays(info);
    *info = dfr_analysis_info_t{};
    info->priv = new_info_priv();
    info->mode = mode;
    info->n = n;
    SPT_HIP(hipMalloc(&info->row_ctr, sizeof(int)));
    SPT_HIP(hipMemset(info->row_ctr, 0, sizeof(int)));
    if (n == 0) return SPTRSV_OK;

    // 1. levels: sync-free max-plus pass (0 = unsolved).
    SPT_HIP(hipMalloc(&info->level, sizeof(int) * (size_t)n));
    int lanes = 2, wg = 256;
    solve_config_for_analysis(m, &lanes, &wg);
    // Levels are integers, so every correct pass gives the same array; the
    // kernel is chosen for speed.  Rows of moderate length (mean >= 6, none
    // over 256 entries) take the lean one-row-per-wave pass with its lanes
    // covering a typical row and 4 resident waves per CU, which keeps the
    // dependency hand-off short.  Short rows (mean < 6, none over 64
    // entries) take one lane per row on the general engine: 64 consecutive
    // rows per wave keep many more rows in flight, which a natural-order walk
    // of a grid-like matrix needs to find ready rows.  Other matrices keep
    // the reference's tile rule.
    int kind = KK_SIMPLE;
    unsigned dyn_lds = 0;
    const int avg = n > 0 ? m->nnz / n : 0;
    if (avg >= 6 && pv->max_row <= 256) {
        kind = KK_FAST_SIMPLE;
        lanes = 4;
        while (lanes < avg && lanes < kWave) lanes <<= 1;
        wg = 256;
        dyn_lds = lds_pad_for(4, wg);
    } else if (avg < 6 && pv->max_row <= 64) {
        lanes = 1;
    }
    {
        SolveArgs A{};
        A.n = n;
        A.k = 1;
        A.ia = m->ia;
        A.ja = m->ja;
        A.pub = info->level;
        A.L = lanes;
        int lg = 0;
        while ((1 << (lg + 1)) <= lanes) ++lg;
        A.lgL = lg;
        A.ticket = pv->counters;
        A.err = pv->counters + 1;
        A.wd_ticks = pv->watchdog_ticks;
        A.dyn_lds = dyn_lds;
        hipLaunchKernelGGL((k_prepare<uint32_t>), dim3(grid_for(n)), dim3(256), 0, s,
                           reinterpret_cast<uint32_t*>(info->level), (size_t)n, 0u, pv->counters, nullptr,
                           (size_t)0, nullptr, nullptr, 0);
        SPT_HIP(hipGetLastError());
        int u = (pv->max_row > 2 * lanes) ? 4 : 1;
        hipError_t e = launch_solve<PolLevel, 1>(kind, 1, wg, u, A, s);
        if (e != hipSuccess) return hip_fail(e, "level pass", __FILE__, __LINE__);
    }
    // 2. nlevs
    int* d_max = nullptr;
    SPT_HIP(hipMalloc(&d_max, sizeof(int) * 2));
    SPT_HIP(hipMemsetAsync(d_max, 0, sizeof(int) * 2, s));
    hipLaunchKernelGGL(k_max, dim3(std::min<unsigned>(grid_for(n), 1024)), dim3(256), 0, s,
                       info->level, n, d_max);
    SPT_HIP(hipGetLastError());
    int h_max[2] = {0, 0};
    SPT_HIP(hipMemcpyAsync(h_max, d_max, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    SPT_HIP(hipMemcpyAsync(&h_max[1], pv->counters + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    SPT_HIP(hipStreamSynchronize(s));
    (void)hipFree(d_max);
    if (h_max[1]) {
        int zero = 0;
        (void)hipMemcpy(pv->counters + 1, &zero, sizeof(int), hipMemcpyHostToDevice);
        return fail(SPTRSV_E_WATCHDOG, "level analysis: a dependency wait exceeded its bound");
    }
    const int nlevs = h_max[0];
    if (nlevs <= 0 || (long long)nlevs * 7 > (1LL << 30))
        return fail(SPTRSV_E_INVALID_CSR, "level analysis produced an invalid level count");
    info->nlevs = nlevs;
    info->lev_ctr = nlevs;
    const int nb = 7 * nlevs;

    // 3. keys and histograms
    int *key = nullptr, *tk = nullptr, *tv = nullptr, *bucket_cnt = nullptr;
    SPT_HIP(hipMalloc(&key, sizeof(int) * (size_t)n));
    SPT_HIP(hipMalloc(&tk, sizeof(int) * (size_t)n));
    SPT_HIP(hipMalloc(&tv, sizeof(int) * (size_t)n));
    SPT_HIP(hipMalloc(&info->iorder, sizeof(int) * (size_t)n));
    SPT_HIP(hipMalloc(&bucket_cnt, sizeof(int) * (size_t)(nb + 1)));
    SPT_HIP(hipMalloc(&info->bucket_off, sizeof(int) * (size_t)(nb + 1)));
    SPT_HIP(hipMalloc(&info->lev_size, sizeof(int) * (size_t)nlevs));
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(nb + 1)), dim3(256), 0, s, bucket_cnt, (size_t)(nb + 1), 0);
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(nlevs)), dim3(256), 0, s, info->lev_size, (size_t)nlevs, 0);
    if (nb <= kKeyLdsBins) {
        const unsigned g = (unsigned)std::max<long long>(1, std::min<long long>(((long long)n + 2047) / 2048, 1024));
        hipLaunchKernelGGL(k_keys_lds, dim3(g), dim3(256), sizeof(int) * (size_t)nb, s, n, m->ia, info->level, nlevs,
                           key, info->iorder, bucket_cnt, pv->counters + 2);
        hipLaunchKernelGGL(k_lev_sizes, dim3(grid_for(nlevs)), dim3(256), 0, s, bucket_cnt, nlevs, info->lev_size);
    } else {
        hipLaunchKernelGGL(k_keys, dim3(grid_for(n)), dim3(256), 0, s, n, m->ia, info->level, nlevs, key,
                           info->iorder, bucket_cnt, info->lev_size, pv->counters + 2);
    }
    SPT_HIP(hipGetLastError());
    {
        int bad = 0;
        SPT_HIP(hipMemcpyAsync(&bad, pv->counters + 2, sizeof(int), hipMemcpyDeviceToHost, s));
        SPT_HIP(hipStreamSynchronize(s));
        if (bad) {
            (void)hipFree(key); (void)hipFree(tk); (void)hipFree(tv); (void)hipFree(bucket_cnt);
            int zero = 0;
            (void)hipMemcpy(pv->counters + 2, &zero, sizeof(int), hipMemcpyHostToDevice);
            return fail(SPTRSV_E_HIP, "level pass left a row without a valid level");
        }
    }

    // 4. iorder = stable sort of rows by key
    int st = radix_sort_pairs(key, info->iorder, tk, tv, n, nb - 1, s);
    (void)hipFree(tk);
    (void)hipFree(tv);
    (void)hipFree(key);
    if (st != SPTRSV_OK) { (void)hipFree(bucket_cnt); return st; }

    // 5. bucket offsets
    SPT_TRY(exclusive_scan(bucket_cnt, info->bucket_off, nb + 1, s));
    SPT_HIP(hipMalloc(&info->inv_iorder, sizeof(int) * (size_t)n));
    hipLaunchKernelGGL(k_inverse, dim3(grid_for(n)), dim3(256), 0, s, info->iorder, n, info->inv_iorder);
    SPT_HIP(hipGetLastError());

    // 6. warp arrays (MULTIROW / ALL), 32-lane reference layout and 64-lane layout
    if (mode == MULTIROW || mode == ALL) {
        int* warp_off = nullptr;
        st = build_warps(bucket_cnt, info->bucket_off, nb, 32, n, &info->ibase_row,
                         &info->ivect_size_warp, &info->n_warps, &warp_off, s);
        if (st == SPTRSV_OK) {
            st = build_warps(bucket_cnt, info->bucket_off, nb, 64, n, &info->ibase_row64,
                             &info->ivect_size_wave, &info->n_waves, nullptr, s);
        }
        if (st == SPTRSV_OK) {
            if (hipMalloc(&info->warp_lev, sizeof(int) * (size_t)(nlevs + 1)) != hipSuccess) {
                st = fail(SPTRSV_E_ALLOC, "warp_lev");
            } else {
                hipLaunchKernelGGL(k_warp_lev, dim3(grid_for(nlevs + 1)), dim3(256), 0, s, warp_off,
                                   nlevs, info->warp_lev);
            }
        }
        if (warp_off) (void)hipFree(warp_off);
    }
    (void)hipFree(bucket_cnt);
    if (st != SPTRSV_OK) return st;
    SPT_HIP(hipStreamSynchronize(s));
    return build_segments(info, m);
}

// ---- analysis files and copies ------------------------------------------------------
unsigned long long fnv1a(const void* p, size_t n, unsigned long long h);
constexpr unsigned long long kFnvBasis = 0xcbf29ce484222325ULL;
const char kAnalysisMagic[8] = {'S', 'P', 'T', 'R', 'A', 'N', 'L', '1'};

// The device arrays of an analysis in file order, with their lengths.
struct InfoArray {
    int** dev;
    long long len;
};
static std::vector<InfoArray> info_arrays(dfr_analysis_info_t* info, bool warps) {
    const long long n = info->n, nl = info->nlevs;
    std::vector<InfoArray> a{{&info->level, n},       {&info->iorder, n},           {&info->inv_iorder, n},
                             {&info->lev_size, nl},   {&info->bucket_off, n ? 7 * nl + 1 : 0}};
    if (warps) {
        a.push_back({&info->warp_lev, nl + 1});
        a.push_back({&info->ibase_row, (long long)info->n_warps + 1});
        a.push_back({&info->ivect_size_warp, info->n_warps});
        a.push_back({&info->ibase_row64, (long long)info->n_waves + 1});
        a.push_back({&info->ivect_size_wave, info->n_waves});
    }
    return a;
}

// FNV-1a of a matrix's structure (row pointers and column indices, copied
// from the device).
static int structure_hash(sp_mat_t* m, unsigned long long* h) {
    std::vector<int> ia((size_t)m->nr + 1), ja((size_t)std::max(m->nnz, 0));
    SPT_HIP(hipMemcpy(ia.data(), m->ia, sizeof(int) * ia.size(), hipMemcpyDeviceToHost));
    if (!ja.empty()) SPT_HIP(hipMemcpy(ja.data(), m->ja, sizeof(int) * ja.size(), hipMemcpyDeviceToHost));
    unsigned long long g = fnv1a(ia.data(), sizeof(int) * ia.size(), kFnvBasis);
    *h = fnv1a(ja.data(), sizeof(int) * ja.size(), g);
    return SPTRSV_OK;
}

static bool has_warps(const dfr_analysis_info_t* info) { return info->mode == MULTIROW || info->mode == ALL; }

// Levels of a level-ordered copy (row p is row iorder[p]) and its identity
// order.
__global__ void k_copy_levels(const int* __restrict__ level, const int* __restrict__ iorder, int n,
                              int* __restrict__ out_level, int* __restrict__ out_iorder, int* __restrict__ out_inv) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        out_level[p] = level[iorder[p]];
        out_iorder[p] = p;
        out_inv[p] = p;
    }
}

// The analysis of `copy`, the level-ordered copy of the matrix `src` was made
// for, derived from `src` instead of recomputed: the copy's levels are src's
// permuted, its iorder is the identity (src's order is a stable sort by key,
// which the copy's rows already follow), and its buckets and warp arrays,
// which index positions in iorder, are src's.  Only the segment table, which
// names entry positions, is built for the copy.
int level_copy_analysis(const dfr_analysis_info_t* src, sp_mat_t* copy, dfr_analysis_info_t* dst) {
    if (!src || !owns_info(src) || !src->iorder || !src->level || !dst || !copy || src->n != copy->nr)
        return fail(SPTRSV_E_NOT_ANALYSED, "level-ordered copy: no level analysis to derive from");
    const bool warps = has_warps(src);
    dfr_analysis_info_t t = *src;
    for (const InfoArray& a : info_arrays(&t, warps)) *a.dev = nullptr;
    t.row_ctr = nullptr;
    t.priv = nullptr;
    free_info_arrays(dst);
    *dst = t;
    dst->priv = new_info_priv();
    int st = SPTRSV_OK;
    if (hipMalloc(&dst->row_ctr, sizeof(int)) != hipSuccess || hipMemset(dst->row_ctr, 0, sizeof(int)) != hipSuccess)
        st = fail(SPTRSV_E_ALLOC, "row_ctr");
    const std::vector<InfoArray> from = info_arrays(const_cast<dfr_analysis_info_t*>(src), warps),
                                 to = info_arrays(dst, warps);
    for (size_t i = 0; i < to.size() && st == SPTRSV_OK; ++i) {
        if (to[i].len == 0) continue;
        const size_t bytes = sizeof(int) * (size_t)to[i].len;
        if (hipMalloc(to[i].dev, bytes) != hipSuccess) {
            st = fail(SPTRSV_E_ALLOC, "analysis arrays");
            break;
        }
        const bool derived = to[i].dev == &dst->level || to[i].dev == &dst->iorder || to[i].dev == &dst->inv_iorder;
        if (!derived && hipMemcpy(*to[i].dev, *from[i].dev, bytes, hipMemcpyDeviceToDevice) != hipSuccess)
            st = fail(SPTRSV_E_HIP, "analysis arrays");
    }
    if (st == SPTRSV_OK && dst->n > 0) {
        hipLaunchKernelGGL(k_copy_levels, dim3(grid_for(dst->n)), dim3(256), 0, nullptr, src->level, src->iorder,
                           dst->n, dst->level, dst->iorder, dst->inv_iorder);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
            st = fail(SPTRSV_E_HIP, "copy levels");
    }
    if (st == SPTRSV_OK) st = build_segments(dst, copy);
    if (st != SPTRSV_OK) free_info_arrays(dst);
    return st;
}

}  // namespace sptrsv

using namespace sptrsv;

extern "C" {

int sptrsv_save_analysis(const char* path, sp_mat_t* mat, dfr_analysis_info_t* info) {
    if (!path || !mat || !mat->priv || !info) return fail(SPTRSV_E_INVALID_ARG, "null argument");
    if (!owns_info(info)) return fail(SPTRSV_E_NOT_ANALYSED, "info holds no analysis of this library");
    if (info->n != mat->nr) return fail(SPTRSV_E_INVALID_ARG, "the analysis belongs to a matrix of another size");
    const bool warps = has_warps(info);
    long long head[8] = {(long long)info->mode, info->n, mat->nnz, info->nlevs, info->n_warps, info->n_waves, 0, 0};
    unsigned long long sh = 0;
    SPT_TRY(structure_hash(mat, &sh));
    std::vector<std::vector<int>> host;
    unsigned long long h = fnv1a(head, sizeof(head), kFnvBasis);
    h = fnv1a(&sh, sizeof(sh), h);
    for (const InfoArray& a : info_arrays(info, warps)) {
        host.emplace_back((size_t)a.len);
        if (a.len > 0) SPT_HIP(hipMemcpy(host.back().data(), *a.dev, sizeof(int) * (size_t)a.len, hipMemcpyDeviceToHost));
        h = fnv1a(host.back().data(), sizeof(int) * host.back().size(), h);
    }
    FILE* f = fopen(path, "wb");
    if (!f) return fail(SPTRSV_E_IO, std::string("cannot create ") + path);
    bool ok = fwrite(kAnalysisMagic, 1, 8, f) == 8 && fwrite(head, sizeof(head), 1, f) == 1 &&
              fwrite(&sh, 8, 1, f) == 1 && fwrite(&h, 8, 1, f) == 1;
    for (const auto& v : host)
        if (ok && !v.empty()) ok = fwrite(v.data(), sizeof(int), v.size(), f) == v.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? SPTRSV_OK : fail(SPTRSV_E_IO, std::string("write error on ") + path);
}

int sptrsv_load_analysis(const char* path, sp_mat_t* mat, dfr_analysis_info_t* info) {
    if (!path || !mat || !mat->priv || !info) return fail(SPTRSV_E_INVALID_ARG, "null argument");
    FILE* f = fopen(path, "rb");
    if (!f) return fail(SPTRSV_E_IO, std::string("cannot open ") + path);
    char magic[8];
    long long head[8];
    unsigned long long sh = 0, h = 0;
    const bool hdr = fread(magic, 1, 8, f) == 8 && memcmp(magic, kAnalysisMagic, 8) == 0 &&
                     fread(head, sizeof(head), 1, f) == 1 && fread(&sh, 8, 1, f) == 1 && fread(&h, 8, 1, f) == 1;
    const long long mode = hdr ? head[0] : -1, n = hdr ? head[1] : -1, nlevs = hdr ? head[3] : -1;
    if (!hdr || mode < ORDER || mode > ALL || n < 0 || n > 0x7ffffffeLL || nlevs < 0 || nlevs > n ||
        head[4] < 0 || head[5] < 0 || head[4] > 0x7ffffffeLL || head[5] > 0x7ffffffeLL) {
        fclose(f);
        return fail(SPTRSV_E_IO, std::string("not an analysis file: ") + path);
    }
    if (n != mat->nr || head[2] != mat->nnz) {
        fclose(f);
        return fail(SPTRSV_E_INVALID_ARG, "the analysis file belongs to a matrix of another size");
    }
    dfr_analysis_info_t t{};
    t.mode = (MODE)mode;
    t.n = (int)n;
    t.nlevs = t.lev_ctr = (int)nlevs;
    t.n_warps = (int)head[4];
    t.n_waves = (int)head[5];
    const bool warps = has_warps(&t);
    std::vector<std::vector<int>> host;
    unsigned long long g = fnv1a(head, sizeof(head), kFnvBasis);
    g = fnv1a(&sh, sizeof(sh), g);
    bool ok = true;
    for (const InfoArray& a : info_arrays(&t, warps)) {
        host.emplace_back((size_t)a.len);
        if (ok && a.len > 0) ok = fread(host.back().data(), sizeof(int), (size_t)a.len, f) == (size_t)a.len;
        if (ok) g = fnv1a(host.back().data(), sizeof(int) * host.back().size(), g);
    }
    fclose(f);
    if (!ok || g != h) return fail(SPTRSV_E_IO, std::string(ok ? "checksum mismatch in " : "truncated analysis file ") + path);
    unsigned long long mh = 0;
    SPT_TRY(structure_hash(mat, &mh));
    if (mh != sh) return fail(SPTRSV_E_INVALID_ARG, "the analysis file belongs to a matrix of another structure");
    // install: the caller's previous arrays go first
    free_info_arrays(info);
    *info = t;
    info->priv = new_info_priv();
    int st = SPTRSV_OK;
    if (hipMalloc(&info->row_ctr, sizeof(int)) != hipSuccess || hipMemset(info->row_ctr, 0, sizeof(int)) != hipSuccess)
        st = fail(SPTRSV_E_ALLOC, "row_ctr");
    size_t i = 0;
    for (const InfoArray& a : info_arrays(info, warps)) {
        const std::vector<int>& v = host[i++];
        if (st != SPTRSV_OK || a.len == 0) continue;
        if (hipMalloc(a.dev, sizeof(int) * v.size()) != hipSuccess ||
            hipMemcpy(*a.dev, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
            st = fail(SPTRSV_E_ALLOC, "analysis arrays");
    }
    if (st == SPTRSV_OK) st = build_segments(info, mat);
    if (st != SPTRSV_OK) free_info_arrays(info);
    return st;
}

int sptrsv_copy_analysis(sp_mat_t* src_mat, dfr_analysis_info_t* src, sp_mat_t* dst_mat, dfr_analysis_info_t* dst) {
    if (!src_mat || !src_mat->priv || !dst_mat || !dst_mat->priv || !src || !dst || src == dst)
        return fail(SPTRSV_E_INVALID_ARG, "bad arguments");
    if (!owns_info(src)) return fail(SPTRSV_E_NOT_ANALYSED, "src holds no analysis of this library");
    if (src_mat->nr != dst_mat->nr || src_mat->nnz != dst_mat->nnz || src->n != src_mat->nr)
        return fail(SPTRSV_E_INVALID_ARG, "the matrices differ in size");
    unsigned long long hs = 0, hd = 0;
    SPT_TRY(structure_hash(src_mat, &hs));
    SPT_TRY(structure_hash(dst_mat, &hd));
    if (hs != hd) return fail(SPTRSV_E_INVALID_ARG, "the matrices differ in structure");
    const bool warps = has_warps(src);
    dfr_analysis_info_t t = *src;
    for (const InfoArray& a : info_arrays(&t, warps)) *a.dev = nullptr;
    t.row_ctr = nullptr;
    free_info_arrays(dst);
    *dst = t;
    dst->priv = new_info_priv();
    int st = SPTRSV_OK;
    if (hipMalloc(&dst->row_ctr, sizeof(int)) != hipSuccess || hipMemset(dst->row_ctr, 0, sizeof(int)) != hipSuccess)
        st = fail(SPTRSV_E_ALLOC, "row_ctr");
    const std::vector<InfoArray> from = info_arrays(src, warps), to = info_arrays(dst, warps);
    for (size_t i = 0; i < to.size() && st == SPTRSV_OK; ++i) {
        if (to[i].len == 0) continue;
        const size_t bytes = sizeof(int) * (size_t)to[i].len;
        if (hipMalloc(to[i].dev, bytes) != hipSuccess ||
            hipMemcpy(*to[i].dev, *from[i].dev, bytes, hipMemcpyDeviceToDevice) != hipSuccess)
            st = fail(SPTRSV_E_ALLOC, "analysis arrays");
    }
    if (st == SPTRSV_OK) st = build_segments(dst, dst_mat);
    if (st != SPTRSV_OK) free_info_arrays(dst);
    return st;
}

int multirow_analysis_base_GPU(dfr_analysis_info_t** info, sp_mat_t* gpu_L, MODE mode) {
    if (!info || !gpu_L || !gpu_L->priv) return fail(SPTRSV_E_INVALID_ARG, "null info or matrix");
    if (mode != ORDER && mode != MULTIROW && mode != ALL && mode != SIMPLE)
        return fail(SPTRSV_E_INVALID_ARG, "unknown MODE");
    if (!*info) {
        *info = static_cast<dfr_analysis_info_t*>(calloc(1, sizeof(dfr_analysis_info_t)));
        if (!*info) return fail(SPTRSV_E_ALLOC, "info struct");
    }
    int st = analyse(*info, gpu_L, mode);
    if (st != SPTRSV_OK) free_info_arrays(*info);
    return st;
}

int sptrsv_analysis_segments(const dfr_analysis_info_t* info, int* n_split_rows, int* n_segments) {
    if (!info || !n_segments) return fail(SPTRSV_E_INVALID_ARG, "null info or count");
    const SegTable* t = segments_of(info);
    *n_segments = t ? t->n_segs : 0;
    if (n_split_rows) *n_split_rows = t ? t->n_split : 0;
    return SPTRSV_OK;
}

int run_analysis(dfr_analysis_info_t* info, sp_mat_t* gpu_L) {
    if (!info || !gpu_L) return fail(SPTRSV_E_INVALID_ARG, "null info or matrix");
    if (info->mode == SIMPLE) return SPTRSV_E_NO_ANALYSIS_NEEDED;
    return multirow_analysis_base_GPU(&info, gpu_L, info->mode);
}

}  // extern "C"
