/*
 * oracle.c -- CPU restatement of the HCL-Fing/SPTRSV hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  The product (sptrsv_amd/) never links, loads or calls it, and has no
 * CPU fallback.
 *
 * The reference is CUDA C++ and cannot be built or run here (no nvcc, no
 * CUDA headers or libraries such as cuSPARSE and NVML, no NVIDIA GPU), so
 * there is no oracle/_ref.
 * Each function below restates the reference file:line it cites and is
 * pinned in tests/test_oracle.py as follows:
 *
 *   oracle_fwd_f64      scipy.sparse.linalg.spsolve_triangular on random
 *                       systems; the exact x == 1 known answer of
 *                       test.cu:13-16 / matrix_properties.h:14 on the 2D
 *                       5-point Laplacian (4/-1 values) and on unit-valued L.
 *   oracle_fwd_f32      exact x == 1 known answer in FP32; bit for bit
 *                       against a numpy float32 row loop (every product,
 *                       difference and the reciprocal rounded to float32)
 *                       on random systems.
 *   oracle_levels       closed forms lev = x+y+1 (2D 5-point, nlevs 127 at
 *                       64x64) and x+y+z+1 (3D 7-point, nlevs 382 at 128^3);
 *                       Kahn peeling and memoised DFS longest path in Python
 *                       on random DAGs; hand-derived golden fixture.
 *   oracle_row_class    every boundary of analysis_csr.cu:13-31 spelled out.
 *   oracle_analysis     iorder vs Python sorted() on (key, row); bucket and
 *                       warp arrays vs the hand-derived golden fixture and
 *                       structural invariants (every row in exactly one warp,
 *                       one bucket per warp, row counts, sentinel).
 *   oracle_coo_to_csr / oracle_extract_lower
 *                       scipy.io.mmread + scipy.sparse on .mtx fixtures,
 *                       hand-derived fixture for the break-at-first-upper
 *                       rule of main.cu:226-232.
 *   oracle_validate_x   direct known answers.
 * tests/test_oracle.py also plants each mutation listed at
 * oracle_set_mutation() and checks that one of these pins fails.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int g_mutation = 0;
/* Planted mistakes (tests only):
 *  1 diagonal read from the first entry of the row instead of the last
 *  2 off-diagonal products added instead of subtracted
 *  3 level max skips the last off-diagonal entry of each row
 *  4 ties in iorder broken by descending row index
 *  5 class boundary shifted (4 off-diagonals -> class 3)
 *  6 warps per bucket rounded down instead of up
 *  7 lower extraction does not stop at the first col >= i
 *  8 symmetric entries not mirrored */
void oracle_set_mutation(int m) { g_mutation = m; }
int oracle_get_mutation(void) { return g_mutation; }

int oracle_check_lower_csr(int n, const int* ia, const int* ja) {
    if (n < 0 || ia[0] != 0) return 1;
    for (int i = 0; i < n; ++i) {
        int s = ia[i], e = ia[i + 1];
        if (e <= s) return i + 1;               /* every row needs its diagonal */
        if (ja[e - 1] != i) return i + 1;       /* diagonal last */
        for (int p = s; p < e - 1; ++p)
            if (ja[p] < 0 || ja[p] >= i) return i + 1;
    }
    return 0;
}

/* solve_csr_simple.cu:44-45 (piv = 1/diag, left_sum seeded with b_i),
 * :60-87 (left_sum -= val*x over the off-diagonal entries),
 * :101 (x = left_sum * piv).  Serial order: storage order within the row. */
void oracle_fwd_f64(int n, const int* ia, const int* ja, const double* a,
                    const double* b, double* x, int k) {
    for (int i = 0; i < n; ++i) {
        int s = ia[i], e = ia[i + 1];
        int dpos = (g_mutation == 1) ? s : e - 1;
        double piv = 1.0 / a[dpos];
        for (int r = 0; r < k; ++r) {
            double left = b[(size_t)i * k + r];
            for (int p = s; p < e - 1; ++p) {
                int pp = (g_mutation == 1) ? p + 1 : p;
                double prod = a[pp] * x[(size_t)ja[pp] * k + r];
                if (g_mutation == 2) left += prod; else left -= prod;
            }
            x[(size_t)i * k + r] = left * piv;
        }
    }
}

void oracle_fwd_f32(int n, const int* ia, const int* ja, const float* a,
                    const float* b, float* x, int k) {
    for (int i = 0; i < n; ++i) {
        int s = ia[i], e = ia[i + 1];
        int dpos = (g_mutation == 1) ? s : e - 1;
        float piv = 1.0f / a[dpos];
        for (int r = 0; r < k; ++r) {
            float left = b[(size_t)i * k + r];
            for (int p = s; p < e - 1; ++p) {
                int pp = (g_mutation == 1) ? p + 1 : p;
                float prod = a[pp] * x[(size_t)ja[pp] * k + r];
                if (g_mutation == 2) left = left + prod; else left = left - prod;
            }
            x[(size_t)i * k + r] = left * piv;
        }
    }
}

/* analysis_csr.cu:147-181: my_level = max(level[col]) over the off-diagonal
 * entries (:150-169, :172-175), stored as 1 + my_level (:179-181).  A
 * diagonal-only row therefore gets level 1.  Rows are visited in increasing
 * order, which is a topological order of a lower-triangular matrix. */
int oracle_levels(int n, const int* ia, const int* ja, int* lev) {
    int nlevs = 0;
    for (int i = 0; i < n; ++i) {
        int m = 0;
        int stop = ia[i + 1] - ((g_mutation == 3) ? 2 : 1);
        for (int p = ia[i]; p < stop; ++p)
            if (lev[ja[p]] > m) m = lev[ja[p]];
        int l = 1 + m;
        lev[i] = l;
        if (l > nlevs) nlevs = l;
    }
    return nlevs;
}

/* vect_size_calculator, analysis_csr.cu:12-31 (input = off-diagonal count,
 * computed at :309-316 as ia[i+1]-ia[i]-1). */
int oracle_row_class(int nnz_row) {
    if (nnz_row == 0) return 6;
    if (nnz_row == 1) return 0;
    if (nnz_row <= 2) return 1;
    if (nnz_row <= ((g_mutation == 5) ? 3 : 4)) return 2;
    if (nnz_row <= 8) return 3;
    if (nnz_row <= 16) return 4;
    return 5;
}

/* get_size_from_pos (analysis_csr.cu:84-101) followed by get_warps_per_bucks
 * (:66-71): size 2^class, the diagonal-only class counts as size 1, and
 * warps = ceil(count * size / lanes). */
static int warps_for_bucket(int cls, int count, int lanes) {
    int size = (cls == 6) ? 1 : (1 << cls);
    if (size > lanes) size = lanes;
    long long t = (long long)count * size;
    if (g_mutation == 6) return (int)(t / lanes);
    return (int)((t + lanes - 1) / lanes);
}

int oracle_count_warps(int n, const int* ia, const int* ja, int lanes) {
    int* lev = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    int nlevs = oracle_levels(n, ia, ja, lev);
    int nb = 7 * nlevs;
    int* cnt = (int*)calloc(nb > 0 ? nb : 1, sizeof(int));
    for (int i = 0; i < n; ++i)
        cnt[7 * (lev[i] - 1) + oracle_row_class(ia[i + 1] - ia[i] - 1)]++;
    int w = 0;
    for (int b = 0; b < nb; ++b) w += warps_for_bucket(b % 7, cnt[b], lanes);
    free(cnt);
    free(lev);
    return w;
}

int oracle_analysis(int n, const int* ia, const int* ja, int lanes,
                    int* lev, int* key, int* iorder,
                    int* bucket_off, int* warp_off,
                    int* ibase_row, int* ivect_size_warp) {
    int nlevs = oracle_levels(n, ia, ja, lev);
    int nb = 7 * nlevs;
    /* key = 7*(lev-1)+class: subst_one (:34-38) then get_index_from_size
     * (:53-57) applied at :322-332. */
    for (int i = 0; i < n; ++i)
        key[i] = 7 * (lev[i] - 1) + oracle_row_class(ia[i + 1] - ia[i] - 1);
    /* bucket counts (stable_sort + reduce_by_key + scatter, :343-348), the
     * entry at index nb stays 0, then exclusive scan over nb+1 (:374). */
    for (int b = 0; b <= nb; ++b) bucket_off[b] = 0;
    for (int i = 0; i < n; ++i) bucket_off[key[i]]++;
    int acc = 0;
    for (int b = 0; b <= nb; ++b) { int c = bucket_off[b]; bucket_off[b] = acc; acc += c; }
    /* iorder: stable_sort_by_key of (key, 0..n-1) (:350-352).  A counting
     * sort that scatters rows in increasing order is stable. */
    int* cursor = (int*)malloc(sizeof(int) * (nb + 1));
    memcpy(cursor, bucket_off, sizeof(int) * (nb + 1));
    if (g_mutation == 4) {
        for (int b = 0; b < nb; ++b) cursor[b] = bucket_off[b + 1] - 1;
        for (int i = 0; i < n; ++i) iorder[cursor[key[i]]--] = i;
    } else {
        for (int i = 0; i < n; ++i) iorder[cursor[key[i]]++] = i;
    }
    free(cursor);
    /* warps per bucket and their exclusive scan (:363-370). */
    acc = 0;
    for (int b = 0; b < nb; ++b) {
        int c = bucket_off[b + 1] - bucket_off[b];
        int w = warps_for_bucket(b % 7, c, lanes);
        warp_off[b] = acc;
        acc += w;
    }
    warp_off[nb] = acc;
    int n_warps = acc;
    /* kernel_base (:243-270): per bucket, warp x gets base row
     * bucket_start + x*(lanes/size) and size 0 (diagonal only) or 2^class;
     * the sentinel ibase_row[n_warps] = n (:249). */
    for (int b = 0; b < nb; ++b) {
        int cls = b % 7;
        int size = (cls == 6) ? 0 : (1 << cls);
        if (size > lanes) size = lanes;
        int sz = size ? size : 1;
        for (int x = 0; x < warp_off[b + 1] - warp_off[b]; ++x) {
            ivect_size_warp[warp_off[b] + x] = size;
            ibase_row[warp_off[b] + x] = bucket_off[b] + x * (lanes / sz);
        }
    }
    ibase_row[n_warps] = n;
    return n_warps;
}

int oracle_coo_csr_nnz(int nnz, const int* ri, const int* ci, int symmetric) {
    int t = nnz;
    if (symmetric && g_mutation != 8)
        for (int i = 0; i < nnz; ++i) if (ri[i] != ci[i]) ++t;
    return t;
}

/* main.cu:124-212: count per row (+ mirrored count, :160-165), exclusive scan
 * (:167-176), then fill in input order; for a symmetric input each
 * off-diagonal entry goes to its row and to the transposed row (:186-204). */
void oracle_coo_to_csr(int m, int nnz, const int* ri, const int* ci,
                       const double* v, int symmetric,
                       int* row_ptr, int* col, double* val) {
    int sym = symmetric && g_mutation != 8;
    int* cnt = (int*)calloc(m + 1, sizeof(int));
    for (int i = 0; i < nnz; ++i) {
        cnt[ri[i]]++;
        if (sym && ri[i] != ci[i]) cnt[ci[i]]++;
    }
    row_ptr[0] = 0;
    for (int i = 0; i < m; ++i) row_ptr[i + 1] = row_ptr[i] + cnt[i];
    memset(cnt, 0, sizeof(int) * (m + 1));
    for (int i = 0; i < nnz; ++i) {
        int r = ri[i], c = ci[i];
        int off = row_ptr[r] + cnt[r]++;
        col[off] = c;
        val[off] = v[i];
        if (sym && r != c) {
            off = row_ptr[c] + cnt[c]++;
            col[off] = r;
            val[off] = v[i];
        }
    }
    free(cnt);
}

/* main.cu:216-246.  Entries with col < i are copied in storage order and the
 * scan of the row stops at the first col >= i (:226-232); then the diagonal
 * is appended (:235-237).  The reference writes 1.0 for every value. */
int oracle_extract_lower(int m, const int* row_ptr, const int* col,
                         const double* val, int unit,
                         int* l_ptr, int* l_col, double* l_val) {
    int p = 0;
    l_ptr[0] = 0;
    for (int i = 0; i < m; ++i) {
        double diag = 1.0;
        if (!unit)
            for (int j = row_ptr[i]; j < row_ptr[i + 1]; ++j)
                if (col[j] == i) { diag = val[j]; break; }
        for (int j = row_ptr[i]; j < row_ptr[i + 1]; ++j) {
            if (col[j] < i) {
                l_col[p] = col[j];
                l_val[p] = unit ? 1.0 : val[j];
                ++p;
            } else if (g_mutation != 7) {
                break;
            }
        }
        l_col[p] = i;
        l_val[p] = diag;
        ++p;
        l_ptr[i + 1] = p;
    }
    return p;
}

/* matrix_properties.h:7-29: count entries with |1 - x| != 0. */
int oracle_validate_x(const double* x, int n) {
    int err = 0;
    for (int i = 0; i < n; ++i)
        if (fabs(1.0 - x[i]) != 0.0) ++err;
    return err;
}
