/*
 * oracle.h -- CPU restatement of the HCL-Fing/SPTRSV algorithms, used ONLY as
 * the checker of the HIP library (by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py).  Nothing in the product links or calls this.
 *
 * Every function cites the reference file:line it restates.  How each one is
 * pinned (closed forms, brute force, independent libraries, hand-derived
 * vectors) is listed in the header of oracle.c.
 *
 * Conventions: CSR lower triangle, int32 row pointer / column indices, the
 * diagonal is the LAST entry of every row (solve_csr_simple.cu:45,60).
 */
#ifndef SPTRSV_ORACLE_H
#define SPTRSV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Planted-mistake switch: tests set a non-zero value to check that the pins
 * of the oracle catch a plausible bug.  0 = faithful restatement. */
void oracle_set_mutation(int m);
int oracle_get_mutation(void);

/* 0 if ia/ja describe a lower-triangular CSR with the diagonal last in every
 * row; otherwise the first offending row + 1 (as a positive number). */
int oracle_check_lower_csr(int n, const int* ia, const int* ja);

/* Serial forward substitution x = L^{-1} b (FP64), k right-hand sides stored
 * row-major (x[i*k + r]).  solve_csr_simple.cu:44-45,82,101: the row sum runs
 * over the off-diagonal entries in storage order starting from b_i, and the
 * result is multiplied by 1/L_ii. */
void oracle_fwd_f64(int n, const int* ia, const int* ja, const double* a,
                    const double* b, double* x, int k);

/* Same in FP32 arithmetic (VALUE_TYPE=float build of the reference). */
void oracle_fwd_f32(int n, const int* ia, const int* ja, const float* a,
                    const float* b, float* x, int k);

/* Level of every row: lev[i] = 1 + max over off-diagonal j of lev[j], 1 for a
 * diagonal-only row (analysis_csr.cu:147-181).  Returns nlevs = max lev. */
int oracle_levels(int n, const int* ia, const int* ja, int* lev);

/* Row class from the off-diagonal count (vect_size_calculator,
 * analysis_csr.cu:12-31): 0->6, 1->0, 2->1, 3..4->2, 5..8->3, 9..16->4, >16->5 */
int oracle_row_class(int nnz_offdiag);

/* Number of warps the reference's analysis would emit for lanes-wide warps
 * (32 in the reference); needed to size the arrays of oracle_analysis. */
int oracle_count_warps(int n, const int* ia, const int* ja, int lanes);

/* Full multirow analysis (multirow_analysis_base_GPU, analysis_csr.cu:272-410):
 *   lev[n]          1-based levels
 *   key[n]          7*(lev-1)+class (analysis_csr.cu:53-57, :322-332)
 *   iorder[n]       stable sort of 0..n-1 by key (:350-352)
 *   bucket_off[7*nlevs+1] exclusive scan of rows per key (:343-348, :374)
 *   warp_off[7*nlevs+1]   exclusive scan of warps per key (:363-370)
 *   ibase_row[n_warps+1], ivect_size_warp[n_warps]   (kernel_base :243-270)
 * lanes: 32 reproduces the reference; 64 gives the same layout for wave64.
 * Arrays must be sized by the caller (nlevs from oracle_levels, n_warps from
 * oracle_count_warps).  Returns n_warps. */
int oracle_analysis(int n, const int* ia, const int* ja, int lanes,
                    int* lev, int* key, int* iorder,
                    int* bucket_off, int* warp_off,
                    int* ibase_row, int* ivect_size_warp);

/* COO (0-based) -> CSR in file order, with symmetric mirroring of
 * off-diagonal entries when symmetric != 0 (main.cu:124-212).
 * row_ptr[m+1], col[nnz_out], val[nnz_out]; nnz_out = nnz + mirrored count
 * (query it with oracle_coo_csr_nnz). */
int oracle_coo_csr_nnz(int nnz, const int* ri, const int* ci, int symmetric);
void oracle_coo_to_csr(int m, int nnz, const int* ri, const int* ci,
                       const double* v, int symmetric,
                       int* row_ptr, int* col, double* val);

/* Lower-triangle extraction (main.cu:216-246): per row, take entries with
 * col < i in storage order up to the first col >= i, then append the
 * diagonal.  unit != 0 gives the reference's all-ones values (:228,:236);
 * unit == 0 keeps A's values and uses A's diagonal entry (1.0 if absent).
 * Returns nnz(L); col/val must hold nnz(A) + m entries. */
int oracle_extract_lower(int m, const int* row_ptr, const int* col,
                         const double* val, int unit,
                         int* l_ptr, int* l_col, double* l_val);

/* validate_x (matrix_properties.h:7-29): number of i with x[i] != 1. */
int oracle_validate_x(const double* x, int n);

#ifdef __cplusplus
}
#endif
#endif
