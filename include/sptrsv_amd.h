/*
 * sptrsv_amd.h -- C-ABI of the MI355X (gfx950) sparse lower-triangular solver.
 *
 * Drop-in for the host interface of HCL-Fing/SPTRSV.  Every entry point below
 * names the reference declaration it replaces (file:line, paths relative to
 * the reference repository).  Names, argument order and argument meaning are
 * kept; the differences are deliberate and listed per function:
 *   - out-parameters are caller-allocated structs (the reference passes them
 *     by value and loses them, interface.cu:9-43);
 *   - every function returns an int status (0 = OK, negative = SPTRSV_E_*)
 *     instead of printing and continuing (common.h:163-167);
 *   - precision is chosen per call (suffix _f32, or the mixed entry points),
 *     not per binary (Makefile:41);
 *   - stream arguments are hipStream_t (NULL = the default stream).
 * Plain C types only: no HIP headers are needed to include this file.
 */
#ifndef SPTRSV_AMD_H
#define SPTRSV_AMD_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same typedef as hip_runtime_api.h, so callers need not include HIP. */
typedef struct ihipStream_t* hipStream_t;

/* ---- status codes ---------------------------------------------------- */
#define SPTRSV_OK 0
#define SPTRSV_E_NO_ANALYSIS_NEEDED (-1) /* run_analysis with SIMPLE (interface.cu:29) */
#define SPTRSV_E_INVALID_ARG (-2)
#define SPTRSV_E_INVALID_CSR (-3)        /* not lower triangular / diagonal not last */
#define SPTRSV_E_ALLOC (-4)
#define SPTRSV_E_HIP (-5)
#define SPTRSV_E_WATCHDOG (-6)           /* a dependency wait exceeded its bound */
#define SPTRSV_E_NOT_ANALYSED (-7)       /* ORDER/MULTIROW solve without analysis */
#define SPTRSV_E_UNAVAILABLE (-8)        /* optional runtime (amdsmi, rocSPARSE) absent */
#define SPTRSV_E_IO (-9)                 /* Matrix Market read/parse failure */
#define SPTRSV_E_CAPACITY (-10)          /* request larger than a fixed limit */
#define SPTRSV_E_NOT_CONVERGED (-11)     /* iterative solve missed its certified bound */

const char* sptrsv_status_string(int status);
/* Detail of the last error raised on the calling thread ("" if none). */
const char* sptrsv_last_error(void);

/* ---- types (dfr_syncfree.h:7-41) -------------------------------------- */

/* dfr_syncfree.h:7-12 */
typedef enum MODE { SIMPLE = 0, ORDER = 1, MULTIROW = 2, ALL = 3 } MODE;

/* dfr_syncfree.h:14-23.  Device CSR of a lower-triangular matrix, int32
 * indices, diagonal stored last in every row.  a holds FP64 values; a32 is
 * the FP32 copy used by the FP32 and mixed-precision solves.  priv is owned
 * by the library (ticket counters, watchdog word, workspaces). */
typedef struct {
    int nr;
    int nc;
    int nnz;
    int* ia;
    int* ja;
    double* a;
    float* a32;
    void* priv;
} sp_mat_t;

/* dfr_syncfree.h:25-41.  All pointers are device pointers owned by the
 * library and released by free_device_analysis.  Fields the reference
 * declares but never sets (lev_ctr, lev_size, warp_lev, inv_iorder) are
 * filled here.  Fields after row_ctr are additions:
 *   n            rows analysed
 *   level        1-based dependency level of every row [n]
 *   bucket_off   exclusive scan of rows per key 7*(level-1)+class [7*nlevs+1]
 *   n_waves, ibase_row64, ivect_size_wave
 *                the same (level, class) packing for 64-lane wavefronts
 *                (ibase_row / ivect_size_warp keep the reference's 32-lane
 *                layout, analysis_csr.cu:243-270, for drop-in parity). */
typedef struct {
    MODE mode;
    int lev_ctr;          /* number of non-empty levels (= nlevs) */
    int nlevs;
    int* lev_size;        /* rows per level [nlevs] */
    int* warp_lev;        /* first 32-lane warp of each level [nlevs+1] */
    int n_warps;
    int* inv_iorder;      /* position of each row in iorder [n] */
    int* iorder;          /* rows sorted by (level, class, row) [n] */
    int* ibase_row;       /* [n_warps+1] */
    int* ivect_size_warp; /* [n_warps] */
    int* row_ctr;         /* ticket counter [1] */
    int n;
    int* level;
    int* bucket_off;
    int n_waves;
    int* ibase_row64;
    int* ivect_size_wave;
    void* priv;
} dfr_analysis_info_t;

/* ---- host interface (header/interface.h:6-9) ------------------------- */

/* interface.h:6 / interface.cu:9-26.  Uploads a host CSR lower triangle
 * (rowptr[n+1], colidx[nnz], val[nnz], nnz = rowptr[n]-rowptr[0]) into the
 * caller-allocated *gpu_L, builds the FP32 copy on the device, validates the
 * structure (SPTRSV_E_INVALID_CSR).  mode is accepted for signature parity. */
int allocate_memory(const int* csrRowPtrL, const int* csrColIdxL,
                    const double* csrValL, int n, sp_mat_t* gpu_L, MODE mode);
/* FP32 input values (VALUE_TYPE=float builds): a = NULL, a32 set. */
int allocate_memory_f32(const int* csrRowPtrL, const int* csrColIdxL,
                        const float* csrValL, int n, sp_mat_t* gpu_L, MODE mode);

/* Upper-triangular U (diagonal first in every row, other columns in (i, n)):
 * no counterpart in the reference (which solves L only).  U is stored as the
 * lower triangle of the reversed index order, so every entry point below --
 * analysis, the three solvers, mixed precision, Jacobi, autotuning, captured
 * solves -- then solves U x = b with b and x in the caller's order. */
int allocate_memory_upper(const int* csrRowPtrU, const int* csrColIdxU, const double* csrValU, int n,
                          sp_mat_t* gpu_U, MODE mode);
/* Transpose: no counterpart in the reference.  Builds, on the device, a new
 * handle *dst that solves A^T x = b, where A is what src solves (L, or U from
 * allocate_memory_upper) -- e.g. the L^T half of an incomplete-Cholesky
 * preconditioner from the L already uploaded.  dst is independent of src (its
 * own arrays; analyse it and free it like any other handle). */
int sptrsv_transpose(sp_mat_t* src, sp_mat_t* dst);
/* New values on the same sparsity pattern, keeping the handle's structure,
 * its analyses and tuned kernel choices: no counterpart in the reference
 * (whose analysis is redone per matrix).  vals is laid out like the value
 * array the handle was created from (allocate_memory*, sptrsv_coo_to_lower,
 * sptrsv_spsv_create; a transpose handle like its source's), count entries,
 * in host memory or, with on_device != 0, device memory.  Stream-ordered
 * after the device is idle; captured solves of the handle are dropped and
 * re-captured on their next call. */
int sptrsv_update_values(sp_mat_t* mat, const double* vals, long long count, int on_device);

/* ---- spsv-style handle (no counterpart in the reference) --------------
 * The create / analysis / solve / destroy sequence of the vendor sparse
 * libraries' SpSV, over a device CSR (0-based, columns in any order).  Only
 * the triangle named by fill is read; DIAG_UNIT ignores stored diagonal
 * entries, DIAG_NON_UNIT uses each row's first (i, i) entry and rejects a
 * row without one (SPTRSV_E_INVALID_CSR).  op(A) is A or A^T.  The handle
 * owns device copies; the caller's arrays are not referenced afterwards. */
#define SPTRSV_FILL_LOWER 0
#define SPTRSV_FILL_UPPER 1
#define SPTRSV_DIAG_NON_UNIT 0
#define SPTRSV_DIAG_UNIT 1
#define SPTRSV_OP_NONE 0
#define SPTRSV_OP_TRANSPOSE 1
typedef struct sptrsv_spsv sptrsv_spsv_t;
int sptrsv_spsv_create(sptrsv_spsv_t** out, int n, int nnz, const int* d_ia, const int* d_ja,
                       const double* d_a, int fill, int diag, int op);
/* Level analysis; tune != 0 also autotunes the kernel for (precision, k). */
int sptrsv_spsv_analysis(sptrsv_spsv_t* h, int precision, int k, int tune);
/* op(A) x = alpha * b for k right-hand sides (device b, x; float* for
 * SPTRSV_PREC_F32).  Stream-ordered, like sptrsv_solve. */
int sptrsv_spsv_solve(sptrsv_spsv_t* h, int precision, double alpha, const void* b, void* x, int k,
                      hipStream_t stream);
/* The handle's matrix, for the sp_mat_t entry points (sptrsv_check, ...). */
sp_mat_t* sptrsv_spsv_matrix(sptrsv_spsv_t* h);
void sptrsv_spsv_destroy(sptrsv_spsv_t* h);
/* interface.h:7 / interface.cu:28-38.  Runs the analysis for info->mode
 * (set by the caller) into the caller-allocated *info.  Returns
 * SPTRSV_E_NO_ANALYSIS_NEEDED (-1) for SIMPLE, like the reference. */
int run_analysis(dfr_analysis_info_t* info, sp_mat_t* gpu_L);

/* interface.h:8 / interface.cu:40-95.  Host b[n] in, host x[n] out.
 * info == NULL or info->mode == SIMPLE runs csr_L_solve_simple; ORDER and
 * MULTIROW run their solvers; ALL runs the autotuned choice. */
int run_solver(sp_mat_t* gpu_L, dfr_analysis_info_t* info, const double* b,
               double* x, int n);

/* interface.h:9 / interface.cu:97-102.  Frees every device array of gpu_L
 * and info (either may be NULL); the structs themselves stay the caller's. */
void free_device_analysis(sp_mat_t* gpu_L, dfr_analysis_info_t* info);

/* ---- test driver (header/test.h:4) ---------------------------------- */

/* test.h:4 / test.cu:4-30, called by main.cu:259 with the host CSR of the
 * lower triangle: uploads it, runs the ALL analysis, builds b = L * 1 and runs
 * the simple, order and multirow solvers once each under BENCH_RUN_SOLVE
 * timing (common.h:103-124) with the exact x == 1 check (validate_x,
 * matrix_properties.h:7-29), printing one line per solver like the
 * reference.  Returns the summed count of x != 1 (0 = all passed) or a
 * negative status; the reference returns void, so a caller that ignores the
 * result is unchanged. */
int test_solve_L_analysis_multirow(const char* filename, int* csrRowPtrL, int* csrColIdxL, double* csrValL,
                                   int n);
/* The same for a VALUE_TYPE=float build (FP32 values, the FP32 solvers). */
int test_solve_L_analysis_multirow_f32(const char* filename, int* csrRowPtrL, int* csrColIdxL, float* csrValL,
                                       int n);

/* ---- analysis (header/analysis.h:20) --------------------------------- */

/* analysis.h:20 / analysis_csr.cu:272-410.  GPU level-set analysis.  *info
 * may point to a caller struct or be NULL (then one is calloc'ed and
 * returned; release it with free() after free_device_analysis).  ORDER
 * stops after iorder; MULTIROW/ALL also build the warp arrays. */
int multirow_analysis_base_GPU(dfr_analysis_info_t** info, sp_mat_t* gpu_L, MODE mode);

/* Analysis reuse: no counterpart in the reference, which recomputes the
 * analysis (analysis_csr.cu:272-410 into dfr_syncfree.h:25-41) on every run.
 * sptrsv_save_analysis writes the analysis of mat to a file: magic, mode and
 * sizes, an FNV-1a checksum of mat's structure (ia, ja), an FNV-1a checksum
 * of the whole content, then level, iorder, inv_iorder, lev_size, bucket_off
 * and (MULTIROW / ALL) warp_lev and both warp layouts.
 * sptrsv_load_analysis reads such a file into *info (the caller's struct; an
 * analysis it held is freed first) for a matrix of the same structure:
 * SPTRSV_E_IO for a damaged or truncated file, SPTRSV_E_INVALID_ARG for a
 * file of another structure.  sptrsv_copy_analysis gives dst_mat, which must
 * have the structure of src_mat (checked), its own copy of src's analysis,
 * e.g. for another system on the same pattern.  Free with
 * free_device_analysis as usual. */
int sptrsv_save_analysis(const char* path, sp_mat_t* mat, dfr_analysis_info_t* info);
int sptrsv_load_analysis(const char* path, sp_mat_t* mat, dfr_analysis_info_t* info);
int sptrsv_copy_analysis(sp_mat_t* src_mat, dfr_analysis_info_t* src, sp_mat_t* dst_mat,
                         dfr_analysis_info_t* dst);

/* Long-row split of a MULTIROW / ALL analysis (no counterpart in the
 * reference, whose multirow kernel walks every row with at most 32 lanes,
 * solve_csr_multirow.cu:21-114): a row of more than 256 off-diagonal entries
 * keeps its first 128 and every further run of 256 entries becomes a segment
 * that a wave of its own folds; the row adds the segments' partial sums in a
 * fixed order, so results stay bit-reproducible.  Reports how many rows are
 * split (may be NULL) and the number of segments. */
int sptrsv_analysis_segments(const dfr_analysis_info_t* info, int* n_split_rows, int* n_segments);

/* ---- solvers (header/solver.h:5-8) ----------------------------------- */
/* b and x are device pointers.  x is fully overwritten.  All solves are
 * stream-ordered, allocate nothing after the first call for a given matrix
 * and right-hand-side count, and can be captured in a hipGraph. */

/* solver.h:5 / solve_csr_simple.cu:141-157.  Natural row order, no
 * analysis.  is_solved is accepted for signature parity and unused
 * (readiness is carried by x itself); it may be NULL. */
int csr_L_solve_simple(sp_mat_t* mat, const double* b, double* x, int n, int* is_solved);
/* solver.h:6 / solve_csr_order.cu:171-189.  Rows in analysed iorder. */
int csr_L_solve_order(sp_mat_t* mat, dfr_analysis_info_t* info, const double* b,
                      double* x, int n, hipStream_t stream);
/* solver.h:7 / solve_csr_multirow.cu:116-130.  (level, class) packed rows. */
int csr_L_solve_multirow(sp_mat_t* mat, dfr_analysis_info_t* info, const double* b,
                         double* x, int n, hipStream_t stream);

/* FP32 versions (VALUE_TYPE=float builds of the same three solvers). */
int csr_L_solve_simple_f32(sp_mat_t* mat, const float* b, float* x, int n, int* is_solved);
int csr_L_solve_order_f32(sp_mat_t* mat, dfr_analysis_info_t* info, const float* b,
                          float* x, int n, hipStream_t stream);
int csr_L_solve_multirow_f32(sp_mat_t* mat, dfr_analysis_info_t* info, const float* b,
                             float* x, int n, hipStream_t stream);

/* ---- kernel selection ------------------------------------------------- */
#define SPTRSV_KERNEL_AUTO 0
#define SPTRSV_KERNEL_SIMPLE 1
#define SPTRSV_KERNEL_ORDER 2
#define SPTRSV_KERNEL_MULTIROW 3
/* MIXED only: the FP32 bulk solve of each refinement step is `sweeps`
 * Jacobi sweeps (d <- D^-1 (r - N d), N the strict lower part) instead of a
 * sync-free substitution, `outer` refinement steps (FP64 residual, FP32
 * sweeps, FP64 update), then an FP64 residual certifies the result on the
 * device: |x - x*|_inf <= |r|_inf / delta with delta = min_i (|L_ii| -
 * sum_j |L_ij|) > 0 (strict row diagonal dominance).  A miss is reported by
 * sptrsv_check as SPTRSV_E_NOT_CONVERGED. */
#define SPTRSV_KERNEL_JACOBI 4

#define SPTRSV_PREC_F64 0    /* plain FP64 sync-free solve */
#define SPTRSV_PREC_F32 1    /* plain FP32 sync-free solve */
#define SPTRSV_PREC_MIXED 2  /* FP32 bulk + FP64 residual/update refinement */

/* One launch configuration of the sync-free solvers.  lanes = lanes per row
 * for SIMPLE/ORDER (1..64, power of two; 0 = from nnz/n as the reference
 * does, solve_csr_simple.cu:121-137); wg = threads per workgroup (256 or
 * 1024, 0 = 256); fused = 1 runs the refinement in the same sync-free pass as
 * the bulk solve, 0 runs separate passes (MIXED only); variant = 0 packs
 * 64/lanes rows per wave (rows of a wave may depend on each other, natural
 * order shares results through LDS), 1 is the lean path where no wave holds
 * two dependent rows (SIMPLE/ORDER: one row per wave; MULTIROW: its
 * single-level waves), one right-hand side only; 2 is a persistent walk of
 * the same rows (block-claimed, the grid capped by cu_waves); 3 is the
 * in-degree walk (SIMPLE/ORDER, one right-hand side): each row waits on a
 * counter of its unsolved dependencies, which solved rows decrement through
 * the column pattern, then folds its entries exactly as variant 1 does (same
 * results, bit for bit); cu_waves caps the waves
 * resident per CU (0 = hardware limit): every resident wave polls, and past a
 * few thousand polling waves the dependency hand-off latency grows several
 * fold, so deep dependency chains run fastest with a small cap.  With lanes
 * = 0 the lanes (and, when variant is 0, the variant) come from the tuned
 * shape of the kernel, else from the default shape; MULTIROW takes its lanes
 * per row from the analysis and ignores a nonzero lanes. */
typedef struct {
    int kernel;
    int lanes;
    int wg;
    int fused;
    int variant;
    int cu_waves;
    int sweeps;   /* JACOBI: FP32 sweeps per refinement step */
    int outer;    /* JACOBI: refinement steps */
    int window;   /* sync-free: entries per lane kept in flight, 1, 4 or 16
                     (16: one right-hand side, wg 256); 0 = by longest row.
                     JACOBI: entries of a row one lane sums (0 = 8) */
    int sweeps_next; /* JACOBI: sweeps of each refinement step after the
                        first (0 = sweeps): later steps solve for a small
                        correction and need less relative accuracy */
    int inner;       /* JACOBI with variant 2 (block sweeps, one right-hand
                        side): Jacobi iterations on each tile's diagonal
                        block per sweep, in LDS (0 = 4) */
    int claim_rounds; /* persistent variant (2): a workgroup claims blocks of
                         waves * claim_rounds work items with one atomic, a
                         wave taking claim_rounds of them in turn (0 = 1) */
    int split_rows;   /* MULTIROW with one right-hand side: 0 or 1 splits each
                         row of more than 256 off-diagonal entries (the row
                         keeps 128, every further run of 256 is folded by a
                         wave of its own and added back in a fixed order);
                         -1 walks whole rows */
    int reorder;      /* ORDER / MULTIROW, FP64 or fused MIXED: 1 solves on a
                         copy of the matrix with its rows stored in the
                         analysis's level order (built on first use, kept
                         with the matrix), so the rows a wave holds are
                         adjacent in memory; b is read and x written in the
                         caller's order by the same kernels, and the results
                         are bit-identical to reorder = 0.  The copy takes
                         16 bytes per entry, is made again for another
                         analysis and dropped by sptrsv_update_values; not
                         with variant 3 or the allocation-free solve */
} sptrsv_config_t;

/* Relative accuracy the JACOBI path certifies (default 1e-10). */
int sptrsv_set_tolerance(sp_mat_t* mat, double rel_tol);
/* Row diagonal-dominance margin delta of the matrix (computed on first use;
 * <= 0 means the JACOBI path cannot certify and is not offered). */
int sptrsv_dominance_margin(sp_mat_t* mat, double* delta);

/* Generic solve with k right-hand sides (row-major n x k, x[i*k + r]).
 * precision selects the arithmetic; for SPTRSV_PREC_F32, b and x are float*,
 * otherwise double*.  refine_steps (MIXED only): number of refinement
 * corrections, >= 1.  cfg may be NULL (AUTO). */
int sptrsv_solve(sp_mat_t* mat, dfr_analysis_info_t* info, int precision,
                 const sptrsv_config_t* cfg, const void* b, void* x, int n,
                 int k, int refine_steps, hipStream_t stream);

/* sptrsv_solve captured once into a hipGraph per (precision, b, x, n, k,
 * refine_steps, resolved configuration) and replayed on `stream`: one
 * launch per solve instead of one per pass (a Jacobi solve has ~30).  The
 * first call for a key runs one eager solve, then captures; later calls
 * only launch.  The matrix owns the graphs (freed with it). */
int sptrsv_solve_graph(sp_mat_t* mat, dfr_analysis_info_t* info, int precision,
                       const sptrsv_config_t* cfg, const void* b, void* x, int n, int k,
                       int refine_steps, hipStream_t stream);
/* Mixed-precision solve (FP32 bulk + FP64 residual and update), k RHS. */
int csr_L_solve_mixed(sp_mat_t* mat, dfr_analysis_info_t* info, const double* b,
                      double* x, int n, int k, int refine_steps, hipStream_t stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the fused mixed-precision solve of sptrsv_solve
 * (SPTRSV_PREC_MIXED, fused = 1, refine_steps = 1), with its published
 * words in a fixed per-device scratch array of SPTRSV_NOALLOC_MAX_WORDS
 * 64-bit words (32 MiB, held by every GPU the library runs on until the
 * process ends) instead of the matrix's workspace.  It calls no hipMalloc,
 * hipFree or other allocating routine.  n * k above the maximum is
 * SPTRSV_E_CAPACITY; a MULTIROW solve that splits long rows (one right-hand
 * side) also keeps its segment partials there, after the published words,
 * and must fit both.  Sync-free kernels only: a named JACOBI configuration
 * is SPTRSV_E_INVALID_ARG, as is a handle stored in reversed order
 * (allocate_memory_upper, or the transpose of a lower-triangular handle), the
 * in-degree walk (variant 3) and a named level-ordered copy (reorder = 1).
 * An automatic call (cfg NULL or kernel AUTO) takes the tuned shape, with
 * two substitutions: a tuned JACOBI choice becomes the default sync-free
 * shape, and a tuned level-ordered copy walks the matrix as given (same
 * bits).
 * Each call re-initialises the scratch it uses.  Every call on one GPU
 * shares that GPU's scratch, so calls that overlap in time on one GPU (two
 * streams, two threads) race and must be serialised by the caller. */
#define SPTRSV_NOALLOC_MAX_WORDS (1 << 22)
int sptrsv_solve_mixed_noalloc(sp_mat_t* mat, dfr_analysis_info_t* info, const sptrsv_config_t* cfg,
                               const double* b, double* x, int n, int k, hipStream_t stream);

/* Allocates (outside any capture) the workspaces the solves above use for k
 * right-hand sides, so later calls allocate nothing. */
int sptrsv_reserve(sp_mat_t* mat, int k);

/* Waits for the stream and reports SPTRSV_E_WATCHDOG if any dependency wait
 * of an earlier solve on this matrix exceeded its bound (the rows involved
 * then hold NaN); clears the flag.  Wait bound in milliseconds: default
 * 20000, environment SPTRSV_WATCHDOG_MS overrides. */
int sptrsv_check(sp_mat_t* mat, hipStream_t stream);
int sptrsv_set_watchdog_ms(sp_mat_t* mat, double ms);

/* Polling back-off of the sync-free waits: 0 spins (default), 1..4 sleep
 * about 64, 128, 256, 512 cycles between polling turns without progress. */
int sptrsv_set_backoff(sp_mat_t* mat, int level);

/* Extra LDS bytes (0..65536) reserved by every solve workgroup, which caps
 * the workgroups resident per CU (160 KB of LDS each): fewer waves polling
 * ahead of the dependency front.  0 (default) = no cap. */
int sptrsv_set_lds_pad(sp_mat_t* mat, int bytes);

/* Per-row timeline of the next solves on this matrix (profiling aid): while
 * enabled, each solve writes three wall-clock stamps per row and right-hand
 * side -- wave start, dependencies folded and reduced, result stored -- to a
 * device buffer of 3*n*k int64.  _read copies count entries to host (ticks of
 * the device wall clock; rate in kHz returned in *rate_khz). */
int sptrsv_trace(sp_mat_t* mat, int k, int enable);
int sptrsv_trace_read(sp_mat_t* mat, int64_t* host, size_t count, double* rate_khz);

/* Times every candidate configuration on this matrix (and the given
 * precision) and stores the fastest in the matrix, where AUTO finds it.
 * best may be NULL.  reps >= 1 timed solves per candidate. */
int sptrsv_autotune(sp_mat_t* mat, dfr_analysis_info_t* info, int precision,
                    int k, int reps, sptrsv_config_t* best, float* best_ms);
/* The heuristic choice used when no tuning result is stored. */
int sptrsv_default_config(const sp_mat_t* mat, const dfr_analysis_info_t* info,
                           int precision, int k, sptrsv_config_t* out);

/* ---- batched independent systems (one persistent launch) -------------- */
typedef struct sptrsv_batch sptrsv_batch_t;

/* Host CSR arrays of nmat independent lower-triangular systems. */
int sptrsv_batch_create(int nmat, const int* n, const int* const* ia,
                        const int* const* ja, const double* const* a,
                        sptrsv_batch_t** out);
int sptrsv_batch_rows(const sptrsv_batch_t* bt, int64_t* total_rows, int64_t* total_nnz);
/* Per-matrix kernel choice: tune = 1 times the candidates on the device,
 * tune = 0 uses the feature heuristic. */
int sptrsv_batch_tune(sptrsv_batch_t* bt, int precision, int tune);
/* Device b / x concatenated over matrices (matrix m starts at row_off[m]). */
int sptrsv_batch_solve(sptrsv_batch_t* bt, int precision, const double* b, double* x,
                       hipStream_t stream);
/* Captures the whole batched solve in a hipGraph once, then replays it. */
int sptrsv_batch_solve_graph(sptrsv_batch_t* bt, int precision, const double* b,
                             double* x, hipStream_t stream);
int sptrsv_batch_row_offsets(const sptrsv_batch_t* bt, int64_t* row_off /* nmat+1 */);
int sptrsv_batch_choice(const sptrsv_batch_t* bt, int m, sptrsv_config_t* cfg);
/* How a MIXED batched solve runs (set by sptrsv_batch_tune, or by the
 * caller after tuning): serial 1 runs the Jacobi members, then the sync-free
 * ones, on the caller's stream; 0 runs them concurrently on two streams.
 * separate 1 solves the sync-free members as one matrix of their own, with
 * its own level analysis and tuned kernel (rows gathered from and scattered
 * back to the batch order); 0 launches them as ticketed tasks. */
int sptrsv_batch_schedule(const sptrsv_batch_t* bt, int* serial, int* separate);
int sptrsv_batch_set_schedule(sptrsv_batch_t* bt, int serial, int separate);
/* Members sharing a structure (same n, row lengths and columns; values may
 * differ): *n_patterns distinct structures, and pattern_of[m] (nmat entries,
 * may be NULL) the first member with m's structure.  sptrsv_batch_tune times
 * the sync-free launch once per structure and gives its lanes choice to every
 * member of that structure (the work, hence the time, depends on the
 * structure only); value-dependent choices (Jacobi) stay per member. */
int sptrsv_batch_patterns(const sptrsv_batch_t* bt, int* n_patterns, int* pattern_of);
/* sptrsv_check for the batch: watchdog, or a Jacobi member that missed its
 * certified bound (each member is certified separately). */
int sptrsv_batch_check(sptrsv_batch_t* bt, hipStream_t stream);
void sptrsv_batch_destroy(sptrsv_batch_t* bt);

/* ---- vendor baseline (replaces csr_L_solve_cusparse_v2, solver.h:8) ---- */
typedef struct sptrsv_rocsparse sptrsv_rocsparse_t;
int csr_L_rocsparse_create(sp_mat_t* mat, sptrsv_rocsparse_t** out, float* analysis_ms);
int csr_L_solve_rocsparse(sptrsv_rocsparse_t* h, const double* b, double* x, hipStream_t stream);
/* rocSPARSE csritsv (FP64 Jacobi with a residual stopping test, host
 * driven) as a second baseline: x starts at 0; iterates until
 * |b - L x|_inf <= tol or max_iter stopping tests (free_iter untested
 * iterations before each test); *iters gets the Jacobi iterations performed,
 * (k + 1) * free_iter + k for the k rocSPARSE reports. */
int csr_L_rocsparse_itsv_create(sp_mat_t* mat, sptrsv_rocsparse_t** out, float* analysis_ms);
int csr_L_solve_rocsparse_itsv(sptrsv_rocsparse_t* h, const double* b, double* x, double tol,
                               int max_iter, int free_iter, int* iters, hipStream_t stream);
void csr_L_rocsparse_destroy(sptrsv_rocsparse_t* h);

/* ---- power (replaces nvmlAPIRun/nvmlAPIEnd, nvmlPower.hpp:17-18) ------ */
/* Starts a sampling thread that reads the socket power of GPU `device`
 * through amdsmi every period_us microseconds.  Returns
 * SPTRSV_E_UNAVAILABLE when amdsmi cannot be loaded or read. */
int amdsmiAPIRun(const char* label, int device, int period_us);
/* Stops the thread; mean power in mW over the samples (< 0 on error). */
float amdsmiAPIEnd(void);
int amdsmiAPISamples(void);

/* ---- timing / bench (BENCH_RUN_SOLVE, common.h:103-124) --------------- */
typedef struct {
    float mean_ms;
    float stdev_ms;
    float min_ms;
    float power_mw;   /* < 0 when power could not be sampled */
    int errors;       /* validate_x count (x != 1), -1 when not checked */
    int reps;
} sptrsv_bench_t;

/* Runs the selected solve reps times on device b/x (each rep re-clears x),
 * times each with hip events, samples power around the loop and, when
 * check_ones != 0, copies x back and counts x[i] != 1 (validate_x,
 * matrix_properties.h:7-29). */
int sptrsv_bench_solve(sp_mat_t* mat, dfr_analysis_info_t* info, int precision,
                       const sptrsv_config_t* cfg, const void* b, void* x, int n,
                       int k, int refine_steps, int reps, int check_ones,
                       sptrsv_bench_t* out);

/* ---- Matrix Market (header/mmio.h:25-52) ------------------------------ */
typedef char MM_typecode[4];
#define MM_MAX_LINE_LENGTH 1025
#define MM_MAX_TOKEN_LENGTH 64
#define mm_is_matrix(t) ((t)[0] == 'M')
#define mm_is_sparse(t) ((t)[1] == 'C')
#define mm_is_coordinate(t) ((t)[1] == 'C')
#define mm_is_dense(t) ((t)[1] == 'A')
#define mm_is_array(t) ((t)[1] == 'A')
#define mm_is_complex(t) ((t)[2] == 'C')
#define mm_is_real(t) ((t)[2] == 'R')
#define mm_is_pattern(t) ((t)[2] == 'P')
#define mm_is_integer(t) ((t)[2] == 'I')
#define mm_is_symmetric(t) ((t)[3] == 'S')
#define mm_is_general(t) ((t)[3] == 'G')
#define mm_is_skew(t) ((t)[3] == 'K')
#define mm_is_hermitian(t) ((t)[3] == 'H')
/* mmio.h:25-27 */
int mm_read_banner(FILE* f, MM_typecode* matcode);
int mm_read_mtx_crd_size(FILE* f, int* M, int* N, int* nz);
int mm_is_valid(MM_typecode matcode);
/* Writes a 4-letter code as text into buf (>= 64 bytes); returns buf. */
char* mm_typecode_to_str_r(const MM_typecode matcode, char* buf);

/* main.cu:66-246 as a library call: reads a coordinate .mtx file (real,
 * integer or pattern; general, symmetric or hermitian; complex and
 * non-square rejected), builds A's CSR in file order with symmetric
 * mirroring, then extracts the lower triangle with an explicit diagonal.
 * unit != 0 gives the reference's all-ones values (main.cu:228,236).
 * *ia, *ja, *a are malloc'ed; release with sptrsv_free_host. */
int sptrsv_read_mtx_lower(const char* path, int unit, int* n, int* nnz,
                          int** ia, int** ja, double** a);
/* The entry section of main.cu:131-155 read by nthreads threads (<= 0: the
 * host's cores, at most 16) over line-aligned chunks: 0-based COO in file
 * order, not mirrored; pattern files give 1.0; *symmetric is 1 for symmetric
 * and hermitian files.  Same entries, errors and values as a sequential read.
 * Arrays are malloc'ed (sptrsv_free_host). */
int sptrsv_read_mtx_coo(const char* path, int nthreads, int* n, int* nz, int* symmetric,
                        int** row, int** col, double** val);
/* main.cu:124-246 on the device: COO (host arrays, file order) -> CSR of A
 * with symmetric mirroring -> lower triangle with an explicit diagonal, the
 * same matrix sptrsv_read_mtx_lower builds on the host, uploaded as *gpu_L
 * (an allocate_memory handle). */
int sptrsv_coo_to_lower(const int* row, const int* col, const double* val, int nz, int n,
                        int symmetric, int unit, sp_mat_t* gpu_L);
/* Binary CSR cache (no counterpart in the reference): magic, sizes and an
 * FNV-1a checksum, then ia, ja, a.  A load that fails the checksum is
 * SPTRSV_E_IO.  Loaded arrays are malloc'ed (sptrsv_free_host). */
int sptrsv_save_csr(const char* path, int n, const int* ia, const int* ja, const double* a);
int sptrsv_load_csr(const char* path, int* n, int* nnz, int** ia, int** ja, double** a);
void sptrsv_free_host(void* p);

/* ---- synthetic inputs of the benchmark workloads ----------------------- */
/* Host arrays malloc'ed by the library (sptrsv_free_host).  b is n*k.
 *  id 1: 2D 5-point Laplacian lower triangle, 64x64 (param = grid side),
 *        diag 4, off -1, b = L*1 (x = 1 exactly), k = 1
 *  id 2: random lower, n = param (1e6), Poisson(8) off-diagonals in
 *        [max(0,i-5000), i), U(-1,1), diag 1+sum|off|, b ~ U(-1,1)
 *  id 3: power-law lower, n = param (2e6), Zipf(1.8) off-diagonal count
 *        capped at min(2000, i), columns in [0,i), diag 1+sum|off|
 *  id 4: 3D 7-point lower, side = param (128), unit diagonal, off -1/6,
 *        k = 16 RHS U(-1,1)
 * Deterministic for a given (id, param, seed) and any thread count. */
int sptrsv_generate(int id, int param, uint64_t seed, int* n, int* nnz,
                    int** ia, int** ja, double** a, int* k, double** b);
/* id 5 member m of a batch: n in [nmin,nmax], even m banded (bandwidth
 * U[1,32]), odd m random (Poisson(6) off-diagonals in [0,i)). */
int sptrsv_generate_batch_member(int m, int nmin, int nmax, uint64_t seed, int* n,
                                 int* nnz, int** ia, int** ja, double** a, double** b);

/* ---- device memory helpers for callers without a HIP runtime --------- */
int sptrsv_device_count(int* count);
int sptrsv_device_name(int device, char* buf, int len);
int sptrsv_malloc(void** p, size_t bytes);
int sptrsv_free(void* p);
int sptrsv_memcpy_htod(void* dst, const void* src, size_t bytes);
int sptrsv_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int sptrsv_memset(void* dst, int value, size_t bytes);
int sptrsv_synchronize(void);
/* Streaming device-to-device copy (bytes and pointers 16-byte aligned) with
 * the library's own kernel: a known-size reference for traffic counters. */
int sptrsv_copy_device(void* dst, const void* src, size_t bytes);
/* Timed region helpers (hip events on `stream`): start records, stop
 * records and waits, returning milliseconds since start. */
int sptrsv_timer_start(hipStream_t stream);
int sptrsv_timer_stop(hipStream_t stream, float* ms);
/* Writes 512 MB of scratch to evict the 256 MB Infinity Cache. */
int sptrsv_flush_cache(hipStream_t stream);

/* Kernel timing: while enabled, every sync-free solve kernel launch is
 * bracketed by hip events on its own stream, and the Jacobi sweeps of each
 * refinement step by one pair around the step's consecutive sweeps.  enable:
 * 0 off, 1 both, 2 sync-free launches only, 3 Jacobi sweeps only.  _read
 * waits for them and returns the summed duration and the number of kernel
 * launches covered, then clears. */
int sptrsv_kernel_timer(int enable);
int sptrsv_kernel_timer_read(float* total_ms, int* launches);

/* One-way latency (ns) of a dependency hand-off between two workgroups on
 * different XCDs (agent-scope store -> polling load), median of `rounds`
 * ping-pong rounds; the t_dep of the latency-bound fraction nlevs*t_dep/t. */
int sptrsv_measure_dep_latency(int rounds, float* one_way_ns);

#ifdef __cplusplus
}
#endif
#endif
