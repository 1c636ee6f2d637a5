/*
 * sptrsv_capi.h -- C ABI of the MI355X (gfx950) sparse upper-triangular solver.
 *
 * Plain C types only (pointers, sizes, int status codes).  The C++ drop-in headers
 * (core.h, utils.h, sptrsv_syncfree_opencl.h) are thin wrappers over these calls, and
 * any FFI (ctypes, cgo, JNI) can bind them directly.
 *
 * Model: a *plan* owns one upper-triangular U resident in device memory (CSR, CSC,
 * level-ordered CSR, analysis arrays).  Solves take DEVICE pointers to b and x and run
 * asynchronously on the plan's stream; sptrsv_plan_sync() waits and reports device-side
 * failures (watchdog).  sptrsv_solve_host() is the synchronous host-pointer convenience.
 *
 * Layouts: CSR rows hold the diagonal FIRST and ascending columns (the reference's CPU
 * convention, src/core.cpp:18-32); CSC columns hold ascending rows with the diagonal LAST
 * (src/utils.cpp:128-159).  Multi-RHS device arrays are row-major n x m (x[i*m + k], the
 * reference's GPU SpTRSM layout, src/kernel_full.cl:163) unless a call says otherwise.
 *
 * Every function returns SPTRSV_OK (0) or a negative SPTRSV_ERR_* code; the matching
 * message is available from sptrsv_last_error() on the calling thread.
 */
#ifndef SPTRSV_CAPI_H
#define SPTRSV_CAPI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPTRSV_OK 0
#define SPTRSV_ERR_BAD_INPUT -1   /* malformed matrix or argument (nothing launched) */
#define SPTRSV_ERR_HIP -2         /* HIP runtime error */
#define SPTRSV_ERR_DEADLOCK -3    /* device watchdog: a dependency wait exceeded its limit */
#define SPTRSV_ERR_IO -4          /* file missing / unknown extension / short read */
#define SPTRSV_ERR_NOMEM -5       /* device or host allocation failed */
#define SPTRSV_ERR_UNSUPPORTED -6 /* size or option outside what the kernels support */
#define SPTRSV_ERR_NO_DEVICE -7   /* no HIP device visible */

/* Solver variants (kernel families). */
#define SPTRSV_VAR_AUTO 0         /* selector picks */
#define SPTRSV_VAR_CSC_PUSH 1     /* wave per CSC column, in-degree counters + atomic left_sum
                                     (the reference's algorithm, src/kernel_full.cl:54-228) */
#define SPTRSV_VAR_ROW_PULL 2     /* lanes per CSR row in natural (descending) order, x published
                                     through its own value; exact reference summation order */
#define SPTRSV_VAR_LEVEL_PULL 3   /* same kernel over level-ordered rows (level-permuted CSR) */
#define SPTRSV_VAR_LEVELSET 4     /* bulk-synchronous: one launch per level (src/core.cpp:273-376) */
#define SPTRSV_VAR_CHAIN 5        /* one wave walks 64-row windows in order: window tile in LDS,
                                     64 broadcast steps per window, recent x in an LDS ring; for
                                     factors whose dependency graph is (nearly) a chain; m == 1 */
#define SPTRSV_VAR_SUPERNODAL 6   /* one workgroup per supernode: dense blocks on MFMA, the
                                     diagonal block applied through its precomputed inverse
                                     (needs sptrsv_plan_build_supernodes) */
#define SPTRSV_VAR_TILE 7         /* one workgroup per tile of T consecutive rows (params.block;
                                     0 = 16384) with the tile's x in LDS: the tile's rows run level
                                     by level with a barrier between levels, dependencies inside the
                                     tile read from LDS, later tiles polled; m == 1, FP64, exact
                                     reference summation order */
#define SPTRSV_VAR_BLOCK 8        /* blocks of rows in a topological order, each solved in one
                                     workgroup: 8x8x8 / 32x16 bricks of a natural-order grid factor
                                     (8 compute waves), else 64-row chunks of bands of 16 levels;
                                     dependencies inside a block through LDS, on other blocks
                                     polled by a poller wave; params.occupancy = resident blocks
                                     per CU (0 = 3 for grid bricks, 4 otherwise); FP64, m <= 4 on
                                     grid factors, m == 1 otherwise; exact reference summation
                                     order */
#define SPTRSV_VAR_PANEL 9        /* many right-hand sides on a natural-order grid factor: one
                                     workgroup per (8x8x8 / 16x16 brick, panel of 16 columns of
                                     X), bricks in dependency order; in-brick dependencies from
                                     LDS, the upstream bricks' face cells read once after their
                                     epoch flags; no sentinel fill of x; FP64, any m; exact
                                     reference summation order */
#define SPTRSV_VAR_COUNT 10

/* Value precision of the solve. */
#define SPTRSV_FP64 0
#define SPTRSV_FP32 1
#define SPTRSV_BF16 2  /* supernodal variant only: BF16 rectangular products on MFMA with FP32
                          accumulation and a power-of-two scale per supernode block; FP32 triangle */

typedef struct {
    int variant;      /* SPTRSV_VAR_* */
    int group;        /* lanes per row for the pull kernels (1,2,4,...,64); 0 = derive from m
                         (m lanes up to 64). m == 1: lanes splitting a row's entries; LEVEL_PULL
                         with m > 1: 8..64 lanes, each taking up to 8 columns */
    int unroll;       /* entries polled per step in the pull kernels (1, 2 or 4); 0 = default */
    int precision;    /* SPTRSV_FP64 / SPTRSV_FP32 */
    int watchdog_us;  /* per-wait device watchdog in microseconds; 0 = default (20 s) */
    int occupancy;    /* pull kernels: at most this many resident workgroups per CU; < 0 = the
                         hardware limit; 0 = default: LEVEL_PULL about three mean level widths of
                         rows in flight, ROW_PULL the hardware limit */
    int block;        /* pull kernels: threads per workgroup (64..1024, power of two); 0 =
                         default: LEVEL_PULL 64 rows per workgroup, ROW_PULL 256 / 1024 threads.
                         TILE: rows per tile (1..16384; 0 = 16384) */
    int exact;        /* AUTO only: nonzero restricts the choice to kernels whose FP64 x is
                         bitwise equal to the serial solver's (BLOCK, LEVEL_PULL): the drop-in
                         replacements of base_gauss_upper use it */
} sptrsv_params_t;

typedef struct sptrsv_plan_s* sptrsv_plan_t;

/* ---- device / memory ---------------------------------------------------------------- */
const char* sptrsv_last_error(void);
const char* sptrsv_status_string(int code);
int sptrsv_device_init(int device);
int sptrsv_device_info(char* name, int name_len, int* compute_units, size_t* total_mem);
int sptrsv_device_sync(void);
int sptrsv_malloc(void** ptr, size_t bytes);
int sptrsv_free(void* ptr);
int sptrsv_memcpy_h2d(void* dst, const void* src, size_t bytes);
int sptrsv_memcpy_d2h(void* dst, const void* src, size_t bytes);
int sptrsv_memcpy_d2d(void* dst, const void* src, size_t bytes);
/* Self-test of the kernels' division by a reused reciprocal (no reference counterpart): for
 * device arrays n, d of `count` doubles, fast[i] = the kernels' n[i] / d[i] with the reciprocal
 * of d[i] computed once, ieee[i] = the compiled division n[i] / d[i]; the two must be bitwise
 * equal for every input.  Synchronous. */
int sptrsv_division_check(const double* n, const double* d, double* fast, double* ieee, int64_t count);

/* ---- host-side validation (no device needed) ---------------------------------------- */
/* Checks an upper CSR: rowptr[0]==0, monotone, rowptr[n]==nnz, first entry of each row is
 * its diagonal (finite, non-zero), remaining columns strictly ascending and < n. */
int sptrsv_validate_csr(int n, int64_t nnz, const int64_t* rowptr, const int* col, const double* val);
/* Checks an upper CSC: colptr monotone, rows strictly ascending, last entry of each column
 * is its diagonal (finite, non-zero). */
int sptrsv_validate_csc(int n, int64_t nnz, const int* colptr, const int* rowidx, const double* val);

/* ---- format conversion (device) ------------------------------------------------------- */
/* CSR (any pattern, n x n) -> CSC with rows ascending in every column: the reference's
 * transpose (src/utils.cpp:128-159) as a stable device radix sort.  Host pointers. */
int sptrsv_csr_to_csc(int n, int64_t nnz, const int64_t* rowptr, const int* col, const double* val,
                      int* colptr /* n+1 */, int* rowidx /* nnz */, double* cval /* nnz */);

/* ---- plans ---------------------------------------------------------------------------- */
/* Uploads U (host CSR, validated), builds on the device: CSC (stable transpose), in-degree,
 * levels, level sets and the level-ordered CSR.  nnz must be < 2^31. */
int sptrsv_plan_create_csr(int n, int64_t nnz, const int64_t* rowptr, const int* col,
                           const double* val, sptrsv_plan_t* plan);
/* Same from a host CSC (diagonal last) -- the reference GPU entry's input. */
int sptrsv_plan_create_csc(int n, int64_t nnz, const int* colptr, const int* rowidx,
                           const double* val, sptrsv_plan_t* plan);
int sptrsv_plan_destroy(sptrsv_plan_t plan);
int sptrsv_plan_info(sptrsv_plan_t plan, int* n, int64_t* nnz, int* n_levels);
/* Host wall time of the plan's analysis, in ms (no reference counterpart; the reference counts
 * its analysis inside barrier_upper, src/core.cpp:375): transpose, in-degree, levels, level sets
 * and level-ordered CSR, i.e. everything sptrsv_plan_create_csr does after the upload.
 * levels_ms: the level computation alone. */
int sptrsv_plan_analysis_time(sptrsv_plan_t plan, double* analysis_ms, double* levels_ms);

/* Supernodal layout.  snodes[0..sn] are row boundaries (reference .snodes semantics,
 * src/utils.cpp:84-111); every supernode's rows must hold a dense triangle followed by one
 * common set of columns (the structure src/core.cpp:53-78 assumes).  sn <= 0: detect the
 * fundamental supernodes from the pattern.  Supernodes wider than 128 rows are split (the
 * split keeps the structure).  Diagonal blocks are inverted on the device. */
int sptrsv_plan_build_supernodes(sptrsv_plan_t plan, int sn, const int* snodes);
/* Relaxed supernodes (no reference counterpart): consecutive rows are merged while the block,
 * padded with explicit zeros to a dense triangle plus one common column set, stays within
 * max_width rows (0 -> 128) and (1 + max_fill) times the rows' own entries (0 -> 1.0).  Any
 * factor qualifies; a banded factor becomes a chain of wide dense blocks solved on MFMA.  AUTO
 * builds this layout by itself, once, for a chain-like factor with at least 16 entries per row
 * on average when the blocks average at least 32 rows (and then solves FP64 through it). */
int sptrsv_plan_build_supernodes_relaxed(sptrsv_plan_t plan, int max_width, double max_fill);
/* Explicit zeros stored by the current supernodal layout (0 for exact supernodes). */
int sptrsv_plan_supernodal_padding(sptrsv_plan_t plan, int64_t* explicit_zeros);
/* Number of supernodes in use and (optionally, sn+1 entries) their boundaries. */
int sptrsv_plan_get_supernodes(sptrsv_plan_t plan, int* sn, int* snodes);

/* The reference's padded supernodal layout (ccs2ccs_pad, src/core.cpp:136-161) built on the
 * device from the plan's CSR: row d of every supernode gets d zeros at the supernode's leading
 * columns, so each supernode is a dense trapezoid.  snodes: sn + 1 host ints from 0 to n.
 * Outputs are device arrays: d_val_pad, d_col_pad (uint64, the reference's type) of *nnz_pad
 * entries, d_rowptr_pad n + 1 ints; *nnz_pad = nnz + sum_k w_k (w_k - 1) / 2.  With all three
 * outputs null, only *nnz_pad is computed.  Synchronous. */
int sptrsv_ccs2ccs_pad_device(sptrsv_plan_t plan, int sn, const int* snodes, double* d_val_pad,
                              uint64_t* d_col_pad, int* d_rowptr_pad, int64_t* nnz_pad);

/* Analysis results (device-computed, copied to host). */
int sptrsv_plan_get_indegree(sptrsv_plan_t plan, int* indeg /* n */);
int sptrsv_plan_get_levels(sptrsv_plan_t plan, int* levels /* n, 1-based */);
/* Block order of the BLOCK variant (no reference counterpart), built here if no block solve
 * built it yet: *n_blocks blocks, *kind 1 = bricks of a natural-order grid factor, 2 = level
 * bands; block_ptr (n_blocks + 1 positions) and perm (position -> row, n) may be null. */
int sptrsv_plan_get_blocks(sptrsv_plan_t plan, int* n_blocks, int* kind, int* block_ptr, int* perm);
int sptrsv_plan_get_level_sets(sptrsv_plan_t plan, int* level_ptr /* n_levels+1 */,
                               int* level_rows /* n */);
int sptrsv_plan_get_csc(sptrsv_plan_t plan, int* colptr /* n+1 */, int* rowidx /* nnz */,
                        double* val /* nnz */);

/* The variant SPTRSV_VAR_AUTO runs for m right-hand sides at `precision` (no reference
 * counterpart).  With a selector network installed (sptrsv_selector_set), its choice.  Else, in
 * this order: SUPERNODAL in FP64 on a plan with supernodes -- the caller's
 * (sptrsv_plan_build_supernodes), or, for a factor of 16 or more entries per row, its own
 * fill-free supernodes or else relaxed ones when they average 32 rows or more (built here on
 * first use); CHAIN for m == 1 on a chain-like factor (levels hold < 4 rows on average); BLOCK
 * for FP64 m <= 3 on a natural-order grid factor; PANEL for FP64 m >= 32 on one; LEVEL_PULL
 * otherwise.  With params.exact only the bitwise kernels are chosen.  For 2 <= m <= 3 on a
 * grid factor the choice checks the block layout's limits and so builds that layout (on the
 * host, once per plan; the block solve would build it on its first call). */
int sptrsv_plan_auto_variant(sptrsv_plan_t plan, int m, int precision, int* variant);
/* Learned choice of the variant for AUTO (no reference counterpart).  The 24 features of the
 * plan's factor at m right-hand sides (the batched selector's: log2 n, log2 nnz, row-length
 * and level-width statistics, offsets, supernodes, log2 m, diagonal dominance); computed once
 * per plan.  sptrsv_selector_set installs a 24 -> h1 -> h2 -> out ReLU network (parameters in
 * the sptrsv_mlp_* layout, inputs normalised as (f - mu) / sd, output k scoring variant
 * classes[k]) that AUTO then consults for every plan: it runs the best-scored variant valid for
 * the plan, m and precision (exact solves: only the bitwise ones); without a network, or when
 * no scored variant is valid, AUTO uses its hand rule.  sptrsv_selector_clear removes it. */
int sptrsv_plan_features(sptrsv_plan_t plan, int m, float* features);
int sptrsv_selector_set(int in, int h1, int h2, int out, const float* params, const float* mu, const float* sd,
                        const int* classes);
int sptrsv_selector_clear(void);

/* Solve U X = B for m right-hand sides.  d_b, d_x: device, row-major n x m (FP64), not
 * overlapping (most kernels mark x's unknown values in x itself before they solve).
 * Asynchronous on the plan's stream. */
int sptrsv_solve(sptrsv_plan_t plan, const sptrsv_params_t* params, const double* d_b,
                 double* d_x, int m);
/* No-allocation variant of sptrsv_solve with SPTRSV_VAR_CSC_PUSH, FP64.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  The push kernel's work memory
 * (left_sum n*m doubles, consumed in-degrees n ints, ticket and status words) lives in a fixed
 * array inside the library, one per GPU, held until the process ends; the call itself allocates
 * nothing.  Limits: n <= SPTRSV_NOALLOC_MAX_ROWS and n*m <= SPTRSV_NOALLOC_MAX_ELEMS, else
 * SPTRSV_ERR_UNSUPPORTED with nothing launched.  Every call on one GPU shares the array: calls
 * that overlap on one GPU race (keep them in one stream order).  Asynchronous on the plan's
 * stream; sptrsv_plan_sync reports a watchdog stop as for sptrsv_solve. */
#define SPTRSV_NOALLOC_MAX_ROWS (1 << 21)
#define SPTRSV_NOALLOC_MAX_ELEMS (1 << 21)
int sptrsv_solve_push_noalloc(sptrsv_plan_t plan, const double* d_b, double* d_x, int m);
/* Waits for the plan's stream; returns SPTRSV_ERR_DEADLOCK if a watchdog fired. */
int sptrsv_plan_sync(sptrsv_plan_t plan);
/* Host pointers, synchronous.  col_major != 0: b and x are column-major (x[i + k*n]). */
int sptrsv_solve_host(sptrsv_plan_t plan, const sptrsv_params_t* params, const double* b,
                      double* x, int m, int col_major);

/* HIP-event timing of the main solve kernel of every solve while enabled. */
int sptrsv_plan_timing(sptrsv_plan_t plan, int enable);
int sptrsv_plan_timing_read(sptrsv_plan_t plan, double* total_ms, int* launches);

/* Forward substitution with the transpose: solve U^T X = B (lower triangular) on the same
 * plan, m right-hand sides, device row-major n x m, asynchronous on the plan's stream.  The
 * reference's forward solve base_gauss_lower (src/core.cpp:4-16) takes L in CSR with the
 * diagonal last; L = U^T is exactly that layout of U's CSC copy.  A companion plan of
 * J U^T J (J reverses the index order, which makes it upper triangular) is built on the first
 * call and kept; params apply to its solve (SUPERNODAL is refused: that layout is built for U).
 * Sums run in descending column order of L, so x agrees with base_gauss_lower to rounding. */
int sptrsv_solve_transposed(sptrsv_plan_t plan, const sptrsv_params_t* params, const double* d_b,
                            double* d_x, int m);
/* Full Cholesky solve U^T U X = B: the transposed solve, then the solve with U (params apply
 * to both; SUPERNODAL only to the second).  No reference counterpart: the reference solves
 * one triangle per call. */
int sptrsv_solve_cholesky(sptrsv_plan_t plan, const sptrsv_params_t* params, const double* d_b,
                          double* d_x, int m);

/* The reference's loaders (src/utils.cpp:9-111) through the C ABI: .matrix (binary CSR),
 * .mtx (Matrix Market lower factor, read as its upper transpose) and .txt matrices, .snodes and
 * .txt supernode files.  Arrays are allocated with malloc; release them with sptrsv_free_host. */
int sptrsv_read_matrix_file(const char* path, int* n, int64_t* nnz, int64_t** rowptr, int** col,
                            double** val);
int sptrsv_read_snodes_file(const char* path, int* sn, int** snodes /* sn + 1 */);
void sptrsv_free_host(void* p);

/* Residual r = b - U x on the device, FP64, row-major n x m. */
int sptrsv_residual(sptrsv_plan_t plan, const double* d_b, const double* d_x, double* d_r, int m);

/* Mixed precision with FP64 iterative refinement (no reference counterpart).
 * x_0 = solve_low(b); then repeat  r = b - U x (FP64 residual on the GPU),  d = solve_low(r),
 * x += d (FP64).  With r_k = max over columns of ||d_k||_inf / ||x_k||_inf (r_0 = 1), the error
 * left after correction k is estimated as r_k * min(1, r_k / r_{k-1}); the loop stops when that
 * estimate is <= tol, when a correction shrinks by less than half (k >= 2: the FP64 residual
 * floor), or after max_iters corrections.  solve_low is the kernel chosen by params->variant in
 * opts->precision (FP32; BF16 only with the supernodal variant).  Synchronous; *iters =
 * corrections applied, *ratio = the error estimate of the returned x (converged when <= tol).
 * Reaching max_iters or the floor without meeting tol is not an error: check *ratio. */
typedef struct {
    int max_iters;    /* 0 -> 30 */
    int precision;    /* SPTRSV_FP32 or SPTRSV_BF16 */
    double tol;       /* 0 -> 1e-13 (estimated relative error of x) */
    int fallback;     /* nonzero: when the iteration ends without meeting tol, x is recomputed
                         by the FP64 solve (AUTO variant, or SUPERNODAL for a supernodal one) */
    int reserved;
} sptrsv_refine_t;
int sptrsv_solve_refined(sptrsv_plan_t plan, const sptrsv_params_t* params, const sptrsv_refine_t* opts,
                         const double* d_b, double* d_x, int m, int* iters, double* ratio);

/* How a refined solve ended (no reference counterpart). */
#define SPTRSV_REFINE_CONVERGED 0  /* estimate <= tol */
#define SPTRSV_REFINE_FLOOR 1      /* corrections stopped shrinking by half (FP64 residual floor,
                                      or a contraction too slow for the inner precision) */
#define SPTRSV_REFINE_MAX_ITERS 2  /* max_iters corrections applied, estimate still > tol */
#define SPTRSV_REFINE_DIVERGED 3   /* a correction or x not finite, or corrections growing:
                                      the inner solve cannot resolve this factor (kappa * u >= 1) */
typedef struct {
    int iters;        /* corrections applied */
    int outcome;      /* SPTRSV_REFINE_* of the reduced-precision iteration */
    int fell_back;    /* 1: x comes from the FP64 fallback solve */
    int variant;      /* kernel family of the reduced-precision solves */
    double estimate;  /* error estimate of the last reduced-precision iterate (+inf when a norm
                         was not finite) */
} sptrsv_refine_info_t;
/* sptrsv_solve_refined with the outcome reported (info may be NULL). */
int sptrsv_solve_refined_ex(sptrsv_plan_t plan, const sptrsv_params_t* params, const sptrsv_refine_t* opts,
                            const double* d_b, double* d_x, int m, sptrsv_refine_info_t* info);

/* 1-norm condition estimate kappa_1(U) = ||U||_1 * est(||U^-1||_1) (no reference counterpart):
 * Hager's method as refined by Higham (LAPACK xLACON's iteration: at most 5 pairs of FP64
 * solves with U and U^T on the device, plus the alternating-sign test vector).  Synchronous.
 * *norm_inv receives the estimate of ||U^-1||_1 (a lower bound, usually exact or within a
 * small factor); either output may be NULL. */
int sptrsv_condest(sptrsv_plan_t plan, double* kappa, double* norm_inv);

/* ---- batches of independent systems (no reference counterpart) -------------------------
 * The systems are given as one block-diagonal upper CSR: system s owns rows
 * sys_ptr[s] .. sys_ptr[s+1]-1 and only references its own columns.  Row pointers are 64-bit
 * end to end (a batch may hold more than 2^31 nonzeros); column indices are global rows (int).
 * Creation validates every system, computes its levels and selector features on the host
 * (systems in parallel) and uploads the matrix.  Every system is solved by
 * one persistent workgroup with its own configuration (SPTRSV_BCFG_*), systems popped from a
 * queue in a caller-given order; sptrsv_batch_solve can replay the launch as a captured
 * hipGraph.  Device x/b are row-major (n_total x m). */
#define SPTRSV_BCFG_LS1 0   /* workgroup level-set, 1 lane per row */
#define SPTRSV_BCFG_LS4 1   /* 4 lanes per row (entries when m == 1, RHS columns otherwise) */
#define SPTRSV_BCFG_LS16 2
#define SPTRSV_BCFG_LS64 3
#define SPTRSV_BCFG_CHAIN 4 /* 64-row windows in order (m <= 32; m > 1: waves over the RHS) */
#define SPTRSV_BCFG_COUNT 5
typedef struct sptrsv_batch_s* sptrsv_batch_t;
int sptrsv_batch_create(int nsys, const int* sys_ptr, int64_t nnz, const int64_t* rowptr, const int* col,
                        const double* val, sptrsv_batch_t* batch);
int sptrsv_batch_destroy(sptrsv_batch_t batch);
int sptrsv_batch_info(sptrsv_batch_t batch, int* nsys, int* n, int64_t* nnz);
/* 24 selector features per system (row-major nsys x 24), computed on the host when the batch
 * is created: log2 n, log2 nnz, row length mean / max / std, levels per row, log2 levels,
 * log2 mean level width, max level width / n, level width std / mean, diagonal-only rows / n,
 * mean and max column offset (relative to n), fractions of offsets equal to 1, at most 64 and
 * above 4096, rows with an entry in the next column / n, supernodes / n, mean and max
 * supernode width, rows in supernodes of width >= 16 / n, mean largest offset per row / n^2,
 * log2 m, log2 of the mean |diagonal| over the mean |off-diagonal|. */
int sptrsv_batch_features(sptrsv_batch_t batch, int m, float* feats);
/* Per-system configuration and processing order (either may be NULL). */
int sptrsv_batch_set_configs(sptrsv_batch_t batch, const int* cfg, const int* order);
int sptrsv_batch_solve(sptrsv_batch_t batch, int precision, const double* d_b, double* d_x, int m, int use_graph);
int sptrsv_batch_sync(sptrsv_batch_t batch);
/* FP32 batched solves with FP64 refinement (residual over all systems, stopping rule of
 * sptrsv_solve_refined over the inf-norms of all systems together; opts->precision FP32 or 0).
 * Synchronous; *estimate = estimated relative error of x. */
int sptrsv_batch_solve_refined(sptrsv_batch_t batch, const sptrsv_refine_t* opts, const double* d_b, double* d_x,
                               int m, int* iters, double* estimate);
int sptrsv_batch_solve_refined_ex(sptrsv_batch_t batch, const sptrsv_refine_t* opts, const double* d_b,
                                  double* d_x, int m, sptrsv_refine_info_t* info);
/* r = b - U x over every system, FP64, row-major (n_total x m), asynchronous on the batch's
 * stream (sptrsv_batch_sync waits). */
int sptrsv_batch_residual(sptrsv_batch_t batch, const double* d_b, const double* d_x, double* d_r, int m);
/* Autotuner primitive: every system on `cfg` in one launch; device microseconds per system. */
int sptrsv_batch_time_config(sptrsv_batch_t batch, int cfg, int precision, const double* d_b, double* d_x, int m,
                             double* sys_us);

/* ---- kernel-selector network (no reference counterpart) ---------------------------------
 * in -> h1 -> h2 -> out, ReLU, masked softmax cross-entropy, FP32, forward/backward/Adam on the
 * GPU.  Parameters: W1[in][h1], b1[h1], W2[h1][h2], b2[h2], W3[h2][out], b3[out]. */
typedef struct sptrsv_mlp_s* sptrsv_mlp_t;
int sptrsv_mlp_create(int in, int h1, int h2, int out, unsigned seed, sptrsv_mlp_t* mlp);
int sptrsv_mlp_destroy(sptrsv_mlp_t mlp);
int sptrsv_mlp_num_params(sptrsv_mlp_t mlp, int* count);
int sptrsv_mlp_get_params(sptrsv_mlp_t mlp, float* host);
int sptrsv_mlp_set_params(sptrsv_mlp_t mlp, const float* host);
/* per-sample losses (B) and the gradient of the mean loss (nparams); mask may be NULL */
int sptrsv_mlp_loss_grad(sptrsv_mlp_t mlp, const float* X, const int* y, const float* mask, int B, float* losses,
                         float* grad);
int sptrsv_mlp_forward(sptrsv_mlp_t mlp, const float* X, int B, float* logits);
int sptrsv_mlp_train(sptrsv_mlp_t mlp, const float* X, const int* y, const float* mask, int B, int epochs, float lr,
                     float* final_loss);

#ifdef __cplusplus
}
#endif
#endif
