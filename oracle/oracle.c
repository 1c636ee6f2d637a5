/*
 * oracle.c -- CPU restatement of the reference's upper-triangular solve path.
 *
 * TEST INFRASTRUCTURE ONLY (checker for tests/ and smoke(), timed CPU baseline for
 * bench.py).  The shipped library never links or calls it.  Every function restates one
 * reference function; the citation is given above it.  Arithmetic order is kept identical
 * to the reference so that the bitwise identities the reference implies (serial == serial
 * "new" == OpenMP level-set) hold here too.  Build: oracle/Makefile (gcc -O2, no FMA
 * contraction, OpenMP).
 */
#include "oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

double orc_wtime(void) { return omp_get_wtime(); }
int orc_max_threads(void) { return omp_get_max_threads(); }

/* src/core.cpp:18-32 (base_gauss_upper).  RHS loop outermost; for every row the
 * off-diagonal products are summed in ascending position order, then (b - sum) is
 * DIVIDED by the diagonal, which is the first entry of the row.  The last row uses the
 * last stored value (src/core.cpp:22). */
double orc_base_gauss_upper(int n, const double* val, const uint64_t* rowptr, const int* col,
                            double* x, const double* b, int rhs) {
    double t0 = omp_get_wtime();
    for (int rh = 0; rh < rhs; ++rh) {
        const int64_t off = (int64_t)rh * n;
        x[off + n - 1] = b[off + n - 1] / val[rowptr[n] - 1];
        for (int row = n - 2; row >= 0; --row) {
            double acc = 0.0;
            for (uint64_t p = rowptr[row] + 1; p < rowptr[row + 1]; ++p)
                acc += val[p] * x[off + col[p]];
            x[off + row] = (b[off + row] - acc) / val[rowptr[row]];
        }
    }
    return omp_get_wtime() - t0;
}

/* src/core.cpp:34-51 (base_gauss_upper_new).  Same per-(row, rhs) arithmetic as above
 * with the RHS loop inside the row loop, so the result is bitwise identical. */
double orc_base_gauss_upper_new(int n, const double* val, const uint64_t* rowptr, const int* col,
                                double* x, const double* b, int rhs) {
    double t0 = omp_get_wtime();
    for (int rh = 0; rh < rhs; ++rh) {
        const int64_t off = (int64_t)rh * n;
        x[off + n - 1] = b[off + n - 1] / val[rowptr[n] - 1];
    }
    for (int row = n - 2; row >= 0; --row) {
        for (int rh = 0; rh < rhs; ++rh) {
            const int64_t off = (int64_t)rh * n;
            double acc = 0.0;
            for (uint64_t p = rowptr[row] + 1; p < rowptr[row + 1]; ++p)
                acc += val[p] * x[off + col[p]];
            x[off + row] = (b[off + row] - acc) / val[rowptr[row]];
        }
    }
    return omp_get_wtime() - t0;
}

/* src/core.cpp:4-16 (base_gauss_lower).  Lower CSR, diagonal LAST in each row, one RHS.
 * The reference divides x[0] by val[col[0]]; for a valid lower factor row 0 holds only its
 * diagonal, col[0] == 0, and that is val[0] -- restated as the row's last entry. */
double orc_base_gauss_lower(int n, const double* val, const uint64_t* rowptr, const int* col,
                            double* x, const double* b) {
    double t0 = omp_get_wtime();
    (void)col;
    x[0] = b[0] / val[rowptr[1] - 1];
    for (int row = 1; row < n; ++row) {
        double acc = 0.0;
        for (uint64_t p = rowptr[row]; p + 1 < rowptr[row + 1]; ++p)
            acc += val[p] * x[col[p]];
        x[row] = (b[row] - acc) / val[rowptr[row + 1] - 1];
    }
    return omp_get_wtime() - t0;
}

/* src/core.cpp:244-270 (setLevelUpAndGetMaxLevel).  Reverse sweep; a row holding only
 * its diagonal is level 1, otherwise 1 + max(1, max level of its off-diagonal columns).
 * Returns the maximum level. */
int orc_set_level_up(int n, const int* col, const uint64_t* rowptr, uint64_t* levels) {
    for (int i = 0; i < n; ++i) levels[i] = 1;
    uint64_t top = 1;
    for (int row = n - 1; row >= 0; --row) {
        if (rowptr[row + 1] - rowptr[row] > 1) {
            uint64_t deepest = 1;
            for (uint64_t p = rowptr[row] + 1; p < rowptr[row + 1]; ++p)
                if (levels[col[p]] > deepest) deepest = levels[col[p]];
            levels[row] += deepest;
        }
        if (levels[row] > top) top = levels[row];
    }
    return (int)top;
}

/* src/core.cpp:282-287 (rowsByLevel inside barrier_upper).  Rows of level l, ascending,
 * land in level_rows[level_ptr[l-1] .. level_ptr[l]) for l = 1..max_level; level_ptr has
 * max_level + 1 entries with level_ptr[0] = 0 and level_ptr[max_level] = n. */
void orc_level_sets(int n, const uint64_t* levels, int max_level, int* level_ptr, int* level_rows) {
    memset(level_ptr, 0, sizeof(int) * (size_t)(max_level + 1));
    for (int i = 0; i < n; ++i) level_ptr[levels[i]]++;
    for (int l = 1; l <= max_level; ++l) level_ptr[l] += level_ptr[l - 1];
    int* cursor = (int*)malloc(sizeof(int) * (size_t)(max_level + 1));
    for (int l = 1; l <= max_level; ++l) cursor[l] = level_ptr[l - 1];
    for (int i = 0; i < n; ++i) level_rows[cursor[levels[i]]++] = i;
    free(cursor);
}

/* src/core.cpp:289-332 (tasksByLevel inside barrier_upper).  For every level l = 1..L the
 * rows are cut into nthreads nnz-balanced contiguous ranges: a row joins the current task
 * unless it would push the task past floor(level_nnz / nthreads), in which case it opens
 * the next task; the last task takes every remaining row.  tasks is (max_level+1) x
 * (nthreads+1), row l holds the prefix-summed boundaries, row 0 is unused (level 0 is
 * empty in the reference). */
void orc_tasks_by_level(const uint64_t* rowptr, const int* level_ptr, const int* level_rows,
                        int max_level, int nthreads, int* tasks) {
    const int T = nthreads;
    memset(tasks, 0, sizeof(int) * (size_t)(max_level + 1) * (size_t)(T + 1));
    for (int l = 1; l <= max_level; ++l) {
        int* tk = tasks + (size_t)l * (T + 1);
        const int lo = level_ptr[l - 1], hi = level_ptr[l];
        uint64_t level_nnz = 0;
        for (int k = lo; k < hi; ++k) level_nnz += rowptr[level_rows[k] + 1] - rowptr[level_rows[k]];
        const uint64_t cap = level_nnz / (uint64_t)T;
        int cur = 1;
        uint64_t filled = 0;
        for (int k = lo; k < hi; ++k) {
            const uint64_t len = rowptr[level_rows[k] + 1] - rowptr[level_rows[k]];
            if (cur == T) { tk[cur]++; continue; }
            if (len + filled > cap) { filled = len; cur++; tk[cur]++; }
            else { tk[cur]++; filled += len; }
        }
        for (int t = 1; t <= T; ++t) tk[t] += tk[t - 1];
    }
}

static int orc_build_schedule(int n, const int* col, const uint64_t* rowptr, int nthreads,
                              int** level_ptr_out, int** level_rows_out, int** tasks_out) {
    uint64_t* levels = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
    int L = orc_set_level_up(n, col, rowptr, levels);
    int* level_ptr = (int*)malloc(sizeof(int) * (size_t)(L + 1));
    int* level_rows = (int*)malloc(sizeof(int) * (size_t)n);
    orc_level_sets(n, levels, L, level_ptr, level_rows);
    int* tasks = (int*)malloc(sizeof(int) * (size_t)(L + 1) * (size_t)(nthreads + 1));
    orc_tasks_by_level(rowptr, level_ptr, level_rows, L, nthreads, tasks);
    free(levels);
    *level_ptr_out = level_ptr;
    *level_rows_out = level_rows;
    *tasks_out = tasks;
    return L;
}

/* src/core.cpp:273-376 (barrier_upper).  Level analysis + task split, then for every RHS
 * an OpenMP team walks the levels in order; each level is an `omp for schedule(static,1)`
 * over its tasks with the loop's implicit barrier between levels.  Per-row arithmetic is
 * the serial solver's, so x is bitwise equal to orc_base_gauss_upper.  Returns analysis +
 * solve seconds like the reference (src/core.cpp:375). */
double orc_barrier_upper(int n, double* x, const double* b, const double* val, const int* col,
                         const uint64_t* rowptr, int nthreads, int rhs,
                         double* t_analysis, double* t_solve) {
    int *level_ptr, *level_rows, *tasks;
    double t0 = omp_get_wtime();
    const int L = orc_build_schedule(n, col, rowptr, nthreads, &level_ptr, &level_rows, &tasks);
    const double ta = omp_get_wtime() - t0;
    const int T = nthreads;
    double t1 = omp_get_wtime();
    for (int rh = 0; rh < rhs; ++rh) {
        const int64_t off = (int64_t)rh * n;
#pragma omp parallel num_threads(nthreads)
        for (int l = 1; l <= L; ++l) {
            const int* tk = tasks + (size_t)l * (T + 1);
            const int base = level_ptr[l - 1];
#pragma omp for schedule(static, 1)
            for (int t = 0; t < T; ++t) {
                for (int k = tk[t]; k < tk[t + 1]; ++k) {
                    const int row = level_rows[base + k];
                    double acc = 0.0;
                    for (uint64_t p = rowptr[row] + 1; p < rowptr[row + 1]; ++p)
                        acc += val[p] * x[off + col[p]];
                    x[off + row] = (b[off + row] - acc) / val[rowptr[row]];
                }
            }
        }
    }
    const double ts = omp_get_wtime() - t1;
    free(level_ptr); free(level_rows); free(tasks);
    if (t_analysis) *t_analysis = ta;
    if (t_solve) *t_solve = ts;
    return ta + ts;
}

/* src/core.cpp:378-483 (barrier_upper_new).  Same schedule with the RHS loop innermost
 * (one parallel region for all RHS).  Bitwise equal to orc_base_gauss_upper. */
double orc_barrier_upper_new(int n, double* x, const double* b, const double* val, const int* col,
                             const uint64_t* rowptr, int nthreads, int rhs,
                             double* t_analysis, double* t_solve) {
    int *level_ptr, *level_rows, *tasks;
    double t0 = omp_get_wtime();
    const int L = orc_build_schedule(n, col, rowptr, nthreads, &level_ptr, &level_rows, &tasks);
    const double ta = omp_get_wtime() - t0;
    const int T = nthreads;
    double t1 = omp_get_wtime();
#pragma omp parallel num_threads(nthreads)
    for (int l = 1; l <= L; ++l) {
        const int* tk = tasks + (size_t)l * (T + 1);
        const int base = level_ptr[l - 1];
#pragma omp for schedule(static, 1)
        for (int t = 0; t < T; ++t) {
            for (int k = tk[t]; k < tk[t + 1]; ++k) {
                const int row = level_rows[base + k];
                for (int rh = 0; rh < rhs; ++rh) {
                    const int64_t off = (int64_t)rh * n;
                    double acc = 0.0;
                    for (uint64_t p = rowptr[row] + 1; p < rowptr[row + 1]; ++p)
                        acc += val[p] * x[off + col[p]];
                    x[off + row] = (b[off + row] - acc) / val[rowptr[row]];
                }
            }
        }
    }
    const double ts = omp_get_wtime() - t1;
    free(level_ptr); free(level_rows); free(tasks);
    if (t_analysis) *t_analysis = ta;
    if (t_solve) *t_solve = ts;
    return ta + ts;
}

/* src/kernel_full.cl:40-52 (sptrsv_syncfree_opencl_analyser).  in-degree of row i = number
 * of CSC entries whose row index is i, INCLUDING the diagonal (= nnz of CSR row i). */
void orc_indegree_csc(int n, int64_t nnz, const int* rowidx, int* indeg) {
    memset(indeg, 0, sizeof(int) * (size_t)n);
    for (int64_t k = 0; k < nnz; ++k) indeg[rowidx[k]]++;
}

/* src/utils.cpp:128-159 (transpose).  Counting sort of CSR(U) into CSC(U): column counts,
 * exclusive prefix, then rows are scattered in ascending row order, so inside each column
 * rows ascend and the diagonal is the last entry.  colptr_t must be zeroed (n+1). */
double orc_transpose(int n, uint64_t nz, const double* val, const int* col, const uint64_t* rowptr,
                     double* val_t, uint64_t* row_t, int* colptr_t) {
    double t0 = omp_get_wtime();
    for (uint64_t k = 0; k < nz; ++k) colptr_t[col[k] + 1]++;
    int64_t run = 0;
    for (int c = 1; c <= n; ++c) {
        const int cnt = colptr_t[c];
        colptr_t[c] = (int)run;
        run += cnt;
    }
    for (int row = 0; row < n; ++row) {
        for (uint64_t p = rowptr[row]; p < rowptr[row + 1]; ++p) {
            const int dst = colptr_t[col[p] + 1]++;
            val_t[dst] = val[p];
            row_t[dst] = (uint64_t)row;
        }
    }
    return omp_get_wtime() - t0;
}

/* src/core.cpp:67-78 (get_rect_part_upper): row i of a supernode of width w starting at s
 * skips its w - (i - s) triangle entries and subtracts the rest, ascending. */
static void orc_rect_part(int s, int w, double* x, const double* val, const int* col,
                          const uint64_t* rowptr, int64_t off) {
    for (int i = s; i < s + w; ++i) {
        double acc = 0.0;
        for (uint64_t p = rowptr[i] + (uint64_t)(w - (i - s)); p < rowptr[i + 1]; ++p)
            acc += val[p] * x[off + col[p]];
        x[off + i] -= acc;
    }
}

/* src/core.cpp:53-65 (get_tr_part_upper): dense back-substitution inside the supernode's
 * triangle; row j uses its diagonal (first entry) and the next (s + w - 1 - j) entries. */
static void orc_tri_part(int s, int w, double* x, const double* val, const int* col,
                         const uint64_t* rowptr, int64_t off) {
    x[off + s + w - 1] = x[off + s + w - 1] / val[rowptr[s + w - 1]];
    int used = 0;
    for (int j = s + w - 2; j >= s; --j) {
        double acc = 0.0;
        for (uint64_t p = rowptr[j] + 1; p <= rowptr[j] + 1 + (uint64_t)used; ++p)
            acc += val[p] * x[off + col[p]];
        x[off + j] = (x[off + j] - acc) / val[rowptr[j]];
        used++;
    }
}

/* src/core.cpp:80-92 (supernodal_upper).  IN PLACE: x holds b on entry.  Per RHS, for
 * supernodes from last to first: rectangular update, then the triangle. */
double orc_supernodal_upper(int sn, const int* snodes, double* x, const double* val, const int* col,
                            const uint64_t* rowptr, int n, int rhs) {
    double t0 = omp_get_wtime();
    for (int rh = 0; rh < rhs; ++rh) {
        const int64_t off = (int64_t)rh * n;
        for (int k = sn - 1; k >= 0; --k) {
            orc_rect_part(snodes[k], snodes[k + 1] - snodes[k], x, val, col, rowptr, off);
            orc_tri_part(snodes[k], snodes[k + 1] - snodes[k], x, val, col, rowptr, off);
        }
    }
    return omp_get_wtime() - t0;
}

/* src/core.cpp:94-133 (supernodal_upper_new).  Same arithmetic per (row, rhs) with the
 * RHS loops moved inside each supernode; bitwise equal to orc_supernodal_upper. */
double orc_supernodal_upper_new(int sn, const int* snodes, double* x, const double* val,
                                const int* col, const uint64_t* rowptr, int n, int rhs) {
    double t0 = omp_get_wtime();
    for (int k = sn - 1; k >= 0; --k) {
        const int s = snodes[k], w = snodes[k + 1] - snodes[k];
        for (int i = s; i < s + w; ++i)
            for (int rh = 0; rh < rhs; ++rh) {
                const int64_t off = (int64_t)rh * n;
                double acc = 0.0;
                for (uint64_t p = rowptr[i] + (uint64_t)(w - (i - s)); p < rowptr[i + 1]; ++p)
                    acc += val[p] * x[off + col[p]];
                x[off + i] -= acc;
            }
        for (int rh = 0; rh < rhs; ++rh) {
            const int64_t off = (int64_t)rh * n;
            x[off + s + w - 1] = x[off + s + w - 1] / val[rowptr[s + w - 1]];
        }
        int used = 0;
        for (int j = s + w - 2; j >= s; --j) {
            for (int rh = 0; rh < rhs; ++rh) {
                const int64_t off = (int64_t)rh * n;
                double acc = 0.0;
                for (uint64_t p = rowptr[j] + 1; p <= rowptr[j] + 1 + (uint64_t)used; ++p)
                    acc += val[p] * x[off + col[p]];
                x[off + j] = (x[off + j] - acc) / val[rowptr[j]];
            }
            used++;
        }
    }
    return omp_get_wtime() - t0;
}

/* src/utils.cpp:201-206: extra storage of the padded format = sum over supernodes of
 * w (w - 1) / 2. */
uint64_t orc_pad_extra(int sn, const int* snodes) {
    uint64_t extra = 0;
    for (int k = 0; k < sn; ++k) {
        const uint64_t w = (uint64_t)(snodes[k + 1] - snodes[k]);
        extra += w * (w - 1) / 2;
    }
    return extra;
}

/* src/core.cpp:136-161 (ccs2ccs_pad).  Row d of each supernode gets d explicit zeros in
 * front (at the supernode's leading columns), turning the supernode into a dense row-major
 * w x (w + r) block.  rowptr_pad must hold n+1 ints with rowptr_pad[0] == 0. */
double orc_ccs2ccs_pad(const double* val, const int* col, const uint64_t* rowptr, double* val_pad,
                       uint64_t* col_pad, int* rowptr_pad, const int* snodes, int sn,
                       int64_t* nnz_pad) {
    double t0 = omp_get_wtime();
    int64_t out = 0;
    rowptr_pad[0] = 0;
    for (int k = 0; k < sn; ++k) {
        const int s = snodes[k], w = snodes[k + 1] - snodes[k];
        const int first_col = col[rowptr[s]];
        for (int d = 0; d < w; ++d) {
            for (int z = 0; z < d; ++z, ++out) {
                val_pad[out] = 0.0;
                col_pad[out] = (uint64_t)(first_col + z);
            }
            for (uint64_t p = rowptr[s + d]; p < rowptr[s + d + 1]; ++p, ++out) {
                val_pad[out] = val[p];
                col_pad[out] = (uint64_t)col[p];
            }
            rowptr_pad[s + d + 1] = (int)out;
        }
    }
    *nnz_pad = out;
    return omp_get_wtime() - t0;
}

/* src/core.cpp:164-201 (supernodal_blas_upper).  IN PLACE: x holds b on entry.  Per RHS,
 * for supernodes from last to first: gather x at the rectangular columns, c = R * g
 * (cblas_dgemm, row-major, lda = w + r), x[block] -= c, then an upper non-unit triangular
 * solve of the w x w block (cblas_dtrsm).  The BLAS calls are restated as plain loops
 * (ascending inner products); MKL's internal order is not specified. */
double orc_supernodal_blas_upper(int n, int sn, const int* snodes, double* x, const double* val_pad,
                                 const uint64_t* col_pad, const int* rowptr_pad, int rhs) {
    double* g = (double*)calloc((size_t)n, sizeof(double));
    double* c = (double*)calloc((size_t)n, sizeof(double));
    double t0 = omp_get_wtime();
    for (int rh = 0; rh < rhs; ++rh) {
        const int64_t off = (int64_t)rh * n;
        for (int k = sn - 1; k >= 0; --k) {
            const int s = snodes[k], w = snodes[k + 1] - snodes[k];
            const int64_t head = rowptr_pad[s];
            const int64_t lda = (rowptr_pad[s + w] - rowptr_pad[s]) / w;
            const int64_t r = rowptr_pad[s + 1] - rowptr_pad[s] - w;
            for (int64_t q = 0; q < r; ++q) g[q] = x[off + (int64_t)col_pad[head + w + q]];
            for (int d = 0; d < w; ++d) {
                double acc = 0.0;
                const double* rrow = val_pad + head + w + (int64_t)d * lda;
                for (int64_t q = 0; q < r; ++q) acc += rrow[q] * g[q];
                c[d] = acc;
            }
            for (int d = 0; d < w; ++d) { x[off + s + d] -= c[d]; c[d] = 0.0; }
            const double* tri = val_pad + head;
            for (int d = w - 1; d >= 0; --d) {
                double acc = 0.0;
                for (int e = d + 1; e < w; ++e) acc += tri[(int64_t)d * lda + e] * x[off + s + e];
                x[off + s + d] = (x[off + s + d] - acc) / tri[(int64_t)d * lda + d];
            }
        }
    }
    double t = omp_get_wtime() - t0;
    free(g);
    free(c);
    return t;
}

/* src/core.cpp:164-201 (supernodal_blas_upper) with the reference's two BLAS calls made
 * through a CBLAS library loaded at run time (dlopen of `blas_path`: MKL's libmkl_rt where the
 * host has it, as the reference links MKL, else an OpenBLAS).  Same call sequence and
 * arguments as src/core.cpp:186-194: cblas_dgemm(RowMajor, NoTrans, NoTrans, w, 1, r, 1, R,
 * lda, g, 1, 1, c, 1), x -= c, cblas_dtrsm(RowMajor, Left, Upper, NoTrans, NonUnit, w, 1, 1,
 * T, lda, x, 1).  The CBLAS enumerators are the standard ones (RowMajor 101, NoTrans 111,
 * Upper 121, NonUnit 131, Left 141).  Returns the solve seconds, or -1 if the library or its
 * symbols cannot be loaded. */
typedef void (*orc_dgemm_t)(int, int, int, int, int, int, double, const double*, int, const double*, int, double,
                            double*, int);
typedef void (*orc_dtrsm_t)(int, int, int, int, int, int, int, double, const double*, int, double*, int);

double orc_supernodal_blas_upper_cblas(const char* blas_path, int n, int sn, const int* snodes, double* x,
                                       const double* val_pad, const uint64_t* col_pad, const int* rowptr_pad,
                                       int rhs) {
    /* MKL's runtime would start its own OpenMP runtime next to the one already in the process
     * and abort; the calls are matrix-vector sized, so its sequential layer is the right one */
    setenv("MKL_THREADING_LAYER", "SEQUENTIAL", 0);
    void* h = dlopen(blas_path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1.0;
    int (*set_layer)(int) = (int (*)(int))dlsym(h, "MKL_Set_Threading_Layer");
    if (set_layer) set_layer(1); /* MKL_THREADING_SEQUENTIAL */
    orc_dgemm_t gemm = (orc_dgemm_t)dlsym(h, "cblas_dgemm");
    orc_dtrsm_t trsm = (orc_dtrsm_t)dlsym(h, "cblas_dtrsm");
    if (!gemm || !trsm) return -1.0;
    double* g = (double*)calloc((size_t)n, sizeof(double));
    double* c = (double*)calloc((size_t)n, sizeof(double));
    const double t0 = omp_get_wtime();
    for (int rh = 0; rh < rhs; ++rh) {
        const int64_t off = (int64_t)rh * n;
        for (int k = sn - 1; k >= 0; --k) {
            const int s = snodes[k], w = snodes[k + 1] - snodes[k];
            const int64_t head = rowptr_pad[s];
            const int lda = (int)((rowptr_pad[s + w] - rowptr_pad[s]) / w);
            const int r = (int)(rowptr_pad[s + 1] - rowptr_pad[s] - w);
            for (int q = 0; q < r; ++q) g[q] = x[off + (int64_t)col_pad[head + w + q]];
            gemm(101, 111, 111, w, 1, r, 1.0, val_pad + head + w, lda, g, 1, 1.0, c, 1);
            for (int d = 0; d < w; ++d) {
                x[off + s + d] -= c[d];
                c[d] = 0.0;
            }
            trsm(101, 141, 121, 111, 131, w, 1, 1.0, val_pad + head, lda, x + off + s, 1);
        }
    }
    const double t = omp_get_wtime() - t0;
    free(g);
    free(c);
    return t;
}

/* src/utils.cpp:355-380 (check_result): ||x_ref - x||_2 / ||x_ref||_2 over `count` values
 * of one common layout.  The reference's col_major branch accumulates the norm with a
 * mixed index (x2[rh*n+i] * x2[k], src/utils.cpp:363); callers here bring both vectors to
 * one layout first and use the plain formula. */
double orc_rel_error(int64_t count, const double* x_ref, const double* x) {
    double num = 0.0, den = 0.0;
    for (int64_t i = 0; i < count; ++i) {
        const double d = x_ref[i] - x[i];
        num += d * d;
        den += x_ref[i] * x_ref[i];
    }
    return sqrt(num) / sqrt(den);
}

/* Residual r = b - U x for CSR U (diagonal first), column-major RHS.  No reference
 * counterpart: the FP64 residual of the refinement loop; products summed in ascending
 * position order. */
void orc_spmv_residual(int n, const double* val, const uint64_t* rowptr, const int* col,
                       const double* x, const double* b, double* r, int rhs) {
    for (int rh = 0; rh < rhs; ++rh) {
        const int64_t off = (int64_t)rh * n;
        for (int row = 0; row < n; ++row) {
            double acc = 0.0;
            for (uint64_t p = rowptr[row]; p < rowptr[row + 1]; ++p)
                acc += val[p] * x[off + col[p]];
            r[off + row] = b[off + row] - acc;
        }
    }
}
