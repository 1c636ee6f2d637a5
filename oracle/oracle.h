/*
 * oracle.h -- CPU restatement of the reference's upper-triangular solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links or calls this code.
 * It is used by tests/ (as the checker), by __graft_entry__.smoke() (as the checker) and
 * by bench.py's cpu_baseline leg (as the timed CPU baseline).  Each function names the
 * reference file:line it restates.
 *
 * Conventions (same as the reference):
 *   - CSR of an upper-triangular U, columns ascending within a row, diagonal FIRST.
 *     rowptr is uint64 (include/core.h:10), column indices int.
 *   - CSC of U (from transpose), rows ascending within a column, diagonal LAST.
 *   - Multi right-hand sides on the CPU are column-major: x[i + rh*n] (src/core.cpp:26).
 *   - Levels are 1-based (src/core.cpp:244-270).
 *
 * Parity pins (every function is checked by tests/test_oracle.py):
 *   orc_base_gauss_upper(_new)  exact rational x on the csr_5/csr_6 fixtures, dense
 *                               back-substitution and scipy solve_triangular on random U
 *   orc_base_gauss_lower        dense forward substitution, scipy solve_triangular
 *   orc_set_level_up            closed forms (banded, 7-point grid), fixture levels
 *   orc_level_sets              fixture level sets, sortedness/partition invariants
 *   orc_tasks_by_level          partition invariant + hand-computed small case
 *   orc_barrier_upper(_new)     bitwise identity with orc_base_gauss_upper (same per-row
 *                               arithmetic, any thread count)
 *   orc_indegree_csc            closed forms, row-length identity from CSR
 *   orc_transpose               scipy.sparse CSR->CSC
 *   orc_supernodal_upper(_new), orc_ccs2ccs_pad, orc_supernodal_blas_upper
 *                               dense solve on csr_6 + its supernode file, random
 *                               supernodal matrices vs scipy
 *   orc_rel_error               closed form on hand vectors
 *   orc_spmv_residual           dense matvec
 */
#ifndef SPTRSV_ORACLE_H
#define SPTRSV_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

double orc_base_gauss_upper(int n, const double* val, const uint64_t* rowptr, const int* col,
                            double* x, const double* b, int rhs);
double orc_base_gauss_upper_new(int n, const double* val, const uint64_t* rowptr, const int* col,
                                double* x, const double* b, int rhs);
double orc_base_gauss_lower(int n, const double* val, const uint64_t* rowptr, const int* col,
                            double* x, const double* b);

int orc_set_level_up(int n, const int* col, const uint64_t* rowptr, uint64_t* levels);
void orc_level_sets(int n, const uint64_t* levels, int max_level, int* level_ptr, int* level_rows);
void orc_tasks_by_level(const uint64_t* rowptr, const int* level_ptr, const int* level_rows,
                        int max_level, int nthreads, int* tasks);
double orc_barrier_upper(int n, double* x, const double* b, const double* val, const int* col,
                         const uint64_t* rowptr, int nthreads, int rhs,
                         double* t_analysis, double* t_solve);
double orc_barrier_upper_new(int n, double* x, const double* b, const double* val, const int* col,
                             const uint64_t* rowptr, int nthreads, int rhs,
                             double* t_analysis, double* t_solve);

void orc_indegree_csc(int n, int64_t nnz, const int* rowidx, int* indeg);
double orc_transpose(int n, uint64_t nz, const double* val, const int* col, const uint64_t* rowptr,
                     double* val_t, uint64_t* row_t, int* colptr_t);

double orc_supernodal_upper(int sn, const int* snodes, double* x, const double* val, const int* col,
                            const uint64_t* rowptr, int n, int rhs);
double orc_supernodal_upper_new(int sn, const int* snodes, double* x, const double* val,
                                const int* col, const uint64_t* rowptr, int n, int rhs);
uint64_t orc_pad_extra(int sn, const int* snodes);
double orc_ccs2ccs_pad(const double* val, const int* col, const uint64_t* rowptr, double* val_pad,
                       uint64_t* col_pad, int* rowptr_pad, const int* snodes, int sn,
                       int64_t* nnz_pad);
double orc_supernodal_blas_upper(int n, int sn, const int* snodes, double* x, const double* val_pad,
                                 const uint64_t* col_pad, const int* rowptr_pad, int rhs);

double orc_rel_error(int64_t count, const double* x_ref, const double* x);
void orc_spmv_residual(int n, const double* val, const uint64_t* rowptr, const int* col,
                       const double* x, const double* b, double* r, int rhs);

double orc_wtime(void);
int orc_max_threads(void);

/* src/core.cpp:164-201 with cblas_dgemm / cblas_dtrsm from a BLAS loaded at run time. */
double orc_supernodal_blas_upper_cblas(const char* blas_path, int n, int sn, const int* snodes, double* x,
                                       const double* val_pad, const uint64_t* col_pad, const int* rowptr_pad,
                                       int rhs);

#ifdef __cplusplus
}
#endif
#endif
