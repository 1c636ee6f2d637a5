// core.h -- drop-in replacement of the reference's CPU solver header (include/core.h:9-19).
// Same names, parameter types and meaning; each call runs on the gfx950 GPU through the C ABI
// (sptrsv_capi.h).  Return value: elapsed seconds like the reference (omp_get_wtime deltas in
// the reference; device-event time of the solve here), or a negative SPTRSV_ERR_* code cast
// to double on failure (the reference returned -1 or called exit()).
#pragma once
#include <cstddef>
#include <cstdint>

#include "sptrsv_capi.h"
#include "utils.h"

// Forward substitution, lower CSR with the diagonal LAST in each row, one RHS
// (reference src/core.cpp:4-16).
double base_gauss_lower(int n, double* val, uint64_t* row_index, int* col, double* x, double* b);
// Backward substitution, upper CSR with the diagonal FIRST, column-major X/B (x[i + rh*n])
// (reference src/core.cpp:18-32 and :34-51).
double base_gauss_upper(int n, double* val, uint64_t* row_index, int* col, double* x, double* b, int rhs);
double base_gauss_upper_new(int n, double* val, uint64_t* row_index, int* col, double* x, double* b, int rhs);
// Supernodal solve IN PLACE (x holds b on entry); `row` = CSR column indices, `col_index` =
// CSR row pointers (reference src/core.cpp:80-133, called at src/utils.cpp:185,190).
double supernodal_upper(size_t sn, int* supernodes, double* x, double* val, int* row, uint64_t* col_index, int n, int rhs);
double supernodal_upper_new(size_t sn, int* supernodes, double* x, double* val, int* row, uint64_t* col_index, int n, int rhs);
// Supernodal solve on the padded format from ccs2ccs_pad, IN PLACE; parameter 7 = padded
// column indices, parameter 8 = padded int row pointers (reference src/core.cpp:164-242).
double supernodal_blas_upper(int n, int nz, size_t sn, int* supernodes, double* x, double* val, uint64_t* col_index, int* row, int rhs);
double supernodal_blas_upper_new(int n, int nz, size_t sn, int* supernodes, double* x, double* val, uint64_t* col_index, int* row, int rhs);
// Padded supernodal format (reference src/core.cpp:136-161).
double ccs2ccs_pad(double* val, int* row, uint64_t* col_index, double* val_pad, uint64_t* col_index_pad, int* row_pad, int* nodes, int& sn, int& nnz);
// 1-based levels, returns the number of levels (reference src/core.cpp:244-270).
int setLevelUpAndGetMaxLevel(int n, int* col, uint64_t* row_index, uint64_t* levelsUp);
// Level-set solve; returns analysis + solve seconds (reference src/core.cpp:273-483).
double barrier_upper(int n, double* x, double* b, double* val, int* col, uint64_t* row, int num_of_threads, int rhs);
double barrier_upper_new(int n, double* x, double* b, double* val, int* col, uint64_t* row, int num_of_threads, int rhs);
