// sptrsv_syncfree_opencl.h -- drop-in replacement of the reference's GPU solver header
// (include/sptrsv_syncfree_opencl.h:11-37).  The names keep the reference's `_opencl` suffix
// so callers link unchanged; the implementation is HIP on gfx950 only.
// X/B are row-major n x rhs (x[i*rhs + k]).  Returns the solve kernel time in seconds, -1 for
// a non-square call (m != n), or a negative SPTRSV_ERR_* code.
#pragma once
#define VALUE_TYPE double

// CSC input, diagonal last in each column (reference src/sptrsv_syncfree_opencl.cpp:6-361).
double sptrsv_syncfree_opencl(int* cscColPtrTR, int* cscRowIdxTR, VALUE_TYPE* cscValTR, int m, int n,
                              int nnzTR, VALUE_TYPE* x, VALUE_TYPE* b, int rhs);
// CSC input, one RHS (reference src/sptrsv_syncfree2_opencl.cpp:5-329).
double sptrsv_syncfree2_opencl(int* cscColPtrTR, int* cscRowIdxTR, const VALUE_TYPE* cscValTR, const int m,
                               const int n, const int nnzTR, VALUE_TYPE* x, const VALUE_TYPE* b);
// CSR input with the diagonal first, one RHS, thread-per-row (reference
// src/sptrsv_syncfree3_opencl.cpp:5-266).
double sptrsv_syncfree3_opencl(int* csrColIdx, int* csrRowPtr, const VALUE_TYPE* csrVal, const int m,
                               const int n, const int nnz, VALUE_TYPE* x, const VALUE_TYPE* b);
