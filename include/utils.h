// utils.h -- drop-in replacement of the reference's loader/checker header (include/utils.h:14-31).
#pragma once
#include <cstdint>
#include <cstdio>

// Reads an upper factor into malloc'd CSR arrays (reference src/utils.cpp:9-82):
//   .matrix  binary [int32 n][u64 nz][int32 col[nz]][u64 rowptr[n+1]][f64 val[nz]]
//   .mtx     Matrix Market coordinates sorted by column (a lower factor L read as CSR of L^T)
//   .txt     "n nz" then val[nz], col[nz], rowptr[n+1]
// Returns seconds, or -1 for an unknown extension / unreadable file.
double read_csr(const char* filename, int* n, unsigned long long* nz, uint64_t** row, int** col, double** val);
// Supernode boundaries snodes[0..sn] from .snodes (binary) or .txt (reference src/utils.cpp:84-111).
void read_snodes(const char* filename, int* sn, int** snodes);
// x = b = 1.0 for n*rhs values, malloc'd (reference src/utils.cpp:113-126).
double fill_x_b(int n, double** x, double** b, int rhs);
// CSR(U) -> CSC(U), rows ascending in every column; the overload the reference defines and
// calls (src/utils.cpp:128-159, call at :276).  col_index_t must be zeroed (n+1 ints).
double transpose(int n, unsigned long long nz, double*& val, int*& row, uint64_t*& col_index,
                 double*& val_t, uint64_t*& row_t, int*& col_index_t);
// Prints the relative L2 error of x2 against x1 to stderr (reference src/utils.cpp:355-380);
// col_major: x2 is column-major n x rhs and x1 row-major.
void check_result(int n, double* x1, double* x2, int rhs, bool col_major);
