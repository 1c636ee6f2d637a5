// Exercises the reference's loader entry points (include/utils.h) from C++, the way the
// reference's own gtest does (test/test_sptrsv.cpp:14-42): read a matrix + supernode file and
// print what was read.  No GPU needed.
#include <cstdio>
#include <cstdlib>

#include "utils.h"

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    int n = 0;
    unsigned long long nz = 0;
    uint64_t* row = nullptr;
    int* col = nullptr;
    double* val = nullptr;
    const double t = read_csr(argv[1], &n, &nz, &row, &col, &val);
    printf("read %s\n", t < 0 ? "failed" : "ok");
    if (t < 0) return 0;
    printf("n %d nz %llu\n", n, nz);
    printf("row");
    for (int i = 0; i <= n; ++i) printf(" %llu", (unsigned long long)row[i]);
    printf("\ncol");
    for (unsigned long long k = 0; k < nz; ++k) printf(" %d", col[k]);
    printf("\nval");
    for (unsigned long long k = 0; k < nz; ++k) printf(" %.17g", val[k]);
    printf("\n");
    int sn = 0;
    int* snodes = nullptr;
    read_snodes(argv[2], &sn, &snodes);
    printf("sn %d snodes", sn);
    for (int i = 0; snodes && i <= sn; ++i) printf(" %d", snodes[i]);
    printf("\n");
    double *x, *b;
    fill_x_b(n, &x, &b, 2);
    double s = 0;
    for (int i = 0; i < 2 * n; ++i) s += x[i] + b[i];
    printf("fill %.1f\n", s);
    x[0] += 1e-3;
    check_result(n, b, x, 2, false);
    free(x); free(b); free(row); free(col); free(val); free(snodes);
    return 0;
}
