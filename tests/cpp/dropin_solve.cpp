// Runs every solver entry point of the reference's C++ API (include/core.h,
// include/sptrsv_syncfree_opencl.h) on one matrix, the way the reference's `run()` calls them
// (src/utils.cpp:161-353), and writes each result as raw doubles to <out_dir>/<name>.bin.
// Needs the GPU.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "core.h"
#include "sptrsv_syncfree_opencl.h"
#include "utils.h"

static void dump(const std::string& dir, const char* name, const double* x, size_t count, double t) {
    if (t < 0) {  // the entry point reported a failure (negative status instead of a time)
        fprintf(stderr, "%s failed: %g\n", name, t);
        exit(4);
    }
    std::string path = dir + "/" + name + ".bin";
    FILE* f = fopen(path.c_str(), "wb");
    fwrite(x, sizeof(double), count, f);
    fclose(f);
    printf("%s %.6g\n", name, t);
}

int main(int argc, char** argv) {
    if (argc < 5) return 2;
    const char* mtx = argv[1];
    const char* snf = argv[2];
    const int rhs = atoi(argv[3]);
    const std::string out = argv[4];
    int n = 0;
    unsigned long long nz = 0;
    uint64_t* row;
    int* col;
    double* val;
    if (read_csr(mtx, &n, &nz, &row, &col, &val) < 0) return 3;
    int sn = 0;
    int* snodes = nullptr;
    read_snodes(snf, &sn, &snodes);
    const size_t cnt = (size_t)n * rhs;
    double *x, *b;
    fill_x_b(n, &x, &b, rhs);
    // b = 1 + i % 7 (column-major) so the columns differ
    for (size_t i = 0; i < cnt; ++i) b[i] = 1.0 + (double)((i * 7) % 11);

    std::vector<double> xs(cnt);
    // each output starts as NaN: a call that leaves x untouched cannot pass on an earlier result
    auto poison = [&]() { std::fill(xs.begin(), xs.end(), std::nan("")); };
    poison();
    double t = base_gauss_upper(n, val, row, col, xs.data(), b, rhs);
    dump(out, "base", xs.data(), cnt, t);
    poison();
    t = base_gauss_upper_new(n, val, row, col, xs.data(), b, rhs);
    dump(out, "base_new", xs.data(), cnt, t);
    poison();
    t = barrier_upper(n, xs.data(), b, val, col, row, 4, rhs);
    dump(out, "barrier", xs.data(), cnt, t);

    std::vector<uint64_t> lv((size_t)n);
    const int L = setLevelUpAndGetMaxLevel(n, col, row, lv.data());
    std::vector<double> lvd((size_t)n);
    for (int i = 0; i < n; ++i) lvd[(size_t)i] = (double)lv[(size_t)i];
    dump(out, "levels", lvd.data(), (size_t)n, (double)L);

    if (sn > 0) {
        memcpy(xs.data(), b, cnt * sizeof(double));
        t = supernodal_upper((size_t)sn, snodes, xs.data(), val, col, row, n, rhs);
        dump(out, "custom", xs.data(), cnt, t);
        int extra = 0;
        for (int k = 0; k < sn; ++k) extra += (snodes[k + 1] - snodes[k]) * (snodes[k + 1] - snodes[k] - 1) / 2;
        std::vector<double> vpad(nz + extra);
        std::vector<uint64_t> cpad(nz + extra);
        std::vector<int> rpad((size_t)n + 1, 0);
        int nz_pad = 0;
        ccs2ccs_pad(val, col, row, vpad.data(), cpad.data(), rpad.data(), snodes, sn, nz_pad);
        memcpy(xs.data(), b, cnt * sizeof(double));
        t = supernodal_blas_upper(n, nz_pad, (size_t)sn, snodes, xs.data(), vpad.data(), cpad.data(), rpad.data(), rhs);
        dump(out, "blas", xs.data(), cnt, t);
    }

    // GPU sync-free entry points: CSC from the reference's transpose, row-major B
    double* val_t = new double[nz]();
    int* col_t = new int[n + 1]();
    uint64_t* row_t = new uint64_t[nz]();
    transpose(n, nz, val, col, row, val_t, row_t, col_t);
    std::vector<int> row_t_int(nz);
    for (unsigned long long k = 0; k < nz; ++k) row_t_int[k] = (int)row_t[k];
    std::vector<double> brow(cnt), xrow(cnt);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < rhs; ++k) brow[(size_t)i * rhs + k] = b[i + (size_t)k * n];
    auto poison_row = [&]() { std::fill(xrow.begin(), xrow.end(), std::nan("")); };
    poison_row();
    t = sptrsv_syncfree_opencl(col_t, row_t_int.data(), val_t, n, n, (int)nz, xrow.data(), brow.data(), rhs);
    dump(out, "syncfree", xrow.data(), cnt, t);
    if (rhs == 1) {
        poison_row();
        t = sptrsv_syncfree2_opencl(col_t, row_t_int.data(), val_t, n, n, (int)nz, xrow.data(), brow.data());
        dump(out, "syncfree2", xrow.data(), cnt, t);
        std::vector<int> rp_int((size_t)n + 1);
        for (int i = 0; i <= n; ++i) rp_int[(size_t)i] = (int)row[i];
        poison_row();
        t = sptrsv_syncfree3_opencl(col, rp_int.data(), val, n, n, (int)nz, xrow.data(), brow.data());
        dump(out, "write_first", xrow.data(), cnt, t);
        // forward substitution on L = U^T (lower CSR, diagonal last)
        std::vector<uint64_t> lrow((size_t)n + 1);
        for (int i = 0; i <= n; ++i) lrow[(size_t)i] = (uint64_t)col_t[i];
        poison();
        t = base_gauss_lower(n, val_t, lrow.data(), row_t_int.data(), xs.data(), b);
        dump(out, "lower", xs.data(), (size_t)n, t);
    }
    const double bad = sptrsv_syncfree_opencl(col_t, row_t_int.data(), val_t, n, n + 1, (int)nz, xrow.data(), brow.data(), rhs);
    printf("nonsquare %.1f\n", bad);
    delete[] val_t; delete[] col_t; delete[] row_t;
    free(x); free(b); free(row); free(col); free(val); free(snodes);
    return 0;
}
