// dropin.cpp -- the reference's C++ entry points (include/core.h, include/utils.h,
// include/sptrsv_syncfree_opencl.h) implemented on top of the C ABI (sptrsv_capi.h).
// Solves, level analysis and the transpose run on the GPU; the loaders are host file I/O.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "core.h"
#include "sptrsv_capi.h"
#include "sptrsv_syncfree_opencl.h"
#include "utils.h"

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void report(const char* where, int rc) {
    fprintf(stderr, "%s: %s (%s)\n", where, sptrsv_status_string(rc), sptrsv_last_error());
}

// One GPU solve of an upper CSR (diagonal first).  Returns device solve seconds or the
// negative status.
// exact != 0: AUTO picks only kernels whose x is bitwise equal to the serial solver's.
double gpu_upper_csr(const char* where, int n, const uint64_t* rowptr, const int* col, const double* val,
                     double* x, const double* b, int rhs, int col_major, int variant, int exact = 0) {
    std::vector<int64_t> rp(rowptr, rowptr + n + 1);
    sptrsv_plan_t plan = nullptr;
    int rc = sptrsv_plan_create_csr(n, rp[n], rp.data(), col, val, &plan);
    if (rc != SPTRSV_OK) {
        report(where, rc);
        return rc;
    }
    sptrsv_params_t prm{};
    prm.variant = variant;
    prm.exact = exact;
    sptrsv_plan_timing(plan, 1);
    rc = sptrsv_solve_host(plan, &prm, b, x, rhs, col_major);
    double ms = 0.0;
    if (rc == SPTRSV_OK) rc = sptrsv_plan_timing_read(plan, &ms, nullptr);
    sptrsv_plan_destroy(plan);
    if (rc != SPTRSV_OK) {
        report(where, rc);
        return rc;
    }
    return ms / 1000.0;
}

// Supernodal solve through the given supernode partition (SUPERNODAL variant); a partition
// whose rows do not have the supernodal structure is reported and solved by the AUTO variant.
double gpu_upper_supernodal(const char* where, int n, const uint64_t* rowptr, const int* col, const double* val,
                            double* x, const double* b, int rhs, size_t sn, const int* snodes) {
    std::vector<int64_t> rp(rowptr, rowptr + n + 1);
    sptrsv_plan_t plan = nullptr;
    int rc = sptrsv_plan_create_csr(n, rp[n], rp.data(), col, val, &plan);
    if (rc != SPTRSV_OK) {
        report(where, rc);
        return rc;
    }
    sptrsv_params_t prm{};
    prm.variant = SPTRSV_VAR_SUPERNODAL;
    if (sptrsv_plan_build_supernodes(plan, (int)sn, snodes) != SPTRSV_OK) {
        std::fprintf(stderr, "%s: %s; solving without supernodes\n", where, sptrsv_last_error());
        prm.variant = SPTRSV_VAR_AUTO;
    }
    sptrsv_plan_timing(plan, 1);
    rc = sptrsv_solve_host(plan, &prm, b, x, rhs, 1);
    double ms = 0.0;
    if (rc == SPTRSV_OK) rc = sptrsv_plan_timing_read(plan, &ms, nullptr);
    sptrsv_plan_destroy(plan);
    if (rc != SPTRSV_OK) {
        report(where, rc);
        return rc;
    }
    return ms / 1000.0;
}

// Supernode rows of the padded format hold d leading zeros (row d of a supernode); drop them
// to get back the plain upper CSR (the inverse of ccs2ccs_pad, src/core.cpp:136-161).
void unpad(int n, size_t sn, const int* snodes, const double* val_pad, const uint64_t* col_pad,
           const int* rowptr_pad, std::vector<uint64_t>& rp, std::vector<int>& col, std::vector<double>& val) {
    rp.assign((size_t)n + 1, 0);
    col.clear();
    val.clear();
    for (size_t k = 0; k < sn; ++k) {
        for (int r = snodes[k]; r < snodes[k + 1]; ++r) {
            const int d = r - snodes[k];
            for (int p = rowptr_pad[r] + d; p < rowptr_pad[r + 1]; ++p) {
                col.push_back((int)col_pad[p]);
                val.push_back(val_pad[p]);
            }
            rp[(size_t)r + 1] = col.size();
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// core.h
// ---------------------------------------------------------------------------------------------

double base_gauss_lower(int n, double* val, uint64_t* row_index, int* col, double* x, double* b) {
    // L x = b with L lower (diagonal last) is U y = c with U = J L J (J reverses indices): row
    // n-1-i of L read backwards is row i of U, diagonal first, columns n-1-j ascending.
    const uint64_t nnz = row_index[n];
    std::vector<uint64_t> rp((size_t)n + 1, 0);
    std::vector<int> c(nnz);
    std::vector<double> v(nnz);
    std::vector<double> bb((size_t)n), yy((size_t)n);
    uint64_t out = 0;
    for (int i = 0; i < n; ++i) {
        const int src = n - 1 - i;
        for (uint64_t p = row_index[src + 1]; p-- > row_index[src];) {
            c[out] = n - 1 - col[p];
            v[out] = val[p];
            ++out;
        }
        rp[(size_t)i + 1] = out;
        bb[(size_t)i] = b[src];
    }
    const double t = gpu_upper_csr("base_gauss_lower", n, rp.data(), c.data(), v.data(), yy.data(), bb.data(), 1, 0,
                                   SPTRSV_VAR_AUTO);
    if (t >= 0)
        for (int i = 0; i < n; ++i) x[n - 1 - i] = yy[(size_t)i];
    return t;
}

// The serial solver's result is the contract here: AUTO restricted to the kernels that sum in
// its order and divide like it (block kernel on grid factors, level-ordered pull otherwise).
double base_gauss_upper(int n, double* val, uint64_t* row_index, int* col, double* x, double* b, int rhs) {
    return gpu_upper_csr("base_gauss_upper", n, row_index, col, val, x, b, rhs, 1, SPTRSV_VAR_AUTO, 1);
}

double base_gauss_upper_new(int n, double* val, uint64_t* row_index, int* col, double* x, double* b, int rhs) {
    return gpu_upper_csr("base_gauss_upper_new", n, row_index, col, val, x, b, rhs, 1, SPTRSV_VAR_AUTO, 1);
}

double supernodal_upper(size_t sn, int* supernodes, double* x, double* val, int* row, uint64_t* col_index, int n,
                        int rhs) {
    std::vector<double> b(x, x + (size_t)n * rhs);
    return gpu_upper_supernodal("supernodal_upper", n, col_index, row, val, x, b.data(), rhs, sn, supernodes);
}

double supernodal_upper_new(size_t sn, int* supernodes, double* x, double* val, int* row, uint64_t* col_index, int n,
                            int rhs) {
    return supernodal_upper(sn, supernodes, x, val, row, col_index, n, rhs);
}

double supernodal_blas_upper(int n, int nz, size_t sn, int* supernodes, double* x, double* val, uint64_t* col_index,
                             int* row, int rhs) {
    (void)nz;
    std::vector<uint64_t> rp;
    std::vector<int> c;
    std::vector<double> v;
    unpad(n, sn, supernodes, val, col_index, row, rp, c, v);
    std::vector<double> b(x, x + (size_t)n * rhs);
    return gpu_upper_supernodal("supernodal_blas_upper", n, rp.data(), c.data(), v.data(), x, b.data(), rhs, sn,
                                supernodes);
}

double supernodal_blas_upper_new(int n, int nz, size_t sn, int* supernodes, double* x, double* val,
                                 uint64_t* col_index, int* row, int rhs) {
    // The reference marks its multi-RHS variant as not working and refuses it
    // (src/core.cpp:203, src/utils.cpp:215-220); here it is the same solve as above.
    return supernodal_blas_upper(n, nz, sn, supernodes, x, val, col_index, row, rhs);
}

double ccs2ccs_pad(double* val, int* row, uint64_t* col_index, double* val_pad, uint64_t* col_index_pad, int* row_pad,
                   int* nodes, int& sn, int& nnz) {
    // Layout conversion only (no arithmetic): row d of every supernode gets d zeros at the
    // supernode's leading columns (src/core.cpp:136-161).
    const double t0 = now_s();
    int64_t out = 0;
    row_pad[0] = 0;
    for (int k = 0; k < sn; ++k) {
        const int s = nodes[k], w = nodes[k + 1] - nodes[k];
        const int lead = row[col_index[s]];
        for (int d = 0; d < w; ++d) {
            for (int z = 0; z < d; ++z, ++out) {
                val_pad[out] = 0.0;
                col_index_pad[out] = (uint64_t)(lead + z);
            }
            for (uint64_t p = col_index[s + d]; p < col_index[s + d + 1]; ++p, ++out) {
                val_pad[out] = val[p];
                col_index_pad[out] = (uint64_t)row[p];
            }
            row_pad[s + d + 1] = (int)out;
        }
    }
    nnz = (int)out;
    return now_s() - t0;
}

int setLevelUpAndGetMaxLevel(int n, int* col, uint64_t* row_index, uint64_t* levelsUp) {
    std::vector<int64_t> rp(row_index, row_index + n + 1);
    std::vector<double> ones((size_t)rp[n], 1.0);  // values do not influence levels
    // the plan validates the pattern; give every diagonal a non-zero value
    sptrsv_plan_t plan = nullptr;
    int rc = sptrsv_plan_create_csr(n, rp[n], rp.data(), col, ones.data(), &plan);
    if (rc != SPTRSV_OK) {
        report("setLevelUpAndGetMaxLevel", rc);
        return rc;
    }
    std::vector<int> lv((size_t)n);
    int L = 0;
    rc = sptrsv_plan_get_levels(plan, lv.data());
    if (rc == SPTRSV_OK) rc = sptrsv_plan_info(plan, nullptr, nullptr, &L);
    sptrsv_plan_destroy(plan);
    if (rc != SPTRSV_OK) {
        report("setLevelUpAndGetMaxLevel", rc);
        return rc;
    }
    for (int i = 0; i < n; ++i) levelsUp[i] = (uint64_t)lv[(size_t)i];
    return L;
}

static double barrier_common(const char* where, int n, double* x, double* b, double* val, int* col, uint64_t* row,
                             int rhs) {
    // analysis + solve, like the reference's return value (src/core.cpp:375)
    std::vector<int64_t> rp(row, row + n + 1);
    const double t0 = now_s();
    sptrsv_plan_t plan = nullptr;
    int rc = sptrsv_plan_create_csr(n, rp[n], rp.data(), col, val, &plan);
    if (rc != SPTRSV_OK) {
        report(where, rc);
        return rc;
    }
    const double t_analysis = now_s() - t0;
    sptrsv_params_t prm{};
    prm.variant = SPTRSV_VAR_LEVELSET;
    sptrsv_plan_timing(plan, 1);
    rc = sptrsv_solve_host(plan, &prm, b, x, rhs, 1);
    double ms = 0.0;
    if (rc == SPTRSV_OK) rc = sptrsv_plan_timing_read(plan, &ms, nullptr);
    sptrsv_plan_destroy(plan);
    if (rc != SPTRSV_OK) {
        report(where, rc);
        return rc;
    }
    return t_analysis + ms / 1000.0;
}

double barrier_upper(int n, double* x, double* b, double* val, int* col, uint64_t* row, int num_of_threads, int rhs) {
    (void)num_of_threads;  // the device schedules whole levels; no host thread count
    return barrier_common("barrier_upper", n, x, b, val, col, row, rhs);
}

double barrier_upper_new(int n, double* x, double* b, double* val, int* col, uint64_t* row, int num_of_threads,
                         int rhs) {
    (void)num_of_threads;
    return barrier_common("barrier_upper_new", n, x, b, val, col, row, rhs);
}

// ---------------------------------------------------------------------------------------------
// utils.h
// ---------------------------------------------------------------------------------------------

static std::string extension_of(const char* filename) {
    std::string f(filename ? filename : "");
    const size_t dot = f.find_last_of('.');
    return dot == std::string::npos ? std::string() : f.substr(dot + 1);
}

// The reference's reader (src/utils.cpp:9-82) with a status: SPTRSV_ERR_IO for a file that cannot
// be opened, an unknown extension or a short read, SPTRSV_ERR_BAD_INPUT for malformed contents.
static int read_csr_status(const char* filename, int* n, unsigned long long* nz, uint64_t** row, int** col,
                           double** val) {
    const std::string ext = extension_of(filename);
    *row = nullptr;
    *col = nullptr;
    *val = nullptr;
    if (ext == "matrix") {
        FILE* f = fopen(filename, "rb");
        if (!f) {
            fprintf(stderr, "read_csr: cannot open %s\n", filename);
            return SPTRSV_ERR_IO;
        }
        bool ok = fread(n, sizeof(int), 1, f) == 1 && fread(nz, sizeof(unsigned long long), 1, f) == 1 && *n > 0;
        if (ok) {
            *col = (int*)malloc(*nz * sizeof(int));
            *row = (uint64_t*)malloc(((size_t)*n + 1) * sizeof(uint64_t));
            *val = (double*)malloc(*nz * sizeof(double));
            ok = fread(*col, sizeof(int), *nz, f) == *nz &&
                 fread(*row, sizeof(uint64_t), (size_t)*n + 1, f) == (size_t)*n + 1 &&
                 fread(*val, sizeof(double), *nz, f) == *nz;
        }
        fclose(f);
        if (!ok) {
            fprintf(stderr, "read_csr: short read in %s\n", filename);
            return SPTRSV_ERR_IO;
        }
    } else if (ext == "mtx") {
        // coordinate entries sorted by column: a lower factor L read as the CSR of L^T
        std::ifstream file(filename);
        if (!file) {
            fprintf(stderr, "read_csr: cannot open %s\n", filename);
            return SPTRSV_ERR_IO;
        }
        while (file.peek() == '%') file.ignore(1 << 20, '\n');
        int n2 = 0;
        file >> *n >> n2 >> *nz;
        if (!file || *n <= 0) return SPTRSV_ERR_BAD_INPUT;
        *col = (int*)malloc(*nz * sizeof(int));
        *row = (uint64_t*)calloc((size_t)*n + 1, sizeof(uint64_t));
        *val = (double*)malloc(*nz * sizeof(double));
        for (unsigned long long k = 0; k < *nz; ++k) {
            int r = 0, c = 0;
            file >> r >> c >> (*val)[k];
            if (!file) return SPTRSV_ERR_IO;
            if (c < 1 || c > *n) return SPTRSV_ERR_BAD_INPUT;
            (*col)[k] = r - 1;
            (*row)[c]++;  // entries of CSR row c-1
        }
        for (int i = 0; i < *n; ++i) (*row)[i + 1] += (*row)[i];
    } else if (ext == "txt") {
        std::ifstream file(filename);
        if (!file) {
            fprintf(stderr, "read_csr: cannot open %s\n", filename);
            return SPTRSV_ERR_IO;
        }
        file >> *n >> *nz;
        if (!file || *n <= 0) return SPTRSV_ERR_BAD_INPUT;
        *col = (int*)malloc(*nz * sizeof(int));
        *row = (uint64_t*)malloc(((size_t)*n + 1) * sizeof(uint64_t));
        *val = (double*)malloc(*nz * sizeof(double));
        for (unsigned long long k = 0; k < *nz; ++k) file >> (*val)[k];
        for (unsigned long long k = 0; k < *nz; ++k) file >> (*col)[k];
        for (int i = 0; i <= *n; ++i) file >> (*row)[i];
        if (!file) return SPTRSV_ERR_IO;
    } else {
        fprintf(stderr, "\nUnknown extension of matrix file %s .Exit\n", ext.c_str());
        return SPTRSV_ERR_IO;
    }
    return SPTRSV_OK;
}

double read_csr(const char* filename, int* n, unsigned long long* nz, uint64_t** row, int** col, double** val) {
    const double t0 = now_s();
    return read_csr_status(filename, n, nz, row, col, val) == SPTRSV_OK ? now_s() - t0 : -1;
}

void read_snodes(const char* filename, int* sn, int** snodes) {
    const std::string ext = extension_of(filename);
    *snodes = nullptr;
    *sn = 0;
    if (ext == "snodes") {
        FILE* f = fopen(filename, "rb");
        if (!f) {
            fprintf(stderr, "read_snodes: cannot open %s\n", filename);
            return;
        }
        if (fread(sn, sizeof(int), 1, f) == 1 && *sn > 0) {
            *snodes = (int*)malloc(((size_t)*sn + 1) * sizeof(int));
            if (fread(*snodes, sizeof(int), (size_t)*sn + 1, f) != (size_t)*sn + 1) {
                free(*snodes);
                *snodes = nullptr;
                *sn = 0;
            }
        }
        fclose(f);
    } else if (ext == "txt") {
        std::ifstream file(filename);
        if (!file) {
            fprintf(stderr, "read_snodes: cannot open %s\n", filename);
            return;
        }
        file >> *sn;
        if (!file || *sn <= 0) {
            *sn = 0;
            return;
        }
        *snodes = (int*)malloc(((size_t)*sn + 1) * sizeof(int));
        for (int i = 0; i <= *sn; ++i) file >> (*snodes)[i];
    } else {
        fprintf(stderr, "\nUnknown extension of file %s .Exit\n", ext.c_str());
    }
}

double fill_x_b(int n, double** x, double** b, int rhs) {
    const double t0 = now_s();
    const size_t count = (size_t)n * (size_t)rhs;
    *x = (double*)malloc(count * sizeof(double));
    *b = (double*)malloc(count * sizeof(double));
    for (size_t i = 0; i < count; ++i) (*b)[i] = (*x)[i] = 1.0;
    return now_s() - t0;
}

double transpose(int n, unsigned long long nz, double*& val, int*& row, uint64_t*& col_index, double*& val_t,
                 uint64_t*& row_t, int*& col_index_t) {
    const double t0 = now_s();
    std::vector<int64_t> rp(col_index, col_index + n + 1);
    std::vector<int> ri(nz ? nz : 1);
    int rc = sptrsv_csr_to_csc(n, (int64_t)nz, rp.data(), row, val, col_index_t, ri.data(), val_t);
    if (rc != SPTRSV_OK) {
        report("transpose", rc);
        return rc;
    }
    for (unsigned long long k = 0; k < nz; ++k) row_t[k] = (uint64_t)ri[k];
    return now_s() - t0;
}

void check_result(int n, double* x1, double* x2, const int rhs, bool col_major) {
    double num = 0.0, den = 0.0;
    const size_t N = (size_t)n;
    for (size_t i = 0; i < N; ++i)
        for (int k = 0; k < rhs; ++k) {
            const double ref = x1[i * rhs + k];
            const double got = col_major ? x2[i + (size_t)k * N] : x2[i * rhs + k];
            num += (ref - got) * (ref - got);
            den += ref * ref;
        }
    fprintf(stderr, "Relative error: %.3e\n", std::sqrt(num) / std::sqrt(den));
}

// ---------------------------------------------------------------------------------------------
// sptrsv_syncfree_opencl.h
// ---------------------------------------------------------------------------------------------

static double gpu_upper_csc(const char* where, const int* colptr, const int* rowidx, const double* val, int n,
                            int nnz, double* x, const double* b, int rhs, int variant) {
    sptrsv_plan_t plan = nullptr;
    int rc = sptrsv_plan_create_csc(n, nnz, colptr, rowidx, val, &plan);
    if (rc != SPTRSV_OK) {
        report(where, rc);
        return rc;
    }
    sptrsv_params_t prm{};
    prm.variant = variant;
    sptrsv_plan_timing(plan, 1);
    rc = sptrsv_solve_host(plan, &prm, b, x, rhs, 0);
    double ms = 0.0;
    if (rc == SPTRSV_OK) rc = sptrsv_plan_timing_read(plan, &ms, nullptr);
    sptrsv_plan_destroy(plan);
    if (rc != SPTRSV_OK) {
        report(where, rc);
        return rc;
    }
    return ms / 1000.0;
}

double sptrsv_syncfree_opencl(int* cscColPtrTR, int* cscRowIdxTR, VALUE_TYPE* cscValTR, int m, int n, int nnzTR,
                              VALUE_TYPE* x, VALUE_TYPE* b, int rhs) {
    if (m != n) {
        printf("This is not a square matrix, return.\n");
        return -1;
    }
    return gpu_upper_csc("sptrsv_syncfree_opencl", cscColPtrTR, cscRowIdxTR, cscValTR, n, nnzTR, x, b, rhs,
                         SPTRSV_VAR_AUTO);
}

double sptrsv_syncfree2_opencl(int* cscColPtrTR, int* cscRowIdxTR, const VALUE_TYPE* cscValTR, const int m,
                               const int n, const int nnzTR, VALUE_TYPE* x, const VALUE_TYPE* b) {
    if (m != n) {
        printf("This is not a square matrix, return.\n");
        return -1;
    }
    return gpu_upper_csc("sptrsv_syncfree2_opencl", cscColPtrTR, cscRowIdxTR, cscValTR, n, nnzTR, x, b, 1,
                         SPTRSV_VAR_CSC_PUSH);
}

double sptrsv_syncfree3_opencl(int* csrColIdx, int* csrRowPtr, const VALUE_TYPE* csrVal, const int m, const int n,
                               const int nnz, VALUE_TYPE* x, const VALUE_TYPE* b) {
    if (m != n) {
        printf("This is not a square matrix, return.\n");
        return -1;
    }
    std::vector<uint64_t> rp(csrRowPtr, csrRowPtr + n + 1);
    (void)nnz;
    // the write-first kernel computes each row in the serial order: the exact AUTO choice
    return gpu_upper_csr("sptrsv_syncfree3_opencl", n, rp.data(), csrColIdx, csrVal, x, b, 1, 0,
                         SPTRSV_VAR_AUTO, 1);
}

// ---- C-ABI access to the loaders (for FFI callers; arrays are malloc'ed, free with
// sptrsv_free_host) ----------------------------------------------------------------------------
extern "C" int sptrsv_read_matrix_file(const char* path, int* n, int64_t* nnz, int64_t** rowptr, int** col,
                                       double** val) {
    if (!path || !n || !nnz || !rowptr || !col || !val) return SPTRSV_ERR_BAD_INPUT;
    unsigned long long nz = 0;
    uint64_t* row = nullptr;
    *rowptr = nullptr;
    *col = nullptr;
    *val = nullptr;
    const int rc = read_csr_status(path, n, &nz, &row, col, val);
    if (rc != SPTRSV_OK) {
        free(row);
        free(*col);
        free(*val);
        *col = nullptr;
        *val = nullptr;
        return rc;
    }
    static_assert(sizeof(uint64_t) == sizeof(int64_t), "row pointer width");
    *nnz = (int64_t)nz;
    *rowptr = reinterpret_cast<int64_t*>(row);
    return SPTRSV_OK;
}

extern "C" int sptrsv_read_snodes_file(const char* path, int* sn, int** snodes) {
    if (!path || !sn || !snodes) return SPTRSV_ERR_BAD_INPUT;
    const std::string ext = extension_of(path);
    FILE* f = fopen(path, "rb");
    if (!f || (ext != "snodes" && ext != "txt")) {
        if (f) fclose(f);
        *snodes = nullptr;
        *sn = 0;
        return SPTRSV_ERR_IO;
    }
    fclose(f);
    read_snodes(path, sn, snodes);
    return *snodes ? SPTRSV_OK : SPTRSV_ERR_BAD_INPUT;
}

extern "C" void sptrsv_free_host(void* p) { free(p); }
