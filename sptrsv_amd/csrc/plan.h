// plan.h -- device-resident state of one upper-triangular factor.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "common.h"

struct sptrsv_plan_s {
    int n = 0;
    int64_t nnz = 0;
    int n_levels = 0;
    int max_row_len = 0;

    // CSR, natural row order, diagonal first (HBM: 4(n+1) + 12 nnz bytes)
    int* rowptr = nullptr;
    int* col = nullptr;
    double* val = nullptr;
    float* val32 = nullptr;  // FP32 copy, built on first FP32 solve

    // CSC, rows ascending, diagonal last
    int* colptr = nullptr;
    int* rowidx = nullptr;
    double* cval = nullptr;
    float* cval32 = nullptr;

    // analysis
    int* indeg = nullptr;       // nnz of each row incl. diagonal (reference analyser)
    int* levels = nullptr;      // 1-based level of each row
    int* level_ptr = nullptr;   // n_levels + 1
    int* level_rows = nullptr;  // rows grouped by level, ascending inside a level

    // CSR with rows stored in level order (row k of it is original row level_rows[k])
    int* lrowptr = nullptr;
    int* lcol = nullptr;
    double* lval = nullptr;
    float* lval32 = nullptr;

    // work buffers
    int* ticket = nullptr;      // [0]: ticket counter, [1]: status word
    int* indeg_work = nullptr;  // in-degree consumed by the push kernel
    double* left_sum = nullptr;
    size_t left_sum_elems = 0;
    float* b32 = nullptr;       // FP32 staging of b / x
    float* x32 = nullptr;
    size_t stage_elems = 0;

    std::vector<int> h_level_ptr;  // host copy, drives the bulk-synchronous launches

    hipStream_t stream = nullptr;
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    size_t events_used = 0;
    uint64_t ticks_per_ms = 100000;
};

namespace sptrsv {
int analyse(sptrsv_plan_s* p);           // analysis.hip
int ensure_fp32(sptrsv_plan_s* p);       // analysis.hip
int device_transpose(int n, int64_t nnz, const int* rowptr, const int* col, const double* val,
                     int* colptr, int* rowidx, double* cval, hipStream_t s);  // analysis.hip
int begin_timed(sptrsv_plan_s* p, hipEvent_t* stop_out);
void end_timed(sptrsv_plan_s* p, hipEvent_t stop);
}  // namespace sptrsv
