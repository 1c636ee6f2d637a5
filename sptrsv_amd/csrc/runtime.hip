// runtime.hip -- device/memory C ABI, input validation, plan lifecycle and timing.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "plan.h"

namespace sptrsv {

static thread_local std::string g_last_error;
static std::mutex g_dev_mutex;
static int g_device = -1;
static uint64_t g_ticks_per_ms = 0;
static int g_compute_units = 0;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int ensure_device() {
    std::lock_guard<std::mutex> lk(g_dev_mutex);
    if (g_device >= 0) return SPTRSV_OK;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return fail(SPTRSV_ERR_NO_DEVICE, "no HIP device visible (the solver needs a gfx950 GPU)");
    SPT_HIP(hipSetDevice(0));
    int khz = 0;
    SPT_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0));
    g_ticks_per_ms = khz > 0 ? (uint64_t)khz : 100000;
    int cus = 0;
    SPT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    g_compute_units = cus > 0 ? cus : 256;
    g_device = 0;
    return SPTRSV_OK;
}

uint64_t wallclock_ticks_per_ms() { return g_ticks_per_ms ? g_ticks_per_ms : 100000; }
int device_compute_units() { return g_compute_units ? g_compute_units : 256; }

int begin_timed(sptrsv_plan_s* p, hipEvent_t* stop_out) {
    *stop_out = nullptr;
    if (!p->timing) return SPTRSV_OK;
    if (p->events_used == p->events.size()) {
        hipEvent_t a, b;
        SPT_HIP(hipEventCreate(&a));
        SPT_HIP(hipEventCreate(&b));
        p->events.emplace_back(a, b);
    }
    auto& pr = p->events[p->events_used++];
    SPT_HIP(hipEventRecord(pr.first, p->stream));
    *stop_out = pr.second;
    return SPTRSV_OK;
}

void end_timed(sptrsv_plan_s* p, hipEvent_t stop) {
    if (stop) (void)hipEventRecord(stop, p->stream);
}

namespace {
__global__ void k_zero_ints(int* out, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = 0;
}
}  // namespace

// Zeroes `count` ints with a kernel on stream `s`, ordered before everything later on that
// stream.  (hipMemset runs on the null stream, which a non-blocking stream does not wait for,
// and may return before it has run: the plans' streams need their control words zeroed in
// stream order.)
int zero_ints_async(int* p, int64_t count, hipStream_t s) {
    if (count <= 0) return SPTRSV_OK;
    const int g = (int)std::min<int64_t>((count + 255) / 256, 65536);
    k_zero_ints<<<g, 256, 0, s>>>(p, count);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

namespace {
__global__ void k_division_check(const double* n, const double* d, double* fast, double* ieee, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        fast[i] = div_by(n[i], d[i], div_recip(d[i]));
        ieee[i] = n[i] / d[i];
    }
}
}  // namespace

// A plan that holds no factor: stream, status word, timing events and the refinement / FP32
// staging buffers, for solvers that keep their own matrix arrays (the batched solver).
int plan_create_control(int n, int64_t nnz, sptrsv_plan_s** out) {
    SPT_CHECK(ensure_device());
    auto* p = new sptrsv_plan_s();
    p->n = n;
    p->nnz = nnz;
    *out = p;
    SPT_CHECK(sptrsv_malloc((void**)&p->ticket, 64 * sizeof(int)));
    SPT_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    SPT_CHECK(zero_ints_async(p->ticket, 64, p->stream));
    SPT_HIP(hipGetDevice(&p->device));
    p->ticks_per_ms = wallclock_ticks_per_ms();
    return SPTRSV_OK;
}
}  // namespace sptrsv

using namespace sptrsv;

extern "C" {

const char* sptrsv_last_error(void) { return g_last_error.c_str(); }

const char* sptrsv_status_string(int code) {
    switch (code) {
        case SPTRSV_OK: return "ok";
        case SPTRSV_ERR_BAD_INPUT: return "bad input";
        case SPTRSV_ERR_HIP: return "HIP runtime error";
        case SPTRSV_ERR_DEADLOCK: return "device watchdog: dependency wait timed out";
        case SPTRSV_ERR_IO: return "I/O error";
        case SPTRSV_ERR_NOMEM: return "out of memory";
        case SPTRSV_ERR_UNSUPPORTED: return "unsupported size or option";
        case SPTRSV_ERR_NO_DEVICE: return "no HIP device";
        default: return "unknown status";
    }
}

int sptrsv_device_init(int device) {
    (void)device;
    return ensure_device();
}

int sptrsv_device_info(char* name, int name_len, int* compute_units, size_t* total_mem) {
    SPT_CHECK(ensure_device());
    hipDeviceProp_t prop;
    SPT_HIP(hipGetDeviceProperties(&prop, 0));
    if (name && name_len > 0) {
        snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (total_mem) *total_mem = prop.totalGlobalMem;
    return SPTRSV_OK;
}

int sptrsv_device_sync(void) {
    SPT_CHECK(ensure_device());
    SPT_HIP(hipDeviceSynchronize());
    return SPTRSV_OK;
}

int sptrsv_malloc(void** ptr, size_t bytes) {
    SPT_CHECK(ensure_device());
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
        *ptr = nullptr;
        return fail(SPTRSV_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    return SPTRSV_OK;
}

int sptrsv_free(void* ptr) {
    if (ptr) SPT_HIP(hipFree(ptr));
    return SPTRSV_OK;
}

int sptrsv_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    SPT_CHECK(ensure_device());
    SPT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SPTRSV_OK;
}

int sptrsv_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    SPT_CHECK(ensure_device());
    SPT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return SPTRSV_OK;
}

int sptrsv_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    SPT_CHECK(ensure_device());
    SPT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return SPTRSV_OK;
}

int sptrsv_validate_csr(int n, int64_t nnz, const int64_t* rowptr, const int* col, const double* val) {
    if (n <= 0) return fail(SPTRSV_ERR_BAD_INPUT, "n must be positive");
    if (!rowptr || !col || !val) return fail(SPTRSV_ERR_BAD_INPUT, "null CSR array");
    if (rowptr[0] != 0) return fail(SPTRSV_ERR_BAD_INPUT, "rowptr[0] != 0");
    if (rowptr[n] != nnz) return fail(SPTRSV_ERR_BAD_INPUT, "rowptr[n] != nnz");
    for (int i = 0; i < n; ++i) {
        const int64_t a = rowptr[i], b = rowptr[i + 1];
        if (b <= a)
            return fail(SPTRSV_ERR_BAD_INPUT, "row " + std::to_string(i) + " is empty (no diagonal)");
        if (col[a] != i)
            return fail(SPTRSV_ERR_BAD_INPUT,
                        "row " + std::to_string(i) + ": first entry is not the diagonal (col " +
                            std::to_string(col[a]) + ")");
        if (!(val[a] != 0.0) || !std::isfinite(val[a]))
            return fail(SPTRSV_ERR_BAD_INPUT, "row " + std::to_string(i) + ": zero or non-finite diagonal");
        int prev = i;
        for (int64_t p = a + 1; p < b; ++p) {
            if (col[p] <= prev || col[p] >= n)
                return fail(SPTRSV_ERR_BAD_INPUT,
                            "row " + std::to_string(i) + ": column " + std::to_string(col[p]) +
                                " not strictly ascending above the diagonal");
            prev = col[p];
        }
    }
    return SPTRSV_OK;
}

int sptrsv_validate_csc(int n, int64_t nnz, const int* colptr, const int* rowidx, const double* val) {
    if (n <= 0) return fail(SPTRSV_ERR_BAD_INPUT, "n must be positive");
    if (!colptr || !rowidx || !val) return fail(SPTRSV_ERR_BAD_INPUT, "null CSC array");
    if (colptr[0] != 0) return fail(SPTRSV_ERR_BAD_INPUT, "colptr[0] != 0");
    if ((int64_t)colptr[n] != nnz) return fail(SPTRSV_ERR_BAD_INPUT, "colptr[n] != nnz");
    for (int j = 0; j < n; ++j) {
        const int a = colptr[j], b = colptr[j + 1];
        if (b <= a)
            return fail(SPTRSV_ERR_BAD_INPUT, "column " + std::to_string(j) + " is empty (no diagonal)");
        if (rowidx[b - 1] != j)
            return fail(SPTRSV_ERR_BAD_INPUT, "column " + std::to_string(j) + ": last entry is not the diagonal");
        if (!(val[b - 1] != 0.0) || !std::isfinite(val[b - 1]))
            return fail(SPTRSV_ERR_BAD_INPUT, "column " + std::to_string(j) + ": zero or non-finite diagonal");
        int prev = -1;
        for (int p = a; p < b - 1; ++p) {
            if (rowidx[p] <= prev || rowidx[p] >= j)
                return fail(SPTRSV_ERR_BAD_INPUT,
                            "column " + std::to_string(j) + ": row " + std::to_string(rowidx[p]) +
                                " not strictly ascending below the diagonal");
            prev = rowidx[p];
        }
    }
    return SPTRSV_OK;
}

static int plan_alloc_common(sptrsv_plan_s* p) {
    const size_t n = (size_t)p->n, nnz = (size_t)p->nnz;
    SPT_CHECK(sptrsv_malloc((void**)&p->rowptr, (n + 1) * sizeof(int)));
    SPT_CHECK(sptrsv_malloc((void**)&p->col, nnz * sizeof(int)));
    SPT_CHECK(sptrsv_malloc((void**)&p->val, nnz * sizeof(double)));
    SPT_CHECK(sptrsv_malloc((void**)&p->colptr, (n + 1) * sizeof(int)));
    SPT_CHECK(sptrsv_malloc((void**)&p->rowidx, nnz * sizeof(int)));
    SPT_CHECK(sptrsv_malloc((void**)&p->cval, nnz * sizeof(double)));
    SPT_CHECK(sptrsv_malloc((void**)&p->indeg, n * sizeof(int)));
    SPT_CHECK(sptrsv_malloc((void**)&p->indeg_work, n * sizeof(int)));
    SPT_CHECK(sptrsv_malloc((void**)&p->levels, n * sizeof(int)));
    SPT_CHECK(sptrsv_malloc((void**)&p->level_rows, n * sizeof(int)));
    SPT_CHECK(sptrsv_malloc((void**)&p->lrowptr, (n + 1) * sizeof(int)));
    SPT_CHECK(sptrsv_malloc((void**)&p->lcol, nnz * sizeof(int)));
    SPT_CHECK(sptrsv_malloc((void**)&p->lval, nnz * sizeof(double)));
    SPT_CHECK(sptrsv_malloc((void**)&p->ticket, 64 * sizeof(int)));
    SPT_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    SPT_HIP(hipGetDevice(&p->device));
    p->ticks_per_ms = wallclock_ticks_per_ms();
    return SPTRSV_OK;
}

int sptrsv_plan_destroy(sptrsv_plan_t p) {
    if (!p) return SPTRSV_OK;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->tr) sptrsv_plan_destroy(p->tr);
    if (p->tr_buf) (void)hipFree(p->tr_buf);
    supernodal_free(p);
    tile_free(p);
    block_free(p);
    panel_free(p);
    levelset_free(p);
    chain_free(p);
    void* bufs[] = {p->rowptr, p->col, p->val, p->val32, p->colptr, p->rowidx, p->cval, p->cval32,
                    p->indeg, p->levels, p->level_ptr, p->level_rows, p->lrowptr, p->lcol,
                    p->lval, p->lval32, p->ticket, p->indeg_work, p->left_sum, p->b32, p->x32,
                    p->ref_r, p->ref_r32, p->ref_d32, p->ref_norms};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (auto& e : p->events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (p->stream && !p->borrowed_stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return SPTRSV_OK;
}

int sptrsv_plan_create_csr(int n, int64_t nnz, const int64_t* rowptr, const int* col,
                           const double* val, sptrsv_plan_t* out) {
    if (!out) return fail(SPTRSV_ERR_BAD_INPUT, "null plan output");
    *out = nullptr;
    if (nnz >= ((int64_t)1 << 31))
        return fail(SPTRSV_ERR_UNSUPPORTED, "nnz >= 2^31 is not supported (int32 device indices)");
    SPT_CHECK(sptrsv_validate_csr(n, nnz, rowptr, col, val));
    SPT_CHECK(ensure_device());
    auto* p = new sptrsv_plan_s();
    p->n = n;
    p->nnz = nnz;
    int rc = plan_alloc_common(p);
    if (rc != SPTRSV_OK) { sptrsv_plan_destroy(p); return rc; }
    std::vector<int> rp32((size_t)n + 1);
    int max_len = 0;
    for (int i = 0; i <= n; ++i) rp32[i] = (int)rowptr[i];
    for (int i = 0; i < n; ++i) max_len = std::max(max_len, rp32[i + 1] - rp32[i]);
    p->max_row_len = max_len;
    {  // every row's first off-diagonal column is a dependency: the depth of that forest bounds
       // the level count from below (one pass, rows in descending order)
        std::vector<int> depth((size_t)n, 1);
        int deepest = n > 0 ? 1 : 0;
        for (int r = n - 1; r >= 0; --r)
            if (rp32[r + 1] - rp32[r] > 1) {
                depth[r] = 1 + depth[col[rp32[r] + 1]];
                deepest = std::max(deepest, depth[r]);
            }
        p->depth_lb = deepest;
    }
    int64_t dims[3] = {0, 0, 0};
    p->grid = detect_grid(n, rp32.data(), col, dims);
    if (p->grid) std::copy(dims, dims + 3, p->grid_dims);
    hipError_t e = hipMemcpy(p->rowptr, rp32.data(), ((size_t)n + 1) * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->col, col, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->val, val, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        sptrsv_plan_destroy(p);
        return fail(SPTRSV_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
    }
    const auto t0 = std::chrono::steady_clock::now();
    rc = analyse(p);
    p->analysis_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc != SPTRSV_OK) { sptrsv_plan_destroy(p); return rc; }
    *out = p;
    return SPTRSV_OK;
}

int sptrsv_plan_create_csc(int n, int64_t nnz, const int* colptr, const int* rowidx,
                           const double* val, sptrsv_plan_t* out) {
    if (!out) return fail(SPTRSV_ERR_BAD_INPUT, "null plan output");
    *out = nullptr;
    SPT_CHECK(sptrsv_validate_csc(n, nnz, colptr, rowidx, val));
    // CSC(U) -> CSR(U) on the host is the same counting sort as the reference's transpose
    // (src/utils.cpp:128-159) applied the other way; it only feeds the upload.
    std::vector<int64_t> rp((size_t)n + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) rp[(size_t)rowidx[k] + 1]++;
    for (int i = 0; i < n; ++i) rp[i + 1] += rp[i];
    std::vector<int64_t> cur(rp.begin(), rp.end() - 1);
    std::vector<int> c((size_t)nnz);
    std::vector<double> v((size_t)nnz);
    // rows must get their diagonal first: walk columns ascending, so within each row the
    // entries arrive in ascending column order, and the diagonal (column == row) is the
    // smallest column of its row in an upper-triangular matrix.
    for (int j = 0; j < n; ++j)
        for (int k = colptr[j]; k < colptr[j + 1]; ++k) {
            const int64_t dst = cur[rowidx[k]]++;
            c[dst] = j;
            v[dst] = val[k];
        }
    return sptrsv_plan_create_csr(n, nnz, rp.data(), c.data(), v.data(), out);
}

int sptrsv_plan_info(sptrsv_plan_t p, int* n, int64_t* nnz, int* n_levels) {
    if (!p) return fail(SPTRSV_ERR_BAD_INPUT, "null plan");
    if (n) *n = p->n;
    if (nnz) *nnz = p->nnz;
    if (n_levels) *n_levels = p->n_levels;
    return SPTRSV_OK;
}

int sptrsv_plan_analysis_time(sptrsv_plan_t p, double* analysis_ms, double* levels_ms) {
    if (!p) return fail(SPTRSV_ERR_BAD_INPUT, "null plan");
    if (analysis_ms) *analysis_ms = p->analysis_ms;
    if (levels_ms) *levels_ms = p->levels_ms;
    return SPTRSV_OK;
}

int sptrsv_plan_get_indegree(sptrsv_plan_t p, int* indeg) {
    if (!p || !indeg) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    SPT_HIP(hipStreamSynchronize(p->stream));
    SPT_HIP(hipMemcpy(indeg, p->indeg, (size_t)p->n * sizeof(int), hipMemcpyDeviceToHost));
    return SPTRSV_OK;
}

int sptrsv_plan_get_levels(sptrsv_plan_t p, int* levels) {
    if (!p || !levels) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    SPT_HIP(hipStreamSynchronize(p->stream));
    SPT_HIP(hipMemcpy(levels, p->levels, (size_t)p->n * sizeof(int), hipMemcpyDeviceToHost));
    return SPTRSV_OK;
}

int sptrsv_plan_get_level_sets(sptrsv_plan_t p, int* level_ptr, int* level_rows) {
    if (!p) return fail(SPTRSV_ERR_BAD_INPUT, "null plan");
    SPT_HIP(hipStreamSynchronize(p->stream));
    if (level_ptr)
        SPT_HIP(hipMemcpy(level_ptr, p->level_ptr, ((size_t)p->n_levels + 1) * sizeof(int),
                          hipMemcpyDeviceToHost));
    if (level_rows)
        SPT_HIP(hipMemcpy(level_rows, p->level_rows, (size_t)p->n * sizeof(int), hipMemcpyDeviceToHost));
    return SPTRSV_OK;
}

int sptrsv_plan_get_csc(sptrsv_plan_t p, int* colptr, int* rowidx, double* val) {
    if (!p) return fail(SPTRSV_ERR_BAD_INPUT, "null plan");
    SPT_HIP(hipStreamSynchronize(p->stream));
    if (colptr) SPT_HIP(hipMemcpy(colptr, p->colptr, ((size_t)p->n + 1) * sizeof(int), hipMemcpyDeviceToHost));
    if (rowidx) SPT_HIP(hipMemcpy(rowidx, p->rowidx, (size_t)p->nnz * sizeof(int), hipMemcpyDeviceToHost));
    if (val) SPT_HIP(hipMemcpy(val, p->cval, (size_t)p->nnz * sizeof(double), hipMemcpyDeviceToHost));
    return SPTRSV_OK;
}

int sptrsv_plan_timing(sptrsv_plan_t p, int enable) {
    if (!p) return fail(SPTRSV_ERR_BAD_INPUT, "null plan");
    SPT_HIP(hipStreamSynchronize(p->stream));
    p->timing = enable != 0;
    p->events_used = 0;
    if (p->tr) {
        p->tr->timing = p->timing;
        p->tr->events_used = 0;
    }
    return SPTRSV_OK;
}

int sptrsv_plan_timing_read(sptrsv_plan_t p, double* total_ms, int* launches) {
    if (!p) return fail(SPTRSV_ERR_BAD_INPUT, "null plan");
    SPT_HIP(hipStreamSynchronize(p->stream));
    double tot = 0.0;
    int count = 0;
    for (sptrsv_plan_s* q : {p, p->tr}) {  // the companion's launches are this plan's too
        if (!q) continue;
        for (size_t i = 0; i < q->events_used; ++i) {
            float ms = 0.f;
            SPT_HIP(hipEventElapsedTime(&ms, q->events[i].first, q->events[i].second));
            tot += ms;
        }
        count += (int)q->events_used;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = count;
    return SPTRSV_OK;
}

int sptrsv_plan_sync(sptrsv_plan_t p) {
    if (!p) return fail(SPTRSV_ERR_BAD_INPUT, "null plan");
    SPT_HIP(hipStreamSynchronize(p->stream));
    int status = 0;
    for (sptrsv_plan_s* q : {p, p->tr}) {  // the companion plan keeps its own status word
        if (!q) continue;
        int st = 0;
        SPT_HIP(hipMemcpy(&st, q->ticket + 1, sizeof(int), hipMemcpyDeviceToHost));
        if (st != 0) {
            const int zero = 0;  // re-arm for the next solve
            SPT_HIP(hipMemcpy(q->ticket + 1, &zero, sizeof(int), hipMemcpyHostToDevice));
            status = st;
        }
    }
    if (status != 0) {
        return fail(SPTRSV_ERR_DEADLOCK,
                    "device watchdog fired: a dependency was never published (cyclic or "
                    "malformed input, or the wait limit is too small)");
    }
    return SPTRSV_OK;
}

}  // extern "C"

extern "C" int sptrsv_division_check(const double* n, const double* d, double* fast, double* ieee, int64_t count) {
    using namespace sptrsv;
    if (!n || !d || !fast || !ieee || count < 0) return fail(SPTRSV_ERR_BAD_INPUT, "division check: bad arguments");
    SPT_CHECK(ensure_device());
    if (count == 0) return SPTRSV_OK;
    const int g = (int)std::min<int64_t>((count + 255) / 256, 65536);
    k_division_check<<<g, 256>>>(n, d, fast, ieee, count);
    SPT_HIP(hipGetLastError());
    SPT_HIP(hipDeviceSynchronize());
    return SPTRSV_OK;
}
