// solve.hip -- gfx950 sync-free / level-set triangular solve kernels for U X = B.
//
// Families (SPTRSV_VAR_*):
//   ROW_PULL / LEVEL_PULL  k_pull: G lanes per CSR row (lanes split the RHS columns),
//       rows handed out by wave tickets in descending order (ROW) or level order (LEVEL,
//       over the level-ordered CSR).  A lane waits for each needed x value by polling it:
//       unknown values hold a signalling-NaN sentinel, so x is published by one relaxed
//       agent-scope store with no separate flag and no L2 write-back.  Products are summed
//       in ascending column order without FMA contraction and divided by the diagonal --
//       the reference serial solver's arithmetic (src/core.cpp:18-32), so FP64 results are
//       bitwise equal to it.
//   ROW_PULL with m == 1 and group > 1  k_pull_split: G lanes split the ENTRIES of a row
//       (long rows), partial sums reduced across the group with cross-lane shuffles.
//   CSC_PUSH  k_push: one wave per CSC column, the reference's algorithm
//       (src/kernel_full.cl:54-228): wait for in-degree == 1, x = (b - left_sum) / diag,
//       scatter x * U(:,j) into left_sum with native FP64 atomics, release, decrement the
//       in-degree of every touched row.
//   LEVELSET  k_levelset (levelset.hip): bulk-synchronous, one cooperative launch walking the
//       levels with a grid barrier between them, each level's rows staged in LDS in chunks
//       (src/core.cpp:273-376 without the OpenMP task split); exact order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "plan.h"

// Work memory of sptrsv_solve_push_noalloc: one fixed array per GPU, part of the code object's
// own device memory and held until the process ends.  Layout: bytes [0, 256) ticket and status
// words, then SPTRSV_NOALLOC_MAX_ROWS consumed in-degrees (int), then SPTRSV_NOALLOC_MAX_ELEMS
// left_sum values (double) -- 24 MiB in all.  Host code never uses this name as an address: the
// per-device address comes from hipGetSymbolAddress (noalloc_scratch below).
constexpr size_t kNoAllocIndegOff = 256;
constexpr size_t kNoAllocSumOff = kNoAllocIndegOff + (size_t)SPTRSV_NOALLOC_MAX_ROWS * sizeof(int);
constexpr size_t kNoAllocBytes = kNoAllocSumOff + (size_t)SPTRSV_NOALLOC_MAX_ELEMS * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace sptrsv {
namespace {

template <typename T>
struct PullArgs {
    int n;             // rows
    int m;             // RHS count (row stride of b and x)
    int k0;            // first RHS column handled by this launch
    int kc;            // number of RHS columns handled by this launch
    const int* rowptr; // CSR (natural order or level order)
    const int* col;
    const T* val;
    const int* perm;   // level order: CSR row k is original row perm[k]; null otherwise
    const T* b;
    T* x;
    int* ctrl;         // [0] ticket, [1] status
    uint64_t limit;    // watchdog ticks
};

// Workgroup size of the pull kernels: 256 threads for one lane per row (finer tickets keep the
// dependency wavefront dense), 1024 when several lanes share a row (fewer ticket atomics).
constexpr int kPullBlock = 1024;
constexpr int kPullBlockRow = 256;

// The row's entries are taken in windows of UNROLL.  A window's indices and values are loaded
// once; each pass polls only the window's entries still holding the sentinel (an x value, once
// seen, never changes), then adds the ready prefix in ascending order.  Re-polling only the
// missing values keeps the wait traffic to about one load per waiting lane and pass.
template <typename T, int G, int KPL, int UNROLL, bool LEVELED, int BS>
__global__ __launch_bounds__(BS) void k_pull(PullArgs<T> a) {
#pragma clang fp contract(off)
    using U = typename Bits<T>::U;
    const int lane = threadIdx.x & (kWave - 1);
    const int t = block_ticket(&a.ctrl[0]);
    const int slot = t * (BS / G) + (int)threadIdx.x / G;
    const int gl = lane % G;
    bool done = slot >= a.n || gl >= a.kc;
    int row_out = 0, p = 0, end = 0, cnt = 0, q = 0;
    T diag = T(1);
    T acc[KPL], bv[KPL];
    int c[UNROLL];
    T v[UNROLL];
    U xb[UNROLL][KPL];
    const size_t m = (size_t)a.m;
#pragma unroll
    for (int j = 0; j < KPL; ++j) acc[j] = bv[j] = T(0);
    auto load_window = [&]() {
        cnt = min(UNROLL, end - p);
        q = 0;
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            c[u] = 0;
            v[u] = T(0);
            if (u < cnt) {
                c[u] = a.col[p + u];
                v[u] = a.val[p + u];
            }
#pragma unroll
            for (int j = 0; j < KPL; ++j) xb[u][j] = Bits<T>::sentinel;
        }
    };
    if (!done) {
        const int row_csr = LEVELED ? slot : a.n - 1 - slot;
        row_out = LEVELED ? a.perm[slot] : row_csr;
        const int s = a.rowptr[row_csr];
        diag = a.val[s];
        p = s + 1;
        end = a.rowptr[row_csr + 1];
#pragma unroll
        for (int j = 0; j < KPL; ++j)
            if (gl + j * G < a.kc) bv[j] = a.b[(size_t)row_out * m + a.k0 + gl + j * G];
        load_window();
    }
    wait_vmem();
    Watchdog wd;
    wd.arm(&a.ctrl[1], a.limit);
    int idle = 0, streak = 0;
    while (true) {
        bool progressed = false;
        if (!done) {
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
#pragma unroll
                for (int j = 0; j < KPL; ++j) {
                    const int k = gl + j * G;
                    if (u >= q && u < cnt && k < a.kc && xb[u][j] == Bits<T>::sentinel)
                        xb[u][j] = poll_bits(a.x + (size_t)c[u] * m + a.k0 + k);
                }
            }
            bool stop = false;
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                if (u >= q && u < cnt && !stop) {
                    bool ready = true;
#pragma unroll
                    for (int j = 0; j < KPL; ++j)
                        if (gl + j * G < a.kc && xb[u][j] == Bits<T>::sentinel) ready = false;
                    if (ready) {
#pragma unroll
                        for (int j = 0; j < KPL; ++j) acc[j] = acc[j] + v[u] * from_bits<T>(xb[u][j]);
                        q = u + 1;
                        progressed = true;
                    } else {
                        stop = true;
                    }
                }
            }
            if (q == cnt) {
                p += cnt;
                if (p < end) {
                    load_window();
                } else {
#pragma unroll
                    for (int j = 0; j < KPL; ++j) {
                        const int k = gl + j * G;
                        if (k < a.kc) publish(a.x + (size_t)row_out * m + a.k0 + k, (bv[j] - acc[j]) / diag);
                    }
                    done = true;
                }
                progressed = true;
            }
        }
        if (__all(done)) break;
        // Back off while nothing moves: a wave far ahead of the dependency front polls less,
        // which leaves the fabric to the polls at the front (ic0_128 m=1: 0.70 -> 0.675 ms).
        if (__any(progressed)) streak = 0;
        if (!__any(progressed)) {
            ++streak;
            if (streak < 2) __builtin_amdgcn_s_sleep(1);
            else if (streak < 8) __builtin_amdgcn_s_sleep(4);
            else __builtin_amdgcn_s_sleep(12);
            if ((++idle & 63) == 0 && !done && wd.expired()) done = true;
        }
    }
}

// m == 1, G lanes per row splitting the row's entries (lane e of a group takes entries
// e, e+G, ...).  Rows of one wave can depend on each other, so a group publishes from inside
// the polling loop: every iteration all lanes reduce their partial sums across the group
// (cross-lane shuffles need every lane active), and a group whose lanes have all finished
// their share publishes once.
template <typename T, int G, bool LEVELED>
__global__ __launch_bounds__(kPullBlock) void k_pull_split(PullArgs<T> a) {
    using U = typename Bits<T>::U;
    const int lane = threadIdx.x & (kWave - 1);
    const int t = block_ticket(&a.ctrl[0]);
    const int slot = t * (kPullBlock / G) + (int)threadIdx.x / G;
    const int gl = lane % G;
    const bool has_row = slot < a.n;
    bool share_done = !has_row, published = !has_row;
    int row_out = 0, p = 0, end = 0;
    T diag = T(1), acc = T(0), bv = T(0);
    if (has_row) {
        const int row_csr = LEVELED ? slot : a.n - 1 - slot;
        row_out = LEVELED ? a.perm[slot] : row_csr;
        const int s = a.rowptr[row_csr];
        diag = a.val[s];
        p = s + 1 + gl;
        end = a.rowptr[row_csr + 1];
        bv = a.b[row_out];
    }
    wait_vmem();
    const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << ((lane / G) * G));
    Watchdog wd;
    wd.arm(&a.ctrl[1], a.limit);
    int idle = 0;
    while (true) {
        bool progressed = false;
        if (!share_done) {
            while (p < end) {
                const U xb = poll_bits(a.x + a.col[p]);
                if (xb == Bits<T>::sentinel) break;
                acc += a.val[p] * from_bits<T>(xb);
                p += G;
                progressed = true;
            }
            if (p >= end) share_done = true;
        }
        const bool group_done = (__ballot(share_done) & gmask) == gmask;
        T tot = acc;
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) tot += __shfl_xor(tot, off, G);
        if (group_done && !published) {
            if (gl == 0) publish(a.x + row_out, (bv - tot) / diag);
            published = true;
            progressed = true;
        }
        if (__all(published)) break;
        if (!__any(progressed)) {
            __builtin_amdgcn_s_sleep(1);
            if ((++idle & 63) == 0 && !share_done && wd.expired()) share_done = true;
        }
    }
}

template <typename T>
struct PushArgs {
    int n, m;
    const int* colptr;
    const int* rowidx;
    const T* val;
    int* indeg;      // consumed
    T* left_sum;     // n*m, zeroed
    const T* b;
    T* x;
    int* ctrl;
    uint64_t limit;
};

// The reference's sync-free algorithm (src/kernel_full.cl:54-228): column j waits until its
// in-degree counter reads 1 (only its diagonal is left), forms x_j = (b_j - left_sum_j) / diag,
// adds val * x_j to the left_sum of every row it feeds and decrements their counters.  No
// fences: left_sum and the counters are device-scope atomics (performed past the XCDs' L2s), a
// column reads its own left_sum with a device-scope load once its counter reads 1, and a
// column's additions are waited for (vmcnt 0: performed) before its decrements are issued.
//
// m = 1: one lane per column.  Persistent waves take tickets of 64 columns in descending order;
// a wave only waits on columns of smaller tickets, held by running waves.
template <typename T>
__global__ __launch_bounds__(256) void k_push1(PushArgs<T> a) {
    const int lane = threadIdx.x & (kWave - 1);
    while (true) {
        const int t = wave_ticket(&a.ctrl[0]);
        if (t * kWave >= a.n) break;
        const int j = a.n - 1 - (t * kWave + lane);
        bool done = j < 0;
        int cs = 0, ce = 0;
        T diag = T(1);
        if (!done) {
            cs = a.colptr[j];
            ce = a.colptr[j + 1] - 1;
            diag = a.val[ce];
        }
        Watchdog wd;
        wd.arm(&a.ctrl[1], a.limit);
        int idle = 0;
        while (true) {
            const bool ready = !done && __hip_atomic_load(&a.indeg[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 1;
            if (ready) {
                const T xj = (a.b[j] - from_bits<T>(poll_bits(a.left_sum + j))) / diag;
                a.x[j] = xj;
                for (int e = cs; e < ce; ++e) atomicAdd(&a.left_sum[a.rowidx[e]], a.val[e] * xj);
            }
            wait_vmem();  // this round's additions are performed
            if (ready) {
                for (int e = cs; e < ce; ++e)
                    __hip_atomic_fetch_add(&a.indeg[a.rowidx[e]], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                done = true;
            }
            if (__all(done)) break;
            if (!__any(ready)) {
                __builtin_amdgcn_s_sleep(1);
                if ((++idle & 63) == 0 && !done && wd.expired()) done = true;
            }
        }
        if (__hip_atomic_load(&a.ctrl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;  // watchdog
    }
}

// One wave per column, descending by ticket; its lanes take the right-hand sides.
template <typename T>
__global__ __launch_bounds__(64) void k_push(PushArgs<T> a) {
    const int lane = threadIdx.x & (kWave - 1);
    const int t = wave_ticket(&a.ctrl[0]);
    const int j = a.n - 1 - t;
    if (j < 0) return;
    const int cs = a.colptr[j], ce = a.colptr[j + 1] - 1;
    const T diag = a.val[ce];
    if (lane == 0) {
        Watchdog wd;
        wd.arm(&a.ctrl[1], a.limit);
        int idle = 0;
        while (__hip_atomic_load(&a.indeg[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 1) {
            __builtin_amdgcn_s_sleep(1);
            if ((++idle & 63) == 0 && wd.expired()) break;
        }
    }
    __builtin_amdgcn_wave_barrier();
    const size_t m = (size_t)a.m;
    if (a.m == 1) {  // one value: lane 0 forms it, the lanes share the column's entries
        T xj = T(0);
        if (lane == 0) {
            xj = (a.b[j] - from_bits<T>(poll_bits(a.left_sum + j))) / diag;
            a.x[j] = xj;
        }
        xj = __shfl(xj, 0, kWave);
        for (int e = cs + lane; e < ce; e += kWave) atomicAdd(&a.left_sum[a.rowidx[e]], a.val[e] * xj);
    } else {  // lane k holds x_j of right-hand side k (64 at a time) in a register
        for (int k0 = 0; k0 < a.m; k0 += kWave) {
            const int k = k0 + lane;
            if (k < a.m) {
                const size_t at = (size_t)j * m + k;
                const T xk = (a.b[at] - from_bits<T>(poll_bits(a.left_sum + at))) / diag;
                a.x[at] = xk;
                for (int e = cs; e < ce; ++e) atomicAdd(&a.left_sum[(size_t)a.rowidx[e] * m + k], a.val[e] * xk);
            }
        }
    }
    wait_vmem();  // the column's additions are performed before its decrements
    for (int e = cs + lane; e < ce; e += kWave)
        __hip_atomic_fetch_add(&a.indeg[a.rowidx[e]], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// launch of the push solve: lane per column at m = 1 on factors with wide levels, wave per
// column otherwise (on a near-chain a wave's 64 columns would resolve one after another, each
// round paying its own wait for the additions)
template <typename T>
void launch_push(const PushArgs<T>& a, hipStream_t s, int n_levels) {
    if (a.m == 1 && (int64_t)n_levels * 4 <= a.n)
        k_push1<T><<<std::max(1, std::min((a.n + 255) / 256, device_compute_units() * 8)), 256, 0, s>>>(a);
    else
        k_push<T><<<a.n, 64, 0, s>>>(a);
}

template <typename T>
__global__ void k_fill_bits(T* out, typename Bits<T>::U bits, int64_t count) {
    using U = typename Bits<T>::U;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < count; i += (int64_t)gridDim.x * blockDim.x) reinterpret_cast<U*>(out)[i] = bits;
}

__global__ void k_copy_int(const int* __restrict__ in, int* __restrict__ out, int64_t count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < count; i += (int64_t)gridDim.x * blockDim.x) out[i] = in[i];
}

__global__ void k_reset_ticket(int* ctrl) {
    if (threadIdx.x == 0) ctrl[0] = 0;
}

// ticket and status of a solve whose control words are not the plan's own
__global__ void k_reset_ctrl(int* ctrl) {
    if (threadIdx.x < 2) ctrl[threadIdx.x] = 0;
}

// a watchdog that fired in `ctrl` is reported through the plan's status word (sptrsv_plan_sync)
__global__ void k_carry_status(const int* ctrl, int* plan_ctrl) {
    if (threadIdx.x == 0 && ctrl[1] != 0) plan_ctrl[1] = ctrl[1];
}

template <typename Tin, typename Tout>
__global__ void k_convert(const Tin* __restrict__ in, Tout* __restrict__ out, int64_t count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < count; i += (int64_t)gridDim.x * blockDim.x) out[i] = (Tout)in[i];
}

// r = b - U x, FP64, G lanes per row over the RHS columns; products in ascending order.
__global__ __launch_bounds__(256) void k_residual(int n, int m, const int* __restrict__ rowptr,
                                                  const int* __restrict__ col, const double* __restrict__ val,
                                                  const double* __restrict__ x, const double* __restrict__ b,
                                                  double* __restrict__ r) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)n * m;
    for (int64_t q = gid; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int row = (int)(q / m), k = (int)(q % m);
        double acc = 0.0;
        for (int e = rowptr[row]; e < rowptr[row + 1]; ++e) acc = fma(val[e], x[(size_t)col[e] * m + k], acc);
        r[q] = b[q] - acc;
    }
}


inline int grid_for(int64_t count, int block = 256, int cap = 65536) {
    int64_t g = (count + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// out = J in for row-major n x m (J reverses the row order)
__global__ void k_reverse_rows(const double* __restrict__ in, double* __restrict__ out, int64_t n, int m) {
    const int64_t count = n * m;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / m, k = e - i * m;
        out[(n - 1 - i) * m + k] = in[e];
    }
}

// U^T x = b is J (J U^T J) J x = b, and J U^T J is upper triangular: its row i' holds column
// n-1-i' of U, entries reversed, so the diagonal (last in a column of the CSC copy) comes first
// and the columns n-1-r ascend.  The companion plan is built once from the CSC copy and runs
// on this plan's stream.
int ensure_transposed(sptrsv_plan_s* P) {
    if (P->tr) return SPTRSV_OK;
    const int n = P->n;
    const int64_t nnz = P->nnz;
    std::vector<int> cp((size_t)n + 1), ri((size_t)nnz);
    std::vector<double> cv((size_t)nnz);
    SPT_HIP(hipStreamSynchronize(P->stream));
    SPT_HIP(hipMemcpy(cp.data(), P->colptr, cp.size() * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(ri.data(), P->rowidx, ri.size() * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(cv.data(), P->cval, cv.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int64_t> rp((size_t)n + 1, 0);
    std::vector<int> c((size_t)nnz);
    std::vector<double> v((size_t)nnz);
    int64_t k = 0;
    for (int ip = 0; ip < n; ++ip) {
        const int j = n - 1 - ip;
        for (int e = cp[j + 1] - 1; e >= cp[j]; --e) {
            c[(size_t)k] = n - 1 - ri[(size_t)e];
            v[(size_t)k] = cv[(size_t)e];
            ++k;
        }
        rp[(size_t)ip + 1] = k;
    }
    sptrsv_plan_t t = nullptr;
    SPT_CHECK(sptrsv_plan_create_csr(n, nnz, rp.data(), c.data(), v.data(), &t));
    (void)hipStreamSynchronize(t->stream);
    (void)hipStreamDestroy(t->stream);
    t->stream = P->stream;
    t->borrowed_stream = true;
    t->timing = P->timing;
    P->tr = t;
    return SPTRSV_OK;
}

int ensure_tr_buf(sptrsv_plan_s* P, int m) {
    const size_t need = 3 * (size_t)P->n * (size_t)m;
    if (P->tr_elems >= need) return SPTRSV_OK;
    SPT_HIP(hipStreamSynchronize(P->stream));
    if (P->tr_buf) SPT_HIP(hipFree(P->tr_buf));
    P->tr_buf = nullptr;
    P->tr_elems = 0;
    SPT_HIP(hipMalloc((void**)&P->tr_buf, need * sizeof(double)));
    P->tr_elems = need;
    return SPTRSV_OK;
}

// Dynamic LDS that lets at most `occ` workgroups of a pull kernel reside on one CU (160 KiB of
// LDS per CU; occ <= 0 = no cap).  Fewer resident rows far ahead of the dependency front means
// fewer lanes polling x values that cannot be ready yet: every poll of an unpublished value is
// a device-scope load that crosses the fabric, and thousands of them in flight lengthen the
// latency of the loads that matter.
inline size_t occupancy_lds(int occ) {
    if (occ <= 0) return 0;
    return (size_t)(163840 / occ - 512) & ~(size_t)255;
}

// Default pull workgroup: 64 rows (one wave at one lane per row), at most 1024 threads.
inline int default_block(int G) { return std::min(1024, 64 * G); }

// Default residency: about three mean level widths of rows in flight over the whole chip
// (measured on the 128^3 IC(0) factor, whose mean level holds 5,490 rows: one 64-row workgroup
// per CU beats the hardware limit 2.9x at one right-hand side).  Wide-level factors get the
// hardware limit back, since their rows do not wait.
inline int default_occupancy(const sptrsv_plan_s* P, int rows_per_block, int bs) {
    const double width = (double)P->n / std::max(P->n_levels, 1);
    const double rows_per_unit = (double)device_compute_units() * rows_per_block;
    const int occ = std::max(1, (int)std::lround(3.0 * width / rows_per_unit));
    const int hw = std::max(1, 2048 / bs);  // 32 waves per CU
    return occ >= hw ? -1 : occ;
}

inline int next_pow2(int v) {
    int g = 1;
    while (g < v) g <<= 1;
    return g;
}

// One pull launch with workgroup size `bs` (64..1024 threads, BS / G rows per workgroup).
template <typename T, int G, int K, int U, bool LEVELED>
void run_pull(int bs, int n, size_t lds, hipStream_t s, const PullArgs<T>& a) {
    const dim3 grid((unsigned)(((int64_t)n * G + bs - 1) / bs));
    switch (bs) {
        case 64: k_pull<T, G, K, U, LEVELED, 64><<<grid, 64, lds, s>>>(a); break;
        case 128: k_pull<T, G, K, U, LEVELED, 128><<<grid, 128, lds, s>>>(a); break;
        case 256: k_pull<T, G, K, U, LEVELED, 256><<<grid, 256, lds, s>>>(a); break;
        case 512: k_pull<T, G, K, U, LEVELED, 512><<<grid, 512, lds, s>>>(a); break;
        default: k_pull<T, G, K, U, LEVELED, 1024><<<grid, 1024, lds, s>>>(a); break;
    }
}

template <typename T, bool LEVELED>
int launch_pull(sptrsv_plan_s* P, const PullArgs<T>& base, int m, int group, int unroll, int occ, int block) {
    // lanes per row: m == 1 -> 1 (or `group` lanes splitting entries); m <= 64 -> next pow2 of
    // m; m > 64 -> 64 lanes with up to 4 columns each, launched in chunks of 256 columns.
    // LEVEL_PULL with m > 1 takes `group` (8..64) as the lanes per row instead: fewer lanes
    // per row put more rows in flight per wave, each lane then takes up to 8 columns.
    hipStream_t s = P->stream;
    const int n = P->n;
    if (m == 1 && group > 1) {
        const int G = next_pow2(std::min(group, 64));
        PullArgs<T> a = base;
        a.k0 = 0;
        a.kc = 1;
        hipEvent_t stop;
        SPT_CHECK(begin_timed(P, &stop));
        k_reset_ticket<<<1, 64, 0, s>>>(a.ctrl);
        const dim3 grid((unsigned)(((int64_t)n * G + kPullBlock - 1) / kPullBlock));
        switch (G) {
            case 2: k_pull_split<T, 2, LEVELED><<<grid, kPullBlock, 0, s>>>(a); break;
            case 4: k_pull_split<T, 4, LEVELED><<<grid, kPullBlock, 0, s>>>(a); break;
            case 8: k_pull_split<T, 8, LEVELED><<<grid, kPullBlock, 0, s>>>(a); break;
            case 16: k_pull_split<T, 16, LEVELED><<<grid, kPullBlock, 0, s>>>(a); break;
            case 32: k_pull_split<T, 32, LEVELED><<<grid, kPullBlock, 0, s>>>(a); break;
            default: k_pull_split<T, 64, LEVELED><<<grid, kPullBlock, 0, s>>>(a); break;
        }
        end_timed(P, stop);
        SPT_HIP(hipGetLastError());
        return SPTRSV_OK;
    }
    for (int k0 = 0; k0 < m; k0 += 256) {
        const int kc = std::min(256, m - k0);
        PullArgs<T> a = base;
        a.k0 = k0;
        a.kc = kc;
        int G = kc == 1 ? 1 : std::min(64, next_pow2(kc));
        if (LEVELED && kc > 1 && group > 0) {
            G = std::min(G, std::max(8, next_pow2(std::min(group, 64))));
            while ((kc + G - 1) / G > 8) G *= 2;
        }
        const int kpl = next_pow2((kc + G - 1) / G);  // 1..8; 8 and G < 64 only under `group`
        // Row order needs the dependency distance in flight, not a few levels: its default is
        // the hardware limit with 256- or 1024-thread workgroups.
        int bs = block > 0 ? next_pow2(std::min(std::max(block, 64), 1024))
                           : (LEVELED ? default_block(G) : (G == 1 ? kPullBlockRow : kPullBlock));
        bs = std::max(bs, G);
        const size_t lds = occupancy_lds(occ != 0 ? occ : LEVELED ? default_occupancy(P, bs / G, bs) : -1);
        hipEvent_t stop;
        SPT_CHECK(begin_timed(P, &stop));
        k_reset_ticket<<<1, 64, 0, s>>>(a.ctrl);
        if (G == 1) {
            if (unroll == 1) run_pull<T, 1, 1, 1, LEVELED>(bs, n, lds, s, a);
            else if (unroll == 2) run_pull<T, 1, 1, 2, LEVELED>(bs, n, lds, s, a);
            else run_pull<T, 1, 1, 4, LEVELED>(bs, n, lds, s, a);
        } else if (G == 2) run_pull<T, 2, 1, 4, LEVELED>(bs, n, lds, s, a);
        else if (G == 4) run_pull<T, 4, 1, 4, LEVELED>(bs, n, lds, s, a);
        else if (G < 64 && kpl > 1) {
            if constexpr (LEVELED) {
                if (G == 8) {
                    if (kpl == 2) run_pull<T, 8, 2, 4, true>(bs, n, lds, s, a);
                    else if (kpl == 4) run_pull<T, 8, 4, 4, true>(bs, n, lds, s, a);
                    else run_pull<T, 8, 8, 4, true>(bs, n, lds, s, a);
                } else if (G == 16) {
                    if (kpl == 2) run_pull<T, 16, 2, 4, true>(bs, n, lds, s, a);
                    else if (kpl == 4) run_pull<T, 16, 4, 4, true>(bs, n, lds, s, a);
                    else run_pull<T, 16, 8, 4, true>(bs, n, lds, s, a);
                } else {
                    if (kpl == 2) run_pull<T, 32, 2, 4, true>(bs, n, lds, s, a);
                    else if (kpl == 4) run_pull<T, 32, 4, 4, true>(bs, n, lds, s, a);
                    else run_pull<T, 32, 8, 4, true>(bs, n, lds, s, a);
                }
            }
        } else if (G == 8) run_pull<T, 8, 1, 4, LEVELED>(bs, n, lds, s, a);
        else if (G == 16) run_pull<T, 16, 1, 4, LEVELED>(bs, n, lds, s, a);
        else if (G == 32) run_pull<T, 32, 1, 4, LEVELED>(bs, n, lds, s, a);
        else if (kpl == 1) run_pull<T, 64, 1, 4, LEVELED>(bs, n, lds, s, a);
        else if (kpl == 2) run_pull<T, 64, 2, 4, LEVELED>(bs, n, lds, s, a);
        else run_pull<T, 64, 4, 4, LEVELED>(bs, n, lds, s, a);
        end_timed(P, stop);
        SPT_HIP(hipGetLastError());
    }
    return SPTRSV_OK;
}

template <typename T>
int solve_typed(sptrsv_plan_s* P, int variant, int group, int unroll, int occ, int block, const T* d_b, T* d_x, int m,
                const T* val, const T* cval, const T* lval, uint64_t limit) {
    hipStream_t s = P->stream;
    const int n = P->n;
    const int64_t nm = (int64_t)n * m;
    switch (variant) {
        case SPTRSV_VAR_ROW_PULL:
        case SPTRSV_VAR_LEVEL_PULL: {
            k_fill_bits<T><<<grid_for(nm), 256, 0, s>>>(d_x, Bits<T>::sentinel, nm);
            PullArgs<T> a{};
            a.n = n;
            a.m = m;
            a.b = d_b;
            a.x = d_x;
            a.ctrl = P->ticket;
            a.limit = limit;
            if (variant == SPTRSV_VAR_ROW_PULL) {
                a.rowptr = P->rowptr;
                a.col = P->col;
                a.val = val;
                a.perm = nullptr;
                return launch_pull<T, false>(P, a, m, group, unroll, occ, block);
            }
            a.rowptr = P->lrowptr;
            a.col = P->lcol;
            a.val = lval;
            a.perm = P->level_rows;
            return launch_pull<T, true>(P, a, m, group, unroll, occ, block);
        }
        case SPTRSV_VAR_CSC_PUSH: {
            const size_t need = (size_t)nm * sizeof(T);
            if (P->left_sum_elems * sizeof(double) < need) {
                if (P->left_sum) SPT_HIP(hipFree(P->left_sum));
                P->left_sum = nullptr;
                SPT_HIP(hipMalloc((void**)&P->left_sum, need));
                P->left_sum_elems = (need + sizeof(double) - 1) / sizeof(double);
            }
            T* ls = reinterpret_cast<T*>(P->left_sum);
            k_fill_bits<T><<<grid_for(nm), 256, 0, s>>>(ls, 0, nm);
            k_copy_int<<<grid_for(n), 256, 0, s>>>(P->indeg, P->indeg_work, n);
            k_reset_ticket<<<1, 64, 0, s>>>(P->ticket);
            PushArgs<T> a{n, m, P->colptr, P->rowidx, cval, P->indeg_work, ls, d_b, d_x, P->ticket, limit};
            hipEvent_t stop;
            SPT_CHECK(begin_timed(P, &stop));
            launch_push<T>(a, s, P->n_levels);
            end_timed(P, stop);
            SPT_HIP(hipGetLastError());
            return SPTRSV_OK;
        }
        case SPTRSV_VAR_LEVELSET:
            return levelset_solve<T>(P, lval, d_b, d_x, m, limit);
        default:
            return fail(SPTRSV_ERR_UNSUPPORTED, "solver variant " + std::to_string(variant) + " not available");
    }
}

}  // namespace

int convert_d2f(const double* in, float* out, int64_t count, hipStream_t s) {
    k_convert<double, float><<<grid_for(count), 256, 0, s>>>(in, out, count);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

int convert_f2d(const float* in, double* out, int64_t count, hipStream_t s) {
    k_convert<float, double><<<grid_for(count), 256, 0, s>>>(in, out, count);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

// AUTO: the learned selector's choice when a network is installed (sptrsv_selector_set);
// otherwise this hand rule: one right-hand side over a factor whose levels hold fewer than four rows on average
// (a dependency chain) goes to the single-wave chain kernel, which keeps each window of rows in
// LDS instead of paying a device-wide round trip per level; one to three right-hand sides in
// FP64 over a natural-order grid factor to the block kernel (bricks solved in LDS; with four,
// its per-step work outgrows the level-ordered pull kernel's); a plan with a supernodal layout
// built by the caller, in FP64, to the supernodal MFMA kernel, and so a factor of 16 or more
// entries per row whose relaxed supernodes hold 32 rows or more on average (built here once, on
// first use); 32 or more in FP64 over a grid
// factor to the brick-panel kernel (on the 128^3 factor faster per solve than LEVEL_PULL from
// m = 32 on, slower at 8 and 16); everything else to LEVEL_PULL.
// With `exact`, only kernels whose FP64 x is bitwise equal to the serial solver qualify.
int default_variant(sptrsv_plan_s* P, int m, int precision, bool exact) {
    int picked = 0;
    if (selector_pick(P, m, precision, exact, &picked)) return picked;  // the learned selector
    const bool chain_like = (int64_t)P->n_levels * 4 > (int64_t)P->n;
    // a supernodal layout the caller built (sptrsv_plan_build_supernodes): the MFMA kernel
    if (!exact && P->sn > 0 && P->relaxed_state != 1 && precision == SPTRSV_FP64) return SPTRSV_VAR_SUPERNODAL;
    if (!exact && precision == SPTRSV_FP64 && relaxed_supernodes_pay(P)) return SPTRSV_VAR_SUPERNODAL;
    if (!exact && m == 1 && chain_like) return SPTRSV_VAR_CHAIN;
    if (m <= 3 && precision == SPTRSV_FP64 && P->grid && block_supports(P, m)) return SPTRSV_VAR_BLOCK;
    if (m >= 32 && precision == SPTRSV_FP64 && panel_supports(P)) return SPTRSV_VAR_PANEL;
    return SPTRSV_VAR_LEVEL_PULL;
}

int ensure_stage(sptrsv_plan_s* P, int64_t nm) {
    if ((int64_t)P->stage_elems >= nm) return SPTRSV_OK;
    SPT_HIP(hipStreamSynchronize(P->stream));
    if (P->b32) SPT_HIP(hipFree(P->b32));
    if (P->x32) SPT_HIP(hipFree(P->x32));
    P->b32 = P->x32 = nullptr;
    P->stage_elems = 0;
    SPT_HIP(hipMalloc((void**)&P->b32, nm * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&P->x32, nm * sizeof(float)));
    P->stage_elems = (size_t)nm;
    return SPTRSV_OK;
}

// FP32 values, device FP32 b and x (the inner solve of the refinement, and sptrsv_solve at FP32).
int solve_fp32_device(sptrsv_plan_s* P, const sptrsv_params_t* prm, const float* d_b32, float* d_x32, int m) {
    const int variant = prm->variant;
    const uint64_t limit = prm->watchdog_us > 0
                               ? std::max<uint64_t>(1, P->ticks_per_ms * (uint64_t)prm->watchdog_us / 1000)
                               : P->ticks_per_ms * 20000ull;
    SPT_CHECK(ensure_fp32(P));
    if (variant == SPTRSV_VAR_CHAIN) {
        if (m != 1) return fail(SPTRSV_ERR_UNSUPPORTED, "chain variant solves one right-hand side");
        return chain_solve32(P, d_b32, d_x32);
    }
    if (variant == SPTRSV_VAR_TILE || variant == SPTRSV_VAR_BLOCK || variant == SPTRSV_VAR_SUPERNODAL ||
        variant == SPTRSV_VAR_PANEL)
        return fail(SPTRSV_ERR_UNSUPPORTED, "this variant has no FP32 form through sptrsv_solve");
    const int unroll = prm->unroll > 0 ? prm->unroll : 4;
    return solve_typed<float>(P, variant, prm->group, unroll, prm->occupancy, prm->block, d_b32, d_x32, m, P->val32,
                              P->cval32, P->lval32, limit);
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on device `dev`, looked up once per device with that device
// current (each GPU holds its own copy of the array at its own address).
int noalloc_scratch(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(SPTRSV_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        int cur = -1;
        SPT_HIP(hipGetDevice(&cur));
        if (cur != dev) SPT_HIP(hipSetDevice(dev));
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !ptr)
            return fail(SPTRSV_ERR_HIP, std::string("hipGetSymbolAddress(scratch) -> ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return SPTRSV_OK;
}

}  // namespace sptrsv

using namespace sptrsv;

extern "C" {

int sptrsv_plan_auto_variant(sptrsv_plan_t P, int m, int precision, int* variant) {
    if (!P || !variant) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    if (m < 1) return fail(SPTRSV_ERR_BAD_INPUT, "m must be >= 1");
    *variant = default_variant(P, m, precision, false);
    return SPTRSV_OK;
}

int sptrsv_solve(sptrsv_plan_t P, const sptrsv_params_t* params, const double* d_b, double* d_x, int m) {
    if (!P || !d_b || !d_x) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    if (m < 1) return fail(SPTRSV_ERR_BAD_INPUT, "m must be >= 1");
    sptrsv_params_t prm{};
    if (params) prm = *params;
    int variant = prm.variant == SPTRSV_VAR_AUTO ? default_variant(P, m, prm.precision, prm.exact != 0) : prm.variant;
    if (variant < 1 || variant >= SPTRSV_VAR_COUNT)
        return fail(SPTRSV_ERR_UNSUPPORTED, "solver variant " + std::to_string(variant) + " not available");
    const uint64_t limit = prm.watchdog_us > 0 ? std::max<uint64_t>(1, P->ticks_per_ms * (uint64_t)prm.watchdog_us / 1000)
                                            : P->ticks_per_ms * 20000ull;
    const int unroll = prm.unroll > 0 ? prm.unroll : 4;
    if (variant == SPTRSV_VAR_SUPERNODAL) return supernodal_solve(P, prm.precision, d_b, d_x, m, limit);
    if (variant == SPTRSV_VAR_TILE) return tile_solve(P, prm.precision, prm.block, d_b, d_x, m, limit);
    if (variant == SPTRSV_VAR_BLOCK) return block_solve(P, prm.precision, prm.occupancy, d_b, d_x, m, limit);
    if (variant == SPTRSV_VAR_PANEL) return panel_solve(P, prm.precision, d_b, d_x, m, limit);
    if (variant == SPTRSV_VAR_CHAIN && prm.precision == SPTRSV_FP64) return chain_solve(P, prm.precision, d_b, d_x, m);
    if (prm.precision == SPTRSV_FP64)
        return solve_typed<double>(P, variant, prm.group, unroll, prm.occupancy, prm.block, d_b, d_x, m, P->val, P->cval, P->lval, limit);
    if (prm.precision == SPTRSV_FP32) {
        SPT_CHECK(ensure_fp32(P));
        const int64_t nm = (int64_t)P->n * m;
        SPT_CHECK(ensure_stage(P, nm));
        sptrsv_params_t low = prm;
        low.variant = variant;
        k_convert<double, float><<<grid_for(nm), 256, 0, P->stream>>>(d_b, P->b32, nm);
        SPT_CHECK(solve_fp32_device(P, &low, P->b32, P->x32, m));
        k_convert<float, double><<<grid_for(nm), 256, 0, P->stream>>>(P->x32, d_x, nm);
        SPT_HIP(hipGetLastError());
        return SPTRSV_OK;
    }
    return fail(SPTRSV_ERR_BAD_INPUT, "unknown precision");
}

// The CSC push solve (k_push) with its work memory in the library's fixed per-GPU array: no
// allocation.  Every call on a GPU shares that array, and each call clears what must start at
// zero (left_sum, the control words) and copies the in-degrees in, all with kernels on the
// plan's stream: calls that overlap on one GPU race, calls in one stream order do not.
int sptrsv_solve_push_noalloc(sptrsv_plan_t P, const double* d_b, double* d_x, int m) {
    if (!P || !d_b || !d_x) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    if (m < 1) return fail(SPTRSV_ERR_BAD_INPUT, "m must be >= 1");
    const int64_t n = P->n, nm = n * m;
    if (n > SPTRSV_NOALLOC_MAX_ROWS || nm > SPTRSV_NOALLOC_MAX_ELEMS)
        return fail(SPTRSV_ERR_UNSUPPORTED, "no-allocation solve: n = " + std::to_string(n) + ", n*m = " +
                                                std::to_string(nm) + " exceed the fixed work array (n <= " +
                                                std::to_string(SPTRSV_NOALLOC_MAX_ROWS) + ", n*m <= " +
                                                std::to_string(SPTRSV_NOALLOC_MAX_ELEMS) + ")");
    unsigned char* base = nullptr;
    SPT_CHECK(noalloc_scratch(P->device, &base));
    int* ctrl = reinterpret_cast<int*>(base);
    int* indeg = reinterpret_cast<int*>(base + kNoAllocIndegOff);
    double* ls = reinterpret_cast<double*>(base + kNoAllocSumOff);
    hipStream_t s = P->stream;
    const uint64_t limit = P->ticks_per_ms * 20000ull;
    k_fill_bits<double><<<grid_for(nm), 256, 0, s>>>(ls, 0, nm);
    k_copy_int<<<grid_for(n), 256, 0, s>>>(P->indeg, indeg, n);
    k_reset_ctrl<<<1, 64, 0, s>>>(ctrl);
    PushArgs<double> a{P->n, m, P->colptr, P->rowidx, P->cval, indeg, ls, d_b, d_x, ctrl, limit};
    launch_push<double>(a, s, P->n_levels);
    k_carry_status<<<1, 64, 0, s>>>(ctrl, P->ticket);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

int sptrsv_solve_host(sptrsv_plan_t P, const sptrsv_params_t* params, const double* b, double* x, int m,
                      int col_major) {
    if (!P || !b || !x) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    if (m < 1) return fail(SPTRSV_ERR_BAD_INPUT, "m must be >= 1");
    const size_t n = (size_t)P->n, count = n * (size_t)m;
    std::string err;
    double *db = nullptr, *dx = nullptr;
    SPT_HIP(hipMalloc((void**)&db, count * sizeof(double)));
    if (hipMalloc((void**)&dx, count * sizeof(double)) != hipSuccess) {
        (void)hipFree(db);
        return fail(SPTRSV_ERR_NOMEM, "device allocation of x failed");
    }
    int rc = SPTRSV_OK;
    if (col_major && m > 1) {
        // column-major host <-> row-major device
        std::vector<double> tmp(count);
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < m; ++k) tmp[i * m + k] = b[i + k * n];
        if (hipMemcpy(db, tmp.data(), count * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(SPTRSV_ERR_HIP, "upload of b failed");
        if (rc == SPTRSV_OK) rc = sptrsv_solve(P, params, db, dx, m);
        if (rc == SPTRSV_OK) rc = sptrsv_plan_sync(P);
        if (rc == SPTRSV_OK && hipMemcpy(tmp.data(), dx, count * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(SPTRSV_ERR_HIP, "download of x failed");
        if (rc == SPTRSV_OK)
            for (size_t i = 0; i < n; ++i)
                for (int k = 0; k < m; ++k) x[i + k * n] = tmp[i * m + k];
    } else {
        if (hipMemcpy(db, b, count * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(SPTRSV_ERR_HIP, "upload of b failed");
        if (rc == SPTRSV_OK) rc = sptrsv_solve(P, params, db, dx, m);
        if (rc == SPTRSV_OK) rc = sptrsv_plan_sync(P);
        if (rc == SPTRSV_OK && hipMemcpy(x, dx, count * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(SPTRSV_ERR_HIP, "download of x failed");
    }
    (void)hipFree(db);
    (void)hipFree(dx);
    return rc;
}


int sptrsv_solve_transposed(sptrsv_plan_t P, const sptrsv_params_t* params, const double* d_b, double* d_x,
                            int m) {
    if (!P || !d_b || !d_x) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    if (m < 1) return fail(SPTRSV_ERR_BAD_INPUT, "m must be >= 1");
    if (params && params->variant == SPTRSV_VAR_SUPERNODAL)
        return fail(SPTRSV_ERR_UNSUPPORTED, "transposed solve: the supernodal layout is built for U only");
    SPT_CHECK(ensure_transposed(P));
    SPT_CHECK(ensure_tr_buf(P, m));
    const int64_t n = P->n, nm = n * m;
    double* rb = P->tr_buf;
    double* rx = P->tr_buf + nm;
    k_reverse_rows<<<grid_for(nm), 256, 0, P->stream>>>(d_b, rb, n, m);
    SPT_HIP(hipGetLastError());
    SPT_CHECK(sptrsv_solve(P->tr, params, rb, rx, m));
    k_reverse_rows<<<grid_for(nm), 256, 0, P->stream>>>(rx, d_x, n, m);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

int sptrsv_solve_cholesky(sptrsv_plan_t P, const sptrsv_params_t* params, const double* d_b, double* d_x, int m) {
    if (!P || !d_b || !d_x) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    if (m < 1) return fail(SPTRSV_ERR_BAD_INPUT, "m must be >= 1");
    SPT_CHECK(ensure_tr_buf(P, m));
    double* y = P->tr_buf + 2 * (int64_t)P->n * m;
    sptrsv_params_t fwd{};
    if (params) fwd = *params;
    if (fwd.variant == SPTRSV_VAR_SUPERNODAL) fwd.variant = SPTRSV_VAR_AUTO;  // U^T has no supernodal layout
    SPT_CHECK(sptrsv_solve_transposed(P, &fwd, d_b, y, m));
    return sptrsv_solve(P, params, y, d_x, m);
}

int sptrsv_residual(sptrsv_plan_t P, const double* d_b, const double* d_x, double* d_r, int m) {
    if (!P || !d_b || !d_x || !d_r) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    const int64_t nm = (int64_t)P->n * m;
    k_residual<<<grid_for(nm), 256, 0, P->stream>>>(P->n, m, P->rowptr, P->col, P->val, d_x, d_b, d_r);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

}  // extern "C"
