// chain.hip -- single-wave window solver for factors whose dependency graph is (nearly) a chain.
//
// When almost every row depends on the next one (random Zipf factors: 948,744 levels for
// n = 1e6; banded Cholesky factors: n levels) there is no parallelism across rows to exploit,
// and any dependency that travels through HBM costs microseconds.  This kernel keeps the whole
// recurrence inside ONE wave:
//   * rows are taken in windows of 64 (lane l owns row hi-1-l), from the last window down;
//   * a window's coefficients on its own rows go to a 64x64 LDS tile T[s][lane];
//   * the window is then solved in 64 steps: at step s lane s forms x = (b - acc) * (1/diag),
//     v_readlane broadcasts it, every lane does acc += T[s][lane] * x;
//   * x of the last 4096 rows is kept in an LDS ring; older x values are read from HBM.  They
//     were published windows earlier, so the next window's row data and those old values are
//     prefetched into registers before the 64 steps and arrive while they run.
// The sum order differs from the serial reference (outside entries first, then the window in
// descending column order) and x is multiplied by the reciprocal diagonal like the reference's
// GPU kernel (src/kernel_full.cl:83,98), fused as x = fma(-acc, 1/d, b/d) to shorten the
// dependent chain per row: x agrees with the serial solver to rounding.
#include <hip/hip_runtime.h>

#include "common.h"
#include "plan.h"

namespace sptrsv {
namespace {

constexpr int kRing = 4096;  // rows of x kept in LDS
constexpr int kPre = 8;      // entries per row prefetched into registers

template <typename T>
struct ChainArgs {
    int n;
    const int* rowptr;
    const int* col;
    const T* val;
    const T* b;
    T* x;
};

template <typename T>
__device__ __forceinline__ T bcast(T v, int src) {
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)u, src);
        const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), src);
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), src));
    }
}

// Row data of one window, as prefetched into registers.
template <typename T>
struct RowPre {
    int start, end;   // CSR range of the row (start = diagonal)
    T diag, bv;
    int c[kPre];
    T v[kPre];
    T old[kPre];      // x of entries older than the ring (else unused)
};

template <typename T>
__device__ __forceinline__ void prefetch_row(const ChainArgs<T>& a, int hi, int lane, RowPre<T>& R) {
    const int r = hi - 1 - lane;
    R.start = R.end = 0;
    R.diag = T(1);
    R.bv = T(0);
    if (r >= 0) {
        R.start = a.rowptr[r];
        R.end = a.rowptr[r + 1];
        R.diag = a.val[R.start];
        R.bv = a.b[r];
    }
#pragma unroll
    for (int q = 0; q < kPre; ++q) {
        const int p = R.start + 1 + q;
        R.c[q] = -1;
        R.v[q] = T(0);
        R.old[q] = T(0);
        if (r >= 0 && p < R.end) {
            R.c[q] = a.col[p];
            R.v[q] = a.val[p];
        }
    }
}

// x of entries beyond the ring: published at least kRing rows (64 windows) earlier
template <typename T>
__device__ __forceinline__ void prefetch_old(const ChainArgs<T>& a, int hi, RowPre<T>& R) {
#pragma unroll
    for (int q = 0; q < kPre; ++q)
        if (R.c[q] >= hi + kRing) R.old[q] = from_bits<T>(poll_bits(a.x + R.c[q]));
}

template <typename T>
__global__ __launch_bounds__(64) void k_chain(ChainArgs<T> a) {
    __shared__ T tile[64][64];
    __shared__ T ring[kRing];
    const int lane = threadIdx.x;
    // three windows in flight: cur (being solved), nxt (row data here, old x requested this
    // iteration), nx2 (row data requested this iteration); every load is issued before a
    // 64-step loop and consumed after it.
    RowPre<T> cur, nxt, nx2;
    int hi = a.n;
    prefetch_row(a, hi, lane, cur);
    prefetch_old(a, hi, cur);
    if (hi - 64 > 0) prefetch_row(a, hi - 64, lane, nxt);
    int windows = 0;
    int prow = -1;        // row whose x is published one window late (see below)
    T pending = T(0);
    while (hi > 0) {
        const int r = hi - 1 - lane;
        const int hn = hi - 64, hn2 = hi - 128;
        // ---- stage this window: tile column and the sums over outside columns ----
#pragma unroll 8
        for (int s = 0; s < 64; ++s) tile[s][lane] = T(0);
        T acc = T(0);
#pragma unroll
        for (int q = 0; q < kPre; ++q) {
            const int c = cur.c[q];
            if (c < 0) continue;
            if (c < hi) tile[hi - 1 - c][lane] = cur.v[q];
            else if (c < hi + kRing) acc += cur.v[q] * ring[c & (kRing - 1)];
            else acc += cur.v[q] * cur.old[q];
        }
        for (int p = cur.start + 1 + kPre; p < cur.end; ++p) {  // long rows: the rest directly
            const int c = a.col[p];
            const T v = a.val[p];
            if (c < hi) tile[hi - 1 - c][lane] = v;
            else if (c < hi + kRing) acc += v * ring[c & (kRing - 1)];
            else acc += v * from_bits<T>(poll_bits(a.x + c));
        }
        const T inv = T(1) / cur.diag;
        const T binv = cur.bv * inv;
        __builtin_amdgcn_wave_barrier();
        // Old x of the next window first (its column ids arrived a window ago), then the previous
        // window's x to HBM, then the row data two windows ahead.  gfx950 retires vector loads
        // and stores through one in-order counter, so this order lets every later wait find its
        // operations already retired.  HBM readers of x are >= 63 windows later; recent x is
        // read from the LDS ring.
        if (hn > 0) prefetch_old(a, hn, nxt);
        if (prow >= 0) publish(a.x + prow, pending);
        if (hn2 > 0) prefetch_row(a, hn2, lane, nx2);
        // ---- 64 dependent steps ----
        T mine = T(0);
#pragma unroll 16
        for (int s = 0; s < 64; ++s) {
            const T xs = fma(-acc, inv, binv);  // (b - acc) / diag with one rounding less
            const T xb = bcast(xs, s);
            if (lane == s) mine = xs;
            acc += tile[s][lane] * xb;
        }
        if (r >= 0) ring[r & (kRing - 1)] = mine;
        prow = r;
        pending = mine;
        __builtin_amdgcn_wave_barrier();
        cur = nxt;
        nxt = nx2;
        // every 32 windows make sure the published x (read back from HBM 63 windows later) landed
        if ((++windows & 31) == 0) wait_vmem();
        hi = hn;
    }
    if (prow >= 0) publish(a.x + prow, pending);
}

}  // namespace

int chain_solve(sptrsv_plan_s* P, int precision, const double* d_b, double* d_x, int m) {
    if (m != 1) return fail(SPTRSV_ERR_UNSUPPORTED, "chain variant solves one right-hand side");
    if (precision != SPTRSV_FP64) return fail(SPTRSV_ERR_UNSUPPORTED, "chain variant is FP64 through sptrsv_solve");
    ChainArgs<double> a{P->n, P->rowptr, P->col, P->val, d_b, d_x};
    hipEvent_t stop;
    SPT_CHECK(begin_timed(P, &stop));
    k_chain<double><<<1, 64, 0, P->stream>>>(a);
    end_timed(P, stop);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

int chain_solve32(sptrsv_plan_s* P, const float* d_b, float* d_x) {
    ChainArgs<float> a{P->n, P->rowptr, P->col, P->val32, d_b, d_x};
    k_chain<float><<<1, 64, 0, P->stream>>>(a);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

}  // namespace sptrsv
