// common.h -- shared host/device helpers of the gfx950 solver library.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdint.h>

#include <string>

#include "sptrsv_capi.h"

// ---- host-side error plumbing ------------------------------------------------------------
namespace sptrsv {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define SPT_HIP(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::sptrsv::fail(SPTRSV_ERR_HIP, std::string(#expr " -> ") +          \
                                                      hipGetErrorString(e_));          \
    } while (0)

#define SPT_CHECK(expr)            \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != SPTRSV_OK)      \
            return rc_;            \
    } while (0)

int ensure_device();
uint64_t wallclock_ticks_per_ms();
int device_compute_units();

// ---- device-side helpers -------------------------------------------------------------------
constexpr int kWave = 64;

// Values of x that are still unknown hold a signalling-NaN bit pattern.  IEEE arithmetic
// never produces a signalling NaN (invalid operations give the default quiet NaN), so a
// computed x can never be mistaken for "not ready", and the value itself is the ready flag.
constexpr uint64_t kSentinel64 = 0x7FF4DEADBEEF5AA5ull;
constexpr uint32_t kSentinel32 = 0x7FA5DEADu;

template <typename T> struct Bits;
template <> struct Bits<double> {
    using U = uint64_t;
    static constexpr U sentinel = kSentinel64;
};
template <> struct Bits<float> {
    using U = uint32_t;
    static constexpr U sentinel = kSentinel32;
};

// Relaxed agent-scope load/store: `global_load/store ... sc1` on gfx950 -- coherent across
// the XCDs' L2s without the L2 write-back/invalidate that acquire/release would add.
template <typename T>
__device__ __forceinline__ typename Bits<T>::U poll_bits(const T* p) {
    using U = typename Bits<T>::U;
    return __hip_atomic_load(reinterpret_cast<const U*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void publish(T* p, T v) {
    using U = typename Bits<T>::U;
    __hip_atomic_store(reinterpret_cast<U*>(p), __builtin_bit_cast(U, v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T from_bits(typename Bits<T>::U u) {
    return __builtin_bit_cast(T, u);
}

// Division by a value whose reciprocal is reused, bitwise equal to the compiled IEEE division
// n / d.  The compiler's FP64 division is: dd = div_scale(d, d, n), r = rcp(dd) refined by two
// Newton steps, nn = div_scale(n, d, n), q = nn * r, rem = fma(-dd, q, nn),
// div_fixup(div_fmas(rem, r, q), d, n).  When neither div_scale rescales (dd = d, nn = n, the
// fmas flag clear) that is div_fixup(fma(fma(-d, n*r, n), r, n*r), d, n) with r depending on d
// alone: div_recip(d) computes r once, div_by(n, d, r) the rest.  div_scale rescales only near
// the ends of the exponent range (zero, denormal or huge operands, a quotient exponent beyond
// about +-768), so div_by takes this path only for biased exponents of n and d in [200, 1800]
// whose difference is at most 600 -- far inside -- and the ordinary division otherwise.  In
// that range div_fixup returns its operand (operands and quotient are finite, normal and
// nonzero, and q + rem * r carries the quotient's sign), so it is left out.
__device__ __forceinline__ double div_recip(double d) {
    const double r0 = __builtin_amdgcn_rcp(d);
    const double r1 = __builtin_fma(r0, __builtin_fma(-d, r0, 1.0), r0);
    return __builtin_fma(r1, __builtin_fma(-d, r1, 1.0), r1);
}
// The quotient without the fallback; `plain` tells whether it is the IEEE one.
__device__ __forceinline__ double div_by_plain(double n, double d, double r, bool& plain) {
#pragma clang fp contract(off)
    const int en = (int)((__builtin_bit_cast(uint64_t, n) >> 52) & 0x7FF);
    const int ed = (int)((__builtin_bit_cast(uint64_t, d) >> 52) & 0x7FF);
    plain = en >= 200 && en <= 1800 && ed >= 200 && ed <= 1800 && en - ed <= 600 && ed - en <= 600;
    const double q = n * r;
    return __builtin_fma(__builtin_fma(-d, q, n), r, q);
}
__device__ __forceinline__ double div_by(double n, double d, double r) {
    bool plain;
    double out = div_by_plain(n, d, r, plain);
    if (!plain) {
        double dd = d;
        asm volatile("" : "+v"(dd));  // a real branch: the compiler would otherwise divide in every lane
        out = n / dd;
    }
    return out;
}

// Wave-level work ticket: lane 0 takes the next ticket, every lane gets it.  Tickets are
// handed out in launch order of the waves, so a wave only ever waits on work whose ticket
// is smaller -- held by a wave that is already resident -- which guarantees progress.
__device__ __forceinline__ int wave_ticket(int* counter) {
    int t = 0;
    if ((threadIdx.x & (kWave - 1)) == 0) t = atomicAdd(counter, 1);
    return __shfl(t, 0, kWave);
}

// s_waitcnt vmcnt(0): on gfx950 one in-order counter covers loads AND stores, so a load whose
// first use sits inside a polling loop would make every later use wait for the wave's
// outstanding x stores.  Loads needed inside such loops are completed before the loop.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Workgroup-level ticket: one atomic per workgroup instead of one per wave (a single shared
// counter serialises its atomics, so fewer, larger tickets matter at millions of rows).
__device__ __forceinline__ int block_ticket(int* counter) {
    __shared__ int s_ticket;
    if (threadIdx.x == 0) s_ticket = atomicAdd(counter, 1);
    __syncthreads();
    return s_ticket;
}

// Watchdog: a lane that has waited longer than `limit` wall-clock ticks sets the status
// word and gives up; every waiting lane also gives up once the status word is non-zero,
// so a malformed input ends the kernel instead of hanging the device.
struct Watchdog {
    uint64_t start;
    uint64_t limit;
    int* status;
    __device__ __forceinline__ void arm(int* s, uint64_t lim) {
        status = s;
        limit = lim;
        start = wall_clock64();
    }
    __device__ __forceinline__ bool expired() {
        if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return true;
        if (wall_clock64() - start > limit) {
            __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return true;
        }
        return false;
    }
};

// r = b - U x (FP64 products and sums, rounded once to R): G lanes per (row, right-hand side),
// striding over the row's entries (coalesced reads of a row), then a shuffle reduction.
// Fixed reduction order: the same x gives the same r on every run.
template <typename IdxT, typename R, int G>
__global__ __launch_bounds__(256) void k_residual(int n, int m, const IdxT* __restrict__ rowptr,
                                                  const int* __restrict__ col, const double* __restrict__ val,
                                                  const double* __restrict__ x, const double* __restrict__ b,
                                                  R* __restrict__ r) {
    const int64_t total = (int64_t)n * m;
    const int g = threadIdx.x & (G - 1);
    const int64_t step = (int64_t)gridDim.x * (blockDim.x / G);
    // every lane of a G-group has the same q: the shuffles below never mix groups
    for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; q < total; q += step) {
        const int row = (int)(q / m), k = (int)(q % m);
        double acc = 0.0;
        for (IdxT e = rowptr[row] + g; e < rowptr[row + 1]; e += G) acc = fma(val[e], x[(size_t)col[e] * m + k], acc);
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
        if (g == 0) r[q] = (R)(b[q] - acc);
    }
}

// lanes per (row, right-hand side) for the residual: about half the mean row length, fewer
// as the right-hand sides already spread the work
inline int residual_lanes(int64_t n, int64_t nnz, int m) {
    const int64_t avg = n > 0 ? nnz / n : 1;
    int g = 1;
    while (g < 64 && 2 * g <= avg / (2 * m)) g *= 2;
    return g;
}

template <typename IdxT, typename R>
void launch_residual(int n, int64_t nnz, int m, const IdxT* rowptr, const int* col, const double* val,
                     const double* x, const double* b, R* r, hipStream_t s) {
    const int G = residual_lanes(n, nnz, m);
    const int64_t threads = (int64_t)n * m * G;
    const int grid = (int)std::min<int64_t>(std::max<int64_t>((threads + 255) / 256, 1), 1 << 20);
    switch (G) {
        case 1: k_residual<IdxT, R, 1><<<grid, 256, 0, s>>>(n, m, rowptr, col, val, x, b, r); break;
        case 2: k_residual<IdxT, R, 2><<<grid, 256, 0, s>>>(n, m, rowptr, col, val, x, b, r); break;
        case 4: k_residual<IdxT, R, 4><<<grid, 256, 0, s>>>(n, m, rowptr, col, val, x, b, r); break;
        case 8: k_residual<IdxT, R, 8><<<grid, 256, 0, s>>>(n, m, rowptr, col, val, x, b, r); break;
        case 16: k_residual<IdxT, R, 16><<<grid, 256, 0, s>>>(n, m, rowptr, col, val, x, b, r); break;
        case 32: k_residual<IdxT, R, 32><<<grid, 256, 0, s>>>(n, m, rowptr, col, val, x, b, r); break;
        default: k_residual<IdxT, R, 64><<<grid, 256, 0, s>>>(n, m, rowptr, col, val, x, b, r); break;
    }
}

}  // namespace sptrsv
