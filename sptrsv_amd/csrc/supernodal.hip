// supernodal.hip -- supernodal U X = B on FP64 MFMA (gfx950).
//
// Reference semantics: src/core.cpp:53-201 (supernodal_upper, ccs2ccs_pad,
// supernodal_blas_upper): rows of a supernode share a dense triangle plus one common set of
// "rectangular" columns; the solve goes supernode by supernode from the last one, first
// x_s -= R_s x(rect) (the reference's cblas_dgemm), then the triangular solve with the
// diagonal block (its cblas_dtrsm).
//
// gfx950 design: the diagonal block of every supernode is inverted once, on the device, when
// the layout is built, so a supernode's solve is two dense products on MFMA:
//     W   = B_s - R_s * X(rect, :)       (w x m, K = r)      v_mfma_f64_16x16x4f64
//     X_s = Tinv_s * W                   (w x m, K = w)      v_mfma_f64_16x16x4f64
// v_mfma_f64_16x16x4f64 lane layout (measured on gfx950): A(i,k) in lane i + 16k, B(k,j) in
// lane j + 16k, result element v of lane l is D(l/16 + 4v, l%16).
// FP64: one workgroup per (supernode, panel of up to 64 columns of X), handed out by tickets
// from the last supernode down; it waits (acquire) for the same panel of the supernodes owning
// its rectangular columns -- earlier in ticket order, so already running -- and publishes its
// own (supernode, panel) flag (release) after writing X_s.  W stays in the accumulators.
// Reduced precision (k_sn_solve_low): the same decomposition with BF16 / FP32 MFMA.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "common.h"
#include "plan.h"

namespace sptrsv {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kMaxW = 128;  // widest supernode the kernel holds; wider ones are split

// column j of Tinv by back-substitution on T e_j (one thread per column)
__global__ void k_sn_invert(const int* __restrict__ start, const int64_t* __restrict__ offt,
                            const double* __restrict__ T, double* __restrict__ Ti) {
    const int k = blockIdx.x;
    const int s0 = start[k], w = start[k + 1] - s0;
    const int64_t base = offt[k];
    for (int j = threadIdx.x; j < w; j += blockDim.x) {
        for (int i = j + 1; i < w; ++i) Ti[base + (int64_t)i * w + j] = 0.0;
        Ti[base + (int64_t)j * w + j] = 1.0 / T[base + (int64_t)j * w + j];
        for (int i = j - 1; i >= 0; --i) {
            double acc = 0.0;
            for (int e = i + 1; e <= j; ++e) acc = fma(T[base + (int64_t)i * w + e], Ti[base + (int64_t)e * w + j], acc);
            Ti[base + (int64_t)i * w + j] = -acc / T[base + (int64_t)i * w + i];
        }
    }
}

struct SnArgs {
    int sn;
    int m;
    int npanel;        // column panels per supernode: ceil(ceil(m / 16) / NW)
    const int* start;
    const int* rcount;
    const int64_t* offt;
    const int64_t* offr;
    const int* offc;
    const int* cols;
    const double* tinv;
    const double* rect;
    const double* tri;  // diagonal blocks T_s (residual mode)
    const int* dep_ptr;
    const int* dep;
    int* flag;         // sn x npanel: epoch of the last solve that published (supernode, panel)
    int epoch;
    const double* b;
    double* x;
    float* r32;        // residual mode: r = b - U x, rounded to FP32
    int* ctrl;
    uint64_t limit;
};

constexpr int kSnCChunk = 1024;    // rectangular column indices staged in LDS at a time

// Workgroup barrier that waits for this wave's LDS operations only (__syncthreads would also
// drain the global loads issued ahead).
__device__ __forceinline__ void lds_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt not waited
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}
constexpr int kSnLdt = kMaxW + 1;  // LDS row stride (doubles) of a 16-row block of Tinv

// One workgroup of NW waves per (supernode k, panel of NW 16-column blocks of X), tickets from
// the last supernode down.  Wave 0 waits for the same panel of the supernodes owning k's
// rectangular columns, then every wave J of the panel computes, for its 16 columns,
//   W   = B_s - R_s * X(rect)    row tiles I of 16 in acc[I]; K in chunks of KC: the chunk of
//                                R goes through a double-buffered LDS tile shared by the waves,
//                                the gathered X rows straight into B-operand registers, and
//                                the next chunk is loaded while the current one is multiplied
//   X_s = Tinv_s * W             in place, row block by row block (X[I] needs W rows >= 16 I);
//                                Tinv row blocks staged through the same LDS.
// The f64 16x16x4 result layout (lane (q, c) holds rows q + 4v of column c) is exactly its
// B-operand layout for the K-step of rows 16 I2 + 4 v (lane (q, c) holds row 4 v + q), so
// Tinv * W reads W straight from the accumulators: element v of acc[I2] in the same lane.
// RESID: the FP64 residual of the refinement, r_s = B_s - T_s X_s - R_s X(rect) (no waits; the
// triangle is one more K segment with columns s0 .. s0 + w - 1), stored as FP32.
template <int NW, bool RESID>
__global__ __launch_bounds__(64 * NW, NW == 1 ? 1 : NW == 8 ? 1 : 8 / NW) void k_sn_solve(SnArgs a) {
    constexpr int NT = 64 * NW;
    constexpr int kSnKC = NW >= 4 ? 32 : 16;  // K chunk (register budget: 2 waves per SIMD)
    constexpr int kSnLdr = kSnKC + 1;         // LDS row stride (doubles) of a K chunk of R
    constexpr int kSnBuf = 2 * kMaxW * kSnLdr;
    static_assert(kSnBuf >= 2 * 16 * kSnLdt, "Tinv blocks must fit the R double buffer");
    constexpr int kPerR = (kMaxW * kSnKC + NT - 1) / NT;  // R chunk elements loaded per thread
    constexpr int kPerT = (16 * kMaxW + NT - 1) / NT;  // Tinv block elements loaded per thread
    __shared__ double sbuf[kSnBuf];
    __shared__ int sC[kSnCChunk];
    __shared__ int s_ticket;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (!RESID) {
        if (tid == 0) s_ticket = atomicAdd(&a.ctrl[0], 1);
        __syncthreads();
    }
    const int t = RESID ? (int)blockIdx.x : s_ticket;
    const int k = a.sn - 1 - t / a.npanel;
    const int panel = t % a.npanel;
    if (k < 0) return;
    const int s0 = a.start[k], w = a.start[k + 1] - s0, r = a.rcount[k];
    const int c16 = lane & 15, quad = lane >> 4;
    const size_t m = (size_t)a.m;
    const int jj = 16 * (panel * NW + wv) + c16;
    const bool jv = jj < a.m;
    const int nI = (w + 15) >> 4;
    v4d acc[8];
#pragma unroll
    for (int I = 0; I < 8; ++I) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 16 * I + quad + 4 * v;
            acc[I][v] = (I < nI && i < w && jv) ? a.b[(size_t)(s0 + i) * m + jj] : 0.0;
        }
    }

    // ---- W = B_s - A X(cols) over K segments (solve: R_s; residual: T_s, then R_s) --------
    // One wave (NW == 1) has registers to spare at its occupancy: it keeps TWO chunks of R and
    // X in flight (sets A and B) and waits for a chunk one iteration after issuing it.
    constexpr bool kDeep = NW == 1;
    double rreg[kPerR], rreg2[kDeep ? kPerR : 1];
    constexpr int KS = kSnKC / 4;  // MFMA K-steps per chunk
    double bv[KS], bn[KS];
    // The solve's first chunks of R (one, or two with one wave) do not depend on X: they load
    // while wave 0 waits for the flags.
    bool preloaded = !RESID;
    if (!RESID) {
        const double* Rk = a.rect + a.offr[k];
        auto pre_r = [&](double* rr, int kg) {
#pragma unroll
            for (int q = 0; q < kPerR; ++q) {
                const int e = tid + q * NT, row = e / kSnKC, kc = kg + e % kSnKC;
                rr[q] = (row < w && kc < r) ? Rk[(size_t)row * r + kc] : 0.0;
            }
        };
        pre_r(rreg, 0);
        if (kDeep && min(kSnCChunk, r) > kSnKC) pre_r(rreg2, kSnKC);
    }
    if (!RESID && wv == 0) {
        Watchdog wd;
        wd.arm(&a.ctrl[1], a.limit);
        const int d1 = a.dep_ptr[k + 1];
        int idle = 0;
        bool give_up = false;
        for (int base = a.dep_ptr[k]; base < d1 && !give_up; base += kWave) {
            const int q = base + lane;
            bool ready = q >= d1;
            const int* f = ready ? nullptr : &a.flag[(size_t)a.dep[q] * a.npanel + panel];
            while (!__all(ready)) {
                if (!ready) ready = __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == a.epoch;
                if (__all(ready)) break;
                __builtin_amdgcn_s_sleep(2);
                if ((++idle & 63) == 0 && __any(wd.expired())) {
                    give_up = true;
                    break;
                }
            }
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");

    // one K segment: A is w x K row-major (leading dimension lda); column kk of the segment is
    // X row sC[kk] (cols != nullptr, staged in LDS by blocks) or row s0 + kk (cols == nullptr)
    auto segment = [&](const double* A, int lda, int K, const int* cols) {
        auto load_r_to = [&](double* rr, int kg) {  // A[:, kg .. kg+KC) into registers (zeros outside)
#pragma unroll
            for (int q = 0; q < kPerR; ++q) {
                const int e = tid + q * NT, row = e / kSnKC, kc = kg + e % kSnKC;
                rr[q] = (row < w && kc < K) ? A[(size_t)row * lda + kc] : 0.0;
            }
        };
        // A goes to LDS negated (the store waits for its loads anyway), so acc = B + (-A) X needs
        // no negation of the X loads, which would make each of them wait as it lands
        auto store_r_from = [&](const double* rr, double* dst) {
#pragma unroll
            for (int q = 0; q < kPerR; ++q) {
                const int e = tid + q * NT;
                if (e / kSnKC < kMaxW) dst[(e / kSnKC) * kSnLdr + e % kSnKC] = -rr[q];
            }
        };
        auto load_r = [&](int kg) { load_r_to(rreg, kg); };
        auto store_r = [&](double* dst) { store_r_from(rreg, dst); };
        for (int cb = 0; cb < K; cb += kSnCChunk) {
            const int cn = min(kSnCChunk, K - cb);
            const int nch = (cn + kSnKC - 1) / kSnKC;
            auto load_x = [&](double* out, int kl) {  // X rows of local chunk kl
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    const int kk = kl + 4 * s + quad;
                    const int row = cols ? sC[kk < cn ? kk : 0] : s0 + cb + kk;
                    out[s] = (kk < cn && jv) ? a.x[(size_t)row * m + jj] : 0.0;
                }
            };
            __syncthreads();
            if (cols)
                for (int q = tid; q < cn; q += NT) sC[q] = cols[cb + q];
            __syncthreads();
            if constexpr (kDeep) {
                // chunk c: R in LDS buffer c & 1, X in set (c & 1 ? bn : bv) ... rotated below:
                // at the top of iteration c, chunk c is in LDS with its X in xa, chunk c + 1 is
                // in flight in rb / xb; the iteration multiplies c, stores c + 1 to LDS, and
                // issues c + 2 into the set c used.
                double xa[KS], xb[KS];
                const bool have = preloaded && cb == 0;  // chunks 0 and 1 loaded before the wait
                load_x(xa, 0);
                if (!have) load_r_to(rreg, cb);
                if (nch > 1) {
                    load_x(xb, kSnKC);
                    if (!have) load_r_to(rreg2, cb + kSnKC);
                }
                store_r_from(rreg, sbuf);
                lds_barrier();
                auto step = [&](int c, double (&xc)[KS], double (&xn)[KS], double* rc, const double* rn) {
                    const double* cur = sbuf + (c & 1) * (kMaxW * kSnLdr);
                    double* nxt = sbuf + ((c + 1) & 1) * (kMaxW * kSnLdr);
#pragma unroll
                    for (int s = 0; s < KS; ++s) {
#pragma unroll
                        for (int I = 0; I < 8; ++I) {
                            if (I < nI) {
                                const double av = cur[(16 * I + c16) * kSnLdr + 4 * s + quad];
                                acc[I] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, xc[s], acc[I], 0, 0, 0);
                            }
                        }
                    }
                    if (c + 1 < nch) store_r_from(rn, nxt);
                    if (c + 2 < nch) {
                        load_x(xc, kSnKC * (c + 2));
                        load_r_to(rc, cb + kSnKC * (c + 2));
                    }
                    lds_barrier();
                    (void)xn;
                };
                for (int c = 0; c < nch; c += 2) {
                    step(c, xa, xb, rreg, rreg2);
                    if (c + 1 < nch) step(c + 1, xb, xa, rreg2, rreg);
                }
                continue;
            }
            load_x(bn, 0);
            if (!(preloaded && cb == 0)) load_r(cb);  // chunk 0 loaded before the wait
            store_r(sbuf);
            __syncthreads();
            for (int c = 0; c < nch; ++c) {
                const double* cur = sbuf + (c & 1) * (kMaxW * kSnLdr);
                double* nxt = sbuf + ((c + 1) & 1) * (kMaxW * kSnLdr);
#pragma unroll
                for (int s = 0; s < KS; ++s) bv[s] = bn[s];
                if (c + 1 < nch) {
                    load_x(bn, kSnKC * (c + 1));
                    load_r(cb + kSnKC * (c + 1));
                }
#pragma unroll
                for (int s = 0; s < KS; ++s) {
#pragma unroll
                    for (int I = 0; I < 8; ++I) {
                        if (I < nI) {
                            const double av = cur[(16 * I + c16) * kSnLdr + 4 * s + quad];
                            acc[I] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[s], acc[I], 0, 0, 0);
                        }
                    }
                }
                if (c + 1 < nch) store_r(nxt);
                __syncthreads();
            }
        }
    };
    if (RESID) segment(a.tri + a.offt[k], w, w, nullptr);
    segment(a.rect + a.offr[k], r, r, a.cols + a.offc[k]);
    if (RESID) {
#pragma unroll
        for (int I = 0; I < 8; ++I) {
            if (I < nI) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = 16 * I + quad + 4 * v;
                    if (i < w && jv) a.r32[(size_t)(s0 + i) * m + jj] = (float)acc[I][v];
                }
            }
        }
        return;
    }

    // ---- X_s = Tinv_s W, row block by row block ------------------------------------------
    const double* Ti = a.tinv + a.offt[k];
    double treg[kPerT];
    auto load_t = [&](int I) {  // Tinv rows 16 I .. 16 I + 15, all columns
#pragma unroll
        for (int q = 0; q < kPerT; ++q) {
            const int e = tid + q * NT, row = 16 * I + e / kMaxW, col = e % kMaxW;
            treg[q] = (e < 16 * kMaxW && row < w && col < w) ? Ti[(size_t)row * w + col] : 0.0;
        }
    };
    auto store_t = [&](double* dst) {
#pragma unroll
        for (int q = 0; q < kPerT; ++q) {
            const int e = tid + q * NT;
            if (e < 16 * kMaxW) dst[(e / kMaxW) * kSnLdt + e % kMaxW] = treg[q];
        }
    };
    __syncthreads();
    load_t(0);
    store_t(sbuf);
    __syncthreads();
#pragma unroll
    for (int I = 0; I < 8; ++I) {
        if (I < nI) {
            const double* cur = sbuf + (I & 1) * (16 * kSnLdt);
            double* nxt = sbuf + ((I + 1) & 1) * (16 * kSnLdt);
            if (I + 1 < nI) load_t(I + 1);
            v4d xt = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int I2 = I; I2 < 8; ++I2) {
                if (I2 < nI) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int kk = 16 * I2 + 4 * v;
                        if (kk < w) {
                            const double av = cur[c16 * kSnLdt + kk + quad];
                            xt = __builtin_amdgcn_mfma_f64_16x16x4f64(av, acc[I2][v], xt, 0, 0, 0);
                        }
                    }
                }
            }
            acc[I] = xt;
            if (I + 1 < nI) store_t(nxt);
            __syncthreads();
        }
    }
#pragma unroll
    for (int I = 0; I < 8; ++I) {
        if (I < nI) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = 16 * I + quad + 4 * v;
                if (i < w && jv) a.x[(size_t)(s0 + i) * m + jj] = acc[I][v];
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (tid == 0)
        __hip_atomic_store(&a.flag[(size_t)k * a.npanel + panel], a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_reset_ticket_sn(int* ctrl) {
    if (threadIdx.x == 0) ctrl[0] = 0;
}

// One right-hand side.  The MFMA kernel's 16-column tiles would leave 15 of 16 columns idle, and
// it loads R and Tinv after its wait, one memory latency per chunk; on a chain of supernodes
// (a banded factor in blocks of 128 rows) those latencies add up along the whole chain.  Here a
// workgroup of 512 threads per supernode gives each row four lanes (lane q of row i = t / 4
// takes columns q, q + 4, q + 8, ...) and loads everything that does not depend on X before it
// waits: b_s, the 4 * kSgPre or fewer columns of R_s with their column indices, and row i of
// Tinv_s (columns j >= i).  x is self-publishing (filled with the sentinel before the launch,
// each value stored once with a relaxed device-scope store), so there is no flag: each lane
// polls x at its columns and, as each value arrives, turns its R entry into the product; once
// all have arrived
//     W_i   = b_i - sum of the products      in a fixed order, four partial sums per row
//                                            reduced across the quad,
//     X_s,i = Tinv_s(i, :) W                 W passed through LDS, reduced the same way,
// and X_s is published.  Every rectangular column is loaded before the wait, so the kernel
// needs r <= 4 * kSgPre for every supernode (a band of width <= 128 in blocks of 128 rows);
// the solve takes the MFMA kernel otherwise, and when x overlaps b (the fill would destroy b).
// The sums run in a fixed order: the same inputs give the same X on every run.
constexpr int kSgPre = 32;  // R columns per lane loaded before the wait
constexpr int kSgT = kMaxW / 4;  // Tinv columns per lane

// x := sentinel, and the ticket counter reset
__global__ void k_fill_sentinel_sn(double* x, int64_t n, int* ctrl) {
    const uint64_t s = Bits<double>::sentinel;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctrl[0] = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        reinterpret_cast<uint64_t*>(x)[i] = s;
}

__global__ __launch_bounds__(512, 1) void k_sn_gemv(SnArgs a) {
    using U = uint64_t;
    constexpr U S = Bits<double>::sentinel;
    __shared__ double sW[kMaxW];
    __shared__ int s_ticket;
    const int tid = threadIdx.x;
    if (tid == 0) s_ticket = atomicAdd(&a.ctrl[0], 1);
    __syncthreads();
    const int k = a.sn - 1 - s_ticket;
    if (k < 0) return;
    const int s0 = a.start[k], w = a.start[k + 1] - s0, r = a.rcount[k];
    const int i = tid >> 2, q = tid & 3;
    const bool row = i < w;
    const double* R = a.rect + a.offr[k] + (size_t)(row ? i : 0) * r;
    const int* C = a.cols + a.offc[k];
    const double* Ti = a.tinv + a.offt[k] + (size_t)(row ? i : 0) * w;

    // ---- independent of X: b, R, column indices, Tinv row ---------------------------------
    const double bi = (row && q == 0) ? a.b[s0 + i] : 0.0;
    double rv[kSgPre], tv[kSgT];
    int cv[kSgPre];
#pragma unroll
    for (int u = 0; u < kSgPre; ++u) {
        const int c = q + 4 * u;
        const bool ok = row && c < r;
        rv[u] = ok ? R[c] : 0.0;
        cv[u] = ok ? C[c] : -1;
    }
#pragma unroll
    for (int u = 0; u < kSgT; ++u) {
        const int j = q + 4 * u;
        tv[u] = (row && j >= i && j < w) ? Ti[j] : 0.0;
    }

    // ---- x at the rectangular columns, as it is published: rv[u] becomes the product -------
    uint32_t left = 0;  // bit u: column u still to arrive
#pragma unroll
    for (int u = 0; u < kSgPre; ++u) left |= cv[u] >= 0 ? 1u << u : 0u;
    {
        Watchdog wd;
        wd.arm(&a.ctrl[1], a.limit);
        int idle = 0;
        while (__any(left != 0)) {
#pragma unroll
            for (int u = 0; u < kSgPre; ++u) {
                if (left & (1u << u)) {
                    const U v = poll_bits(a.x + cv[u]);
                    if (v != S) {
                        rv[u] *= from_bits<double>(v);
                        left &= ~(1u << u);
                    }
                }
            }
            if (!__any(left != 0)) break;
            __builtin_amdgcn_s_sleep(1);
            if ((++idle & 63) == 0 && __any(left != 0 && wd.expired())) break;
        }
    }

    // ---- W = b - R X(rect) -----------------------------------------------------------------
    double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
    for (int u = 0; u < kSgPre; u += 2) {
        acc0 += rv[u];
        acc1 += rv[u + 1];
    }
    double s = acc0 + acc1;
    s += __shfl_xor(s, 1, 4);
    s += __shfl_xor(s, 2, 4);
    if (q == 0) sW[i] = row ? bi - s : 0.0;  // zeros past w: Tinv's padding multiplies them
    __syncthreads();

    // ---- X_s = Tinv_s W ----------------------------------------------------------------------
    double t0 = 0.0, t1 = 0.0;
#pragma unroll
    for (int u = 0; u < kSgT; u += 2) {
        t0 = fma(tv[u], sW[q + 4 * u], t0);
        t1 = fma(tv[u + 1], sW[q + 4 * u + 4], t1);
    }
    double xs = t0 + t1;
    xs += __shfl_xor(xs, 1, 4);
    xs += __shfl_xor(xs, 2, 4);
    if (row && q == 0) publish(a.x + s0 + i, xs);
}

template <typename X>
int upload(X** dst, const std::vector<X>& src) {
    SPT_HIP(hipMalloc((void**)dst, std::max<size_t>(src.size(), 1) * sizeof(X)));
    if (!src.empty()) SPT_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(X), hipMemcpyHostToDevice));
    return SPTRSV_OK;
}


// ---- reduced precision (inner solves of the FP64 refinement) ------------------------------

typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x16bf16_1k (measured): A(i,k) is element k%4 of lane i + 16*(k/4), B(k,j) is
// element k%4 of lane j + 16*(k/4); result element v of lane l is D(4*(l/16) + v, l%16).
// v_mfma_f32_16x16x4f32: A(i,k) lane i + 16k, B(k,j) lane j + 16k, D like the bf16 form.

__device__ __forceinline__ short to_bf16(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (values are finite)
    return (short)(u >> 16);
}

// one block per supernode: scale = 2^ceil(log2 max|R|), R16 = bf16(R / scale), rows padded to
// a multiple of 16 columns (zeros) so every MFMA k-step reads whole aligned groups
__global__ void k_sn_pack16(const int* __restrict__ start, const int* __restrict__ rcount,
                            const int64_t* __restrict__ offr, const int64_t* __restrict__ offr16,
                            const double* __restrict__ R, uint16_t* __restrict__ R16, float* __restrict__ rscale) {
    __shared__ double red[256];
    const int k = blockIdx.x;
    const int w = start[k + 1] - start[k], r = rcount[k], rp = (r + 15) & ~15;
    const double* src = R + offr[k];
    double mx = 0.0;
    for (int q = threadIdx.x; q < w * r; q += blockDim.x) mx = fmax(mx, fabs(src[q]));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const double scale = red[0] > 0.0 ? exp2(ceil(log2(red[0]))) : 1.0;
    if (threadIdx.x == 0) rscale[k] = (float)scale;
    uint16_t* dst = R16 + offr16[k];
    for (int q = threadIdx.x; q < w * rp; q += blockDim.x) {
        const int d = q / rp, c = q % rp;
        dst[q] = c < r ? (uint16_t)to_bf16((float)(src[(size_t)d * r + c] / scale)) : (uint16_t)0;
    }
}

__global__ void k_sn_to_float(const double* __restrict__ in, float* __restrict__ out, int64_t count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < count; i += (int64_t)gridDim.x * blockDim.x) out[i] = (float)in[i];
}

struct SnLowArgs {
    int sn;
    int m;
    int npanel;
    const int* start;
    const int* rcount;
    const int64_t* offt;
    const int64_t* offr16;
    const int* offc;
    const int* cols;
    const float* tinv;
    const uint16_t* rect16;
    const float* rscale;
    const int* dep_ptr;
    const int* dep;
    int* flag;
    int epoch;
    const float* b;
    float* x;
    int* ctrl;
    uint64_t limit;
};

// Reduced-precision solve, same decomposition and waits as k_sn_solve:
//   W   = B_s - scale_s * (R16_s * bf16(X(rect)))   v_mfma_f32_16x16x16bf16_1k, FP32 accumulate;
//                                                   K chunks of 32 of R16 double-buffered in LDS
//   X_s = Tinv_s * W                                v_mfma_f32_16x16x4f32; W of each wave through
//                                                   LDS (the f32 result layout is not its B layout)
template <int NW>
__global__ __launch_bounds__(64 * NW, NW == 1 ? 1 : 8 / NW) void k_sn_solve_low(SnLowArgs a) {
    constexpr int NT = 64 * NW;
    constexpr int KC = 32;                         // bf16 K chunk: 2 MFMA steps of 16
    constexpr int LDR = KC + 8;                    // LDS row stride in bf16 (80 B, 8-byte aligned)
    constexpr int G = KC / 4;                      // 4-element groups per chunk row
    constexpr int kPerR = (kMaxW * G + NT - 1) / NT;
    constexpr int KS = KC / 16;
    constexpr int LDT = kMaxW + 1;                 // LDS row stride (floats) of a Tinv row block
    static_assert(2 * kMaxW * LDR * 2 >= 2 * 16 * LDT * 4, "Tinv blocks must fit the R buffer");
    __shared__ __align__(16) uint16_t sR[2 * kMaxW * LDR];
    __shared__ float sW[NW][kMaxW][17];
    __shared__ int sC[kSnCChunk];
    __shared__ int s_ticket;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) s_ticket = atomicAdd(&a.ctrl[0], 1);
    __syncthreads();
    const int t = s_ticket;
    const int k = a.sn - 1 - t / a.npanel;
    const int panel = t % a.npanel;
    if (k < 0) return;
    const int s0 = a.start[k], w = a.start[k + 1] - s0, r = a.rcount[k];
    const int c16 = lane & 15, quad = lane >> 4;
    const size_t m = (size_t)a.m;
    const int jj = 16 * (panel * NW + wv) + c16;
    const bool jv = jj < a.m;
    const int nI = (w + 15) >> 4;
    const int rp = (r + 15) & ~15;
    const uint16_t* R16 = a.rect16 + a.offr16[k];
    const int* C = a.cols + a.offc[k];
    v4f acc[8];
#pragma unroll
    for (int I = 0; I < 8; ++I) acc[I] = v4f{0.f, 0.f, 0.f, 0.f};
    const v4s zero = {0, 0, 0, 0};
    v4s rreg[kPerR];
    float bv[KS][4], bn[KS][4];
    auto load_r = [&](int kg) {  // R16[:, kg .. kg+KC) in 4-element groups (zeros outside)
#pragma unroll
        for (int q = 0; q < kPerR; ++q) {
            const int g = tid + q * NT, row = g / G, kq = kg + 4 * (g % G);
            rreg[q] = (row < w && kq < rp) ? *reinterpret_cast<const v4s*>(R16 + (size_t)row * rp + kq) : zero;
        }
    };
    auto store_r = [&](uint16_t* dst) {
#pragma unroll
        for (int q = 0; q < kPerR; ++q) {
            const int g = tid + q * NT;
            if (g / G < kMaxW) *reinterpret_cast<v4s*>(dst + (g / G) * LDR + 4 * (g % G)) = rreg[q];
        }
    };
    // independent of X: the first chunk of R16 and the first block of column indices load
    // while wave 0 waits for the flags
    load_r(0);
    for (int q = tid; q < min(kSnCChunk, r); q += NT) sC[q] = C[q];
    if (wv == 0) {
        Watchdog wd;
        wd.arm(&a.ctrl[1], a.limit);
        const int d1 = a.dep_ptr[k + 1];
        int idle = 0;
        bool give_up = false;
        for (int base = a.dep_ptr[k]; base < d1 && !give_up; base += kWave) {
            const int q = base + lane;
            bool ready = q >= d1;
            const int* f = ready ? nullptr : &a.flag[(size_t)a.dep[q] * a.npanel + panel];
            while (!__all(ready)) {
                if (!ready) ready = __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == a.epoch;
                if (__all(ready)) break;
                __builtin_amdgcn_s_sleep(2);
                if ((++idle & 63) == 0 && __any(wd.expired())) {
                    give_up = true;
                    break;
                }
            }
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");

    for (int cb = 0; cb < r; cb += kSnCChunk) {
        const int cn = min(kSnCChunk, r - cb);
        const int nch = (cn + KC - 1) / KC;
        auto load_x = [&](int kl) {  // gathered X rows of local chunk kl into bn
#pragma unroll
            for (int s = 0; s < KS; ++s) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int kk = kl + 16 * s + 4 * quad + e;
                    bn[s][e] = (kk < cn && jv) ? a.x[(size_t)sC[kk < cn ? kk : 0] * m + jj] : 0.f;
                }
            }
        };
        if (cb > 0) {  // the first block's indices and chunk were loaded before the wait
            __syncthreads();
            for (int q = tid; q < cn; q += NT) sC[q] = C[cb + q];
            __syncthreads();
        }
        load_x(0);
        if (cb > 0) load_r(cb);
        store_r(sR);
        __syncthreads();
        for (int c = 0; c < nch; ++c) {
            const uint16_t* cur = sR + (c & 1) * (kMaxW * LDR);
            uint16_t* nxt = sR + ((c + 1) & 1) * (kMaxW * LDR);
#pragma unroll
            for (int s = 0; s < KS; ++s)
#pragma unroll
                for (int e = 0; e < 4; ++e) bv[s][e] = bn[s][e];
            if (c + 1 < nch) {
                load_x(KC * (c + 1));
                load_r(cb + KC * (c + 1));
            }
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                v4s bb;
#pragma unroll
                for (int e = 0; e < 4; ++e) bb[e] = to_bf16(bv[s][e]);
#pragma unroll
                for (int I = 0; I < 8; ++I) {
                    if (I < nI) {
                        const v4s av = *reinterpret_cast<const v4s*>(cur + (16 * I + c16) * LDR + 16 * s + 4 * quad);
                        acc[I] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bb, acc[I], 0, 0, 0);
                    }
                }
            }
            if (c + 1 < nch) store_r(nxt);
            __syncthreads();
        }
    }
    // W = B_s - scale * acc, this wave's 16 columns, into LDS (rows 4 quad + v of each tile)
    const float sc = a.rscale[k];
#pragma unroll
    for (int I = 0; I < 8; ++I) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 16 * I + 4 * quad + v;
            if (i < kMaxW) sW[wv][i][c16] = (I < nI && i < w && jv) ? a.b[(size_t)(s0 + i) * m + jj] - sc * acc[I][v] : 0.f;
        }
    }
    // X_s = Tinv_s W, Tinv row blocks double-buffered in LDS
    float* sT = reinterpret_cast<float*>(sR);
    const float* Ti = a.tinv + a.offt[k];
    constexpr int kPerT = (16 * kMaxW + NT - 1) / NT;
    float treg[kPerT];
    auto load_t = [&](int I) {
#pragma unroll
        for (int q = 0; q < kPerT; ++q) {
            const int e = tid + q * NT, row = 16 * I + e / kMaxW, col = e % kMaxW;
            treg[q] = (e < 16 * kMaxW && row < w && col < w) ? Ti[(size_t)row * w + col] : 0.f;
        }
    };
    auto store_t = [&](float* dst) {
#pragma unroll
        for (int q = 0; q < kPerT; ++q) {
            const int e = tid + q * NT;
            if (e < 16 * kMaxW) dst[(e / kMaxW) * LDT + e % kMaxW] = treg[q];
        }
    };
    __syncthreads();
    load_t(0);
    store_t(sT);
    __syncthreads();
    for (int I = 0; I < nI; ++I) {
        const float* cur = sT + (I & 1) * (16 * LDT);
        float* nxt = sT + ((I + 1) & 1) * (16 * LDT);
        if (I + 1 < nI) load_t(I + 1);
        v4f xt = {0.f, 0.f, 0.f, 0.f};
        for (int kk = 16 * I; kk < w; kk += 4)
            xt = __builtin_amdgcn_mfma_f32_16x16x4f32(cur[c16 * LDT + kk + quad], sW[wv][kk + quad][c16], xt, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 16 * I + 4 * quad + v;
            if (i < w && jv) a.x[(size_t)(s0 + i) * m + jj] = xt[v];
        }
        if (I + 1 < nI) store_t(nxt);
        __syncthreads();
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (tid == 0)
        __hip_atomic_store(&a.flag[(size_t)k * a.npanel + panel], a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Ready flags for sn x npanel (supernode, column panel) pairs; a new array starts at 0, below
// every epoch in use.
int ensure_sn_flags(sptrsv_plan_s* P, int nblk) {
    if (nblk <= P->sn_flag_blocks) return SPTRSV_OK;
    SPT_HIP(hipStreamSynchronize(P->stream));
    if (P->sn_flag) SPT_HIP(hipFree(P->sn_flag));
    P->sn_flag = nullptr;
    const size_t count = (size_t)P->sn * nblk;
    SPT_HIP(hipMalloc((void**)&P->sn_flag, count * sizeof(int)));
    SPT_CHECK(zero_ints_async(P->sn_flag, (int64_t)count, P->stream));  // in stream order, before the solve
    P->sn_flag_blocks = nblk;
    return SPTRSV_OK;
}

int ensure_sn_low(sptrsv_plan_s* P) {
    if (P->sn_rect16) return SPTRSV_OK;
    const int S = P->sn;
    std::vector<int64_t> off16(S);
    int64_t tot16 = 0, tot_t = 0;
    for (int k = 0; k < S; ++k) {
        const int w = P->h_snodes[k + 1] - P->h_snodes[k];
        off16[k] = tot16;
        tot16 += (int64_t)w * ((P->h_sn_r[k] + 15) & ~15);
        tot_t += (int64_t)w * w;
    }
    SPT_CHECK(upload(&P->sn_offr16, off16));
    SPT_HIP(hipMalloc((void**)&P->sn_rect16, std::max<int64_t>(tot16, 1) * sizeof(uint16_t)));
    SPT_HIP(hipMalloc((void**)&P->sn_rscale, (size_t)S * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&P->sn_tinv32, std::max<int64_t>(tot_t, 1) * sizeof(float)));
    k_sn_pack16<<<S, 256, 0, P->stream>>>(P->sn_start, P->sn_r, P->sn_offr, P->sn_offr16, P->sn_rect,
                                          P->sn_rect16, P->sn_rscale);
    k_sn_to_float<<<(unsigned)std::min<int64_t>((tot_t + 255) / 256, 65536), 256, 0, P->stream>>>(
        P->sn_tinv, P->sn_tinv32, tot_t);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

}  // namespace

void supernodal_free(sptrsv_plan_s* p) {
    void* bufs[] = {p->sn_start, p->sn_r, p->sn_offt, p->sn_offr, p->sn_offc, p->sn_cols, p->sn_tinv,
                    p->sn_rect, p->sn_tinv32, p->sn_rect16, p->sn_rscale, p->sn_dep_ptr, p->sn_dep, p->sn_flag,
                    p->sn_tri};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    p->sn_start = p->sn_r = p->sn_offc = p->sn_cols = p->sn_dep_ptr = p->sn_dep = p->sn_flag = nullptr;
    p->sn_flag_blocks = 1;
    p->sn_offt = p->sn_offr = nullptr;
    p->sn_tinv = p->sn_rect = p->sn_tri = nullptr;
    p->sn_tinv32 = nullptr;
    p->sn_rect16 = nullptr;
    p->sn_rscale = nullptr;
    if (p->sn_offr16) (void)hipFree(p->sn_offr16);
    p->sn_offr16 = nullptr;
    p->h_sn_r.clear();
    p->sn_rmax = 0;
    p->sn_levels = 0;
    p->sn = 0;
    p->h_snodes.clear();
}

// Waves per workgroup for nblk 16-column blocks of X: above 4 blocks one workgroup of 8 waves
// takes all of them (up to 128 columns), so each supernode's R is streamed from HBM once.
inline int sn_waves(int nblk) { return nblk > 4 ? 8 : nblk >= 4 ? 4 : nblk >= 2 ? 2 : 1; }

int supernodal_solve(sptrsv_plan_s* P, int precision, const double* d_b, double* d_x, int m, uint64_t limit) {
    if (P->sn == 0)
        return fail(SPTRSV_ERR_UNSUPPORTED, "supernodal variant needs sptrsv_plan_build_supernodes first");
    if (precision != SPTRSV_FP64)
        return fail(SPTRSV_ERR_UNSUPPORTED, "supernodal variant: FP64 only through sptrsv_solve");
    const int nblk = (m + 15) / 16;
    // One column on a factor whose supernode dependency levels hold at most 256 supernodes on
    // average (chain-like: the one-workgroup-per-CU residency of k_sn_gemv costs nothing there)
    // and whose rectangular blocks fit its prefetch: k_sn_gemv.  Else k_sn_solve, NW waves per
    // panel of NW 16-column blocks; its one-wave form keeps four supernodes per CU resident.
    const uintptr_t xb = (uintptr_t)d_x, bb = (uintptr_t)d_b, len = (uintptr_t)P->n * m * sizeof(double);
    const bool apart = xb + len <= bb || bb + len <= xb;  // the sentinel fill of x must not touch b
    const bool gemv = m == 1 && apart && P->sn_rmax <= 4 * kSgPre && P->sn <= 256 * (int64_t)P->sn_levels;
    const int nw = sn_waves(nblk);
    const int npanel = gemv ? 1 : (nblk + nw - 1) / nw;
    SPT_CHECK(ensure_sn_flags(P, npanel));
    P->sn_epoch++;
    SnArgs a{P->sn, m, npanel, P->sn_start, P->sn_r, P->sn_offt, P->sn_offr, P->sn_offc, P->sn_cols, P->sn_tinv,
             P->sn_rect, P->sn_tri, P->sn_dep_ptr, P->sn_dep, P->sn_flag, P->sn_epoch, d_b, d_x, nullptr,
             P->ticket, limit};
    if (gemv) k_fill_sentinel_sn<<<(unsigned)std::min<int64_t>(((int64_t)P->n + 255) / 256, 65536), 256, 0, P->stream>>>(
        d_x, P->n, P->ticket);
    else k_reset_ticket_sn<<<1, 64, 0, P->stream>>>(P->ticket);
    hipEvent_t stop;
    SPT_CHECK(begin_timed(P, &stop));
    const unsigned grid = (unsigned)((int64_t)P->sn * npanel);
    if (gemv) k_sn_gemv<<<grid, 512, 0, P->stream>>>(a);
    else if (nw == 8) k_sn_solve<8, false><<<grid, 512, 0, P->stream>>>(a);
    else if (nw == 4) k_sn_solve<4, false><<<grid, 256, 0, P->stream>>>(a);
    else if (nw == 2) k_sn_solve<2, false><<<grid, 128, 0, P->stream>>>(a);
    else k_sn_solve<1, false><<<grid, 64, 0, P->stream>>>(a);
    end_timed(P, stop);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

// r = b - U x in FP64 on MFMA through the supernodal blocks, stored as FP32 (refinement).
int supernodal_residual32(sptrsv_plan_s* P, const double* d_b, const double* d_x, float* d_r32, int m) {
    if (P->sn == 0)
        return fail(SPTRSV_ERR_UNSUPPORTED, "supernodal variant needs sptrsv_plan_build_supernodes first");
    const int nblk = (m + 15) / 16;
    const int nw = sn_waves(nblk);
    const int npanel = (nblk + nw - 1) / nw;
    SnArgs a{P->sn, m, npanel, P->sn_start, P->sn_r, P->sn_offt, P->sn_offr, P->sn_offc, P->sn_cols, P->sn_tinv,
             P->sn_rect, P->sn_tri, P->sn_dep_ptr, P->sn_dep, P->sn_flag, 0, d_b, const_cast<double*>(d_x), d_r32,
             P->ticket, 0};
    const unsigned grid = (unsigned)((int64_t)P->sn * npanel);
    if (nw == 8) k_sn_solve<8, true><<<grid, 512, 0, P->stream>>>(a);
    else if (nw == 4) k_sn_solve<4, true><<<grid, 256, 0, P->stream>>>(a);
    else if (nw == 2) k_sn_solve<2, true><<<grid, 128, 0, P->stream>>>(a);
    else k_sn_solve<1, true><<<grid, 64, 0, P->stream>>>(a);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

int supernodal_solve_low(sptrsv_plan_s* P, int precision, const float* d_b32, float* d_x32, int m, uint64_t limit) {
    if (P->sn == 0)
        return fail(SPTRSV_ERR_UNSUPPORTED, "supernodal variant needs sptrsv_plan_build_supernodes first");
    if (precision != SPTRSV_BF16)
        return fail(SPTRSV_ERR_UNSUPPORTED, "reduced-precision supernodal solve is BF16 (rectangular) + FP32");
    SPT_CHECK(ensure_sn_low(P));
    const int nblk = (m + 15) / 16;
    const int nw = nblk >= 4 ? 4 : nblk >= 2 ? 2 : 1;
    const int npanel = (nblk + nw - 1) / nw;
    SPT_CHECK(ensure_sn_flags(P, npanel));
    P->sn_epoch++;
    SnLowArgs a{P->sn, m, npanel, P->sn_start, P->sn_r, P->sn_offt, P->sn_offr16, P->sn_offc, P->sn_cols,
                P->sn_tinv32, P->sn_rect16, P->sn_rscale, P->sn_dep_ptr, P->sn_dep, P->sn_flag, P->sn_epoch,
                d_b32, d_x32, P->ticket, limit};
    k_reset_ticket_sn<<<1, 64, 0, P->stream>>>(P->ticket);
    hipEvent_t stop;
    SPT_CHECK(begin_timed(P, &stop));
    const unsigned grid = (unsigned)((int64_t)P->sn * npanel);
    if (nw == 4) k_sn_solve_low<4><<<grid, 256, 0, P->stream>>>(a);
    else if (nw == 2) k_sn_solve_low<2><<<grid, 128, 0, P->stream>>>(a);
    else k_sn_solve_low<1><<<grid, 64, 0, P->stream>>>(a);
    end_timed(P, stop);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

}  // namespace sptrsv

using namespace sptrsv;

namespace sptrsv {
namespace {

// Relaxed supernodes: consecutive rows are merged while the merged block, padded with explicit
// zeros to a dense triangle plus one common column set (the union of the rows' columns right of
// the block), holds at most (1 + max_fill) times the rows' own entries and at most max_width
// rows.  A band of width 100 (the exact Cholesky factor of a 2D Laplacian) merges into blocks of
// 128 rows with w(w-1)/2 explicit zeros each.
std::vector<int> relaxed_bounds(int n, const std::vector<int>& rp, const std::vector<int>& col, int max_width,
                                double max_fill) {
    // block [s0, s0 + w): cnt[c] = block rows referencing column c; uni = distinct referenced
    // columns right of the block (>= s0 + w); orig = the rows' own entries
    std::vector<int> bounds{0}, cnt((size_t)n, 0), touched;
    int s0 = 0, uni = 0;
    int64_t orig = 0;
    for (int i = 0; i < n;) {
        const int64_t w = i - s0;
        // row i would join: column i moves inside the block, row i's columns (all > i) join the
        // union unless already referenced
        int uni2 = uni - (cnt[(size_t)i] > 0 ? 1 : 0);
        for (int p = rp[i] + 1; p < rp[i + 1]; ++p) uni2 += cnt[(size_t)col[p]] == 0 ? 1 : 0;
        const int64_t orig2 = orig + (rp[i + 1] - rp[i]);
        const int64_t padded = (w + 1) * (w + 2) / 2 + (w + 1) * (int64_t)uni2;
        if (w > 0 && (w + 1 > max_width || (double)(padded - orig2) > max_fill * (double)orig2)) {
            bounds.push_back(i);  // close the block before row i and start a new one at it
            for (int c : touched) cnt[(size_t)c] = 0;
            touched.clear();
            s0 = i;
            uni = 0;
            orig = 0;
            continue;
        }
        for (int p = rp[i] + 1; p < rp[i + 1]; ++p)
            if (cnt[(size_t)col[p]]++ == 0) touched.push_back(col[p]);
        uni = uni2;
        orig = orig2;
        ++i;
    }
    bounds.push_back(n);
    return bounds;
}

}  // namespace

// Supernodal layout from boundaries.  exact: every supernode must have the structure of the
// reference's supernodes (a dense triangle followed by one common column set in every row,
// src/core.cpp:53-78); otherwise rows are padded with explicit zeros to that structure.
int build_sn_layout(sptrsv_plan_s* P, const std::vector<int>& rp, const std::vector<int>& col,
                    const std::vector<double>& val, const std::vector<int>& bounds, bool exact) {
    const int n = P->n;
    // split wide supernodes (the first part's extra columns join its rectangular set)
    std::vector<int> sb;
    for (size_t k = 0; k + 1 < bounds.size(); ++k)
        for (int q = bounds[k]; q < bounds[k + 1]; q += kMaxW) sb.push_back(q);
    sb.push_back(n);
    const int S = (int)sb.size() - 1;

    std::vector<int> sn_of((size_t)n);
    for (int k = 0; k < S; ++k)
        for (int i = sb[k]; i < sb[k + 1]; ++i) sn_of[(size_t)i] = k;
    std::vector<int> rcount(S), offc(S), cols, dep_ptr(S + 1, 0), dep;
    std::vector<int64_t> offt(S), offr(S);
    int64_t tot_t = 0, tot_r = 0;
    int wmax = 0;
    std::vector<int> mark((size_t)n, -1), uni;
    for (int k = 0; k < S; ++k) {
        const int s0 = sb[k], w = sb[k + 1] - s0;
        wmax = std::max(wmax, w);
        if (exact) {
            const int len0 = rp[s0 + 1] - rp[s0];
            bool ok = len0 >= w;
            for (int t = 0; ok && t < w; ++t) ok = col[rp[s0] + t] == s0 + t;
            const int r = ok ? len0 - w : 0;
            for (int d = 1; ok && d < w; ++d) {
                const int row = s0 + d, a = rp[row];
                ok = rp[row + 1] - a == (w - d) + r;
                for (int t = 0; ok && t < w - d; ++t) ok = col[a + t] == row + t;
                for (int q = 0; ok && q < r; ++q) ok = col[a + (w - d) + q] == col[rp[s0] + w + q];
            }
            if (!ok)
                return fail(SPTRSV_ERR_BAD_INPUT, "rows " + std::to_string(s0) + ".." + std::to_string(s0 + w - 1) +
                                                      " do not form a supernode (dense triangle + common columns)");
        }
        // rectangular columns: the union of the rows' columns right of the block, ascending
        uni.clear();
        for (int d = 0; d < w; ++d)
            for (int p = rp[s0 + d]; p < rp[s0 + d + 1]; ++p)
                if (col[p] >= s0 + w && mark[(size_t)col[p]] != k) {
                    mark[(size_t)col[p]] = k;
                    uni.push_back(col[p]);
                }
        std::sort(uni.begin(), uni.end());
        const int r = (int)uni.size();
        rcount[k] = r;
        offt[k] = tot_t;
        offr[k] = tot_r;
        offc[k] = (int)cols.size();
        tot_t += (int64_t)w * w;
        tot_r += (int64_t)w * r;
        int last = -1;
        for (int c : uni) {
            cols.push_back(c);
            if (sn_of[(size_t)c] != last) {
                last = sn_of[(size_t)c];
                dep.push_back(last);
            }
        }
        dep_ptr[k + 1] = (int)dep.size();
    }
    std::vector<double> T((size_t)tot_t, 0.0), R((size_t)tot_r, 0.0);
    for (int k = 0; k < S; ++k) {
        const int s0 = sb[k], w = sb[k + 1] - s0, r = rcount[k];
        const int* cb = cols.data() + offc[k];
        for (int d = 0; d < w; ++d) {
            for (int p = rp[s0 + d]; p < rp[s0 + d + 1]; ++p) {
                const int c = col[p];
                if (c < s0 + w) {
                    T[(size_t)(offt[k] + (int64_t)d * w + (c - s0))] = val[(size_t)p];
                } else {
                    const int q = (int)(std::lower_bound(cb, cb + r, c) - cb);
                    R[(size_t)(offr[k] + (int64_t)d * r + q)] = val[(size_t)p];
                }
            }
        }
    }

    supernodal_free(P);
    std::vector<int> flags(S, 0);
    double* d_T = nullptr;
    SPT_CHECK(upload(&P->sn_start, sb));
    SPT_CHECK(upload(&P->sn_r, rcount));
    SPT_CHECK(upload(&P->sn_offt, offt));
    SPT_CHECK(upload(&P->sn_offr, offr));
    SPT_CHECK(upload(&P->sn_offc, offc));
    SPT_CHECK(upload(&P->sn_cols, cols));
    SPT_CHECK(upload(&P->sn_rect, R));
    SPT_CHECK(upload(&P->sn_dep_ptr, dep_ptr));
    SPT_CHECK(upload(&P->sn_dep, dep));
    SPT_CHECK(upload(&P->sn_flag, flags));
    SPT_CHECK(upload(&d_T, T));
    SPT_HIP(hipMalloc((void**)&P->sn_tinv, std::max<size_t>((size_t)tot_t, 1) * sizeof(double)));
    k_sn_invert<<<S, 128, 0, P->stream>>>(P->sn_start, P->sn_offt, d_T, P->sn_tinv);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(P->stream);
    P->sn_tri = d_T;  // kept for the FP64 residual
    if (e != hipSuccess) return fail(SPTRSV_ERR_HIP, std::string("supernode inversion: ") + hipGetErrorString(e));
    P->sn = S;
    P->h_sn_r = rcount;
    P->sn_rmax = rcount.empty() ? 0 : *std::max_element(rcount.begin(), rcount.end());
    {  // dependency levels of the supernodes (a supernode depends on later ones only)
        std::vector<int> lv(S, 1);
        int nl = S > 0 ? 1 : 0;
        for (int k = S - 1; k >= 0; --k) {
            for (int q = dep_ptr[k]; q < dep_ptr[k + 1]; ++q) lv[k] = std::max(lv[k], lv[dep[q]] + 1);
            nl = std::max(nl, lv[k]);
        }
        P->sn_levels = nl;
    }
    P->sn_wmax = wmax;
    P->sn_epoch = 0;
    P->sn_flag_blocks = 1;
    P->h_snodes = sb;
    int64_t stored = tot_t;
    for (int k = 0; k < S; ++k) stored -= (int64_t)(sb[k + 1] - sb[k]) * (sb[k + 1] - sb[k] - 1) / 2;
    P->sn_padded = stored + tot_r - P->nnz;
    return SPTRSV_OK;
}

int download_csr(sptrsv_plan_s* P, std::vector<int>& rp, std::vector<int>& col, std::vector<double>& val) {
    rp.resize((size_t)P->n + 1);
    col.resize((size_t)P->nnz);
    val.resize((size_t)P->nnz);
    SPT_HIP(hipStreamSynchronize(P->stream));
    SPT_HIP(hipMemcpy(rp.data(), P->rowptr, rp.size() * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(col.data(), P->col, col.size() * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(val.data(), P->val, val.size() * sizeof(double), hipMemcpyDeviceToHost));
    return SPTRSV_OK;
}

// AUTO on a factor of 16 or more entries per row: supernodes of at least 32 rows on average
// (without fill if the factor has such, else relaxed) turn its dependency chain of n rows into
// one of n / w blocks, each solved on FP64 MFMA (and its level sets into far fewer block
// levels).  Built once.
bool relaxed_supernodes_pay(sptrsv_plan_s* P) {
    if (P->relaxed_state == 0) {
        P->relaxed_state = -1;
        if ((double)P->nnz / std::max(P->n, 1) >= 16.0 && P->sn == 0) {
            std::vector<int> rp, col;
            std::vector<double> val;
            if (download_csr(P, rp, col, val) == SPTRSV_OK) {
                // supernodes without fill first (the factor's own, where it has them wide
                // enough); else relaxed ones that pad up to as many entries as the rows hold
                const auto wide = [&](const std::vector<int>& b) { return (double)P->n / (double)(b.size() - 1) >= 32.0; };
                std::vector<int> b = relaxed_bounds(P->n, rp, col, kMaxW, 0.0);
                if (!wide(b)) b = relaxed_bounds(P->n, rp, col, kMaxW, 1.0);
                if (wide(b) && build_sn_layout(P, rp, col, val, b, false) == SPTRSV_OK) P->relaxed_state = 1;
            }
        }
    }
    return P->relaxed_state == 1 && P->sn > 0;
}

}  // namespace sptrsv

extern "C" int sptrsv_plan_build_supernodes(sptrsv_plan_t P, int sn_in, const int* snodes_in) {
    if (!P) return fail(SPTRSV_ERR_BAD_INPUT, "null plan");
    const int n = P->n;
    std::vector<int> rp, col;
    std::vector<double> val;
    SPT_CHECK(download_csr(P, rp, col, val));
    std::vector<int> bounds;
    if (sn_in > 0) {
        if (!snodes_in) return fail(SPTRSV_ERR_BAD_INPUT, "null supernode array");
        if (snodes_in[0] != 0 || snodes_in[sn_in] != n)
            return fail(SPTRSV_ERR_BAD_INPUT, "supernode boundaries must start at 0 and end at n");
        for (int k = 0; k < sn_in; ++k)
            if (snodes_in[k + 1] <= snodes_in[k])
                return fail(SPTRSV_ERR_BAD_INPUT, "supernode boundaries must increase");
        bounds.assign(snodes_in, snodes_in + sn_in + 1);
    } else {
        // fundamental supernodes: row i+1 continues row i's supernode when
        // pattern(i) == {i} U pattern(i+1)
        bounds.push_back(0);
        for (int i = 0; i + 1 < n; ++i) {
            const int li = rp[i + 1] - rp[i], lj = rp[i + 2] - rp[i + 1];
            bool cont = li == lj + 1 && li >= 2 && col[rp[i] + 1] == i + 1;
            for (int t = 0; cont && t < lj; ++t) cont = col[rp[i] + 1 + t] == col[rp[i + 1] + t];
            if (!cont) bounds.push_back(i + 1);
        }
        bounds.push_back(n);
    }
    P->relaxed_state = -1;  // an explicit layout replaces any relaxed one AUTO may build
    return build_sn_layout(P, rp, col, val, bounds, true);
}

extern "C" int sptrsv_plan_build_supernodes_relaxed(sptrsv_plan_t P, int max_width, double max_fill) {
    if (!P) return fail(SPTRSV_ERR_BAD_INPUT, "null plan");
    if (max_width < 0 || max_fill < 0) return fail(SPTRSV_ERR_BAD_INPUT, "negative relaxation limit");
    std::vector<int> rp, col;
    std::vector<double> val;
    SPT_CHECK(download_csr(P, rp, col, val));
    const int w = max_width == 0 ? kMaxW : std::min(max_width, kMaxW);
    const std::vector<int> bounds = relaxed_bounds(P->n, rp, col, w, max_fill == 0 ? 1.0 : max_fill);
    P->relaxed_state = -1;
    return build_sn_layout(P, rp, col, val, bounds, false);
}

extern "C" int sptrsv_plan_supernodal_padding(sptrsv_plan_t P, int64_t* explicit_zeros) {
    if (!P || !explicit_zeros) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    *explicit_zeros = P->sn > 0 ? P->sn_padded : 0;
    return SPTRSV_OK;
}

extern "C" int sptrsv_plan_get_supernodes(sptrsv_plan_t P, int* sn, int* snodes) {
    if (!P) return fail(SPTRSV_ERR_BAD_INPUT, "null plan");
    if (sn) *sn = P->sn;
    if (snodes)
        for (size_t i = 0; i < P->h_snodes.size(); ++i) snodes[i] = P->h_snodes[i];
    return SPTRSV_OK;
}

namespace sptrsv {
namespace {

// Row r of supernode k (rows s_k .. s_{k+1}-1, d = r - s_k) gets d zeros at the supernode's
// leading columns s_k .. s_k + d - 1, then its own entries: every supernode becomes a dense
// trapezoid of width w_k + r_k.  Its padded start is rowptr[r] + pad[k] + d (d - 1) / 2, where
// pad[k] counts the zeros of the earlier supernodes.
__global__ void k_pad_rows(int n, const int* __restrict__ rowptr, const int* __restrict__ col,
                           const double* __restrict__ val, int sn, const int* __restrict__ snodes,
                           const int64_t* __restrict__ pad, double* __restrict__ val_pad,
                           uint64_t* __restrict__ col_pad, int* __restrict__ rowptr_pad) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        int lo = 0, hi = sn;  // supernode of r: snodes[k] <= r < snodes[k + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (snodes[mid] <= r) lo = mid;
            else hi = mid;
        }
        const int s = snodes[lo];
        const int64_t d = r - s;
        const int64_t start = rowptr[r] + pad[lo] + d * (d - 1) / 2;
        for (int64_t z = 0; z < d; ++z) {
            val_pad[start + z] = 0.0;
            col_pad[start + z] = (uint64_t)(s + z);
        }
        const int p0 = rowptr[r], p1 = rowptr[r + 1];
        for (int p = p0; p < p1; ++p) {
            val_pad[start + d + (p - p0)] = val[p];
            col_pad[start + d + (p - p0)] = (uint64_t)col[p];
        }
        rowptr_pad[r + 1] = (int)(start + d + (p1 - p0));
        if (r == 0) rowptr_pad[0] = 0;
    }
}

}  // namespace
}  // namespace sptrsv

extern "C" int sptrsv_ccs2ccs_pad_device(sptrsv_plan_t P, int sn, const int* snodes, double* d_val_pad,
                                         uint64_t* d_col_pad, int* d_rowptr_pad, int64_t* nnz_pad) {
    if (!P || !snodes || sn < 1) return fail(SPTRSV_ERR_BAD_INPUT, "null plan or empty partition");
    if (snodes[0] != 0 || snodes[sn] != P->n)
        return fail(SPTRSV_ERR_BAD_INPUT, "supernode boundaries must run from 0 to n");
    std::vector<int64_t> pad((size_t)sn + 1, 0);
    for (int k = 0; k < sn; ++k) {
        const int64_t w = (int64_t)snodes[k + 1] - snodes[k];
        if (w < 1) return fail(SPTRSV_ERR_BAD_INPUT, "supernode boundaries must increase");
        pad[(size_t)k + 1] = pad[(size_t)k] + w * (w - 1) / 2;
    }
    const int64_t total = P->nnz + pad[(size_t)sn];
    if (total >= ((int64_t)1 << 31)) return fail(SPTRSV_ERR_UNSUPPORTED, "padded nnz >= 2^31");
    if (nnz_pad) *nnz_pad = total;
    if (!d_val_pad && !d_col_pad && !d_rowptr_pad) return SPTRSV_OK;  // size query
    if (!d_val_pad || !d_col_pad || !d_rowptr_pad) return fail(SPTRSV_ERR_BAD_INPUT, "null output");
    int* d_sn = nullptr;
    int64_t* d_pad = nullptr;
    SPT_HIP(hipMalloc((void**)&d_sn, ((size_t)sn + 1) * sizeof(int)));
    if (hipMalloc((void**)&d_pad, ((size_t)sn + 1) * sizeof(int64_t)) != hipSuccess) {
        (void)hipFree(d_sn);
        return fail(SPTRSV_ERR_NOMEM, "device allocation failed");
    }
    hipError_t e = hipMemcpyAsync(d_sn, snodes, ((size_t)sn + 1) * sizeof(int), hipMemcpyHostToDevice, P->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_pad, pad.data(), ((size_t)sn + 1) * sizeof(int64_t), hipMemcpyHostToDevice, P->stream);
    if (e == hipSuccess) {
        const int threads = 256;
        const int blocks = (int)std::min<int64_t>((P->n + threads - 1) / threads, 65536);
        k_pad_rows<<<blocks, threads, 0, P->stream>>>(P->n, P->rowptr, P->col, P->val, sn, d_sn, d_pad, d_val_pad,
                                                      d_col_pad, d_rowptr_pad);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(P->stream);  // the temporaries are freed next
    (void)hipFree(d_sn);
    (void)hipFree(d_pad);
    if (e != hipSuccess) return fail(SPTRSV_ERR_HIP, std::string("ccs2ccs_pad_device: ") + hipGetErrorString(e));
    return SPTRSV_OK;
}
