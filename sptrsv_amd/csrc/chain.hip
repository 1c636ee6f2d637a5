// chain.hip -- solver for factors whose dependency graph is (nearly) a chain.
//
// When almost every row depends on the next one (random Zipf factors: 948,744 levels for
// n = 1e6; banded factors: n levels) there is no parallelism across rows, and any dependency
// that travels through HBM costs microseconds; a row-by-row walk inside one wave still pays a
// dependent multiply-add and a cross-lane broadcast per row.  This solver shortens the chain
// 32-fold with precomputed inverses:
//   * rows are taken in windows of 32 from the last row up; window k holds rows hi-32 .. hi-1
//     (hi = n - 32 k); lane l and lane l + 32 both own row hi - 1 - l (halves h = 0, 1);
//   * the window's triangle T_k (its rows restricted to its own columns) is inverted once, when
//     the layout is built (one wave per window, all windows in parallel), and kept as a packed
//     lower-triangular 32 x 32 block A_k = T_k^-1 in window-lane order;
//   * per window the solve forms g = b - (products with every column outside the window, all
//     solved before) and x = A_k g: a 32 x 32 product split between the two halves of the wave.
// Every window has a fixed-size record, built once: A_k, its "near" entries (columns less than
// 2048 rows above the window, whose x is still in an LDS ring; stored as ring offsets) and its
// "old" entries (columns further up, whose x is read back from HBM).  One workgroup: wave 0 is
// the only wave on the dependency chain and touches LDS only (x goes to the LDS ring); waves
// 1..6 stage records through a three-step pipeline (record -> x of its old entries -> LDS slot,
// 12 slots) and wave 7 copies finished windows from the ring to x in HBM, so that memory latency
// stays off the chain.  Windows whose entries do not fit a record take a direct path.
// x is formed in a different order than the serial solver's sums (A_k g instead of
// substitution): it agrees with it to rounding (relative error <= 1e-12 in FP64).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "common.h"
#include "plan.h"

namespace sptrsv {
namespace {

constexpr int kW = 32;                    // rows per window
constexpr int kPack = kW * (kW + 1) / 2;  // entries of a packed lower-triangular inverse
constexpr int kRing = 2048;               // x values kept in LDS
constexpr int kOld = 64;                  // old entries a record holds
constexpr int kSlots = 12;                // staged windows
constexpr int kStagers = 6;               // stager waves
constexpr int kChainThreads = 64 * (2 + kStagers);  // + the solving wave and the writer wave
constexpr int kWriteChunk = 8;            // windows the writer copies to HBM at a time
constexpr int kMaxNear = 448;             // near entries a record may hold (record <= 10 KiB)

// packed position of A[l][j] (j <= l): column-major by j
__host__ __device__ constexpr int wpos(int l, int j) { return j * kW - j * (j - 1) / 2 + (l - j); }

// Byte layout of one window's record for value type T and near capacity E (multiple of 64).
template <typename T>
struct RecLayout {
    int E;
    __host__ __device__ int inv() const { return 0; }
    __host__ __device__ int nval() const { return kPack * (int)sizeof(T); }
    __host__ __device__ int oval() const { return nval() + E * (int)sizeof(T); }
    __host__ __device__ int ocol() const { return oval() + kOld * (int)sizeof(T); }
    __host__ __device__ int ncol() const { return ocol() + kOld * 4; }  // int16 ring offsets
    __host__ __device__ int nrp() const { return ncol() + E * 2; }      // int16 x 33
    __host__ __device__ int orp() const { return nrp() + 34 * 2; }      // int16 x 33
    __host__ __device__ int flags() const { return orp() + 34 * 2; }    // int: 1 = direct path
    __host__ __device__ int bytes() const { return (flags() + 4 + 15) & ~15; }
};

// ---- layout: the inverse of every window's triangle ----------------------------------------
// One wave per window: lane l < 32 scatters row l's in-window entries into T (LDS), then lane j
// computes column j of T^-1 by forward substitution in window-lane order (T[l][i], i <= l).
template <typename TA>
__global__ __launch_bounds__(64) void k_chain_invert(int n, int nwin, const int* __restrict__ rowptr,
                                                     const int* __restrict__ col, const double* __restrict__ val,
                                                     char* __restrict__ rec, int rec_bytes) {
    __shared__ double T[kW][kW + 1];
    __shared__ double A[kW][kW + 1];
    const int k = blockIdx.x;
    if (k >= nwin) return;
    const int lane = threadIdx.x;
    const int hi = n - kW * k, lo = max(0, hi - kW), rows = hi - lo;
    for (int e = lane; e < kW * (kW + 1); e += 64) {
        (&T[0][0])[e] = 0.0;
        (&A[0][0])[e] = 0.0;
    }
    __syncthreads();
    if (lane < kW) {
        if (lane < rows) {
            const int r = hi - 1 - lane;
            for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
                const int c = col[p];
                if (c >= hi) break;  // columns ascend: the rest is outside the window
                T[lane][hi - 1 - c] = val[p];
            }
        } else {
            T[lane][lane] = 1.0;  // rows past the matrix start: identity, never stored
        }
    }
    __syncthreads();
    if (lane < kW) {
        const int j = lane;
        A[j][j] = 1.0 / T[j][j];
        for (int l = j + 1; l < kW; ++l) {
            double acc = 0.0;
            for (int i = j; i < l; ++i) acc = fma(T[l][i], A[i][j], acc);
            A[l][j] = -acc / T[l][l];
        }
    }
    __syncthreads();
    TA* dst = reinterpret_cast<TA*>(rec + (size_t)k * rec_bytes);
    for (int e = lane; e < kPack; e += 64) {
        int j = 0, base = 0;  // invert wpos: column j holds kW - j entries
        while (base + (kW - j) <= e) {
            base += kW - j;
            ++j;
        }
        dst[e] = (TA)A[j + (e - base)][j];
    }
}

template <typename T>
struct ChainArgs {
    int n;
    int nwin;
    int rec_bytes;
    RecLayout<T> L;
    const char* rec;    // nwin records
    const int* rowptr;  // direct path
    const int* col;
    const T* val;
    const T* b;
    T* x;
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));  // a native 16-byte register piece

// What a stager holds for one window between its pipeline steps.
template <typename T, int NP>
struct Staged {
    u32x4 piece[NP];  // the record, 16-byte pieces, lane-strided
    T bv;             // b of row hi - 1 - lane (lanes < 32)
    int ocol;         // column of old entry `lane` (-1: none)
    T oldx;           // x of that column
};

// Step 1: the record, b and the old columns.  Loads are unconditional (indices clamped into
// range) so that the registers are not spilled and each wait falls at the first use.
template <typename T, int NP>
__device__ __forceinline__ void stage_load(const ChainArgs<T>& a, int k, int lane, Staged<T, NP>& s) {
    const int kk = min(k, a.nwin - 1);
    const char* r = a.rec + (size_t)kk * a.rec_bytes;
    const u32x4* src = reinterpret_cast<const u32x4*>(r);
    const int npieces = a.rec_bytes / 16;
#pragma unroll
    for (int q = 0; q < NP; ++q) s.piece[q] = src[min(lane + 64 * q, npieces - 1)];
    const int row = max(0, a.n - kW * kk - 1 - (lane & (kW - 1)));
    s.bv = a.b[row];
    s.ocol = reinterpret_cast<const int*>(r + a.L.ocol())[lane];
}

// Step 2: x of the old entries.  They were solved at least 2048 rows (64 windows) before the
// window; `published` (windows whose x stores have completed) is waited for to make sure.
template <typename T, int NP>
__device__ __forceinline__ void stage_old(const ChainArgs<T>& a, int k, Staged<T, NP>& s, int* published) {
    while (__hip_atomic_load(published, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < min(k, a.nwin) - 63)
        __builtin_amdgcn_s_sleep(2);
    s.oldx = from_bits<T>(poll_bits(s.ocol >= 0 ? a.x + s.ocol : a.b));  // unused entries read b[0]
}

// Step 3: into the window's LDS slot once the slot's previous window is solved.
template <typename T, int NP>
__device__ __forceinline__ void stage_deposit(const ChainArgs<T>& a, int k, int lane, const Staged<T, NP>& s,
                                              char* slot, T* slot_b, T* slot_oldx, int* slot_full, int* done) {
    if (k >= a.nwin) return;
    while (__hip_atomic_load(done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < k - kSlots + 1)
        __builtin_amdgcn_s_sleep(1);
    u32x4* dst = reinterpret_cast<u32x4*>(slot);
    const int npieces = a.rec_bytes / 16;
#pragma unroll
    for (int q = 0; q < NP; ++q)
        if (lane + 64 * q < npieces) dst[lane + 64 * q] = s.piece[q];
    if (lane < kW) slot_b[lane] = s.bv;
    slot_oldx[lane] = s.oldx;
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) __hip_atomic_store(&slot_full[k % kSlots], k + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <typename T, int NP>
__global__ __launch_bounds__(kChainThreads, 1) void k_chain(ChainArgs<T> a) {
#pragma clang fp contract(off)
    constexpr int kSlotBytes = NP * 64 * 16;
    __shared__ T ring[kRing];
    __shared__ __align__(16) char slot[kSlots][kSlotBytes];
    __shared__ T slot_b[kSlots][kW];
    __shared__ T slot_oldx[kSlots][kOld];
    __shared__ T gbuf[2][kW];
    __shared__ int slot_full[kSlots];  // window + 1 staged in the slot
    __shared__ int done;               // windows solved
    __shared__ int written;            // windows copied from the ring to HBM (stores issued)
    __shared__ int published;          // windows whose x stores have completed
    // the wave's role as a scalar: the branches below are then uniform, and the compiler keeps
    // the roles' code apart instead of linearising it (which would carry the stagers' pending
    // loads into the solving wave's waits)
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (tid < kSlots) slot_full[tid] = 0;
    if (tid == 0) done = written = published = 0;
    __syncthreads();

    if (wave == kStagers + 1) {
        // ---- writer: finished windows from the ring to x, kWriteChunk at a time --------------
        int w0 = 0;
        while (w0 < a.nwin) {
            int d;
            while ((d = __hip_atomic_load(&done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) <= w0)
                __builtin_amdgcn_s_sleep(2);
            const int w1 = min(d, w0 + kWriteChunk);
            const int rlo = max(0, a.n - kW * w1), rhi = a.n - kW * w0;
            for (int r = rlo + lane; r < rhi; r += 64) a.x[r] = ring[r & (kRing - 1)];
            w0 = w1;
            if (lane == 0) __hip_atomic_store(&written, w0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            wait_vmem();  // the stores have completed: old-x readers may use them
            if (lane == 0) __hip_atomic_store(&published, w0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        return;
    }

    if (wave > 0) {
        // ---- stagers: wave w stages windows w-1, w-1+kStagers, ..., three in flight ---------
        const int s0 = wave - 1;
        Staged<T, NP> p0, p1, p2;
        stage_load(a, s0, lane, p0);
        stage_load(a, s0 + kStagers, lane, p1);
        stage_old(a, s0, p0, &published);
        for (int k = s0; k < a.nwin; k += kStagers) {
            stage_load(a, k + 2 * kStagers, lane, p2);
            stage_old(a, k + kStagers, p1, &published);
            stage_deposit(a, k, lane, p0, slot[k % kSlots], slot_b[k % kSlots], slot_oldx[k % kSlots], slot_full,
                          &done);
            p0 = p1;
            p1 = p2;
        }
        return;
    }

    // ---- wave 0: the windows in order -------------------------------------------------------
    const int l = lane & (kW - 1), h = lane >> 5;
    int hi = a.n;
    for (int k = 0; k < a.nwin; ++k) {
        const int r = hi - 1 - l;
        const bool valid = r >= 0;
        const int sl = k % kSlots;
        while (__hip_atomic_load(&slot_full[sl], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != k + 1)
            __builtin_amdgcn_s_sleep(1);
        const char* rec = slot[sl];
        T acc = T(0);
        if (*reinterpret_cast<const int*>(rec + a.L.flags()) == 0) {
            const T* nval = reinterpret_cast<const T*>(rec + a.L.nval());
            const short* ncol = reinterpret_cast<const short*>(rec + a.L.ncol());
            const short* nrp = reinterpret_cast<const short*>(rec + a.L.nrp());
            const T* oval = reinterpret_cast<const T*>(rec + a.L.oval());
            const short* orp = reinterpret_cast<const short*>(rec + a.L.orp());
            const T* oldx = slot_oldx[sl];
            for (int e = nrp[l] + h; e < nrp[l + 1]; e += 2) acc += nval[e] * ring[(hi + ncol[e]) & (kRing - 1)];
            for (int e = orp[l] + h; e < orp[l + 1]; e += 2) acc += oval[e] * oldx[e];
        } else if (valid) {
            // entries beyond the record's capacity: straight from the CSR; x of old columns (at
            // least 64 windows back) is read from HBM once the writer has published it
            while (__hip_atomic_load(&published, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < k - 63)
                __builtin_amdgcn_s_sleep(2);
            for (int p = a.rowptr[r] + 1 + h; p < a.rowptr[r + 1]; p += 2) {
                const int c = a.col[p];
                if (c < hi) continue;
                const T xv = c < hi + kRing ? ring[c & (kRing - 1)] : from_bits<T>(poll_bits(a.x + c));
                acc += a.val[p] * xv;
            }
        }
        const T g = slot_b[sl][l] - (acc + __shfl_xor(acc, 32));
        if (h == 0) gbuf[k & 1][l] = g;
        __builtin_amdgcn_wave_barrier();
        const T* A = reinterpret_cast<const T*>(rec + a.L.inv());
        const T* gb = gbuf[k & 1];
        T s = T(0);
#pragma unroll
        for (int jj = 0; jj < kW / 2; ++jj) {
            const int j = (kW / 2) * h + jj;
            const T av = A[j <= l ? wpos(l, j) : wpos(j, j)];
            s += (j <= l ? av : T(0)) * gb[j];
        }
        const T xv = s + __shfl_xor(s, 32);
        // the ring entries about to be replaced (window k - 64) must have been copied to HBM
        while (__hip_atomic_load(&written, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < k - 63)
            __builtin_amdgcn_s_sleep(1);
        if (h == 0 && valid) ring[r & (kRing - 1)] = xv;
        // release: the slot's LDS reads have returned and x is in the ring, so the stager may
        // refill the slot and the writer may copy the window (this wave has no HBM accesses in
        // flight here, so the release waits on LDS only)
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) __hip_atomic_store(&done, k + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        hi -= kW;
    }
}

// Host side: every window's near and old entries, the record capacity, and the records.
template <typename T>
int build_records(sptrsv_plan_s* P, char** out, int* rec_bytes, int* near_cap) {
    const int n = P->n, nwin = (n + kW - 1) / kW;
    std::vector<int> rp((size_t)n + 1), col((size_t)P->nnz);
    std::vector<double> val((size_t)P->nnz);
    SPT_HIP(hipStreamSynchronize(P->stream));
    SPT_HIP(hipMemcpy(rp.data(), P->rowptr, rp.size() * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(col.data(), P->col, col.size() * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(val.data(), P->val, val.size() * sizeof(double), hipMemcpyDeviceToHost));
    // near capacity: the largest window's near count, in steps of 64, at most kMaxNear
    int emax = 0;
    for (int k = 0; k < nwin; ++k) {
        const int hi = n - kW * k, lo = std::max(0, hi - kW);
        int cnt = 0;
        for (int r = lo; r < hi; ++r)
            for (int p = rp[r] + 1; p < rp[r + 1]; ++p) cnt += (col[p] >= hi && col[p] < hi + kRing) ? 1 : 0;
        emax = std::max(emax, cnt);
    }
    RecLayout<T> L{std::min(kMaxNear, std::max(64, (emax + 63) & ~63))};
    const int rb = L.bytes();
    std::vector<char> host((size_t)nwin * rb, 0);
    for (int k = 0; k < nwin; ++k) {
        char* R = host.data() + (size_t)k * rb;
        const int hi = n - kW * k, lo = std::max(0, hi - kW);
        T* nv = reinterpret_cast<T*>(R + L.nval());
        T* ov = reinterpret_cast<T*>(R + L.oval());
        int* oc = reinterpret_cast<int*>(R + L.ocol());
        short* nc = reinterpret_cast<short*>(R + L.ncol());
        short* nrp = reinterpret_cast<short*>(R + L.nrp());
        short* orp = reinterpret_cast<short*>(R + L.orp());
        int* flags = reinterpret_cast<int*>(R + L.flags());
        for (int q = 0; q < kOld; ++q) oc[q] = -1;
        int ne = 0, no = 0;
        for (int l = 0; l < kW; ++l) {
            nrp[l] = (short)std::min(ne, L.E);
            orp[l] = (short)std::min(no, kOld);
            const int r = hi - 1 - l;
            if (r < lo) continue;
            for (int p = rp[r] + 1; p < rp[r + 1]; ++p) {
                const int c = col[p];
                if (c < hi) continue;
                if (c < hi + kRing) {
                    if (ne < L.E) {
                        nv[ne] = (T)val[(size_t)p];
                        nc[ne] = (short)(c - hi);
                    }
                    ++ne;
                } else {
                    if (no < kOld) {
                        ov[no] = (T)val[(size_t)p];
                        oc[no] = c;
                    }
                    ++no;
                }
            }
        }
        nrp[kW] = (short)std::min(ne, L.E);
        orp[kW] = (short)std::min(no, kOld);
        *flags = (ne > L.E || no > kOld) ? 1 : 0;
    }
    char* d = nullptr;
    SPT_HIP(hipMalloc((void**)&d, host.size()));
    SPT_HIP(hipMemcpy(d, host.data(), host.size(), hipMemcpyHostToDevice));
    k_chain_invert<T><<<nwin, 64, 0, P->stream>>>(n, nwin, P->rowptr, P->col, P->val, d, rb);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(P->stream);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return fail(SPTRSV_ERR_HIP, std::string("chain layout: ") + hipGetErrorString(e));
    }
    *out = d;
    *rec_bytes = rb;
    *near_cap = L.E;
    return SPTRSV_OK;
}

template <typename T, int NP>
int launch_chain(sptrsv_plan_s* P, const char* rec, int rec_bytes, int near_cap, const T* val, const T* b, T* x) {
    const int nwin = (P->n + kW - 1) / kW;
    ChainArgs<T> a{P->n, nwin, rec_bytes, RecLayout<T>{near_cap}, rec, P->rowptr, P->col, val, b, x};
    k_chain<T, NP><<<1, kChainThreads, 0, P->stream>>>(a);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

// The kernel instance is picked by the record size: NP 16-byte pieces per stager lane.
template <typename T>
int run_chain(sptrsv_plan_s* P, const char* rec, int rec_bytes, int near_cap, const T* val, const T* b, T* x) {
    if (P->n == 0) return SPTRSV_OK;  // no windows: nothing to stage
    const int np = (rec_bytes / 16 + 63) / 64;
    if (np > 10) return fail(SPTRSV_ERR_UNSUPPORTED, "chain layout: record larger than an LDS slot");
    if (np <= 4) return launch_chain<T, 4>(P, rec, rec_bytes, near_cap, val, b, x);
    if (np <= 6) return launch_chain<T, 6>(P, rec, rec_bytes, near_cap, val, b, x);
    if (np <= 8) return launch_chain<T, 8>(P, rec, rec_bytes, near_cap, val, b, x);
    return launch_chain<T, 10>(P, rec, rec_bytes, near_cap, val, b, x);
}

}  // namespace

void chain_free(sptrsv_plan_s* P) {
    if (P->chain_rec) (void)hipFree(P->chain_rec);
    if (P->chain_rec32) (void)hipFree(P->chain_rec32);
    P->chain_rec = P->chain_rec32 = nullptr;
}

int chain_solve(sptrsv_plan_s* P, int precision, const double* d_b, double* d_x, int m) {
    if (m != 1) return fail(SPTRSV_ERR_UNSUPPORTED, "chain variant solves one right-hand side");
    if (precision != SPTRSV_FP64) return fail(SPTRSV_ERR_UNSUPPORTED, "chain variant is FP64 through sptrsv_solve");
    if (!P->chain_rec) SPT_CHECK(build_records<double>(P, &P->chain_rec, &P->chain_rec_bytes, &P->chain_near));
    hipEvent_t stop;
    SPT_CHECK(begin_timed(P, &stop));
    SPT_CHECK(run_chain<double>(P, P->chain_rec, P->chain_rec_bytes, P->chain_near, P->val, d_b, d_x));
    end_timed(P, stop);
    return SPTRSV_OK;
}

int chain_solve32(sptrsv_plan_s* P, const float* d_b, float* d_x) {
    if (!P->chain_rec32) SPT_CHECK(build_records<float>(P, &P->chain_rec32, &P->chain_rec32_bytes, &P->chain_near32));
    return run_chain<float>(P, P->chain_rec32, P->chain_rec32_bytes, P->chain_near32, P->val32, d_b, d_x);
}

}  // namespace sptrsv
