// chain.hip -- solver for factors whose dependency graph is (nearly) a chain.
//
// When almost every row depends on the next one (random Zipf factors: 948,744 levels for
// n = 1e6; banded factors: n levels) there is no parallelism across rows, and any dependency
// that travels through HBM costs microseconds; a row-by-row walk inside one wave still pays a
// dependent multiply-add and a cross-lane broadcast per row.  This solver shortens the chain
// 32-fold with precomputed inverses:
//   * rows are taken in windows of 32 from the last row up; window k holds rows hi-32 .. hi-1
//     (hi = n - 32 k); lane L = l + 32 h owns row hi - 1 - l in both halves h = 0, 1;
//   * the window's triangle T_k (its rows restricted to its own columns) is inverted once, when
//     the layout is built (one wave per window, all windows in parallel): A_k = T_k^-1;
//   * per window the solve forms g = b - (products with every column outside the window, all
//     solved before) and x = A_k g, a 32 x 32 product split between the two halves of the wave.
// Every window has a fixed-size record, built once, laid out lane-major so that each lane reads
// its own part with conflict-free LDS loads: the lane's 16 entries of A_k, its "near" entries
// (columns less than 2048 rows above the window, whose x is still in an LDS ring; stored as ring
// indices) and its "old" entries (columns further up, whose x is read back from HBM).  A row's
// outside entries alternate between its two lanes.
// One workgroup: wave 0 is the only wave on the dependency chain.  It reads LDS only, holds
// the next window's record in registers while it solves the current one, and sums across
// halves with v_permlane32_swap, so a window costs two dependent LDS round trips (the ring
// gather and the broadcast of g).  Waves 1..6 stage records three windows deep (record ->
// x of its old entries -> LDS slot, 10 slots) and fold the old entries into b; wave 7 copies
// finished windows from the ring to x in HBM.  Windows whose entries do not fit a record take a
// direct path through the CSR.
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
constexpr int kRing = 2048;               // x values kept in LDS; ring[kRing] is a constant 0
constexpr int kNear = 6;                  // near entries per lane (12 per row); the layout assumes 6
constexpr int kOldSlots = 2;              // old entries per lane (4 per row)
constexpr int kSlots = 10;                // staged windows
constexpr int kStagers = 6;               // stager waves
constexpr int kChainThreads = 64 * (2 + kStagers);  // + the solving wave and the writer wave
constexpr int kWriteChunk = 16;           // windows the writer copies to HBM at a time

// Byte layout of one window's record for value type T; lane L = l + 32 h:
//   A     T[8][64][2]      A_k[l][16 h + 2 i + t] at ((i * 64 + L) * 2 + t), 0 where the column
//                          index exceeds l (A_k is lower triangular in lane order)
//   nval  T[3][64][2]      near entry e of lane L at (((e >> 1) * 64 + L) * 2 + (e & 1)), 0 for
//                          padding
//   ncol  u16[64][4]       its ring index (column & (kRing - 1)) at (L * 4 + e) for e < 4, kRing
//                          for padding
//   ncol2 u16[64][2]       the same for e = 4, 5 at (L * 2 + e - 4)
//   flags int[4]           [0] = 1: the window takes the direct path
//   --- the part above is staged into an LDS slot ---
//   oval  T[kOldSlots][64] old entries, folded into b by the stager
//   ocol  int[kOldSlots][64]   their columns (-1: none)
template <typename T>
struct RecLayout {
    static constexpr int nval() { return 16 * 64 * (int)sizeof(T); }
    static constexpr int ncol() { return nval() + kNear * 64 * (int)sizeof(T); }
    static constexpr int ncol2() { return ncol() + 64 * 4 * 2; }
    static constexpr int flags() { return ncol2() + 64 * 2 * 2; }
    static constexpr int staged() { return flags() + 16; }
    static constexpr int oval() { return staged(); }
    static constexpr int ocol() { return oval() + kOldSlots * 64 * (int)sizeof(T); }
    static constexpr int bytes() { return ocol() + kOldSlots * 64 * 4; }
    static constexpr int pieces() { return (staged() / 16 + 63) / 64; }  // 16-byte pieces per stager lane
};
static_assert(RecLayout<double>::staged() % 16 == 0 && RecLayout<double>::bytes() % 16 == 0, "record alignment");
static_assert(RecLayout<float>::staged() % 16 == 0 && RecLayout<float>::bytes() % 16 == 0, "record alignment");
static_assert(kNear == 6, "ncol and ncol2 hold six ring indices per lane");

// position of A_k[l][j] in the record, in elements of T
__host__ __device__ constexpr int apos(int l, int j) {
    return ((((j & 15) >> 1) * 64 + l + 32 * (j >> 4)) * 2) + (j & 1);
}

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
    // lane L writes A[l][16 h + jj], jj = 0..15 (zeros above the diagonal)
    TA* dst = reinterpret_cast<TA*>(rec + (size_t)k * rec_bytes);
    const int l = lane & (kW - 1), h = lane >> 5;
    for (int jj = 0; jj < 16; ++jj) {
        const int j = 16 * h + jj;
        dst[apos(l, j)] = (TA)(j <= l ? A[l][j] : 0.0);
    }
}

template <typename T>
struct ChainArgs {
    int n;
    int nwin;
    const char* rec;    // nwin records of RecLayout<T>::bytes()
    const int* rowptr;  // direct path
    const int* col;
    const T* val;
    const T* b;
    T* x;
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));  // a native 16-byte register piece
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
template <typename T> struct Pair;
template <> struct Pair<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct Pair<float> { typedef float type __attribute__((ext_vector_type(2))); };

// (value of lane L & 31) + (value of lane (L & 31) + 32), the same in every lane: the two halves
// are exchanged with v_permlane32_swap (no LDS round trip)
__device__ __forceinline__ double half_sum(double v) {
    const unsigned lo = __double2loint(v), hi = __double2hiint(v);
    const auto pl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto ph = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return __hiloint2double(ph[0], pl[0]) + __hiloint2double(ph[1], pl[1]);
}
__device__ __forceinline__ float half_sum(float v) {
    const unsigned u = __float_as_uint(v);
    const auto p = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}

// What a stager holds for one window between its pipeline steps.
template <typename T, int NP>
struct Staged {
    u32x4 piece[NP];    // the staged part of the record, 16-byte pieces, lane-strided
    T bv;               // b of row hi - 1 - (lane & 31)
    int ocol[kOldSlots];
    T oval[kOldSlots];
    T oldx[kOldSlots];  // x of those columns
};

// Step 1: the record, b and the old entries.  Loads are unconditional (indices clamped into
// range) so that the registers are not spilled and each wait falls at the first use.
template <typename T, int NP>
__device__ __forceinline__ void stage_load(const ChainArgs<T>& a, int k, int lane, Staged<T, NP>& s) {
    using L = RecLayout<T>;
    const int kk = min(k, a.nwin - 1);
    const char* r = a.rec + (size_t)kk * L::bytes();
    const u32x4* src = reinterpret_cast<const u32x4*>(r);
    constexpr int npieces = L::staged() / 16;
#pragma unroll
    for (int q = 0; q < NP; ++q) s.piece[q] = src[min(lane + 64 * q, npieces - 1)];
    const int row = max(0, a.n - kW * kk - 1 - (lane & (kW - 1)));
    s.bv = a.b[row];
#pragma unroll
    for (int q = 0; q < kOldSlots; ++q) {
        s.ocol[q] = reinterpret_cast<const int*>(r + L::ocol())[q * 64 + lane];
        s.oval[q] = reinterpret_cast<const T*>(r + L::oval())[q * 64 + lane];
    }
}

// Step 2: x of the old entries.  They were solved at least 2048 rows (64 windows) before the
// window; `published` (windows whose x stores have completed) is waited for to make sure.
template <typename T, int NP>
__device__ __forceinline__ void stage_old(const ChainArgs<T>& a, int k, Staged<T, NP>& s, int* published) {
    while (__hip_atomic_load(published, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < min(k, a.nwin) - 63)
        __builtin_amdgcn_s_sleep(2);
#pragma unroll
    for (int q = 0; q < kOldSlots; ++q) {
        const T v = from_bits<T>(poll_bits(s.ocol[q] >= 0 ? a.x + s.ocol[q] : a.b));  // unused slots read b[0]
        s.oldx[q] = s.ocol[q] >= 0 ? v : T(0);
    }
}

// Step 3: into the window's LDS slot once the slot's previous window is solved; b minus the old
// entries' products goes to slot_s (row l's b in lane l, only the products in lane l + 32).
template <typename T, int NP>
__device__ __forceinline__ void stage_deposit(const ChainArgs<T>& a, int k, int lane, const Staged<T, NP>& s,
                                              char* slot, T* slot_s, int* slot_full, int* done) {
    if (k >= a.nwin) return;
    while (__hip_atomic_load(done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < k - kSlots + 1)
        __builtin_amdgcn_s_sleep(1);
    u32x4* dst = reinterpret_cast<u32x4*>(slot);
    constexpr int npieces = RecLayout<T>::staged() / 16;
#pragma unroll
    for (int q = 0; q < NP; ++q)
        if (lane + 64 * q < npieces) dst[lane + 64 * q] = s.piece[q];
    const bool valid = a.n - kW * k - 1 - (lane & (kW - 1)) >= 0;
    T acc = (lane < kW && valid) ? s.bv : T(0);
#pragma unroll
    for (int q = 0; q < kOldSlots; ++q) acc = fma(-s.oval[q], s.oldx[q], acc);
    slot_s[lane] = acc;
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) __hip_atomic_store(&slot_full[k % kSlots], k + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The solving wave's registers for one window.
template <typename T>
struct Win {
    T A[16];      // A_k[l][16 h + jj]
    T nv[kNear];  // near entries
    int nc[kNear];
    T s;          // b minus the old entries' products (lane share)
    int flags;
};

template <typename T>
__device__ __forceinline__ void win_load(const char* slot, const T* slot_s, int lane, Win<T>& w) {
    using L = RecLayout<T>;
    using P = typename Pair<T>::type;
    const P* a2 = reinterpret_cast<const P*>(slot);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const P v = a2[i * 64 + lane];
        w.A[2 * i] = v.x;
        w.A[2 * i + 1] = v.y;
    }
    // 15 LDS reads in all: issued behind the 8 reads of g, they stay within the 15 that the
    // LDS counter can leave outstanding, so the product with g waits for none of them
    const P* nv = reinterpret_cast<const P*>(slot + L::nval());
#pragma unroll
    for (int i = 0; i < kNear / 2; ++i) {
        const P v = nv[i * 64 + lane];
        w.nv[2 * i] = v.x;
        w.nv[2 * i + 1] = v.y;
    }
    const u16x4 c = reinterpret_cast<const u16x4*>(slot + L::ncol())[lane];
    const u16x2 c2 = reinterpret_cast<const u16x2*>(slot + L::ncol2())[lane];
    w.nc[0] = c.x;
    w.nc[1] = c.y;
    w.nc[2] = c.z;
    w.nc[3] = c.w;
    w.nc[4] = c2.x;
    w.nc[5] = c2.y;
    w.s = slot_s[lane];
    w.flags = *reinterpret_cast<const int*>(slot + L::flags());
}

// One window on the solving wave: g from the record in `cur` (or the CSR), the next window's
// record into `nxt`, x = A_k g into the ring.  SB: bytes of an LDS slot.
template <typename T, int SB>
__device__ __forceinline__ void solve_window(const ChainArgs<T>& a, int k, int lane, int& hi, int& wr,
                                             const Win<T>& cur, Win<T>& nxt, T* ring, T* gbuf, char* slot,
                                             T* slot_s, int* slot_full, int* done, int* written, int* published) {
    using P = typename Pair<T>::type;
    const int l = lane & (kW - 1), h = lane >> 5;
    const int r = hi - 1 - l;
    const bool valid = r >= 0;
    // whether the next window is staged: read now, looked at after the gather below (its wait
    // covers this read too)
    const int sl = (k + 1) % kSlots;
    int staged = __hip_atomic_load(&slot_full[sl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    // g: this lane's share of b minus the products with columns outside the window
    T acc = cur.s;
    if (cur.flags == 0) {
        T part = T(0);
#pragma unroll
        for (int e = 0; e < kNear; ++e) part = fma(cur.nv[e], ring[cur.nc[e]], part);
        acc -= part;
    } else if (valid) {
        // entries beyond the record's capacity: straight from the CSR, alternating between the
        // row's two lanes; x of old columns (at least 64 windows back) is read from HBM once the
        // writer has published it
        while (__hip_atomic_load(published, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < k - 63)
            __builtin_amdgcn_s_sleep(2);
        T part = T(0);
        for (int p = a.rowptr[r] + 1 + h; p < a.rowptr[r + 1]; p += 2) {
            const int c = a.col[p];
            if (c < hi) continue;
            const T xv = c < hi + kRing ? ring[c & (kRing - 1)] : from_bits<T>(poll_bits(a.x + c));
            part = fma(a.val[p], xv, part);
        }
        acc -= part;
    }
    const T g = half_sum(acc);
    if (k + 1 < a.nwin) {
        while (staged != k + 2) {
            __builtin_amdgcn_s_sleep(1);
            staged = __hip_atomic_load(&slot_full[sl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    // The slot's contents were stored before its flag was (the stager's release), and this
    // wave's LDS reads execute in issue order: the reads below see them.  The empty asm keeps
    // the compiler from moving them above the flag test.
    __asm__ volatile("" ::: "memory");
    T* gk = gbuf + (k & 1) * kW;
    if (h == 0) gk[l] = g;
    __builtin_amdgcn_wave_barrier();
    const P* gb = reinterpret_cast<const P*>(gk + 16 * h);
    P gv[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) gv[i] = gb[i];
    // the next window's record, read while the broadcast of g is in flight (after the last
    // window: a read of a slot that is not used)
    win_load(slot + sl * SB, slot_s + sl * 64, lane, nxt);
    // x = A_k g: 16 columns per half, four partial sums
    T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
#pragma unroll
    for (int i = 0; i < 8; i += 2) {
        s0 = fma(cur.A[2 * i], gv[i].x, s0);
        s1 = fma(cur.A[2 * i + 1], gv[i].y, s1);
        s2 = fma(cur.A[2 * i + 2], gv[i + 1].x, s2);
        s3 = fma(cur.A[2 * i + 3], gv[i + 1].y, s3);
    }
    const T xv = half_sum((s0 + s1) + (s2 + s3));
    // the ring entries about to be replaced (window k - 64) must have been copied to HBM
    if (wr < k - 63) {
        while ((wr = __hip_atomic_load(written, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) < k - 63)
            __builtin_amdgcn_s_sleep(1);
    }
    if (h == 0 && valid) ring[r & (kRing - 1)] = xv;
    // This wave's LDS accesses complete in issue order, so the ring store and every read of the
    // window's slot precede the store below: a relaxed store suffices, and the empty asm keeps
    // the compiler from moving memory accesses across it.
    __asm__ volatile("" ::: "memory");
    if (lane == 0) __hip_atomic_store(done, k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    hi -= kW;
}

template <typename T, int NP>
__global__ __launch_bounds__(kChainThreads, 1) void k_chain(ChainArgs<T> a) {
    constexpr int kSlotBytes = NP * 64 * 16;
    __shared__ T ring[kRing + 1];
    __shared__ __align__(16) char slot[kSlots][kSlotBytes];
    __shared__ T slot_s[kSlots][64];
    __shared__ __align__(16) T gbuf[2][kW];
    __shared__ int slot_full[kSlots];  // window + 1 staged in the slot
    __shared__ int done;               // windows solved
    __shared__ int written;            // windows copied from the ring to HBM (stores issued)
    __shared__ int published;          // windows whose x stores have completed
    // the wave's role as a scalar: the branches below are then uniform, and the compiler keeps
    // the roles' code apart instead of linearising it
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (tid < kSlots) slot_full[tid] = 0;
    if (tid == 0) {
        done = written = published = 0;
        ring[kRing] = T(0);  // what padding entries read
    }
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
            stage_deposit(a, k, lane, p0, slot[k % kSlots], slot_s[k % kSlots], slot_full, &done);
            p0 = p1;
            p1 = p2;
        }
        return;
    }

    // ---- wave 0: the windows in order, two per loop pass so that the two register sets of
    // window records alternate roles without copies ---------------------------------------------
    while (__hip_atomic_load(&slot_full[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != 1)
        __builtin_amdgcn_s_sleep(1);
    Win<T> w0, w1;
    win_load(slot[0], slot_s[0], lane, w0);
    int wr = 0;  // last value of `written` seen
    int hi = a.n;
    for (int k = 0; k < a.nwin; k += 2) {
        solve_window<T, kSlotBytes>(a, k, lane, hi, wr, w0, w1, ring, &gbuf[0][0], &slot[0][0], &slot_s[0][0],
                                    slot_full, &done, &written, &published);
        if (k + 1 >= a.nwin) break;
        solve_window<T, kSlotBytes>(a, k + 1, lane, hi, wr, w1, w0, ring, &gbuf[0][0], &slot[0][0], &slot_s[0][0],
                                    slot_full, &done, &written, &published);
    }
}

// Host side: every window's near and old entries and the records.
template <typename T>
int build_records(sptrsv_plan_s* P, char** out) {
    using L = RecLayout<T>;
    const int n = P->n, nwin = (n + kW - 1) / kW;
    std::vector<int> rp((size_t)n + 1), col((size_t)P->nnz);
    std::vector<double> val((size_t)P->nnz);
    SPT_HIP(hipStreamSynchronize(P->stream));
    SPT_HIP(hipMemcpy(rp.data(), P->rowptr, rp.size() * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(col.data(), P->col, col.size() * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(val.data(), P->val, val.size() * sizeof(double), hipMemcpyDeviceToHost));
    const int rb = L::bytes();
    std::vector<char> host((size_t)nwin * rb, 0);
    for (int k = 0; k < nwin; ++k) {
        char* R = host.data() + (size_t)k * rb;
        const int hi = n - kW * k, lo = std::max(0, hi - kW);
        T* nv = reinterpret_cast<T*>(R + L::nval());
        unsigned short* nc = reinterpret_cast<unsigned short*>(R + L::ncol());
        unsigned short* nc2 = reinterpret_cast<unsigned short*>(R + L::ncol2());
        auto ncref = [&](int e, int lane) -> unsigned short& { return e < 4 ? nc[lane * 4 + e] : nc2[lane * 2 + e - 4]; };
        int* flags = reinterpret_cast<int*>(R + L::flags());
        T* ov = reinterpret_cast<T*>(R + L::oval());
        int* oc = reinterpret_cast<int*>(R + L::ocol());
        for (int q = 0; q < 64; ++q)
            for (int e = 0; e < kNear; ++e) ncref(e, q) = (unsigned short)kRing;
        for (int q = 0; q < kOldSlots * 64; ++q) oc[q] = -1;
        auto npos = [](int e, int lane) { return (((e >> 1) * 64 + lane) * 2) + (e & 1); };
        bool direct = false;
        for (int l = 0; l < kW && !direct; ++l) {
            const int r = hi - 1 - l;
            if (r < lo) break;
            int ne = 0, no = 0;  // entry i of a kind goes to lane l + 32 (i & 1), slot i >> 1
            for (int p = rp[r] + 1; p < rp[r + 1]; ++p) {
                const int c = col[p];
                if (c < hi) continue;
                if (c < hi + kRing) {
                    const int e = ne >> 1, L2 = l + 32 * (ne & 1);
                    if (e >= kNear) {
                        direct = true;
                        break;
                    }
                    nv[npos(e, L2)] = (T)val[(size_t)p];
                    ncref(e, L2) = (unsigned short)(c & (kRing - 1));
                    ++ne;
                } else {
                    const int q = no >> 1, L2 = l + 32 * (no & 1);
                    if (q >= kOldSlots) {
                        direct = true;
                        break;
                    }
                    ov[q * 64 + L2] = (T)val[(size_t)p];
                    oc[q * 64 + L2] = c;
                    ++no;
                }
            }
        }
        if (direct) {  // the solving wave reads the CSR for this window: no near or old entries
            for (int q = 0; q < 64; ++q)
                for (int e = 0; e < kNear; ++e) {
                    nv[npos(e, q)] = T(0);
                    ncref(e, q) = (unsigned short)kRing;
                }
            for (int q = 0; q < kOldSlots * 64; ++q) {
                ov[q] = T(0);
                oc[q] = -1;
            }
        }
        *flags = direct ? 1 : 0;
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
    return SPTRSV_OK;
}

template <typename T>
int run_chain(sptrsv_plan_s* P, const char* rec, const T* val, const T* b, T* x) {
    if (P->n == 0) return SPTRSV_OK;  // no windows: nothing to stage
    const int nwin = (P->n + kW - 1) / kW;
    ChainArgs<T> a{P->n, nwin, rec, P->rowptr, P->col, val, b, x};
    k_chain<T, RecLayout<T>::pieces()><<<1, kChainThreads, 0, P->stream>>>(a);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
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
    if (!P->chain_rec) SPT_CHECK(build_records<double>(P, &P->chain_rec));
    hipEvent_t stop;
    SPT_CHECK(begin_timed(P, &stop));
    SPT_CHECK(run_chain<double>(P, P->chain_rec, P->val, d_b, d_x));
    end_timed(P, stop);
    return SPTRSV_OK;
}

int chain_solve32(sptrsv_plan_s* P, const float* d_b, float* d_x) {
    if (!P->chain_rec32) SPT_CHECK(build_records<float>(P, &P->chain_rec32));
    return run_chain<float>(P, P->chain_rec32, P->val32, d_b, d_x);
}

}  // namespace sptrsv
