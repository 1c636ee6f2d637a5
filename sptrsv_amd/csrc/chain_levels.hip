// chain_levels.hip -- level analysis for factors whose dependency graph is (nearly) a chain.
//
// The sync-free sweep (k_levels) pays a device round trip per level, so a random Zipf factor
// with 948,744 levels (n = 1e6) takes seconds.  Here one wave walks the rows in 32-row windows
// from the last row up, as the chain solver does, with max-plus in place of multiply-add.  The
// reference's rule (src/core.cpp:244-270),
//   level(r) = 1 + max(level(c) : c an off-diagonal column of row r), 1 without any,
// unrolled inside a window (lane l <-> row hi - 1 - l) reads
//   level_l = max_j (D[l][j] + e_j)
// where e_j is the deepest level among row j's columns outside the window (0 if none) and
// D[l][j] = 1 + the longest path from row l to row j through the window's own entries
// (D[l][l] = 1, kNoPath where there is none).  D is built once per window (one wave per window,
// all windows in parallel), with each lane's near columns (within 2048 rows above the window:
// their levels are in an LDS ring) and old columns (further up: read back from the levels
// array).  The solving wave loads each window's record kLAhead windows ahead and touches LDS
// only for the ring and the broadcast of e; a second wave reads every record's lines about
// kLTouch windows ahead of it, so that those loads hit L2.  Levels are integers: the result is
// exactly the sweep's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "common.h"
#include "plan.h"

namespace sptrsv {
namespace {

constexpr int kLW = 32;            // rows per window
constexpr int kLRing = 2048;       // levels kept in LDS; ring[kLRing] is a constant 0
constexpr int kLNear = 6;          // near columns per lane (12 per row)
constexpr int kLOld = 2;           // old columns per lane (4 per row)
constexpr int kLAhead = 6;         // windows between a record's load and its use
constexpr int kLTouch = 48;        // windows between a record's first read (into L2) and its use
constexpr unsigned kLDirect = 0xFFFF;  // nc[0] of every lane of a window on the direct path
constexpr int kNoPath = -(1 << 30);

// Record of one window, lane L = l + 32 h (5,632 bytes):
//   D   int[4][64][4]   D[l][16 h + 4 i + t] at ((i * 64 + L) * 4 + t)
//   nc  u16[64][8]      near columns' ring indices (column & (kLRing - 1)) at (L * 8 + e), e < 6;
//                       kLRing for padding; kLDirect in entry 0 of every lane: direct path
//   oc  int[64][2]      old columns at (L * 2 + q), -1 for none
constexpr int kLRecD = 0;
constexpr int kLRecNc = 4 * 64 * 16;
constexpr int kLRecOc = kLRecNc + 64 * 16;
constexpr int kLRecBytes = kLRecOc + 64 * 8;

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// ---- layout: one wave per window --------------------------------------------------------------
__global__ __launch_bounds__(64) void k_level_records(int n, int nwin, const int* __restrict__ rowptr,
                                                      const int* __restrict__ col, char* __restrict__ rec) {
    __shared__ unsigned dep[kLW];            // bit i: row l has an entry in column of lane i
    __shared__ int D[kLW][kLW + 1];
    __shared__ unsigned short nc[64][8];
    __shared__ int oc[64][2];
    __shared__ int direct;
    const int k = blockIdx.x;
    if (k >= nwin) return;
    const int lane = threadIdx.x;
    const int hi = n - kLW * k, lo = max(0, hi - kLW), rows = hi - lo;
    for (int e = 0; e < 8; ++e) nc[lane][e] = (unsigned short)kLRing;
    oc[lane][0] = oc[lane][1] = -1;
    if (lane == 0) direct = 0;
    __syncthreads();
    if (lane < kLW) {
        unsigned m = 0;
        if (lane < rows) {
            const int r = hi - 1 - lane;
            int ne = 0, no = 0;  // entry i of a kind goes to lane l + 32 (i & 1), slot i >> 1
            for (int p = rowptr[r] + 1; p < rowptr[r + 1]; ++p) {
                const int c = col[p];
                if (c < hi) {
                    m |= 1u << (hi - 1 - c);
                } else if (c < hi + kLRing) {
                    const int e = ne >> 1, L2 = lane + 32 * (ne & 1);
                    if (e < kLNear)
                        nc[L2][e] = (unsigned short)(c & (kLRing - 1));
                    else
                        atomicOr(&direct, 1);
                    ++ne;
                } else {
                    const int q = no >> 1, L2 = lane + 32 * (no & 1);
                    if (q < kLOld)
                        oc[L2][q] = c;
                    else
                        atomicOr(&direct, 1);
                    ++no;
                }
            }
        }
        dep[lane] = m;
    }
    __syncthreads();
    if (lane < kLW) {  // lane j: column j of D, rows in increasing order
        const int j = lane;
        for (int l = 0; l < j; ++l) D[l][j] = kNoPath;
        D[j][j] = 1;
        for (int l = j + 1; l < kLW; ++l) {
            unsigned mm = dep[l] & ~((1u << j) - 1u);  // dependencies i >= j (all below l)
            int best = kNoPath;
            while (mm) {
                const int i = __builtin_ctz(mm);
                mm &= mm - 1u;
                if (D[i][j] > kNoPath) best = max(best, D[i][j] + 1);
            }
            D[l][j] = best;
        }
    }
    __syncthreads();
    char* R = rec + (size_t)k * kLRecBytes;
    const int l = lane & (kLW - 1), h = lane >> 5;
    int* Dg = reinterpret_cast<int*>(R + kLRecD);
    for (int jj = 0; jj < 16; ++jj) Dg[((jj >> 2) * 64 + lane) * 4 + (jj & 3)] = D[l][16 * h + jj];
    unsigned short* ncg = reinterpret_cast<unsigned short*>(R + kLRecNc);
    int* ocg = reinterpret_cast<int*>(R + kLRecOc);
    for (int e = 0; e < 8; ++e)
        ncg[lane * 8 + e] = direct ? (unsigned short)(e == 0 ? kLDirect : kLRing) : nc[lane][e];
    for (int q = 0; q < 2; ++q) ocg[lane * 2 + q] = direct ? -1 : oc[lane][q];
}

// max of (value of lane L & 31, value of lane (L & 31) + 32), the same in every lane
__device__ __forceinline__ int half_max(int v) {
    const auto p = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
    return max((int)p[0], (int)p[1]);
}

struct LWin {
    i32x4 D[4];  // D[l][16 h + 4 i + t]
    u32x4 nc;    // eight ring indices, two per word
    i32x2 oc;    // old columns
};

__device__ __forceinline__ void lwin_load(const char* rec, int nwin, int k, int lane, LWin& w) {
    const char* R = rec + (size_t)min(k, nwin - 1) * kLRecBytes;
    const i32x4* d = reinterpret_cast<const i32x4*>(R + kLRecD);
#pragma unroll
    for (int i = 0; i < 4; ++i) w.D[i] = d[i * 64 + lane];
    w.nc = reinterpret_cast<const u32x4*>(R + kLRecNc)[lane];
    w.oc = reinterpret_cast<const i32x2*>(R + kLRecOc)[lane];
}

// deepest level among a lane's old columns (written at least 62 windows earlier by this wave,
// and every store of this wave completes in order with its loads)
__device__ __forceinline__ int old_levels(const int* levels, const LWin& w) {
    const int a = __hip_atomic_load(&levels[max(w.oc.x, 0)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int b = __hip_atomic_load(&levels[max(w.oc.y, 0)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return max(w.oc.x >= 0 ? a : 0, w.oc.y >= 0 ? b : 0);
}

// One window: the record of window k + kLAhead into `fill`, the old levels of window
// k + kLAhead / 2 (whose record is `half`) into `half_old`, levels of window k from `cur`.
__device__ __forceinline__ void level_window(int n, int nwin, int k, int lane, const int* __restrict__ rowptr,
                                             const int* __restrict__ col, const char* rec, int* levels,
                                             int* ring, int* gbuf, const LWin& cur, int cur_old, LWin& fill,
                                             const LWin& half, int& half_old, int& deepest) {
    const int l = lane & (kLW - 1), h = lane >> 5;
    const int hi = n - kLW * k;
    const int r = hi - 1 - l;
    const bool valid = r >= 0;
    int e = cur_old;
    const unsigned w0 = cur.nc.x;
    if (__builtin_amdgcn_readfirstlane(w0 & 0xFFFFu) != kLDirect) {
        const int a0 = ring[w0 & 0xFFFFu], a1 = ring[w0 >> 16];
        const int a2 = ring[cur.nc.y & 0xFFFFu], a3 = ring[cur.nc.y >> 16];
        const int a4 = ring[cur.nc.z & 0xFFFFu], a5 = ring[cur.nc.z >> 16];
        e = max(max(e, max(a0, a1)), max(max(a2, a3), max(a4, a5)));
    } else if (valid) {
        // columns beyond the record's capacity: straight from the CSR, alternating between the
        // row's two lanes, sixteen entries of the row per pass with their loads in flight together
        const int end = rowptr[r + 1];
        for (int p0 = rowptr[r] + 1; p0 < end; p0 += 16) {
            int cc[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int q = p0 + h + 2 * u;
                cc[u] = q < end ? col[q] : -1;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int c = cc[u];
                if (c >= hi) {
                    const int v = c < hi + kLRing ? ring[c & (kLRing - 1)]
                                                  : __hip_atomic_load(&levels[c], __ATOMIC_RELAXED,
                                                                      __HIP_MEMORY_SCOPE_AGENT);
                    e = max(e, v);
                }
            }
        }
    }
    e = half_max(e);
    int* gk = gbuf + (k & 1) * kLW;
    if (h == 0) gk[l] = e;
    __builtin_amdgcn_wave_barrier();
    const i32x4* gb = reinterpret_cast<const i32x4*>(gk + 16 * h);
    int c0 = kNoPath, c1 = kNoPath;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const i32x4 g = gb[i];
        c0 = max(c0, max(cur.D[i].x + g.x, cur.D[i].y + g.y));
        c1 = max(c1, max(cur.D[i].z + g.z, cur.D[i].w + g.w));
    }
    const int lev = half_max(max(c0, c1));
    if (h == 0 && valid) {
        ring[r & (kLRing - 1)] = lev;
        levels[r] = lev;
        deepest = max(deepest, lev);
    }
    lwin_load(rec, nwin, k + kLAhead, lane, fill);
    half_old = old_levels(levels, half);
}

__global__ __launch_bounds__(128) void k_chain_levels(int n, int nwin, const int* __restrict__ rowptr,
                                                      const int* __restrict__ col, const char* rec, int* levels,
                                                      int* ctrl) {
    __shared__ int ring[kLRing + 1];
    __shared__ __align__(16) int gbuf[2][kLW];
    __shared__ int done;  // windows finished by the solving wave
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (tid == 0) {
        ring[kLRing] = 0;  // what padding entries read
        done = 0;
    }
    __syncthreads();
    if (wave == 1) {
        // one word of every 128-byte line of each record, kLTouch windows ahead of the solving wave
        unsigned touch = 0;
        for (int w = 0; w < nwin; ++w) {
            while (w > __hip_atomic_load(&done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) + kLTouch)
                __builtin_amdgcn_s_sleep(4);
            const unsigned* R = reinterpret_cast<const unsigned*>(rec + (size_t)w * kLRecBytes);
            touch |= R[min(lane * 32, kLRecBytes / 4 - 1)];  // 64 lanes x 128 bytes >= a record
        }
        __asm__ volatile("" ::"v"(touch));  // a use of the words read: the reads stay
        return;
    }
    static_assert(kLAhead == 6, "the loop below rotates six record sets");
    LWin w[6];
    int o[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) lwin_load(rec, nwin, i, lane, w[i]);
#pragma unroll
    for (int i = 0; i < 3; ++i) o[i] = old_levels(levels, w[i]);  // none: the first 64 windows have no old columns
    int deepest = 0;
    for (int k0 = 0; k0 < nwin; k0 += 6) {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int k = k0 + q;
            if (k < nwin) {
                level_window(n, nwin, k, lane, rowptr, col, rec, levels, ring, &gbuf[0][0], w[q], o[q], w[q],
                             w[(q + 3) % 6], o[(q + 3) % 6], deepest);
                if (lane == 0) __hip_atomic_store(&done, k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int v = __shfl_xor(deepest, off, 64);
        deepest = v > deepest ? v : deepest;
    }
    if (lane == 0) ctrl[2] = deepest;
}

}  // namespace

// Levels of a chain-like factor; leaves the deepest level in ctrl[2] (as k_levels does).
int chain_levels(sptrsv_plan_s* p, int* ctrl) {
    const int n = p->n, nwin = (n + kLW - 1) / kLW;
    if (n == 0) return SPTRSV_OK;
    char* rec = nullptr;
    SPT_HIP(hipMalloc((void**)&rec, (size_t)nwin * kLRecBytes));
    k_level_records<<<nwin, 64, 0, p->stream>>>(n, nwin, p->rowptr, p->col, rec);
    k_chain_levels<<<1, 128, 0, p->stream>>>(n, nwin, p->rowptr, p->col, rec, p->levels, ctrl);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    (void)hipFree(rec);
    if (e != hipSuccess) return fail(SPTRSV_ERR_HIP, std::string("chain level analysis: ") + hipGetErrorString(e));
    return SPTRSV_OK;
}

}  // namespace sptrsv
