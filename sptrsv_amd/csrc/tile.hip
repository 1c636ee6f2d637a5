// tile.hip -- tiled sync-free SpTRSV (m == 1) for gfx950.
//
// A dependency that crosses the GPU costs a device-scope store-then-poll round trip (about
// 2 us); one inside a CU costs an LDS access after a workgroup barrier.  Tile t is the rows
// [t T, (t+1) T) (T = 16,384 by default, so the tile's x fits in LDS); one workgroup per tile,
// tiles taken from the last one down by workgroup tickets.  Inside a tile the rows run in
// steps: a step is up to kTileThreads rows of one level of the tile (levels ascending), with a
// barrier between steps.  Every dependency inside the tile has a lower level, so it was solved
// in an earlier step and is read from LDS; a dependency in a later tile is read from the
// self-publishing x (signalling-NaN sentinel until stored).  On a natural-order 3D stencil
// factor with T = one plane, only the plane-to-plane dependency crosses the GPU: the critical
// path holds about one remote hop per plane plus one barrier per level, instead of one remote
// hop per level.
//
// Keeping the steps short is about gfx950's memory counters: one in-order counter (at most 63
// outstanding loads) covers every vector load of a wave, and waiting on any load waits on all
// older ones.  So
//   * every position is one 80-byte record (row, length, first four off-diagonal entries,
//     diagonal) read with five 16-byte loads plus its b (gathered into position order per
//     solve), kIdxLead steps ahead, into a register ring (the loop is unrolled by kIdxSlots);
//     no load depends on another;
//   * a poller wave requests the remote x values kPollLead steps ahead (its remote lists
//     kListLead steps ahead) and hands them to the compute threads through an LDS ring;
//   * the step barrier waits for LDS operations only (__syncthreads would drain every
//     outstanding load and store of the wave).
// Each row sums its products in ascending column order and divides by the diagonal -- the
// reference serial solver's arithmetic (src/core.cpp:18-32) -- so FP64 x is bitwise equal to
// it.  A tile's remote columns belong to later tiles (earlier tickets, already running), which
// guarantees progress.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "common.h"
#include "plan.h"

namespace sptrsv {
namespace {

constexpr int kTileThreads = 256;   // compute threads: rows per step (one thread each)
constexpr int kTileMaxRows = 16384; // x of the tile in LDS: 128 KiB
constexpr int kTilePre = 4;         // off-diagonal entries per row held in the record
constexpr int kPollE = 4;           // remote-list entries per poller lane and step (<= 256 per step)
constexpr int kPollLead = 5;        // remote values requested this many steps ahead
constexpr int kListLead = 8;        // remote lists loaded this many steps ahead
constexpr int kSlots = 10;          // poller register slots (> kListLead)
constexpr int kIdxLead = 6;         // compute threads load a step's records this many steps ahead
constexpr int kIdxSlots = 8;        // their register ring (> kIdxLead)
constexpr int kRing = kTileThreads * kTilePre;  // remote-value slots per step
constexpr int kLenMask = (1 << 24) - 1;         // len: row length | ring bits << 24

// One position (row slot of a step) = five 16-byte words:
//   w0 {row, len | ring bits << 24, e0 (first entry in the tile-ordered CSR), c0}
//   w1 {c1, c2, c3, 0}   w2 {diag, unused}   w3 {v0, v1}   w4 {v2, v3}
struct alignas(16) TileRec {
    int row, len, e0, c0;
    int c1, c2, c3, pad;
    double diag, spare;
    double v0, v1;
    double v2, v3;
};
static_assert(sizeof(TileRec) == 80, "tile record layout");

struct TileArgs {
    int n;
    int rows;          // T
    int n_tiles;
    int cap;           // remote-list entries per step (multiple of 64, <= 64 * kPollE)
    const TileRec* rec;
    const double* bt;  // b gathered into position order (per solve)
    const int* col;    // tile-ordered CSR entries (beyond the record's four)
    const double* val;
    const int2* rlist; // per step: cap (remote column, ring slot); column -1 = none
    double* x;
    int* ctrl;
    uint64_t limit;
};

struct RecRegs {
    int4 w0, w1, w2, w3, w4;
    double b;
};

__device__ __forceinline__ void load_rec(const TileArgs& a, int p, RecRegs& q) {
    const int4* w = reinterpret_cast<const int4*>(a.rec + p);
    q.w0 = w[0];
    q.w1 = w[1];
    q.w2 = w[2];
    q.w3 = w[3];
    q.w4 = w[4];
    q.b = a.bt[p];
}

__device__ __forceinline__ double dbl(int lo, int hi) {
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Workgroup barrier that waits for this wave's LDS operations only.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0), vmcnt / expcnt not waited
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

__device__ __forceinline__ uint64_t wait_bits(const TileArgs& a, int c, uint64_t bits, Watchdog& wd, int& idle) {
    while (bits == Bits<double>::sentinel) {
        __builtin_amdgcn_s_sleep(1);
        if ((++idle & 63) == 0 && wd.expired()) return 0;
        bits = poll_bits(a.x + c);
    }
    return bits;
}

// Workgroup of kTileThreads compute threads + one poller wave.  Iteration it (0..S, S steps)
// ends with a barrier; the compute threads solve step it - 1 in it; the poller loads the remote
// list of step it + kListLead, requests the remote values of step it + kPollLead and completes
// those of step it into ring[it % 2].  The step tables come in as restrict-qualified
// parameters: their uniform reads compile to scalar loads (lgkmcnt), not vector loads.
__global__ __launch_bounds__(kTileThreads + kWave) void k_tile(TileArgs a, const int* __restrict__ sptr,
                                                              const int* __restrict__ step) {
#pragma clang fp contract(off)
    extern __shared__ double xs[];
    __shared__ uint64_t ring[2][kRing];
    __shared__ int s_ticket;
    const int tid = threadIdx.x;
    if (tid == 0) s_ticket = atomicAdd(&a.ctrl[0], 1);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(a.n_tiles - 1 - s_ticket);
    if (t < 0) return;
    const int r0 = t * a.rows, r1 = min(a.n, r0 + a.rows);
    const int s0 = sptr[t];
    const int S = sptr[t + 1] - s0;
    Watchdog wd;
    wd.arm(&a.ctrl[1], a.limit);
    int idle = 0;
    if (tid >= kTileThreads) {
        // ---- poller wave ------------------------------------------------------------------
        const int lane = tid - kTileThreads;
        const int E = a.cap / kWave;
        int2 pl[kSlots][kPollE];
        uint64_t pv[kSlots][kPollE];
        auto load_list = [&](int i, int sl) {
#pragma unroll
            for (int e = 0; e < kPollE; ++e) {
                const bool ok = i < S && e < E;
                const size_t j = ok ? (size_t)(s0 + i) * a.cap + lane + kWave * e : 0;
                pl[sl][e] = ok ? a.rlist[j] : int2{-1, 0};
            }
        };
        auto request = [&](int sl) {
#pragma unroll
            for (int e = 0; e < kPollE; ++e)
                pv[sl][e] = pl[sl][e].x >= 0 ? poll_bits(a.x + pl[sl][e].x) : Bits<double>::sentinel;
        };
#pragma unroll
        for (int i = 0; i < kListLead; ++i) load_list(i, i);
#pragma unroll
        for (int i = 0; i < kPollLead; ++i) request(i);
        for (int base = 0; base <= S; base += kSlots) {
#pragma unroll
            for (int j = 0; j < kSlots; ++j) {
                const int it = base + j;
                if (it <= S) {
                    load_list(it + kListLead, (j + kListLead) % kSlots);
                    request((j + kPollLead) % kSlots);
                    if (it < S) {
#pragma unroll
                        for (int e = 0; e < kPollE; ++e) {
                            const int c = pl[j][e].x;
                            if (c >= 0)
                                __hip_atomic_store(&ring[it & 1][pl[j][e].y], wait_bits(a, c, pv[j][e], wd, idle),
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                    }
                    lds_barrier();
                }
            }
        }
        return;
    }
    // ---- compute threads ------------------------------------------------------------------
    RecRegs q[kIdxSlots];
#pragma unroll
    for (int i = 0; i < kIdxLead; ++i) {
        const int pb = step[s0 + min(i, S)], pe = step[s0 + min(i + 1, S)];
        load_rec(a, pb + tid < pe ? pb + tid : 0, q[i]);
    }
    lds_barrier();  // iteration 0: the poller completes step 0's remote values
    for (int base = 0; base < S; base += kIdxSlots) {
#pragma unroll
        for (int j = 0; j < kIdxSlots; ++j) {
            const int i = base + j;
            if (i < S) {
                const int ia = i + kIdxLead;
                if (ia < S) {
                    const int pb = step[s0 + ia], pe = step[s0 + ia + 1];
                    load_rec(a, pb + tid < pe ? pb + tid : pb, q[(j + kIdxLead) % kIdxSlots]);
                }
                if (step[s0 + i] + tid < step[s0 + i + 1]) {
                    const RecRegs& r = q[j];
                    const int row = r.w0.x, len = r.w0.y & kLenMask, rbits = r.w0.y >> 24, e0 = r.w0.z;
                    const int cs[kTilePre] = {r.w0.w, r.w1.x, r.w1.y, r.w1.z};
                    const double vs[kTilePre] = {dbl(r.w3.x, r.w3.y), dbl(r.w3.z, r.w3.w), dbl(r.w4.x, r.w4.y),
                                                 dbl(r.w4.z, r.w4.w)};
                    double acc = 0.0;
#pragma unroll
                    for (int k = 0; k < kTilePre; ++k) {
                        if (k + 1 < len) {
                            const int c = cs[k];
                            double xv;
                            if (c < r1)
                                xv = xs[c - r0];
                            else if ((rbits >> k) & 1)
                                xv = from_bits<double>(__hip_atomic_load(&ring[i & 1][tid * kTilePre + k],
                                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
                            else
                                xv = from_bits<double>(wait_bits(a, c, poll_bits(a.x + c), wd, idle));
                            acc = acc + vs[k] * xv;
                        }
                    }
                    for (int e = e0 + kTilePre + 1; e < e0 + len; ++e) {
                        const int c = a.col[e];
                        const double xv =
                            c < r1 ? xs[c - r0] : from_bits<double>(wait_bits(a, c, poll_bits(a.x + c), wd, idle));
                        acc = acc + a.val[e] * xv;
                    }
                    const double xi = (r.b - acc) / dbl(r.w2.x, r.w2.y);
                    xs[row - r0] = xi;
                    publish(a.x + row, xi);
                }
                lds_barrier();
            }
        }
    }
}

// b in position order (per solve)
__global__ void k_gather_b(const double* __restrict__ b, const int* __restrict__ row, double* __restrict__ bt,
                           int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        bt[i] = b[row[i]];
}

__global__ void k_fill_sentinel(double* x, int64_t n) {
    const uint64_t s = Bits<double>::sentinel;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        reinterpret_cast<uint64_t*>(x)[i] = s;
}

__global__ void k_reset_ticket_tile(int* ctrl) {
    if (threadIdx.x == 0) ctrl[0] = 0;
}

template <typename X>
int upload(X** dst, const std::vector<X>& src) {
    SPT_HIP(hipMalloc((void**)dst, std::max<size_t>(src.size(), 1) * sizeof(X)));
    if (!src.empty()) SPT_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(X), hipMemcpyHostToDevice));
    return SPTRSV_OK;
}

// Host layout: rows of each tile sorted by (level, row); steps = runs of one level cut at
// kTileThreads rows; records and the tile-ordered CSR follow the positions; the remote list
// of a step holds its record entries whose column lies in a later tile (up to cap of them go
// through the poller -- ring bit set -- the rest are read by their own thread).
int build_tiles(sptrsv_plan_s* P, int T) {
    const int n = P->n;
    const int64_t nnz = P->nnz;
    std::vector<int> rp((size_t)n + 1), col((size_t)nnz), lev((size_t)n);
    std::vector<double> val((size_t)nnz);
    SPT_HIP(hipStreamSynchronize(P->stream));
    SPT_HIP(hipMemcpy(rp.data(), P->rowptr, rp.size() * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(col.data(), P->col, col.size() * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(val.data(), P->val, val.size() * sizeof(double), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(lev.data(), P->levels, lev.size() * sizeof(int), hipMemcpyDeviceToHost));
    const int nt = (n + T - 1) / T;
    std::vector<int> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    std::vector<int> sptr((size_t)nt + 1, 0), step, stile;
    for (int t = 0; t < nt; ++t) {
        const int lo = t * T, hi = std::min(n, lo + T);
        std::stable_sort(order.begin() + lo, order.begin() + hi, [&](int x, int y) { return lev[x] < lev[y]; });
        sptr[t] = (int)step.size();
        for (int p = lo; p < hi;) {
            int q = p + 1;
            while (q < hi && lev[order[q]] == lev[order[p]] && q - p < kTileThreads) ++q;
            step.push_back(p);
            stile.push_back(t);
            p = q;
        }
    }
    sptr[nt] = (int)step.size();
    step.push_back(n);
    const int nsteps = (int)stile.size();
    std::vector<TileRec> rec((size_t)n);
    std::vector<int> tcol((size_t)nnz);
    std::vector<double> tval((size_t)nnz);
    int64_t pos = 0;
    for (int p = 0; p < n; ++p) {
        const int r = order[p];
        const int len = rp[r + 1] - rp[r];
        TileRec& R = rec[(size_t)p];
        R = TileRec{};
        R.row = r;
        R.len = len;
        R.e0 = (int)pos;
        std::copy(col.begin() + rp[r], col.begin() + rp[r + 1], tcol.begin() + pos);
        std::copy(val.begin() + rp[r], val.begin() + rp[r + 1], tval.begin() + pos);
        pos += len;
        R.diag = val[rp[r]];
        int* cs[kTilePre] = {&R.c0, &R.c1, &R.c2, &R.c3};
        double* vs[kTilePre] = {&R.v0, &R.v1, &R.v2, &R.v3};
        for (int k = 0; k < kTilePre; ++k)
            if (k + 1 < len) {
                *cs[k] = col[rp[r] + 1 + k];
                *vs[k] = val[rp[r] + 1 + k];
            }
    }
    auto rec_col = [](const TileRec& R, int k) { return k == 0 ? R.c0 : k == 1 ? R.c1 : k == 2 ? R.c2 : R.c3; };
    int maxr = 1;
    for (int q = 0; q < nsteps; ++q) {
        const int hi = std::min(n, (stile[q] + 1) * T);
        int cnt = 0;
        for (int p = step[q]; p < step[q + 1]; ++p)
            for (int k = 0; k < kTilePre; ++k) cnt += k + 1 < rec[(size_t)p].len && rec_col(rec[(size_t)p], k) >= hi;
        maxr = std::max(maxr, cnt);
    }
    const int cap = std::min(kWave * kPollE, (maxr + kWave - 1) / kWave * kWave);
    std::vector<int2> rlist((size_t)nsteps * cap, int2{-1, 0});
    for (int q = 0; q < nsteps; ++q) {
        const int hi = std::min(n, (stile[q] + 1) * T);
        int cnt = 0;
        for (int p = step[q]; p < step[q + 1]; ++p) {
            TileRec& R = rec[(size_t)p];
            const int len = R.len & kLenMask;
            for (int k = 0; k < kTilePre; ++k) {
                const int c = rec_col(R, k);
                if (k + 1 < len && c >= hi && cnt < cap) {
                    rlist[(size_t)q * cap + cnt] = int2{c, (p - step[q]) * kTilePre + k};
                    R.len |= 1 << (24 + k);
                    ++cnt;
                }
            }
        }
    }
    tile_free(P);
    SPT_CHECK(upload(&P->tile_sptr, sptr));
    SPT_CHECK(upload(&P->tile_step, step));
    TileRec* d_rec = nullptr;
    SPT_CHECK(upload(&d_rec, rec));
    P->tile_rec = d_rec;
    SPT_CHECK(upload(&P->tile_col, tcol));
    SPT_CHECK(upload(&P->tile_val, tval));
    SPT_CHECK(upload(&P->tile_row, order));
    SPT_HIP(hipMalloc((void**)&P->tile_bt, std::max<size_t>((size_t)n, 1) * sizeof(double)));
    int2* d_list = nullptr;
    SPT_CHECK(upload(&d_list, rlist));
    P->tile_rlist = d_list;
    P->tile_cap = cap;
    P->n_tiles = nt;
    P->tile_rows = T;
    return SPTRSV_OK;
}

}  // namespace

void tile_free(sptrsv_plan_s* p) {
    void* bufs[] = {p->tile_sptr, p->tile_step, p->tile_rec, p->tile_col, p->tile_val, p->tile_rlist,
                    p->tile_row, p->tile_bt};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    p->tile_sptr = p->tile_step = p->tile_col = p->tile_row = nullptr;
    p->tile_bt = nullptr;
    p->tile_rec = nullptr;
    p->tile_val = nullptr;
    p->tile_rlist = nullptr;
    p->n_tiles = 0;
    p->tile_rows = 0;
    p->tile_cap = 0;
}

int tile_solve(sptrsv_plan_s* P, int precision, int rows, const double* d_b, double* d_x, int m, uint64_t limit) {
    if (m != 1) return fail(SPTRSV_ERR_UNSUPPORTED, "tile variant solves one right-hand side");
    if (precision != SPTRSV_FP64) return fail(SPTRSV_ERR_UNSUPPORTED, "tile variant is FP64");
    if (rows < 0 || rows > kTileMaxRows)
        return fail(SPTRSV_ERR_BAD_INPUT, "tile rows must be 1.." + std::to_string(kTileMaxRows) + " (0: default)");
    const int T = std::min(rows > 0 ? rows : kTileMaxRows, std::max(P->n, 1));
    if (P->tile_rows != T) SPT_CHECK(build_tiles(P, T));
    const int64_t n = P->n;
    TileRec* rec = static_cast<TileRec*>(P->tile_rec);
    const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
    k_fill_sentinel<<<g, 256, 0, P->stream>>>(d_x, n);
    k_gather_b<<<g, 256, 0, P->stream>>>(d_b, P->tile_row, P->tile_bt, n);
    k_reset_ticket_tile<<<1, 64, 0, P->stream>>>(P->ticket);
    TileArgs a{P->n, T, P->n_tiles, P->tile_cap, rec, P->tile_bt, P->tile_col, P->tile_val,
               static_cast<const int2*>(P->tile_rlist), d_x, P->ticket, limit};
    hipEvent_t stop;
    SPT_CHECK(begin_timed(P, &stop));
    k_tile<<<P->n_tiles, kTileThreads + kWave, (size_t)T * sizeof(double), P->stream>>>(a, P->tile_sptr,
                                                                                          P->tile_step);
    end_timed(P, stop);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

}  // namespace sptrsv
