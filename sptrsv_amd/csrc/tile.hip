// tile.hip -- tiled sync-free SpTRSV (m == 1) for gfx950.
//
// A dependency that crosses the GPU through HBM costs a few microseconds (store, then a
// device-scope poll that misses the XCD's L2); one that stays inside a CU costs an LDS
// round trip.  This kernel therefore gives each 1024-thread workgroup a tile of up to 960
// consecutive rows (taken from the last rows down by workgroup tickets):
//   * the tile's CSR entries are staged into LDS once (coalesced), values + 16-bit slots;
//   * 15 compute waves own one row per lane and wait on LDS slots only: a slot holds either
//     a row of this tile (published by its lane) or a column outside the tile;
//   * wave 15 is a poller: it fetches the tile's outside columns from HBM (device-scope
//     polls of the self-publishing x values) into their LDS slots as they become ready.
// Each row sums its products in ascending column order and divides by the diagonal -- the
// reference serial solver's arithmetic (src/core.cpp:18-32) -- so FP64 x is bitwise equal to
// it.  A tile's outside columns all belong to earlier tiles (earlier tickets, already
// running), which guarantees progress.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "common.h"
#include "plan.h"

namespace sptrsv {
namespace {

constexpr int kTileBlock = 1024;
constexpr int kTileRows = 960;   // 15 compute waves
constexpr int kTileExt = 2048;   // outside columns per tile
constexpr int kTileNnz = 4096;   // staged entries per tile (incl. diagonals)
constexpr uint16_t kDiagSlot = 0xFFFF;

struct TileArgs {
    int n_tiles;
    const int* lo;
    const int* hi;
    const int* ext_ptr;
    const int* ext;
    const uint16_t* lidx;
    const int* rowptr;
    const double* val;
    const double* b;
    double* x;
    int* ctrl;
    uint64_t limit;
};

__device__ __forceinline__ uint64_t lds_poll(const double* p) {
    return __hip_atomic_load(reinterpret_cast<const uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_publish(double* p, uint64_t bits) {
    __hip_atomic_store(reinterpret_cast<uint64_t*>(p), bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__global__ __launch_bounds__(kTileBlock) void k_tile(TileArgs a) {
#pragma clang fp contract(off)
    __shared__ double xs[kTileRows + kTileExt];
    __shared__ double sval[kTileNnz];
    __shared__ uint16_t sidx[kTileNnz];
    __shared__ double sb[kTileRows];
    __shared__ int sext[kTileExt];
    const int tid = threadIdx.x;
    const int t = block_ticket(&a.ctrl[0]);
    if (t >= a.n_tiles) return;
    const int lo = a.lo[t], hi = a.hi[t];
    const int e0 = a.rowptr[lo], e1 = a.rowptr[hi];
    for (int i = tid; i < kTileRows + kTileExt; i += kTileBlock) reinterpret_cast<uint64_t*>(xs)[i] = kSentinel64;
    for (int q = tid; q < e1 - e0; q += kTileBlock) {
        sval[q] = a.val[e0 + q];
        sidx[q] = a.lidx[e0 + q];
    }
    // b of the tile's rows and the outside column ids, so the loops below never wait on a
    // global load (which on gfx950 would also wait for the wave's outstanding stores)
    for (int q = tid; q < hi - lo; q += kTileBlock) sb[q] = a.b[lo + q];
    {
        const int x0 = a.ext_ptr[t], ne = a.ext_ptr[t + 1] - x0;
        for (int q = tid; q < ne; q += kTileBlock) sext[q] = a.ext[x0 + q];
    }
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63;
    Watchdog wd;
    wd.arm(&a.ctrl[1], a.limit);
    int idle = 0;
    if (wave == kTileBlock / kWave - 1) {
        // poller: lane l owns outside columns l, l+64, ... (at most 32) and delivers each one
        // as soon as it is published, in any order; up to eight polls in flight per lane
        const int ne = a.ext_ptr[t + 1] - a.ext_ptr[t];
        const int nq = lane < ne ? (ne - lane + 63) / 64 : 0;
        uint32_t pend = nq >= 32 ? 0xFFFFFFFFu : ((1u << nq) - 1u);
        while (true) {
            bool progressed = false;
            if (pend) {
                uint32_t rest = pend;
                int qs[8];
                uint64_t bits[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    qs[u] = -1;
                    bits[u] = kSentinel64;
                    if (rest) {
                        qs[u] = __builtin_ctz(rest);
                        rest &= rest - 1;
                        bits[u] = poll_bits(a.x + sext[lane + 64 * qs[u]]);
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (qs[u] >= 0 && bits[u] != kSentinel64) {
                        lds_publish(&xs[kTileRows + lane + 64 * qs[u]], bits[u]);
                        pend &= ~(1u << qs[u]);
                        progressed = true;
                    }
                }
            }
            if (__all(pend == 0)) break;
            if (!__any(progressed)) {
                __builtin_amdgcn_s_sleep(1);
                if ((++idle & 63) == 0 && pend && wd.expired()) pend = 0;
            }
        }
        return;
    }
    const int rows = hi - lo;
    bool done = tid >= rows;
    int r = 0, p = 0, end = 0;
    double diag = 1.0, acc = 0.0, bb = 0.0;
    if (!done) {
        r = hi - 1 - tid;
        const int s = a.rowptr[r] - e0;
        end = a.rowptr[r + 1] - e0;
        diag = sval[s];
        p = s + 1;
        bb = sb[r - lo];
    }
    while (true) {
        bool progressed = false;
        if (!done) {
            while (p < end) {
                const uint64_t bits = lds_poll(&xs[sidx[p]]);
                if (bits == kSentinel64) break;
                acc = acc + sval[p] * from_bits<double>(bits);
                ++p;
                progressed = true;
            }
            if (p >= end) {
                const double xr = (bb - acc) / diag;
                lds_publish(&xs[r - lo], __builtin_bit_cast(uint64_t, xr));
                publish(a.x + r, xr);
                done = true;
                progressed = true;
            }
        }
        if (__all(done)) break;
        if (!__any(progressed)) {
            // an idle wave backs off so the waves on the dependency chain keep the SIMD and LDS
            __builtin_amdgcn_s_sleep(1);
            if ((++idle & 255) == 0 && !done && wd.expired()) done = true;
        }
    }
}

__global__ void k_reset_ticket_tile(int* ctrl) {
    if (threadIdx.x == 0) ctrl[0] = 0;
}

__global__ void k_fill_sentinel(double* x, int64_t count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < count; i += (int64_t)gridDim.x * blockDim.x) reinterpret_cast<uint64_t*>(x)[i] = kSentinel64;
}

template <typename X>
int upload(X** dst, const std::vector<X>& src) {
    SPT_HIP(hipMalloc((void**)dst, std::max<size_t>(src.size(), 1) * sizeof(X)));
    if (!src.empty()) SPT_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(X), hipMemcpyHostToDevice));
    return SPTRSV_OK;
}

// Greedy tiling from the last row down: a tile grows while it has <= kTileRows rows,
// <= kTileNnz entries and <= kTileExt distinct outside columns.
int build_tiles(sptrsv_plan_s* P) {
    const int n = P->n;
    const int64_t nnz = P->nnz;
    std::vector<int> rp((size_t)n + 1), col((size_t)nnz);
    SPT_HIP(hipStreamSynchronize(P->stream));
    SPT_HIP(hipMemcpy(rp.data(), P->rowptr, rp.size() * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(col.data(), P->col, col.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> tlo, thi, ext_ptr(1, 0), ext;
    std::vector<uint16_t> lidx((size_t)nnz, kDiagSlot);
    std::vector<int> stamp((size_t)n, -1), slot_of((size_t)n, -1), first_need((size_t)n, -1);
    int hi = n;
    while (hi > 0) {
        const int tnum = (int)tlo.size();
        int lo = hi;
        std::vector<int> cur;
        while (lo > 0 && hi - lo < kTileRows) {
            const int r = lo - 1;
            if (rp[hi] - rp[r] > kTileNnz) break;
            int add = 0;
            for (int p = rp[r] + 1; p < rp[r + 1]; ++p)
                if (col[p] >= hi && stamp[col[p]] != tnum) ++add;
            if ((int)cur.size() + add > kTileExt) break;
            for (int p = rp[r] + 1; p < rp[r + 1]; ++p)
                if (col[p] >= hi && stamp[col[p]] != tnum) {
                    stamp[col[p]] = tnum;
                    cur.push_back(col[p]);
                }
            lo = r;
        }
        if (lo == hi)
            return fail(SPTRSV_ERR_UNSUPPORTED, "row " + std::to_string(hi - 1) +
                                                    " is too long for the tile variant (> 4096 entries or > 2048 outside columns)");
        // order by first need: rows run from the top of the tile down, so a column read by a
        // higher row is needed earlier; slot k goes to poller lane k % 64 as its (k/64)-th item
        for (int r = hi - 1; r >= lo; --r)
            for (int p = rp[r] + 1; p < rp[r + 1]; ++p)
                if (col[p] >= hi && first_need[col[p]] < r) first_need[col[p]] = r;
        std::sort(cur.begin(), cur.end(), [&](int u, int v) {
            return first_need[u] != first_need[v] ? first_need[u] > first_need[v] : u > v;
        });
        for (int c : cur) first_need[c] = -1;
        for (size_t k = 0; k < cur.size(); ++k) slot_of[cur[k]] = (int)k;
        for (int r = lo; r < hi; ++r)
            for (int p = rp[r] + 1; p < rp[r + 1]; ++p)
                lidx[p] = (uint16_t)(col[p] < hi ? col[p] - lo : kTileRows + slot_of[col[p]]);
        tlo.push_back(lo);
        thi.push_back(hi);
        ext.insert(ext.end(), cur.begin(), cur.end());
        ext_ptr.push_back((int)ext.size());
        hi = lo;
    }
    tile_free(P);
    SPT_CHECK(upload(&P->tile_lo, tlo));
    SPT_CHECK(upload(&P->tile_hi, thi));
    SPT_CHECK(upload(&P->tile_ext_ptr, ext_ptr));
    SPT_CHECK(upload(&P->tile_ext, ext));
    SPT_CHECK(upload(&P->tile_lidx, lidx));
    P->n_tiles = (int)tlo.size();
    return SPTRSV_OK;
}

}  // namespace

void tile_free(sptrsv_plan_s* p) {
    void* bufs[] = {p->tile_lo, p->tile_hi, p->tile_ext_ptr, p->tile_ext, p->tile_lidx};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    p->tile_lo = p->tile_hi = p->tile_ext_ptr = p->tile_ext = nullptr;
    p->tile_lidx = nullptr;
    p->n_tiles = 0;
}

int tile_solve(sptrsv_plan_s* P, int precision, const double* d_b, double* d_x, int m, uint64_t limit) {
    if (m != 1) return fail(SPTRSV_ERR_UNSUPPORTED, "tile variant solves one right-hand side");
    if (precision != SPTRSV_FP64) return fail(SPTRSV_ERR_UNSUPPORTED, "tile variant is FP64");
    if (P->n_tiles == 0) SPT_CHECK(build_tiles(P));
    const int64_t n = P->n;
    k_fill_sentinel<<<(unsigned)std::min<int64_t>((n + 255) / 256, 65536), 256, 0, P->stream>>>(d_x, n);
    k_reset_ticket_tile<<<1, 64, 0, P->stream>>>(P->ticket);
    TileArgs a{P->n_tiles, P->tile_lo, P->tile_hi, P->tile_ext_ptr, P->tile_ext, P->tile_lidx,
               P->rowptr, P->val, d_b, d_x, P->ticket, limit};
    hipEvent_t stop;
    SPT_CHECK(begin_timed(P, &stop));
    k_tile<<<P->n_tiles, kTileBlock, 0, P->stream>>>(a);
    end_timed(P, stop);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

}  // namespace sptrsv
