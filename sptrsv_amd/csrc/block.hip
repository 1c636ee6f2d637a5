// block.hip -- block pull (FP64; m == 1, or up to 4 right-hand sides on grid factors): blocks
// of rows solved inside one workgroup, the block's internal dependencies handed over through
// LDS.
//
// In the row and level pull kernels every dependency is a device-scope store-then-poll round
// trip (1-2 us), so the solve costs about one round trip per level of the dependency DAG.
// Here the rows are cut into blocks, the blocks are taken in a topological order (every
// dependency of a block lies in the same block or in an earlier one), and each block is
// solved by W compute waves and one poller wave of one workgroup:
//   * a compute wave holds a row per lane and reads every value it needs from LDS: the
//     block's own rows from their slots as soon as their lanes store them, the values of
//     earlier blocks from slots the poller fills; it never waits on global memory;
//   * the poller wave holds the block's list of external values (at most kMaxExt) and polls
//     the self-publishing x for them until all have arrived, copying each into its LDS slot;
//   * a row's value goes to x at its step only if another block reads it; the others are
//     stored after the wave's last step, off the dependency chain.
// Grid factors use 512-row blocks (8 compute waves), other factors 64-row blocks, four of them
// per workgroup under one ticket.
// A dependency inside the block costs an LDS round trip; only a dependency on another block
// costs a device round trip (the poller's), and a compute wave never waits on a memory round
// trip of its own.
//
// Blocks: for a factor whose off-diagonal offsets are those of a natural-order 2D or 3D grid
// ({1, nx} or {1, nx, nx*ny}), 32x16 or 8x8x8 bricks ordered by brick anti-diagonal (a path
// through a 128^3 grid crosses ~45 bricks instead of 381 levels), rows inside a brick by level;
// otherwise (or if the brick order is not topological) bands of 16 levels with rows in
// descending order inside a band, cut into chunks of the block size.  A block is cut short
// where its external list would exceed kMaxExt.
//
// Each lane sums its row's products in ascending column order and divides by the diagonal,
// the reference serial solver's arithmetic (src/core.cpp:18-32), so x is bitwise equal to it.
// Workgroups hold tickets in block order and wait only on earlier blocks, whose workgroups
// are already resident, which guarantees progress; every wait is bounded by the watchdog.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "common.h"
#include "plan.h"

namespace sptrsv {

// Grid of a natural-order stencil factor: the distinct off-diagonal offsets are {1, nx} or
// {1, nx, nx*ny} with n a multiple of the largest.  Returns false for anything else.
bool detect_grid(int n, const int* rp, const int* col, int64_t* dims) {
    int seen[3], k = 0;  // distinct offsets so far (one scan, at most three kept)
    for (int r = 0; r < n; ++r)
        for (int e = rp[r] + 1; e < rp[r + 1]; ++e) {
            const int d = col[e] - r;
            if ((k > 0 && d == seen[0]) || (k > 1 && d == seen[1]) || (k > 2 && d == seen[2])) continue;
            if (k == 3) return false;
            seen[k++] = d;
        }
    std::vector<int> o(seen, seen + k);
    std::sort(o.begin(), o.end());
    if (o.empty() || o[0] != 1) return false;
    if (o.size() == 2 && o[1] > 1 && n % o[1] == 0) {
        dims[0] = o[1];
        dims[1] = n / o[1];
        dims[2] = 1;
        return true;
    }
    if (o.size() == 3 && o[1] > 1 && o[2] > o[1] && o[2] % o[1] == 0 && n % o[2] == 0) {
        dims[0] = o[1];
        dims[1] = o[2] / o[1];
        dims[2] = n / o[2];
        return true;
    }
    return false;
}

namespace {

constexpr int kBlockRows = 64;
constexpr int kExtPerLane = 8;
constexpr int kMaxExt = kWave * kExtPerLane;  // external values per block
constexpr int kBand = 16;                      // levels per band of the generic block order
constexpr int kNB = 4;                         // 64-row blocks per workgroup (one ticket each)
constexpr int kBigRows = 8 * kWave;            // rows of a grid factor's block: 8 compute waves
constexpr int kBlockMaxM = 4;                  // right-hand sides per lane (grid factors)
constexpr int kPolled = 1 << 30;               // step word flag: another block reads this row

struct BlockArgs {
    int nblk;
    int m;                          // right-hand sides (row stride of b and x)
    const int* __restrict__ bptr;   // nblk + 1 block boundaries (positions)
    const int* __restrict__ eptr;   // nblk + 1 offsets into ecol
    const int* __restrict__ ecol;   // external values of each block (rows of earlier blocks)
    const int* __restrict__ bperm;  // position -> row
    const int* __restrict__ brp;    // CSR in position order (diagonal first)
    const int* __restrict__ bslot;  // LDS slot of each entry's x value (< 0: -(column + 1), read
                                    // from x: entries of a row beyond the block's kMaxExt)
    const int* __restrict__ bstep;  // step of each position inside its block (| kPolled)
    const double* __restrict__ bval;
    const double* __restrict__ b;
    double* x;
    int* ctrl;
    uint64_t limit;
};

__device__ __forceinline__ uint64_t lds_get(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_put(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// A workgroup solves NB consecutive blocks of W * 64 rows: waves 0..NB*W-1 are the compute
// waves (W per block, 64 rows each), waves NB*W..NB*W+NB-1 the blocks' pollers, and one
// ticket covers all NB blocks (a single ticket counter serialises its atomics at ~12 ns each,
// which bounds how many blocks per microsecond can start).
// M > 1: M right-hand sides per lane, the block's LDS slots held once per column (column c's
// slots at c * (R + XM), so consecutive lanes read consecutive banks whatever M); the poller
// then keeps at most XJ = 4 external rows per lane (256 per block).
template <int UNROLL, bool SHORT, int NB, int W, int M = 1>
__global__ __launch_bounds__((W + 1) * NB * kWave) void k_block(BlockArgs a) {
#pragma clang fp contract(off)
    using U = uint64_t;
    constexpr U S = Bits<double>::sentinel;
    constexpr int R = W * kWave;                     // rows per block
    constexpr int XJ = M == 1 ? kExtPerLane : 4;     // external rows per poller lane
    constexpr int XM = XJ * kWave;                   // external rows per block
    __shared__ U slots[NB * (R + XM) * M];
    __shared__ int s_ticket;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x) / kWave;  // wave index, uniform
    const bool poller = w >= NB * W;
    const int blk = poller ? w - NB * W : w / W;
    const int sw = poller ? 0 : w % W;  // compute wave of its block
    U* slot = slots + blk * (R + XM) * M;
    const size_t mm = M == 1 ? 1 : (size_t)a.m;
    if (threadIdx.x == 0) s_ticket = atomicAdd(&a.ctrl[0], 1);
#pragma unroll
    for (int k = 0; k < M; ++k)
        if (!poller) lds_put(&slot[k * (R + XM) + sw * kWave + lane], S);
#pragma unroll
    for (int j = 0; j < XJ; ++j)
#pragma unroll
        for (int k = 0; k < M; ++k)
            if (poller) lds_put(&slot[k * (R + XM) + R + lane + kWave * j], S);
    __syncthreads();
    const int t = s_ticket * NB + blk;
    if (t >= a.nblk) return;  // a whole wave; no barrier follows
    Watchdog wd;
    wd.arm(&a.ctrl[1], a.limit);
    int idle = 0, streak = 0;
    if (poller) {
        const int e0 = a.eptr[t], ne = a.eptr[t + 1] - e0;
        int ec[XJ];
        bool need[XJ][M];
        U v[XJ][M];
#pragma unroll
        for (int j = 0; j < XJ; ++j) {
            const int k = lane + kWave * j;
            ec[j] = a.ecol[e0 + k];  // ecol is padded by kMaxExt entries: always in bounds
#pragma unroll
            for (int c = 0; c < M; ++c) {
                need[j][c] = k < ne && (M == 1 || c < a.m);
                v[j][c] = S;
            }
        }
        wait_vmem();
        while (true) {
#pragma unroll
            for (int j = 0; j < XJ; ++j)
#pragma unroll
                for (int c = 0; c < M; ++c)
                    if (need[j][c]) v[j][c] = poll_bits(a.x + (size_t)ec[j] * mm + c);
            bool got = false, left = false;
#pragma unroll
            for (int j = 0; j < XJ; ++j) {
#pragma unroll
                for (int c = 0; c < M; ++c) {
                    if (need[j][c] && v[j][c] != S) {
                        lds_put(&slot[c * (R + XM) + R + lane + kWave * j], v[j][c]);
                        need[j][c] = false;
                        got = true;
                    }
                    left |= need[j][c];
                }
            }
            // wake the compute wave, which sleeps while the values of its step are missing
            if (__any(got)) asm volatile("s_waitcnt lgkmcnt(0)\n\ts_wakeup" ::: "memory");
            if (!__any(left)) break;
            if (__any(got)) streak = 0;
            if (!__any(got)) {
                ++streak;
                if (streak < 2) __builtin_amdgcn_s_sleep(1);
                else if (streak < 8) __builtin_amdgcn_s_sleep(4);
                else __builtin_amdgcn_s_sleep(12);
                if ((++idle & 63) == 0 && wd.expired()) break;
            }
        }
        return;
    }
    // compute wave: rows in steps; a row's step is one more than the largest step among its
    // dependencies inside the block (0 without any), so at step s every in-block value it
    // reads belongs to an earlier step: stored already by this wave, or by another compute wave
    // of the block that may still be behind.  The lanes of step s wait in LDS for whatever is
    // still missing (another wave's rows, external values).
    const int p0 = a.bptr[t], p1 = a.bptr[t + 1];
    const int pos = p0 + sw * kWave + lane;
    const bool valid = pos < p1;
    int row = 0, step = -1, p = 0, end = 0;
    bool polled = false;  // another block reads this row from x: publish it at its step
    double xk[M];         // otherwise its values, written once the wave's steps are done
    double diag = 1.0, bv[M];
    int sl[UNROLL];
    double v[UNROLL];
#pragma unroll
    for (int c = 0; c < M; ++c) bv[c] = 0.0;
    auto load_window = [&](int q) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            sl[u] = 0;
            v[u] = 0.0;
            if (q + u < end) {
                sl[u] = a.bslot[q + u];
                v[u] = a.bval[q + u];
            }
        }
    };
    if (valid) {
        row = a.bperm[pos];
        step = a.bstep[pos] & (kPolled - 1);
        polled = a.bstep[pos] >= kPolled;
        const int s = a.brp[pos];
        diag = a.bval[s];
        p = s + 1;
        end = a.brp[pos + 1];
#pragma unroll
        for (int c = 0; c < M; ++c)
            if (M == 1 || c < a.m) bv[c] = a.b[(size_t)row * mm + c];
        load_window(p);
    }
    wait_vmem();
    int nsteps = step + 1;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) nsteps = max(nsteps, __shfl_xor(nsteps, o));
    bool dead = false;
    // one entry's values (M of them, a.m used) from its slot: in LDS, or (slot < 0) from x
    auto read = [&](int u, U (&xv)[UNROLL][M], bool direct) {
#pragma unroll
        for (int c = 0; c < M; ++c)
            xv[u][c] = (direct && sl[u] < 0) ? poll_bits(a.x + (size_t)(-sl[u] - 1) * mm + c)
                                             : lds_get(&slot[c * (R + XM) + sl[u]]);
    };
    auto ready = [&](int u, int cnt, const U (&xv)[UNROLL][M]) {
        bool ok = true;
#pragma unroll
        for (int c = 0; c < M; ++c) ok &= u >= cnt || (M > 1 && c >= a.m) || xv[u][c] != S;
        return ok;
    };
    for (int st = 0; st < nsteps && !dead; ++st) {
        const bool mine = step == st;
        double acc[M];
#pragma unroll
        for (int c = 0; c < M; ++c) acc[c] = 0.0;
        if constexpr (SHORT) {
            // every row fits one window: read the step's values, wait in LDS for missing
            // external ones, then the ordered sums
            const int cnt = end - p;
            U xv[UNROLL][M];
            while (true) {
                bool ok = true;
                if (mine) {
#pragma unroll
                    for (int u = 0; u < UNROLL; ++u) read(u, xv, false);  // all reads first,
#pragma unroll
                    for (int u = 0; u < UNROLL; ++u) ok &= ready(u, cnt, xv);  // then the checks
                }
                if (__all(ok)) break;
                __builtin_amdgcn_s_sleep(10);
                if ((++idle & 255) == 0 && wd.expired()) {
                    dead = true;
                    break;
                }
            }
            if (mine) {
#pragma unroll
                for (int c = 0; c < M; ++c)
#pragma unroll
                    for (int u = 0; u < UNROLL; ++u) {
                        const double t2 = acc[c] + v[u] * __builtin_bit_cast(double, xv[u][c]);
                        acc[c] = u < cnt ? t2 : acc[c];
                    }
            }
        } else {
            for (int q = p;; q += UNROLL) {
                const bool active = mine && q < end;
                if (!__any(active)) break;
                if (q != p) {  // rows longer than one window: next window from memory
                    if (active) load_window(q);
                    wait_vmem();
                }
                const int cnt = min(UNROLL, end - q);
                U xv[UNROLL][M];
                while (true) {
                    bool ok = true;
                    if (active) {
#pragma unroll
                        for (int u = 0; u < UNROLL; ++u) read(u, xv, true);
#pragma unroll
                        for (int u = 0; u < UNROLL; ++u) ok &= ready(u, cnt, xv);
                    }
                    if (__all(ok)) break;
                    __builtin_amdgcn_s_sleep(10);
                    if ((++idle & 255) == 0 && wd.expired()) {
                        dead = true;
                        break;
                    }
                }
                if (dead) break;
                if (active) {
#pragma unroll
                    for (int c = 0; c < M; ++c)
#pragma unroll
                        for (int u = 0; u < UNROLL; ++u) {
                            const double t2 = acc[c] + v[u] * __builtin_bit_cast(double, xv[u][c]);
                            acc[c] = u < cnt ? t2 : acc[c];
                        }
                }
            }
        }
        if (mine && !dead) {
#pragma unroll
            for (int c = 0; c < M; ++c) {
                if (M == 1 || c < a.m) {
                    const double xi = (bv[c] - acc[c]) / diag;
                    lds_put(&slot[c * (R + XM) + sw * kWave + lane], __builtin_bit_cast(U, xi));
                    if (polled) publish(a.x + (size_t)row * mm + c, xi);
                    xk[c] = xi;
                }
            }
        }
        // the block's other compute waves may sleep on this step's values
        if (W > 1 && __any(mine)) asm volatile("s_waitcnt lgkmcnt(0)\n\ts_wakeup" ::: "memory");
    }
    // rows no other block reads leave the dependency chain: their scattered stores come after
    // the wave's last step instead of inside every step
    if (valid && !polled && !dead) {
#pragma unroll
        for (int c = 0; c < M; ++c)
            if (M == 1 || c < a.m) publish(a.x + (size_t)row * mm + c, xk[c]);
    }
}

// x := sentinel, and the ticket counter reset (one launch before the solve instead of two)
__global__ void k_fill_sentinel_b(double* x, int64_t n, int* ctrl) {
    const uint64_t s = Bits<double>::sentinel;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctrl[0] = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        reinterpret_cast<uint64_t*>(x)[i] = s;
}

template <typename X>
int upload_b(X** dst, const std::vector<X>& src) {
    SPT_HIP(hipMalloc((void**)dst, std::max<size_t>(src.size(), 1) * sizeof(X)));
    if (!src.empty()) SPT_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(X), hipMemcpyHostToDevice));
    return SPTRSV_OK;
}

// Block order (position -> row) and block boundaries.  `kind` reports which order was used:
// 1 = grid bricks, 2 = level bands.
void block_order(int n, const std::vector<int>& rp, const std::vector<int>& col, const std::vector<int>& lev,
                 int R, std::vector<int>& order, std::vector<int>& bptr, int* kind) {
    int64_t d[3];
    order.assign((size_t)n, 0);
    bptr.clear();
    if (R == kBigRows && detect_grid(n, rp.data(), col.data(), d)) {
        // bricks of 512 rows: 8x8x8 in 3D, 32x16 in 2D
        const int64_t bx = d[2] > 1 ? 8 : 32;
        const int64_t by = d[2] > 1 ? 8 : 16;
        const int64_t bz = d[2] > 1 ? 8 : 1;
        const int64_t gx = (d[0] + bx - 1) / bx, gy = (d[1] + by - 1) / by, gz = (d[2] + bz - 1) / bz;
        // bricks by descending anti-diagonal (a brick's dependencies lie in bricks with larger
        // coordinates), then descending brick index; rows inside a brick by level, then
        // descending: each compute wave then holds a run of consecutive levels, so most of its
        // rows' in-brick dependencies are its own rows (by plane, every step would wait on the
        // wave of the neighbouring plane)
        std::vector<int64_t> key((size_t)n);
        for (int r = 0; r < n; ++r) {
            const int64_t i = r % d[0], j = (r / d[0]) % d[1], k = r / (d[0] * d[1]);
            const int64_t I = i / bx, J = j / by, K = k / bz;
            const int64_t diagk = (gx - 1 - I) + (gy - 1 - J) + (gz - 1 - K);
            const int64_t bidx = (gx * gy * gz - 1) - (I + gx * (J + gy * K));
            key[(size_t)r] = diagk * (gx * gy * gz) + bidx;
        }
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            if (key[(size_t)x] != key[(size_t)y]) return key[(size_t)x] < key[(size_t)y];
            if (lev[(size_t)x] != lev[(size_t)y]) return lev[(size_t)x] < lev[(size_t)y];
            return x > y;
        });
        for (int pos = 0; pos < n; ++pos)
            if (pos == 0 || key[(size_t)order[(size_t)pos]] != key[(size_t)order[(size_t)pos - 1]] ||
                pos - bptr.back() == R)
                bptr.push_back(pos);
        bptr.push_back(n);
        // topological check: every dependency in the same or an earlier block
        std::vector<int> blk((size_t)n);
        for (size_t b = 0; b + 1 < bptr.size(); ++b)
            for (int pos = bptr[b]; pos < bptr[b + 1]; ++pos) blk[(size_t)order[(size_t)pos]] = (int)b;
        bool ok = true;
        for (int r = 0; r < n && ok; ++r)
            for (int e = rp[r] + 1; e < rp[r + 1]; ++e)
                if (blk[(size_t)col[e]] > blk[(size_t)r]) {
                    ok = false;
                    break;
                }
        if (ok) {
            *kind = 1;
            return;
        }
        bptr.clear();
    }
    // level bands: a dependency has a lower level, so it lies in an earlier band or, in the
    // same band, at a larger row index -- an earlier position
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const int bx = (lev[(size_t)x] - 1) / kBand, by = (lev[(size_t)y] - 1) / kBand;
        return bx != by ? bx < by : x > y;
    });
    for (int pos = 0; pos < n; pos += R) bptr.push_back(pos);
    bptr.push_back(n);
    *kind = 2;
}

int build_blocks(sptrsv_plan_s* P) {
    const int n = P->n;
    const int64_t nnz = P->nnz;
    std::vector<int> rp((size_t)n + 1), col((size_t)nnz), lev((size_t)n);
    std::vector<double> val((size_t)nnz);
    SPT_HIP(hipStreamSynchronize(P->stream));
    SPT_HIP(hipMemcpy(rp.data(), P->rowptr, rp.size() * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(col.data(), P->col, col.size() * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(val.data(), P->val, val.size() * sizeof(double), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(lev.data(), P->levels, lev.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> order, pre;
    int kind = 0;
    // grid factors: 8x8x8 bricks (32x16 in 2D) on 8 compute waves each; otherwise 64-row blocks
    const int R = P->grid ? kBigRows : kBlockRows;
    P->blk_rows = R;
    block_order(n, rp, col, lev, R, order, pre, &kind);
    std::vector<int> posof((size_t)n);
    for (int pos = 0; pos < n; ++pos) posof[(size_t)order[(size_t)pos]] = pos;
    // cut blocks where the external list would exceed kMaxExt; external lists in order of
    // first use.  A dependency sits at an earlier position, so it is either in the block
    // (position >= the block's start) or external.
    std::vector<int> bptr, eptr, ecol, stamp((size_t)n, -1), extidx((size_t)n, 0);
    std::vector<int> brp((size_t)n + 1), bslot((size_t)nnz), bstep((size_t)n, 0);
    std::vector<double> bval((size_t)nnz);
    int64_t e0 = 0;
    int nb = -1, start = 0, next_pre = 0;
    auto open_block = [&](int pos) {
        ++nb;
        start = pos;
        bptr.push_back(pos);
        eptr.push_back((int)ecol.size());
    };
    for (int pos = 0; pos < n; ++pos) {
        const int r = order[(size_t)pos];
        bool cut = pos == pre[(size_t)next_pre];
        if (cut) ++next_pre;
        if (!cut) {
            int fresh = 0;
            for (int e = rp[r] + 1; e < rp[r + 1]; ++e) {
                const int c = col[e];
                fresh += posof[(size_t)c] < start && stamp[(size_t)c] != nb;
            }
            cut = (int)ecol.size() - eptr.back() + fresh > kMaxExt;
        }
        if (cut) open_block(pos);
        // a row with more external values than a block holds opens a block of its own: its
        // entries beyond the first kMaxExt are read from x directly (slot -(column + 1))
        brp[(size_t)pos] = (int)e0;
        bslot[(size_t)e0] = 0;
        bval[(size_t)e0] = val[rp[r]];
        ++e0;
        for (int e = rp[r] + 1; e < rp[r + 1]; ++e, ++e0) {
            const int c = col[e];
            const int pc = posof[(size_t)c];
            if (pc >= start) {
                bslot[(size_t)e0] = pc - start;
                bstep[(size_t)pos] = std::max(bstep[(size_t)pos], bstep[(size_t)pc] + 1);
            } else if (stamp[(size_t)c] == nb) {
                bslot[(size_t)e0] = R + extidx[(size_t)c];
            } else if ((int)ecol.size() - eptr.back() < kMaxExt) {
                stamp[(size_t)c] = nb;
                extidx[(size_t)c] = (int)ecol.size() - eptr.back();
                ecol.push_back(c);
                bslot[(size_t)e0] = R + extidx[(size_t)c];
            } else {
                bslot[(size_t)e0] = -(c + 1);
            }
            bval[(size_t)e0] = val[e];
        }
    }
    brp[(size_t)n] = (int)e0;
    bptr.push_back(n);
    eptr.push_back((int)ecol.size());
    // rows another block reads from x (a poller's list or a direct read) are published at their
    // step; the rest once their wave is done (steps stay far below kPolled)
    {
        std::vector<char> read_by_other((size_t)n, 0);
        for (int c : ecol) read_by_other[(size_t)c] = 1;
        for (int64_t e = 0; e < nnz; ++e)
            if (bslot[(size_t)e] < 0) read_by_other[(size_t)(-bslot[(size_t)e] - 1)] = 1;
        for (int pos = 0; pos < n; ++pos)
            if (read_by_other[(size_t)order[(size_t)pos]]) bstep[(size_t)pos] |= kPolled;
    }
    ecol.resize(ecol.size() + kMaxExt, 0);  // the poller loads kMaxExt entries from every block's start
    block_free(P);
    SPT_CHECK(upload_b(&P->blk_ptr, bptr));
    SPT_CHECK(upload_b(&P->blk_eptr, eptr));
    SPT_CHECK(upload_b(&P->blk_ecol, ecol));
    SPT_CHECK(upload_b(&P->blk_perm, order));
    SPT_CHECK(upload_b(&P->blk_rp, brp));
    SPT_CHECK(upload_b(&P->blk_col, bslot));
    SPT_CHECK(upload_b(&P->blk_step, bstep));
    SPT_CHECK(upload_b(&P->blk_val, bval));
    P->n_blocks = (int)bptr.size() - 1;
    P->blk_kind = kind;
    P->blk_maxlen = 0;
    for (int r = 0; r < n; ++r) P->blk_maxlen = std::max(P->blk_maxlen, rp[r + 1] - rp[r] - 1);
    P->blk_maxext = 0;
    for (size_t b = 0; b + 1 < eptr.size(); ++b) P->blk_maxext = std::max(P->blk_maxext, eptr[b + 1] - eptr[b]);
    return SPTRSV_OK;
}

}  // namespace

void block_free(sptrsv_plan_s* p) {
    void* bufs[] = {p->blk_ptr, p->blk_eptr, p->blk_ecol, p->blk_perm, p->blk_rp, p->blk_col, p->blk_step, p->blk_val};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    p->blk_ptr = p->blk_eptr = p->blk_ecol = p->blk_perm = p->blk_rp = p->blk_col = p->blk_step = nullptr;
    p->blk_val = nullptr;
    p->n_blocks = 0;
    p->blk_kind = 0;
    p->blk_maxlen = 0;
}

// 512-row blocks (8 compute waves + a poller), m right-hand sides per lane.  They are built
// for grid factors only, whose rows hold at most 3 off-diagonals (three distinct offsets): one
// 3-entry window per row.
void launch_bricks(const BlockArgs& a, int m, unsigned grid, size_t lds, hipStream_t s) {
    if (m == 1) k_block<3, true, 1, 8, 1><<<grid, 9 * kWave, lds, s>>>(a);
    else if (m == 2) k_block<3, true, 1, 8, 2><<<grid, 9 * kWave, lds, s>>>(a);
    else if (m == 3) k_block<3, true, 1, 8, 3><<<grid, 9 * kWave, lds, s>>>(a);
    else k_block<3, true, 1, 8, 4><<<grid, 9 * kWave, lds, s>>>(a);
}

// Several right-hand sides per lane run on 512-row blocks with at most 256 external rows per
// block (each LDS slot then holds M values); builds the block order on first use.
bool block_supports(sptrsv_plan_s* P, int m) {
    if (m == 1) return true;
    if (m < 1 || m > kBlockMaxM || P->n == 0) return false;
    if (P->n_blocks == 0 && build_blocks(P) != SPTRSV_OK) return false;
    return P->blk_rows == kBigRows && P->blk_maxlen <= 3 && P->blk_maxext <= 4 * kWave;
}

int block_solve(sptrsv_plan_s* P, int precision, int occ, const double* d_b, double* d_x, int m, uint64_t limit) {
    if (precision != SPTRSV_FP64) return fail(SPTRSV_ERR_UNSUPPORTED, "block variant is FP64");
    if (m < 1 || m > kBlockMaxM)
        return fail(SPTRSV_ERR_UNSUPPORTED, "block variant solves 1.." + std::to_string(kBlockMaxM) + " right-hand sides");
    if (P->n == 0) return SPTRSV_OK;
    if (P->n_blocks == 0) SPT_CHECK(build_blocks(P));
    const bool big = P->blk_rows == kBigRows;
    if (big && P->blk_maxlen > 3)  // not built: grid factors have three distinct offsets at most
        return fail(SPTRSV_ERR_UNSUPPORTED, "block variant: 512-row blocks with rows of more than 3 off-diagonals");
    if (!block_supports(P, m))
        return fail(SPTRSV_ERR_UNSUPPORTED,
                    "block variant: several right-hand sides need the 512-row blocks of a grid factor with at most " +
                        std::to_string(4 * kWave) + " external rows per block");
    const int64_t nm = (int64_t)P->n * m;
    const unsigned g = (unsigned)std::min<int64_t>((nm + 255) / 256, 65536);
    k_fill_sentinel_b<<<g, 256, 0, P->stream>>>(d_x, nm, P->ticket);
    BlockArgs a{P->n_blocks, m, P->blk_ptr, P->blk_eptr, P->blk_ecol, P->blk_perm, P->blk_rp, P->blk_col, P->blk_step,
                P->blk_val, d_b, d_x, P->ticket, limit};
    // resident workgroups per CU: dynamic LDS sized so that at most that many share a CU (160 KiB
    // of LDS per CU).  64-row blocks: kNB blocks per workgroup, `occ` blocks per CU (default
    // 4); 512-row blocks: one per workgroup, `occ` per CU (default 3)
    const int nb = big ? 1 : kNB;
    const int M = m;  // one kernel instance per right-hand-side count
    const int per_cu = std::max(1, (occ > 0 ? std::min(occ, 32) : big ? 3 : 4) / nb);
    const size_t own = (size_t)nb * (P->blk_rows + (M == 1 ? kMaxExt : 4 * kWave)) * M * 8 + 512;
    const size_t lds = per_cu >= 8 || own * per_cu >= 163840 ? 0 : ((163840 / per_cu - own) & ~(size_t)255);
    const unsigned grid = (unsigned)((P->n_blocks + nb - 1) / nb);
    hipEvent_t stop;
    SPT_CHECK(begin_timed(P, &stop));
    if (big) {
        launch_bricks(a, m, grid, lds, P->stream);
    } else {
        if (P->blk_maxlen <= 4) k_block<4, true, kNB, 1><<<grid, 2 * kNB * kWave, lds, P->stream>>>(a);
        else k_block<4, false, kNB, 1><<<grid, 2 * kNB * kWave, lds, P->stream>>>(a);
    }
    end_timed(P, stop);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

}  // namespace sptrsv

using namespace sptrsv;

extern "C" int sptrsv_plan_get_blocks(sptrsv_plan_t P, int* n_blocks, int* kind, int* block_ptr, int* perm) {
    if (!P || !n_blocks || !kind) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    if (P->n > 0 && P->n_blocks == 0) SPT_CHECK(build_blocks(P));
    *n_blocks = P->n_blocks;
    *kind = P->blk_kind;
    SPT_HIP(hipStreamSynchronize(P->stream));
    if (block_ptr && P->n_blocks > 0)
        SPT_HIP(hipMemcpy(block_ptr, P->blk_ptr, ((size_t)P->n_blocks + 1) * sizeof(int), hipMemcpyDeviceToHost));
    if (perm && P->n > 0) SPT_HIP(hipMemcpy(perm, P->blk_perm, (size_t)P->n * sizeof(int), hipMemcpyDeviceToHost));
    return SPTRSV_OK;
}
