// levelset.hip -- the bulk-synchronous level-set solve (the reference's barrier_upper,
// src/core.cpp:337-376): a level's rows run only after every row of the previous level.
//
// The rows of a level are independent; the reference gives each thread a share of a level and
// puts an OpenMP barrier between levels.  Here every level's rows (consecutive slots of the
// level-ordered CSR, so a level's entries are one contiguous slice) are cut into chunks of at
// most max(16, 256 / G) rows (G lanes per row over the right-hand sides: at most four
// (row, column) items per thread, so a level costs about one round of dependency loads whatever
// m is; smaller chunks only add contention on the ticket and level counters: on the 128^3
// factor at m = 32, 9.4 ms with 16 rows, 10.5 with 4, 25 with 256) and kChunkEntries entries,
// built once per plan and chunk size, and each chunk is one workgroup.  Workgroups take chunks by ticket in level order; a chunk waits until the
// previous level's completion counter has counted all of that level's chunks (held by
// workgroups with smaller tickets, which are running: no co-residency is needed, and no
// cooperative launch), stages its column indices and values in LDS with coalesced loads,
// solves its rows from LDS (G lanes per row over the right-hand sides) with x stored and read
// by device-scope stores and loads, and counts itself done once its stores are performed.  Rows are summed in
// ascending column order and divided by the diagonal, the serial solver's arithmetic
// (src/core.cpp:18-32): x is bitwise equal to it.  A row longer than a chunk holds is read
// from the global CSR directly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "common.h"
#include "plan.h"

namespace sptrsv {

namespace {

constexpr int kLT = 256;             // threads per workgroup
constexpr int kMinChunkRows = 16;    // fewer rows per chunk: more chunks contend on the counters
constexpr int kChunkEntries = 3072;  // staged entries: 36 KiB of LDS (int + double each)
constexpr int kTicketSlot = 32;      // this solve's ticket counter in the plan's control words

template <typename T>
struct LevelSetArgs {
    int m;
    const int* __restrict__ ck_level;  // n_levels + 1 offsets into the chunk list
    const int* __restrict__ ck_slot;   // chunks + 1: first level-ordered slot of each chunk
    const int* __restrict__ ck_lev;    // chunks: level of each chunk
    const int* __restrict__ rowptr;    // level-ordered CSR (diagonal first)
    const int* __restrict__ col;
    const T* __restrict__ val;
    const int* __restrict__ perm;      // slot -> row
    const T* __restrict__ b;
    T* x;
    int* done;                         // n_levels: chunks of each level finished (zeroed per solve)
    int* ticket;
    int* status;
    uint64_t limit;
};

template <typename T, int G>
__global__ __launch_bounds__(kLT) void k_levelset(LevelSetArgs<T> a) {
#pragma clang fp contract(off)
    __shared__ int scol[kChunkEntries];
    __shared__ T sval[kChunkEntries];
    __shared__ int s_c;
    const int tid = threadIdx.x;
    if (tid == 0) s_c = atomicAdd(a.ticket, 1);
    __syncthreads();
    const int c = s_c;
    const int l = a.ck_lev[c];
    const int s0 = a.ck_slot[c], s1 = a.ck_slot[c + 1];
    const int e0 = a.rowptr[s0], e1 = a.rowptr[s1];
    const bool staged = e1 - e0 <= kChunkEntries;
    // the chunk's entries do not depend on the previous level: staged while it finishes
    if (staged)
        for (int e = e0 + tid; e < e1; e += kLT) {
            scol[e - e0] = a.col[e];
            sval[e - e0] = a.val[e];
        }
    if (tid == 0 && l > 0) {
        const int need = a.ck_level[l] - a.ck_level[l - 1];
        Watchdog wd;
        wd.arm(a.status, a.limit);
        int idle = 0;
        while (__hip_atomic_load(&a.done[l - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
            __builtin_amdgcn_s_sleep(1);
            if ((++idle & 63) == 0 && wd.expired()) break;
        }
    }
    __syncthreads();
    const size_t mm = (size_t)a.m;
    for (int t = tid; t < (s1 - s0) * G; t += kLT) {
        const int slot = s0 + t / G;
        const int rs = a.rowptr[slot], re = a.rowptr[slot + 1];
        const int row = a.perm[slot];
        const T diag = staged ? sval[rs - e0] : a.val[rs];
        for (int k = t % G; k < a.m; k += G) {
            T acc = T(0);
            for (int q = rs + 1; q < re; ++q) {
                const int cq = staged ? scol[q - e0] : a.col[q];
                const T vq = staged ? sval[q - e0] : a.val[q];
                // x of earlier levels: device-scope loads (the writers' stores went past
                // their XCD's L2 before they counted their chunk done)
                acc = acc + vq * from_bits<T>(poll_bits(a.x + (size_t)cq * mm + k));
            }
            const size_t at = (size_t)row * mm + k;
            publish(a.x + at, (a.b[at] - acc) / diag);
        }
    }
    __builtin_amdgcn_s_waitcnt(0);  // every wave: this chunk's device-scope stores performed
    __syncthreads();
    if (tid == 0) atomicAdd(&a.done[l], 1);
}

template <typename T>
void launch_levelset(int G, unsigned grid, const LevelSetArgs<T>& a, hipStream_t s) {
    switch (G) {
        case 1: k_levelset<T, 1><<<grid, kLT, 0, s>>>(a); break;
        case 2: k_levelset<T, 2><<<grid, kLT, 0, s>>>(a); break;
        case 4: k_levelset<T, 4><<<grid, kLT, 0, s>>>(a); break;
        case 8: k_levelset<T, 8><<<grid, kLT, 0, s>>>(a); break;
        case 16: k_levelset<T, 16><<<grid, kLT, 0, s>>>(a); break;
        case 32: k_levelset<T, 32><<<grid, kLT, 0, s>>>(a); break;
        default: k_levelset<T, 64><<<grid, kLT, 0, s>>>(a); break;
    }
}

// Chunks of every level: consecutive slots, at most `rows` rows and kChunkEntries entries (a
// longer row is a chunk of its own, read from the global CSR), into chunk set `si`.
int build_chunks(sptrsv_plan_s* P, int si, int rows) {
    const int n = P->n, L = P->n_levels;
    std::vector<int> lrp((size_t)n + 1);
    SPT_HIP(hipStreamSynchronize(P->stream));
    SPT_HIP(hipMemcpy(lrp.data(), P->lrowptr, lrp.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> ck_level(1, 0), ck_slot, ck_lev;
    for (int l = 0; l < L; ++l) {
        const int lo = P->h_level_ptr[(size_t)l], hi = P->h_level_ptr[(size_t)l + 1];
        int s = lo;
        while (s < hi) {
            int e = s + 1;
            while (e < hi && e - s < rows && lrp[(size_t)e + 1] - lrp[(size_t)s] <= kChunkEntries) ++e;
            ck_slot.push_back(s);
            ck_lev.push_back(l);
            s = e;
        }
        ck_level.push_back((int)ck_slot.size());
    }
    ck_slot.push_back(n);
    auto up = [](int** dst, const std::vector<int>& v) {
        SPT_HIP(hipMalloc((void**)dst, std::max<size_t>(v.size(), 1) * sizeof(int)));
        if (!v.empty()) SPT_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
        return SPTRSV_OK;
    };
    SPT_CHECK(up(&P->ls_ck_level[si], ck_level));
    SPT_CHECK(up(&P->ls_ck_slot[si], ck_slot));
    SPT_CHECK(up(&P->ls_ck_lev[si], ck_lev));
    if (!P->ls_done) SPT_HIP(hipMalloc((void**)&P->ls_done, std::max<size_t>((size_t)L, 1) * sizeof(int)));
    P->ls_chunks[si] = (int)ck_lev.size();
    return SPTRSV_OK;
}

}  // namespace

void levelset_free(sptrsv_plan_s* p) {
    for (int si = 0; si < sptrsv_plan_s::kLsSizes; ++si) {
        for (int* q : {p->ls_ck_level[si], p->ls_ck_slot[si], p->ls_ck_lev[si]})
            if (q) (void)hipFree(q);
        p->ls_ck_level[si] = p->ls_ck_slot[si] = p->ls_ck_lev[si] = nullptr;
        p->ls_chunks[si] = 0;
    }
    if (p->ls_done) (void)hipFree(p->ls_done);
    p->ls_done = nullptr;
}

template <typename T>
int levelset_solve(sptrsv_plan_s* P, const T* lval, const T* d_b, T* d_x, int m, uint64_t limit) {
    if (P->n == 0) return SPTRSV_OK;
    int G = 1, si = 0;
    while (G < 64 && G < m) {
        G *= 2;
        ++si;
    }
    if (!P->ls_ck_level[si]) SPT_CHECK(build_chunks(P, si, std::max(kMinChunkRows, kLT / G)));
    // the level counters and the ticket start at zero, in stream order before the solve
    SPT_CHECK(zero_ints_async(P->ls_done, P->n_levels, P->stream));
    SPT_CHECK(zero_ints_async(P->ticket + kTicketSlot, 1, P->stream));
    LevelSetArgs<T> a{m, P->ls_ck_level[si], P->ls_ck_slot[si], P->ls_ck_lev[si], P->lrowptr, P->lcol, lval,
                      P->level_rows, d_b, d_x, P->ls_done, P->ticket + kTicketSlot, P->ticket + 1, limit};
    hipEvent_t stop;
    SPT_CHECK(begin_timed(P, &stop));
    launch_levelset<T>(G, (unsigned)P->ls_chunks[si], a, P->stream);
    end_timed(P, stop);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

template int levelset_solve<double>(sptrsv_plan_s*, const double*, const double*, double*, int, uint64_t);
template int levelset_solve<float>(sptrsv_plan_s*, const float*, const float*, float*, int, uint64_t);

}  // namespace sptrsv
