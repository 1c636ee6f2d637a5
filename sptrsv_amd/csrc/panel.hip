// panel.hip -- brick panels (FP64, many right-hand sides, natural-order grid factors).
//
// With m right-hand sides the pull kernels read each dependency's m values across the fabric:
// a value published by another CU can only be read with a device-scope load, which misses the
// XCD's L2, so every row reads its three neighbours' m values from memory (about three times
// the size of x per solve) on top of b and x themselves.  Here a row's in-brick dependencies
// are read from LDS instead, and only the cells on the faces of the three upstream bricks come
// from memory (a quarter to three eighths of the brick, once each).
//
// Work item = (brick, panel of 16 right-hand-side columns), one 256-thread workgroup each,
// taken by ticket in brick dependency order (panels of one brick are consecutive tickets):
//   1. every load that does not depend on other work items is issued first: b of the lane's
//      cells into registers, the cells' records (diagonal, entries at +1, +nx, +nx*ny) and
//      entry masks into LDS;
//   2. three lanes wait for the epoch flags of this panel of the upstream bricks, then the
//      workgroup copies their face cells (the halo) from x into LDS with device-scope loads;
//   3. steps: at step s the cells at distance s from the brick's far corner are solved, four
//      lanes per cell, four columns per lane; dependencies come from the previous step's LDS
//      buffer or from the halo, b from the lane's registers (overwritten by x); the steps touch
//      LDS and registers only, so no step waits on memory; an LDS-only barrier separates them;
//   4. the brick's x leaves in one burst of stores, device-scope for the face cells that a
//      downstream brick reads; once every wave's stores are performed (vmcnt 0) one lane stores
//      the (brick, panel) flag = this solve's epoch (with at most four panels, the face cells are
//      stored and flagged first and the other cells after the flag).  Flags carry the epoch,
//      so nothing is reset between solves and x needs no sentinel fill.
// Each cell sums its products in ascending column order (+1, +nx, +nx*ny), only over the
// entries the row has, and divides by the diagonal: the reference serial solver's arithmetic
// (src/core.cpp:18-32), so x is bitwise equal to it.
#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>
#include <vector>

#include "common.h"
#include "panel_layout.h"
#include "plan.h"

namespace sptrsv {

namespace {

constexpr int kPanel = 16;  // right-hand-side columns per work item
constexpr int kPT = 256;    // threads: 64 cells per step x 4 lanes, 4 columns per lane
constexpr int kTicketSlot = 16;  // this kernel's ticket counter in the plan's control words
constexpr int kSplitPanels = 4;  // face stores flagged first up to this many panels (m <= 64)

struct PanelArgs {
    int nx, ny, nz;
    int bx, by, bz;
    int gx, gy, gz;
    int nsteps, nhalo, npanel, m;
    int step_ptr_len, ncells;
    unsigned base;                    // tickets taken before this solve
    const int* __restrict__ blist;    // bricks in dependency order
    const int* __restrict__ step_ptr;
    const uint64_t* __restrict__ cells;
    const double* __restrict__ rec;   // 4 per row: diagonal, entry at +1, +nx, +nx*ny
    const unsigned char* __restrict__ mask;  // bit d: the row has its entry in direction d
    const double* __restrict__ b;
    double* x;
    int* flag;                        // per (brick, panel): epoch of the last solve that finished it
    int epoch;
    unsigned* ticket;
    int* status;
    uint64_t limit;
};

__device__ __forceinline__ void barrier_lds_only() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); outstanding stores keep going
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// f(integral_constant<int, 0>) ... f(integral_constant<int, N - 1>), in order
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (N > 0) {
        static_for<N - 1>(f);
        f(std::integral_constant<int, N - 1>{});
    }
}

// Brick shape known at compile time: the cells of step s are positions [start(s), start(s+1))
// of the step-ordered cell table (the order panel_layout() builds), and a thread's cells are
// the positions congruent to its cell lane modulo 64, so every register index below is static.
template <int BX, int BY, int BZ>
struct Brick {
    static constexpr int NS = BX + BY + BZ - 2;
    static constexpr int NC = BX * BY * BZ;
    static constexpr int NJ = (NC + 63) / 64;  // cells per cell lane
    static constexpr int NH = BY * BZ + BX * BZ + (BZ > 1 ? BX * BY : 0);
    static constexpr int width(int s) {
        int w = 0;
        for (int k = 0; k < BZ; ++k)
            for (int j = 0; j < BY; ++j)
                for (int i = 0; i < BX; ++i) w += (BX - 1 - i) + (BY - 1 - j) + (BZ - 1 - k) == s;
        return w;
    }
    static constexpr int start(int s) {
        int p = 0;
        for (int t = 0; t < s; ++t) p += width(t);
        return p;
    }
};

// Value of lane K of this lane's group of four (DPP quad broadcast; every lane of the group
// must be active).
template <int K>
__device__ __forceinline__ double quad_get(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)u, K * 0x55, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), K * 0x55, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// VEC: m a multiple of 4 and b 16-byte aligned, so every lane's 4 columns are all valid or
// all beyond m, and 16-byte aligned: two 16-byte loads per cell instead of four 8-byte ones.  At most 128 VGPRs: four workgroups per CU.
template <int BX, int BY, int BZ, bool VEC, bool SPLIT>
__global__ __launch_bounds__(kPT, 4) void k_panel(PanelArgs a) {
#pragma clang fp contract(off)
    using B = Brick<BX, BY, BZ>;
    constexpr int NJ = B::NJ;
    static_assert(B::NC <= kPanelMaxCells && B::NH <= kPanelMaxHalo, "brick tables");
    static_assert(B::NC <= 64 * 8 && 4 * NJ <= 32, "cell lane registers");
    // 40 KiB in all for 8x8x8 bricks: four workgroups share a CU's 160 KiB
    __shared__ uint64_t scell[B::NC];
    __shared__ double xbuf[2][kPanelMaxWidth * kPanel];
    __shared__ double halo[B::NH * kPanel];
    const int tid = threadIdx.x;
    const int kr = tid >> 2, q = tid & 3;
    int* s_t = reinterpret_cast<int*>(&xbuf[0][0]);  // the ticket's LDS word until step 0
    // the cell table's loads and the ticket's atomic in flight together
    uint64_t cw[(B::NC + kPT - 1) / kPT];
#pragma unroll
    for (int e = 0; e < (B::NC + kPT - 1) / kPT; ++e) cw[e] = a.cells[min(tid + e * kPT, B::NC - 1)];
    if (tid == 0) *s_t = (int)(atomicAdd(a.ticket, 1u) - a.base);
#pragma unroll
    for (int e = 0; e < (B::NC + kPT - 1) / kPT; ++e)
        if (tid + e * kPT < B::NC) scell[tid + e * kPT] = cw[e];
    __syncthreads();
    const int t = *s_t;
    const int bl = t / a.npanel, p = t % a.npanel;
    const int lin = a.blist[bl];
    const int I = lin % a.gx, J = (lin / a.gx) % a.gy, K = lin / (a.gx * a.gy);
    const int c0 = p * kPanel + 4 * q;  // this lane's first column
    const size_t mm = (size_t)a.m;

    // 1. every load of the work item that does not depend on other work items: for each cell of
    //    this cell lane its b (4 columns), record value q (the group of four lanes holds the
    //    whole record), and 4 bits in one word: its entry mask and whether it lies in the grid
    auto cell_row = [&](uint64_t w, bool& in) {
        const int gi = I * BX + (int)(w & 31), gj = J * BY + (int)((w >> 5) & 31), gk = K * BZ + (int)((w >> 10) & 31);
        in = gi < a.nx && gj < a.ny && gk < a.nz;
        return in ? gi + a.nx * (gj + a.ny * gk) : 0;
    };
    double bv[NJ][4];
    double rv[NJ];
    int mkp = 0;
    // (static_for, not a loop: every index into the register arrays is a constant, so they stay
    // in registers)
    static_for<NJ>([&](auto jj) __attribute__((always_inline)) {
        constexpr int j = decltype(jj)::value;
        const int c = j * 64 + kr;
        bool in;
        const int r0 = cell_row(scell[c < B::NC ? c : 0], in);
        const bool act = c < B::NC && in;
        const int r = act ? r0 : 0;
        rv[j] = a.rec[4 * (size_t)r + q];
        const int mk = a.mask[r];
        mkp |= (act ? mk | 8 : 0) << (4 * j);
        if constexpr (VEC) {
            const double2* bp = reinterpret_cast<const double2*>(a.b + (size_t)r * mm + min(c0, a.m - 4));
            const double2 b01 = bp[0], b23 = bp[1];
            bv[j][0] = b01.x;
            bv[j][1] = b01.y;
            bv[j][2] = b23.x;
            bv[j][3] = b23.y;
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) bv[j][u] = a.b[(size_t)r * mm + min(c0 + u, a.m - 1)];
        }
    });

    // 2. upstream bricks of this panel, then their face cells
    if (tid < 3) {
        int up = -1;
        if (tid == 0 && I + 1 < a.gx) up = lin + 1;
        if (tid == 1 && J + 1 < a.gy) up = lin + a.gx;
        if (tid == 2 && K + 1 < a.gz) up = lin + a.gx * a.gy;
        if (up >= 0) {
            Watchdog wd;
            wd.arm(a.status, a.limit);
            const int* f = &a.flag[(size_t)up * a.npanel + p];
            int idle = 0;
            while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != a.epoch) {
                // back off: a long wait (the start of the solve) polls the fabric less often
                if (idle < 8) __builtin_amdgcn_s_sleep(2);
                else if (idle < 32) __builtin_amdgcn_s_sleep(8);
                else __builtin_amdgcn_s_sleep(24);
                if ((++idle & 63) == 0 && wd.expired()) break;
            }
        }
    }
    __syncthreads();
    {
        // sixteen lanes per face cell, one column each: every load instruction reads four whole
        // 128-byte rows of x
        constexpr int kRows = kPT / kPanel;  // face cells per pass
        constexpr int kPass = (B::NH + kRows - 1) / kRows;
        constexpr int fx = BY * BZ, fy = BX * BZ;
        const int col = tid & (kPanel - 1);
        uint64_t hv[kPass];
#pragma unroll
        for (int j = 0; j < kPass; ++j) {
            const int h = tid / kPanel + j * kRows;
            int gi, gj, gk;
            if (h < fx) {
                gi = (I + 1) * BX;
                gj = J * BY + h % BY;
                gk = K * BZ + h / BY;
            } else if (h < fx + fy) {
                gi = I * BX + (h - fx) % BX;
                gj = (J + 1) * BY;
                gk = K * BZ + (h - fx) / BX;
            } else {
                gi = I * BX + (h - fx - fy) % BX;
                gj = J * BY + (h - fx - fy) / BX;
                gk = (K + 1) * BZ;
            }
            // a cell outside the grid (or a column beyond m) is never read: load x[0] instead
            const bool ok = h < B::NH && gi < a.nx && gj < a.ny && gk < a.nz && p * kPanel + col < a.m;
            hv[j] = poll_bits(a.x + (ok ? (size_t)(gi + a.nx * (gj + a.ny * gk)) * mm + p * kPanel + col : 0));
        }
#pragma unroll
        for (int j = 0; j < kPass; ++j) {
            const int h = tid / kPanel + j * kRows;
            if (h < B::NH) halo[h * kPanel + col] = from_bits<double>(hv[j]);
        }
    }
    __syncthreads();

    // 3. steps, LDS and registers only: the cell of step s at position c of the table belongs
    //    to cell lane c % 64; its b (then x), record and mask are entry c / 64 of that lane's
    //    registers, one of two static entries per step
    static_for<B::NS>([&](auto step) __attribute__((always_inline)) {
        constexpr int s = decltype(step)::value;
        constexpr int ps = B::start(s), pe = B::start(s + 1);
        constexpr int j0 = ps / 64, j1 = (pe - 1) / 64;
        const int c = (j0 == j1 || kr >= ps % 64) ? j0 * 64 + kr : j1 * 64 + kr;
        const bool hi = j0 != j1 && c >= j1 * 64;
        const int bits = (mkp >> (4 * (hi ? j1 : j0))) & 15;
        const bool act = c >= ps && c < pe && (bits & 8);
        const int msk = bits & 7;
        // the cell's record from its group of four lanes (all active here)
        const double rj = hi ? rv[j1] : rv[j0];
        const double dg = quad_get<0>(rj), v[3] = {quad_get<1>(rj), quad_get<2>(rj), quad_get<3>(rj)};
        double* cur = xbuf[s & 1];
        const double* prev = xbuf[(s + 1) & 1];
        if (act) {
            const uint64_t w = scell[c];
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if ((msk >> d) & 1) {
                    const int sl = (int)((w >> (16 + 9 * d)) & 511);
                    const double* src = sl < kPanelSlotHalo ? prev + sl * kPanel : halo + (sl - kPanelSlotHalo) * kPanel;
#pragma unroll
                    for (int u = 0; u < 4; ++u) acc[u] = acc[u] + v[d] * src[4 * q + u];
                }
            }
            // one reciprocal for the lane's four columns; the four quotients side by side, then
            // (rarely) the ordinary division for a column outside the reciprocal's range
            const double rdg = div_recip(dg);
            double num[4], xv[4];
            bool plain[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                num[u] = (hi ? bv[j1][u] : bv[j0][u]) - acc[u];
                xv[u] = div_by_plain(num[u], dg, rdg, plain[u]);
            }
            if (!(plain[0] && plain[1] && plain[2] && plain[3])) {
                double dd = dg;
                asm volatile("" : "+v"(dd));  // the division stays inside the branch
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (!plain[u]) xv[u] = num[u] / dd;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                cur[(c - ps) * kPanel + 4 * q + u] = xv[u];
                // both candidate registers written, each with a select: a store through a
                // selected address would put the whole register array in scratch memory
                if constexpr (j0 == j1) {
                    bv[j0][u] = xv[u];
                } else {
                    bv[j0][u] = hi ? bv[j0][u] : xv[u];
                    bv[j1][u] = hi ? xv[u] : bv[j1][u];
                }
            }
        }
        barrier_lds_only();
    });

    // 4. x of the brick, through LDS so that sixteen lanes hold one cell's sixteen columns and
    //    every store instruction writes four whole 128-byte rows: the cells that a downstream
    //    brick reads (local coordinate 0 in a direction with a downstream brick) with
    //    device-scope stores, the others with ordinary ones (read only after the kernel); the
    //    flag once every wave's stores are performed.  With SPLIT the face cells go first and
    //    the flag follows their stores alone; the other cells are staged through LDS again and
    //    stored after the flag, so that their stores do not hold back the downstream bricks
    //    (worth it with few panels in flight; with many, the second staging costs more).
    //    parts: 1 = face cells, 2 = the others, 3 = both
    auto put = [&](const int parts) __attribute__((always_inline)) {
        const int wv = tid / 64, ln = tid & 63;
        double* stage = &xbuf[0][0] + wv * 16 * kPanel;  // this wave's 16 cells (2 KiB)
        const int col = ln & (kPanel - 1);
        static_for<NJ>([&](auto jj) __attribute__((always_inline)) {
            constexpr int j = decltype(jj)::value;
#pragma unroll
            for (int u = 0; u < 4; ++u) stage[(kr & 15) * kPanel + 4 * q + u] = bv[j][u];
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): the wave's LDS writes are done
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int cl = ln / kPanel + 4 * i;  // the wave's cell (0..15) this lane stores
                const int c = j * 64 + wv * 16 + cl;
                const double v = stage[cl * kPanel + col];
                const uint64_t w = scell[c < B::NC ? c : 0];
                bool in;
                const int r = cell_row(w, in);
                if (c < B::NC && in && p * kPanel + col < a.m) {
                    double* xp = a.x + (size_t)r * mm + p * kPanel + col;
                    const bool face =
                        ((w & 31) == 0 && I > 0) || (((w >> 5) & 31) == 0 && J > 0) || (((w >> 10) & 31) == 0 && K > 0);
                    if (face && (parts & 1)) publish(xp, v);
                    else if (!face && (parts & 2)) *xp = v;
                }
            }
            __builtin_amdgcn_s_waitcnt(0xC07F);  // the reads are done before the next slot's writes
            __builtin_amdgcn_wave_barrier();
        });
    };
    put(SPLIT ? 1 : 3);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&a.flag[(size_t)lin * a.npanel + p], a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (SPLIT) put(2);
}

// Records of a grid factor: per row its diagonal and the entries at +1, +nx, +nx*ny (0 where
// absent, with the mask bit clear).  Rows hold only those offsets (detected at plan creation).
__global__ void k_panel_records(int n, int nx, int nxy, const int* __restrict__ rowptr, const int* __restrict__ col,
                                const double* __restrict__ val, double* rec, unsigned char* mask) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        double v[4] = {val[rowptr[r]], 0.0, 0.0, 0.0};
        int msk = 0;
        for (int e = rowptr[r] + 1; e < rowptr[r + 1]; ++e) {
            const int d = col[e] - r;
            const int k = d == 1 ? 0 : d == nx ? 1 : 2;
            v[1 + k] = val[e];
            msk |= 1 << k;
        }
        for (int u = 0; u < 4; ++u) rec[4 * (size_t)r + u] = v[u];
        mask[r] = (unsigned char)msk;
    }
}

template <typename X>
int upload_p(X** dst, const std::vector<X>& src) {
    SPT_HIP(hipMalloc((void**)dst, std::max<size_t>(src.size(), 1) * sizeof(X)));
    if (!src.empty()) SPT_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(X), hipMemcpyHostToDevice));
    return SPTRSV_OK;
}

int build_panel(sptrsv_plan_s* P) {
    PanelLayout L;
    if (!P->grid || !panel_layout(P->grid_dims[0], P->grid_dims[1], P->grid_dims[2], &L))
        return fail(SPTRSV_ERR_UNSUPPORTED, "panel variant: the factor is not a natural-order 2D/3D grid factor");
    // the kernels' compile-time step boundaries must be the table's
    auto same_steps = [&](auto brick) {
        using Bk = decltype(brick);
        if (L.nsteps != Bk::NS || L.ncells != Bk::NC) return false;
        for (int s = 0; s <= L.nsteps; ++s)
            if (L.step_ptr[(size_t)s] != Bk::start(s)) return false;
        return true;
    };
    if (!same_steps(Brick<8, 8, 8>{}) && !same_steps(Brick<16, 16, 1>{}))
        return fail(SPTRSV_ERR_UNSUPPORTED, "panel variant: no kernel for this brick shape");
    panel_free(P);
    SPT_CHECK(upload_p(&P->pn_blist, L.blist));
    SPT_CHECK(upload_p(&P->pn_step_ptr, L.step_ptr));
    SPT_CHECK(upload_p(&P->pn_cells, L.cells));
    SPT_HIP(hipMalloc((void**)&P->pn_rec, std::max<size_t>((size_t)P->n * 4, 1) * sizeof(double)));
    SPT_HIP(hipMalloc((void**)&P->pn_mask, std::max<size_t>((size_t)P->n, 1)));
    if (P->n > 0) {
        const int g = (int)std::min<int64_t>(((int64_t)P->n + 255) / 256, 65536);
        k_panel_records<<<g, 256, 0, P->stream>>>(P->n, (int)L.nx, (int)(L.nx * L.ny), P->rowptr, P->col, P->val,
                                                  P->pn_rec, P->pn_mask);
        SPT_HIP(hipGetLastError());
    }
    P->pn_bx = L.bx;
    P->pn_by = L.by;
    P->pn_bz = L.bz;
    P->pn_gx = L.gx;
    P->pn_gy = L.gy;
    P->pn_gz = L.gz;
    P->pn_nsteps = L.nsteps;
    P->pn_nhalo = L.nhalo;
    P->pn_ncells = L.ncells;
    P->pn_nbricks = (int)L.blist.size();
    P->pn_built = true;
    return SPTRSV_OK;
}

int ensure_panel_flags(sptrsv_plan_s* P, int npanel) {
    const size_t need = (size_t)P->pn_nbricks * npanel;
    if (need <= P->pn_flag_elems) return SPTRSV_OK;
    SPT_HIP(hipStreamSynchronize(P->stream));
    if (P->pn_flag) SPT_HIP(hipFree(P->pn_flag));
    P->pn_flag = nullptr;
    P->pn_flag_elems = 0;
    SPT_HIP(hipMalloc((void**)&P->pn_flag, need * sizeof(int)));
    // below every epoch in use (epochs start at 1); zeroed in stream order, before the solve
    SPT_CHECK(zero_ints_async(P->pn_flag, (int64_t)need, P->stream));
    P->pn_flag_elems = need;
    return SPTRSV_OK;
}

}  // namespace

void panel_free(sptrsv_plan_s* p) {
    void* bufs[] = {p->pn_blist, p->pn_step_ptr, p->pn_cells, p->pn_rec, p->pn_mask, p->pn_flag};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    p->pn_blist = p->pn_step_ptr = nullptr;
    p->pn_cells = nullptr;
    p->pn_rec = nullptr;
    p->pn_mask = nullptr;
    p->pn_flag = nullptr;
    p->pn_flag_elems = 0;
    p->pn_built = false;
}

bool panel_supports(sptrsv_plan_s* P) {
    PanelLayout L;
    return P->grid && P->n > 0 && panel_layout(P->grid_dims[0], P->grid_dims[1], P->grid_dims[2], &L);
}

int panel_solve(sptrsv_plan_s* P, int precision, const double* d_b, double* d_x, int m, uint64_t limit) {
    if (precision != SPTRSV_FP64) return fail(SPTRSV_ERR_UNSUPPORTED, "panel variant is FP64");
    if (P->n == 0) return SPTRSV_OK;
    if (!P->pn_built) SPT_CHECK(build_panel(P));
    const int npanel = (m + kPanel - 1) / kPanel;
    const int64_t items = (int64_t)P->pn_nbricks * npanel;
    if (items >= ((int64_t)1 << 31)) return fail(SPTRSV_ERR_UNSUPPORTED, "panel variant: too many work items");
    SPT_CHECK(ensure_panel_flags(P, npanel));
    if (++P->pn_epoch <= 0) P->pn_epoch = 1;  // flags hold epochs of earlier solves, all different
    PanelArgs a{(int)P->grid_dims[0], (int)P->grid_dims[1], (int)P->grid_dims[2], P->pn_bx, P->pn_by, P->pn_bz,
                P->pn_gx, P->pn_gy, P->pn_gz, P->pn_nsteps, P->pn_nhalo, npanel, m, P->pn_nsteps + 1, P->pn_ncells, P->pn_tickets, P->pn_blist,
                P->pn_step_ptr, P->pn_cells, P->pn_rec, P->pn_mask, d_b, d_x, P->pn_flag, P->pn_epoch,
                reinterpret_cast<unsigned*>(P->ticket + kTicketSlot), P->ticket + 1, limit};
    const bool b888 = P->pn_bx == 8 && P->pn_by == 8 && P->pn_bz == 8;
    const bool b16 = P->pn_bx == 16 && P->pn_by == 16 && P->pn_bz == 1;
    if (!b888 && !b16) return fail(SPTRSV_ERR_UNSUPPORTED, "panel variant: no kernel for this brick shape");
    // 16-byte loads and stores where every lane's 4 columns are 16-byte aligned
    const bool vec = m % 4 == 0 && ((uintptr_t)d_b & 15) == 0;
    hipEvent_t stop;
    SPT_CHECK(begin_timed(P, &stop));
    const unsigned g = (unsigned)items;
    const bool split = npanel <= kSplitPanels;
    if (b888) {
        if (split) {
            if (vec) k_panel<8, 8, 8, true, true><<<g, kPT, 0, P->stream>>>(a);
            else k_panel<8, 8, 8, false, true><<<g, kPT, 0, P->stream>>>(a);
        } else {
            if (vec) k_panel<8, 8, 8, true, false><<<g, kPT, 0, P->stream>>>(a);
            else k_panel<8, 8, 8, false, false><<<g, kPT, 0, P->stream>>>(a);
        }
    } else {
        if (split) {
            if (vec) k_panel<16, 16, 1, true, true><<<g, kPT, 0, P->stream>>>(a);
            else k_panel<16, 16, 1, false, true><<<g, kPT, 0, P->stream>>>(a);
        } else {
            if (vec) k_panel<16, 16, 1, true, false><<<g, kPT, 0, P->stream>>>(a);
            else k_panel<16, 16, 1, false, false><<<g, kPT, 0, P->stream>>>(a);
        }
    }
    end_timed(P, stop);
    SPT_HIP(hipGetLastError());
    P->pn_tickets += (unsigned)items;  // every workgroup takes exactly one ticket
    return SPTRSV_OK;
}

}  // namespace sptrsv
