// analysis.hip -- device-side analysis of an upper-triangular factor:
//   CSR -> CSC transpose (stable; diagonal last)          src/utils.cpp:128-159
//   in-degree per row (incl. diagonal)                     src/kernel_full.cl:40-52
//   1-based levels (dependency depth)                      src/core.cpp:244-270
//   level sets + level pointers (rows ascending per level) src/core.cpp:282-287
//   level-ordered CSR (rows stored level by level)         (no counterpart: feeds the
//                                                           level-scheduled sync-free kernel)
// All integer results are exact; the tests compare them bit for bit with the oracle.  Sorting
// and prefix sums are this library's own (radix.hip).
#include <hip/hip_runtime.h>

#include <chrono>
#include <string>
#include <vector>

#include "common.h"
#include "plan.h"

namespace sptrsv {

namespace {

__global__ void k_iota(int* out, int64_t count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < count; i += (int64_t)gridDim.x * blockDim.x) out[i] = (int)i;
}

__global__ void k_fill_int(int* out, int value, int64_t count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < count; i += (int64_t)gridDim.x * blockDim.x) out[i] = value;
}

// row id of every CSR position
__global__ void k_expand_rows(const int* __restrict__ rowptr, int n, int* __restrict__ row_of) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    for (; r < n; r += gridDim.x * blockDim.x)
        for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) row_of[p] = r;
}

// entries of every row: the in-degree the reference's analyser counts as a histogram over the
// CSC row indices (every entry of row r, the diagonal included, has row index r)
__global__ void k_row_lengths(const int* __restrict__ rowptr, int n, int* __restrict__ out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    for (; r < n; r += gridDim.x * blockDim.x) out[r] = rowptr[r + 1] - rowptr[r];
}

__global__ void k_gather_csc(const int* __restrict__ perm, const int* __restrict__ row_of,
                             const double* __restrict__ val, int64_t nnz, int* __restrict__ rowidx,
                             double* __restrict__ cval) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
        const int src = perm[k];
        rowidx[k] = row_of[src];
        cval[k] = val[src];
    }
}

// Tickets of the level sweep: 64 consecutive rows each, a wave's lane l on row hi - l.
// Default order: descending rows (ticket t = rows n - 1 - 64 t down).  On a natural-order grid
// factor, descending rows put about eight k-planes of 128^2 rows in flight, and each plane's
// lines depend on one another in a chain; there a ticket is a segment of 64 rows of one i-line
// (segment g = i / 64), and tickets go by their anti-diagonal distance
// (segs - 1 - g) + (ny - 1 - j) + (nz - 1 - k) from the last segment, so that every segment
// whose dependencies are known can be in flight and each distance costs one device round trip.
// That order is used only after a check that every column of every row lies in the row's own
// ticket or an earlier one (k_grid_order_check).
struct LevelTickets {
    int count;         // tickets
    int nx, segs;      // grid order: row length of a line and 64-row segments per line
    const int* item;   // grid order: line * segs + segment at each position (nullptr: descending rows)
};

__device__ __forceinline__ void ticket_rows(const LevelTickets& T, int n, int t, int* lo, int* hi) {
    if (T.item == nullptr) {
        *hi = n - 1 - 64 * t;
        *lo = max(*hi - 63, 0);
        return;
    }
    const int e = T.item[t], line = e / T.segs, seg = e - line * T.segs;
    *lo = line * T.nx + 64 * seg;
    *hi = min(*lo + 63, line * T.nx + T.nx - 1);
}

// flag[0] = 1 if some column of some row lies in a later ticket of the grid order
__global__ void k_grid_order_check(int n, const int* __restrict__ rowptr, const int* __restrict__ col, int nx,
                                   int segs, const int* __restrict__ pos_of, int* flag) {
    auto ticket = [&](int r) {
        const int line = r / nx, i = r - line * nx;
        return pos_of[line * segs + i / 64];
    };
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const int tr = ticket(r);
        for (int p = rowptr[r] + 1; p < rowptr[r + 1]; ++p)
            if (ticket(col[p]) > tr) {
                atomicOr(flag, 1);
                break;
            }
    }
}

// Level of every row by a sync-free max-plus sweep: waves take tickets (above) loop after loop
// (persistent: a finished wave takes the next ticket instead of making way for a new wave); a
// row's level is 1 + the deepest level among its off-diagonal columns (1 for a diagonal-only
// row), exactly the reference's rule (src/core.cpp:244-270).  level 0 marks "not yet known".
// Each lane walks its row from the far end, up to kLevelWin entries at a time.  Columns above
// the ticket's rows ("external") lie in earlier tickets: their levels are polled from memory,
// a window at a time.  Columns inside the ticket ("internal": consecutive rows of a grid line
// form a chain) are read from the wave's level slots in LDS: between two device rounds the wave
// runs LDS-only rounds until no lane progresses, so a chain inside a ticket costs LDS round
// trips, not device ones.  A wave waits only on rows of earlier tickets, held by running waves,
// and on its own lanes' rows (columns lie above their row): the sweep cannot deadlock.
constexpr int kLevelWin = 8;

__global__ __launch_bounds__(256) void k_levels(int n, LevelTickets T, const int* __restrict__ rowptr,
                                                const int* __restrict__ col, int* levels,
                                                int* ctrl, uint64_t limit) {
    __shared__ int s_lv[256];
    const int lane = threadIdx.x & (kWave - 1);
    volatile int* lv_own = s_lv + (threadIdx.x & ~(kWave - 1));  // this wave's 64 slots
    int mine = 0;
    Watchdog wd;
    wd.arm(&ctrl[1], limit);
    while (true) {
        const int t = wave_ticket(&ctrl[0]);
        if (t >= T.count) break;
        int lo, hi;
        ticket_rows(T, n, t, &lo, &hi);
        const int row = hi - lane;
        lv_own[lane] = 0;
        __builtin_amdgcn_wave_barrier();
        bool done = row < lo;
        int first = 0, p = 0, deepest = 0;
        bool has = false;
        if (!done) {
            first = rowptr[row] + 1;
            p = rowptr[row + 1];
            has = p > first;
        }
        // the current window: entries [p - cnt, p) with columns c[0] > c[1] > ... (largest
        // first); c[w0 ..] not yet taken.  External columns form a prefix of the window.
        int c[kLevelWin];
        int cnt = 0, w0 = 0;  // cnt 0: columns not read yet
        int idle = 0;
        auto load_window = [&]() {
            cnt = min(kLevelWin, p - first);
            w0 = 0;
#pragma unroll
            for (int j = 0; j < kLevelWin; ++j) c[j] = col[max(p - cnt, p - 1 - j)];
        };
        auto ext_left = [&]() {
            bool e = false;
#pragma unroll
            for (int j = 0; j < kLevelWin; ++j) e = e || (j >= w0 && j < cnt && c[j] > hi);
            return e;
        };
        auto consumed = [&](int mx) {
            deepest = mx;
            p -= cnt;
            cnt = 0;
            w0 = 0;
        };
        // device round: the window's external entries, polled from memory
        auto take_ext = [&]() {
            int lv[kLevelWin], ne = w0;
#pragma unroll
            for (int j = 0; j < kLevelWin; ++j) {
                const bool use = j >= w0 && j < cnt && c[j] > hi;
                lv[j] = use ? __hip_atomic_load(&levels[c[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 1;
                ne = use ? j + 1 : ne;
            }
            bool all = true;
            int mx = deepest;
#pragma unroll
            for (int j = 0; j < kLevelWin; ++j) {
                all = all && lv[j] != 0;
                mx = lv[j] > mx ? lv[j] : mx;
            }
            if (all) {
                if (ne >= cnt) consumed(mx);
                else {
                    deepest = mx;
                    w0 = ne;
                }
            }
            return all;
        };
        // LDS round: the rest of the window, all internal, from the wave's slots
        auto take_int = [&]() {
            bool all = true;
            int mx = deepest;
#pragma unroll
            for (int j = 0; j < kLevelWin; ++j) {
                const int v = (j >= w0 && j < cnt) ? lv_own[hi - c[j]] : 1;
                all = all && v != 0;
                mx = v > mx ? v : mx;
            }
            if (all) consumed(mx);
            return all;
        };
        // "simple" rows (the whole row in one window, no internal column but row + 1): their
        // level is max(E, level(row + 1) + 1) with E = 1 + the deepest external level, so along a
        // run of them level - lane is a running maximum of E - lane, which a segmented prefix
        // scan over the wave resolves at once (a grid line's chain in six shuffle steps)
        bool simple = false, link = false;
        if (!done) {
            const bool whole = p - first <= kLevelWin;
            if (has) load_window();
            int nint = 0;
            bool next_only = true;
#pragma unroll
            for (int j = 0; j < kLevelWin; ++j)
                if (j < cnt && c[j] <= hi) {
                    ++nint;
                    next_only = next_only && c[j] == row + 1;
                }
            simple = whole && (nint == 0 || (nint == 1 && next_only));
            link = simple && nint == 1;
        }
        auto finish_at = [&](int lev) {
            lv_own[lane] = lev;
            __hip_atomic_store(&levels[row], lev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            mine = lev > mine ? lev : mine;
            done = true;
        };
        auto finish = [&]() {
            const int lev = has ? 1 + deepest : 1;
            lv_own[lane] = lev;
            __hip_atomic_store(&levels[row], lev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            mine = lev > mine ? lev : mine;
            done = true;
        };
        while (true) {
            bool progressed = false;
            if (!done && p > first) {
                if (cnt == 0) load_window();
                if (ext_left()) progressed = take_ext();
            }
            // scan round: key = level - lane (known) or E - lane (simple, externals taken);
            // a lane starts a segment unless it links to lane - 1; lanes of unknown level poison
            // their segment
            if (__any(simple && !done)) {
                const bool ready = simple && !done && (p <= first || !ext_left());
                int key = done ? lv_own[lane] - lane : (has ? 1 + deepest : 1) - lane;
                int fl = (done || !ready || !link ? 1 : 0) | (!done && !ready ? 2 : 0);  // 1 start, 2 poison
#pragma unroll
                for (int d = 1; d < kWave; d <<= 1) {
                    const int k2 = __shfl_up(key, d, kWave), f2 = __shfl_up(fl, d, kWave);
                    if (lane >= d && !(fl & 1)) {
                        key = key > k2 ? key : k2;
                        fl |= f2;
                    }
                }
                if (ready && !(fl & 2)) {
                    finish_at(key + lane);
                    progressed = true;
                }
                __builtin_amdgcn_wave_barrier();
            }
            // LDS rounds until no lane progresses
            while (true) {
                bool lp = false;
                if (!done && p > first) {
                    if (cnt == 0) load_window();
                    if (!ext_left()) lp = take_int();
                }
                if (!done && p <= first) {
                    finish();
                    lp = true;
                }
                __builtin_amdgcn_wave_barrier();
                if (!__any(lp)) break;
                progressed = true;
            }
            if (__all(done)) break;
            if (!__any(progressed)) {
                __builtin_amdgcn_s_sleep(1);
                if ((++idle & 63) == 0 && !done && wd.expired()) done = true;
            }
        }
        if (__hip_atomic_load(&ctrl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;  // watchdog
    }
    // one atomic per wave for the maximum level
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const int o = __shfl_xor(mine, off, kWave);
        mine = o > mine ? o : mine;
    }
    if (lane == 0) atomicMax(&ctrl[2], mine);
}

__global__ void k_level_rowlen(const int* __restrict__ rowptr, const int* __restrict__ level_rows, int n,
                               int* __restrict__ lens) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    for (; k < n; k += gridDim.x * blockDim.x) {
        const int r = level_rows[k];
        lens[k + 1] = rowptr[r + 1] - rowptr[r];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) lens[0] = 0;
}

// copy row level_rows[k] into slot k of the level-ordered CSR; G lanes per row (G a power of two
// dividing the wave, about the mean row length)
__global__ void k_level_gather(const int* __restrict__ rowptr, const int* __restrict__ col,
                               const double* __restrict__ val, const int* __restrict__ level_rows,
                               const int* __restrict__ lrowptr, int n, int G, int* __restrict__ lcol,
                               double* __restrict__ lval) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int sub = (int)(t & (G - 1));
    const int64_t stride = (int64_t)gridDim.x * blockDim.x / G;
    for (int64_t k = t / G; k < n; k += stride) {
        const int r = level_rows[k];
        const int a = rowptr[r], len = rowptr[r + 1] - a, d = lrowptr[k];
        for (int e = sub; e < len; e += G) {
            lcol[d + e] = col[a + e];
            lval[d + e] = val[a + e];
        }
    }
}

__global__ void k_to_float(const double* __restrict__ in, float* __restrict__ out, int64_t count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < count; i += (int64_t)gridDim.x * blockDim.x) out[i] = (float)in[i];
}

inline int grid_for(int64_t count, int block = 256, int cap = 65536) {
    int64_t g = (count + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

size_t align_up(size_t b) { return (b + 255) & ~(size_t)255; }

int key_bits(int64_t kmax) {  // bits of the largest key
    int b = 1;
    while (b < 31 && ((int64_t)1 << b) <= kmax) ++b;
    return b;
}

}  // namespace

size_t transpose_scratch_bytes(int64_t nnz) {
    return 3 * align_up((size_t)std::max<int64_t>(nnz, 1) * sizeof(int)) + sort_scratch_bytes(nnz);
}

// Stable CSR -> CSC on the device: sort the CSR positions by column (stable, so rows stay
// ascending inside every column -- the reference's counting-sort order), gather rows and values,
// and take colptr from the sorted columns.  scratch: transpose_scratch_bytes(nnz) bytes, or
// nullptr to allocate them here.
int device_transpose(int n, int64_t nnz, const int* rowptr, const int* col, const double* val,
                     int* colptr, int* rowidx, double* cval, hipStream_t s, void* scratch) {
    DevBuf own;
    if (!scratch) {
        SPT_HIP(hipMalloc(&own.p, transpose_scratch_bytes(nnz)));
        scratch = own.p;
    }
    const size_t z = align_up((size_t)std::max<int64_t>(nnz, 1) * sizeof(int));
    char* at = static_cast<char*>(scratch);
    int* row_of = reinterpret_cast<int*>(at);
    int* col_sorted = reinterpret_cast<int*>(at + z);
    int* pos_sorted = reinterpret_cast<int*>(at + 2 * z);
    k_expand_rows<<<grid_for(n), 256, 0, s>>>(rowptr, n, row_of);
    SPT_CHECK(sort_pairs(col, nullptr, nnz, key_bits(n), col_sorted, pos_sorted, at + 3 * z, s));
    if (nnz > 0) k_gather_csc<<<grid_for(nnz), 256, 0, s>>>(pos_sorted, row_of, val, nnz, rowidx, cval);
    SPT_CHECK(key_bounds(col_sorted, nnz, n, 0, colptr, s));
    if (own.p) SPT_HIP(hipStreamSynchronize(s));
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

int analyse(sptrsv_plan_s* p) {
    const int n = p->n;
    const int64_t nnz = p->nnz;
    hipStream_t s = p->stream;

    // one work buffer for the transpose and then the level sort
    const size_t zn = align_up((size_t)std::max(n, 1) * sizeof(int));
    const size_t level_bytes = zn + sort_scratch_bytes(n) + scan_scratch_bytes((int64_t)n + 1);
    DevBuf work;
    const size_t work_bytes = std::max(transpose_scratch_bytes(nnz), level_bytes);
    SPT_HIP(hipMalloc(&work.p, work_bytes));
    SPT_CHECK(device_transpose(n, nnz, p->rowptr, p->col, p->val, p->colptr, p->rowidx, p->cval, s, work.p));

    // ---- in-degree: entries per row ------------------------------------------------------
    k_row_lengths<<<grid_for(n), 256, 0, s>>>(p->rowptr, n, p->indeg);

    // ---- levels --------------------------------------------------------------------------
    k_fill_int<<<grid_for(n), 256, 0, s>>>(p->levels, 0, n);
    k_fill_int<<<1, 64, 0, s>>>(p->ticket, 0, 64);
    // The sync-free sweep costs about one device round trip (~2 us) per level, the windowed
    // sweep of one wave about 0.3 us per 32 rows.  depth_lb bounds the level count from below:
    // the windowed sweep is taken once that bound makes it the cheaper one.
    SPT_HIP(hipStreamSynchronize(s));
    const auto tl = std::chrono::steady_clock::now();
    if (p->depth_lb >= 2048 && (int64_t)n <= 256 * (int64_t)p->depth_lb) {
        SPT_CHECK(chain_levels(p, p->ticket));
    } else {
        LevelTickets T{(n + kWave - 1) / kWave, 0, 0, nullptr};
        int widest = 0;  // grid order: most segments at one anti-diagonal distance
        if (p->grid && p->grid_dims[0] * p->grid_dims[1] * p->grid_dims[2] == n &&
            2 * p->grid_dims[1] * p->grid_dims[2] * ((p->grid_dims[0] + 63) / 64) * (int64_t)sizeof(int) <=
                (int64_t)work_bytes) {
            // line segments by anti-diagonal distance (items line * segs + segment), and each
            // item's position, in the work buffer (free between the transpose and the level sort)
            const int nx = (int)p->grid_dims[0], ny = (int)p->grid_dims[1], nz = (int)p->grid_dims[2];
            const int segs = (nx + 63) / 64, items = ny * nz * segs;
            std::vector<int> tab(2 * (size_t)items);
            std::vector<int> start((size_t)(segs + ny + nz), 0);  // counting sort by distance
            auto dist = [&](int g, int j, int k) { return (segs - 1 - g) + (ny - 1 - j) + (nz - 1 - k); };
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int g = 0; g < segs; ++g) ++start[(size_t)dist(g, j, k) + 1];
            for (size_t d = 1; d < start.size(); ++d) widest = std::max(widest, start[d]);
            for (size_t d = 1; d < start.size(); ++d) start[d] += start[d - 1];
            for (int k = nz - 1; k >= 0; --k)
                for (int j = ny - 1; j >= 0; --j)
                    for (int g = segs - 1; g >= 0; --g) {
                        const int e = (j + ny * k) * segs + g, q = start[(size_t)dist(g, j, k)]++;
                        tab[(size_t)q] = e;
                        tab[(size_t)items + (size_t)e] = q;
                    }
            int* d_tab = static_cast<int*>(work.p);
            SPT_HIP(hipMemcpy(d_tab, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice));
            k_grid_order_check<<<grid_for(n), 256, 0, s>>>(n, p->rowptr, p->col, nx, segs, d_tab + items,
                                                           p->ticket + 3);
            int bad = 0;
            SPT_HIP(hipMemcpyAsync(&bad, p->ticket + 3, sizeof(int), hipMemcpyDeviceToHost, s));
            SPT_HIP(hipStreamSynchronize(s));
            k_fill_int<<<1, 64, 0, s>>>(p->ticket + 3, 0, 1);
            if (!bad) T = LevelTickets{items, nx, segs, d_tab};
        }
        const uint64_t limit = p->ticks_per_ms * 60000ull;  // a level sweep never waits a minute
        // persistent waves.  Row order: 8 per CU (32 were slower: the polls of rows in flight
        // contend), fewer for small factors.  Grid order: about half the widest anti-diagonal,
        // one wave per workgroup -- on 128^3 (256 segments at the widest) 8, 4, 2, 1 and 0.5
        // waves per CU took 1.81, 1.53, 1.41, 1.27 and 1.19 ms, 0.25 per CU 1.89 ms
        if (T.item) {
            const int waves = std::min(T.count, std::max(32, std::min(widest / 2, device_compute_units())));
            k_levels<<<waves, 64, 0, s>>>(n, T, p->rowptr, p->col, p->levels, p->ticket, limit);
        } else {
            const int blocks = std::min((T.count + 3) / 4, device_compute_units() * 2);
            k_levels<<<blocks, 256, 0, s>>>(n, T, p->rowptr, p->col, p->levels, p->ticket, limit);
        }
    }
    int ctrl[3] = {0, 0, 0};
    SPT_HIP(hipMemcpyAsync(ctrl, p->ticket, sizeof(ctrl), hipMemcpyDeviceToHost, s));
    SPT_HIP(hipStreamSynchronize(s));
    p->levels_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tl).count();
    if (ctrl[1] != 0)
        return fail(SPTRSV_ERR_DEADLOCK, "level analysis watchdog fired");
    const int L = ctrl[2];
    p->n_levels = L;

    // ---- level sets: row ids stably sorted by level (rows ascending per level), and the
    // level pointers from the sorted levels (level_ptr[0] = 0) ------------------------------
    int* lev_sorted = static_cast<int*>(work.p);
    char* rest = static_cast<char*>(work.p) + zn;
    SPT_HIP(hipMalloc((void**)&p->level_ptr, ((size_t)L + 1) * sizeof(int)));
    SPT_CHECK(sort_pairs(p->levels, nullptr, n, key_bits(L), lev_sorted, p->level_rows, rest, s));
    SPT_CHECK(key_bounds(lev_sorted, n, L, 1, p->level_ptr, s));

    // ---- level-ordered CSR ---------------------------------------------------------------
    k_level_rowlen<<<grid_for(n), 256, 0, s>>>(p->rowptr, p->level_rows, n, p->lrowptr);
    SPT_CHECK(scan_ints(p->lrowptr, p->lrowptr, (int64_t)n + 1, true, rest, s));
    int G = 1;
    while (G < kWave && (int64_t)G * n < nnz) G *= 2;
    k_level_gather<<<grid_for((int64_t)n * G), 256, 0, s>>>(p->rowptr, p->col, p->val, p->level_rows, p->lrowptr,
                                                             n, G, p->lcol, p->lval);
    p->h_level_ptr.resize((size_t)L + 1);
    SPT_HIP(hipMemcpyAsync(p->h_level_ptr.data(), p->level_ptr, ((size_t)L + 1) * sizeof(int),
                           hipMemcpyDeviceToHost, s));
    SPT_HIP(hipStreamSynchronize(s));
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

int ensure_fp32(sptrsv_plan_s* p) {
    if (p->val32) return SPTRSV_OK;
    const int64_t nnz = p->nnz;
    SPT_HIP(hipMalloc((void**)&p->val32, nnz * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&p->cval32, nnz * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&p->lval32, nnz * sizeof(float)));
    k_to_float<<<grid_for(nnz), 256, 0, p->stream>>>(p->val, p->val32, nnz);
    k_to_float<<<grid_for(nnz), 256, 0, p->stream>>>(p->cval, p->cval32, nnz);
    k_to_float<<<grid_for(nnz), 256, 0, p->stream>>>(p->lval, p->lval32, nnz);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

}  // namespace sptrsv

extern "C" int sptrsv_csr_to_csc(int n, int64_t nnz, const int64_t* rowptr, const int* col, const double* val,
                                 int* colptr, int* rowidx, double* cval) {
    using namespace sptrsv;
    if (n <= 0 || nnz < 0 || !rowptr || !col || !val || !colptr || !rowidx || !cval)
        return fail(SPTRSV_ERR_BAD_INPUT, "bad transpose arguments");
    if (nnz >= ((int64_t)1 << 31)) return fail(SPTRSV_ERR_UNSUPPORTED, "nnz >= 2^31");
    if (rowptr[0] != 0 || rowptr[n] != nnz) return fail(SPTRSV_ERR_BAD_INPUT, "rowptr does not span nnz");
    for (int64_t k = 0; k < nnz; ++k)
        if (col[k] < 0 || col[k] >= n) return fail(SPTRSV_ERR_BAD_INPUT, "column index out of range");
    for (int i = 0; i < n; ++i)
        if (rowptr[i + 1] < rowptr[i]) return fail(SPTRSV_ERR_BAD_INPUT, "rowptr not monotone");
    SPT_CHECK(ensure_device());
    std::vector<int> rp((size_t)n + 1);
    for (int i = 0; i <= n; ++i) rp[i] = (int)rowptr[i];
    DevBuf d_rp, d_col, d_val, d_cp, d_ri, d_cv;
    const size_t z = (size_t)std::max<int64_t>(nnz, 1);
    SPT_HIP(hipMalloc(&d_rp.p, rp.size() * sizeof(int)));
    SPT_HIP(hipMalloc(&d_col.p, z * sizeof(int)));
    SPT_HIP(hipMalloc(&d_val.p, z * sizeof(double)));
    SPT_HIP(hipMalloc(&d_cp.p, ((size_t)n + 1) * sizeof(int)));
    SPT_HIP(hipMalloc(&d_ri.p, z * sizeof(int)));
    SPT_HIP(hipMalloc(&d_cv.p, z * sizeof(double)));
    SPT_HIP(hipMemcpy(d_rp.p, rp.data(), rp.size() * sizeof(int), hipMemcpyHostToDevice));
    SPT_HIP(hipMemcpy(d_col.p, col, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice));
    SPT_HIP(hipMemcpy(d_val.p, val, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice));
    SPT_CHECK(device_transpose(n, nnz, (int*)d_rp.p, (int*)d_col.p, (double*)d_val.p, (int*)d_cp.p,
                               (int*)d_ri.p, (double*)d_cv.p, nullptr));
    SPT_HIP(hipMemcpy(colptr, d_cp.p, ((size_t)n + 1) * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(rowidx, d_ri.p, (size_t)nnz * sizeof(int), hipMemcpyDeviceToHost));
    SPT_HIP(hipMemcpy(cval, d_cv.p, (size_t)nnz * sizeof(double), hipMemcpyDeviceToHost));
    return SPTRSV_OK;
}
