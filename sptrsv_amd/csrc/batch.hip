// batch.hip -- many independent upper-triangular systems in one persistent launch.
//
// The systems are stored as one block-diagonal CSR with 64-bit row pointers (a batch of
// thousands of large systems passes 2^31 nonzeros), owned by the batch; the level analysis of
// every system runs on the host when the batch is created (systems in parallel).  Each system
// carries a configuration:
//   LS_G1 / LS_G4 / LS_G16 / LS_G64   one workgroup walks the system's levels with a barrier
//       between levels; G lanes per row (over the right-hand sides when m > 1, over the row's
//       entries when m == 1 -- then the partial sums meet through cross-lane shuffles)
//   CHAIN   (m <= 32) the 64-row window recurrence of chain.hip: m == 1 on wave 0 alone,
//       m > 1 with the window tile shared in LDS and the waves taking the right-hand sides
// Persistent workgroups (2 per CU) pop systems from a queue in a given order (longest
// predicted time first) until it is empty; systems never wait on each other, so there is no
// cross-workgroup dependency at all.  With a timing buffer the kernel also records every
// system's device time (s_memrealtime) -- the autotuner sweeps configurations with it, one
// launch per configuration.  The whole solve is replayable as a captured hipGraph.
#include <hip/hip_runtime.h>
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.h"
#include "plan.h"

struct sptrsv_batch_s {
    sptrsv_plan_s* ctl = nullptr;  // stream, status word, timing, refinement and FP32 staging
    int nsys = 0;
    int n = 0;
    int64_t nnz = 0;
    std::vector<int> h_sys_ptr;
    int64_t* rowptr = nullptr;  // n + 1, block-diagonal CSR, diagonal first in every row
    int* col = nullptr;
    double* val = nullptr;
    float* val32 = nullptr;     // FP32 copy, built on the first FP32 solve
    int* sys_ptr = nullptr;   // nsys + 1 row offsets
    int* lvl_off = nullptr;   // nsys + 1 offsets into lvl_ptr
    int* lvl_ptr = nullptr;   // per system: L_s + 1 absolute positions into rows[]
    int* rows = nullptr;      // rows ordered by (system, level, row)
    int* cfg = nullptr;       // nsys
    int* order = nullptr;     // nsys, processing order
    int* queue = nullptr;     // [0] next system
    std::vector<int> h_cfg;
    std::vector<float> h_features;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    const void* g_b = nullptr;
    void* g_x = nullptr;
    int g_m = 0, g_prec = -1;
    unsigned long long* timer = nullptr;  // nsys start/stop ticks
    int workgroups = 0;
};

namespace sptrsv {
namespace {

constexpr int kBRing = 2048;     // x values kept in LDS by the chain configuration
constexpr int kBChainMaxM = 32;  // chain configuration up to this many right-hand sides

template <typename T>
struct BatchArgs {
    int nsys;
    int m;
    const int* sys_ptr;
    const int* lvl_off;
    const int* lvl_ptr;
    const int* rows;
    const int* cfg;
    const int* order;
    const int64_t* rowptr;
    const int* col;
    const T* val;
    const T* b;
    T* x;
    int* queue;
    unsigned long long* timer;  // optional: 2 per system
};

template <typename T>
__device__ __forceinline__ T bcast_b(T v, int src) {
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)u, src);
        const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), src);
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), src));
    }
}

// Window recurrence over rows [lo, hi) (one wave; the system's own rows only).  Same
// arithmetic as chain.hip without its prefetch pipeline: systems here are short.
template <typename T>
__device__ void chain_system(const BatchArgs<T>& a, int lo, int top, T (*tile)[64], T* ring) {
    const int lane = threadIdx.x & 63;
    for (int hi = top; hi > lo; hi -= 64) {
        const int r = hi - 1 - lane;
        const bool valid = r >= lo;
#pragma unroll 8
        for (int s = 0; s < 64; ++s) tile[s][lane] = T(0);
        T acc = T(0), diag = T(1), bv = T(0);
        if (valid) {
            const int64_t st = a.rowptr[r], en = a.rowptr[r + 1];
            diag = a.val[st];
            bv = a.b[r];
            for (int64_t p = st + 1; p < en; ++p) {
                const int c = a.col[p];
                const T v = a.val[p];
                if (c < hi) tile[hi - 1 - c][lane] = v;
                else if (c < hi + kBRing) acc += v * ring[c & (kBRing - 1)];
                else acc += v * a.x[c];  // written >= 32 windows ago by this wave (see below)
            }
        }
        const T inv = T(1) / diag;
        const T binv = bv * inv;
        __builtin_amdgcn_wave_barrier();
        T mine = T(0);
#pragma unroll 16
        for (int s = 0; s < 64; ++s) {
            const T xs = fma(-acc, inv, binv);
            const T xb = bcast_b(xs, s);
            if (lane == s) mine = xs;
            acc += tile[s][lane] * xb;
        }
        if (valid) {
            ring[r & (kBRing - 1)] = mine;
            a.x[r] = mine;
        }
        __builtin_amdgcn_wave_barrier();
    }
    wait_vmem();
}

// The same window recurrence with m right-hand sides (m <= 32): the window's coefficient tile
// is built once by wave 0 and shared; wave w solves columns w, w + 4, ...; the LDS ring holds
// the last kBRing / m rows of each column (>= 64, so the previous window is always in it).
template <typename T>
__device__ void chain_system_m(const BatchArgs<T>& a, int lo, int top, T (*tile)[64], T* ring, T* sinv) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = a.m;
    int R = kBRing;
    while (R * m > kBRing) R >>= 1;
    for (int hi = top; hi > lo; hi -= 64) {
        const int r = hi - 1 - lane;
        const bool valid = r >= lo;
        if (wave == 0) {
#pragma unroll 8
            for (int s = 0; s < 64; ++s) tile[s][lane] = T(0);
            T diag = T(1);
            if (valid) {
                const int64_t st = a.rowptr[r], en = a.rowptr[r + 1];
                diag = a.val[st];
                for (int64_t p = st + 1; p < en; ++p) {
                    const int c = a.col[p];
                    if (c < hi) tile[hi - 1 - c][lane] = a.val[p];
                }
            }
            sinv[lane] = T(1) / diag;
        }
        __syncthreads();
        const T inv = sinv[lane];
        for (int k = wave; k < m; k += 4) {
            T* rk = ring + k * R;
            T acc = T(0), bv = T(0);
            if (valid) {
                const int64_t st = a.rowptr[r], en = a.rowptr[r + 1];
                bv = a.b[(size_t)r * m + k];
                for (int64_t p = st + 1; p < en; ++p) {
                    const int c = a.col[p];
                    if (c < hi) continue;
                    const T v = a.val[p];
                    acc += v * (c < hi + R ? rk[c & (R - 1)] : a.x[(size_t)c * m + k]);
                }
            }
            const T binv = bv * inv;
            T mine = T(0);
#pragma unroll 16
            for (int s = 0; s < 64; ++s) {
                const T xs = fma(-acc, inv, binv);
                const T xb = bcast_b(xs, s);
                if (lane == s) mine = xs;
                acc += tile[s][lane] * xb;
            }
            if (valid) {
                rk[r & (R - 1)] = mine;
                a.x[(size_t)r * m + k] = mine;
            }
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_batch(BatchArgs<T> a) {
#pragma clang fp contract(off)
    __shared__ T tile[64][64];
    __shared__ T ring[kBRing];
    __shared__ T sinv[64];
    __shared__ int s_sys;
    const int tid = threadIdx.x;
    const size_t m = (size_t)a.m;
    while (true) {
        if (tid == 0) {
            const int q = atomicAdd(a.queue, 1);
            s_sys = q < a.nsys ? a.order[q] : -1;
        }
        __syncthreads();
        const int s = s_sys;
        if (s < 0) break;
        const int cfg = a.cfg[s];
        const int lo = a.sys_ptr[s], hi = a.sys_ptr[s + 1];
        unsigned long long t0 = 0;
        if (a.timer && tid == 0) t0 = wall_clock64();
        if (cfg == SPTRSV_BCFG_CHAIN && a.m == 1) {
            if (tid < 64) chain_system(a, lo, hi, tile, ring);
        } else if (cfg == SPTRSV_BCFG_CHAIN && a.m <= kBChainMaxM) {
            chain_system_m(a, lo, hi, tile, ring, sinv);
        } else {
            const int G = cfg == SPTRSV_BCFG_LS1 ? 1 : cfg == SPTRSV_BCFG_LS4 ? 4 : cfg == SPTRSV_BCFG_LS16 ? 16 : 64;
            const int groups = 256 / G, grp = tid / G, gl = tid % G;
            const int l0 = a.lvl_off[s], l1 = a.lvl_off[s + 1];
            for (int l = l0; l + 1 < l1; ++l) {
                const int rs = a.lvl_ptr[l], re = a.lvl_ptr[l + 1];
                for (int base = rs; base < re; base += groups) {
                    const int idx = base + grp;
                    const bool on = idx < re;
                    const int row = on ? a.rows[idx] : 0;
                    const int64_t st = on ? a.rowptr[row] : 0, en = on ? a.rowptr[row + 1] : 0;
                    if (a.m == 1) {
                        // G lanes split the entries, cross-lane reduction (groups are whole
                        // aligned lane sets, so every lane of a group takes the same path)
                        T acc = T(0);
                        for (int64_t p = st + 1 + gl; p < en; p += G) acc = acc + a.val[p] * a.x[a.col[p]];
                        for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
                        if (on && gl == 0) a.x[row] = (a.b[row] - acc) / a.val[st];
                    } else if (on) {
                        // G lanes over the right-hand sides, ascending entry order (exact)
                        const T d = a.val[st];
                        for (size_t k = gl; k < m; k += G) {
                            T acc = T(0);
                            for (int64_t p = st + 1; p < en; ++p) acc = acc + a.val[p] * a.x[(size_t)a.col[p] * m + k];
                            a.x[(size_t)row * m + k] = (a.b[(size_t)row * m + k] - acc) / d;
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (a.timer) {
            __syncthreads();
            if (tid == 0) {
                a.timer[2 * s] = t0;
                a.timer[2 * s + 1] = wall_clock64();
            }
        }
        __syncthreads();
    }
}

__global__ void k_reset_queue(int* q) {
    if (threadIdx.x == 0) q[0] = 0;
}

template <typename X>
int upload(X** dst, const std::vector<X>& src) {
    SPT_HIP(hipMalloc((void**)dst, std::max<size_t>(src.size(), 1) * sizeof(X)));
    if (!src.empty()) SPT_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(X), hipMemcpyHostToDevice));
    return SPTRSV_OK;
}

__global__ void k_val_to_float(const double* __restrict__ in, float* __restrict__ out, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

inline int grid_for_b(int64_t count) {
    return (int)std::min<int64_t>(std::max<int64_t>((count + 255) / 256, 1), 65536);
}

int ensure_batch_fp32(sptrsv_batch_s* B) {
    if (B->val32) return SPTRSV_OK;
    SPT_HIP(hipMalloc((void**)&B->val32, std::max<int64_t>(B->nnz, 1) * sizeof(float)));
    k_val_to_float<<<grid_for_b(B->nnz), 256, 0, B->ctl->stream>>>(B->val, B->val32, B->nnz);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

template <typename T>
int launch_batch(sptrsv_batch_s* B, const T* val, const T* d_b, T* d_x, int m, unsigned long long* timer) {
    sptrsv_plan_s* P = B->ctl;
    BatchArgs<T> a{B->nsys, m, B->sys_ptr, B->lvl_off, B->lvl_ptr, B->rows, B->cfg, B->order,
                   B->rowptr, B->col, val, d_b, d_x, B->queue, timer};
    k_reset_queue<<<1, 64, 0, P->stream>>>(B->queue);
    k_batch<T><<<B->workgroups, 256, 0, P->stream>>>(a);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

}  // namespace

// The 24 selector features of one system, computed from its CSR and levels (host).
void system_features(const int64_t* rp, const int* col, const double* val, const std::vector<int>& lev, int lo, int hi,
                     int m, float* f) {
    const int n = hi - lo;
    const int64_t e0 = rp[lo], e1 = rp[hi];
    const double nnz = (double)(e1 - e0);
    int L = 0;
    for (int r = lo; r < hi; ++r) L = std::max(L, lev[r]);
    std::vector<int> width(L + 1, 0);
    for (int r = lo; r < hi; ++r) width[lev[r]]++;
    double wmax = 0, wsum = 0, wsq = 0;
    for (int l = 1; l <= L; ++l) {
        wmax = std::max(wmax, (double)width[l]);
        wsum += width[l];
        wsq += (double)width[l] * width[l];
    }
    const double wmean = wsum / L, wstd = std::sqrt(std::max(0.0, wsq / L - wmean * wmean));
    double lmax = 0, lsq = 0, diag_only = 0, off_sum = 0, off_max = 0, off1 = 0, off64 = 0, offfar = 0;
    double has_next = 0, rowmax_off = 0, dsum = 0, osum = 0, noff = 0;
    for (int r = lo; r < hi; ++r) {
        const int64_t len = rp[r + 1] - rp[r];
        lmax = std::max(lmax, (double)len);
        lsq += (double)len * len;
        if (len == 1) diag_only += 1;
        dsum += std::fabs(val[rp[r]]);
        double rmax = 0;
        for (int64_t p = rp[r] + 1; p < rp[r + 1]; ++p) {
            const double o = col[p] - r;
            off_sum += o;
            off_max = std::max(off_max, o);
            rmax = std::max(rmax, o);
            if (o == 1) {
                off1 += 1;
                has_next += 1;
            }
            if (o <= 64) off64 += 1;
            if (o > 4096) offfar += 1;
            osum += std::fabs(val[p]);
            noff += 1;
        }
        rowmax_off += rmax;
    }
    // fundamental supernodes: row r+1 continues row r's supernode when pattern(r) = {r} U pattern(r+1)
    int nsn = 1, cur = 1, snmax = 1;
    double in_wide = 0;
    std::vector<int> widths;
    for (int r = lo; r + 1 < hi; ++r) {
        const int64_t li = rp[r + 1] - rp[r], lj = rp[r + 2] - rp[r + 1];
        bool cont = li == lj + 1 && li >= 2 && col[rp[r] + 1] == r + 1;
        for (int64_t t = 0; cont && t < lj; ++t) cont = col[rp[r] + 1 + t] == col[rp[r + 1] + t];
        if (cont) {
            ++cur;
        } else {
            widths.push_back(cur);
            ++nsn;
            cur = 1;
        }
    }
    widths.push_back(cur);
    for (int w : widths) {
        snmax = std::max(snmax, w);
        if (w >= 16) in_wide += w;
    }
    const double rmean = nnz / n;
    const double nd = std::max(noff, 1.0);
    f[0] = (float)std::log2((double)n);
    f[1] = (float)std::log2(nnz);
    f[2] = (float)rmean;
    f[3] = (float)lmax;
    f[4] = (float)std::sqrt(std::max(0.0, lsq / n - rmean * rmean));
    f[5] = (float)((double)L / n);
    f[6] = (float)std::log2((double)L);
    f[7] = (float)std::log2(wmean);
    f[8] = (float)(wmax / n);
    f[9] = (float)(wstd / wmean);
    f[10] = (float)(diag_only / n);
    f[11] = (float)(off_sum / nd / n);
    f[12] = (float)(off_max / n);
    f[13] = (float)(off1 / nd);
    f[14] = (float)(off64 / nd);
    f[15] = (float)(offfar / nd);
    f[16] = (float)(has_next / n);
    f[17] = (float)((double)nsn / n);
    f[18] = (float)((double)n / nsn);
    f[19] = (float)snmax;
    f[20] = (float)(in_wide / n);
    f[21] = (float)(rowmax_off / n / n);
    f[22] = (float)std::log2((double)m);
    f[23] = (float)std::log2((dsum / n) / std::max(osum / nd, 1e-300));
}

}  // namespace sptrsv

using namespace sptrsv;

extern "C" {

int sptrsv_batch_destroy(sptrsv_batch_t B) {
    if (!B) return SPTRSV_OK;
    if (B->ctl && B->ctl->stream) (void)hipStreamSynchronize(B->ctl->stream);
    if (B->gexec) (void)hipGraphExecDestroy(B->gexec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    void* bufs[] = {B->rowptr, B->col, B->val, B->val32, B->sys_ptr, B->lvl_off, B->lvl_ptr,
                    B->rows, B->cfg, B->order, B->queue, B->timer};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    sptrsv_plan_destroy(B->ctl);
    delete B;
    return SPTRSV_OK;
}

int sptrsv_batch_create(int nsys, const int* sys_ptr, int64_t nnz, const int64_t* rowptr, const int* col,
                        const double* val, sptrsv_batch_t* out) {
    if (!out || !sys_ptr || !rowptr || !col || !val || nsys <= 0) return fail(SPTRSV_ERR_BAD_INPUT, "bad batch arguments");
    *out = nullptr;
    const int n = sys_ptr[nsys];
    if (sys_ptr[0] != 0) return fail(SPTRSV_ERR_BAD_INPUT, "system offsets must start at 0");
    for (int s = 0; s < nsys; ++s)
        if (sys_ptr[s + 1] <= sys_ptr[s]) return fail(SPTRSV_ERR_BAD_INPUT, "empty system in batch");
    if (rowptr[0] != 0 || rowptr[n] != nnz) return fail(SPTRSV_ERR_BAD_INPUT, "rowptr does not span nnz");
    // validation, per system in parallel: upper CSR with the diagonal first (finite, non-zero),
    // columns strictly ascending, every column inside the row's own system
    std::vector<int> bad((size_t)nsys, 0);
#pragma omp parallel for schedule(dynamic, 4)
    for (int s = 0; s < nsys; ++s) {
        for (int r = sys_ptr[s]; r < sys_ptr[s + 1] && !bad[(size_t)s]; ++r) {
            const int64_t a = rowptr[r], e = rowptr[r + 1];
            if (e <= a || col[a] != r || !std::isfinite(val[a]) || val[a] == 0.0) {
                bad[(size_t)s] = 1;
                break;
            }
            for (int64_t p = a + 1; p < e; ++p)
                if (col[p] <= col[p - 1] || col[p] >= sys_ptr[s + 1]) {
                    bad[(size_t)s] = 2;
                    break;
                }
        }
    }
    for (int s = 0; s < nsys; ++s) {
        if (bad[(size_t)s] == 1)
            return fail(SPTRSV_ERR_BAD_INPUT, "system " + std::to_string(s) +
                                                  ": a row does not start with a finite non-zero diagonal");
        if (bad[(size_t)s] == 2)
            return fail(SPTRSV_ERR_BAD_INPUT, "system " + std::to_string(s) +
                                                  " has unsorted columns or references a column of another system");
    }
    // per-system levels (1-based; rows only reference later rows of their own system)
    std::vector<int> lev((size_t)n), lcount((size_t)nsys);
#pragma omp parallel for schedule(dynamic, 4)
    for (int s = 0; s < nsys; ++s) {
        int L = 0;
        for (int r = sys_ptr[s + 1] - 1; r >= sys_ptr[s]; --r) {
            int l = 0;
            for (int64_t p = rowptr[r] + 1; p < rowptr[r + 1]; ++p) l = std::max(l, lev[(size_t)col[p]]);
            lev[(size_t)r] = l + 1;
            L = std::max(L, l + 1);
        }
        lcount[(size_t)s] = L;
    }
    // rows ordered by (system, level, row); lvl_ptr holds, per system, L_s + 1 positions
    std::vector<int> lvl_off((size_t)nsys + 1, 0);
    for (int s = 0; s < nsys; ++s) lvl_off[(size_t)s + 1] = lvl_off[(size_t)s] + lcount[(size_t)s] + 1;
    std::vector<int> lvl_ptr((size_t)lvl_off[(size_t)nsys]), rows((size_t)n);
#pragma omp parallel for schedule(dynamic, 4)
    for (int s = 0; s < nsys; ++s) {
        const int lo = sys_ptr[s], hi = sys_ptr[s + 1], L = lcount[(size_t)s];
        std::vector<int> cnt((size_t)L + 2, 0);
        for (int r = lo; r < hi; ++r) cnt[(size_t)lev[(size_t)r] + 1]++;
        for (int l = 1; l <= L + 1; ++l) cnt[(size_t)l] += cnt[(size_t)l - 1];
        // cnt[l] = rows with level < l: level l occupies positions lo + [cnt[l], cnt[l+1])
        for (int l = 1; l <= L + 1; ++l) lvl_ptr[(size_t)lvl_off[(size_t)s] + l - 1] = lo + cnt[(size_t)l];
        std::vector<int> cur(cnt.begin(), cnt.end());
        for (int r = lo; r < hi; ++r) rows[(size_t)(lo + cur[(size_t)lev[(size_t)r]]++)] = r;
    }
    auto* B = new sptrsv_batch_s();
    int rc = plan_create_control(n, nnz, &B->ctl);
    if (rc != SPTRSV_OK) {
        sptrsv_batch_destroy(B);
        return rc;
    }
    B->nsys = nsys;
    B->n = n;
    B->nnz = nnz;
    B->h_sys_ptr.assign(sys_ptr, sys_ptr + nsys + 1);
    std::vector<int> cfg(nsys, SPTRSV_BCFG_LS1), order(nsys);
    for (int s = 0; s < nsys; ++s) order[s] = s;
    B->workgroups = std::max(1, device_compute_units()) * 2;
    const size_t e = (size_t)std::max<int64_t>(nnz, 1);
    if (hipMalloc((void**)&B->rowptr, ((size_t)n + 1) * sizeof(int64_t)) != hipSuccess ||
        hipMalloc((void**)&B->col, e * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&B->val, e * sizeof(double)) != hipSuccess) {
        sptrsv_batch_destroy(B);
        return fail(SPTRSV_ERR_NOMEM, "device allocation of the batch's matrix");
    }
    if (hipMemcpy(B->rowptr, rowptr, ((size_t)n + 1) * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->col, col, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->val, val, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        sptrsv_batch_destroy(B);
        return fail(SPTRSV_ERR_HIP, "upload of the batch's matrix");
    }
    if ((rc = upload(&B->sys_ptr, B->h_sys_ptr)) || (rc = upload(&B->lvl_off, lvl_off)) ||
        (rc = upload(&B->lvl_ptr, lvl_ptr)) || (rc = upload(&B->rows, rows)) || (rc = upload(&B->cfg, cfg)) ||
        (rc = upload(&B->order, order))) {
        sptrsv_batch_destroy(B);
        return rc;
    }
    B->h_cfg = cfg;
    if (hipMalloc((void**)&B->queue, 64 * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&B->timer, 2 * (size_t)nsys * sizeof(unsigned long long)) != hipSuccess) {
        sptrsv_batch_destroy(B);
        return fail(SPTRSV_ERR_NOMEM, "batch work buffers");
    }
    // selector features (host, once, systems in parallel)
    B->h_features.assign((size_t)nsys * 24, 0.f);
#pragma omp parallel for schedule(dynamic, 4)
    for (int s = 0; s < nsys; ++s)
        system_features(rowptr, col, val, lev, sys_ptr[s], sys_ptr[s + 1], 1, &B->h_features[(size_t)s * 24]);
    *out = B;
    return SPTRSV_OK;
}

int sptrsv_batch_info(sptrsv_batch_t B, int* nsys, int* n, int64_t* nnz) {
    if (!B) return fail(SPTRSV_ERR_BAD_INPUT, "null batch");
    if (nsys) *nsys = B->nsys;
    if (n) *n = B->n;
    if (nnz) *nnz = B->nnz;
    return SPTRSV_OK;
}

int sptrsv_batch_features(sptrsv_batch_t B, int m, float* feats) {
    if (!B || !feats || m < 1) return fail(SPTRSV_ERR_BAD_INPUT, "bad features arguments");
    for (int s = 0; s < B->nsys; ++s) {
        for (int k = 0; k < 24; ++k) feats[(size_t)s * 24 + k] = B->h_features[(size_t)s * 24 + k];
        feats[(size_t)s * 24 + 22] = (float)std::log2((double)m);
    }
    return SPTRSV_OK;
}

int sptrsv_batch_set_configs(sptrsv_batch_t B, const int* cfg, const int* order) {
    if (!B) return fail(SPTRSV_ERR_BAD_INPUT, "null batch");
    if (cfg) {
        for (int s = 0; s < B->nsys; ++s)
            if (cfg[s] < 0 || cfg[s] >= SPTRSV_BCFG_COUNT) return fail(SPTRSV_ERR_BAD_INPUT, "bad configuration id");
        B->h_cfg.assign(cfg, cfg + B->nsys);
        SPT_HIP(hipMemcpy(B->cfg, cfg, (size_t)B->nsys * sizeof(int), hipMemcpyHostToDevice));
    }
    if (order) {
        std::vector<char> seen((size_t)B->nsys, 0);
        for (int q = 0; q < B->nsys; ++q) {
            if (order[q] < 0 || order[q] >= B->nsys || seen[(size_t)order[q]])
                return fail(SPTRSV_ERR_BAD_INPUT, "order must be a permutation of the systems");
            seen[(size_t)order[q]] = 1;
        }
        SPT_HIP(hipMemcpy(B->order, order, (size_t)B->nsys * sizeof(int), hipMemcpyHostToDevice));
    }
    return SPTRSV_OK;
}

// FP64 (precision 0) or FP32 (1) solve of every system, one persistent launch, asynchronous.
// use_graph != 0: the launch is captured once per (b, x, m, precision) and replayed.
int sptrsv_batch_solve(sptrsv_batch_t B, int precision, const double* d_b, double* d_x, int m, int use_graph) {
    if (!B || !d_b || !d_x || m < 1) return fail(SPTRSV_ERR_BAD_INPUT, "bad batch solve arguments");
    sptrsv_plan_s* P = B->ctl;
    if (precision == SPTRSV_FP32) {
        SPT_CHECK(ensure_batch_fp32(B));
        SPT_CHECK(ensure_stage(P, (int64_t)B->n * m));
    } else if (precision != SPTRSV_FP64) {
        return fail(SPTRSV_ERR_BAD_INPUT, "batch precision is FP64 or FP32");
    }
    auto enqueue = [&]() -> int {
        if (precision == SPTRSV_FP64) return launch_batch<double>(B, B->val, d_b, d_x, m, nullptr);
        // FP32 values; b and x are converted on the device around the solve
        return launch_batch<float>(B, B->val32, P->b32, P->x32, m, nullptr);
    };
    if (precision == SPTRSV_FP32) SPT_CHECK(convert_d2f(d_b, P->b32, (int64_t)B->n * m, P->stream));
    if (!use_graph) {
        hipEvent_t stop;
        SPT_CHECK(begin_timed(P, &stop));
        SPT_CHECK(enqueue());
        end_timed(P, stop);
    } else {
        if (!B->gexec || B->g_b != d_b || B->g_x != d_x || B->g_m != m || B->g_prec != precision) {
            if (B->gexec) SPT_HIP(hipGraphExecDestroy(B->gexec));
            if (B->graph) SPT_HIP(hipGraphDestroy(B->graph));
            B->gexec = nullptr;
            B->graph = nullptr;
            SPT_HIP(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue();
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(P->stream, &g);
            if (rc != SPTRSV_OK) return rc;
            if (e != hipSuccess) return fail(SPTRSV_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            B->graph = g;
            SPT_HIP(hipGraphInstantiate(&B->gexec, g, nullptr, nullptr, 0));
            B->g_b = d_b;
            B->g_x = d_x;
            B->g_m = m;
            B->g_prec = precision;
        }
        hipEvent_t stop;
        SPT_CHECK(begin_timed(P, &stop));
        SPT_HIP(hipGraphLaunch(B->gexec, P->stream));
        end_timed(P, stop);
    }
    if (precision == SPTRSV_FP32) SPT_CHECK(convert_f2d(P->x32, d_x, (int64_t)B->n * m, P->stream));
    return SPTRSV_OK;
}

// FP32 batched solves with FP64 refinement: the residual runs over the whole block-diagonal
// CSR, the stopping rule over the inf-norms of all systems together; with opts.fallback a run
// that does not converge is finished by the FP64 batched solve.  Synchronous.
int sptrsv_batch_solve_refined_ex(sptrsv_batch_t B, const sptrsv_refine_t* opts, const double* d_b, double* d_x,
                                  int m, sptrsv_refine_info_t* info) {
    if (!B || !d_b || !d_x || m < 1) return fail(SPTRSV_ERR_BAD_INPUT, "bad batch solve arguments");
    sptrsv_refine_t ro{};
    if (opts) ro = *opts;
    if (ro.precision != 0 && ro.precision != SPTRSV_FP32)
        return fail(SPTRSV_ERR_UNSUPPORTED, "batched refinement runs FP32 inner solves");
    sptrsv_plan_s* P = B->ctl;
    SPT_CHECK(ensure_batch_fp32(B));
    auto solve_low = [&](const float* in, float* out) -> int {
        return launch_batch<float>(B, B->val32, in, out, m, nullptr);
    };
    auto residual = [&](const double* b, const double* x, float* r) -> int {
        launch_residual<int64_t, float>(B->n, B->nnz, m, B->rowptr, B->col, B->val, x, b, r, P->stream);
        SPT_HIP(hipGetLastError());
        return SPTRSV_OK;
    };
    RefineResult res{};
    SPT_CHECK(refine_loop(P, ro, d_b, d_x, m, solve_low, residual, &res));
    int fell_back = 0;
    if (res.outcome != SPTRSV_REFINE_CONVERGED && ro.fallback) {
        SPT_CHECK(launch_batch<double>(B, B->val, d_b, d_x, m, nullptr));
        SPT_CHECK(sptrsv_plan_sync(P));
        fell_back = 1;
    }
    if (info) {
        info->iters = res.iters;
        info->outcome = res.outcome;
        info->fell_back = fell_back;
        info->variant = -1;
        info->estimate = res.estimate;
    }
    return SPTRSV_OK;
}

int sptrsv_batch_solve_refined(sptrsv_batch_t B, const sptrsv_refine_t* opts, const double* d_b, double* d_x, int m,
                               int* iters, double* estimate) {
    sptrsv_refine_info_t info{};
    SPT_CHECK(sptrsv_batch_solve_refined_ex(B, opts, d_b, d_x, m, &info));
    if (iters) *iters = info.iters;
    if (estimate) *estimate = info.estimate;
    return SPTRSV_OK;
}

int sptrsv_batch_sync(sptrsv_batch_t B) {
    if (!B) return fail(SPTRSV_ERR_BAD_INPUT, "null batch");
    return sptrsv_plan_sync(B->ctl);
}

// r = b - U x over every system, FP64, asynchronous on the batch's stream.
int sptrsv_batch_residual(sptrsv_batch_t B, const double* d_b, const double* d_x, double* d_r, int m) {
    if (!B || !d_b || !d_x || !d_r || m < 1) return fail(SPTRSV_ERR_BAD_INPUT, "bad residual arguments");
    launch_residual<int64_t, double>(B->n, B->nnz, m, B->rowptr, B->col, B->val, d_x, d_b, d_r, B->ctl->stream);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

// Autotuner primitive: one launch with every system on configuration `cfg` (systems for which
// cfg is invalid -- CHAIN with m > 32 -- run LS_G1), device time of each system in microseconds.
int sptrsv_batch_time_config(sptrsv_batch_t B, int cfg, int precision, const double* d_b, double* d_x, int m,
                             double* sys_us) {
    if (!B || !sys_us || cfg < 0 || cfg >= SPTRSV_BCFG_COUNT) return fail(SPTRSV_ERR_BAD_INPUT, "bad timing arguments");
    sptrsv_plan_s* P = B->ctl;
    std::vector<int> all((size_t)B->nsys, cfg);
    SPT_HIP(hipMemcpy(B->cfg, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice));
    int rc = SPTRSV_OK;
    if (precision == SPTRSV_FP64) {
        rc = launch_batch<double>(B, B->val, d_b, d_x, m, B->timer);
    } else {
        rc = sptrsv_batch_solve(B, SPTRSV_FP32, d_b, d_x, m, 0);  // FP32 copy and staging, warm up
        if (rc == SPTRSV_OK) rc = launch_batch<float>(B, B->val32, P->b32, P->x32, m, B->timer);
    }
    if (rc == SPTRSV_OK) rc = sptrsv_plan_sync(P);
    std::vector<unsigned long long> t(2 * (size_t)B->nsys);
    if (rc == SPTRSV_OK &&
        hipMemcpy(t.data(), B->timer, t.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(SPTRSV_ERR_HIP, "timer download");
    const double us_per_tick = 1000.0 / (double)P->ticks_per_ms;
    for (int s = 0; s < B->nsys && rc == SPTRSV_OK; ++s) sys_us[s] = (double)(t[2 * s + 1] - t[2 * s]) * us_per_tick;
    // restore the configured choice
    if (hipMemcpy(B->cfg, B->h_cfg.data(), B->h_cfg.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess &&
        rc == SPTRSV_OK)
        rc = fail(SPTRSV_ERR_HIP, "config restore");
    return rc;
}

}  // extern "C"
