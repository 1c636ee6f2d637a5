// refine.hip -- FP64 iterative refinement around reduced-precision solves, and the 1-norm
// condition estimate that tells when that refinement can work.
//
// Refinement (sptrsv_solve_refined_ex):  x_0 = solve_low(b);  repeat  r = b - U x  (FP64
// products, FP64 accumulation, rounded once to the inner precision),  d = solve_low(r),
// x += d  (FP64), with the inf-norms of d and x reduced on the device.  The inner solve
// contracts the error by about kappa(U) * u_low per step, so it converges only while that is
// below 1: the loop reports why it stopped (converged / residual floor / max_iters /
// diverged), and with opts.fallback a run that did not converge is finished by the FP64 solve.
//
// Condition estimate (sptrsv_condest):  kappa_1(U) = ||U||_1 ||U^-1||_1, the second factor
// from Hager's method in Higham's form (the iteration of LAPACK's xLACON, N. J. Higham, ACM
// TOMS 14 (1988) 381-396, Algorithm 4.1): at most five pairs of solves with U and U^T, then
// the alternating test vector x_i = (-1)^i (1 + i / (n - 1)).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>
#include <string>
#include <vector>

#include "common.h"
#include "plan.h"

namespace sptrsv {
namespace {

inline int grid_for_r(int64_t count, int block = 256, int cap = 65536) {
    int64_t g = (count + block - 1) / block;
    return (int)std::min<int64_t>(std::max<int64_t>(g, 1), cap);
}

// max that keeps NaN: once a NaN is seen it stays (fmax would drop it)
__device__ __forceinline__ double max_keep_nan(double acc, double v) {
    return (v > acc || v != v) && acc == acc ? v : acc;
}

// x += d (FP64) and per-column max|d|, max|x|.  Non-negative doubles order like their bit
// patterns, and a positive NaN's pattern is above +inf's, so the bitwise atomicMax keeps a NaN
// as well.  grid.y covers columns in groups of 64; 4 row-waves per block reduce through LDS.
__global__ __launch_bounds__(256) void k_update(int n, int m, const float* __restrict__ d, double* __restrict__ x,
                                                unsigned long long* __restrict__ norms, int first) {
    __shared__ double sd[4][64], sx[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = blockIdx.y * 64 + lane;
    double md = 0.0, mx = 0.0;
    if (k < m) {
        for (int row = blockIdx.x * 4 + wv; row < n; row += gridDim.x * 4) {
            const size_t at = (size_t)row * m + k;
            const double dv = (double)d[at];
            const double xv = first ? dv : x[at] + dv;
            x[at] = xv;
            md = max_keep_nan(md, fabs(dv));
            mx = max_keep_nan(mx, fabs(xv));
        }
    }
    sd[wv][lane] = md;
    sx[wv][lane] = mx;
    __syncthreads();
    if (wv == 0 && k < m) {
        md = max_keep_nan(max_keep_nan(sd[0][lane], sd[1][lane]), max_keep_nan(sd[2][lane], sd[3][lane]));
        mx = max_keep_nan(max_keep_nan(sx[0][lane], sx[1][lane]), max_keep_nan(sx[2][lane], sx[3][lane]));
        atomicMax(&norms[k], __builtin_bit_cast(unsigned long long, fabs(md)));
        atomicMax(&norms[m + k], __builtin_bit_cast(unsigned long long, fabs(mx)));
    }
}

__global__ void k_zero_u64(unsigned long long* p, int count) {
    for (int i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0ull;
}

// ||U||_1: one thread per CSC column sums |u_ij|; the maximum through the bit patterns
__global__ __launch_bounds__(256) void k_norm1(int n, const int* __restrict__ colptr, const double* __restrict__ cval,
                                               unsigned long long* __restrict__ out) {
    double best = 0.0;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int e = colptr[j]; e < colptr[j + 1]; ++e) s += fabs(cval[e]);
        best = max_keep_nan(best, s);
    }
    for (int off = 32; off > 0; off >>= 1) best = max_keep_nan(best, __shfl_xor(best, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __builtin_bit_cast(unsigned long long, fabs(best)));
}

int ensure_refine_buffers(sptrsv_plan_s* P, int64_t nm, int m) {
    if ((int64_t)P->ref_elems >= nm && (int)P->ref_cols >= m) return SPTRSV_OK;
    SPT_HIP(hipStreamSynchronize(P->stream));
    if (P->ref_r32) SPT_HIP(hipFree(P->ref_r32));
    if (P->ref_d32) SPT_HIP(hipFree(P->ref_d32));
    if (P->ref_norms) SPT_HIP(hipFree(P->ref_norms));
    P->ref_r32 = P->ref_d32 = nullptr;
    P->ref_norms = nullptr;
    P->ref_elems = P->ref_cols = 0;
    SPT_HIP(hipMalloc((void**)&P->ref_r32, nm * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&P->ref_d32, nm * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&P->ref_norms, 2 * (size_t)m * sizeof(unsigned long long)));
    P->ref_elems = (size_t)nm;
    P->ref_cols = (size_t)m;
    return SPTRSV_OK;
}

// largest ||d_k|| / ||x_k|| over the columns; +inf when a norm is not finite
int read_ratio(sptrsv_plan_s* P, int m, std::vector<unsigned long long>& h, double* ratio) {
    SPT_CHECK(sptrsv_plan_sync(P));
    SPT_HIP(hipMemcpy(h.data(), P->ref_norms, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    double last = 0.0;
    for (int k = 0; k < m; ++k) {
        const double dn = __builtin_bit_cast(double, h[(size_t)k]);
        const double xn = __builtin_bit_cast(double, h[(size_t)m + k]);
        if (!std::isfinite(dn) || !std::isfinite(xn)) {
            *ratio = std::numeric_limits<double>::infinity();
            return SPTRSV_OK;
        }
        last = std::max(last, xn > 0 ? dn / xn : dn);
    }
    *ratio = last;
    return SPTRSV_OK;
}

}  // namespace

// r = b - U x over the plan's CSR in FP64, stored FP32 (any factor; one thread per (row, rhs)).
int residual32(sptrsv_plan_s* P, const double* d_b, const double* d_x, float* d_r32, int m) {
    const int64_t nm = (int64_t)P->n * m;
    launch_residual<int, float>(P->n, P->nnz, m, P->rowptr, P->col, P->val, d_x, d_b, d_r32, P->stream);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

// FP64 iterative refinement around a reduced-precision solve (see sptrsv_solve_refined).
// Stopping rule: correction k has relative size r_k ~ the error of x before it; with the
// observed contraction c = r_k / r_{k-1} the error left after it is about r_k * c.  Stop when
// that estimate is <= tol (converged), when a correction or x is not finite or corrections grow
// (diverged), or when they shrink by less than half (the FP64 residual floor).
int refine_loop(sptrsv_plan_s* P, const sptrsv_refine_t& ro, const double* d_b, double* d_x, int m,
                const std::function<int(const float*, float*)>& solve_low,
                const std::function<int(const double*, const double*, float*)>& residual, RefineResult* res) {
    const int max_iters = ro.max_iters > 0 ? ro.max_iters : 30;
    const double tol = ro.tol > 0 ? ro.tol : 1e-13;
    const int n = P->n;
    const int64_t nm = (int64_t)n * m;
    hipStream_t s = P->stream;
    SPT_CHECK(ensure_refine_buffers(P, nm, m));
    const dim3 ugrid((unsigned)std::min<int64_t>(std::max<int64_t>((n + 3) / 4, 1), 4096), (unsigned)((m + 63) / 64));
    std::vector<unsigned long long> h((size_t)2 * m);
    SPT_CHECK(convert_d2f(d_b, P->ref_r32, nm, s));
    SPT_CHECK(solve_low(P->ref_r32, P->ref_d32));
    k_zero_u64<<<1, 256, 0, s>>>(P->ref_norms, 2 * m);
    k_update<<<ugrid, 256, 0, s>>>(n, m, P->ref_d32, d_x, P->ref_norms, 1);
    SPT_HIP(hipGetLastError());
    double x0 = 0.0;
    SPT_CHECK(read_ratio(P, m, h, &x0));  // d = x here: only finiteness matters
    RefineResult r{0, SPTRSV_REFINE_MAX_ITERS, 1.0};
    if (!std::isfinite(x0)) {
        r.outcome = SPTRSV_REFINE_DIVERGED;
        r.estimate = std::numeric_limits<double>::infinity();
        *res = r;
        return SPTRSV_OK;
    }
    double prev = 1.0;
    for (int it = 1; it <= max_iters; ++it) {
        SPT_CHECK(residual(d_b, d_x, P->ref_r32));
        SPT_CHECK(solve_low(P->ref_r32, P->ref_d32));
        k_zero_u64<<<1, 256, 0, s>>>(P->ref_norms, 2 * m);
        k_update<<<ugrid, 256, 0, s>>>(n, m, P->ref_d32, d_x, P->ref_norms, 0);
        SPT_HIP(hipGetLastError());
        double last = 0.0;
        SPT_CHECK(read_ratio(P, m, h, &last));
        r.iters = it;
        if (!std::isfinite(last)) {
            r.outcome = SPTRSV_REFINE_DIVERGED;
            r.estimate = last;
            break;
        }
        r.estimate = last * std::min(1.0, last / prev);
        if (!(r.estimate > tol)) {
            r.outcome = SPTRSV_REFINE_CONVERGED;
            break;
        }
        if (it >= 2 && last > prev) {
            r.outcome = SPTRSV_REFINE_DIVERGED;
            break;
        }
        if (it >= 2 && last > 0.5 * prev) {
            r.outcome = SPTRSV_REFINE_FLOOR;
            break;
        }
        prev = last;
    }
    *res = r;
    return SPTRSV_OK;
}

// Variant of the reduced-precision solves: AUTO takes the FP32 choice (never the FP64-only
// tile and block kernels).
int refine_variant(sptrsv_plan_s* P, int requested, int m, int low) {
    return requested == SPTRSV_VAR_AUTO ? default_variant(P, m, low, false) : requested;
}

}  // namespace sptrsv

using namespace sptrsv;

extern "C" {

int sptrsv_solve_refined_ex(sptrsv_plan_t P, const sptrsv_params_t* params, const sptrsv_refine_t* opts,
                            const double* d_b, double* d_x, int m, sptrsv_refine_info_t* info) {
    if (!P || !d_b || !d_x) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    if (m < 1) return fail(SPTRSV_ERR_BAD_INPUT, "m must be >= 1");
    sptrsv_params_t prm{};
    if (params) prm = *params;
    sptrsv_refine_t ro{};
    if (opts) ro = *opts;
    const int low = ro.precision == 0 ? SPTRSV_FP32 : ro.precision;
    const int variant = refine_variant(P, prm.variant, m, low);
    if (variant == SPTRSV_VAR_TILE || variant == SPTRSV_VAR_BLOCK)
        return fail(SPTRSV_ERR_UNSUPPORTED, "refinement: the tile and block variants have no reduced-precision form");
    if (variant == SPTRSV_VAR_CHAIN && m != 1)
        return fail(SPTRSV_ERR_UNSUPPORTED, "refinement: the chain variant solves one right-hand side");
    if (variant == SPTRSV_VAR_SUPERNODAL ? low != SPTRSV_BF16 && low != SPTRSV_FP32 : low != SPTRSV_FP32)
        return fail(SPTRSV_ERR_UNSUPPORTED,
                    "refinement: BF16 inner solves need the supernodal variant, the others run FP32");
    const uint64_t limit = prm.watchdog_us > 0 ? std::max<uint64_t>(1, P->ticks_per_ms * (uint64_t)prm.watchdog_us / 1000)
                                               : P->ticks_per_ms * 20000ull;
    if (variant != SPTRSV_VAR_SUPERNODAL) SPT_CHECK(ensure_fp32(P));
    sptrsv_params_t low_prm = prm;
    low_prm.variant = variant;
    low_prm.precision = SPTRSV_FP32;
    auto solve_low = [&](const float* in, float* out) -> int {
        if (variant == SPTRSV_VAR_SUPERNODAL) return supernodal_solve_low(P, low, in, out, m, limit);
        return solve_fp32_device(P, &low_prm, in, out, m);
    };
    auto residual = [&](const double* b, const double* x, float* r) -> int {
        if (variant == SPTRSV_VAR_SUPERNODAL) return supernodal_residual32(P, b, x, r, m);
        return residual32(P, b, x, r, m);
    };
    RefineResult res{};
    SPT_CHECK(refine_loop(P, ro, d_b, d_x, m, solve_low, residual, &res));
    int fell_back = 0;
    if (res.outcome != SPTRSV_REFINE_CONVERGED && ro.fallback) {
        sptrsv_params_t fb = prm;
        fb.precision = SPTRSV_FP64;
        fb.variant = variant == SPTRSV_VAR_SUPERNODAL ? SPTRSV_VAR_SUPERNODAL : SPTRSV_VAR_AUTO;
        SPT_CHECK(sptrsv_solve(P, &fb, d_b, d_x, m));
        SPT_CHECK(sptrsv_plan_sync(P));
        fell_back = 1;
    }
    if (info) {
        info->iters = res.iters;
        info->outcome = res.outcome;
        info->fell_back = fell_back;
        info->variant = variant;
        info->estimate = res.estimate;
    }
    return SPTRSV_OK;
}

int sptrsv_solve_refined(sptrsv_plan_t P, const sptrsv_params_t* params, const sptrsv_refine_t* opts,
                         const double* d_b, double* d_x, int m, int* iters, double* ratio) {
    sptrsv_refine_info_t info{};
    SPT_CHECK(sptrsv_solve_refined_ex(P, params, opts, d_b, d_x, m, &info));
    if (iters) *iters = info.iters;
    if (ratio) *ratio = info.estimate;
    return SPTRSV_OK;
}

int sptrsv_condest(sptrsv_plan_t P, double* kappa, double* norm_inv) {
    if (!P) return fail(SPTRSV_ERR_BAD_INPUT, "null plan");
    const int n = P->n;
    hipStream_t s = P->stream;
    double *dx = nullptr, *dy = nullptr;
    unsigned long long* dn = nullptr;
    SPT_HIP(hipMalloc((void**)&dx, (size_t)n * sizeof(double)));
    if (hipMalloc((void**)&dy, (size_t)n * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&dn, sizeof(unsigned long long)) != hipSuccess) {
        (void)hipFree(dx);
        (void)hipFree(dy);
        return fail(SPTRSV_ERR_NOMEM, "condest work vectors");
    }
    std::vector<double> x((size_t)n), y((size_t)n);
    sptrsv_params_t prm{};  // AUTO, FP64
    int rc = SPTRSV_OK;
    // y = U^-1 x (transposed = false) or U^-T x
    auto apply = [&](bool transposed) -> int {
        if (hipMemcpy(dx, x.data(), x.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return fail(SPTRSV_ERR_HIP, "condest upload");
        SPT_CHECK(transposed ? sptrsv_solve_transposed(P, &prm, dx, dy, 1) : sptrsv_solve(P, &prm, dx, dy, 1));
        SPT_CHECK(sptrsv_plan_sync(P));
        if (hipMemcpy(y.data(), dy, y.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            return fail(SPTRSV_ERR_HIP, "condest download");
        return SPTRSV_OK;
    };
    auto asum = [&](const std::vector<double>& v) {
        double t = 0.0;
        for (double e : v) t += std::fabs(e);
        return t;
    };
    auto argmax_abs = [&](const std::vector<double>& v) {
        size_t j = 0;
        for (size_t i = 1; i < v.size(); ++i)
            if (std::fabs(v[i]) > std::fabs(v[j])) j = i;
        return j;
    };
    double est = 0.0;
    do {
        std::fill(x.begin(), x.end(), 1.0 / n);
        if ((rc = apply(false))) break;
        est = asum(y);
        if (n == 1) break;
        std::vector<int8_t> sgn((size_t)n);
        for (int i = 0; i < n; ++i) {
            sgn[(size_t)i] = y[(size_t)i] >= 0 ? 1 : -1;
            x[(size_t)i] = sgn[(size_t)i];
        }
        if ((rc = apply(true))) break;
        size_t j = argmax_abs(y);
        for (int iter = 2; iter <= 5; ++iter) {
            std::fill(x.begin(), x.end(), 0.0);
            x[j] = 1.0;
            if ((rc = apply(false))) break;
            const double old = est;
            est = asum(y);
            bool same = true;
            for (int i = 0; i < n && same; ++i) same = (y[(size_t)i] >= 0 ? 1 : -1) == sgn[(size_t)i];
            if (same || est <= old) break;
            for (int i = 0; i < n; ++i) {
                sgn[(size_t)i] = y[(size_t)i] >= 0 ? 1 : -1;
                x[(size_t)i] = sgn[(size_t)i];
            }
            if ((rc = apply(true))) break;
            const size_t jlast = j;
            j = argmax_abs(y);
            if (y[jlast] == std::fabs(y[j])) break;  // the maximum did not move
        }
        if (rc) break;
        // alternating test vector: catches matrices for which the iteration stalls early
        for (int i = 0; i < n; ++i) x[(size_t)i] = (i % 2 ? -1.0 : 1.0) * (1.0 + (double)i / (n - 1));
        if ((rc = apply(false))) break;
        est = std::max(est, 2.0 * asum(y) / (3.0 * n));
    } while (false);
    double norm1 = 0.0;
    if (rc == SPTRSV_OK) {
        k_zero_u64<<<1, 64, 0, s>>>(dn, 1);
        k_norm1<<<grid_for_r(n, 256, 4096), 256, 0, s>>>(n, P->colptr, P->cval, dn);
        unsigned long long bits = 0;
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess ||
            hipMemcpy(&bits, dn, sizeof(bits), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(SPTRSV_ERR_HIP, "condest norm");
        norm1 = __builtin_bit_cast(double, bits);
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    (void)hipFree(dn);
    if (rc != SPTRSV_OK) return rc;
    if (kappa) *kappa = norm1 * est;
    if (norm_inv) *norm_inv = est;
    return SPTRSV_OK;
}

}  // extern "C"
