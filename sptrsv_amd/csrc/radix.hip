// radix.hip -- the analysis's sorting and scanning: a stable LSD radix sort of (key, value)
// int pairs and prefix sums of ints, plus the pointer array of a sorted key list.
//
// Sort: 8-bit digits, one pass per digit of the keys' width.  A pass is three launches over
// tiles of 4,096 items (256 threads x 16): a digit histogram per tile (LDS atomics), an
// exclusive scan of the digit-major histogram (digit d of tile t lands after every smaller digit
// and after digit d of every earlier tile), and the scatter.  The scatter ranks each item among
// the items of its tile with the same digit and a smaller index: a wave holds 64 consecutive
// items per round, eight ballots give each lane the mask of lanes with its digit, the lanes
// below it in that mask are its rank inside the round, a per-wave LDS counter per digit carries
// the rounds, and a prefix over the four waves' counters orders the waves.  The order of equal
// keys is therefore the input order, in every pass: the sort is stable, and its output is the
// unique stably sorted sequence (bit-exact by construction).
#include <hip/hip_runtime.h>

#include "common.h"
#include "plan.h"

namespace sptrsv {

namespace {

constexpr int kST = 256;              // threads per workgroup
constexpr int kSI = 16;               // items per thread
constexpr int kTile = kST * kSI;      // items per tile
constexpr int kWavesPerTile = kST / kWave;

// Inclusive or exclusive prefix sums of one tile of 4,096 ints (16 consecutive per thread); the
// tile's total to sums[tile] when sums is given.  in may equal out.
template <bool INCL>
__global__ __launch_bounds__(kST) void k_scan_tile(const int* in, int* out, int64_t count, int* sums) {
    __shared__ int wsum[kWavesPerTile];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
    const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)tid * kSI;
    int v[kSI];
    int tot = 0;
#pragma unroll
    for (int j = 0; j < kSI; ++j) {
        v[j] = base + j < count ? in[base + j] : 0;
        tot += v[j];
    }
    int x = tot;  // inclusive scan of the thread totals across the wave
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) wsum[w] = x;
    __syncthreads();
    int run = x - tot;
    for (int k = 0; k < w; ++k) run += wsum[k];
    if (sums && tid == kST - 1) sums[blockIdx.x] = run + tot;
#pragma unroll
    for (int j = 0; j < kSI; ++j) {
        if (INCL) run += v[j];
        if (base + j < count) out[base + j] = run;
        if (!INCL) run += v[j];
    }
}

// out[i] += offs[tile of i] (the exclusive scan of the tile totals)
__global__ __launch_bounds__(kST) void k_scan_add(int* out, int64_t count, const int* offs) {
    const int64_t base = (int64_t)blockIdx.x * kTile;
    const int add = offs[blockIdx.x];
    for (int j = threadIdx.x; j < kTile; j += kST)
        if (base + j < count) out[base + j] += add;
}

// hist[d * ntiles + tile] = items of the tile whose digit (key >> shift) & 255 is d
__global__ __launch_bounds__(kST) void k_radix_count(const int* __restrict__ keys, int64_t N, int shift, int ntiles,
                                                     int* __restrict__ hist) {
    __shared__ int bins[256];
    bins[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kTile + threadIdx.x;
#pragma unroll 4
    for (int j = 0; j < kSI; ++j) {
        const int64_t i = base + (int64_t)j * kST;
        if (i < N) atomicAdd(&bins[(keys[i] >> shift) & 255], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = bins[threadIdx.x];
}

// Stable scatter of one tile by digit.  Item i = tile * 4096 + wave * 1024 + round * 64 + lane.
// vals == nullptr: the value of item i is i.
__global__ __launch_bounds__(kST) void k_radix_scatter(const int* __restrict__ keys, const int* __restrict__ vals,
                                                       int64_t N, int shift, int ntiles, const int* __restrict__ offs,
                                                       int* __restrict__ keys_out, int* __restrict__ vals_out) {
    __shared__ int cnt[kWavesPerTile][256];  // per wave and digit: items ranked so far
    __shared__ int gbase[256];               // first output slot of each digit for this tile
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
#pragma unroll
    for (int k = 0; k < kWavesPerTile; ++k) cnt[k][tid] = 0;
    gbase[tid] = offs[(int64_t)tid * ntiles + blockIdx.x];
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)w * (kWave * kSI) + lane;
    const uint64_t below = (1ull << lane) - 1ull;
    int key[kSI], val[kSI], dig[kSI], rnk[kSI];
#pragma unroll
    for (int j = 0; j < kSI; ++j) {
        const int64_t i = base + (int64_t)j * kWave;
        const bool ok = i < N;
        key[j] = ok ? keys[i] : 0;
        val[j] = ok ? (vals ? vals[i] : (int)i) : 0;
        const int d = (key[j] >> shift) & 255;
        dig[j] = d;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        // every lane reads its digit's count, then the lowest lane of each digit adds the round's
        // items of that digit (one wave: its LDS operations complete in program order)
        const int before = cnt[w][d];
        __builtin_amdgcn_wave_barrier();
        if (ok && (peers & below) == 0) cnt[w][d] = before + (int)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rnk[j] = before + (int)__popcll(peers & below);
    }
    __syncthreads();
    {  // waves in order: digit d of wave k starts after digit d of waves 0..k-1
        int run = 0;
#pragma unroll
        for (int k = 0; k < kWavesPerTile; ++k) {
            const int t = cnt[k][tid];
            cnt[k][tid] = run;
            run += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSI; ++j) {
        const int64_t i = base + (int64_t)j * kWave;
        if (i < N) {
            const int pos = gbase[dig[j]] + cnt[w][dig[j]] + rnk[j];
            keys_out[pos] = key[j];
            vals_out[pos] = val[j];
        }
    }
}

// ptr[c] = first position whose key - base is >= c, for c in [0, K]; keys sorted ascending
__global__ void k_key_bounds(const int* __restrict__ sorted, int64_t N, int K, int base, int* __restrict__ ptr) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= N; i += (int64_t)gridDim.x * blockDim.x) {
        const int prev = i == 0 ? -1 : sorted[i - 1] - base;
        const int cur = i == N ? K : sorted[i] - base;
        for (int c = prev + 1; c <= cur; ++c) ptr[c] = (int)i;
    }
}

int64_t tiles_of(int64_t count) { return (count + kTile - 1) / kTile; }

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

size_t scan_scratch_bytes(int64_t count) {
    size_t b = 0;
    for (int64_t nb = tiles_of(count); nb > 1; nb = tiles_of(nb)) b += align256((size_t)nb * sizeof(int));
    return b + 256;
}

int scan_ints(const int* in, int* out, int64_t count, bool inclusive, void* scratch, hipStream_t s) {
    if (count <= 0) return SPTRSV_OK;
    const int64_t nb = tiles_of(count);
    int* sums = nb > 1 ? static_cast<int*>(scratch) : nullptr;
    if (inclusive)
        k_scan_tile<true><<<(unsigned)nb, kST, 0, s>>>(in, out, count, sums);
    else
        k_scan_tile<false><<<(unsigned)nb, kST, 0, s>>>(in, out, count, sums);
    SPT_HIP(hipGetLastError());
    if (nb > 1) {
        char* rest = static_cast<char*>(scratch) + align256((size_t)nb * sizeof(int));
        SPT_CHECK(scan_ints(sums, sums, nb, false, rest, s));
        k_scan_add<<<(unsigned)nb, kST, 0, s>>>(out, count, sums);
        SPT_HIP(hipGetLastError());
    }
    return SPTRSV_OK;
}

size_t sort_scratch_bytes(int64_t N) {
    const int64_t hist = 256 * std::max<int64_t>(tiles_of(N), 1);
    return 2 * align256((size_t)std::max<int64_t>(N, 1) * sizeof(int)) + align256((size_t)hist * sizeof(int)) +
           scan_scratch_bytes(hist);
}

int sort_pairs(const int* keys_in, const int* vals_in, int64_t N, int bits, int* keys_out, int* vals_out,
               void* scratch, hipStream_t s) {
    if (N <= 0) return SPTRSV_OK;
    if (N >= ((int64_t)1 << 31)) return fail(SPTRSV_ERR_UNSUPPORTED, "sort: 2^31 items or more");
    const int passes = std::max(1, (bits + 7) / 8);
    const int64_t nt = tiles_of(N);
    char* at = static_cast<char*>(scratch);
    int* kalt = reinterpret_cast<int*>(at);
    at += align256((size_t)N * sizeof(int));
    int* valt = reinterpret_cast<int*>(at);
    at += align256((size_t)N * sizeof(int));
    int* hist = reinterpret_cast<int*>(at);
    at += align256((size_t)(256 * nt) * sizeof(int));
    const int* ks = keys_in;
    const int* vs = vals_in;
    for (int pass = 0; pass < passes; ++pass) {
        // the last pass writes the outputs; earlier ones alternate so no pass reads what it writes
        const bool to_out = (passes - 1 - pass) % 2 == 0;
        int* kd = to_out ? keys_out : kalt;
        int* vd = to_out ? vals_out : valt;
        k_radix_count<<<(unsigned)nt, kST, 0, s>>>(ks, N, 8 * pass, (int)nt, hist);
        SPT_HIP(hipGetLastError());
        SPT_CHECK(scan_ints(hist, hist, 256 * nt, false, at, s));
        k_radix_scatter<<<(unsigned)nt, kST, 0, s>>>(ks, vs, N, 8 * pass, (int)nt, hist, kd, vd);
        SPT_HIP(hipGetLastError());
        ks = kd;
        vs = vd;
    }
    return SPTRSV_OK;
}

int key_bounds(const int* sorted, int64_t N, int K, int base, int* ptr, hipStream_t s) {
    const int64_t g = std::min<int64_t>((N + 1 + 255) / 256, 65536);
    k_key_bounds<<<(unsigned)g, 256, 0, s>>>(sorted, N, K, base, ptr);
    SPT_HIP(hipGetLastError());
    return SPTRSV_OK;
}

}  // namespace sptrsv
