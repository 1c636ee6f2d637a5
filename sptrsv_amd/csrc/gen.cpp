// gen.cpp -- synthetic batches of independent upper-triangular systems, generated natively.
//
// The batched-solve workload holds thousands of systems with up to 65,536 rows each, about five
// billion nonzeros in all: too many to draw in Python.  This generator draws every system from
// its own random streams (xoshiro256**, seeded through SplitMix64 from (seed, system, stream)),
// so systems are generated in parallel and independently of each other, in two passes:
//   sptrsv_gen_batch_sizes  rows, nonzeros and kind of every system (structure draws only);
//   sptrsv_gen_batch_fill   the block-diagonal CSR into caller-allocated arrays (the same
//                           structure draws, then the value draws from a second stream).
// System kinds (equally likely, n ~ log-uniform[nmin, nmax]):
//   banded      bandwidth ~ U{4..64}: row i holds columns i .. min(i + bw, n - 1);
//   zipf        per row Poisson(8) draws of column offsets ~ Zipf(1.5) capped at n - 1 - i,
//               zero offsets and duplicates dropped;
//   supernodal  S = round(n / 72) supernodes of widths ~ U{16..128}, parent of s ~ U{s+1..S-1},
//               each with dense blocks to 1..6 distinct random ancestors on its path to the root.
// Values: off-diagonal U(-1,1) and diagonal 1 + sum |off-diagonal| (banded, zipf); supernodal
// triangles with diagonal U[1,2] and strict upper part N(0, 0.1), rectangular blocks N(0, 0.1).
// Diagonal first in every row, columns ascending, columns global (row offset of the system).
// Distributions follow published methods: Poisson by Knuth's product of uniforms, Zipf by
// Devroye's rejection (Non-Uniform Random Variate Generation, 1986, X.6.1), normal by
// Marsaglia's polar method.
#include <omp.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "sptrsv_capi.h"
#include "sptrsv_gen.h"

namespace {

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t system, uint64_t stream) {
        uint64_t x = seed * 0x2545F4914F6CDD1Dull ^ (system + 1) * 0x9E3779B97F4A7C15ull ^ (stream + 1) * 0xD1B54A32D192ED03ull;
        for (auto& w : s) w = splitmix(x);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {  // xoshiro256**
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }              // [0, 1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    int integer(int lo, int hi) { return lo + (int)((next() >> 33) % (uint64_t)(hi - lo + 1)); }  // [lo, hi]
    int poisson(double lam) {
        const double L = std::exp(-lam);
        int k = 0;
        double p = 1.0;
        do {
            ++k;
            p *= uniform();
        } while (p > L);
        return k - 1;
    }
    int64_t zipf(double a) {
        const double am1 = a - 1.0, b = std::pow(2.0, am1);
        for (;;) {
            const double u = 1.0 - uniform(), v = uniform();  // u in (0, 1]
            const double xf = std::floor(std::pow(u, -1.0 / am1));
            if (xf > 4.0e18 || xf < 1.0) continue;
            const double t = std::pow(1.0 + 1.0 / xf, am1);
            if (v * xf * (t - 1.0) / (b - 1.0) <= t / b) return (int64_t)xf;
        }
    }
    double normal(double sigma) {
        for (;;) {
            const double x = 2.0 * uniform() - 1.0, y = 2.0 * uniform() - 1.0, r = x * x + y * y;
            if (r > 0.0 && r < 1.0) return sigma * x * std::sqrt(-2.0 * std::log(r) / r);
        }
    }
};

// Structure of one system, drawn from its structure stream.
struct System {
    int kind = 0;
    int n = 0;
    int bw = 0;                                // banded
    std::vector<int> zrp, zcol;                // zipf: local CSR of the off-diagonals
    std::vector<int> start;                    // supernodal: S + 1 boundaries
    std::vector<std::vector<int>> anc;         // supernodal: chosen ancestors, ascending
    std::vector<int> rect;                     // supernodal: rectangular width per supernode
    int64_t nnz = 0;
};

void draw_structure(System& sys, uint64_t seed, int s, int nmin, int nmax) {
    Rng g(seed, (uint64_t)s, 0);
    const double ln0 = std::log((double)nmin), ln1 = std::log((double)nmax);
    int n = (int)std::exp(g.uniform(ln0, ln1));
    n = std::max(nmin, std::min(nmax, n));
    sys.kind = g.integer(0, 2);
    if (sys.kind == SPTRSV_GEN_BANDED) {
        sys.n = n;
        sys.bw = g.integer(4, 64);
        int64_t t = 0;
        for (int i = 0; i < n; ++i) t += std::min<int64_t>(sys.bw + 1, n - i);
        sys.nnz = t;
    } else if (sys.kind == SPTRSV_GEN_ZIPF) {
        sys.n = n;
        sys.zrp.assign((size_t)n + 1, 0);
        sys.zcol.clear();
        std::vector<int> row;
        for (int i = 0; i < n; ++i) {
            const int c = g.poisson(8.0);
            row.clear();
            for (int k = 0; k < c; ++k) {
                const int64_t off = std::min<int64_t>(g.zipf(1.5), (int64_t)(n - 1 - i));
                if (off > 0) row.push_back(i + (int)off);
            }
            std::sort(row.begin(), row.end());
            row.erase(std::unique(row.begin(), row.end()), row.end());
            sys.zcol.insert(sys.zcol.end(), row.begin(), row.end());
            sys.zrp[(size_t)i + 1] = (int)sys.zcol.size();
        }
        sys.nnz = (int64_t)n + (int64_t)sys.zcol.size();
    } else {
        const int S = std::max(1, (int)std::lround(n / 72.0));
        sys.start.assign((size_t)S + 1, 0);
        for (int k = 0; k < S; ++k) sys.start[(size_t)k + 1] = sys.start[(size_t)k] + g.integer(16, 128);
        std::vector<int> parent((size_t)S, -1);
        for (int k = 0; k + 1 < S; ++k) parent[(size_t)k] = g.integer(k + 1, S - 1);
        sys.anc.assign((size_t)S, {});
        sys.rect.assign((size_t)S, 0);
        std::vector<int> path;
        int64_t t = 0;
        for (int k = 0; k < S; ++k) {
            path.clear();
            for (int p = parent[(size_t)k]; p >= 0; p = parent[(size_t)p]) path.push_back(p);
            const int want = g.integer(1, 6), take = std::min<int>(want, (int)path.size());
            for (int q = 0; q < take; ++q) std::swap(path[(size_t)q], path[(size_t)g.integer(q, (int)path.size() - 1)]);
            std::vector<int> chosen(path.begin(), path.begin() + take);
            std::sort(chosen.begin(), chosen.end());
            int r = 0;
            for (int a : chosen) r += sys.start[(size_t)a + 1] - sys.start[(size_t)a];
            sys.anc[(size_t)k] = std::move(chosen);
            sys.rect[(size_t)k] = r;
            const int64_t w = sys.start[(size_t)k + 1] - sys.start[(size_t)k];
            t += w * (w + 1) / 2 + w * r;
        }
        sys.n = sys.start[(size_t)S];
        sys.nnz = t;
    }
}

// Writes system `s` at rows [r0, r0 + n), entries [e0, e0 + nnz) of the batch CSR.
void fill_system(const System& sys, uint64_t seed, int s, int r0, int64_t e0, int64_t* rowptr, int* col, double* val) {
    Rng g(seed, (uint64_t)s, 1);
    int64_t e = e0;
    const int n = sys.n;
    if (sys.kind == SPTRSV_GEN_BANDED || sys.kind == SPTRSV_GEN_ZIPF) {
        for (int i = 0; i < n; ++i) {
            rowptr[r0 + i] = e;
            const int64_t d = e++;
            col[d] = r0 + i;
            double sum = 0.0;
            if (sys.kind == SPTRSV_GEN_BANDED) {
                const int last = std::min(i + sys.bw, n - 1);
                for (int j = i + 1; j <= last; ++j, ++e) {
                    col[e] = r0 + j;
                    val[e] = g.uniform(-1.0, 1.0);
                    sum += std::fabs(val[e]);
                }
            } else {
                for (int q = sys.zrp[(size_t)i]; q < sys.zrp[(size_t)i + 1]; ++q, ++e) {
                    col[e] = r0 + sys.zcol[(size_t)q];
                    val[e] = g.uniform(-1.0, 1.0);
                    sum += std::fabs(val[e]);
                }
            }
            val[d] = 1.0 + sum;
        }
    } else {
        const int S = (int)sys.anc.size();
        for (int k = 0; k < S; ++k) {
            const int s0 = sys.start[(size_t)k], w = sys.start[(size_t)k + 1] - s0;
            for (int d = 0; d < w; ++d) {
                rowptr[r0 + s0 + d] = e;
                col[e] = r0 + s0 + d;
                val[e++] = g.uniform(1.0, 2.0);
                for (int j = d + 1; j < w; ++j, ++e) {
                    col[e] = r0 + s0 + j;
                    val[e] = g.normal(0.1);
                }
                for (int a : sys.anc[(size_t)k])
                    for (int c = sys.start[(size_t)a]; c < sys.start[(size_t)a + 1]; ++c, ++e) {
                        col[e] = r0 + c;
                        val[e] = g.normal(0.1);
                    }
            }
        }
    }
}

}  // namespace

extern "C" {

int sptrsv_gen_batch_sizes(int count, int first, uint64_t seed, int nmin, int nmax, int* n, int64_t* nnz, int* kind) {
    if (count <= 0 || first < 0 || nmin < 1 || nmax < nmin || !n || !nnz || !kind) return SPTRSV_ERR_BAD_INPUT;
#pragma omp parallel for schedule(dynamic, 8)
    for (int s = 0; s < count; ++s) {
        System sys;
        draw_structure(sys, seed, first + s, nmin, nmax);
        n[s] = sys.n;
        nnz[s] = sys.nnz;
        kind[s] = sys.kind;
    }
    return SPTRSV_OK;
}

int sptrsv_gen_batch_fill(int count, int first, uint64_t seed, int nmin, int nmax, const int* n, const int64_t* nnz,
                          int64_t* rowptr, int* col, double* val) {
    if (count <= 0 || first < 0 || nmin < 1 || nmax < nmin || !n || !nnz || !rowptr || !col || !val)
        return SPTRSV_ERR_BAD_INPUT;
    std::vector<int64_t> r0((size_t)count + 1, 0), e0((size_t)count + 1, 0);
    for (int s = 0; s < count; ++s) {
        r0[(size_t)s + 1] = r0[(size_t)s] + n[s];
        e0[(size_t)s + 1] = e0[(size_t)s] + nnz[s];
    }
    if (r0[(size_t)count] >= ((int64_t)1 << 31)) return SPTRSV_ERR_UNSUPPORTED;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : bad)
    for (int s = 0; s < count; ++s) {
        System sys;
        draw_structure(sys, seed, first + s, nmin, nmax);
        if (sys.n != n[s] || sys.nnz != nnz[s]) {
            bad |= 1;
            continue;
        }
        fill_system(sys, seed, first + s, (int)r0[(size_t)s], e0[(size_t)s], rowptr, col, val);
    }
    rowptr[r0[(size_t)count]] = e0[(size_t)count];
    return bad ? SPTRSV_ERR_BAD_INPUT : SPTRSV_OK;
}

}  // extern "C"
