// CPU emulation of the PANEL kernel's indexing (sptrsv_amd/csrc/panel.hip) on the tables of
// sptrsv_amd/csrc/panel_layout.h: work items in ticket order, the halo read from x, steps with
// two LDS buffers, four lanes per cell and four columns per lane (a cell at table position c
// on cell lane c % 64, its b and x in register slot c / 64), the halo and the stores sixteen
// lanes per cell, exactly the kernel's address arithmetic.  Checks that every value a cell reads was written before (in-brick: by the
// previous step of the same brick; halo: by an earlier work item), that no read or write leaves
// its buffer, and that x is bitwise equal to the serial back substitution (src/core.cpp:18-32).
// Built with -fsanitize=address,undefined by tests/test_capi_cpu.py.
//   panel_emulate nx ny nz m seed drop   (drop: fraction of off-diagonal entries removed)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "panel_layout.h"

using namespace sptrsv;

namespace {
constexpr int kPanel = 16, kPT = 256;

int fail(const char* what, long a, long b) {
    std::printf("FAIL %s %ld %ld\n", what, a, b);
    return 1;
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 7) return 2;
    const int64_t nx = std::atol(argv[1]), ny = std::atol(argv[2]), nz = std::atol(argv[3]);
    const int m = std::atoi(argv[4]);
    const unsigned seed = (unsigned)std::atoi(argv[5]);
    const double drop = std::atof(argv[6]);
    const int64_t n = nx * ny * nz;
    const int64_t off[3] = {1, nx, nx * ny};
    // factor: per row diagonal + entries at +1, +nx, +nx*ny inside the grid, some dropped
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(-0.3, 0.3), D(2.0, 3.0), Z(0.0, 1.0);
    std::vector<double> rec((size_t)n * 4, 0.0);
    std::vector<unsigned char> mask((size_t)n, 0);
    for (int64_t r = 0; r < n; ++r) {
        const int64_t i = r % nx, j = (r / nx) % ny, k = r / (nx * ny);
        rec[(size_t)r * 4] = D(rng);
        const bool in[3] = {i + 1 < nx, j + 1 < ny, k + 1 < nz};
        for (int d = 0; d < 3; ++d)
            if (in[d] && Z(rng) >= drop) {
                rec[(size_t)r * 4 + 1 + d] = U(rng);
                mask[(size_t)r] |= (unsigned char)(1 << d);
            }
    }
    std::vector<double> b((size_t)n * m);
    for (auto& v : b) v = U(rng) * 10.0;
    // serial back substitution, ascending column order
    std::vector<double> ref((size_t)n * m);
    for (int64_t r = n - 1; r >= 0; --r)
        for (int c = 0; c < m; ++c) {
            double acc = 0.0;
            for (int d = 0; d < 3; ++d)
                if (mask[(size_t)r] >> d & 1) acc = acc + rec[(size_t)r * 4 + 1 + d] * ref[(size_t)(r + off[d]) * m + c];
            ref[(size_t)r * m + c] = (b[(size_t)r * m + c] - acc) / rec[(size_t)r * 4];
        }

    PanelLayout L;
    if (!panel_layout(nx, ny, nz, &L)) return fail("layout", nx, ny);
    if ((int)L.cells.size() != L.ncells || (int)L.step_ptr.size() != L.nsteps + 1) return fail("tables", L.ncells, L.nsteps);
    const int npanel = (m + kPanel - 1) / kPanel;
    const int64_t items = (int64_t)L.blist.size() * npanel;
    std::vector<double> x((size_t)n * m, NAN), truth((size_t)n * m, NAN);
    std::vector<int64_t> done_item((size_t)L.blist.size() * npanel, -1);  // ticket that finished (brick, panel)
    std::vector<unsigned char> face_done((size_t)n * npanel, 0);  // (row, panel) stored before its flag
    std::vector<int> brick_pos((size_t)L.blist.size());
    for (size_t q = 0; q < L.blist.size(); ++q) brick_pos[(size_t)L.blist[q]] = (int)q;
    std::vector<double> xbuf(2 * (kPT / 4) * kPanel), halo((size_t)kPanelMaxHalo * kPanel);
    std::vector<int> xbuf_cell(2 * (kPT / 4)), halo_ok(kPanelMaxHalo);
    for (int64_t t = 0; t < items; ++t) {
        const int bl = (int)(t / npanel), p = (int)(t % npanel);
        const int lin = L.blist[(size_t)bl];
        const int I = lin % L.gx, J = (lin / L.gx) % L.gy, K = lin / (L.gx * L.gy);
        // upstream bricks must have smaller tickets (the kernel waits on them)
        const int ups[3] = {I + 1 < L.gx ? lin + 1 : -1, J + 1 < L.gy ? lin + L.gx : -1,
                            K + 1 < L.gz ? lin + L.gx * L.gy : -1};
        for (int up : ups)
            if (up >= 0 && (done_item[(size_t)up * npanel + p] < 0 || done_item[(size_t)up * npanel + p] >= t))
                return fail("upstream not finished", t, up);
        // halo
        const int fx = L.by * L.bz, fy = L.bx * L.bz;
        for (int h = 0; h < kPanelMaxHalo; ++h) halo_ok[h] = 0;
        for (int tid = 0; tid < kPT; ++tid) {
            const int col = tid & (kPanel - 1);  // sixteen lanes per face cell
            for (int j = 0; j < (kPanelMaxHalo + kPT / kPanel - 1) / (kPT / kPanel); ++j) {
                const int h = tid / kPanel + j * (kPT / kPanel);
                int64_t gi, gj, gk;
                if (h < fx) { gi = (int64_t)(I + 1) * L.bx; gj = (int64_t)J * L.by + h % L.by; gk = (int64_t)K * L.bz + h / L.by; }
                else if (h < fx + fy) { gi = (int64_t)I * L.bx + (h - fx) % L.bx; gj = (int64_t)(J + 1) * L.by; gk = (int64_t)K * L.bz + (h - fx) / L.bx; }
                else { gi = (int64_t)I * L.bx + (h - fx - fy) % L.bx; gj = (int64_t)J * L.by + (h - fx - fy) / L.bx; gk = (int64_t)(K + 1) * L.bz; }
                int64_t hx, hy, hz;
                if (h < L.nhalo) {
                    panel_halo_cell(L, h, I, J, K, &hx, &hy, &hz);
                    if (hx != gi || hy != gj || hz != gk) return fail("halo coordinates", h, lin);
                }
                const bool ok = h < L.nhalo && gi < nx && gj < ny && gk < nz;
                if (!ok) continue;
                const int64_t row = gi + nx * (gj + ny * gk);
                if (row < 0 || row >= n) return fail("halo row", row, h);
                // the flag covers only the face stores: a halo cell must be one of them
                if (!face_done[(size_t)row * npanel + p]) return fail("halo cell not stored before the flag", row, h);
                halo_ok[h] = 1;
                if (p * kPanel + col < m) halo[(size_t)h * kPanel + col] = x[(size_t)row * m + p * kPanel + col];
            }
        }
        // each lane's b (then x) registers: cell slot j of cell lane kr is table position j * 64 + kr
        const int NJ = (L.ncells + 63) / 64;
        std::vector<double> bvr((size_t)kPT * NJ * 4, 0.0);
        for (int tid = 0; tid < kPT; ++tid)
            for (int j = 0; j < NJ; ++j) {
                const int ci = j * 64 + (tid >> 2), c0 = p * kPanel + 4 * (tid & 3);
                if (ci >= L.ncells) continue;
                const uint64_t c = L.cells[(size_t)ci];
                const int64_t gi = (int64_t)I * L.bx + panel_cell_i(c), gj = (int64_t)J * L.by + panel_cell_j(c),
                              gk = (int64_t)K * L.bz + panel_cell_k(c);
                if (!(gi < nx && gj < ny && gk < nz)) continue;
                const int64_t r = gi + nx * (gj + ny * gk);
                for (int u = 0; u < 4; ++u)
                    if (c0 + u < m) bvr[((size_t)tid * NJ + j) * 4 + u] = b[(size_t)r * m + c0 + u];
            }
        // steps
        for (int& c : xbuf_cell) c = -1;
        for (int s = 0; s < L.nsteps; ++s) {
            const int cur = s & 1, prv = (s + 1) & 1;
            std::vector<double> out((size_t)kPT * 4);
            std::vector<int> out_pos((size_t)kPT * 4, -1);
            std::vector<int> out_cell(kPT / 4, -1);
            const int s0 = L.step_ptr[(size_t)s], cnt = L.step_ptr[(size_t)s + 1] - s0;
            if (cnt > kPT / 4 || cnt > kPanelMaxWidth) return fail("step width", s, cnt);
            std::vector<int> covered((size_t)cnt, 0);
            for (int tid = 0; tid < kPT; ++tid) {
                const int kr = tid >> 2, q = tid & 3, c0 = p * kPanel + 4 * q;
                // the kernel's cell of this lane at step s: table position ci with ci % 64 == kr
                const int ps = s0, pe = s0 + cnt, j0 = ps / 64, j1 = (pe - 1) / 64;
                const int ci = (j0 == j1 || kr >= ps % 64) ? j0 * 64 + kr : j1 * 64 + kr;
                if (ci < ps || ci >= pe) continue;
                if (ci % 64 != kr || (ci / 64 != j0 && ci / 64 != j1)) return fail("cell lane", s, ci);
                if (q == 0) ++covered[(size_t)(ci - ps)];
                const int kr_pos = ci - ps;  // position inside the step: the LDS row it writes
                const uint64_t c = L.cells[(size_t)ci];
                const int64_t gi = (int64_t)I * L.bx + panel_cell_i(c), gj = (int64_t)J * L.by + panel_cell_j(c),
                              gk = (int64_t)K * L.bz + panel_cell_k(c);
                if (!(gi < nx && gj < ny && gk < nz)) continue;
                const int64_t r = gi + nx * (gj + ny * gk);
                double acc[4] = {0, 0, 0, 0};
                for (int d = 0; d < 3; ++d) {
                    if (!(mask[(size_t)r] >> d & 1)) continue;
                    const int sl = panel_cell_slot(c, d);
                    const int64_t dep = r + off[d];
                    const double* src;
                    if (sl < kPanelSlotHalo) {
                        if (xbuf_cell[(size_t)prv * (kPT / 4) + sl] != (int)dep) return fail("in-brick slot", r, sl);
                        src = &xbuf[((size_t)prv * (kPT / 4) + sl) * kPanel];
                    } else {
                        const int h = sl - kPanelSlotHalo;
                        if (h >= L.nhalo || !halo_ok[h]) return fail("halo slot", r, h);
                        src = &halo[(size_t)h * kPanel];
                    }
                    for (int u = 0; u < 4; ++u)
                        if (c0 + u < m) {
                            if (src[4 * q + u] != truth[(size_t)dep * m + c0 + u]) return fail("stale value", r, dep);
                            acc[u] = acc[u] + rec[(size_t)r * 4 + 1 + d] * src[4 * q + u];
                        }
                }
                for (int u = 0; u < 4; ++u)
                    if (c0 + u < m) {
                        double& reg = bvr[((size_t)tid * NJ + ci / 64) * 4 + u];
                        if (reg != b[(size_t)r * m + c0 + u]) return fail("b register", r, ci);
                        const double xv = (reg - acc[u]) / rec[(size_t)r * 4];
                        out[(size_t)tid * 4 + u] = xv;
                        truth[(size_t)r * m + c0 + u] = xv;
                        reg = xv;
                    }
                out_cell[kr_pos] = (int)r;
                for (int u = 0; u < 4; ++u) out_pos[(size_t)tid * 4 + u] = kr_pos;
            }
            for (int k = 0; k < cnt; ++k)
                if (covered[(size_t)k] != 1) return fail("cell not covered once", s, k);
            for (int k = 0; k < kPT / 4; ++k) xbuf_cell[(size_t)cur * (kPT / 4) + k] = out_cell[k];
            for (int tid = 0; tid < kPT; ++tid) {
                const int q = tid & 3;
                for (int u = 0; u < 4; ++u)
                    if (out_pos[(size_t)tid * 4 + u] >= 0)
                        xbuf[((size_t)cur * (kPT / 4) + out_pos[(size_t)tid * 4 + u]) * kPanel + 4 * q + u] = out[(size_t)tid * 4 + u];
            }
        }
        // store phase: per wave and cell slot, the quads' registers through a 16 x 16 staging
        // tile, then sixteen lanes per cell, one column each
        for (int wv = 0; wv < kPT / 64; ++wv)
            for (int j = 0; j < NJ; ++j) {
                double stage[16 * kPanel];
                for (int ln = 0; ln < 64; ++ln) {
                    const int tid = wv * 64 + ln, kr = tid >> 2, q = tid & 3;
                    for (int u = 0; u < 4; ++u) stage[(kr & 15) * kPanel + 4 * q + u] = bvr[((size_t)tid * NJ + j) * 4 + u];
                }
                for (int ln = 0; ln < 64; ++ln)
                    for (int i = 0; i < 4; ++i) {
                        const int cl = ln / kPanel + 4 * i, ci = j * 64 + wv * 16 + cl, col = ln & (kPanel - 1);
                        if (ci >= L.ncells || p * kPanel + col >= m) continue;
                        const uint64_t c = L.cells[(size_t)ci];
                        const int64_t gi = (int64_t)I * L.bx + panel_cell_i(c), gj = (int64_t)J * L.by + panel_cell_j(c),
                                      gk = (int64_t)K * L.bz + panel_cell_k(c);
                        if (!(gi < nx && gj < ny && gk < nz)) continue;
                        const int64_t r = gi + nx * (gj + ny * gk);
                        const size_t at = (size_t)r * m + p * kPanel + col;
                        if (!std::isnan(x[at])) return fail("x written twice", r, col);
                        x[at] = stage[cl * kPanel + col];
                        // the kernel's face test: local coordinate 0 towards a downstream brick
                        const bool face = (panel_cell_i(c) == 0 && I > 0) || (panel_cell_j(c) == 0 && J > 0) ||
                                          (panel_cell_k(c) == 0 && K > 0);
                        if (face) face_done[(size_t)r * npanel + p] = 1;
                    }
            }
        done_item[(size_t)lin * npanel + p] = t;
    }
    for (size_t e = 0; e < x.size(); ++e)
        if (std::memcmp(&x[e], &ref[e], sizeof(double)) != 0) return fail("x differs from the serial solve", (long)e, 0);
    std::printf("ok n=%ld m=%d bricks=%zu steps=%d halo=%d items=%ld\n", (long)n, m, L.blist.size(), L.nsteps, L.nhalo,
                (long)items);
    return 0;
}
