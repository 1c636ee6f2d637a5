// panel_layout.h -- brick geometry of the PANEL variant (host code, no HIP): a natural-order
// grid factor with off-diagonal offsets {1, nx, nx*ny} is cut into bx x by x bz bricks; a
// brick's cells are solved in steps (a cell's step is its distance from the brick's far
// corner), and every dependency of a cell at step s lies either in the same brick at step
// s - 1 or on a face of one of the three upstream bricks (I+1, J, K), (I, J+1, K), (I, J, K+1).
//
// The kernel and the CPU emulation in tests/cpp/panel_emulate.cpp both read these tables.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <vector>

namespace sptrsv {

constexpr int kPanelMaxCells = 512;  // cells per brick
constexpr int kPanelMaxWidth = 48;   // cells in one step of a brick
constexpr int kPanelMaxHalo = 192;   // face cells of the upstream bricks read by one brick
constexpr int kPanelMaxSteps = 64;  // steps of a brick
constexpr int kPanelSlotHalo = 64;   // dependency slot >= this: halo cell (slot - 64)

// Cell word: bits 0-4 / 5-9 / 10-14 the local coordinates i, j, k; bits 16-24 / 25-33 / 34-42
// the slots of the dependencies at +1, +nx, +nx*ny: < 64 the position of that cell inside the
// previous step, >= 64 its halo index + 64.
inline uint64_t panel_cell_pack(int i, int j, int k, int sx, int sy, int sz) {
    return (uint64_t)i | ((uint64_t)j << 5) | ((uint64_t)k << 10) | ((uint64_t)sx << 16) | ((uint64_t)sy << 25) |
           ((uint64_t)sz << 34);
}
inline int panel_cell_i(uint64_t c) { return (int)(c & 31); }
inline int panel_cell_j(uint64_t c) { return (int)((c >> 5) & 31); }
inline int panel_cell_k(uint64_t c) { return (int)((c >> 10) & 31); }
inline int panel_cell_slot(uint64_t c, int d) { return (int)((c >> (16 + 9 * d)) & 511); }

struct PanelLayout {
    int64_t nx = 0, ny = 0, nz = 0;  // grid
    int bx = 0, by = 0, bz = 0;      // brick
    int gx = 0, gy = 0, gz = 0;      // bricks per dimension
    int nsteps = 0, ncells = 0, width = 0, nhalo = 0;
    std::vector<int> blist;          // linear brick ids I + gx (J + gy K) in dependency order
    std::vector<int> step_ptr;       // nsteps + 1 offsets into cells
    std::vector<uint64_t> cells;     // ncells cell words, grouped by step
};

// Halo index of the face cell that a brick cell reads in direction d (0: +1, 1: +nx,
// 2: +nx*ny) when that neighbour lies in the next brick: faces x, y, z one after another.
inline int panel_halo_index(const PanelLayout& L, int d, int i, int j, int k) {
    if (d == 0) return j + L.by * k;
    if (d == 1) return L.by * L.bz + i + L.bx * k;
    return L.by * L.bz + L.bx * L.bz + i + L.bx * j;
}

// Global grid coordinates of halo cell h of brick (I, J, K): the neighbour cell in the
// upstream brick.
inline void panel_halo_cell(const PanelLayout& L, int h, int I, int J, int K, int64_t* gi, int64_t* gj, int64_t* gk) {
    const int fx = L.by * L.bz, fy = L.bx * L.bz;
    if (h < fx) {
        *gi = (int64_t)(I + 1) * L.bx;
        *gj = (int64_t)J * L.by + h % L.by;
        *gk = (int64_t)K * L.bz + h / L.by;
    } else if (h < fx + fy) {
        h -= fx;
        *gi = (int64_t)I * L.bx + h % L.bx;
        *gj = (int64_t)(J + 1) * L.by;
        *gk = (int64_t)K * L.bz + h / L.bx;
    } else {
        h -= fx + fy;
        *gi = (int64_t)I * L.bx + h % L.bx;
        *gj = (int64_t)J * L.by + h / L.bx;
        *gk = (int64_t)(K + 1) * L.bz;
    }
}

// Bricks of 8x8x8 (3D) or 16x16 (2D).  Returns false for a grid the tables cannot describe.
inline bool panel_layout(int64_t nx, int64_t ny, int64_t nz, PanelLayout* L) {
    if (nx < 2 || ny < 2 || nz < 1) return false;
    L->nx = nx;
    L->ny = ny;
    L->nz = nz;
    L->bx = nz > 1 ? 8 : 16;
    L->by = nz > 1 ? 8 : 16;
    L->bz = nz > 1 ? 8 : 1;
    const int64_t gx = (nx + L->bx - 1) / L->bx, gy = (ny + L->by - 1) / L->by, gz = (nz + L->bz - 1) / L->bz;
    if (gx * gy * gz >= ((int64_t)1 << 30)) return false;
    L->gx = (int)gx;
    L->gy = (int)gy;
    L->gz = (int)gz;
    L->nsteps = L->bx + L->by + L->bz - 2;
    L->ncells = L->bx * L->by * L->bz;
    L->nhalo = L->by * L->bz + L->bx * L->bz + (nz > 1 ? L->bx * L->by : 0);
    // step of a cell: its distance from the far corner (bx-1, by-1, bz-1); every in-brick
    // dependency (+1 in one coordinate) is one step earlier
    auto step_of = [&](int i, int j, int k) { return (L->bx - 1 - i) + (L->by - 1 - j) + (L->bz - 1 - k); };
    std::vector<std::vector<int>> by_step((size_t)L->nsteps);
    for (int k = L->bz - 1; k >= 0; --k)
        for (int j = L->by - 1; j >= 0; --j)
            for (int i = L->bx - 1; i >= 0; --i) by_step[(size_t)step_of(i, j, k)].push_back(i + L->bx * (j + L->by * k));
    std::vector<int> pos((size_t)L->ncells);
    L->width = 0;
    for (auto& v : by_step) {
        for (size_t q = 0; q < v.size(); ++q) pos[(size_t)v[q]] = (int)q;
        L->width = std::max(L->width, (int)v.size());
    }
    if (L->ncells > kPanelMaxCells || L->width > kPanelMaxWidth || L->nhalo > kPanelMaxHalo || L->nsteps > kPanelMaxSteps)
        return false;
    L->step_ptr.assign(1, 0);
    L->cells.clear();
    for (auto& v : by_step) {
        for (int li : v) {
            const int i = li % L->bx, j = (li / L->bx) % L->by, k = li / (L->bx * L->by);
            int sl[3];
            const int nb[3][3] = {{i + 1, j, k}, {i, j + 1, k}, {i, j, k + 1}};
            for (int d = 0; d < 3; ++d) {
                const bool inside = nb[d][0] < L->bx && nb[d][1] < L->by && nb[d][2] < L->bz;
                sl[d] = inside ? pos[(size_t)(nb[d][0] + L->bx * (nb[d][1] + L->by * nb[d][2]))]
                               : kPanelSlotHalo + panel_halo_index(*L, d, i, j, k);
            }
            L->cells.push_back(panel_cell_pack(i, j, k, sl[0], sl[1], sl[2]));
        }
        L->step_ptr.push_back((int)L->cells.size());
    }
    // bricks by anti-diagonal from the far corner (every upstream brick on an earlier one),
    // then by descending linear id
    std::vector<std::pair<int64_t, int>> key;
    key.reserve((size_t)(gx * gy * gz));
    for (int K = 0; K < L->gz; ++K)
        for (int J = 0; J < L->gy; ++J)
            for (int I = 0; I < L->gx; ++I) {
                const int lin = I + L->gx * (J + L->gy * K);
                const int64_t diag = (int64_t)(L->gx - 1 - I) + (L->gy - 1 - J) + (L->gz - 1 - K);
                key.emplace_back(diag * (int64_t)(1 << 30) + ((1 << 30) - 1 - lin), lin);
            }
    std::sort(key.begin(), key.end());
    L->blist.resize(key.size());
    for (size_t q = 0; q < key.size(); ++q) L->blist[q] = key[q].second;
    return true;
}

}  // namespace sptrsv
