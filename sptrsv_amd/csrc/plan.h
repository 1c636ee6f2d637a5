// plan.h -- device-resident state of one upper-triangular factor.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <utility>
#include <vector>

#include "common.h"

struct sptrsv_plan_s {
    int n = 0;
    int64_t nnz = 0;
    int n_levels = 0;
    int max_row_len = 0;
    double analysis_ms = 0.0;  // host wall time of analyse()
    double levels_ms = 0.0;    // of its level computation
    int depth_lb = 0;   // depth of the forest "row -> its first off-diagonal column": a lower
                        // bound on the level count, found on the host at plan creation
    bool grid = false;  // off-diagonal offsets of a natural-order 2D/3D grid factor (block.hip)
    float feat[24] = {};  // selector features (select.hip), computed on first use
    bool feat_ok = false;
    int64_t grid_dims[3] = {0, 0, 0};  // its nx, ny, nz

    // CSR, natural row order, diagonal first (HBM: 4(n+1) + 12 nnz bytes)
    int* rowptr = nullptr;
    int* col = nullptr;
    double* val = nullptr;
    float* val32 = nullptr;  // FP32 copy, built on first FP32 solve

    // CSC, rows ascending, diagonal last
    int* colptr = nullptr;
    int* rowidx = nullptr;
    double* cval = nullptr;
    float* cval32 = nullptr;

    // analysis
    int* indeg = nullptr;       // nnz of each row incl. diagonal (reference analyser)
    int* levels = nullptr;      // 1-based level of each row
    int* level_ptr = nullptr;   // n_levels + 1
    int* level_rows = nullptr;  // rows grouped by level, ascending inside a level

    // CSR with rows stored in level order (row k of it is original row level_rows[k])
    int* lrowptr = nullptr;
    int* lcol = nullptr;
    double* lval = nullptr;
    float* lval32 = nullptr;

    // work buffers
    int* ticket = nullptr;      // [0]: ticket counter, [1]: status word
    int* indeg_work = nullptr;  // in-degree consumed by the push kernel
    double* left_sum = nullptr;
    size_t left_sum_elems = 0;
    float* b32 = nullptr;       // FP32 staging of b / x
    float* x32 = nullptr;
    size_t stage_elems = 0;

    std::vector<int> h_level_ptr;  // host copy, drives the bulk-synchronous launches

    // supernodal layout (built by sptrsv_plan_build_supernodes): supernode k holds rows
    // [sn_start[k], sn_start[k+1]); its diagonal block is stored INVERTED as a dense row-major
    // w x w block, its rectangular part as a dense row-major w x r block over the columns
    // sn_cols[sn_offc[k] .. + r).
    int sn = 0;
    int sn_wmax = 0;
    int* sn_start = nullptr;     // sn + 1
    int* sn_r = nullptr;         // sn
    int64_t* sn_offt = nullptr;  // sn, offset of the inverse block
    int64_t* sn_offr = nullptr;  // sn, offset of the rectangular block
    int* sn_offc = nullptr;      // sn, offset into sn_cols
    int* sn_cols = nullptr;
    double* sn_tinv = nullptr;
    double* sn_tri = nullptr;     // diagonal blocks (row-major w x w, zero below), for the residual
    double* sn_rect = nullptr;
    float* sn_tinv32 = nullptr;     // mixed precision: FP32 inverse blocks
    uint16_t* sn_rect16 = nullptr;  // mixed precision: BF16 rectangular blocks, scaled per block
    float* sn_rscale = nullptr;     // per-supernode scale of the BF16 block
    int* sn_dep_ptr = nullptr;   // sn + 1
    int* sn_dep = nullptr;       // supernodes owning the rectangular columns
    int* sn_flag = nullptr;      // sn x sn_flag_blocks epochs of the last solve that published
    int sn_flag_blocks = 1;      //   (supernode, column panel)
    int sn_epoch = 0;
    std::vector<int> h_snodes;
    std::vector<int> h_sn_r;        // rectangular column count per supernode
    int sn_rmax = 0;                // the largest of them
    int sn_levels = 0;              // dependency levels of the supernodes
    int64_t* sn_offr16 = nullptr;   // offsets of the BF16 blocks (rows padded to 16 columns)
    int64_t sn_padded = 0;          // explicit zeros stored by the layout (relaxed supernodes)
    int relaxed_state = 0;          // AUTO's relaxed supernodes: 0 not tried, 1 built, -1 not used

    // refinement work buffers (n*m)
    double* ref_r = nullptr;
    float* ref_r32 = nullptr;
    float* ref_d32 = nullptr;
    unsigned long long* ref_norms = nullptr;  // 2*m: max|d_k|, max|x_k| as ordered bit patterns
    size_t ref_elems = 0;
    size_t ref_cols = 0;

    // tiles for the TILE variant (built on first use for a tile size): tile t = rows
    // [t T, (t+1) T); its steps are tile_step[tile_sptr[t] .. tile_sptr[t+1]) (first positions);
    // position p is the 80-byte record tile_rec[p] (row, length, first entries, diagonal),
    // further entries in tile_col / tile_val; tile_rlist: per step, tile_cap (remote column,
    // ring slot) pairs for the poller wave.
    int n_tiles = 0;
    int tile_rows = 0;
    int tile_cap = 0;
    int* tile_sptr = nullptr;
    int* tile_step = nullptr;
    void* tile_rec = nullptr;
    int* tile_col = nullptr;
    double* tile_val = nullptr;
    void* tile_rlist = nullptr;
    int* tile_row = nullptr;      // row of each position
    double* tile_bt = nullptr;    // b in position order (per solve)

    // blocks of the BLOCK variant (built on first use): block k = positions [blk_ptr[k],
    // blk_ptr[k+1]), position p = row blk_perm[p]; CSR in position order whose column entries
    // (blk_col) are LDS slots: < 64 a row of the block, 64 + e its e-th external value;
    // blk_kind 1 = grid bricks, 2 = level bands
    int n_blocks = 0;
    int blk_kind = 0;
    int blk_maxlen = 0;          // most off-diagonal entries in a row
    int blk_rows = 0;            // rows per block the layout was built for (64 or 512)
    int blk_maxext = 0;          // longest external list of a block
    int* blk_ptr = nullptr;
    int* blk_eptr = nullptr;     // per block: its external values blk_ecol[blk_eptr[k] ..)
    int* blk_ecol = nullptr;
    int* blk_perm = nullptr;
    int* blk_rp = nullptr;
    int* blk_col = nullptr;
    int* blk_step = nullptr;     // per position: 1 + the largest step of its in-block dependencies
    double* blk_val = nullptr;

    // brick panels of the PANEL variant (grid factors, built on first use): bricks in dependency
    // order, the brick's cell table by step, per row its record (diagonal, entries at +1, +nx,
    // +nx*ny) and entry mask; per (brick, panel of 16 columns) the epoch of the last solve that
    // finished it.  pn_tickets: tickets taken so far from the plan's control word 16.
    bool pn_built = false;
    int pn_bx = 0, pn_by = 0, pn_bz = 0, pn_gx = 0, pn_gy = 0, pn_gz = 0;
    int pn_nsteps = 0, pn_nhalo = 0, pn_nbricks = 0, pn_ncells = 0;
    int* pn_blist = nullptr;
    int* pn_step_ptr = nullptr;
    uint64_t* pn_cells = nullptr;
    double* pn_rec = nullptr;
    unsigned char* pn_mask = nullptr;
    int* pn_flag = nullptr;
    size_t pn_flag_elems = 0;
    int pn_epoch = 0;
    unsigned pn_tickets = 0;

    // chunks of the level-set solve (levelset.hip, built on first use), one set per chunk size
    // (256 >> s rows for G = 1 << s lanes per row): chunks of level l are
    // ls_ck_level[s][l] .. ls_ck_level[s][l+1]; chunk c is level-ordered slots
    // [ls_ck_slot[s][c], ls_ck_slot[s][c+1]) of level ls_ck_lev[s][c]; ls_done: per level,
    // chunks finished in the current solve
    static constexpr int kLsSizes = 7;
    int* ls_ck_level[kLsSizes] = {};
    int* ls_ck_slot[kLsSizes] = {};
    int* ls_ck_lev[kLsSizes] = {};
    int ls_chunks[kLsSizes] = {};
    int* ls_done = nullptr;

    // chain variant: one record per 32-row window (packed inverse of its triangle, near and old
    // entries), built on first use; FP64 and FP32 values
    char* chain_rec = nullptr;
    char* chain_rec32 = nullptr;

    // companion plan of J U^T J (J reverses the index order) for solves with U^T, built on the
    // first transposed solve and run on this plan's stream; tr_buf holds reversed b, reversed
    // x and the intermediate of the Cholesky solve (n*m each)
    sptrsv_plan_s* tr = nullptr;
    double* tr_buf = nullptr;
    size_t tr_elems = 0;
    bool borrowed_stream = false;  // stream owned by the plan this one is the companion of

    hipStream_t stream = nullptr;
    int device = 0;  // HIP device the plan's memory and stream live on
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    size_t events_used = 0;
    uint64_t ticks_per_ms = 100000;
};

namespace sptrsv {
int analyse(sptrsv_plan_s* p);           // analysis.hip
int chain_levels(sptrsv_plan_s* p, int* ctrl);  // chain_levels.hip: levels of a chain-like factor
int ensure_fp32(sptrsv_plan_s* p);       // analysis.hip
int device_transpose(int n, int64_t nnz, const int* rowptr, const int* col, const double* val,
                     int* colptr, int* rowidx, double* cval, hipStream_t s, void* scratch = nullptr);
size_t transpose_scratch_bytes(int64_t nnz);

// radix.hip: stable radix sort of int (key, value) pairs (keys in [0, 2^bits); vals == nullptr
// sorts the item indices), prefix sums of ints, and the pointer array of a sorted key list
// (ptr[c] = first position whose key - base >= c, c in [0, K]).  in/out of scan_ints may alias.
size_t sort_scratch_bytes(int64_t N);
int sort_pairs(const int* keys_in, const int* vals_in, int64_t N, int bits, int* keys_out, int* vals_out,
               void* scratch, hipStream_t s);
size_t scan_scratch_bytes(int64_t count);
int scan_ints(const int* in, int* out, int64_t count, bool inclusive, void* scratch, hipStream_t s);
int key_bounds(const int* sorted, int64_t N, int K, int base, int* ptr, hipStream_t s);  // analysis.hip
int supernodal_solve(sptrsv_plan_s* p, int precision, const double* d_b, double* d_x, int m,
                     uint64_t limit);  // supernodal.hip
void supernodal_free(sptrsv_plan_s* p);   // supernodal.hip
int supernodal_residual32(sptrsv_plan_s* p, const double* d_b, const double* d_x, float* d_r32, int m);
int residual32(sptrsv_plan_s* p, const double* d_b, const double* d_x, float* d_r32, int m);
struct RefineResult {
    int iters;        // corrections applied
    int outcome;      // SPTRSV_REFINE_*
    double estimate;  // error estimate of the last iterate
};
int refine_loop(sptrsv_plan_s* p, const sptrsv_refine_t& ro, const double* d_b, double* d_x, int m,
                const std::function<int(const float*, float*)>& solve_low,
                const std::function<int(const double*, const double*, float*)>& residual, RefineResult* res);
// AUTO's choice by the installed selector network (select.hip); false: use the hand rule
bool selector_pick(sptrsv_plan_s* p, int m, int precision, bool exact, int* variant);
// AUTO's choice (exact: only kernels bitwise equal to the serial solver)
int default_variant(sptrsv_plan_s* p, int m, int precision, bool exact);  // solve.hip
// FP32 solve of device b32 -> x32 with the variant in params (not AUTO), m right-hand sides
int solve_fp32_device(sptrsv_plan_s* p, const sptrsv_params_t* params, const float* d_b32, float* d_x32,
                      int m);  // solve.hip
int ensure_stage(sptrsv_plan_s* p, int64_t nm);  // solve.hip: FP32 staging buffers b32 / x32
bool relaxed_supernodes_pay(sptrsv_plan_s* p);  // supernodal.hip: builds the relaxed layout once
int plan_create_control(int n, int64_t nnz, sptrsv_plan_s** out);  // runtime.hip: no factor arrays
int supernodal_solve_low(sptrsv_plan_s* p, int precision, const float* d_b32, float* d_x32, int m,
                         uint64_t limit);  // supernodal.hip
int tile_solve(sptrsv_plan_s* p, int precision, int rows, const double* d_b, double* d_x, int m,
               uint64_t limit);  // tile.hip
void tile_free(sptrsv_plan_s* p);         // tile.hip
int block_solve(sptrsv_plan_s* p, int precision, int occ, const double* d_b, double* d_x, int m,
                uint64_t limit);  // block.hip
bool block_supports(sptrsv_plan_s* p, int m);  // block.hip
void block_free(sptrsv_plan_s* p);        // block.hip
bool detect_grid(int n, const int* rowptr, const int* col, int64_t* dims);  // block.hip
int panel_solve(sptrsv_plan_s* p, int precision, const double* d_b, double* d_x, int m,
                uint64_t limit);  // panel.hip
bool panel_supports(sptrsv_plan_s* p);  // panel.hip
void panel_free(sptrsv_plan_s* p);      // panel.hip
template <typename T>
int levelset_solve(sptrsv_plan_s* p, const T* lval, const T* d_b, T* d_x, int m, uint64_t limit);  // levelset.hip
void levelset_free(sptrsv_plan_s* p);   // levelset.hip
int chain_solve(sptrsv_plan_s* p, int precision, const double* d_b, double* d_x, int m);  // chain.hip
int chain_solve32(sptrsv_plan_s* p, const float* d_b, float* d_x);                        // chain.hip
void chain_free(sptrsv_plan_s* p);                                                        // chain.hip
int convert_d2f(const double* in, float* out, int64_t count, hipStream_t s);  // solve.hip
int convert_f2d(const float* in, double* out, int64_t count, hipStream_t s);  // solve.hip
int zero_ints_async(int* p, int64_t count, hipStream_t s);  // runtime.hip: in stream order
// the 24 selector features of rows [lo, hi) of a CSR with 1-based levels lev (batch.hip)
void system_features(const int64_t* rp, const int* col, const double* val, const std::vector<int>& lev, int lo,
                     int hi, int m, float* f);
int begin_timed(sptrsv_plan_s* p, hipEvent_t* stop_out);
void end_timed(sptrsv_plan_s* p, hipEvent_t stop);
}  // namespace sptrsv
