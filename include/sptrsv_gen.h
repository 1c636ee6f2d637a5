/*
 * sptrsv_gen.h -- native generator of the batched-solve workload (no reference counterpart).
 *
 * Batches of independent upper-triangular systems, n ~ log-uniform[nmin, nmax], kinds banded /
 * random Zipf / random supernodal in equal proportion, each system drawn from its own seeded
 * random streams so that the batch is generated in parallel and reproducibly; system `first + s`
 * of a call is system s of its output, so consecutive calls (first = 0, c, 2c, ...) produce
 * consecutive pieces of one larger batch.  Two passes:
 * sptrsv_gen_batch_sizes gives every system's rows, nonzeros and kind; the caller allocates the
 * block-diagonal CSR (rowptr: total rows + 1 int64, col: total nnz int32, val: total nnz double)
 * and sptrsv_gen_batch_fill writes it (diagonal first, columns ascending, global columns).  The
 * arrays feed sptrsv_batch_create directly.  Both calls return SPTRSV_OK or a negative
 * SPTRSV_ERR_* code (sptrsv_capi.h).
 */
#ifndef SPTRSV_GEN_H
#define SPTRSV_GEN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPTRSV_GEN_BANDED 0     /* bandwidth ~ U{4..64}, values U(-1,1), diagonal 1 + sum |off| */
#define SPTRSV_GEN_ZIPF 1       /* Poisson(8) offsets ~ Zipf(1.5) per row, same values rule */
#define SPTRSV_GEN_SUPERNODAL 2 /* widths U{16..128}, 1..6 ancestor blocks, N(0,0.1), diag U[1,2] */

int sptrsv_gen_batch_sizes(int count, int first, uint64_t seed, int nmin, int nmax, int* n, int64_t* nnz,
                           int* kind);
int sptrsv_gen_batch_fill(int count, int first, uint64_t seed, int nmin, int nmax, const int* n, const int64_t* nnz,
                          int64_t* rowptr, int* col, double* val);

#ifdef __cplusplus
}
#endif
#endif
