/*
 * idft_hip.h -- entry points of the HIP library (libidft.so) only.
 */
#ifndef IDFT_HIP_H
#define IDFT_HIP_H

#include "idft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Checks that device 0 is a gfx950 (MI355X) GPU.  No counterpart in the
 * reference (its device set-up lives in DFT-FE); loaders call it before any
 * other entry point so that a missing or wrong GPU fails loudly.
 * Returns IDFT_OK, IDFT_EHIP (no device) or IDFT_EINVAL (not gfx950). */
int idft_hip_init(void);

/* Average device time (HIP events on the context's stream) of one
 * application of the Loewdin cell operator on a block of ncols random
 * columns, FP64 or FP32 (precision = IDFT_FP64 / IDFT_MIXED), over reps
 * repetitions after 2 warm-up applications, taking three different random
 * blocks in turn so that the input is not cache-resident: ms_per_apply for the whole
 * operator (cell-product kernel + node-gather kernel), ms_gemm (may be
 * NULL) for the cell-product kernel alone.  No counterpart in the
 * reference; used by the benchmark to measure the dominant kernel live. */
int idft_profile_apply(idft_ctx* ctx, int32_t ncols, int32_t precision, int32_t reps,
                       double* ms_per_apply, double* ms_gemm);

/* Cell-product kernels (the `kernel` argument below). */
#define IDFT_KERNEL_ROW_PASS 0   /* stored H_e tiles, one wave per 16-row block */
#define IDFT_KERNEL_TILE_PASS 1  /* stored H_e tiles, each read once per row half */
#define IDFT_KERNEL_SUM_FACTOR 2 /* no stored H_e: 1D contractions with the cell's
                                    quadrature weights (sum factorisation) */

/* Sets the operator launch shape used at the given precision: the
 * cell-product kernel and its shape, and the node chunks per node-gather
 * work-group (gather_rows in 1..4096 chunks of 256 / (columns per thread
 * group) consecutive nodes, or 0 for the default of about 16384 gather
 * work-groups).
 *   IDFT_KERNEL_ROW_PASS / IDFT_KERNEL_TILE_PASS: cells_per_wg in 1..64,
 *     vectors_per_tile a multiple of 16 in 16..80; the tile pass keeps
 *     partial sums of all block rows per wave and applies to 16-column tiles
 *     (and 32 in FP32), wider tiles run the row pass.  The stored H_e is
 *     built on first use.
 *   IDFT_KERNEL_SUM_FACTOR: vectors_per_tile = columns per work-group, 2, 4
 *     or 8; cells_per_wg must be 1 (one column per work item). */
int idft_set_apply_shape(idft_ctx* ctx, int32_t precision, int32_t cells_per_wg,
                         int32_t vectors_per_tile, int32_t kernel, int32_t gather_rows);

/* Autotuner of the cell operator: times the sum-factorised kernel at every
 * supported column count per work-group (2 only for blocks narrower than 8
 * columns), then every stored-H shape
 * cells-per-work-group in {1,2,4,8} x vectors-per-tile in {16,32,64} (tiles
 * wider than the block collapse to one pass) x {row pass, tile pass} where
 * it applies, then node chunks per gather work-group in {default,1,2,4,8} with
 * the best.  Each shape is scored on ncols columns at the given precision by
 * the device time of a plain application (as idft_profile_apply) plus that of
 * one whose input carries node and column scales (as in the MINRES steps);
 * the fastest is kept for this context and precision and reported (ms_best:
 * that sum), and the stored H_e is freed when no precision's kernel reads it.  The kernels sum in
 * different orders, so results agree to rounding, not bit for bit, across
 * kernels.  No counterpart in the reference (its cell matrices go through a
 * library batched GEMM). */
int idft_autotune_apply(idft_ctx* ctx, int32_t ncols, int32_t precision, int32_t reps,
                        int32_t* cells_per_wg, int32_t* vectors_per_tile, int32_t* kernel,
                        int32_t* gather_rows, double* ms_best);

/* Multivector handle over the caller's device memory: ptr holds n_nodes *
 * ncols doubles, node-major ([node][col]), on the device of the context, and
 * must outlive the handle; idft_mv_destroy releases the handle only.  With
 * it a caller whose multivectors already live in device memory (DFT-FE's
 * MultiVector on the device, MultiVectorAdjointLinearSolverProblem.h:41-88)
 * drives the contract-2 entry points without a host copy.  No counterpart
 * in the reference (its multivectors are the solver's own objects). */
int idft_mv_wrap(idft_ctx* ctx, double* device_ptr, int32_t ncols, idft_mv** out);
/* vmult (as idft_adj_vmult: Ax = H~ x - eps_i x_i, constrained rows zero)
 * without any device allocation: the cell-product scratch is a fixed device
 * array of the library (64 MiB per GPU, held until the process ends) instead
 * of a buffer allocated on first use.  No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc), typically
 * with Ax and x from idft_mv_wrap.  Requires cells * (p+1)^3 * block <=
 * 8,388,608 (IDFT_EINVAL otherwise) and an operator that needs nothing built
 * on first use: the sum-factorised cell product, or stored cell matrices
 * already built (IDFT_ESTATE otherwise).  Every call on one GPU uses that
 * GPU's one array, so calls that overlap on a GPU (other threads or
 * contexts) race; the array's contents are fully rewritten by each call. */
int idft_adj_vmult_noalloc(idft_ctx* ctx, idft_mv* Ax, idft_mv* x, int32_t block);
/* Device pointer of a multivector handle (owned or wrapped). */
double* idft_mv_device_ptr(const idft_mv* mv);

/* Attainable FMA rate of device 0 (the current device): TFLOP/s of a kernel
 * whose threads run `chains` (4, 8 or 16) independent multiply-add chains,
 * FP64 (precision = IDFT_FP64) or unpacked FP32 (IDFT_MIXED), 32 work-groups
 * of 256 threads per CU.  No counterpart in the reference; the benchmark
 * reports it beside the published peaks, as the rate a latency-limited
 * kernel can reach.  Returns IDFT_EINVAL for other chain counts. */
int idft_fma_probe(int32_t precision, int32_t chains, double* tflops);

/* Device memory held by this library's allocations in the process: current
 * bytes and the peak since the last reset (reset_peak != 0 restarts the
 * peak at the current value after reading).  No counterpart in the
 * reference; used to report the peak HBM of a run. */
int idft_device_memory(int64_t* current_bytes, int64_t* peak_bytes, int32_t reset_peak);

#ifdef __cplusplus
}
#endif
#endif /* IDFT_HIP_H */
