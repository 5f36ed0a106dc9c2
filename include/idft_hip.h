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
 * repetitions after 2 warm-up applications: ms_per_apply for the whole
 * operator (MFMA cell-product kernel + node-gather kernel), ms_gemm (may be
 * NULL) for the MFMA cell-product kernel alone.  No counterpart in the
 * reference; used by the benchmark to measure the dominant kernel live. */
int idft_profile_apply(idft_ctx* ctx, int32_t ncols, int32_t precision, int32_t reps,
                       double* ms_per_apply, double* ms_gemm);

/* Sets the operator launch shape used at the given precision: the
 * cell-product tile (cells_per_wg in 1..64, vectors_per_tile a multiple of
 * 16 in 16..80), tile_pass (1: the cell product reads every stored tile of
 * H_e once and keeps partial sums of all block rows per wave -- used for
 * 16-column tiles, and 32 in FP32; 0: one wave per block row, mirror tiles
 * read again) and the node rows per node-gather work-group (gather_rows in
 * 1..4096, or 0 for the default of about 16384 gather work-groups). */
int idft_set_apply_shape(idft_ctx* ctx, int32_t precision, int32_t cells_per_wg,
                         int32_t vectors_per_tile, int32_t tile_pass, int32_t gather_rows);

/* Autotuner of the cell operator: times every tile shape
 * cells-per-work-group in {1,2,4,8} x vectors-per-tile in {16,32,64} (tiles
 * wider than the block collapse to one pass) x tile_pass {0,1} where it
 * applies, then node rows per gather work-group in {default,1,2,4,8} with
 * the best tile, with idft_profile_apply on ncols columns at the given
 * precision, keeps the fastest for this context and precision, and reports
 * it.  The two cell-product kernels sum in different orders, so results
 * agree to rounding, not bit for bit, across tile_pass.  No counterpart in
 * the reference (its cell matrices go through a library batched GEMM). */
int idft_autotune_apply(idft_ctx* ctx, int32_t ncols, int32_t precision, int32_t reps,
                        int32_t* cells_per_wg, int32_t* vectors_per_tile, int32_t* tile_pass,
                        int32_t* gather_rows, double* ms_best);

#ifdef __cplusplus
}
#endif
#endif /* IDFT_HIP_H */
