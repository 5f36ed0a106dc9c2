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

#ifdef __cplusplus
}
#endif
#endif /* IDFT_HIP_H */
