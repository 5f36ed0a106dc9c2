/*
 * idft_io.h -- v_xc checkpoint files and L-BFGS resume (libidft.so).
 *
 * Host-side file formats around the hot path.  The v_xc values are the
 * child-mesh (linear) nodal potential of the inverse problem, n_vxc values
 * per spin channel, node id = i + NV1 (j + NV1 k) on the (2 n_cells + 1)^3
 * grid of idft_geometry_vxc_points.  Multi-channel arrays are channel-major
 * (n_spin blocks of n_vxc values).
 */
#ifndef IDFT_IO_H
#define IDFT_IO_H

#include "idft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Text records as InverseDFTSolverFunction::writeChildMeshDataToFile
 * (InverseDFTSolverFunction.cpp:593-645): one line per v_xc node,
 * "id x y z v0" (" v1" appended for n_spin = 2), values with 17 significant
 * digits.  n_spin is 1 or 2. */
int idft_write_vxc_text(const idft_desc* d, const char* path, const double* pot, int32_t n_spin);

/* InverseDFTEngine::readVxcDataFromFileWithSearch (InverseDFTEngine.cpp:
 * 2982-3142): reads up to n_vxc records "id x y z v0 [v1]"; a record inside
 * the mesh's bounding box widened by 1e-2 is matched to its nearest v_xc
 * node and assigned when closer than dist_tol ("TOL FOR DIST BETWEEN
 * POINTS", default 1e-3).  Two records nearest to one node, or a node left
 * without a record, is IDFT_EINVAL (the reference's two AssertThrows);
 * a missing file is IDFT_EINVAL. */
int idft_read_vxc_text(const idft_desc* d, const char* path, double dist_tol, int32_t n_spin,
                       double* pot);

/* Binary v_xc file (no counterpart in the reference): the 8 bytes
 * "IDFTVXC1", int32 n_cells, int32 n_spin, double half_width, int64 n_vxc,
 * then the n_spin * n_vxc doubles.  Reading checks that the file's mesh is
 * d's mesh (IDFT_EINVAL otherwise). */
int idft_write_vxc_binary(const idft_desc* d, const char* path, const double* pot,
                          int32_t n_spin);
int idft_read_vxc_binary(const idft_desc* d, const char* path, int32_t n_spin, double* pot);

/* Periodic v_xc output of getForceVector (InverseDFTSolverFunction.cpp:
 * 1250-1254, 647-655): with frequency > 0, evaluation number c (counted
 * from 0) with c % frequency == 0 writes its input potential as text
 * records to "<folder>/<postfix>_<c>" ("WRITE VXC DATA", "FOLDER FOR VXC
 * DATA", "FREQUENCY FOR WRITING VXC").  frequency = 0 turns it off. */
int idft_set_vxc_output(idft_ctx* ctx, const char* folder, const char* postfix,
                        int32_t frequency);

/* L-BFGS checkpoint (no counterpart in the reference, whose restart reads
 * a v_xc file and starts with an empty history).  With path non-NULL,
 * idft_bfgs_solve writes after every `every` iterations a binary file with
 * the iterate, its gradient and loss and the (s, y, 1/y.s) history; with
 * resume != 0 and that file present, idft_bfgs_solve continues from the
 * stored state (history included) instead of the initial guess.  path NULL
 * turns both off. */
int idft_bfgs_set_checkpoint(idft_ctx* ctx, const char* path, int32_t every, int32_t resume);

#ifdef __cplusplus
}
#endif
#endif /* IDFT_IO_H */
