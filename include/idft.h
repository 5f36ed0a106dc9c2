/*
 * idft.h -- C ABI of the MI355X inverse-DFT hot path.
 *
 * Two libraries implement this header:
 *   invdft_amd/_lib/libidft.so   the product (HIP kernels for gfx950 + C++ host)
 *   oracle/_build/liboracle.so   the FP64 CPU restatement used only by tests
 *                                and by the CPU-baseline leg of bench.py
 *
 * Conventions
 *   - Every entry point returns an int status (IDFT_OK == 0, negative on
 *     error); idft_last_error(ctx) gives the message.  No C++ exception
 *     crosses this boundary.  The reference throws (AssertThrow /
 *     dftfe::utils::throwException) where these return an error code.
 *   - Host arrays are caller-owned and copied.  Multivector handles
 *     (idft_mv) live in device memory for the product and are node-major:
 *     value(node, col) at [node * ncols + col], the layout of
 *     dftfe::linearAlgebra::MultiVector (MultiVectorAdjointLinearSolverProblem.cpp:1198-1200).
 *   - One host thread per context; calls are host-synchronous on return.
 *   - Spin-unpolarised, Gamma point only (n_spin == 1, k-weight 1).
 *
 * Discretisation fixed by this ABI (the reference delegates it to deal.II /
 * DFT-FE, which are not part of the reference repository):
 *   domain [-L, L]^3, n_cells^3 uniform hexahedra, degree-p Lagrange basis on
 *   Gauss-Lobatto-Legendre (GLL) nodes, homogeneous Dirichlet boundary for the
 *   orbitals, lumped (GLL-quadrature) mass matrix M, parent quadrature
 *   Gauss (p+3)^3 per cell for V_e, density, loss and adjoint right-hand side;
 *   v_xc lives on a linear (Q1) child mesh with 2 child cells per parent edge
 *   and Gauss 2^3 points per child cell.  Global node (i,j,k) has index
 *   (k*N1 + j)*N1 + i with N1 = n_cells*p + 1; cell (cx,cy,cz) has index
 *   (cz*n_cells + cy)*n_cells + cx; quadrature data are [cell][q] with
 *   q = (qz*Q + qy)*Q + qx, Q = p+3.
 */
#ifndef IDFT_H
#define IDFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  IDFT_OK = 0,
  IDFT_EINVAL = -1,   /* invalid argument or size                        */
  IDFT_EHIP = -2,     /* HIP runtime error                               */
  IDFT_ENOCONV = -3,  /* solver did not converge within its iteration cap */
  IDFT_ENAN = -4,     /* NaN/Inf detected in an input or an iterate       */
  IDFT_ESTATE = -5,   /* call order violated (e.g. vmult before update)   */
  IDFT_ENOMEM = -6    /* allocation failed                                */
};

enum { IDFT_FP64 = 0, IDFT_MIXED = 1 };
/* Gradient mode.
 * REFERENCE: the reference's semantics -- force integrated on the child-mesh
 *   Gauss-2 points (InverseDFTSolverFunction.cpp:1161-1246, 1609-1674) and the
 *   null-space component of p fixed with the consistent-mass quadrature inner
 *   product (MultiVectorAdjointLinearSolverProblem.cpp:1021, 1143-1243).
 * EXACT: the exact gradient of the discrete loss -- force on the parent
 *   quadrature through which v_xc enters H, null-space component fixed in the
 *   lumped-mass inner product in which the orbitals are orthonormal.  Used by
 *   finite-difference checks; differs from REFERENCE by quadrature and
 *   mass-lumping error only. */
enum { IDFT_GRADIENT_REFERENCE = 0, IDFT_GRADIENT_EXACT = 1 };
/* BFGSInverseDFTSolver::LSType (BFGSInverseDFTSolver.h:31). */
enum { IDFT_LS_SECANT_LOSS = 0, IDFT_LS_SECANT_FORCE_NORM = 1, IDFT_LS_CP = 2 };

typedef struct idft_desc {
  int32_t n_cells;            /* parent cells per axis                          */
  int32_t degree;             /* p, 2..5                                        */
  double half_width;          /* L                                              */
  int32_t n_eig;              /* Kohn-Sham states solved (occupied + buffer)    */
  double n_electrons;         /* N_e (spin-unpolarised: 2 per orbital)          */
  int32_t precision;          /* IDFT_FP64 | IDFT_MIXED                         */
  double temperature;         /* TVal [K] for Fermi-Dirac occupancies           */
  int32_t cheb_degree;        /* Chebyshev filter degree                        */
  int32_t max_cheb_passes;    /* maxChebPasses (inverseDFTParameters.cpp:412)   */
  double cheb_tol;            /* DFT-FE chebyshevTolerance                      */
  double initial_tol_cheb;    /* initialTolForChebFiltering (.cpp:411)          */
  double adaptive_factor_cheb;    /* adaptiveFactorForChebFiltering (.cpp:410)  */
  double adaptive_factor_adjoint; /* adaptiveFactorForAdjoint (.cpp:409)        */
  double adjoint_initial_tol; /* inverseAdjointInitialTol (.cpp:416)            */
  int32_t adjoint_max_iter;   /* inverseAdjointMaxIterations (.cpp:417)         */
  double degeneracy_tol;      /* inverseDegeneracyTol (.cpp:424)                */
  double frac_occ_tol;        /* inverseFractionOccTol (.cpp:423)               */
  int32_t additional_states;  /* additionalEigenStatesSolved (.cpp:428)         */
  int32_t block_size;         /* DFT-FE chebyWfcBlockSize (adjoint block)       */
  double rho_tol_constraints; /* rhoTolForConstraints (.cpp:413)                */
  int32_t gradient_mode;      /* IDFT_GRADIENT_*                                */
  int32_t lanczos_steps;      /* steps of the spectrum upper-bound Lanczos      */
  int32_t psi_checkpoint;     /* HIP: 0 keeps the FP64 eigenvector block resident
                                 between the forward and the backward pass; 1
                                 keeps only an FP32 copy of it (the checkpoint),
                                 frees the eigensolver's work buffers, and
                                 recomputes the FP64 block in the backward pass
                                 by one Chebyshev filter pass, orthonormalisation
                                 and Rayleigh-Ritz from the checkpoint (less
                                 device memory, one extra filter pass).  The
                                 oracle keeps its vectors either way.         */
  int32_t check_every;        /* MINRES steps between host convergence checks   */
  uint64_t seed;              /* seed of the initial random subspace            */
  int32_t minres_graph;       /* HIP: replay check_every MINRES steps as one
                                 captured HIP graph (1) or launch each kernel
                                 (0); same results either way                  */
  double relax_ratio;         /* MIXED: adjoint MINRES switches from the FP64 to
                                 the FP32 cell operator once every open
                                 column's residual estimate is below
                                 relax_ratio times the tolerance (0: never)    */
  int32_t minres_fused;       /* HIP: with sum-factorised cell products, take
                                 each MINRES step's alpha from the cell kernel
                                 and run the Lanczos recurrence inside the node
                                 gather (1), or store H~v and run it separately
                                 (0); equal to rounding                        */
} idft_desc;

typedef struct idft_sizes {
  int64_t n_nodes;          /* N = N1^3 (boundary nodes included, held zero)  */
  int64_t n_cells;          /* parent cells                                   */
  int64_t nodes_per_cell;   /* (p+1)^3                                        */
  int64_t quad_per_cell;    /* (p+3)^3                                        */
  int64_t n_vxc;            /* child nodes (2 n_cells + 1)^3                  */
  int64_t child_quad_per_cell; /* child Gauss points inside one parent cell (64) */
} idft_sizes;

typedef struct idft_stats {
  double t_eigen_ms, t_density_ms, t_adjoint_ms, t_force_ms, t_total_ms;
  int32_t cheb_passes;        /* passes in the last solveEigen                  */
  int32_t minres_iters_max;   /* over columns, last evaluation                  */
  int64_t minres_iters_total; /* sum over columns, last evaluation              */
  int32_t refinements;        /* FP64 refinement sweeps (MIXED)                 */
  double max_residual;        /* eigen residual at exit of the last solveEigen  */
  double fermi_energy;
  double cheb_tol_used, adjoint_tol_used;
  double loss_unweighted, l1_density_error, int_rho_veff;
  int64_t n_evaluations;
  double spectrum_upper_bound;
  int64_t applies_fp64;       /* Hamiltonian applications (cumulative)          */
  int64_t applies_fp32;
  int64_t apply_cols_fp64;    /* sum of block widths over those applications    */
  int64_t apply_cols_fp32;
  int32_t adjoint_blocks;     /* blocks of the last backward pass after the
                                 degeneracy-preserving shrink (ISF:930-955)   */
  int32_t psi_recomputes;     /* checkpoint recomputations (cumulative)       */
  int32_t minres_compactions; /* HIP: repacks of the MINRES block to its open
                                 columns (cumulative)                         */
  int64_t minres_column_steps; /* HIP: MINRES steps x columns streamed
                                  (cumulative)                                */
} idft_stats;

typedef struct idft_ctx idft_ctx;
typedef struct idft_mv idft_mv;

/* ---- descriptor, geometry (no context needed) ---------------------------- */
void idft_desc_default(idft_desc* d);
int idft_geometry_sizes(const idft_desc* d, idft_sizes* s);
/* xyz[3*i+c] of every parent quadrature point, layout [cell][q]. */
int idft_geometry_quad_points(const idft_desc* d, double* xyz);
/* Parent FE node coordinates, [node][3]. */
int idft_geometry_node_points(const idft_desc* d, double* xyz);
/* Child (v_xc) node coordinates, [child node][3]. */
int idft_geometry_vxc_points(const idft_desc* d, double* xyz);
/* 1D tables: GLL nodes/weights (p+1) and Gauss nodes/weights (p+3) on [-1,1]. */
int idft_basis_tables(int32_t p, double* gll_x, double* gll_w, double* gauss_x,
                      double* gauss_w);

/* ---- lifecycle ------------------------------------------------------------
 * Replaces InverseDFTSolverFunction::reinit (InverseDFTSolverFunction.cpp:108-450)
 * plus MultiVectorAdjointLinearSolverProblem::reinit
 * (MultiVectorAdjointLinearSolverProblem.cpp:475-538).
 *   pot_base_q       potBase at parent quadrature            [cells*Q^3]
 *   rho_target_q     spin-up target density rho_t            [cells*Q^3]
 *   weight_q         loss weight w                           [cells*Q^3]
 *   rho_target_nodal spin-up target density at FE nodes; nodes with a value
 *                    below rho_tol_constraints form the adjoint constraints
 *                    (InverseDFTEngine.cpp:1547-1573)         [N]          */
int idft_create(const idft_desc* d, const double* pot_base_q,
                const double* rho_target_q, const double* weight_q,
                const double* rho_target_nodal, idft_ctx** out);
void idft_destroy(idft_ctx* ctx);
const char* idft_last_error(const idft_ctx* ctx);
int idft_get_sizes(const idft_ctx* ctx, idft_sizes* s);
int idft_get_stats(const idft_ctx* ctx, idft_stats* s);
/* Backend name: "hip-gfx950" for the product, "cpu-oracle" for the oracle. */
const char* idft_backend(void);

/* ---- contract 1: value / force ------------------------------------------ */
/* InverseDFTSolverFunction::getForceVector (InverseDFTSolverFunction.h:146-148,
 * .cpp:657-1311).  pot [n_vxc] in/out (the reference distributes and zeroes
 * constrained entries; the uniform child mesh has none), force [n_vxc] out,
 * loss [1] out (the weighted loss of the spin channel). */
int idft_get_force(idft_ctx* ctx, double* pot, double* force, double* loss);
/* InverseDFTSolverFunction::dotProduct (.cpp:1691-1705): out[i % block] sums. */
int idft_dot(idft_ctx* ctx, const double* v1, const double* v2, int64_t n,
             int32_t block, double* out);
/* getInitialGuess / setInitialGuess / setSolution (.h:140-150). */
int idft_set_initial_guess(idft_ctx* ctx, const double* pot);
int idft_get_initial_guess(const idft_ctx* ctx, double* pot);
int idft_set_solution(idft_ctx* ctx, const double* pot);
int idft_get_solution(const idft_ctx* ctx, double* pot);
/* Forward / backward split of getForceVector: forward = solveEigen + density
 * + loss (.cpp:683-912); backward = blocked adjoint MINRES + force
 * (.cpp:914-1246).  forward must precede backward. */
int idft_forward(idft_ctx* ctx, const double* pot, double* loss);
int idft_backward(idft_ctx* ctx, double* force);

/* ---- pieces of the forward pass (tests, benchmarks) ---------------------- */
/* v_KS = potBase + interpolated v_xc at parent quadrature, then the cell
 * Hamiltonians H_e = K_e/2 + V_e (InverseDFTSolverFunction.cpp:1355-1388). */
int idft_set_potential(idft_ctx* ctx, const double* pot);
/* Set v_KS at parent quadrature directly and rebuild H_e. */
int idft_set_veff_quad(idft_ctx* ctx, const double* veff_q);
/* The v_KS of the current Hamiltonian at parent quadrature [cells*Q^3]
 * (potBase + interpolated v_xc after idft_set_potential; the reference forms
 * it at InverseDFTSolverFunction.cpp:1360-1381). */
int idft_get_veff_quad(const idft_ctx* ctx, double* veff_q);
/* ChFSI passes on the current Hamiltonian until the max residual over the
 * occupied states is below tol or max_passes is reached
 * (InverseDFTSolverFunction.cpp:1390-1474). */
int idft_eigensolve(idft_ctx* ctx, double tol, int32_t max_passes,
                    double* eigenvalues, double* max_residual, int32_t* passes);
/* Physical orbitals psi = M^{-1/2} x (node-major [N][n_eig]), M-orthonormal. */
int idft_get_eigenvectors(const idft_ctx* ctx, double* psi);
int idft_set_eigenvectors(idft_ctx* ctx, const double* psi);
int idft_get_occupancies(const idft_ctx* ctx, double* occ, double* fermi);
/* Ritz values of the current eigenvector block [n_eig] (after the last
 * Rayleigh-Ritz; IDFT_ESTATE before any). */
int idft_get_eigenvalues(const idft_ctx* ctx, double* eigenvalues);
/* Total density rho = 2 sum_i f_i psi_i^2 at parent quadrature [cells*Q^3]. */
int idft_get_density(idft_ctx* ctx, double* rho_q);
/* y = M^{-1/2} H M^{-1/2} x (Loewdin Hamiltonian), boundary rows/cols zero.
 * precision selects the FP64 or the FP32 cell operator. */
int idft_apply_hamiltonian(idft_ctx* ctx, const double* x, double* y,
                           int32_t ncols, int32_t precision);
/* Scaled Chebyshev filter of the given degree on [a, b], scaled at a0,
 * applied in place to x ([N][ncols]).  precision as above. */
int idft_chebyshev_filter(idft_ctx* ctx, double* x, int32_t ncols,
                          int32_t degree, double a, double b, double a0,
                          int32_t precision);

/* ---- contract 2: the adjoint problem ------------------------------------- */
int idft_mv_create(idft_ctx* ctx, int32_t ncols, idft_mv** out);
void idft_mv_destroy(idft_mv* mv);
int idft_mv_set(idft_ctx* ctx, idft_mv* mv, const double* host);
int idft_mv_get(idft_ctx* ctx, const idft_mv* mv, double* host);

/* MultiVectorAdjointLinearSolverProblem::updateInputPsi
 * (MultiVectorAdjointLinearSolverProblem.h:112-118, .cpp:817-910).
 * psi is borrowed (the handle must outlive the solve), occ = f_i * w_k,
 * u_q = w (rho_t - rho/2) at parent quadrature, pairs = (i, j) for every j in
 * the degeneracy group of i (2*n_pairs int32), eigenvalues [block] are NEGATED
 * IN PLACE as the reference does (.cpp:900-902). */
int idft_adj_update_input_psi(idft_ctx* ctx, idft_mv* psi, const double* occ,
                              const double* u_q, const int32_t* pairs,
                              int32_t n_pairs, double fermi_energy,
                              double* eigenvalues, int32_t block);
/* computeRhs (.h:41-48, .cpp:614-812): rhs <- b~, remembers x for distributeX. */
int idft_adj_compute_rhs(idft_ctx* ctx, idft_mv* ndbc, idft_mv* x, idft_mv* rhs);
/* vmult (.h:54-59, .cpp:1345-1367): Ax = H~ x - eps_i x_i, constrained zero. */
int idft_adj_vmult(idft_ctx* ctx, idft_mv* Ax, idft_mv* x, int32_t block);
/* precondition_Jacobi / precondition_JacobiSqrt (.h:70-88, .cpp:587-612):
 * dst = D src with D = 2 M_ii / K_ii (or sqrt(D)); omega is ignored. */
int idft_adj_precondition_jacobi(idft_ctx* ctx, idft_mv* dst, const idft_mv* src,
                                 double omega);
int idft_adj_precondition_jacobi_sqrt(idft_ctx* ctx, idft_mv* dst,
                                      const idft_mv* src, double omega);
/* distributeX (.h:64, .cpp:1002-1141) on the x remembered by computeRhs. */
int idft_adj_distribute_x(idft_ctx* ctx);
/* multiVectorDotProdQuadWise (.h:101-110, .cpp:1143-1243): out[pair]. */
int idft_adj_dot_quad_wise(idft_ctx* ctx, idft_mv* v1, idft_mv* v2, double* out);
/* dftfe::MultiVectorMinResSolver::solve as called at
 * InverseDFTSolverFunction.cpp:1106-1112: computeRhs, preconditioned MINRES
 * from x = 0 until every column's preconditioned residual norm is <= tol,
 * then distributeX if distribute != 0.  iters[block] receives per-column
 * iteration counts (may be NULL). */
int idft_minres_solve(idft_ctx* ctx, idft_mv* x, idft_mv* ndbc, double tol,
                      int32_t max_iter, int32_t distribute, int32_t* iters);

/* ---- optimiser (host driver over contract 1) ------------------------------ */
/* BFGSInverseDFTSolver::solve (BFGSInverseDFTSolver.cpp:465-627) over this
 * context's getForceVector.  loss_hist / gnorm_hist [max_iter+1] (may be
 * NULL) receive the loss and force norm at the start of each iteration,
 * time_hist_ms [max_iter+1] (may be NULL) the wall-clock time (ms, since the
 * call began) at the start of each iteration and at the end; n_done the
 * iterations completed; the final potential is kept as the solution
 * (idft_get_solution). */
int idft_bfgs_solve(idft_ctx* ctx, int32_t ls_type, int32_t max_iter, double tol,
                    double ls_tol, int32_t history, int32_t n_line_search,
                    double* loss_hist, double* gnorm_hist, double* time_hist_ms,
                    int32_t* n_done);

#ifdef __cplusplus
}
#endif
#endif /* IDFT_H */
