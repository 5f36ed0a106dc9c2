/* idft_setup.h -- the set-up stage that produces the hot path's inputs:
 * the target density from an atomic-orbital (Gaussian) density matrix, the
 * loss weights, the electrostatic base potential by a Poisson solve and the
 * initial v_xc guess.  Implemented by libidft.so (HIP, gfx950) and by the CPU
 * oracle; every entry point is host-synchronous and returns an IDFT_* status.
 *
 * Units are atomic (Bohr, Hartree).  Quadrature arrays are the parent
 * quadrature of the mesh described by an idft_desc, [cells][Q^3] in the
 * order of idft_geometry_quad_points; nodal arrays follow
 * idft_geometry_node_points (parent) and idft_geometry_vxc_points (child). */
#ifndef IDFT_SETUP_H
#define IDFT_SETUP_H

#include "idft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------- atomic-orbital target density
 * The basis of gaussianFunctionManager (src/gaussian/gaussianFunctionManager.cpp):
 * one function per (shell, m), p shells ordered m = 1, -1, 0 and the others
 * m = -l..l (storeBasisFunctions, :904-1005); value
 *   phi(x) = C_lm D_m P_l^|m|(cos theta) Q_m(phi) sum_k c_k N_k r^l exp(-a_k r^2)
 * about the atom (evaluateBasisValue, :416-446), N_k the primitive norm
 * (getNormConsts, :251-262), P without the Condon-Shortley phase. */
typedef struct idft_ao_basis idft_ao_basis;

/* readAtomicCoordsAndBasisFileNames + readGaussianFiles (:660-825): one atom
 * per line "symbol x y z basis-file" of coords_file; basis files are opened
 * relative to basis_dir (NULL or "": as written); angstrom != 0 converts the
 * coordinates to Bohr (the reference passes unit 'A', InverseDFTEngine.cpp:566-570).
 * IDFT_EINVAL on a missing file, a malformed shell or a non-numeric entry. */
int idft_ao_basis_read(const char* coords_file, const char* basis_dir, int32_t angstrom,
                       idft_ao_basis** out);
void idft_ao_basis_destroy(idft_ao_basis* b);
/* Number of basis functions (the density matrix is n x n). */
int idft_ao_basis_size(const idft_ao_basis* b, int32_t* n_basis);
/* readMatrix (:827-852): n rows of n whitespace-separated numbers. */
int idft_ao_read_matrix(const char* path, int32_t n, double* out);
/* Basis values at host points xyz [n_points][3] into values [n_basis][n_points]
 * (with the basis normalisation of the last idft_ao_density_quad call that
 * asked for it, else 1).  For tests; the density entry point never leaves
 * the device. */
int idft_ao_values(const idft_ao_basis* b, int64_t n_points, const double* xyz,
                   double* values);
/* evaluateForQuad + getRhoValue (:1183-1430) on the parent quadrature of the
 * mesh of d.  normalize != 0: each function is rescaled by 1/sqrt(S_ii), S_ii
 * = sum_q phi_i^2 JxW on this quadrature (normalizeBasis; the scale is kept in
 * b).  For each spin s, rho_q[s][q] = sum_ij P_s,ij phi_i phi_j with P_s the
 * symmetric matrix of the upper triangle of P[s] (the reference sums
 * P_ii phi_i^2 + 2 sum_{i<j} P_ij phi_i phi_j).  grad_q (may be NULL):
 * [s][q][3] gradient of rho_s (getRhoGradient, :1481-1516).  s_diag (may be
 * NULL): S_ii before the rescaling. */
int idft_ao_density_quad(const idft_desc* d, idft_ao_basis* b, int32_t normalize, int32_t n_spin,
                         const double* P, double* rho_q, double* grad_q, double* s_diag);

/* ------------------------------------------------------------ loss weights
 * InverseDFTEngine.cpp:3661-3675: w = 1/((f rho)^alpha1 + tau) + (f rho)^alpha2
 * with f the spin factor (2 spin-unpolarised, 1 polarised). */
int idft_loss_weights(int64_t n, const double* rho_spin, double spin_factor, double alpha1,
                      double alpha2, double tau, double* w);

/* ----------------------------------------------------------- electrostatics
 * Potential of a charge density on the degree-p mesh of d (the reference's
 * setPotBasePoissonNuclear / computeHartreePotOnParentQuad,
 * InverseDFTEngine.cpp:2457-2624, a CG solve of DFT-FE's poissonSolverProblem):
 *   -lap phi = 4 pi c,   phi = g on the boundary,
 * c the charge density at parent quadrature (electrons positive), by
 * Jacobi-preconditioned CG on the stiffness matrix to ||r||_D <= tol ||b||_D.
 * boundary_nodal (may be NULL): g at the parent nodes (only boundary entries
 * are read); NULL takes the monopole + dipole field of c about the origin
 * (multipole boundary conditions).  phi_q [cells*Q^3] and phi_nodal [N]
 * (either may be NULL) receive the solution. */
typedef struct idft_poisson_stats {
  int32_t iterations;
  double rel_residual;   /* final ||r||_D / ||b||_D                              */
  double charge;         /* int c                                                */
  double dipole[3];      /* int c x                                              */
} idft_poisson_stats;
int idft_poisson_solve(const idft_desc* d, const double* charge_q, const double* boundary_nodal,
                       double tol, int32_t max_iter, double* phi_q, double* phi_nodal,
                       idft_poisson_stats* st);

/* ------------------------------------------------------ initial v_xc guess
 * setInitialPotL2Proj (InverseDFTEngine.cpp:1044-1311, spin-unpolarised):
 * at each parent quadrature point, with r = spin_factor rho_spin,
 *   v = v_x(r [, sigma]) + r/(r + tau_vx) v_c(r),
 * v_x the LB94 exchange potential (lb94 != 0, needs sigma = |grad r|^2) or
 * Slater exchange, v_c the Perdew-Wang 1992 correlation potential; with
 * fermi_amaldi != 0 the far field blends to the Fermi-Amaldi potential:
 *   v <- r/(r + tau_fa) v - (1 - r/(r + tau_fa)) factor_fa v_H / n_electrons.
 * v_lda_q [cells*Q^3] receives v (d_vxcLDAQuadData).  vxc_child [n_vxc]
 * receives the guess on the child v_xc mesh: L2 projection of v onto the
 * parent nodes (zero boundary), its values at the child Gauss points scaled
 * by (1 - factor_lda), L2 projection onto the child Q1 nodes (:1311-1431). */
typedef struct idft_guess_params {
  int32_t lb94;          /* USE LB94_X IN INITIAL GUESS (default 1)              */
  double tau_vx;         /* TAU FOR WEIGHTS FOR SETTING VX BC (1e-2)             */
  int32_t fermi_amaldi;  /* SET FERMIAMALDI IN THE FAR FIELD AS INPUT (1)        */
  double tau_fa;         /* TAU FOR WEIGHTS FOR SETTING FABC (1e-2)              */
  double factor_fa;      /* FACTOR FOR FERMIAMALDI (1.0)                         */
  double factor_lda;     /* FACTOR FOR LDA VXC (0.0)                             */
  double spin_factor;    /* 2 spin-unpolarised                                   */
  double n_electrons;
} idft_guess_params;
void idft_guess_params_default(idft_guess_params* gp);
/* XC potentials alone at n points (tests): v_x (lb94 selects LB94, sigma may
 * be NULL for Slater) and v_c of the total density r. */
int idft_xc_potentials(int64_t n, const double* r, const double* sigma, int32_t lb94,
                       double* v_x, double* v_c);
/* The two projections of the guess, separately: L2 projection of parent
 * quadrature values onto the parent nodes (consistent mass, zero boundary
 * values), and a parent nodal field to the child Q1 nodes (L2 projection of
 * its values at the child Gauss points, times scale). */
int idft_l2_project_parent(const idft_desc* d, const double* v_q, double* x_nodal);
int idft_parent_to_child(const idft_desc* d, const double* x_nodal, double scale,
                         double* vxc_child);
int idft_initial_vxc(const idft_desc* d, const idft_guess_params* gp, const double* rho_spin_q,
                     const double* sigma_q, const double* v_hartree_q, double* v_lda_q,
                     double* vxc_child);

#ifdef __cplusplus
}
#endif
#endif
