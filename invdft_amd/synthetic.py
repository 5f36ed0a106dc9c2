"""Synthetic inverse-DFT problems (soft-Coulomb wells + seeded Gaussian v_xc).

The reference builds its inputs (potBase by a Poisson solve, rho_target from
a Gaussian/Slater density matrix, the initial v_xc from LB94/LDA;
InverseDFTEngine.cpp:1044-1498, 1722-1786, 2457-2630).  These generators make
inputs of the same shape directly:

  potBase(r) = sum_A -Z_A / sqrt(|r - R_A|^2 + 0.5)
  v_xc0(r)   = sum_k a_k exp(-|r - c_k|^2 / (2 s_k^2)),
               a_k ~ U[-0.05, 0.05] Ha, s_k ~ U[1, 2] Bohr, c_k ~ U[-4, 4]^3
  rho_target = spin-up density of potBase alone (v_xc = 0), computed by the
               library itself in FP64, w = 1.
"""
import numpy as np

from . import abi


def soft_coulomb(xyz, atoms):
    v = np.zeros(len(xyz))
    for z, pos in atoms:
        d2 = np.sum((xyz - np.asarray(pos, dtype=np.float64)) ** 2, axis=1)
        v -= z / np.sqrt(d2 + 0.5)
    return v


def gaussian_field(xyz, seed, n=8, amp=0.05, width=(1.0, 2.0), box=4.0):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-amp, amp, n)
    s = rng.uniform(width[0], width[1], n)
    c = rng.uniform(-box, box, (n, 3))
    v = np.zeros(len(xyz))
    for k in range(n):
        d2 = np.sum((xyz - c[k]) ** 2, axis=1)
        v += a[k] * np.exp(-d2 / (2.0 * s[k] ** 2))
    return v


# Workloads (sizes and physics of each named configuration).
CONFIGS = {
    "adjoint_16c_p4": dict(n_cells=16, degree=4, half_width=12.0, n_eig=6, n_electrons=4.0,
                           atoms=[(3.0, (0.0, 0.0, 0.0)), (1.0, (0.0, 0.0, 3.0))]),
    "inverse_24c_p5": dict(n_cells=24, degree=5, half_width=12.0, n_eig=12, n_electrons=16.0,
                           atoms=[(3.0, (0.0, 0.0, 0.0)), (1.0, (0.0, 0.0, 3.0))],
                           cheb_degree=30),
    "batched_40c_p4": dict(n_cells=40, degree=4, half_width=12.0, n_eig=80, n_electrons=128.0,
                           atoms="random8"),
    "gradcheck_12c_p4": dict(n_cells=12, degree=4, half_width=12.0, n_eig=6, n_electrons=8.0,
                             atoms=[(3.0, (0.0, 0.0, 0.0)), (1.0, (0.0, 0.0, 3.0))]),
    "gradcheck_32c_p5": dict(n_cells=32, degree=5, half_width=12.0, n_eig=32, n_electrons=48.0,
                             atoms="random8"),
}


def random_atoms(seed, n=8, box=3.0, charges=(1, 6, 7, 8)):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-box, box, (n, 3))
    z = rng.choice(np.asarray(charges, dtype=np.float64), n)
    return [(float(z[i]), tuple(pos[i])) for i in range(n)]


class Problem:
    """Inputs of one inverse problem, ready for ``abi.Context``."""

    def __init__(self, desc, pot_base_q, rho_target_q, weight_q, rho_target_nodal, vxc0,
                 eigenvectors=None):
        self.desc = desc
        self.pot_base_q = pot_base_q
        self.rho_target_q = rho_target_q
        self.weight_q = weight_q
        self.rho_target_nodal = rho_target_nodal
        self.vxc0 = vxc0
        self.eigenvectors = eigenvectors

    def context(self, lib, **overrides):
        d = abi.Desc()
        C_copy = bytes(self.desc)
        abi.C.memmove(abi.C.byref(d), C_copy, len(C_copy))
        for k, v in overrides.items():
            setattr(d, k, v)
        ctx = abi.Context(lib, d, self.pot_base_q, self.rho_target_q, self.weight_q,
                          self.rho_target_nodal)
        ctx.set_initial_guess(self.vxc0)
        if self.eigenvectors is not None:
            ctx.set_eigenvectors(self.eigenvectors)
        return ctx


def make_problem(lib, n_cells, degree, half_width, n_eig, n_electrons, atoms, seed=1,
                 target_tol=1e-10, max_passes=200, **desc_kw):
    """Build potBase, rho_target (by an FP64 eigensolve with ``lib``), w and v_xc0."""
    if atoms == "random8":
        atoms = random_atoms(seed)
    desc = abi.default_desc(lib, n_cells=n_cells, degree=degree, half_width=half_width,
                            n_eig=n_eig, n_electrons=n_electrons, seed=seed, **desc_kw)
    qxyz = abi.quad_points(lib, desc)
    pot_base = soft_coulomb(qxyz, atoms)
    nq = len(qxyz)
    sizes = abi.geometry(lib, desc)
    w = np.ones(nq)
    # target: eigensolve of potBase alone, always in FP64
    tdesc = abi.default_desc(lib, n_cells=n_cells, degree=degree, half_width=half_width,
                             n_eig=n_eig, n_electrons=n_electrons, seed=seed, **desc_kw)
    tdesc.precision = abi.FP64
    tctx = abi.Context(lib, tdesc, pot_base, np.zeros(nq), w, np.ones(sizes.n_nodes))
    tctx.set_potential(np.zeros(sizes.n_vxc))
    tctx.eigensolve(target_tol, max_passes)
    rho = tctx.density()
    occ, _ = tctx.occupancies()
    psi = tctx.eigenvectors()
    rho_nodal = (psi ** 2) @ occ  # spin-up density at the GLL nodes
    tctx.close()
    vxc0 = gaussian_field(abi.vxc_points(lib, desc), seed + 1000)
    return Problem(desc, pot_base, 0.5 * rho, w, rho_nodal, vxc0, eigenvectors=psi)
