"""Pins of the oracle's adjoint problem and gradient against mathematics."""
import numpy as np
import pytest

from invdft_amd import abi
from tests.helpers import small_problem


def lumped_mass(olib, desc):
    """Tensor-product GLL lumped mass: M_n = m(i) m(j) m(k), m = sum over the
    cells sharing a 1D node of (h/2) w_a."""
    p, nc, L = desc.degree, desc.n_cells, desc.half_width
    gx, gw, _, _ = abi.basis_tables(olib, p)
    h = 2 * L / nc
    m1 = np.zeros(nc * p + 1)
    for c in range(nc):
        m1[c * p: c * p + p + 1] += 0.5 * h * gw
    return np.einsum("k,j,i->kji", m1, m1, m1).reshape(-1)


@pytest.fixture(scope="module")
def adj(olib):
    prob = small_problem(olib)
    ctx = prob.context(olib)
    ctx.set_potential(prob.vxc0)
    ev, _, _ = ctx.eigensolve(1e-12, 100)
    occ, ef = ctx.occupancies()
    psi = ctx.eigenvectors()
    u = prob.rho_target_q - 0.5 * ctx.density()
    B = psi.shape[1]
    pairs = [v for i in range(B) for j in range(B) if abs(ev[i] - ev[j]) < 0.002 for v in (i, j)]
    mpsi = ctx.mv(B, psi)
    ctx.adj_update_input_psi(mpsi, occ, u, pairs, ef, ev)
    M = lumped_mass(olib, prob.desc)
    xyz = abi.node_points(olib, prob.desc)
    free = np.where(~np.any(np.isclose(np.abs(xyz), prob.desc.half_width), axis=1))[0]
    E = np.zeros((ctx.sizes.n_nodes, len(free)))
    return dict(ctx=ctx, prob=prob, ev=ev, occ=occ, ef=ef, psi=psi, B=B, pairs=pairs, mpsi=mpsi,
                M=M, free=free)


def _dense_H(ctx, free, n_nodes):
    cols = []
    for k0 in range(0, len(free), 400):
        idx = free[k0:k0 + 400]
        E = np.zeros((n_nodes, len(idx)))
        E[idx, np.arange(len(idx))] = 1.0
        cols.append(ctx.apply_hamiltonian(E)[free])
    return np.hstack(cols)


def test_psi_is_mass_orthonormal(adj):
    G = adj["psi"].T @ (adj["M"][:, None] * adj["psi"])
    assert np.max(np.abs(G - np.eye(adj["B"]))) < 1e-12


def test_rhs_orthogonal_to_own_orbital(adj):
    ctx, B = adj["ctx"], adj["B"]
    x, rhs = ctx.mv(B), ctx.mv(B)
    ctx.adj_compute_rhs(None, x, rhs)
    b = rhs.get()
    xt = np.sqrt(adj["M"])[:, None] * adj["psi"]
    for i in range(B):
        if np.linalg.norm(b[:, i]) > 0:
            assert abs(b[:, i] @ xt[:, i]) < 1e-10 * np.linalg.norm(b[:, i])


def test_minres_solves_the_shifted_system(adj):
    ctx, B, free = adj["ctx"], adj["B"], adj["free"]
    H = _dense_H(ctx, free, ctx.sizes.n_nodes)
    x, rhs = ctx.mv(B), ctx.mv(B)
    ctx.adj_compute_rhs(None, x, rhs)
    b = rhs.get()[free]
    sol = ctx.mv(B)
    ctx.minres_solve(sol, None, 1e-13, 3000, distribute=False)
    pt = sol.get()[free]
    xt = (np.sqrt(adj["M"])[:, None] * adj["psi"])[free]
    for i in range(B):
        nb = np.linalg.norm(b[:, i])
        if adj["occ"][i] < 1e-8:  # right-hand side below the tolerance: x stays 0
            assert np.linalg.norm(pt[:, i]) == 0
            continue
        A = H - adj["ev"][i] * np.eye(len(free))
        assert np.linalg.norm(A @ pt[:, i] - b[:, i]) < 1e-9 * nb
        ref = np.linalg.lstsq(A, b[:, i], rcond=1e-12)[0]
        P = np.eye(len(free)) - np.outer(xt[:, i], xt[:, i])
        assert np.linalg.norm(P @ (pt[:, i] - ref)) < 1e-7 * np.linalg.norm(ref)


def test_distribute_x_null_space_identity(adj):
    """distributeX sets <p_i, psi_j> to -f'_i mu1_ij + delta_ij f'_i eta (~0 when gapped),
    assuming <psi_j, psi_j> = 1 in the inner product it uses.  The reference uses
    the consistent-mass quadrature, in which the (lumped-orthonormal) orbitals
    have norm 1 + O(lumping error), so the identity holds to that order: the
    dot product shrinks by the factor |<psi_j, psi_j>_q - 1|."""
    ctx, B, M = adj["ctx"], adj["B"], adj["M"]
    npairs = len(adj["pairs"]) // 2
    raw = ctx.mv(B)
    ctx.minres_solve(raw, None, 1e-13, 3000, distribute=False)
    before = ctx.adj_dot_quad_wise(ctx.mv(B, raw.get() / np.sqrt(M)[:, None]), adj["mpsi"], npairs)
    sol = ctx.mv(B)
    ctx.minres_solve(sol, None, 1e-13, 3000, distribute=True)
    after = ctx.adj_dot_quad_wise(sol, adj["mpsi"], npairs)
    norms = ctx.adj_dot_quad_wise(adj["mpsi"], adj["mpsi"], npairs)
    pr = np.array(adj["pairs"]).reshape(-1, 2)
    for k, (i, j) in enumerate(pr):
        if i == j and adj["occ"][i] > 1e-8:
            assert abs(after[k]) <= 2 * abs(norms[k] - 1.0) * abs(before[k]) + 1e-15


def test_distribute_x_exact_mode_lumped_identity(olib):
    """In the exact-gradient mode the projection uses the lumped mass, in which
    the orbitals are orthonormal, so p_i^T M psi_i = 0 to rounding."""
    prob = small_problem(olib)
    ctx = prob.context(olib, gradient_mode=abi.GRADIENT_EXACT)
    ctx.set_potential(prob.vxc0)
    ev, _, _ = ctx.eigensolve(1e-12, 100)
    occ, ef = ctx.occupancies()
    psi = ctx.eigenvectors()
    u = prob.rho_target_q - 0.5 * ctx.density()
    B = psi.shape[1]
    mpsi = ctx.mv(B, psi)
    ctx.adj_update_input_psi(mpsi, occ, u, [v for i in range(B) for v in (i, i)], ef, ev)
    sol = ctx.mv(B)
    ctx.minres_solve(sol, None, 1e-13, 3000, distribute=True)
    p = sol.get()
    M = lumped_mass(olib, prob.desc)
    for i in range(B):
        if occ[i] > 1e-8:
            assert abs(p[:, i] @ (M * psi[:, i])) < 1e-10 * np.linalg.norm(p[:, i])


def test_exact_gradient_matches_central_differences(olib):
    prob = small_problem(olib)
    ctx = prob.context(olib, gradient_mode=abi.GRADIENT_EXACT, cheb_tol=1e-13,
                       initial_tol_cheb=1e-13, adjoint_initial_tol=1e-14, rho_tol_constraints=0.0)
    v = prob.vxc0
    g, L = ctx.get_force(v)
    rng = np.random.default_rng(3)
    for _ in range(3):
        dvec = rng.standard_normal(v.size)
        dvec /= np.linalg.norm(dvec)
        h = 1e-4
        fd = (ctx.forward(v + h * dvec) - ctx.forward(v - h * dvec)) / (2 * h)
        assert abs(fd - g @ dvec) < 2e-7 * abs(fd) + 1e-14


def test_reference_gradient_is_close_to_exact(olib):
    """The reference's child-quadrature force differs from the exact gradient by
    quadrature / mass-lumping error only: same direction, same scale."""
    prob = small_problem(olib)
    kw = dict(cheb_tol=1e-12, initial_tol_cheb=1e-12)
    ge, _ = prob.context(olib, gradient_mode=abi.GRADIENT_EXACT, **kw).get_force(prob.vxc0)
    gr, _ = prob.context(olib, gradient_mode=abi.GRADIENT_REFERENCE, **kw).get_force(prob.vxc0)
    cos = ge @ gr / (np.linalg.norm(ge) * np.linalg.norm(gr))
    assert cos > 0.9 and 0.5 < np.linalg.norm(gr) / np.linalg.norm(ge) < 2.0


def test_bfgs_decreases_loss(olib):
    prob = small_problem(olib)
    ctx = prob.context(olib, cheb_tol=1e-10, initial_tol_cheb=1e-10)
    loss, gnorm, n = ctx.bfgs_solve(max_iter=3)
    assert n == 3 and loss[-1] < 0.5 * loss[0]


def test_nan_in_density_difference_is_reported(olib):
    prob = small_problem(olib)
    ctx = prob.context(olib)
    ctx.set_potential(prob.vxc0)
    ev, _, _ = ctx.eigensolve(1e-9, 100)
    occ, ef = ctx.occupancies()
    psi = ctx.eigenvectors()
    u = np.zeros(ctx.nq_total)
    u[5] = np.nan
    B = psi.shape[1]
    mpsi = ctx.mv(B, psi)
    ctx.adj_update_input_psi(mpsi, occ, u, [v for i in range(B) for v in (i, i)], ef, ev)
    with pytest.raises(abi.IdftError) as e:
        ctx.minres_solve(ctx.mv(B), None, 1e-10, 100)
    assert e.value.code == abi.IDFT_ENAN
