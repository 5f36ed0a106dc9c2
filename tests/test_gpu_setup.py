"""Set-up stage on the GPU (gfx950) against the oracle, on the same inputs:
atomic-orbital basis values and the LiH target density (with gradient and
the quadrature renormalisation), loss weights, exchange-correlation
potentials, the Poisson solve, both L2 projections and the initial guess;
then the whole set-up feeding getForceVector / BFGS on the LiH demo inputs.
"""
import os

import numpy as np
import pytest

from invdft_amd import abi
from tests.helpers import rel
from tests.setup_abi import (ao_basis, ao_density, ao_values, initial_vxc, l2_project_parent,
                             loss_weights, parent_to_child, poisson, xc_potentials)

pytestmark = pytest.mark.gpu

LIH = os.path.join(os.path.dirname(__file__), "golden", "lih")


def _say(*a):
    print("CASE", *a, flush=True)


@pytest.fixture(scope="module")
def P():
    return np.load(os.path.join(LIH, "lih_matrices.npz"))["density_matrix"]


def _desc(lib, nc, p, L):
    return abi.default_desc(lib, n_cells=nc, degree=p, half_width=L, n_eig=2, n_electrons=2.0)


def test_ao_values_match_oracle(glib, olib):
    _say("ao_values")
    coords = os.path.join(LIH, "AtomicCoords")
    bg, bo = ao_basis(glib, coords, LIH), ao_basis(olib, coords, LIH)
    assert bg.n == bo.n == 176
    rng = np.random.default_rng(5)
    pts = rng.uniform(-3, 3, (300, 3))
    pts[:3] = [[0, 0, 0], [0, 0, 3.0139242], [0, 0, 1.0]]
    vg, vo = ao_values(glib, bg, pts), ao_values(olib, bo, pts)
    scale = np.max(np.abs(vo), axis=1, keepdims=True)
    assert np.max(np.abs(vg - vo) / scale) < 1e-12


@pytest.mark.parametrize("normalize", [False, True])
def test_ao_density_matches_oracle(glib, olib, P, normalize):
    _say("ao_density", normalize)
    coords = os.path.join(LIH, "AtomicCoords")
    d = _desc(olib, 3, 4, 4.0)
    out = {}
    for name, lib in (("g", glib), ("o", olib)):
        b = ao_basis(lib, coords, LIH)
        out[name] = ao_density(lib, d, b, P, normalize=normalize, gradient=True)
    (rg, gg, sg), (ro, go, so) = out["g"], out["o"]
    assert np.allclose(sg, so, rtol=1e-12, atol=0)
    assert rel(rg, ro) < 1e-12
    assert rel(gg, go) < 1e-11


def test_weights_and_xc_match_oracle(glib, olib):
    _say("weights_xc")
    rng = np.random.default_rng(6)
    rho = 10.0 ** rng.uniform(-8, 1, 4000)
    sig = rho ** (8 / 3) * rng.uniform(0, 5, 4000)
    for a1, a2 in ((0.0, 0.0), (0.5, 0.25)):
        assert rel(loss_weights(glib, rho, 2.0, a1, a2, 1e-2), loss_weights(olib, rho, 2.0, a1, a2, 1e-2)) < 1e-15
    for lb in (False, True):
        xg = xc_potentials(glib, rho, sig, lb94=lb)
        xo = xc_potentials(olib, rho, sig, lb94=lb)
        assert rel(xg[0], xo[0]) < 1e-13 and rel(xg[1], xo[1]) < 1e-13


@pytest.mark.parametrize("p", [4, 5])
def test_poisson_matches_oracle(glib, olib, p):
    _say("poisson", p)
    d = _desc(olib, 6, p, 6.0)
    xyz = abi.quad_points(olib, d)
    r2 = np.sum((xyz - [0.3, -0.2, 0.5]) ** 2, axis=1)
    c = 1.5 * (0.5 / np.pi) ** 1.5 * np.exp(-0.5 * r2) - 2.0 * 3 * 2.0 / (4 * np.pi * (np.sum(xyz ** 2, 1) + 2.0) ** 2.5)
    pg, ng, sg = poisson(glib, d, c, tol=1e-12)
    po, no, so = poisson(olib, d, c, tol=1e-12)
    assert abs(sg.charge - so.charge) < 1e-12 and np.allclose(sg.dipole, so.dipole, atol=1e-12)
    assert rel(pg, po) < 1e-10 and rel(ng, no) < 1e-10
    assert sg.rel_residual <= 1e-12


def test_projections_and_guess_match_oracle(glib, olib):
    _say("projections_guess")
    for p in (4, 5):
        d = _desc(olib, 4, p, 4.0)
        xyz = abi.quad_points(olib, d)
        v = np.exp(-np.sum(xyz ** 2, 1) / 3.0) * (1 + 0.2 * xyz[:, 0])
        xg, xo = l2_project_parent(glib, d, v), l2_project_parent(olib, d, v)
        assert rel(xg, xo) < 1e-11
        cg, co = parent_to_child(glib, d, xo, 0.8), parent_to_child(olib, d, xo, 0.8)
        assert rel(cg, co) < 1e-12
        rho = 0.5 * np.exp(-np.sum(xyz ** 2, 1))
        sigma = (2 * rho * 2 * np.linalg.norm(xyz, axis=1)) ** 2
        vh = 2.0 / np.sqrt(np.sum(xyz ** 2, 1) + 1.0)
        gg = initial_vxc(glib, d, rho, sigma, vh, n_electrons=2.0)
        go = initial_vxc(olib, d, rho, sigma, vh, n_electrons=2.0)
        assert rel(gg[0], go[0]) < 1e-13 and rel(gg[1], go[1]) < 1e-11


def test_lih_setup_feeds_the_inverse_problem(glib, olib, P):
    """The reference demo's LiH density matrix through the whole set-up on
    the GPU (target density with gradient, Poisson potBase of the target plus
    soft-Coulomb nuclei, weights, LB94 + Fermi-Amaldi guess), then two BFGS
    iterations of the inverse problem from that guess; the first evaluation
    (loss and force) equals the oracle's on the same inputs."""
    _say("lih_end_to_end")
    coords = os.path.join(LIH, "AtomicCoords")
    kw = dict(n_cells=6, degree=4, half_width=6.0, n_eig=4, n_electrons=4.0)
    d = abi.default_desc(glib, **kw)
    b = ao_basis(glib, coords, LIH)
    rho, grad, _ = ao_density(glib, d, b, P, normalize=True, gradient=True)
    xyz = abi.quad_points(glib, d)
    # nuclei as Plummer charges (soft-Coulomb wells, a^2 = 0.5)
    nuc = np.zeros(len(xyz))
    for Z, R in ((3.0, (0.0, 0.0, 0.0)), (1.0, (0.0, 0.0, 3.0139242))):
        r2 = np.sum((xyz - R) ** 2, 1)
        nuc += Z * 3 * 0.5 / (4 * np.pi * (r2 + 0.5) ** 2.5)
    pot_base, _, st = poisson(glib, d, 2.0 * rho - nuc, tol=1e-11)
    vh, _, _ = poisson(glib, d, 2.0 * rho, tol=1e-11)
    w = loss_weights(glib, rho, 2.0, 0.0, 0.0, 1e-2)
    sigma = 4.0 * np.sum(grad ** 2, axis=1)
    _, v0 = initial_vxc(glib, d, rho, sigma, vh, n_electrons=4.0)
    nodes = abi.node_points(glib, d)
    rho_nodal = np.full(len(nodes), 1.0)  # no adjoint mask in this check
    out = {}
    # eigenpairs to 1e-10 and an adjoint tolerance of 1e-10 (the default rule,
    # min(1e-11, tol_cheb / 1000) = 1e-13, is below what the coarse LiH
    # systems reach in 5000 steps)
    tight = dict(kw, cheb_tol=1e-10, initial_tol_cheb=1e-10, adjoint_initial_tol=1e-10,
                 adaptive_factor_adjoint=1.0)
    for name, lib in (("g", glib), ("o", olib)):
        ctx = abi.Context(lib, abi.default_desc(lib, **tight), pot_base, rho, w, rho_nodal)
        out[name] = ctx.get_force(v0.copy())
        its = ctx.stats()["minres_iters_max"]
        assert its < ctx.desc.adjoint_max_iter, (name, its)
        ctx.close()
    (fg, lg), (fo, lo) = out["g"], out["o"]
    ctx = abi.Context(glib, abi.default_desc(glib, **kw), pot_base, rho, w, rho_nodal)
    ctx.set_initial_guess(v0)
    hist, _, n = ctx.bfgs_solve(ls_type=abi.LS_CP, max_iter=2)
    ctx.close()
    print(f"LiH 6^3 p=4: int rho_t = {np.sum(rho * np.tile(_jxw(d), 216)):.6f} per spin, "
          f"charge {st.charge:.4f}, loss GPU {lg:.10e} oracle {lo:.10e}, BFGS {hist}", flush=True)
    assert abs(lg - lo) < 1e-8 * abs(lo)
    assert rel(fg, fo) < 1e-6
    assert n == 2 and np.all(np.isfinite(hist)) and abs(hist[0] - lg) < 1e-6 * abs(lg)


def _jxw(d):
    from numpy.polynomial.legendre import leggauss
    x, wq = leggauss(d.degree + 3)
    h = 2.0 * d.half_width / d.n_cells
    w1 = wq * h / 2
    return np.einsum("k,j,i->kji", w1, w1, w1).ravel()
