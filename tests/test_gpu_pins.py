"""HIP library on the cases that pin the oracle, against the oracle and the
same independent numpy restatements (gfx950 only).

  * child -> parent v_xc interpolation of trilinear fields (exact);
  * the reference-mode force against the numpy child-Gauss integral;
  * fractional occupancies in a degeneracy group (f' != 0, eta != 0, nonzero
    off-diagonal mu_ij and distributeX coefficients), contract by contract
    and through getForceVector, with the degeneracy-preserving block shrink;
  * the exact-gradient mode against central finite differences there;
  * the L-BFGS step against the explicit inverse-Hessian product rebuilt
    from the driver's checkpoint file;
  * multivector handles over caller-owned device memory.
"""
import struct

import numpy as np
import pytest

from invdft_amd import abi
from tests import pins
from tests.helpers import DEG_ATOMS, bare_context, rel, small_problem, symmetric_vxc
from tests.test_oracle_pins import _adjoint_p, trilinear

pytestmark = pytest.mark.gpu


def _say(*a):
    print("CASE", *a, flush=True)


def test_child_interpolation_reproduces_trilinear_gpu(glib):
    _say("trilinear_interp")
    for nc, p in ((3, 4), (2, 5)):
        c, d = bare_context(glib, nc, p, half_width=3.0)
        c.set_potential(trilinear(abi.vxc_points(glib, d)))
        want = trilinear(abi.quad_points(glib, d))
        assert np.max(np.abs(c.veff_quad() - want)) < 1e-13 * np.max(np.abs(want))


@pytest.mark.parametrize("p,atoms,scale", [(4, None, 1.0), (5, None, 1.0), (4, DEG_ATOMS, 0.1)])
def test_reference_force_equals_child_gauss_integral_gpu(glib, olib, p, atoms, scale):
    _say("ref_force_numpy", p, atoms is not None)
    over = {"degree": p}
    if atoms is not None:
        over["atoms"] = atoms
    prob = small_problem(olib, **over)
    kw = dict(cheb_tol=1e-12, initial_tol_cheb=1e-12)
    v = scale * prob.vxc0
    g_lib, _ = prob.context(glib, **kw).get_force(v)
    ctx = prob.context(glib, **kw)
    ctx.forward(v)
    psi = ctx.eigenvectors()
    occ, fermi = ctx.occupancies()
    ev = ctx.eigenvalues()
    u = prob.rho_target_q - 0.5 * ctx.density()
    tol = min(prob.desc.adjoint_initial_tol, 1e-12 / prob.desc.adaptive_factor_adjoint)
    pv = _adjoint_p(ctx, prob, psi, ev, occ, fermi, u, tol)
    g_np = pins.child_gauss_force(glib, prob.desc, psi, pv)
    assert rel(g_np, g_lib) < 1e-9


def _deg_inputs(olib, T=500.0):
    prob = small_problem(olib, atoms=DEG_ATOMS, temperature=T)
    o = prob.context(olib)
    v = 0.1 * prob.vxc0
    o.set_potential(v)
    ev, _, _ = o.eigensolve(1e-12, 100)
    occ, fermi = o.occupancies()
    return prob, v, o, ev, occ, fermi


def test_fractional_degenerate_adjoint_contract(glib, olib):
    """computeRhs (off-diagonal mu_ij inside the triplet), MINRES and
    distributeX (f'_i mu_ij and eta corrections) on identical inputs."""
    _say("fractional_contract")
    prob, v, o, ev, occ, fermi = _deg_inputs(olib)
    psi = o.eigenvectors()
    u = prob.rho_target_q - 0.5 * o.density()
    B = psi.shape[1]
    pairs = pins.degeneracy_pairs(ev)
    pr = np.array(pairs).reshape(-1, 2)
    off = [k for k, (i, j) in enumerate(pr) if i != j and occ[i] > 0.1]
    assert len(off) == 6  # the triplet's ordered off-diagonal pairs
    fp = pins.fd_occ_der(ev[1:4], fermi, prob.desc.temperature)
    assert np.min(np.abs(fp)) > 10.0 and len(set(np.round(fp, 6))) == 3
    g = prob.context(glib)
    g.set_potential(v)
    res = {}
    for name, c in (("g", g), ("o", o)):
        mpsi = c.mv(B, psi)
        c.adj_update_input_psi(mpsi, occ, u, pairs, fermi, ev.copy())
        rhs = c.mv(B)
        c.adj_compute_rhs(None, c.mv(B), rhs)
        sol = c.mv(B)
        c.minres_solve(sol, None, 1e-13, 5000, distribute=True)
        dots = c.adj_dot_quad_wise(sol, mpsi, len(pairs) // 2)
        res[name] = (rhs.get(), sol.get(), dots, mpsi)
    assert rel(res["g"][0], res["o"][0]) < 1e-11
    assert rel(res["g"][1], res["o"][1]) < 1e-8
    dg, do = res["g"][2], res["o"][2]
    assert np.max(np.abs(dg - do)) < 1e-8 * np.max(np.abs(do))
    # distributeX left nonzero <p_i, psi_j> = -f'_i mu1_ij inside the group
    assert np.min(np.abs(do[off])) > 1e-6 * np.max(np.abs(do))


@pytest.mark.parametrize("mode", [abi.GRADIENT_REFERENCE, abi.GRADIENT_EXACT])
@pytest.mark.parametrize("block", [0, 3])
def test_fractional_degenerate_force(glib, olib, mode, block):
    _say("fractional_force", mode, block)
    prob = small_problem(olib, atoms=DEG_ATOMS)
    kw = dict(cheb_tol=1e-12, initial_tol_cheb=1e-12, gradient_mode=mode, block_size=block)
    v = 0.1 * prob.vxc0
    g = prob.context(glib, **kw)
    o = prob.context(olib, **kw)
    fg, lg = g.get_force(v)
    fo, lo = o.get_force(v)
    assert abs(lg - lo) < 1e-9 * abs(lo)
    assert rel(fg, fo) < 1e-7
    assert g.stats()["adjoint_blocks"] == o.stats()["adjoint_blocks"] == (4 if block else 1)


def test_exactly_degenerate_fractional_gradient_matches_fd_gpu(glib, olib):
    _say("fractional_fd")
    T = 3000.0
    prob = small_problem(olib, atoms=DEG_ATOMS, temperature=T)
    ctx = prob.context(glib, gradient_mode=abi.GRADIENT_EXACT, cheb_tol=1e-13,
                       initial_tol_cheb=1e-13, adjoint_initial_tol=1e-14,
                       rho_tol_constraints=0.0)
    v = symmetric_vxc(glib, prob.desc)
    g, _ = ctx.get_force(v)
    occ, _ = ctx.occupancies()
    assert np.allclose(occ[1:4], 1.0 / 3.0, atol=1e-9)
    rng = np.random.default_rng(3)
    for _ in range(3):
        d = rng.standard_normal(v.size)
        d /= np.linalg.norm(d)
        h = 1e-4
        fd = (ctx.forward(v + h * d) - ctx.forward(v - h * d)) / (2 * h)
        assert abs(fd - g @ d) < 1e-6 * abs(fd), (fd, g @ d)


def _read_lbfgs_checkpoint(path):
    """The driver's checkpoint: 'IDFTLBF1', int64 n, int32 m, double L, x[n],
    g[n], then m records (s[n], y[n], rho), oldest first."""
    with open(path, "rb") as f:
        assert f.read(8) == b"IDFTLBF1"
        n, = struct.unpack("<q", f.read(8))
        m, = struct.unpack("<i", f.read(4))
        L, = struct.unpack("<d", f.read(8))
        x = np.frombuffer(f.read(8 * n), dtype="<f8").copy()
        g = np.frombuffer(f.read(8 * n), dtype="<f8").copy()
        hist = []
        for _ in range(m):
            s = np.frombuffer(f.read(8 * n), dtype="<f8").copy()
            y = np.frombuffer(f.read(8 * n), dtype="<f8").copy()
            rho, = struct.unpack("<d", f.read(8))
            hist.append((s, y, rho))
        assert f.read() == b""
    return L, x, g, hist


def test_lbfgs_step_is_explicit_inverse_hessian_product(glib, olib, tmp_path):
    """The L-BFGS direction equals -H_k g with H_k built explicitly:
    H_0 = gamma I, gamma = s^T y / y^T y of the newest pair, then
    H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T over the pairs, oldest
    first (BFGSInverseDFTSolver.cpp:66-118).  The step the driver takes is
    lambda times that direction, with lambda from one CP secant step on
    f(lambda) = g(x + lambda p).p from lambda = 0 and 1
    (BFGSInverseDFTSolver.cpp:215-303)."""
    _say("lbfgs_explicit")
    prob = small_problem(olib)
    kw = dict(cheb_tol=1e-11, initial_tol_cheb=1e-11)
    ck = {}
    for k in (2, 3):
        c = prob.context(glib, **kw)
        path = str(tmp_path / f"lbfgs_{k}.bin")
        c.set_bfgs_checkpoint(path, every=1, resume=False)
        _, _, n = c.bfgs_solve(max_iter=k)
        assert n == k
        ck[k] = _read_lbfgs_checkpoint(path)
    L2, x2, g2, h2 = ck[2]
    L3, x3, g3, h3 = ck[3]
    assert len(h2) == 2 and len(h3) == 3
    s_new = h3[-1][0]
    assert rel(x3 - x2, s_new) < 1e-12
    for (sa, ya, ra), (sb, yb, rb) in zip(h2, h3[:2]):
        assert rel(sb, sa) < 1e-12 and rel(yb, ya) < 1e-12 and abs(rb - ra) < 1e-12 * abs(ra)
        assert abs(ra * (sa @ ya) - 1.0) < 1e-12
    n = x2.size
    s_last, y_last, _ = h2[-1]
    H = (s_last @ y_last) / (y_last @ y_last) * np.eye(n)
    for s, y, r in h2:
        V = np.eye(n) - r * np.outer(y, s)
        H = V.T @ H @ V + r * np.outer(s, s)
    pdir = -H @ g2
    lam = (s_new @ pdir) / (pdir @ pdir)
    assert rel(s_new, lam * pdir) < 1e-10
    # one CP secant step: f(0) = g.p, f(1) = g(x + p).p
    f0 = g2 @ pdir
    gp, _ = prob.context(glib, **kw).get_force(x2 + pdir)
    f1 = gp @ pdir
    if np.sqrt(abs(f1)) < 1e-6:  # converged at lambda = 1 (ls_tol 1e-6)
        assert abs(lam - 1.0) < 1e-12
        return
    slope = f1 - f0
    if slope > 0:
        slope = -slope
    lam_cp = 1.0 - f1 / slope
    if lam_cp < 0:
        lam_cp = 1.0 + f1 / slope
    assert abs(lam - lam_cp) < 1e-4 * abs(lam_cp), (lam, lam_cp)


def test_mv_wrap_over_device_memory(glib, olib):
    """Contract-2 calls on handles over caller-owned device buffers equal the
    same calls on library-owned handles, and write into the caller's memory."""
    _say("mv_wrap")
    import ctypes as C
    import torch
    prob = small_problem(olib)
    o = prob.context(olib)
    o.set_potential(prob.vxc0)
    ev, _, _ = o.eigensolve(1e-11, 60)
    occ, fermi = o.occupancies()
    psi = o.eigenvectors()
    u = prob.rho_target_q - 0.5 * o.density()
    B = psi.shape[1]
    pairs = pins.degeneracy_pairs(ev)
    g = prob.context(glib)
    g.set_potential(prob.vxc0)
    n = g.sizes.n_nodes
    xin = np.random.default_rng(4).standard_normal((n, B))
    wrap = glib.idft_mv_wrap
    wrap.restype = C.c_int
    wrap.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]

    def wrapped(t):
        h = C.c_void_p()
        g._check(wrap(g.h, C.c_void_p(t.data_ptr()), B, C.byref(h)))
        m = abi.MultiVector.__new__(abi.MultiVector)
        m.ctx, m.ncols, m.h = g, B, h
        return m

    dev = {k: torch.zeros(n * B, dtype=torch.float64, device="cuda") for k in ("psi", "x", "ax",
                                                                             "rhs", "xs")}
    dev["psi"].copy_(torch.from_numpy(psi.reshape(-1)))
    dev["x"].copy_(torch.from_numpy(xin.reshape(-1)))
    torch.cuda.synchronize()
    W = {k: wrapped(t) for k, t in dev.items()}
    g.adj_update_input_psi(W["psi"], occ, u, pairs, fermi, ev.copy())
    g.adj_compute_rhs(None, W["xs"], W["rhs"])
    g.adj_vmult(W["ax"], W["x"])
    rhs_w = dev["rhs"].cpu().numpy().reshape(n, B)
    ax_w = dev["ax"].cpu().numpy().reshape(n, B)
    own = {"psi": g.mv(B, psi), "x": g.mv(B, xin), "ax": g.mv(B), "rhs": g.mv(B), "xs": g.mv(B)}
    g.adj_update_input_psi(own["psi"], occ, u, pairs, fermi, ev.copy())
    g.adj_compute_rhs(None, own["xs"], own["rhs"])
    g.adj_vmult(own["ax"], own["x"])
    assert np.array_equal(rhs_w, own["rhs"].get())
    assert np.array_equal(ax_w, own["ax"].get())
    for m in W.values():
        glib.idft_mv_destroy(m.h)
        m.h = None
    # the caller's buffers survive the handles
    assert torch.equal(dev["rhs"].cpu(), torch.from_numpy(rhs_w.reshape(-1)))
