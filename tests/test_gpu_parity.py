"""HIP library vs the FP64 CPU oracle on identical inputs (gfx950 only)."""
import numpy as np
import pytest

from invdft_amd import abi
from tests.helpers import bare_pair, rel, small_problem

pytestmark = pytest.mark.gpu


def _say(*a):
    print("CASE", *a, flush=True)


def test_backend(glib):
    assert glib.idft_backend() == b"hip-gfx950"


def test_fma_probe(glib):
    """idft_fma_probe: positive rates below the published peaks (FP64 78.6,
    unpacked FP32 78.6 TFLOP/s), more independent chains no slower than
    fewer by more than noise, and IDFT_EINVAL for unsupported chain counts."""
    import ctypes as C
    _say("fma_probe")
    fn = glib.idft_fma_probe
    fn.restype = C.c_int
    fn.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_double)]
    v = C.c_double()
    rates = {}
    for prec in (abi.FP64, abi.MIXED):
        for ch in (4, 16):
            assert fn(prec, ch, C.byref(v)) == 0
            rates[(prec, ch)] = v.value
            assert 1.0 < v.value < 80.0
        assert rates[(prec, 16)] > 0.8 * rates[(prec, 4)]
    print(f"  {rates}", flush=True)
    assert fn(abi.FP64, 5, C.byref(v)) == abi.IDFT_EINVAL


@pytest.mark.parametrize("p", [2, 3, 4, 5])
@pytest.mark.parametrize("ncols", [1, 12, 20, 80])
def test_cell_apply_fp64(glib, olib, p, ncols):
    _say("apply_fp64", p, ncols)
    g, o, s = bare_pair(glib, olib, 3, p, 4)
    x = np.random.default_rng(p * 100 + ncols).standard_normal((s.n_nodes, ncols))
    yg = g.apply_hamiltonian(x)
    yo = o.apply_hamiltonian(x)
    assert rel(yg, yo) < 1e-13


@pytest.mark.parametrize("p", [3, 4, 5])
@pytest.mark.parametrize("ncols", [1, 16, 40])
def test_cell_apply_fp32(glib, olib, p, ncols):
    _say("apply_fp32", p, ncols)
    g, o, s = bare_pair(glib, olib, 3, p, 4, precision=abi.MIXED)
    x = np.random.default_rng(p + ncols).standard_normal((s.n_nodes, ncols))
    yg = g.apply_hamiltonian(x, precision=abi.MIXED)
    yo = o.apply_hamiltonian(x)
    assert rel(yg, yo) < 2e-6


@pytest.mark.parametrize("prec,tol", [(abi.FP64, 1e-11), (abi.MIXED, 1e-4)])
def test_chebyshev_filter(glib, olib, prec, tol):
    _say("filter", prec)
    g, o, s = bare_pair(glib, olib, 3, 4, 6, precision=prec)
    x = np.random.default_rng(5).standard_normal((s.n_nodes, 6))
    a, b, a0 = 2.0, 60.0, -3.0
    xg = g.chebyshev_filter(x, 12, a, b, a0, precision=prec)
    xo = o.chebyshev_filter(x, 12, a, b, a0)
    assert rel(xg, xo) < tol


def _subspace_gap(A, B):
    qa, _ = np.linalg.qr(A)
    qb, _ = np.linalg.qr(B)
    return np.linalg.norm(qa @ (qa.T @ qb) - qb)


def test_eigensolve_density(glib, olib):
    _say("eigensolve")
    prob = small_problem(olib)
    g = prob.context(glib)
    o = prob.context(olib)
    for c in (g, o):
        c.set_potential(prob.vxc0)
    eg, rg, _ = g.eigensolve(1e-11, 60)
    eo, ro, _ = o.eigensolve(1e-11, 60)
    assert rg < 1e-10 and ro < 1e-10
    assert np.max(np.abs(eg - eo)) < 1e-10
    occg, _ = g.occupancies()
    occo, _ = o.occupancies()
    assert np.max(np.abs(occg - occo)) < 1e-9
    assert rel(g.density(), o.density()) < 1e-9
    nocc = int(np.sum(occo > 0.5))
    assert _subspace_gap(g.eigenvectors()[:, :nocc], o.eigenvectors()[:, :nocc]) < 1e-8


@pytest.mark.parametrize("prec", [abi.FP64, abi.MIXED])
@pytest.mark.parametrize("n_cells", [3, 6])
def test_eigensolve_wide_block(glib, olib, prec, n_cells):
    """32 states: the Gram and X Q products of orthonormalisation and
    Rayleigh-Ritz run as library DGEMMs from 24 columns on (the Gram's node
    sum in batched chunks of 8192 nodes plus a remainder: 2197 nodes at 3^3
    cells are one short chunk, 15625 at 6^3 one chunk and a remainder); the
    lowest 8 eigenvalues (the solver's convergence test covers the occupied
    states and a few above them, not the top of the block), occupancies and
    density against the oracle's."""
    _say("eigensolve_wide", prec, n_cells)
    prob = small_problem(olib, n_eig=32, n_cells=n_cells)
    g = prob.context(glib, precision=prec)
    o = prob.context(olib)
    for c in (g, o):
        c.set_potential(prob.vxc0)
    eg, rg, _ = g.eigensolve(1e-10, 80)
    eo, ro, _ = o.eigensolve(1e-10, 80)
    assert rg < 1e-9 and ro < 1e-9
    assert np.max(np.abs(eg[:8] - eo[:8])) < 1e-9
    occg, _ = g.occupancies()
    occo, _ = o.occupancies()
    assert np.max(np.abs(occg - occo)) < 1e-8
    assert rel(g.density(), o.density()) < 1e-8


def _adjoint_setup(ctx, prob, psi, occ, eig, u, pairs, fermi):
    B = psi.shape[1]
    mpsi = ctx.mv(B, psi)
    neg = ctx.adj_update_input_psi(mpsi, occ, u, pairs, fermi, eig)
    return mpsi, neg


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_adjoint_contract(glib, olib, p):
    """Every contract-2 entry point at degrees 2 to 5 (the node-scaled
    sum-factorised cell products and fused gathers of the MINRES steps at
    each)."""
    _say("adjoint_contract", p)
    prob = small_problem(olib, degree=p)
    o = prob.context(olib)
    o.set_potential(prob.vxc0)
    ev, _, _ = o.eigensolve(1e-11, 60)
    occ, fermi = o.occupancies()
    psi = o.eigenvectors()
    rho = o.density()
    u = prob.rho_target_q - 0.5 * rho
    B = psi.shape[1]
    pairs = []
    for i in range(B):
        for j in range(B):
            if abs(ev[i] - ev[j]) < 0.002:
                pairs += [i, j]
    g = prob.context(glib)
    g.set_potential(prob.vxc0)
    res = {}
    for name, c in (("g", g), ("o", o)):
        mpsi, neg = _adjoint_setup(c, prob, psi, occ, ev, u, pairs, fermi)
        assert np.allclose(neg, -ev)
        x = c.mv(B)
        rhs = c.mv(B)
        c.adj_compute_rhs(None, x, rhs)
        xin = c.mv(B, np.random.default_rng(1).standard_normal((psi.shape[0], B)))
        ax = c.mv(B)
        c.adj_vmult(ax, xin)
        pj = c.mv(B)
        c.adj_precondition_jacobi(pj, xin)
        pjs = c.mv(B)
        c.adj_precondition_jacobi_sqrt(pjs, xin)
        dq = c.adj_dot_quad_wise(mpsi, mpsi, len(pairs) // 2)
        sol = c.mv(B)
        its = c.minres_solve(sol, None, 1e-13, 3000, distribute=True)
        res[name] = dict(rhs=rhs.get(), ax=ax.get(), pj=pj.get(), pjs=pjs.get(), dq=dq,
                         sol=sol.get(), its=its, keep=(mpsi, x, rhs, xin, ax, pj, pjs, sol))
    G, O = res["g"], res["o"]
    assert rel(G["rhs"], O["rhs"]) < 1e-11
    assert rel(G["ax"], O["ax"]) < 1e-13
    assert rel(G["pj"], O["pj"]) < 1e-15
    assert rel(G["pjs"], O["pjs"]) < 1e-15
    assert np.max(np.abs(G["dq"] - O["dq"])) < 1e-12
    assert rel(G["sol"], O["sol"]) < 1e-8
    assert np.all(np.abs(G["its"] - O["its"]) <= 3)


@pytest.mark.parametrize("prec", [abi.FP64, abi.MIXED])
def test_minres_graph_replay_matches_launches(glib, olib, prec):
    """MINRES steps replayed from a captured HIP graph (check_every steps per
    replay, host convergence checks between replays) give bit-identical
    solutions and iteration counts to the same steps launched one by one,
    over two solves in a row on each context."""
    _say("minres_graph", prec)
    prob = small_problem(olib)
    o = prob.context(olib)
    o.set_potential(prob.vxc0)
    ev, _, _ = o.eigensolve(1e-11, 60)
    occ, fermi = o.occupancies()
    psi = o.eigenvectors()
    u = prob.rho_target_q - 0.5 * o.density()
    B = psi.shape[1]
    pairs = [v for i in range(B) for j in range(B) if abs(ev[i] - ev[j]) < 0.002 for v in (i, j)]
    out = {}
    for graph in (0, 1):
        g = prob.context(glib, precision=prec, minres_graph=graph)
        g.set_potential(prob.vxc0)
        mpsi, _ = _adjoint_setup(g, prob, psi, occ, ev.copy(), u, pairs, fermi)
        runs = []
        for _ in range(2):
            x = g.mv(B)
            g.adj_compute_rhs(None, x, g.mv(B))
            sol = g.mv(B)
            its = g.minres_solve(sol, None, 1e-12, 3000, distribute=False)
            runs.append((sol.get(), np.array(its)))
        out[graph] = runs
        del mpsi
    assert np.max(out[0][0][1]) > 16  # several replays of the 8-step graph
    for k in range(2):
        assert np.array_equal(out[0][k][1], out[1][k][1])
        assert np.array_equal(out[0][k][0], out[1][k][0])


@pytest.mark.parametrize("prec", [abi.FP64, abi.MIXED])
def test_minres_fused_step_matches_separate(glib, olib, prec):
    """Fused MINRES steps (alpha from the sum-factorised cell kernel's per-cell
    dots, Lanczos recurrence inside the node gather) against the separate
    Lanczos kernel (minres_fused = 0) and against the stored-H row pass,
    which always runs the separate steps: the same solutions to rounding and
    iteration counts within 2, and the oracle's solution."""
    _say("minres_fused", prec)
    prob = small_problem(olib)
    o = prob.context(olib)
    o.set_potential(prob.vxc0)
    ev, _, _ = o.eigensolve(1e-11, 60)
    occ, fermi = o.occupancies()
    psi = o.eigenvectors()
    u = prob.rho_target_q - 0.5 * o.density()
    B = psi.shape[1]
    pairs = [v for i in range(B) for j in range(B) if abs(ev[i] - ev[j]) < 0.002 for v in (i, j)]
    mpsi_o, _ = _adjoint_setup(o, prob, psi, occ, ev.copy(), u, pairs, fermi)
    xo = o.mv(B)
    o.adj_compute_rhs(None, xo, o.mv(B))
    so = o.mv(B)
    its_o = np.array(o.minres_solve(so, None, 1e-12, 3000, distribute=False))
    ref = so.get()
    out = {}
    for name, fused, row_pass in (("fused", 1, False), ("separate", 0, False),
                                  ("row_pass", 1, True)):
        g = prob.context(glib, precision=prec, minres_fused=fused)
        if row_pass:
            g.set_apply_shape(abi.FP64, 1, 16, 0, 0)
            if prec == abi.MIXED:
                g.set_apply_shape(abi.MIXED, 1, 16, 0, 0)
        g.set_potential(prob.vxc0)
        mpsi, _ = _adjoint_setup(g, prob, psi, occ, ev.copy(), u, pairs, fermi)
        x = g.mv(B)
        g.adj_compute_rhs(None, x, g.mv(B))
        sol = g.mv(B)
        its = np.array(g.minres_solve(sol, None, 1e-12, 3000, distribute=False))
        out[name] = (sol.get(), its)
        print(f"  {name}: iterations {its.tolist()}, rel to oracle {rel(out[name][0], ref):.2e}",
              flush=True)
        del mpsi
    print(f"  oracle iterations {its_o.tolist()}", flush=True)
    for name in ("separate", "row_pass"):
        assert rel(out["fused"][0], out[name][0]) < 1e-8
        assert np.all(np.abs(out["fused"][1] - out[name][1]) <= 2)
    assert rel(out["fused"][0], ref) < 1e-8
    assert np.all(np.abs(out["fused"][1] - its_o) <= 3)


@pytest.mark.parametrize("p", [2, 3, 4, 5])
@pytest.mark.parametrize("mode", [abi.GRADIENT_REFERENCE, abi.GRADIENT_EXACT])
def test_get_force(glib, olib, mode, p):
    _say("get_force", mode, p)
    prob = small_problem(olib, degree=p)
    kw = dict(cheb_tol=1e-11, initial_tol_cheb=1e-11, gradient_mode=mode)
    g = prob.context(glib, **kw)
    o = prob.context(olib, **kw)
    fg, lg = g.get_force(prob.vxc0)
    fo, lo = o.get_force(prob.vxc0)
    assert abs(lg - lo) / abs(lo) < 1e-9
    assert rel(fg, fo) < 1e-7


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_mixed_precision_force(glib, olib, p):
    _say("mixed_force", p)
    prob = small_problem(olib, degree=p)
    kw = dict(cheb_tol=1e-10, initial_tol_cheb=1e-10)
    gm = prob.context(glib, precision=abi.MIXED, **kw)
    o = prob.context(olib, **kw)
    fm, lm = gm.get_force(prob.vxc0)
    fo, lo = o.get_force(prob.vxc0)
    assert abs(lm - lo) / abs(lo) < 1e-7
    assert rel(fm, fo) < 1e-6


@pytest.mark.parametrize("p,prec", [(4, abi.FP64), (5, abi.MIXED)])
def test_bfgs_two_iterations(glib, olib, p, prec):
    """Two L-BFGS iterations against the oracle's FP64 driver, in FP64 at
    degree 4 and in mixed precision at degree 5."""
    _say("bfgs", p, prec)
    prob = small_problem(olib, degree=p)
    kw = dict(cheb_tol=1e-11, initial_tol_cheb=1e-11)
    g = prob.context(glib, precision=prec, **kw)
    o = prob.context(olib, **kw)
    lg, gg, ng = g.bfgs_solve(max_iter=2)
    lo, go, no = o.bfgs_solve(max_iter=2)
    assert ng == no == 2
    assert np.allclose(lg, lo, rtol=1e-7, atol=0)
    assert np.allclose(gg, go, rtol=1e-6, atol=0)
    assert lg[-1] < lg[0]


@pytest.mark.parametrize("p,ncols", [(4, 40), (5, 12), (3, 80)])
def test_every_tile_shape_keeps_parity(glib, olib, p, ncols):
    """Every cell-product tile shape on the autotuner's grid gives the same operator."""
    _say("tile_shapes", p, ncols)
    g, o, s = bare_pair(glib, olib, 3, p, 4, precision=abi.MIXED)
    x = np.random.default_rng(7).standard_normal((s.n_nodes, ncols))
    yo = o.apply_hamiltonian(x)
    for cp in (1, 2, 4, 8):
        for vt in (16, 32, 64, 80):
            for prec, tol in ((abi.FP64, 1e-13), (abi.MIXED, 2e-6)):
                # the tile pass applies to 16 (FP64) / 32 (FP32) columns
                for k in (abi.KERNEL_ROW_PASS, abi.KERNEL_TILE_PASS):
                    g.set_apply_shape(prec, cp, vt, kernel=k)
                    assert rel(g.apply_hamiltonian(x, precision=prec), yo) < tol, (cp, vt, k, prec)
    for gr in (1, 2, 3, 7, 4096):
        for prec, tol in ((abi.FP64, 1e-13), (abi.MIXED, 2e-6)):
            g.set_apply_shape(prec, 1, 16, gr)
            assert rel(g.apply_hamiltonian(x, precision=prec), yo) < tol, (gr, prec)
    best = g.autotune_apply(ncols, abi.FP64, reps=2)
    assert best["gather_rows"] in (0, 1, 2, 4, 8) and best["kernel"] in (0, 1, 2)
    if best["kernel"] == abi.KERNEL_SUM_FACTOR:
        assert best["vectors_per_tile"] in (2, 4, 8)
    else:
        assert best["cells_per_wg"] in (1, 2, 4, 8) and best["vectors_per_tile"] % 16 == 0
    assert rel(g.apply_hamiltonian(x), yo) < 1e-13


@pytest.mark.parametrize("p", [2, 3, 4, 5])
@pytest.mark.parametrize("ncols", [1, 3, 12, 20, 80])
def test_sum_factorised_cell_product(glib, olib, p, ncols):
    """The sum-factorised cell product (no stored H_e: 1D contractions with the
    per-cell quadrature weights) against the oracle's dense cell matrices, at
    every column count per work-group, FP64 and FP32; ragged last column groups
    included (3, 20)."""
    _say("sum_factor", p, ncols)
    g, o, s = bare_pair(glib, olib, 3, p, 4, precision=abi.MIXED)
    x = np.random.default_rng(11 * p + ncols).standard_normal((s.n_nodes, ncols))
    yo = o.apply_hamiltonian(x)
    for cg in (2, 4, 8, 16):
        for vc in (1, 2):  # columns per item (the cells_per_wg argument): 1 is built
            for prec, tol in ((abi.FP64, 1e-13), (abi.MIXED, 2e-6)):
                if vc != 1 or cg == 16:
                    with pytest.raises(RuntimeError):
                        g.set_apply_shape(prec, vc, cg, kernel=abi.KERNEL_SUM_FACTOR)
                    continue
                g.set_apply_shape(prec, vc, cg, kernel=abi.KERNEL_SUM_FACTOR)
                assert rel(g.apply_hamiltonian(x, precision=prec), yo) < tol, (cg, vc, prec)
