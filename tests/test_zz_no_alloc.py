"""The allocation-free vmult (idft_adj_vmult_noalloc, include/idft_hip.h)
against idft_adj_vmult, through the C ABI on the GPU (gfx950 only).

Two calls in a row on one context, the second with a narrower block, each
compared bit for bit with idft_adj_vmult on the same inputs (both run the same
kernels; only the cell-product scratch differs).  The library's device
allocations are counted around each allocation-free call and must not change.
"""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import bare_context, small_problem

pytestmark = pytest.mark.gpu

_DP = C.POINTER(C.c_double)


def _bind(glib):
    fn = glib.idft_adj_vmult_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
    mem = glib.idft_device_memory
    mem.restype = C.c_int
    mem.argtypes = [C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int32]
    return fn, mem


def _device_bytes(mem):
    cur, peak = C.c_int64(), C.c_int64()
    assert mem(C.byref(cur), C.byref(peak), 0) == 0
    return cur.value


def test_vmult_noalloc_matches_vmult(glib, olib):
    print("CASE vmult_noalloc: 3^3 cells p=4, blocks 8 then 3", flush=True)
    noalloc, mem = _bind(glib)
    assert glib.idft_hip_init() == 0
    prob = small_problem(olib)
    g = prob.context(glib, cheb_tol=1e-11, initial_tol_cheb=1e-11)
    g.forward(prob.vxc0)
    psi_all = g.eigenvectors()
    occ_all, fermi = g.occupancies()
    ev_all = g.eigenvalues()
    u = prob.rho_target_q - 0.5 * g.density()
    rng = np.random.default_rng(7)
    for b in (psi_all.shape[1], 3):
        ev = ev_all[:b].copy()
        pairs = [k for i in range(b) for j in range(b) if abs(ev[i] - ev[j]) < 0.002 for k in (i, j)]
        mpsi = g.mv(b, psi_all[:, :b])
        g.adj_update_input_psi(mpsi, occ_all[:b], u, pairs, fermi, ev)
        x = rng.standard_normal((psi_all.shape[0], b))
        x1, x2 = g.mv(b, x), g.mv(b, x)
        ax1, ax2 = g.mv(b), g.mv(b)
        g.adj_vmult(ax1, x1)
        before = _device_bytes(mem)
        rc = noalloc(g.h, ax2.h, x2.h, b)
        assert rc == 0, glib.idft_last_error(g.h)
        assert _device_bytes(mem) == before
        a1, a2 = ax1.get(), ax2.get()
        assert np.linalg.norm(a1) > 0
        assert np.array_equal(a1, a2), (b, float(np.max(np.abs(a1 - a2))))
        assert np.array_equal(x1.get(), x2.get())  # constrained rows zeroed alike
    g.close()


def test_vmult_noalloc_refuses_oversized_block(glib):
    """16^3 cells at p = 5 with 10 columns need 8,847,360 scratch doubles,
    more than the fixed 8,388,608: a clear IDFT_EINVAL, nothing launched."""
    print("CASE vmult_noalloc oversized: 16^3 cells p=5, block 10", flush=True)
    noalloc, _ = _bind(glib)
    c, d = bare_context(glib, 16, 5, half_width=8.0, n_eig=10, n_electrons=2.0)
    c.set_veff_quad(np.zeros(c.nq_total))
    b = 10
    n = c.sizes.n_nodes
    mpsi = c.mv(b, np.random.default_rng(3).standard_normal((n, b)))
    c.adj_update_input_psi(mpsi, np.zeros(b), np.zeros(c.nq_total), [k for i in range(b) for k in (i, i)],
                           0.0, np.zeros(b))
    x, ax = c.mv(b), c.mv(b)
    rc = noalloc(c.h, ax.h, x.h, b)
    assert rc == -1
    assert b"exceeds the fixed scratch" in glib.idft_last_error(c.h)
    c.close()
