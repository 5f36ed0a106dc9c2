"""Golden fixture tests/golden/small_problem.npz (written by make_golden.py)."""
import ctypes
import os

import numpy as np
import pytest

from invdft_amd import abi, synthetic

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "small_problem.npz")
TIGHT = dict(cheb_tol=1e-12, initial_tol_cheb=1e-12)


def _problem(lib):
    z = np.load(FIX)  # plain arrays, no pickle
    raw = z["desc"].tobytes()
    # the fixture holds the descriptor bytes it was written with; fields added
    # at the end of the struct since then keep their defaults
    d = abi.default_desc(lib)
    ctypes.memmove(ctypes.byref(d), raw, min(len(raw), ctypes.sizeof(d)))
    prob = synthetic.Problem(d, z["pot_base_q"], z["rho_target_q"], z["weight_q"],
                             z["rho_target_nodal"], z["vxc0"])
    return prob, z


def _check(lib, dens_tol, force_tol):
    prob, z = _problem(lib)
    ctx = prob.context(lib, **TIGHT)
    ctx.set_potential(prob.vxc0)
    ev, _, _ = ctx.eigensolve(1e-12, 200)
    occ, _ = ctx.occupancies()
    conv = occ > prob.desc.frac_occ_tol  # states ChFSI converges (buffer states are not)
    assert np.max(np.abs(ev[conv] - z["eigenvalues"][conv])) < 1e-10
    d = ctx.density()
    assert np.linalg.norm(d - z["density"]) / np.linalg.norm(z["density"]) < dens_tol
    for mode, name in ((abi.GRADIENT_REFERENCE, "reference"), (abi.GRADIENT_EXACT, "exact")):
        f, loss = prob.context(lib, gradient_mode=mode, **TIGHT).get_force(prob.vxc0)
        assert abs(loss - z[f"loss_{name}"][0]) < 1e-10 * abs(z[f"loss_{name}"][0])
        ref = z[f"force_{name}"]
        assert np.linalg.norm(f - ref) / np.linalg.norm(ref) < force_tol


def test_oracle_reproduces_golden(olib):
    _check(olib, 1e-10, 1e-8)


@pytest.mark.gpu
def test_hip_matches_golden(glib):
    _check(glib, 1e-9, 1e-7)
