"""Writes tests/golden/small_problem.npz: inputs and FP64-oracle outputs of a
small inverse problem (3^3 cells, p=4, [-4,4]^3, wells Z=2 at 0 and Z=1 at
(0,0,1.5), 8 states, 4 electrons).  The reference ships no golden vectors for
this path, so these come from the oracle, whose arithmetic is pinned by the
closed-form / dense / finite-difference tests in tests/test_oracle_*.py.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from invdft_amd import abi, synthetic  # noqa: E402
from oracle import oracle  # noqa: E402

KW = dict(n_cells=3, degree=4, half_width=4.0, n_eig=8, n_electrons=4.0,
          atoms=[(2.0, (0.0, 0.0, 0.0)), (1.0, (0.0, 0.0, 1.5))])
TIGHT = dict(cheb_tol=1e-12, initial_tol_cheb=1e-12)


def main():
    olib = oracle.load()
    prob = synthetic.make_problem(olib, target_tol=1e-12, **KW)
    out = dict(pot_base_q=prob.pot_base_q, rho_target_q=prob.rho_target_q,
               weight_q=prob.weight_q, rho_target_nodal=prob.rho_target_nodal, vxc0=prob.vxc0,
               desc=np.frombuffer(bytes(prob.desc), dtype=np.uint8))
    prob.eigenvectors = None  # seeded random start, as the tests do
    ctx = prob.context(olib, **TIGHT)
    ctx.set_potential(prob.vxc0)
    ev, res, _ = ctx.eigensolve(1e-12, 200)
    out["eigenvalues"] = ev
    out["density"] = ctx.density()
    for mode, name in ((abi.GRADIENT_REFERENCE, "reference"), (abi.GRADIENT_EXACT, "exact")):
        c = prob.context(olib, gradient_mode=mode, **TIGHT)
        f, loss = c.get_force(prob.vxc0)
        out[f"force_{name}"] = f
        out[f"loss_{name}"] = np.array([loss])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "small_problem.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
