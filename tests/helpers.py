"""Shared problem builders for the tests (inputs only; no checked logic)."""
import numpy as np

from invdft_amd import abi, synthetic

SMALL = dict(n_cells=3, degree=4, half_width=4.0, n_eig=8, n_electrons=4.0,
             atoms=[(2.0, (0.0, 0.0, 0.0)), (1.0, (0.0, 0.0, 1.5))])

# One Z = 2 soft-Coulomb well at the centre of the cube: levels 1s, a p triplet
# (exactly degenerate by cubic symmetry), 2s, a d-like triplet.  With 4
# electrons the triplet holds 2 of its 6 (f = 1/3 each when it stays
# degenerate).
DEG_ATOMS = [(2.0, (0.0, 0.0, 0.0))]

_cache = {}


def small_problem(olib, **over):
    kw = dict(SMALL)
    kw.update(over)
    key = tuple(sorted((k, str(v)) for k, v in kw.items()))
    if key not in _cache:
        _cache[key] = synthetic.make_problem(olib, target_tol=1e-11, **kw)
    return _cache[key]


def bare_context(lib, n_cells, degree, half_width=4.0, n_eig=4, n_electrons=2.0, **kw):
    """A context with potBase = 0, rho_target = 0, w = 1 (operator tests)."""
    d = abi.default_desc(lib, n_cells=n_cells, degree=degree, half_width=half_width,
                         n_eig=n_eig, n_electrons=n_electrons, **kw)
    s = abi.geometry(lib, d)
    nq = s.n_cells * s.quad_per_cell
    c = abi.Context(lib, d, np.zeros(nq), np.zeros(nq), np.ones(nq), np.ones(s.n_nodes))
    return c, d


def symmetric_vxc(lib, desc):
    """A v_xc with the cube's full symmetry (radial about the centre)."""
    r2 = np.sum(abi.vxc_points(lib, desc) ** 2, axis=1)
    return 0.05 * np.exp(-r2 / 2.0) - 0.02 * np.exp(-r2 / 8.0)


def random_veff(lib, desc, seed):
    xyz = abi.quad_points(lib, desc)
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (3, 3))
    return (-2.0 / np.sqrt(np.sum(xyz ** 2, axis=1) + 0.7)
            + 0.1 * np.sin(xyz @ c[0]) + 0.05 * np.cos(xyz @ c[1]))


def bare_pair(glib, olib, n_cells, degree, n_eig, seed=0, precision=abi.FP64, half_width=4.0):
    """Contexts of both backends on the same random effective potential."""
    d = abi.default_desc(olib, n_cells=n_cells, degree=degree, half_width=half_width,
                         n_eig=n_eig, n_electrons=2.0, precision=precision)
    s = abi.geometry(olib, d)
    nq = s.n_cells * s.quad_per_cell
    v = random_veff(olib, d, seed)
    args = (np.zeros(nq), np.zeros(nq), np.ones(nq), np.ones(s.n_nodes))
    g = abi.Context(glib, d, *args)
    o = abi.Context(olib, d, *args)
    g.set_veff_quad(v)
    o.set_veff_quad(v)
    return g, o, s


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))
