"""Independent numpy restatements used to pin both libraries (test code only).

Each function recomputes one quantity from its mathematical definition and
from the public geometry/basis tables, without calling the code it checks.
"""
import numpy as np

from invdft_amd import abi


def lagrange_matrix(nodes, t):
    """E[q][a] = l_a(t_q), the Lagrange basis on `nodes` evaluated at t."""
    t = np.asarray(t, dtype=np.float64)
    E = np.ones((len(t), len(nodes)))
    for a, xa in enumerate(nodes):
        for b, xb in enumerate(nodes):
            if b != a:
                E[:, a] *= (t - xb) / (xa - xb)
    return E


def lagrange_deriv_matrix(nodes, t):
    """E'[q][a] = l_a'(t_q) (product rule over the factors)."""
    t = np.asarray(t, dtype=np.float64)
    n = len(nodes)
    D = np.zeros((len(t), n))
    for a in range(n):
        for k in range(n):
            if k == a:
                continue
            term = np.full(len(t), 1.0 / (nodes[a] - nodes[k]))
            for b in range(n):
                if b != a and b != k:
                    term *= (t - nodes[b]) / (nodes[a] - nodes[b])
            D[:, a] += term
    return D


def cell_blocks(u, nc, p):
    """Node-major field u[N][B] -> per-cell blocks [cz][cy][cx][az][ay][ax][B]
    (lexicographic nodes, node (i, j, k) at (k N1 + j) N1 + i)."""
    N1 = nc * p + 1
    B = u.shape[1]
    g = u.reshape(N1, N1, N1, B)
    out = np.empty((nc, nc, nc, p + 1, p + 1, p + 1, B))
    for cz in range(nc):
        for cy in range(nc):
            for cx in range(nc):
                out[cz, cy, cx] = g[cz * p:cz * p + p + 1, cy * p:cy * p + p + 1,
                                    cx * p:cx * p + p + 1]
    return out


# Child mesh: 2 linear cells per parent edge; Gauss-2 points of each child cell
# in parent reference coordinates, child cell k in {0, 1} spans [k-1, k].
CHILD_T = np.array([-0.5 - 0.5 / np.sqrt(3.0), -0.5 + 0.5 / np.sqrt(3.0),
                    0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)])


def child_hats():
    """H[q][c]: the 3 child hat functions (nodes at t = -1, 0, 1) at CHILD_T."""
    t = CHILD_T
    H = np.zeros((4, 3))
    H[:, 0] = np.where(t < 0, -t, 0.0)
    H[:, 1] = 1.0 - np.abs(t)
    H[:, 2] = np.where(t > 0, t, 0.0)
    return H


def child_gauss_force(lib, desc, psi, p_adj):
    """g_a = sum over child cells and their Gauss-2 points of
    N_a^child(x) sum_j psi_j(x) p_j(x) JxW, JxW = (h/4)^3 (the child cell has
    edge h/2, Gauss-2 weights are 1).  Reference semantics:
    InverseDFTSolverFunction.cpp:1161-1217 (psi p at child quadrature) and
    1609-1674 (integration against the child shape functions)."""
    nc, p, L = desc.n_cells, desc.degree, desc.half_width
    h = 2.0 * L / nc
    gx, _, _, _ = abi.basis_tables(lib, p)
    E = lagrange_matrix(gx, CHILD_T)  # [4][p+1]
    cps = cell_blocks(psi, nc, p)
    cpp = cell_blocks(p_adj, nc, p)
    vp = np.einsum("za,yb,xc,ZYXabcj->ZYXzyxj", E, E, E, cps, optimize=True)
    vq = np.einsum("za,yb,xc,ZYXabcj->ZYXzyxj", E, E, E, cpp, optimize=True)
    S = np.sum(vp * vq, axis=-1) * (0.25 * h) ** 3  # [cz][cy][cx][qz][qy][qx]
    Hc = child_hats()
    part = np.einsum("zk,yj,xi,ZYXzyx->ZYXkji", Hc, Hc, Hc, S, optimize=True)
    NV1 = 2 * nc + 1
    g = np.zeros((NV1, NV1, NV1))
    for cz in range(nc):
        for cy in range(nc):
            for cx in range(nc):
                g[2 * cz:2 * cz + 3, 2 * cy:2 * cy + 3, 2 * cx:2 * cx + 3] += part[cz, cy, cx]
    return g.reshape(-1)


def lumped_mass(lib, desc):
    """Tensor-product GLL lumped mass: M_n = m(i) m(j) m(k), m = sum over the
    cells sharing a 1D node of (h/2) w_a."""
    p, nc, L = desc.degree, desc.n_cells, desc.half_width
    _, gw, _, _ = abi.basis_tables(lib, p)
    h = 2 * L / nc
    m1 = np.zeros(nc * p + 1)
    for c in range(nc):
        m1[c * p: c * p + p + 1] += 0.5 * h * gw
    return np.einsum("k,j,i->kji", m1, m1, m1).reshape(-1)


def stiffness_diagonal(lib, desc):
    """Diagonal of the assembled stiffness K_ij = int grad N_i . grad N_j:
    per cell (h/2) (d_a m_b m_c + m_a d_b m_c + m_a m_b d_c) with the exact 1D
    integrals d_a = int l_a'^2, m_a = int l_a^2 on [-1, 1] (numpy Gauss-Legendre
    with p + 1 points is exact for these degree-2p integrands)."""
    p, nc, L = desc.degree, desc.n_cells, desc.half_width
    gx, _, _, _ = abi.basis_tables(lib, p)
    t, w = np.polynomial.legendre.leggauss(p + 1)
    E = lagrange_matrix(gx, t)
    Dm = lagrange_deriv_matrix(gx, t)
    m = w @ (E ** 2)
    d = w @ (Dm ** 2)
    h = 2 * L / nc
    loc = 0.5 * h * (np.einsum("c,b,a->cba", m, m, d) + np.einsum("c,b,a->cba", m, d, m)
                     + np.einsum("c,b,a->cba", d, m, m))
    N1 = nc * p + 1
    K = np.zeros((N1, N1, N1))
    for cz in range(nc):
        for cy in range(nc):
            for cx in range(nc):
                K[cz * p:cz * p + p + 1, cy * p:cy * p + p + 1, cx * p:cx * p + p + 1] += loc
    return K.reshape(-1)


def fd_occ(e, mu, T):
    """Fermi-Dirac occupancy 1 / (1 + exp((e - mu) / kT)), kT = T * 3.166811563e-6 Ha."""
    kt = 3.166811563e-6 * T
    return 1.0 / (1.0 + np.exp((np.asarray(e) - mu) / kt))


def fd_occ_der(e, mu, T):
    """d f / d e = -f (1 - f) / kT."""
    kt = 3.166811563e-6 * T
    f = fd_occ(e, mu, T)
    return -f * (1.0 - f) / kt


def degeneracy_pairs(ev, tol=0.002):
    """Flattened (i, j) for every j in the degeneracy group of i, with the
    groups formed as InverseDFTSolverFunction.cpp:37-77 forms them: i joins
    its earlier group (or opens one), then every j within tol of i joins i's
    group."""
    B = len(ev)
    groups, of = [], {}
    for i in range(B):
        g = of.get(i)
        if g is None:
            g = len(groups)
            groups.append({i})
            of[i] = g
        for j in range(B):
            if abs(ev[i] - ev[j]) < tol:
                groups[g].add(j)
                of[j] = g
    return [v for i in range(B) for j in sorted(groups[of[i]]) for v in (i, j)]
