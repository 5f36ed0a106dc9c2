"""Independent numpy restatement of the atomic-orbital (Gaussian) basis, used
to pin the libraries' target-density evaluation.

The reference builds its basis in gaussianFunctionManager.cpp: shells read
from the basis files (readGaussianFiles, :720-825), one function per m with
the p shell ordered m = 1, -1, 0 and every other shell m = -l..l
(storeBasisFunctions, :904-1005), primitive norms getNormConsts (:251-262), and
the value Clm Dm Plm(cos theta) Qm(phi) r^l sum_k c_k N_k exp(-a_k r^2)
(evaluateBasisValue, :416-446) with the Legendre function free of the
Condon-Shortley phase (Plm, :310-320).

Here the angular part comes from scipy's complex spherical harmonics (which
carry that phase, removed explicitly), so the libraries' recurrences are
checked against a separate construction.  Overlaps are integrated exactly:
each r^l Y_lm is a homogeneous polynomial, expanded on Cartesian monomials by
a least-squares fit (exact up to rounding), and a product of two Gaussian
primitives is one Gaussian (Gaussian product theorem), so every 1D factor is a
polynomial times a Gaussian that Gauss-Hermite quadrature integrates exactly.
"""
import math
import os

import numpy as np
from scipy import special

# DFT-FE's Angstrom -> Bohr factor (1 / 0.52917721067 Bohr per Angstrom); the
# demo's coordinates.inp (Bohr) agrees with AtomicCoords (Angstrom) to its 8
# printed digits.
ANG_TO_BOHR = 1.0 / 0.52917721067
L_CHARS = "SPDFGHI"


def read_atoms(coords_file, basis_dir, angstrom=True):
    """[(symbol, xyz (Bohr), basis file path)] from an AtomicCoords file."""
    atoms = []
    with open(coords_file) as f:
        for line in f:
            w = line.split()
            if len(w) < 5:
                continue
            xyz = np.array([float(v) for v in w[1:4]])
            if angstrom:
                xyz = xyz * ANG_TO_BOHR
            atoms.append((w[0], xyz, os.path.join(basis_dir, w[4])))
    return atoms


def read_shells(path):
    """[(l, alphas, coefs)] in file order (first line ignored)."""
    shells = []
    with open(path) as f:
        lines = f.read().splitlines()[1:]
    i = 0
    while i < len(lines):
        w = lines[i].split()
        i += 1
        if not w:
            continue
        lchars, n = w[0].upper(), int(w[1])
        rows = [list(map(float, lines[i + k].split())) for k in range(n)]
        i += n
        for j, ch in enumerate(lchars):
            shells.append((L_CHARS.index(ch), np.array([r[0] for r in rows]),
                           np.array([r[1 + j] for r in rows])))
    return shells


def prim_norm(alpha, l):
    """1 / sqrt(int (r^l e^{-a r^2})^2 r^2 dr), the radial primitive norm."""
    return 1.0 / np.sqrt(special.factorial2(2 * l + 1) * math.sqrt(math.pi) /
                         (2.0 ** (2 * l + 3.5) * alpha ** (l + 1.5)))


def build_basis(atoms):
    """List of dicts (center, l, m, alpha, cn = coef * norm) in the reference's order."""
    basis = []
    for _, xyz, path in atoms:
        for l, a, c in read_shells(path):
            ms = [1, -1, 0] if l == 1 else list(range(-l, l + 1))
            for m in ms:
                basis.append(dict(center=xyz, l=l, m=m, alpha=a, cn=c * prim_norm(a, l)))
    return basis


def real_sh(l, m, xyz):
    """Real spherical harmonic without the Condon-Shortley phase (cos for
    m > 0, sin for m < 0), from scipy's complex Y_l^m."""
    r = np.linalg.norm(xyz, axis=-1)
    theta = np.arccos(np.clip(xyz[..., 2] / np.where(r > 0, r, 1.0), -1.0, 1.0))
    phi = np.arctan2(xyz[..., 1], xyz[..., 0])
    am = abs(m)
    y = special.sph_harm_y(l, am, theta, phi)  # carries (-1)^m
    if m == 0:
        return y.real
    v = np.sqrt(2.0) * (-1.0) ** am * (y.real if m > 0 else y.imag)
    return v


def values(basis, pts):
    """Basis values [n_basis, n_points] at points [n_points, 3] (Bohr)."""
    out = np.empty((len(basis), len(pts)))
    for i, b in enumerate(basis):
        d = pts - b["center"]
        r2 = np.sum(d * d, axis=1)
        rad = np.exp(-np.outer(r2, b["alpha"])) @ b["cn"]
        out[i] = rad * np.sqrt(r2) ** b["l"] * real_sh(b["l"], b["m"], d)
    return out


def _monomials(l):
    return [(a, b, l - a - b) for a in range(l + 1) for b in range(l + 1 - a)]


def solid_harmonic_coeffs(l, m):
    """Coefficients of r^l Y_lm on the monomials x^a y^b z^c (a+b+c = l)."""
    mons = _monomials(l)
    rng = np.random.default_rng(1000 + 10 * l + m)
    p = rng.standard_normal((4 * len(mons) + 8, 3))
    r = np.linalg.norm(p, axis=1)
    f = r ** l * real_sh(l, m, p)
    A = np.stack([p[:, 0] ** a * p[:, 1] ** b * p[:, 2] ** c for a, b, c in mons], axis=1)
    coef, *_ = np.linalg.lstsq(A, f, rcond=None)
    assert np.max(np.abs(A @ coef - f)) < 1e-10 * max(1.0, np.max(np.abs(f)))
    return mons, coef


_GH = np.polynomial.hermite.hermgauss(8)  # exact for polynomial degree <= 15


def _gauss_1d(pa, pb, ea, eb, A, B):
    """int (x-A)^pa (x-B)^pb exp(-ea (x-A)^2 - eb (x-B)^2) dx, exactly."""
    p = ea + eb
    P = (ea * A + eb * B) / p
    pref = np.exp(-ea * eb / p * (A - B) ** 2)
    t, w = _GH
    x = P + t / np.sqrt(p)
    return pref / np.sqrt(p) * np.sum(w * (x - A) ** pa * (x - B) ** pb)


def overlap(basis):
    """Exact overlap matrix S_ij = int phi_i phi_j of the (unnormalised) basis."""
    cache = {}

    def sh(l, m):
        if (l, m) not in cache:
            cache[(l, m)] = solid_harmonic_coeffs(l, m)
        return cache[(l, m)]

    n = len(basis)
    S = np.zeros((n, n))
    for i in range(n):
        bi = basis[i]
        mi, ci = sh(bi["l"], bi["m"])
        for j in range(i, n):
            bj = basis[j]
            mj, cj = sh(bj["l"], bj["m"])
            A, B = bi["center"], bj["center"]
            s = 0.0
            for ai, cni in zip(bi["alpha"], bi["cn"]):
                for aj, cnj in zip(bj["alpha"], bj["cn"]):
                    acc = 0.0
                    for (ma, ca) in zip(mi, ci):
                        if ca == 0.0:
                            continue
                        for (mb, cb) in zip(mj, cj):
                            if cb == 0.0:
                                continue
                            v = 1.0
                            for k in range(3):
                                v *= _gauss_1d(ma[k], mb[k], ai, aj, A[k], B[k])
                            acc += ca * cb * v
                    s += cni * cnj * acc
            S[i, j] = S[j, i] = s
    return S


def density(basis, P, pts):
    """rho(x) = sum_ij P_ij phi_i(x) phi_j(x) with P taken from its upper
    triangle (getRhoValue, gaussianFunctionManager.cpp:1406-1430)."""
    Pu = np.triu(P) + np.triu(P, 1).T
    v = values(basis, pts)
    return np.einsum("ip,ij,jp->p", v, Pu, v)
