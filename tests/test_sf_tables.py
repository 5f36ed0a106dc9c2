"""The sum-factorised cell product's compile-time 1D tables
(invdft_amd/csrc/sf_tables.h: bisection on the Legendre recurrence) against
independent numpy constructions: Gauss points from `leggauss`, GLL nodes from
the roots of P_p', the Lagrange basis from `polyfromroots`, and the exact 1D
mass and stiffness integrated with more Gauss points than the header uses.
The tables define H_e's Kronecker form on the GPU, so a wrong node, sign or
transposed index here shows up as a wrong operator; the GPU parity tests check
the whole product against the oracle's dense cell matrices."""
import os
import shutil
import subprocess

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg
from numpy.polynomial import polynomial as nppoly

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "invdft_amd", "csrc")


@pytest.fixture(scope="module")
def tables(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("host compiler or ROCm headers missing")
    out = tmp_path_factory.mktemp("sft") / "sf_tables_dump"
    subprocess.run([gxx, "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{SRC}", os.path.join(HERE, "sf_tables_dump.cpp"), "-o", str(out)],
                   check=True, capture_output=True)
    txt = subprocess.run([str(out)], check=True, capture_output=True, text=True).stdout
    t = {}
    for line in txt.splitlines():
        f = line.split()
        t[(int(f[0]), f[1])] = np.array([float(v) for v in f[2:]])
    return t


def _gll(p):
    inner = npleg.Legendre.basis(p).deriv().roots()
    return np.concatenate([[-1.0], np.sort(np.real(inner)), [1.0]])


def _basis(p):
    x = _gll(p)
    polys = []
    for a in range(p + 1):
        c = nppoly.polyfromroots(np.delete(x, a))
        polys.append(c / nppoly.polyval(x[a], c))
    return polys


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_basis_at_gauss_points(tables, p):
    g, _ = npleg.leggauss(p + 3)
    ref = np.array([[nppoly.polyval(gq, c) for c in _basis(p)] for gq in g])
    L = tables[(p, "L")].reshape(p + 3, p + 1)
    assert np.max(np.abs(L - ref)) < 1e-13
    assert np.allclose(L.sum(axis=1), 1.0, atol=1e-14)  # partition of unity


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_mass_and_stiffness(tables, p):
    g, w = npleg.leggauss(p + 4)  # exact for the degree-2p integrands
    B = _basis(p)
    V = np.array([[nppoly.polyval(x, c) for c in B] for x in g])
    dV = np.array([[nppoly.polyval(x, nppoly.polyder(c)) for c in B] for x in g])
    M = (V * w[:, None]).T @ V
    D = (dV * w[:, None]).T @ dV
    n1 = p + 1
    assert np.max(np.abs(tables[(p, "M")].reshape(n1, n1) - M)) < 1e-13
    assert np.max(np.abs(tables[(p, "D")].reshape(n1, n1) - D)) < 1e-12 * np.max(np.abs(D))
    assert abs(M.sum() - 2.0) < 1e-13                      # integral of 1 over [-1, 1]
    assert np.max(np.abs(tables[(p, "D")].reshape(n1, n1).sum(axis=1))) < 1e-12  # D 1 = 0


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_tables_centro_symmetric(tables, p):
    """The kernel applies every 1D table in even-odd form: output R-1-i is
    taken as e_i - o_i, which equals row R-1-i of the product only when
    A[R-1-i][C-1-j] = A[i][j] (nodes and Gauss points symmetric about 0)."""
    n1, q = p + 1, p + 3
    for name, shape in (("L", (q, n1)), ("M", (n1, n1)), ("D", (n1, n1))):
        A = tables[(p, name)].reshape(shape)
        assert np.max(np.abs(A[::-1, ::-1] - A)) < 1e-14 * max(1.0, np.max(np.abs(A))), name
        # a wrong pairing (reversing one axis only) is far from symmetric
        assert np.max(np.abs(A[::-1, :] - A)) > 1e-3, name
