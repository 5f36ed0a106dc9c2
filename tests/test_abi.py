"""The C ABI: both libraries export the header's symbols; host-only entry points agree."""
import ctypes
import os
import re

import numpy as np
import pytest

from invdft_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIP_LIB = os.path.join(ROOT, "invdft_amd", "_lib", "libidft.so")


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(idft_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def hip_host():
    """The HIP library opened for its host-only entry points (no GPU calls)."""
    if not os.path.exists(HIP_LIB):
        pytest.fail("HIP library not built; run __graft_entry__.build()")
    return abi.bind(ctypes.CDLL(HIP_LIB, mode=ctypes.RTLD_LOCAL))


def test_hip_library_exports_every_declared_symbol(hip_host):
    for name in (_declared("idft.h") + _declared("idft_hip.h") + _declared("idft_io.h") +
                 _declared("idft_setup.h")):
        assert hasattr(hip_host, name), name


def test_oracle_exports_the_common_headers(olib):
    for name in _declared("idft.h") + _declared("idft_setup.h"):
        assert hasattr(olib, name), name


def test_backend_names(hip_host, olib):
    assert hip_host.idft_backend() == b"hip-gfx950"
    assert olib.idft_backend() == b"cpu-oracle"


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_geometry_agrees(hip_host, olib, p):
    d = abi.default_desc(olib, n_cells=3, degree=p, half_width=5.0)
    for fn in (abi.quad_points, abi.node_points, abi.vxc_points):
        a, b = fn(hip_host, d), fn(olib, d)
        assert a.shape == b.shape
        assert np.max(np.abs(a - b)) < 1e-13
    for tab_h, tab_o in zip(abi.basis_tables(hip_host, p), abi.basis_tables(olib, p)):
        assert np.max(np.abs(tab_h - tab_o)) < 1e-14


def test_defaults_follow_reference_parameters(olib):
    d = abi.default_desc(olib)
    # inverseDFTParameters.cpp:405-424 defaults
    assert d.adaptive_factor_adjoint == 1000.0 and d.adaptive_factor_cheb == 100.0
    assert d.initial_tol_cheb == 1e-6 and d.max_cheb_passes == 100
    assert d.rho_tol_constraints == 1e-6 and d.adjoint_initial_tol == 1e-11
    assert d.adjoint_max_iter == 5000 and d.degeneracy_tol == 0.002 and d.frac_occ_tol == 1e-8


def test_invalid_descriptor_rejected(hip_host, olib):
    for lib in (hip_host, olib):
        d = abi.default_desc(lib, n_cells=0)
        s = abi.Sizes()
        assert lib.idft_geometry_sizes(ctypes.byref(d), ctypes.byref(s)) == abi.IDFT_EINVAL


def test_oracle_call_order_errors(olib):
    from tests.helpers import small_problem
    prob = small_problem(olib)
    ctx = prob.context(olib)
    B = prob.desc.n_eig
    a, x = ctx.mv(B), ctx.mv(B)
    with pytest.raises(abi.IdftError) as e:
        ctx.adj_vmult(a, x)
    assert e.value.code in (abi.IDFT_ESTATE, abi.IDFT_EINVAL)
    with pytest.raises(abi.IdftError) as e:
        ctx.backward()
    assert e.value.code == abi.IDFT_ESTATE
