"""The C++ shim (include/idft_shim.hpp): a program written against the
reference's two interfaces, compiled with g++ and linked with the oracle
(CPU) and with the HIP library (GPU), gives the same contract results."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "shim_driver.cpp")


def _build_and_run(tmp, libdir, libname):
    exe = os.path.join(tmp, f"shim_{libname}")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), SRC,
                    "-L", libdir, f"-l{libname}", f"-Wl,-rpath,{libdir}", "-o", exe],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=600).stdout
    return {k: float(v) for k, v in (line.split() for line in out.strip().splitlines())}


def _oracle(tmp):
    from oracle import oracle
    oracle.load()
    return _build_and_run(tmp, os.path.join(ROOT, "oracle", "_build"), "oracle")


def test_shim_drives_the_oracle(tmp_path):
    r = _oracle(str(tmp_path))
    assert all(np.isfinite(v) for v in r.values())
    assert r["loss"] > 0 and r["force_sumsq"] > 0
    assert r["dot"] == pytest.approx(r["force_sumsq"], rel=1e-12)
    # updateInputPsi negates the caller's eigenvalues in place (MVALSP:898-903)
    assert r["negated_eigenvalue0"] == -r["eigenvalue0"]


@pytest.mark.gpu
def test_shim_drives_the_hip_library_like_the_oracle(tmp_path):
    print("CASE shim driver: oracle then HIP library", flush=True)
    o = _oracle(str(tmp_path))
    g = _build_and_run(str(tmp_path), os.path.join(ROOT, "invdft_amd", "_lib"), "idft")
    tol = {"loss": 1e-9, "force_sumsq": 1e-7, "force_wsum": 1e-6, "dot": 1e-7,
           "negated_eigenvalue0": 1e-10, "eigenvalue0": 1e-10, "rhs_sumsq": 1e-10,
           "vmult_wsum": 1e-12, "jacobi_wsum": 1e-14, "jacobi_sqrt_wsum": 1e-14,
           "solution_sumsq": 1e-7, "solution_wsum": 1e-6}
    for k, t in tol.items():
        assert g[k] == pytest.approx(o[k], rel=t), (k, g[k], o[k])
