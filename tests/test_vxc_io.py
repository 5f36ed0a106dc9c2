"""v_xc checkpoint files and L-BFGS resume (include/idft_io.h).

CPU: the text records of InverseDFTSolverFunction::writeChildMeshDataToFile
and the nearest-node reader of InverseDFTEngine::readVxcDataFromFileWithSearch
(shuffled and perturbed records, duplicates, missing nodes), the binary
variant, through the HIP library's host-only entry points.
GPU: the periodic v_xc output of getForceVector and an L-BFGS run resumed
from its checkpoint.
"""
import ctypes
import os

import numpy as np
import pytest

from invdft_amd import abi

from helpers import small_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIP_LIB = os.path.join(ROOT, "invdft_amd", "_lib", "libidft.so")


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HIP_LIB):
        pytest.fail("HIP library not built; run __graft_entry__.build()")
    return abi.bind(ctypes.CDLL(HIP_LIB, mode=ctypes.RTLD_LOCAL))


def _desc(lib, n_cells=3, half_width=4.0):
    return abi.default_desc(lib, n_cells=n_cells, degree=4, half_width=half_width)


def test_text_records_layout_and_round_trip(host, tmp_path):
    d = _desc(host)
    xyz = abi.vxc_points(host, d)
    pot = np.random.default_rng(0).standard_normal(len(xyz))
    f = tmp_path / "vxc_0"
    abi.write_vxc(host, d, f, pot)
    rec = np.loadtxt(f)
    assert rec.shape == (len(xyz), 5)
    assert np.array_equal(rec[:, 0], np.arange(len(xyz)))
    assert np.max(np.abs(rec[:, 1:4] - xyz)) < 1e-15
    assert np.array_equal(rec[:, 4], pot)  # 17 significant digits round-trip
    assert np.array_equal(abi.read_vxc(host, d, f), pot)


def test_two_spin_channels(host, tmp_path):
    d = _desc(host)
    n = abi.geometry(host, d).n_vxc
    pot = np.random.default_rng(1).standard_normal((2, n))
    f = tmp_path / "vxc2"
    abi.write_vxc(host, d, f, pot)
    assert np.loadtxt(f).shape == (n, 6)
    assert np.array_equal(abi.read_vxc(host, d, f, n_spin=2), pot)
    fb = tmp_path / "vxc2.bin"
    abi.write_vxc(host, d, fb, pot, binary=True)
    assert np.array_equal(abi.read_vxc(host, d, fb, n_spin=2, binary=True), pot)


def _write_records(path, rec):
    with open(path, "w") as fh:
        for r in rec:
            fh.write(" ".join(repr(float(v)) for v in r) + "\n")


def test_nearest_node_search(host, tmp_path):
    """Records in any order, coordinates off by less than the tolerance,
    land on their nodes; a record farther than the tolerance is dropped."""
    d = _desc(host)
    xyz = abi.vxc_points(host, d)
    rng = np.random.default_rng(2)
    pot = rng.standard_normal(len(xyz))
    perm = rng.permutation(len(xyz))
    jitter = rng.uniform(-3e-4, 3e-4, (len(xyz), 3))
    rec = np.column_stack([perm, xyz[perm] + jitter[perm], pot[perm]])
    f = tmp_path / "shuffled"
    _write_records(f, rec)
    assert np.array_equal(abi.read_vxc(host, d, f), pot)
    # the same file read with a tolerance below the jitter loses nodes
    with pytest.raises(abi.IdftError):
        abi.read_vxc(host, d, f, dist_tol=1e-5)


def test_reader_errors(host, tmp_path):
    d = _desc(host)
    xyz = abi.vxc_points(host, d)
    pot = np.arange(len(xyz), dtype=float)
    rec = np.column_stack([np.arange(len(xyz)), xyz, pot])
    dup = rec.copy()
    dup[5, 1:4] = dup[4, 1:4]  # two records nearest to node 4 (and none on node 5)
    f = tmp_path / "dup"
    _write_records(f, dup)
    with pytest.raises(abi.IdftError):
        abi.read_vxc(host, d, f)
    short = tmp_path / "short"
    _write_records(short, rec[:-1])  # last node without a record
    with pytest.raises(abi.IdftError):
        abi.read_vxc(host, d, short)
    with pytest.raises(abi.IdftError):
        abi.read_vxc(host, d, tmp_path / "absent")


def test_binary_checks_the_mesh(host, tmp_path):
    d = _desc(host)
    pot = np.random.default_rng(3).standard_normal(abi.geometry(host, d).n_vxc)
    f = tmp_path / "vxc.bin"
    abi.write_vxc(host, d, f, pot, binary=True)
    assert open(f, "rb").read(8) == b"IDFTVXC1"
    assert np.array_equal(abi.read_vxc(host, d, f, binary=True), pot)
    with pytest.raises(abi.IdftError):
        abi.read_vxc(host, _desc(host, half_width=5.0), f, binary=True)
    with pytest.raises(abi.IdftError):
        abi.read_vxc(host, d, f, n_spin=2, binary=True)


@pytest.mark.gpu
def test_periodic_vxc_output(glib, olib, tmp_path):
    print("CASE vxc_output", flush=True)
    prob = small_problem(olib)
    ctx = prob.context(glib, cheb_tol=1e-10, initial_tol_cheb=1e-10)
    ctx.set_vxc_output(str(tmp_path), "vxc", 2)
    pots = [prob.vxc0 * (1.0 + 0.01 * k) for k in range(3)]
    for p in pots:
        ctx.get_force(p)
    assert sorted(os.listdir(tmp_path)) == ["vxc_0", "vxc_2"]
    for k in (0, 2):
        assert np.array_equal(abi.read_vxc(glib, prob.desc, tmp_path / f"vxc_{k}"), pots[k])


@pytest.mark.gpu
def test_bfgs_resume_from_checkpoint(glib, olib, tmp_path):
    print("CASE bfgs_resume", flush=True)
    prob = small_problem(olib)
    kw = dict(cheb_tol=1e-11, initial_tol_cheb=1e-11)
    la, ga, _ = prob.context(glib, **kw).bfgs_solve(max_iter=3)
    ck = tmp_path / "lbfgs.ckpt"
    b = prob.context(glib, **kw)
    b.set_bfgs_checkpoint(str(ck), every=1, resume=False)
    lb, _, nb = b.bfgs_solve(max_iter=2)
    assert nb == 2 and ck.exists() and open(ck, "rb").read(8) == b"IDFTLBF1"
    assert np.array_equal(lb, la[:3])
    c = prob.context(glib, **kw)
    c.set_bfgs_checkpoint(str(ck), every=1, resume=True)
    lc, gc, nc = c.bfgs_solve(max_iter=1)
    assert nc == 1
    assert lc[0] == la[2] and gc[0] == ga[2]  # the stored state
    assert abs(lc[1] - la[3]) <= 1e-7 * abs(la[3])  # one more step, history included
