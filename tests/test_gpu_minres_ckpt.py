"""MINRES block compaction and eigenvector checkpoint/recompute on the GPU
(gfx950 only).

  * compaction: unoccupied states (f = 0, so b~ = 0) never enter the MINRES
    block and converged columns leave it, so the streamed column-steps are
    well below steps x block width, while the force still equals the oracle's;
    the repacked width is kept even (one finished column rides along);
  * checkpoint (desc.psi_checkpoint = 1): the eigenvector block is kept as an
    FP32 copy between the passes and recomputed by one ChFSI pass for the
    backward pass; the force equals the resident path's to the eigensolver
    tolerance, the library's peak device memory is lower, and a second
    evaluation warm-starts from the checkpoint.
"""
import ctypes as C

import numpy as np
import pytest

from invdft_amd import abi, synthetic
from tests.helpers import rel, small_problem

pytestmark = pytest.mark.gpu


def _say(*a):
    print("CASE", *a, flush=True)


def _mem(glib):
    fn = glib.idft_device_memory
    fn.restype = C.c_int
    fn.argtypes = [C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int32]
    cur, peak = C.c_int64(), C.c_int64()

    def read(reset=False):
        assert fn(C.byref(cur), C.byref(peak), int(reset)) == 0
        return cur.value, peak.value
    return read


@pytest.mark.parametrize("prec", [abi.FP64, abi.MIXED])
def test_compaction_streams_open_columns_only(glib, olib, prec):
    _say("compaction", prec)
    prob = small_problem(olib)  # 8 states, 2 occupied
    kw = dict(cheb_tol=1e-11, initial_tol_cheb=1e-11)
    g = prob.context(glib, precision=prec, **kw)
    fg, lg = g.get_force(prob.vxc0)
    st = g.stats()
    fo, lo = prob.context(olib, **kw).get_force(prob.vxc0)
    B = prob.desc.n_eig
    print(f"  column-steps {st['minres_column_steps']} vs {st['minres_iters_max']} x {B}, "
          f"compactions {st['minres_compactions']}", flush=True)
    assert st["minres_compactions"] >= 1
    assert st["minres_column_steps"] <= 2 * st["minres_iters_max"] + 8 * B
    assert rel(fg, fo) < (1e-7 if prec == abi.FP64 else 1e-6)
    g.close()


def test_compaction_keeps_even_width(glib, olib):
    """One occupied state of eight: after the first check one column is open,
    and the repacked block keeps one finished column as an inert second (even
    width); the force still equals the oracle's, and the column-steps stay
    those of the open column plus that one."""
    _say("compaction even width")
    prob = small_problem(olib, n_electrons=2.0)
    kw = dict(cheb_tol=1e-11, initial_tol_cheb=1e-11)
    g = prob.context(glib, **kw)
    fg, lg = g.get_force(prob.vxc0)
    st = g.stats()
    fo, lo = prob.context(olib, **kw).get_force(prob.vxc0)
    B = prob.desc.n_eig
    print(f"  column-steps {st['minres_column_steps']} vs {st['minres_iters_max']} x {B}, "
          f"compactions {st['minres_compactions']}", flush=True)
    assert st["minres_compactions"] >= 1
    assert st["minres_column_steps"] <= 2 * st["minres_iters_max"] + 8 * B
    assert abs(lg - lo) < 1e-9 * abs(lo)
    assert rel(fg, fo) < 1e-7
    g.close()


def test_checkpoint_recompute_matches_resident(glib, olib):
    _say("psi_checkpoint")
    mem = _mem(glib)
    cfg = dict(synthetic.CONFIGS["gradcheck_12c_p4"])
    prob = synthetic.make_problem(glib, seed=3, target_tol=1e-10, **cfg)
    kw = dict(cheb_tol=1e-10, initial_tol_cheb=1e-10)
    v1 = prob.vxc0
    v2 = 1.05 * prob.vxc0
    out = {}
    for ck in (0, 1):
        mem(reset=True)
        base, _ = mem()
        c = prob.context(glib, psi_checkpoint=ck, **kw)
        f1, l1 = c.get_force(v1)
        f2, l2 = c.get_force(v2)
        _, peak = mem()
        st = c.stats()
        out[ck] = dict(f1=f1, f2=f2, l1=l1, l2=l2, peak=peak - base, rec=st["psi_recomputes"])
        c.close()
    a, b = out[0], out[1]
    print(f"  peak device memory resident {a['peak'] / 2**20:.1f} MiB, checkpointed "
          f"{b['peak'] / 2**20:.1f} MiB; force diff {rel(b['f1'], a['f1']):.2e} / "
          f"{rel(b['f2'], a['f2']):.2e}", flush=True)
    assert b["rec"] == 2 and a["rec"] == 0
    assert abs(b["l1"] - a["l1"]) < 1e-12 * abs(a["l1"])
    assert rel(b["f1"], a["f1"]) < 1e-6 and rel(b["f2"], a["f2"]) < 1e-6
    assert b["peak"] < a["peak"]
