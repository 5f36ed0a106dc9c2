"""Uneven batch of independent inverse problems on one GPU.

Each problem is drawn from a seeded stream: a mesh of U{8..20}^3 cells on
[-12, 12]^3, degree U{3, 4, 5}, U{1..16} occupied orbitals (plus 4 buffer
states in the eigensolve), and U{1..4} soft-Coulomb atoms placed in
[-2, 2]^3 with Z in U{1..8}; its target density is the FP64 ground state of
the atoms alone (set-up, not timed) and it runs `steps` L-BFGS iterations.

``run_concurrent`` drives the problems from `workers` host threads, each with
its own context and therefore its own HIP stream; the library releases the
GIL inside every call (ctypes), so the problems' kernels overlap on the GPU.
``run_sequential`` solves them one after another, for the comparison.
"""
import threading
import time

import numpy as np

from . import abi, synthetic


def draw_problems(count, seed=5):
    """Problem parameters of the uneven batch (seeded)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        nc = int(rng.integers(8, 21))
        p = int(rng.integers(3, 6))
        nocc = int(rng.integers(1, 17))
        na = int(rng.integers(1, 5))
        pos = rng.uniform(-2.0, 2.0, (na, 3))
        z = rng.integers(1, 9, na).astype(float)
        atoms = [(float(z[i]), tuple(float(v) for v in pos[i])) for i in range(na)]
        out.append(dict(n_cells=nc, degree=p, half_width=12.0, n_eig=nocc + 4,
                        n_electrons=2.0 * nocc, atoms=atoms, seed=1000 + k))
    return out


def build_problems(lib, params):
    """Target densities of every problem (FP64 eigensolves on the GPU)."""
    probs = []
    for prm in params:
        kw = dict(prm)
        seed = kw.pop("seed")
        probs.append(synthetic.make_problem(lib, seed=seed, target_tol=1e-9, **kw))
    return probs


def _solve(lib, prob, precision, steps, tol_kw):
    ctx = prob.context(lib, precision=precision, **tol_kw)
    t0 = time.perf_counter()
    loss, _, n = ctx.bfgs_solve(ls_type=abi.LS_CP, max_iter=steps)
    dt = time.perf_counter() - t0
    ctx.close()
    return dt, float(loss[0]), float(loss[-1]), n


def run_concurrent(lib, probs, precision, steps, workers=4, tol_kw=None, progress=None):
    """Per-problem seconds and the wall time of the whole batch; progress(k,
    result) is called as each problem finishes."""
    tol_kw = tol_kw or {}
    res = [None] * len(probs)
    nxt = [0]
    lock = threading.Lock()
    errors = []

    def worker():
        while True:
            with lock:
                k = nxt[0]
                nxt[0] += 1
            if k >= len(probs) or errors:
                return
            try:
                res[k] = _solve(lib, probs[k], precision, steps, tol_kw)
                if progress:
                    with lock:
                        progress(k, res[k])
            except Exception as e:  # surfaced after the join
                errors.append((k, e))
                return

    t0 = time.perf_counter()
    threads = [threading.Thread(target=worker) for _ in range(max(1, workers))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    if errors:
        k, e = errors[0]
        raise RuntimeError(f"problem {k} failed: {e}")
    return res, wall


def run_sequential(lib, probs, precision, steps, tol_kw=None, progress=None):
    return run_concurrent(lib, probs, precision, steps, workers=1, tol_kw=tol_kw,
                          progress=progress)


def summary(res, wall):
    t = np.array([r[0] for r in res])
    return {"problems": len(res), "wall_s": round(wall, 3),
            "problems_per_s": round(len(res) / wall, 4),
            "p50_s": round(float(np.percentile(t, 50)), 4),
            "p99_s": round(float(np.percentile(t, 99)), 4),
            "loss_decreased": int(sum(1 for r in res if r[2] < r[1]))}
