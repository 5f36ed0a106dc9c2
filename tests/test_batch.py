"""Uneven-batch driver: problem draw and concurrent contexts (harness only;
the numbers of each problem are the single-problem path's)."""
import numpy as np

from invdft_amd import abi, batch, synthetic

from helpers import SMALL


def test_draw_is_seeded_and_in_range():
    a, b = batch.draw_problems(96), batch.draw_problems(96)
    assert a == b
    for p in a:
        assert 8 <= p["n_cells"] <= 20 and 3 <= p["degree"] <= 5
        nocc = p["n_electrons"] / 2
        assert 1 <= nocc <= 16 and p["n_eig"] == nocc + 4
        assert 1 <= len(p["atoms"]) <= 4
        for z, pos in p["atoms"]:
            assert 1 <= z <= 8 and max(abs(c) for c in pos) <= 2.0
    assert len({(p["n_cells"], p["degree"]) for p in a}) > 10  # uneven


def test_concurrent_contexts_match_sequential(olib):
    probs = []
    for k, nocc in enumerate((1, 2, 3)):
        kw = dict(SMALL, n_eig=nocc + 4, n_electrons=2.0 * nocc)
        probs.append(synthetic.make_problem(olib, seed=20 + k, target_tol=1e-10, **kw))
    tol = dict(cheb_tol=1e-10, initial_tol_cheb=1e-10)
    conc, wall = batch.run_concurrent(olib, probs, abi.FP64, 1, workers=3, tol_kw=tol)
    seq, _ = batch.run_sequential(olib, probs, abi.FP64, 1, tol_kw=tol)
    for c, s in zip(conc, seq):
        assert c[3] == s[3] == 1
        assert abs(c[2] - s[2]) <= 1e-10 * abs(s[2])
    sm = batch.summary(conc, wall)
    assert sm["problems"] == 3 and sm["p50_s"] <= sm["p99_s"]
    assert np.isfinite(sm["problems_per_s"])
