"""Separate solver instances used from separate host threads at the same time (the C ABI
keeps no process-wide mutable state besides the tuner's lock-protected cache; graphs
are captured in thread-local mode on each solver's own stream): every thread's results
equal the same solves run one after another, bit for bit."""
import threading

import numpy as np
import pytest

from vfs_amd import StereoTgv
from vfs_amd import synth as S

pytestmark = pytest.mark.gpu

P = S.SOLVER_PARAMS
SIZES = [(128, 96, 51), (192, 160, 52), (160, 120, 53)]


def _solver(pair):
    H, W = pair["i0"].shape
    g = StereoTgv()
    g.initialize(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], 3,
                 P["fScale"], 2, 10)
    g.limitRange = P["limitRange"]
    g.copyMaskToDevice(pair["mask"])
    g.loadVectorFields(pair["tv"], pair["cv"])
    g.copyImagesToDevice(pair["i0"], pair["i1"])
    return g


def _run(pair, rounds):
    """Masked and unmasked solves, alternating, on one fresh solver."""
    g = _solver(pair)
    out = []
    for r in range(rounds):
        if r % 2 == 0:
            g.solveStereoForwardMasked()
        else:
            g.solveStereoForward()
        out.append(g.copy1DDisparityToHost().copy())
    g.close()
    return out


def test_solvers_on_separate_threads_match_sequential():
    pairs = [S.make_pair(w, h, seed, radius=h / 2, dmax=w / 21.2) for w, h, seed in SIZES]
    rounds = 4
    ref = [_run(p, rounds) for p in pairs]
    got = [None] * len(pairs)
    errors = []
    start = threading.Barrier(len(pairs))

    def work(k):
        try:
            start.wait(timeout=60)
            got[k] = _run(pairs[k], rounds)
        except Exception as e:  # reported below
            errors.append((k, repr(e)))

    threads = [threading.Thread(target=work, args=(k,)) for k in range(len(pairs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in threads), "a solver thread did not finish"
    assert not errors, errors
    for k in range(len(pairs)):
        for r in range(rounds):
            assert np.array_equal(got[k][r], ref[k][r], equal_nan=True), (k, r)
