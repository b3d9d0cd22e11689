"""Device memory comes back: solvers, batches and servers created, used and closed over
and over leave the device's free memory where one such cycle left it (every arena,
pyramid, graph, event and pinned buffer is released on close)."""
import ctypes

import numpy as np
import pytest

from vfs_amd import Batch, Server, StereoTgv
from vfs_amd import synth as S

pytestmark = pytest.mark.gpu

P = S.SOLVER_PARAMS


def _free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _solver(pair):
    H, W = pair["i0"].shape
    g = StereoTgv()
    g.initialize(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], 3,
                 P["fScale"], 2, 8)
    g.limitRange = P["limitRange"]
    g.copyMaskToDevice(pair["mask"])
    g.loadVectorFields(pair["tv"], pair["cv"])
    g.copyImagesToDevice(pair["i0"], pair["i1"])
    return g


def _cycle(pairs):
    # single solvers, masked and unmasked, graph replay and a warm re-solve
    for p in pairs:
        g = _solver(p)
        g.solveStereoForwardMasked()
        g.solveStereoForward()
        g.solve_warm(1, 2, 4)
        assert np.isfinite(g.copy1DDisparityToHost()[p["mask"] > 0]).all()
        g.close()
    # a batch of all pairs
    members = [_solver(p) for p in pairs]
    b = Batch(members)
    b.solve(True)
    b.close()
    for g in members:
        g.close()
    # a server with one stream per pair, two frames each
    srv = Server(P, hbm_budget_bytes=4 << 30, max_streams=8, max_active=8)
    sids = [srv.add_stream(p["mask"], p["tv"], p["cv"], 3, 2, 8, 1, 2, 4) for p in pairs]
    for k in range(2):
        for s, p in zip(sids, pairs):
            srv.submit(s, p["i0"], p["i1"], tag=k * len(pairs) + s)
    assert len(srv.run_until_idle()) == 2 * len(pairs)
    srv.close()


def test_repeated_create_use_close_releases_device_memory():
    pairs = [S.make_pair(w, h, seed, radius=h / 2, dmax=w / 21.2) for w, h, seed in [(160, 120, 61), (224, 160, 62)]]
    _cycle(pairs)  # process-wide one-time state: tuner cache, runtime pools
    base = _free_bytes()
    for _ in range(4):
        _cycle(pairs)
    lost = base - _free_bytes()
    assert lost < (32 << 20), f"{lost / 2**20:.1f} MiB not released after 4 cycles"


def test_misuse_is_reported_and_leaves_the_solver_usable():
    """Calls out of order or with bad arguments fail with an error and a message (the
    reference exits the process instead); the same solver then solves normally and
    matches a fresh one bit for bit."""
    from vfs_amd import VfsError
    pair = S.make_pair(160, 120, 63, radius=60, dmax=160 / 21.2)
    g = StereoTgv()
    with pytest.raises(VfsError, match="before initialize"):
        g.copyMaskToDevice(pair["mask"])
    H, W = pair["i0"].shape
    g.initialize(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], 3,
                 P["fScale"], 2, 8)
    g.limitRange = P["limitRange"]
    with pytest.raises(VfsError, match="loadVectorFields"):
        g.solveStereoForwardMasked()
    g.loadVectorFields(pair["tv"], pair["cv"])
    with pytest.raises(VfsError, match="copyMaskToDevice"):
        g.solveStereoForwardMasked()
    g.copyMaskToDevice(pair["mask"])
    with pytest.raises(VfsError, match="copyImagesToDevice"):
        g.solveStereoForwardMasked()
    g.copyImagesToDevice(pair["i0"], pair["i1"])
    with pytest.raises(VfsError, match="warm"):
        g.solve_warm(1, 2, 4)
    with pytest.raises(VfsError):
        g.read_plane("no such plane", 0)
    with pytest.raises(VfsError):
        g.read_plane("u", 3)
    with pytest.raises(VfsError):
        g.set_option("no such option", 1)
    with pytest.raises(VfsError):
        g.set_option("per_launch", 3)
    g.solveStereoForwardMasked()
    with pytest.raises(VfsError):
        g.solve_warm(3, 2, 4)  # start level >= nLevels
    u = g.copy1DDisparityToHost()
    fresh = _solver(pair)
    fresh.solveStereoForwardMasked()
    assert np.array_equal(u, fresh.copy1DDisparityToHost())
