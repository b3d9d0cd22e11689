"""Warm-started solves and the continuous-batching server.

Warm start (the serving schedule): the previous result is downscaled to a start level
and refined there and below with few warps; the oracle implements the same schedule
(oracle_solve_warm), so GPU warm solves are parity-checked like cold ones.  The server's
frames must equal a standalone solver running the same cold/warm sequence, and a
stream's results must be identical whether it is served alone or mixed with others."""
import numpy as np
import pytest

from oracle import oracle as O
from vfs_amd import synth as S

P = S.SOLVER_PARAMS


def frames(W, H, seed, n):
    base = S.make_pair(W, H, seed, radius=H / 2, dmax=W / 21.2)
    return base, [base] + [S.drift_pair(base, k, seed) for k in range(1, n)]


def test_oracle_warm_start_tracks_drift():
    base, fr = frames(160, 120, 5, 4)
    o = O.Oracle(160, 120, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], 3,
                 P["fScale"], 5, 20)
    o.copyMaskToDevice(base["mask"])
    o.loadVectorFields(base["tv"], base["cv"])
    m = base["mask"] > 0
    errs = []
    for k, f in enumerate(fr):
        o.copyImagesToDevice(f["i0"], f["i1"])
        if k == 0:
            o.solveStereoForwardMasked()
        else:
            o.solve_warm(1, 2, 20)
        errs.append(np.median(np.abs(o.copy1DDisparityToHost() - f["gt"])[m]))
    assert max(errs) < 0.3, errs
    # drift of at most 0.5 px per frame: the ground truth really moves
    assert np.abs(fr[-1]["gt"] - fr[0]["gt"])[m].max() > 0.4


def test_drift_is_bounded():
    base, fr = frames(96, 80, 9, 4)
    for a, b in zip(fr, fr[1:]):
        assert np.abs(a["gt"] - b["gt"]).max() <= 0.5 + 1e-5


# ------------------------------------------------------------------ GPU


def gpu_solver(base, levels, warps, iters):
    from vfs_amd import StereoTgv
    H, W = base["i0"].shape
    g = StereoTgv()
    g.initialize(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], levels,
                 P["fScale"], warps, iters)
    g.limitRange = P["limitRange"]
    g.copyMaskToDevice(base["mask"])
    g.loadVectorFields(base["tv"], base["cv"])
    return g


def standalone_sequence(base, fr, levels, warps, iters, warm_flags, warm=(1, 2, 5)):
    g = gpu_solver(base, levels, warps, iters)
    out = []
    for f, w in zip(fr, warm_flags):
        g.copyImagesToDevice(f["i0"], f["i1"])
        if w:
            g.solve_warm(*warm)
        else:
            g.solveStereoForwardMasked()
        out.append(g.copy1DDisparityToHost())
    return out


@pytest.mark.gpu
def test_warm_solve_matches_oracle():
    base, fr = frames(160, 120, 11, 3)
    g = gpu_solver(base, 3, 3, 10)
    o = O.Oracle(160, 120, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], 3,
                 P["fScale"], 3, 10)
    o.copyMaskToDevice(base["mask"])
    o.loadVectorFields(base["tv"], base["cv"])
    m = base["mask"] > 0
    for k, f in enumerate(fr):
        g.copyImagesToDevice(f["i0"], f["i1"])
        o.copyImagesToDevice(f["i0"], f["i1"])
        if k == 0:
            g.solveStereoForwardMasked()
            o.solveStereoForwardMasked()
        else:
            g.solve_warm(1, 2, 5)
            o.solve_warm(1, 2, 5)
        err = np.abs(g.copy1DDisparityToHost() - o.copy1DDisparityToHost())[m]
        assert err.mean() <= 5e-4, (k, err.mean())


@pytest.mark.gpu
def test_warm_solves_do_not_depend_on_iterations_per_launch():
    """Cold and warm solves (start levels 1 and 2) give the same disparity whether the
    inner iterations run one, four or eight per launch."""
    base, fr = frames(200, 120, 13, 4)
    out = {}
    for per_launch in (1, 4, 8):
        g = gpu_solver(base, 3, 3, 10)
        g.set_option("per_launch", per_launch)
        res = []
        for k, f in enumerate(fr):
            g.copyImagesToDevice(f["i0"], f["i1"])
            if k == 0:
                g.solveStereoForwardMasked()
            else:
                g.solve_warm(1 + k % 2, 2, 6)
            res.append(g.copy1DDisparityToHost())
        out[per_launch] = res
    for k in range(len(fr)):
        assert np.array_equal(out[1][k], out[4][k]), k
        assert np.array_equal(out[1][k], out[8][k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("iters", [8, 9])
def test_warm_solves_leave_the_same_buffers_whatever_the_iterations_per_launch(iters):
    """After every cold and warm solve, both buffers of each ping-pong pair (u, u_, v_: the
    current one and its partner, whose stale values the next solve's mask-edge reads see)
    hold the same values at every level-0 pixel whether the inner iterations ran one,
    four or eight per launch; also with an odd iteration count, which ends each warp with
    a one-iteration launch."""
    base, fr = frames(200, 120, 17, 5)
    names = ("u", "u_", "v_", "u.other", "u_.other", "v_.other")
    out = {}
    for per_launch in (1, 4, 8):
        g = gpu_solver(base, 3, 3, iters + 1)
        g.set_option("per_launch", per_launch)
        res = []
        for k, f in enumerate(fr):
            g.copyImagesToDevice(f["i0"], f["i1"])
            if k == 0:
                g.solveStereoForwardMasked()
            else:
                g.solve_warm(1 + k % 2, 2, iters)
            res.append({n: g.read_plane(n, 0) for n in names})
        out[per_launch] = res
    for k in range(len(fr)):
        for n in names:
            for pl in (4, 8):
                assert np.array_equal(out[1][k][n], out[pl][k][n], equal_nan=True), (k, n, pl)


def make_server(budget=16 << 30, max_active=64):
    from vfs_amd import Server
    return Server(P, hbm_budget_bytes=budget, max_streams=16, max_active=max_active)


@pytest.mark.gpu
def test_server_matches_standalone_solver():
    base, fr = frames(160, 120, 21, 4)
    srv = make_server()
    sid = srv.add_stream(base["mask"], base["tv"], base["cv"], 3, 3, 10, 1, 2, 5)
    for k, f in enumerate(fr):
        srv.submit(sid, f["i0"], f["i1"], tag=k)
    res = sorted(srv.run_until_idle(), key=lambda r: r["tag"])
    assert [r["warm"] for r in res] == [0, 1, 1, 1]
    ref = standalone_sequence(base, fr, 3, 3, 10, [0, 1, 1, 1])
    for r, u in zip(res, ref):
        assert np.array_equal(r["u"], u), r["tag"]


@pytest.mark.gpu
def test_server_alone_vs_mix():
    """Each stream's results are bit-identical served alone and mixed with streams of
    other resolutions whose frames arrive interleaved."""
    specs = [(160, 120, 31), (96, 80, 32), (224, 160, 33)]
    streams = [frames(W, H, seed, 3) for W, H, seed in specs]
    alone = []
    for base, fr in streams:
        srv = make_server()
        sid = srv.add_stream(base["mask"], base["tv"], base["cv"], 3, 3, 10, 1, 2, 5)
        for k, f in enumerate(fr):
            srv.submit(sid, f["i0"], f["i1"], tag=k)
        alone.append({r["tag"]: r["u"] for r in srv.run_until_idle()})
        srv.close()
    srv = make_server()
    sids = [srv.add_stream(b["mask"], b["tv"], b["cv"], 3, 3, 10, 1, 2, 5) for b, _ in streams]
    rng = np.random.default_rng(0)
    pending = [[(i, k) for k in range(3)] for i in range(len(streams))]
    results = []
    while any(pending):
        i = int(rng.choice([j for j in range(len(pending)) if pending[j]]))
        si, k = pending[i].pop(0)
        f = streams[si][1][k]
        srv.submit(sids[si], f["i0"], f["i1"], tag=100 * si + k)
        for _ in range(int(rng.integers(0, 40))):  # let some ticks run before the next arrival
            srv.step()
            results += srv.collect()
    results += srv.run_until_idle()
    assert len(results) == 9
    st = srv.stats()
    assert st["entries"] > st["ticks"]  # frames shared launches
    for r in results:
        si, k = divmod(r["tag"], 100)
        assert np.array_equal(r["u"], alone[si][k]), (si, k)


@pytest.mark.gpu
def test_server_budget_evicts_and_falls_back_to_cold():
    """A budget that holds two of three streams: streams are evicted and re-materialized,
    frames after an eviction run the full schedule, and every result still equals a
    standalone solver running the same cold/warm sequence."""
    specs = [(160, 120, 41), (160, 120, 42), (160, 120, 43)]
    streams = [frames(W, H, seed, 3) for W, H, seed in specs]
    probe = make_server()
    b0, f0 = streams[0]
    sid = probe.add_stream(b0["mask"], b0["tv"], b0["cv"], 3, 3, 10, 1, 2, 5)
    probe.submit(sid, f0[0]["i0"], f0[0]["i1"], tag=0)
    probe.run_until_idle()
    one = probe.stats()["hbm_used"]  # device bytes of one resident stream
    probe.close()
    assert one > 0
    from vfs_amd import Server
    srv = Server(P, hbm_budget_bytes=int(2.5 * one), max_streams=16, max_active=1)
    sids = [srv.add_stream(b["mask"], b["tv"], b["cv"], 3, 3, 10, 1, 2, 5) for b, _ in streams]
    for k in range(3):
        for si in range(3):
            f = streams[si][1][k]
            srv.submit(sids[si], f["i0"], f["i1"], tag=100 * si + k)
    res = srv.run_until_idle()
    assert len(res) == 9
    assert srv.stats()["resident_streams"] <= 2
    by = {}
    for r in res:
        si, k = divmod(r["tag"], 100)
        by.setdefault(si, {})[k] = r
    assert sum(1 - r["warm"] for r in res) > 3  # evictions forced extra cold starts
    for si, (base, fr) in enumerate(streams):
        flags = [by[si][k]["warm"] for k in range(3)]
        ref = standalone_sequence(base, fr, 3, 3, 10, flags)
        for k in range(3):
            assert np.array_equal(by[si][k]["u"], ref[k]), (si, k, flags)


@pytest.mark.gpu
def test_pinned_submission_equals_copying_submission():
    """Frames submitted from page-locked buffers (no host copy; one buffer backing several
    queued frames of two streams at once) give the results of copying submissions bit for
    bit, and every result carries its energy record."""
    specs = [(160, 120, 51), (96, 80, 52)]
    streams = [frames(W, H, seed, 3) for W, H, seed in specs]
    order = [(si, k) for k in range(5) for si in range(2)]   # frame k of each stream: distinct frame k % 3
    ref = {}
    srv = make_server()
    sids = [srv.add_stream(b["mask"], b["tv"], b["cv"], 3, 3, 10, 1, 2, 5) for b, _ in streams]
    for si, k in order:
        f = streams[si][1][k % 3]
        srv.submit(sids[si], f["i0"], f["i1"], tag=10 * si + k)
    for r in srv.run_until_idle():
        ref[r["tag"]] = r["u"]
    srv.close()
    srv = make_server()
    sids = [srv.add_stream(b["mask"], b["tv"], b["cv"], 3, 3, 10, 1, 2, 5) for b, _ in streams]
    pinned = [[srv.pinned_frame(sids[si], f["i0"], f["i1"]) for f in fr] for si, (_, fr) in enumerate(streams)]
    for si, k in order:   # all ten queued at once: frames k and k + 3 share a buffer
        srv.submit_pinned(sids[si], pinned[si][k % 3], tag=10 * si + k)
    res = srv.run_until_idle()
    assert len(res) == 10
    for r in res:
        assert np.array_equal(r["u"], ref[r["tag"]]), r["tag"]
        assert r["nonfinite"] == 0 and r["energy"] > 0
    srv.close()
