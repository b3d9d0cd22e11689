"""GPU parity at the stated workloads: the 64-pair mixed-size batch, 16-pair reduced
storage batches at 848x800, served FP16 camera streams over 100 frames, and the
reference's own demo configuration (test_vfs/main.cpp:132-145: 11 levels at fScale 1.2,
limitRange 0.1; warps and iterations reduced so that the FP64 oracle finishes in about
a second).  Each test is a single pass; the oracle runs only where it is compared.
"""
import functools

import numpy as np
import pytest

from oracle import oracle as O
from vfs_amd import Batch, Server, StereoTgv
from vfs_amd import synth as S

pytestmark = pytest.mark.gpu

P = S.SOLVER_PARAMS


def _solver(pair, levels, warps, iters, storage="f32", fscale=None, limit=None):
    H, W = pair["i0"].shape
    g = StereoTgv()
    g.set_option("storage", storage)
    g.initialize(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], levels,
                 P["fScale"] if fscale is None else fscale, warps, iters)
    g.limitRange = P["limitRange"] if limit is None else limit
    g.copyMaskToDevice(pair["mask"])
    g.loadVectorFields(pair["tv"], pair["cv"])
    g.copyImagesToDevice(pair["i0"], pair["i1"])
    return g


def _tail(name, err, mean_max, p999_max, frac_over_1px):
    q = float(np.percentile(err, 99.9))
    over = int((err > 1.0).sum())
    print(f"{name}: mean {err.mean():.3e} px, p99 {np.percentile(err, 99):.3e}, p99.9 {q:.3e}, "
          f"over 1 px {over} of {err.size}, max {err.max():.3f}")
    assert err.mean() <= mean_max, (name, err.mean())
    assert q <= p999_max, (name, q)
    assert over <= frac_over_1px * err.size, (name, over)


@functools.lru_cache(maxsize=None)
def _pair(W, H, seed, radius, dmax):
    return S.make_pair(W, H, seed, radius, dmax)


@functools.lru_cache(maxsize=None)
def _oracle_u(W, H, seed, radius, dmax, levels, warps, iters, fscale=None, limit=None, precision="f64"):
    params = dict(P)
    if fscale is not None:
        params["fScale"] = fscale
    o = O.solve_pair(_pair(W, H, seed, radius, dmax), levels, warps, iters, params, precision=precision, limit=limit)
    return o.copy1DDisparityToHost().astype(np.float64)


def test_batched_64_mixed_pairs_equal_solo_and_follow_the_oracle():
    """64 pairs, sizes 424x400 / 640x480 / 848x800 in turn (seeds 100 + i), L5 W10 I50,
    one batch: every member equals the same pair solved alone (its own launch choices)
    bit for bit, and one pair of each size meets the FP64 oracle (640x480 includes the
    tensor's column limit inside the mask: columns 480..639 of level 0)."""
    sizes = [(424, 400), (640, 480), (848, 800)]
    keys = [(sizes[i % 3][0], sizes[i % 3][1], 100 + i, sizes[i % 3][1] / 2.0, sizes[i % 3][0] / 21.2)
            for i in range(64)]
    solvers = [_solver(_pair(*k), 5, 10, 50) for k in keys]
    b = Batch(solvers)
    b.solve(True)
    batch_u = [s.copy1DDisparityToHost() for s in solvers]
    print("batch tuning:", [b.level_tuning(l, with_per_launch=True) for l in range(5)])
    for i, (s, k) in enumerate(zip(solvers, keys)):
        p = _pair(*k)
        s.copyImagesToDevice(p["i0"], p["i1"])
        s.solveStereoForwardMasked()
        assert np.array_equal(s.copy1DDisparityToHost(), batch_u[i]), (i, k[:2])
    # One pair of each size against the FP64 oracle.  Every pair meets the 0.05 px bar;
    # the regression bounds are those of the reference's own float arithmetic on the same
    # input: the FP32 oracle (the algorithm in float) drifts from the FP64 one by a
    # seed-dependent amount (0.0015 px mean at the 848x800 seed 2 pair, 0.0054 px at seed
    # 102), and the GPU may drift by at most 1.5 times that in the mean (or 0.005 px) and
    # twice that in the tail.
    for i in range(3):
        k = keys[i]
        m = _pair(*k)["mask"] > 0
        ref = _oracle_u(*k, 5, 10, 50)
        err = np.abs(batch_u[i] - ref)
        e32 = np.abs(_oracle_u(*k, 5, 10, 50, precision="f32") - ref)[m]
        p32 = float(np.percentile(e32, 99.9))
        print(f"batch member {i} {k[0]}x{k[1]}: FP32 oracle drift mean {e32.mean():.3e} px, p99.9 {p32:.3e}, "
              f"over 1 px {(e32 > 1).sum()}")
        assert err[m].mean() <= 0.05, err[m].mean()
        _tail(f"batch member {i} {k[0]}x{k[1]}", err[m], max(0.005, 1.5 * e32.mean()), max(0.6, 2 * p32),
              max(3e-4, 2 * (e32 > 1).mean()))
        if k[0] == 640:
            # 640x480: the tensor grid covers columns x < 32 ceil(480 / 32) = 480 only
            # (calcTensor.cu:81, 227), so inside the mask columns 480..559 keep
            # a = b = c = 0 and the threshold there steps with tau unpreconditioned; the
            # float arithmetic is sensitive in that band (FP32 oracle 0.34 px mean there,
            # 0.002 px elsewhere).  Outside the band the fixed regression bound holds.
            band = np.zeros_like(m)
            band[:, 480:] = True
            _tail("640x480 outside the zero-tensor band", err[m & ~band], 0.005, 0.6, 3e-4)
    b.close()
    for s in solvers:
        s.close()


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_reduced_storage_16_pairs_848x800(storage):
    """16 pairs at 848x800 (seeds 200 + i) in one batch with 16-bit storage of images,
    derivatives, tensor and duals: member 0 meets the 0.05 px bar against the FP64 oracle
    and equals its solo solve bit for bit."""
    keys = [(848, 800, 200 + i, 400.0, 40.0) for i in range(16)]
    solvers = [_solver(_pair(*k), 5, 10, 50, storage=storage) for k in keys]
    Batch(solvers).solve(True)
    u0 = solvers[0].copy1DDisparityToHost()
    p = _pair(*keys[0])
    solvers[0].copyImagesToDevice(p["i0"], p["i1"])
    solvers[0].solveStereoForwardMasked()
    assert np.array_equal(solvers[0].copy1DDisparityToHost(), u0)
    err = np.abs(u0 - _oracle_u(*keys[0], 5, 10, 50))[p["mask"] > 0]
    _tail(f"{storage} 16x848x800 member 0", err, 0.05, 1.5, 2e-3)
    for s in solvers:
        s.close()


def test_served_fp16_streams_follow_the_fp64_warm_schedule():
    """Serving: four FP16 streams (424x400, 848x800, 424x400, 848x800; seeds 300 + s), 100
    frames each, all submitted at once so that the server runs them by continuous
    batching.  Streams 0 and 1 carry the drifting sequence frame by frame; the FP64 oracle
    runs the same sequence (cold first frame, then warm W2 I20 from level 1) and the
    served frames stay within the 0.05 px bar at every tenth frame."""
    n_frames = 100
    sizes = [(424, 400), (848, 800), (424, 400), (848, 800)]
    bases = [S.make_pair(W, H, 300 + s, H / 2.0, W / 21.2) for s, (W, H) in enumerate(sizes)]
    srv = Server(P, hbm_budget_bytes=16 << 30, max_streams=4, max_active=64, storage="f16")
    sids = [srv.add_stream(b["mask"], b["tv"], b["cv"], 5, 10, 50, 1, 2, 20) for b in bases]
    frames = {}
    for s in range(4):
        seq = [bases[s]] + [S.drift_pair(bases[s], k, 300 + s) for k in (range(1, n_frames) if s < 2 else (1, 2, 3))]
        frames[s] = seq
    for k in range(n_frames):
        for s in range(4):
            f = frames[s][k] if s < 2 else frames[s][k % 4]
            srv.submit(sids[s], f["i0"], f["i1"], tag=k * 4 + s, want_result=s < 2)
    res = srv.run_until_idle(timeout_s=300)
    assert len(res) == 4 * n_frames
    assert all(r["nonfinite"] == 0 and r["energy"] > 0 for r in res)
    assert sum(r["warm"] for r in res) == 4 * (n_frames - 1)
    st = srv.stats()
    print(f"served: {st}, ops per launch {st['entries'] / max(st['ticks'], 1):.2f}")
    got = {r["tag"]: r["u"] for r in res if r["u"] is not None}
    srv.close()
    for s in range(2):
        W, H = sizes[s]
        o = O.Oracle(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], 5,
                     P["fScale"], 10, 50)
        o.set_limit_range(P["limitRange"])
        o.copyMaskToDevice(bases[s]["mask"])
        o.loadVectorFields(bases[s]["tv"], bases[s]["cv"])
        m = bases[s]["mask"] > 0
        for k in range(n_frames):
            f = frames[s][k]
            o.copyImagesToDevice(f["i0"], f["i1"])
            if k == 0:
                o.solveStereoForwardMasked()
            else:
                o.solve_warm(1, 2, 20)
            if k % 10 == 0 or k == n_frames - 1:
                err = np.abs(got[k * 4 + s] - o.copy1DDisparityToHost())[m]
                print(f"stream {s} ({W}x{H}) frame {k}: mean {err.mean():.3e} px, max {err.max():.3f}")
                assert err.mean() <= 0.05, (s, k, err.mean())


def test_reference_demo_configuration_848x800_11_levels():
    """test_vfs/main.cpp:132-145: 848x800, nLevel 11, fScale 1.2, limitRange 0.1 (its 100
    warps x 100 iterations reduced to W10 I20): non-power-of-2 pyramids down to 134x127,
    the upscale by fScale, and a clamp that binds."""
    k = (848, 800, 2, 400.0, 40.0)
    p = _pair(*k)
    g = _solver(p, 11, 10, 20, fscale=1.2, limit=0.1)
    g.solveStereoForwardMasked()
    dims = [g.level_dims(l)[:2] for l in range(11)]
    print("levels:", dims, "tuning:", [g.level_tuning(l, with_per_launch=True) for l in range(11)])
    assert dims[1] == (706, 666) and dims[10] == (134, 127)
    err = np.abs(g.copy1DDisparityToHost() - _oracle_u(*k, 11, 10, 20, fscale=1.2, limit=0.1))[p["mask"] > 0]
    _tail("demo config L11 fScale 1.2 limit 0.1 W10 I20", err, 0.005, 0.6, 3e-4)
