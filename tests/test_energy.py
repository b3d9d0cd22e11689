"""GPU energy monitor (OP_ENERGY / OP_EFINAL) against the oracle's energy function.

The FP64 reduction on the GPU is checked on the planes it reduced: the same planes read
back from the GPU and fed to the oracle's k_energy give the same sums to 1e-9 relative
(the per-pixel terms are evaluated in double on both sides; only the order of the sums
differs).  The records of a solve follow the FP64 oracle's own solve within the solve's
FP32 drift, do not depend on launch choices or batching, and a zero trajectory vector
inside the mask is reported as VFS_ERR_NUMERIC.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vfs_amd import Batch, StereoTgv, VfsError
from vfs_amd import synth as S

pytestmark = pytest.mark.gpu

P = S.SOLVER_PARAMS


def _gpu(pair, levels, warps, iters, masked=True, limit=None, storage="f32", energy=1, **opts):
    H, W = pair["i0"].shape
    g = StereoTgv()
    g.set_option("storage", storage)
    g.set_option("energy", energy)
    for k, v in opts.items():
        g.set_option(k, v)
    g.initialize(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], levels,
                 P["fScale"], warps, iters)
    g.limitRange = P["limitRange"] if limit is None else limit
    g.copyMaskToDevice(pair["mask"])
    g.loadVectorFields(pair["tv"], pair["cv"])
    g.copyImagesToDevice(pair["i0"], pair["i1"])
    if masked:
        g.solveStereoForwardMasked()
    else:
        g.solveStereoForward()
    return g


def _oracle(pair, levels, warps, iters, masked=True, limit=None):
    o = O.solve_pair(pair, levels, warps, iters, P, masked=masked, limit=limit)
    return o


@pytest.mark.parametrize("masked", [True, False])
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_energy_matches_oracle_function_on_the_gpu_planes(masked, storage):
    """One level, one warp: after the solve the planes the reduction read are still there
    (the end of the warp keeps v_, Iu, Iz, u_last and the tensor; masked, the pre-median u
    is the other u buffer; unmasked with a limit that never clamps, u itself)."""
    pair = S.make_pair(112, 96, 23, radius=48, dmax=112 / 21.2)
    g = _gpu(pair, 1, 1, 6, masked=masked, limit=1e6, storage=storage)
    rec = g.energy(0, 0)
    u = g.read_plane("u.other" if masked else "u", 0)
    planes = {n: g.read_plane(n, 0) for n in ("v_", "Iu", "Iz", "u_last", "a", "b", "c")}
    # the solver's parameters are floats (stereotgv.h:107-109)
    f = lambda v: float(np.float32(v))
    ref = O.k_energy(u, u, planes["v_"], planes["Iu"], planes["Iz"], planes["u_last"], planes["a"], planes["b"],
                     planes["c"], f(P["lam"]), f(P["alpha0"]), f(P["alpha1"]), pair["mask"] if masked else None)
    print(f"{storage} masked={masked}: gpu {rec}, oracle function {ref}")
    assert rec["pixels"] == ref["pixels"] > 0 and rec["nonfinite"] == 0
    # masked: the exact planes; unmasked, the warp end rewrites u as u_last + (u - u_last),
    # which can move u by one float rounding
    tol = 1e-9 if masked else 1e-6
    assert abs(rec["energy"] - ref["energy"]) <= tol * ref["energy"]
    assert abs(rec["data"] - ref["data"]) <= tol * ref["data"]
    assert rec["residual"] > 0


def test_energy_records_follow_the_oracle_solve():
    """Every (level, warp) record of an L3 W3 I8 solve (option "energy" 2) against the FP64
    oracle's record of the same warp."""
    pair = S.make_pair(128, 96, 29, radius=48, dmax=128 / 21.2)
    L, Wi, It = 3, 3, 8
    g = _gpu(pair, L, Wi, It, energy=2)
    o = _oracle(pair, L, Wi, It)
    for lvl in range(L):
        for wi in range(Wi):
            r, q = g.energy(lvl, wi), o.energy(lvl, wi)
            print(f"L{lvl} W{wi}: gpu E {r['energy']:.6e} D {r['data']:.6e} R {r['residual']:.4e} | "
                  f"fp64 E {q['energy']:.6e} D {q['data']:.6e} R {q['residual']:.4e}")
            assert r["pixels"] == q["pixels"] and r["nonfinite"] == q["nonfinite"] == 0
            assert abs(r["energy"] - q["energy"]) <= 1e-3 * q["energy"]
            assert abs(r["data"] - q["data"]) <= 5e-3 * q["data"]
            assert abs(r["residual"] - q["residual"]) <= 0.2 * q["residual"] + 1e-6 * q["pixels"]


def test_energy_does_not_depend_on_launch_choices_or_batching():
    """The reduction order is fixed by the level's pixel order: one or four iterations per
    launch, another tile shape and a batch with another pair give identical records (the
    solves themselves are bit-identical in every launch form)."""
    pair = S.make_pair(128, 96, 31, radius=48, dmax=128 / 21.2)
    other = S.make_pair(160, 120, 32, radius=60, dmax=160 / 21.2)
    ref = _gpu(pair, 3, 2, 10, energy=2)
    recs = [[ref.energy(l, w) for w in range(2)] for l in range(3)]
    for opts in ({"per_launch": 1}, {"per_launch": 4}, {"shape": 3}):
        g = _gpu(pair, 3, 2, 10, energy=2, **opts)
        assert [[g.energy(l, w) for w in range(2)] for l in range(3)] == recs, opts
    members = []
    for p in (other, pair):
        H, W = p["i0"].shape
        s = StereoTgv()
        s.set_option("energy", 2)
        s.initialize(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], 3,
                     P["fScale"], 2, 10)
        s.limitRange = P["limitRange"]
        s.copyMaskToDevice(p["mask"])
        s.loadVectorFields(p["tv"], p["cv"])
        s.copyImagesToDevice(p["i0"], p["i1"])
        members.append(s)
    Batch(members).solve(True)
    assert [[members[1].energy(l, w) for w in range(2)] for l in range(3)] == recs


def test_energy_option_changes_no_result():
    pair = S.make_pair(96, 80, 37, radius=40, dmax=96 / 21.2)
    u1 = _gpu(pair, 2, 2, 5, energy=1).copy1DDisparityToHost()
    g0 = _gpu(pair, 2, 2, 5, energy=0)
    assert np.array_equal(g0.copy1DDisparityToHost(), u1)
    with pytest.raises(VfsError):
        g0.energy(0, 1)
    g2 = _gpu(pair, 2, 2, 5, energy=2)
    assert np.array_equal(g2.copy1DDisparityToHost(), u1)


def test_zero_trajectory_vector_is_reported():
    """A zero trajectory vector at a mask pixel makes its derivative NaN
    (computeDerivatives.cu:99-104): the level-0 record counts it, the oracle's counts the
    same pixel, and the copy-out (still filled) returns VFS_ERR_NUMERIC."""
    pair = S.make_pair(96, 80, 41, radius=40, dmax=96 / 21.2)
    pair["tv"] = pair["tv"].copy()
    pair["tv"][40, 30] = 0.0
    assert pair["mask"][40, 30] > 0
    g = _gpu(pair, 2, 2, 5)
    o = _oracle(pair, 2, 2, 5)
    assert g.energy(0, 1)["nonfinite"] == o.energy(0, 1)["nonfinite"] == 1
    with pytest.raises(VfsError, match=r"\(-5\).*non-finite"):
        g.copy1DDisparityToHost()
    with pytest.raises(VfsError, match=r"\(-5\)"):
        g.synchronize()
    # a solve of a clean field clears the condition
    clean = S.make_pair(96, 80, 41, radius=40, dmax=96 / 21.2)
    g.loadVectorFields(clean["tv"], clean["cv"])
    g.solveStereoForwardMasked()
    assert np.isfinite(g.copy1DDisparityToHost()).all()
