"""Persistent launches (a warp's inner iterations in one launch of resident workgroups
that exchange tile borders every iteration) against the other launch forms, through the
C ABI: every plane bit for bit, in every tile shape, storage type, solve form and
iteration count, since each pixel runs the same code on the same operands.
"""
import numpy as np
import pytest

from vfs_amd import StereoTgv
from vfs_amd import synth as S

from test_gpu_parity import make_gpu, small_pair

pytestmark = pytest.mark.gpu

PLANES = ("u", "u_", "v_", "warpUV", "Iu", "Iz")
PERSIST_SHAPES = (7, 8, 11, 12, 13, 14, 17, 18)


def solve(pair, levels, warps, iters, masked=True, persist=0, storage="f32", per_launch=None, graph=True):
    g = make_gpu(pair, levels, warps, iters, storage=storage, graph=graph)
    g.set_option("persist", persist)
    if per_launch is not None:
        g.set_option("per_launch", per_launch)
    (g.solveStereoForwardMasked if masked else g.solveStereoForward)()
    return g


def planes_of(g, levels=(0,), flow=True):
    out = {f"{n}@{l}": g.read_plane(n, l) for n in PLANES for l in levels}
    if flow:
        out["flow"] = g.copyDisparityToHost()
    return out


def set_persistent(g, shape, masked=True, levels=3):
    """Fixes persistent launches in `shape` at every level where the device can hold all
    their tiles at once (at most 1024); returns those levels (at least one)."""
    from vfs_amd import VfsError
    done = []
    for lvl in range(levels):
        try:
            g.set_level_tuning(lvl, shape, per_launch=StereoTgv.PERSISTENT, masked=masked)
            done.append(lvl)
        except VfsError:
            w, h, _ = g.level_dims(lvl)
            tw, th = StereoTgv.SHAPES[shape]
            assert -(-w // tw) * -(-h // th) > 256, (shape, lvl, "persistent launch refused")
    assert done, shape
    return done


def assert_same(got, ref, what):
    for n, v in ref.items():
        assert np.array_equal(got[n], v, equal_nan=True), (what, n, float(np.nanmax(np.abs(got[n] - v))))


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("masked", [True, False])
def test_persistent_equals_one_iteration_per_launch(masked, storage):
    """Option persist 1 (the shape with the most tiles at each level) against one
    iteration per launch: 15 iterations per warp = one persistent launch of 14 + one."""
    pair = small_pair(200, 120, seed=37)
    ref = planes_of(solve(pair, 3, 2, 15, masked, persist=0, storage=storage, per_launch=1), (0, 1))
    g = solve(pair, 3, 2, 15, masked, persist=1, storage=storage)
    print(f"persist masked {masked} storage {storage}", flush=True)
    for lvl in range(3):
        assert g.level_tuning(lvl, masked, with_per_launch=True)[2] == StereoTgv.PERSISTENT
    assert_same(planes_of(g, (0, 1)), ref, "persist 1")
    assert g.get_option("persist_reruns") == 0, hex(g.get_option("persist_why"))


@pytest.mark.parametrize("masked", [True, False])
def test_every_persistent_shape(masked):
    """Every tile shape of the persistent launch at every level (8x8, 4x4, 36x25, 27x25,
    14x13, 7x7, 53x25, 53x50: one to three staged positions per work-item, partial tiles
    at the right and bottom edges) equals one iteration per launch."""
    pair = small_pair(200, 120, seed=41)
    ref = planes_of(solve(pair, 3, 2, 11, masked, persist=0, per_launch=1))
    for shape in PERSIST_SHAPES:
        g = make_gpu(pair, 3, 2, 11)
        print(f"persistent shape {shape} masked {masked}", flush=True)
        levels = set_persistent(g, shape, masked)
        (g.solveStereoForwardMasked if masked else g.solveStereoForward)()
        for lvl in levels:
            assert g.level_tuning(lvl, masked, with_per_launch=True)[::2] == (StereoTgv.SHAPES[shape],
                                                                              StereoTgv.PERSISTENT)
        assert_same(planes_of(g), ref, f"shape {shape}")
        assert g.get_option("persist_reruns") == 0, hex(g.get_option("persist_why"))


@pytest.mark.parametrize("masked", [True, False])
def test_persistent_odd_geometry(masked):
    """An odd-width pair (157x94: row padding, partial tiles in both directions; its
    trajectory field has a zero vector inside the mask, so the planes are compared
    without the numeric check of the copy-outs, NaN included)."""
    pair = small_pair(157, 94, seed=43)
    ref = planes_of(solve(pair, 3, 2, 9, masked, persist=0, per_launch=1), flow=False)
    for shape in PERSIST_SHAPES:
        g = make_gpu(pair, 3, 2, 9)
        print(f"odd geometry persistent shape {shape} masked {masked}", flush=True)
        set_persistent(g, shape, masked)
        (g.solveStereoForwardMasked if masked else g.solveStereoForward)()
        assert_same(planes_of(g, flow=False), ref, f"odd geometry, shape {shape}")


@pytest.mark.parametrize("iters", [1, 2, 3, 4, 130])
def test_persistent_iteration_counts(iters):
    """Iteration counts with no persistent launch (1), a launch of only its last two
    iterations (2), a remainder (3), and beyond the 128-entry step table (130 = 128 + 2)."""
    pair = small_pair(128, 96, seed=47)
    ref = planes_of(solve(pair, 2, 2, iters, persist=0, per_launch=1))
    print(f"persistent iterations {iters}", flush=True)
    assert_same(planes_of(solve(pair, 2, 2, iters, persist=1)), ref, f"{iters} iterations")


def test_persistent_with_mask_in_column_0():
    """A mask pixel in column 0 of a level forbids multi-iteration launches there (the
    reference's flat-index left neighbour); the other levels still run persistent ones."""
    pair = small_pair(160, 96, seed=51)
    m = pair["mask"].copy()
    m[40:44, 0:3] = 1.0
    pair = dict(pair, mask=m)
    ref = planes_of(solve(pair, 3, 2, 10, persist=0, per_launch=1))
    g = solve(pair, 3, 2, 10, persist=1)
    ks = [g.level_tuning(l, True, with_per_launch=True)[2] for l in range(3)]
    print(f"per_launch by level with column-0 mask: {ks}", flush=True)
    assert ks[0] != StereoTgv.PERSISTENT
    assert_same(planes_of(g), ref, "mask in column 0")


def test_persistent_then_warm_start_and_graphs():
    """A warm start after a persistent cold solve continues from the same buffers as after
    any other launch form; direct launches equal graph replays; a second cold solve (the
    launch counter advanced, the same images copied in again: a solve without new images
    calibrates I1 again, as the reference's does) gives the same result again."""
    pair = small_pair(200, 120, seed=53)
    outs = []
    for persist, graph in ((0, True), (1, True), (1, False)):
        g = solve(pair, 3, 3, 12, persist=persist, per_launch=1 if persist == 0 else None, graph=graph)
        first = planes_of(g)
        g.copyImagesToDevice(pair["i0"], pair["i1"])
        g.solveStereoForwardMasked()
        again = planes_of(g)
        assert_same(again, first, f"second cold solve persist {persist}")
        g.solve_warm(1, 2, 10)
        outs.append((first, planes_of(g)))
        print(f"warm after persist {persist} graph {graph}", flush=True)
    for first, warm in outs[1:]:
        assert_same(first, outs[0][0], "cold")
        assert_same(warm, outs[0][1], "warm")


def test_persistent_848x800_workload():
    """The 848x800 L5 W10 I50 masked solve (53x50 tiles at level 0, three staged positions
    per work-item) with the tuner's choices equals one iteration per launch."""
    pair = S.make_pair(848, 800, seed=2, radius=400, dmax=40)
    ref = planes_of(solve(pair, 5, 10, 50, persist=0, per_launch=1))
    g = solve(pair, 5, 10, 50, persist=1)
    tun = [g.level_tuning(l, True, with_per_launch=True) for l in range(5)]
    print(f"848x800 persistent tuning: {tun}", flush=True)
    assert tun[0][::2] == (StereoTgv.SHAPES[18], StereoTgv.PERSISTENT)
    assert_same(planes_of(g), ref, "848x800")
    assert g.get_option("persist_reruns") == 0, hex(g.get_option("persist_why"))
    # back-to-back solves (no launch gives up: every tile waits for all its neighbours'
    # progress, so none overwrites exchange words a neighbour has not read)
    g.time_solves(True, 10)
    assert g.get_option("persist_reruns") == 0, hex(g.get_option("persist_why"))


# blocked persistent launches (blocks of K iterations with recomputed halos, rings
# exchanged between blocks): the tile shapes of each K
BLOCKED = ((2, (6, 7, 8, 10, 12, 13)), (4, (6, 7, 10, 12, 13)), (8, (6, 10)))


def set_blocked(g, shape, K, masked=True, levels=3):
    """Fixes blocked persistent launches (blocks of K) in `shape` at every level where the
    device holds all their tiles at once; returns those levels (at least one)."""
    from vfs_amd import VfsError
    done = []
    for lvl in range(levels):
        try:
            g.set_level_tuning(lvl, shape, per_launch=StereoTgv.PERSISTENT | K, masked=masked)
            done.append(lvl)
        except VfsError:
            w, h, _ = g.level_dims(lvl)
            tw, th = StereoTgv.SHAPES[shape]
            assert -(-w // tw) * -(-h // th) > 256, (shape, K, lvl, "blocked persistent launch refused")
    assert done, (shape, K)
    return done


@pytest.mark.parametrize("masked", [True, False])
def test_every_blocked_persistent_form(masked):
    """Blocks of 2, 4 and 8 iterations in every tile shape that has them, at every level
    (19 iterations per warp: one launch of 18, whose last block of 4 or 8 is shortened to 2,
    and one iteration in another launch), equal one iteration per launch."""
    pair = small_pair(200, 120, seed=59)
    ref = planes_of(solve(pair, 3, 2, 19, masked, persist=0, per_launch=1), (0, 1))
    for K, shapes in BLOCKED:
        for shape in shapes:
            g = make_gpu(pair, 3, 2, 19)
            print(f"blocked persistent K {K} shape {shape} masked {masked}", flush=True)
            levels = set_blocked(g, shape, K, masked)
            (g.solveStereoForwardMasked if masked else g.solveStereoForward)()
            for lvl in levels:
                assert g.level_tuning(lvl, masked, with_per_launch=True)[::2] == (StereoTgv.SHAPES[shape],
                                                                                  StereoTgv.PERSISTENT | K)
            assert_same(planes_of(g, (0, 1)), ref, f"K {K} shape {shape}")
            assert g.get_option("persist_reruns") == 0, hex(g.get_option("persist_why"))


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_blocked_persistent_storage(storage):
    """Half-precision planes: p and q rounded to the storage type between iterations, the
    rings carrying the unrounded values (every dual step reads them rounded)."""
    pair = small_pair(200, 120, seed=61)
    ref = planes_of(solve(pair, 3, 2, 17, persist=0, storage=storage, per_launch=1), (0, 1))
    for K, shape in ((4, 10), (8, 6), (2, 13)):
        g = make_gpu(pair, 3, 2, 17, storage=storage)
        print(f"blocked persistent {storage} K {K} shape {shape}", flush=True)
        set_blocked(g, shape, K)
        g.solveStereoForwardMasked()
        assert_same(planes_of(g, (0, 1)), ref, f"{storage} K {K} shape {shape}")


@pytest.mark.parametrize("masked", [True, False])
def test_blocked_persistent_odd_geometry(masked):
    """157x94 (row padding, partial tiles narrower than a ring in both directions)."""
    pair = small_pair(157, 94, seed=67)
    ref = planes_of(solve(pair, 3, 2, 12, masked, persist=0, per_launch=1), flow=False)
    for K, shape in ((2, 8), (4, 7), (4, 12), (8, 6), (8, 10)):
        g = make_gpu(pair, 3, 2, 12)
        print(f"odd geometry blocked K {K} shape {shape} masked {masked}", flush=True)
        set_blocked(g, shape, K, masked)
        (g.solveStereoForwardMasked if masked else g.solveStereoForward)()
        assert_same(planes_of(g, flow=False), ref, f"odd geometry, K {K} shape {shape}")


@pytest.mark.parametrize("iters", [2, 4, 6, 9, 130])
def test_blocked_persistent_iteration_counts(iters):
    """Fewer iterations than a block (2: one shortened block), one block (4), a full block
    of 4 and a shortened one or one shortened block of 8 (6), the same and one iteration in
    another launch (9), beyond the 128-entry step table (130)."""
    pair = small_pair(128, 96, seed=71)
    ref = planes_of(solve(pair, 2, 2, iters, persist=0, per_launch=1))
    for K, shape in ((4, 7), (8, 6)):
        g = make_gpu(pair, 2, 2, iters)
        set_blocked(g, shape, K, levels=2)
        print(f"blocked persistent iterations {iters} K {K}", flush=True)
        g.solveStereoForwardMasked()
        assert_same(planes_of(g), ref, f"{iters} iterations, K {K}")


def test_tuned_848x800_solve_equals_one_iteration_per_launch():
    """The tuner's mixture at 848x800 (whole-warp persistent launches at level 0, blocked
    or whole-warp ones and eight-iteration launches below, each warp's derivatives in its
    first launch) equals one iteration per launch bit for bit, with no re-run."""
    pair = S.make_pair(848, 800, seed=5, radius=400, dmax=40)
    ref = planes_of(solve(pair, 5, 10, 50, persist=0, per_launch=1))
    g = make_gpu(pair, 5, 10, 50)
    g.solveStereoForwardMasked()
    tun = [g.level_tuning(l, True, with_per_launch=True) for l in range(5)]
    print(f"848x800 tuned launch choices: {tun}", flush=True)
    assert tun[0][2] & StereoTgv.PERSISTENT
    assert_same(planes_of(g), ref, "tuned 848x800")
    assert g.get_option("persist_reruns") == 0, hex(g.get_option("persist_why"))


@pytest.mark.parametrize("iters", [6, 15])
def test_persistent_launches_that_give_up_run_again(iters):
    """A solve whose persistent launches give up (forced from the host: option
    persist_give_up sets their abort flag, every workgroup leaves at once) runs again
    without them at the next call that waits for it, and that call reads the re-run's
    buffers, whichever copy-out it is: 6 iterations per warp are one persistent launch (one
    role rotation) against a four- and a two-iteration launch (two) in the re-run."""
    pair = small_pair(200, 120, seed=73)
    ref = solve(pair, 3, 2, iters, persist=0, per_launch=1)
    ref_planes, ref_1d = planes_of(ref, (0, 1)), ref.copy1DDisparityToHost()
    g = make_gpu(pair, 3, 2, iters)
    for lvl in range(3):
        g.set_level_tuning(lvl, 4 if lvl < 2 else 6, per_launch=4)
    set_persistent(g, 7)
    g.solveStereoForwardMasked()
    assert g.get_option("persist_reruns") == 0, hex(g.get_option("persist_why"))
    for round_, first in enumerate(("1d", "flow", "planes")):
        g.set_option("persist_give_up", 1)
        g.copyImagesToDevice(pair["i0"], pair["i1"])
        g.solveStereoForwardMasked()
        print(f"give up: iterations {iters}, read first: {first}", flush=True)
        if first == "1d":
            assert np.array_equal(g.copy1DDisparityToHost(), ref_1d, equal_nan=True)
        elif first == "flow":
            assert np.array_equal(g.copyDisparityToHost(), ref_planes["flow"], equal_nan=True)
        assert_same(planes_of(g, (0, 1)), ref_planes, f"re-run, {first} read first")
        assert g.get_option("persist_reruns") == round_ + 1
    # a parameter changed before the results are read does not reach the re-run
    g.set_option("persist_give_up", 1)
    g.copyImagesToDevice(pair["i0"], pair["i1"])
    g.solveStereoForwardMasked()
    limit = g.limitRange
    g.limitRange = limit * 0.5
    assert_same(planes_of(g, (0, 1)), ref_planes, "re-run after a parameter change")
    assert g.get_option("persist_reruns") == 4
    g.limitRange = limit
    # and the next solve runs its persistent launches again
    g.copyImagesToDevice(pair["i0"], pair["i1"])
    g.solveStereoForwardMasked()
    assert_same(planes_of(g, (0, 1)), ref_planes, "after the re-runs")
    assert g.get_option("persist_reruns") == 4


def test_persistent_level_tuning_rejects_what_cannot_run():
    """VFS_PERSISTENT | K takes K = 2, 4 or 8 in a shape that has the blocked form, no
    persistent choice while option persist is 0; the solver stays usable after the errors."""
    from vfs_amd import VfsError
    pair = small_pair(128, 96, seed=79)
    ref = planes_of(solve(pair, 2, 1, 6, persist=0, per_launch=1))
    g = make_gpu(pair, 2, 1, 6)
    with pytest.raises(VfsError):
        g.set_level_tuning(0, 7, per_launch=StereoTgv.PERSISTENT | 3)
    with pytest.raises(VfsError):
        g.set_level_tuning(0, 8, per_launch=StereoTgv.PERSISTENT | 4)  # 4x4 tiles, blocks of 4
    with pytest.raises(VfsError):
        g.set_level_tuning(0, 0, per_launch=StereoTgv.PERSISTENT)  # 64x8: no persistent form
    g.set_option("persist", 0)
    with pytest.raises(VfsError):
        g.set_level_tuning(0, 7, per_launch=StereoTgv.PERSISTENT)
    g.set_option("persist", -1)
    g.set_level_tuning(0, 7, per_launch=StereoTgv.PERSISTENT | 2)
    g.solveStereoForwardMasked()
    assert_same(planes_of(g), ref, "after the rejected choices")
