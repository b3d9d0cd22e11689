"""vfs_copy_warped_image_noalloc against vfs_copy_warped_image through the C ABI: the same
warped image bit for bit, on two calls in a row (the second from a smaller solver, so the
shared staging array holds the first call's larger plane), in FP16 and BF16 storage and
in FP32 (no staging); an image beyond the staging limit is rejected."""
import ctypes

import numpy as np
import pytest

from vfs_amd import StereoTgv, lib
from vfs_amd import synth as S

pytestmark = pytest.mark.gpu

P = S.SOLVER_PARAMS


def _solved(W, H, seed, storage):
    pair = S.make_pair(W, H, seed, radius=H / 2, dmax=W / 21.2)
    g = StereoTgv()
    g.set_option("storage", storage)
    g.initialize(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], 3,
                 P["fScale"], 2, 6)
    g.limitRange = P["limitRange"]
    g.copyMaskToDevice(pair["mask"])
    g.loadVectorFields(pair["tv"], pair["cv"])
    g.copyImagesToDevice(pair["i0"], pair["i1"])
    g.solveStereoForwardMasked()
    return g


def _noalloc(g):
    out = np.full((g.height, g.width), np.nan, np.float32)
    rc = lib().vfs_copy_warped_image_noalloc(g._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), g.width)
    assert rc == 0, (rc, lib().vfs_last_error(g._h))
    return out


@pytest.mark.parametrize("storage", ["f16", "bf16", "f32"])
def test_no_alloc_copy_equals_the_allocating_copy(storage):
    big = _solved(240, 200, 51, storage)
    small = _solved(120, 96, 52, storage)
    for g in (big, small, big):
        want = g.copyWarpedImageToHost()
        got = _noalloc(g)
        assert np.array_equal(got, want)
        assert np.isfinite(got).all()


def test_no_alloc_copy_rejects_planes_beyond_the_staging():
    g = StereoTgv()
    g.set_option("storage", "f16")
    # stride 2304 x 2048 rows = 4.5 Mi floats > 4 Mi
    g.initialize(2304, 2048, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], 1,
                 P["fScale"], 1, 1)
    out = np.zeros((2048, 2304), np.float32)
    rc = lib().vfs_copy_warped_image_noalloc(g._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 2304)
    assert rc == -1
    assert b"exceeds" in lib().vfs_last_error(g._h)
    g.close()
