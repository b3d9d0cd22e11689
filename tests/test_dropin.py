"""The C++ drop-in consumer (tests/dropin/test_vfs_dropin.cpp) replays the reference
demo's call sequence (test_vfs/main.cpp:104-231) through include/stereotgv/stereotgv.h
and libvfs.  Its outputs must equal the Python mirror's on the same inputs and meet
the parity bar against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from vfs_amd import synth as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "dropin", "test_vfs_dropin")


def write_pgm(path, img):
    h, w = img.shape
    with open(path, "wb") as f:
        f.write(f"P5 {w} {h} 255\n".encode())
        f.write(np.ascontiguousarray(img, np.uint8).tobytes())


def write_flo(path, field):
    h, w = field.shape[:2]
    with open(path, "wb") as f:
        np.array([202021.25], np.float32).tofile(f)
        np.array([w, h], np.int32).tofile(f)
        np.ascontiguousarray(field, np.float32).tofile(f)


def read_flo(path):
    with open(path, "rb") as f:
        tag = np.fromfile(f, np.float32, 1)[0]
        w, h = np.fromfile(f, np.int32, 2)
        assert tag == np.float32(202021.25)
        return np.fromfile(f, np.float32, w * h * 2).reshape(h, w, 2)


def test_dropin_binary_built():
    assert os.path.exists(BIN), "tests/dropin/test_vfs_dropin not built"
    r = subprocess.run([BIN], stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    assert r.returncode == 2 and b"usage" in r.stderr


def test_flo_roundtrip(tmp_path):
    f = np.random.default_rng(0).normal(size=(7, 9, 2)).astype(np.float32)
    write_flo(tmp_path / "a.flo", f)
    assert np.array_equal(read_flo(tmp_path / "a.flo"), f)


@pytest.mark.gpu
def test_dropin_matches_python_mirror_and_oracle(tmp_path):
    from oracle import oracle as O
    from vfs_amd import StereoTgv

    W, H, L, WI, IT = 192, 160, 3, 4, 15
    pair = S.make_pair(W, H, seed=31, radius=H / 2, dmax=W / 21.2)
    i0 = np.clip(pair["i0"] * 255 + 0.5, 0, 255).astype(np.uint8)
    i1 = np.clip(pair["i1"] * 255 + 0.5, 0, 255).astype(np.uint8)
    mask8 = (pair["mask"] * 255).astype(np.uint8)
    write_pgm(tmp_path / "i0.pgm", i0)
    write_pgm(tmp_path / "i1.pgm", i1)
    write_pgm(tmp_path / "mask.pgm", mask8)
    write_flo(tmp_path / "tv.flo", pair["tv"])
    write_flo(tmp_path / "cv.flo", pair["cv"])
    out = tmp_path / "out"
    out.mkdir()
    print(f"dropin: running {BIN} on {W}x{H} L{L} W{WI} I{IT}", flush=True)
    r = subprocess.run([BIN, str(tmp_path), str(out), str(L), "2.0", str(WI), str(IT), "1.0", "0"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    flow = read_flo(out / "disparity.flo")
    u = np.fromfile(out / "u.raw", np.float32).reshape(H, W)
    depth = np.fromfile(out / "depth.raw", np.float32).reshape(H, W)

    P = S.SOLVER_PARAMS
    mask = mask8.astype(np.float32) * np.float32(1.0 / 255.0)
    g = StereoTgv()
    g.initialize(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], L, 2.0, WI,
                 IT)
    g.limitRange = 1.0
    g.copyMaskToDevice(mask)
    g.loadVectorFields(pair["tv"], pair["cv"])
    g.copyImagesToDevice(i0, i1)
    g.solveStereoForwardMasked()
    assert np.array_equal(u, g.copy1DDisparityToHost())
    assert np.array_equal(flow, g.copyDisparityToHost())
    # the consumer scales the intrinsics in float, as test_vfs/main.cpp:208-211 does
    f32 = np.float32
    s = f32(848.0) / f32(W)
    dg, _ = g.copyStereoToHost(f32(285.722) / s, f32(286.759) / s, f32(420.135) / s, f32(403.394) / s, -0.00659769,
                               0.0473251, -0.0458264, 0.00897725, -0.0641854, -0.000218299, 0.000111253)
    assert np.array_equal(depth, dg, equal_nan=True)
    vis = np.fromfile(out / "vis.raw", np.float32).reshape(H, W, 3)
    assert np.array_equal(vis, g.copyDisparityVisToHost(40.0))

    o = O.Oracle(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], L, 2.0, WI,
                 IT)
    o.set_limit_range(1.0)
    o.copyMaskToDevice(mask)
    o.loadVectorFields(pair["tv"], pair["cv"])
    o.copyImagesToDevice(i0, i1)
    o.solveStereoForwardMasked()
    m = mask > 0
    err = np.abs(u - o.copy1DDisparityToHost())[m]
    assert err.mean() <= 5e-4, err.mean()


@pytest.mark.gpu
def test_dropin_on_the_reference_fixtures(tmp_path):
    """The consumer on the reference's own wang data (PNG frames and mask, .flo fields;
    equalize at 848x800 then halve, L5 W20 I10, limitRange 0.1, wang intrinsics: the run
    that wrote test_vfs/depth16.png): the C++ I/O and preprocessing give the same
    disparity as the Python mirror of them, and the depth meets the fixture pin."""
    from test_fixtures import D, compare, inputs, INTRINSICS
    from vfs_amd import StereoTgv
    from vfs_amd import imageio as io

    for src, dst in (("wang1.png", "i0.png"), ("wang2.png", "i1.png"), ("maskHalf.png", "mask.png"),
                     ("translationVectorWangHalf.flo", "tv.flo"), ("calibrationVectorWangHalf.flo", "cv.flo")):
        os.symlink(os.path.join(D, src), tmp_path / dst)
    out = tmp_path / "out"
    out.mkdir()
    print(f"dropin: running {BIN} on the wang fixtures", flush=True)
    r = subprocess.run([BIN, str(tmp_path), str(out), "5", "2.0", "20", "10", "0.1", "1", "1", "1"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    W, H = 424, 400
    u = np.fromfile(out / "u.raw", np.float32).reshape(H, W)
    depth = np.fromfile(out / "depth.raw", np.float32).reshape(H, W)
    i0, i1, mask, tv, cv = inputs()
    P = S.SOLVER_PARAMS
    g = StereoTgv()
    g.initialize(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], 5, 2.0, 20, 10)
    g.limitRange = 0.1
    g.copyMaskToDevice(mask)
    g.loadVectorFields(tv, cv)
    g.copyImagesToDevice(i0, i1)
    g.solveStereoForwardMasked()
    assert np.array_equal(u, g.copy1DDisparityToHost())
    compare("dropin depth16", io.to_u16(depth, 256), io.read_png_gray(os.path.join(D, "depth16.png")), mask > 0)
