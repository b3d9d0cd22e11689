"""FlowToHSV colour coding (flowToHsv.cu:3-121) behind copyDisparityVisToHost
(stereotgv.cu:551-559).

Oracle pins: the colour wheel's anchor colours are the primaries and secondaries of the
Middlebury flow colour code at the published segment boundaries (red-yellow 15,
yellow-green 6, green-cyan 4, cyan-blue 11, blue-magenta 13, magenta-red 6); zero flow
is white; a flow of magnitude m <= 1 blends the wheel colour towards white by m and a
longer one is the wheel colour darkened to 0.75; the angle convention is
atan2(-dv, -du) over 3.14159 (flow pointing to -x lands mid-wheel, on cyan-blue); the
output channel order is (B, G, R); the coding is continuous in angle except at +-pi,
where the reference's 1-based wheel positions (entries 1..54, then 0) step from entry 0
back to entry 1."""
import numpy as np
import pytest

from oracle import oracle as O
from vfs_amd import synth as S

ANCHORS = {0: (255, 0, 0), 15: (255, 255, 0), 21: (0, 255, 0), 25: (0, 255, 255), 36: (0, 0, 255),
           49: (255, 0, 255)}


def flow_at_position(fk, mag):
    """Flow whose wheel position is fk (1-based, 1 <= fk < 55) with magnitude mag."""
    a = (fk - 1) / 54.0 * 2 - 1
    ang = a * 3.14159
    # a = atan2(-dv, -du) / 3.14159  ->  (-du, -dv) = mag (cos, sin)
    return np.array([-mag * np.cos(ang), -mag * np.sin(ang)])


def test_wheel_anchor_colours():
    for k, rgb in ANCHORS.items():
        assert tuple(O.wheel_entry(k).astype(int)) == rgb, k
    # monotone ramps between anchors: every entry differs from its neighbour in one channel
    for k in range(55):
        d = np.abs(O.wheel_entry((k + 1) % 55) - O.wheel_entry(k))
        assert (d > 0).sum() == 1, k


def test_zero_flow_is_white_and_magnitude_rules():
    assert np.allclose(O.k_flow_hsv(np.zeros((1, 1, 2)), 1.0), 1.0)
    for k, rgb in ANCHORS.items():
        if k == 0:
            continue
        pos = k  # position fk = k (fraction 0) reads wheel entry k
        c = np.array(rgb[::-1]) / 255.0  # (B, G, R)
        far = O.k_flow_hsv(flow_at_position(pos, 2.0).reshape(1, 1, 2), 1.0)[0, 0]
        assert np.allclose(far, 0.75 * c, atol=1e-5), (k, far)
        near = O.k_flow_hsv(flow_at_position(pos, 0.4).reshape(1, 1, 2), 1.0)[0, 0]
        assert np.allclose(near, 1 - 0.4 * (1 - c), atol=1e-5), (k, near)
        # the flow scale divides the flow
        scaled = O.k_flow_hsv((flow_at_position(pos, 0.4) * 40).reshape(1, 1, 2), 40.0)[0, 0]
        assert np.allclose(scaled, near, atol=1e-6)


def test_flow_to_minus_x_is_mid_wheel():
    out = O.k_flow_hsv(np.array([[[-3.0, 0.0]]]), 1.0)[0, 0]
    # position 28: entry 28, the fourth of the cyan-blue ramp (G = 255 - floor(255 * 3 / 11) = 186)
    assert np.allclose(out, 0.75 * np.array([255, 186, 0]) / 255.0, atol=1e-6)


def test_colour_is_continuous_in_angle_except_at_the_wrap():
    # flow angle theta: the wheel angle is atan2(-dv, -du) = theta - pi (mod 2 pi)
    theta = np.linspace(1e-4, 2 * np.pi - 1e-4, 20001)
    uv = np.stack([np.cos(theta), np.sin(theta)], -1).reshape(1, -1, 2) * 2.0
    c = O.k_flow_hsv(uv, 1.0)[0]
    assert np.abs(np.diff(c, axis=0)).max() < 0.01
    # across theta = 0 (wheel angle +-pi) the colour steps from entry 0 to entry 1
    # (5e-5 rad from +-pi, outside the 2.7e-6 rad where the reference reads past its table)
    lo = O.k_flow_hsv(np.array([[[2.0, -1e-4]]]), 1.0)[0, 0]   # wheel angle just below +pi
    hi = O.k_flow_hsv(np.array([[[2.0, 1e-4]]]), 1.0)[0, 0]    # just above -pi
    assert np.allclose(lo, 0.75 * O.wheel_entry(0)[::-1] / 255, atol=1e-3)
    assert np.allclose(hi, 0.75 * O.wheel_entry(1)[::-1] / 255, atol=1e-3)


@pytest.mark.gpu
def test_gpu_disparity_vis_matches_oracle():
    from vfs_amd import StereoTgv
    P = S.SOLVER_PARAMS
    pair = S.make_pair(160, 120, 51, radius=60, dmax=8.0)
    g = StereoTgv()
    g.initialize(160, 120, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], 3,
                 P["fScale"], 3, 10)
    g.copyMaskToDevice(pair["mask"])
    g.loadVectorFields(pair["tv"], pair["cv"])
    g.copyImagesToDevice(pair["i0"], pair["i1"])
    g.solveStereoForwardMasked()
    flow = g.copyDisparityToHost()
    for scale in (1.0, 4.0, 40.0):
        vis = g.copyDisparityVisToHost(scale)
        ref = O.k_flow_hsv(flow.astype(np.float64), scale)
        assert vis.shape == (120, 160, 3)
        assert np.abs(vis - ref).max() < 1e-4, (scale, np.abs(vis - ref).max())
