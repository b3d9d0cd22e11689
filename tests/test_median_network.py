"""The device median's comparator network (kMedian25 in vfs_amd/csrc/tgv_device.h) against
the definition it stands for: the 13th smallest of 25 values, the value medianFilter.cu's
partial selection sort leaves at position 12.  A comparator network selects the k-th
smallest of every input if it does so for every 0/1 input (the 0-1 principle), so all
2^25 0/1 windows are checked, bit-parallel; random real windows (ties, negative values,
the window's corner zeros) are checked against sorting as well."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def network():
    src = open(os.path.join(ROOT, "vfs_amd", "csrc", "tgv_device.h")).read()
    body = re.search(r"kMedian25\[(\d+)\]\[2\]\s*=\s*\{(.*?)\};", src, re.S)
    pairs = [(int(a), int(b)) for a, b in re.findall(r"\{(\d+),\s*(\d+)\}", body.group(2))]
    assert len(pairs) == int(body.group(1))
    return pairs


def run(pairs, m):
    m = list(m)
    for a, b in pairs:
        x, y = m[a], m[b]
        m[a], m[b] = np.minimum(x, y), np.maximum(x, y)
    return m[12]


def test_network_selects_the_median_of_every_01_window():
    pairs = network()
    assert all(0 <= a < 25 and 0 <= b < 25 and a != b for a, b in pairs)
    vals = np.arange(1 << 25, dtype=np.uint32)
    wires = [((vals >> v) & 1).astype(np.uint8) for v in range(25)]
    ones = np.zeros(1 << 25, dtype=np.uint8)
    for w in wires:
        ones += w
    got = run(pairs, wires)
    # the 13th smallest of a 0/1 window is 1 exactly when at most 12 values are 0
    assert np.array_equal(got, (ones >= 13).astype(np.uint8))


def test_network_matches_sorting_on_real_windows():
    pairs = network()
    rng = np.random.default_rng(5)
    w = rng.normal(size=(25, 20000)).astype(np.float32)
    w[:, :5000] = np.round(w[:, :5000])          # many ties
    w[[0, 4, 20, 24], 5000:10000] = 0.0          # the corner zeros of an edge window
    got = run(pairs, [w[i] for i in range(25)])
    assert np.array_equal(got, np.sort(w, axis=0)[12])
