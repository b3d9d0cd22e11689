"""The persistent op's work-item order (vfs::persist_rank) on the host: every staged
position taken once, the tile's border (the positions whose strips neighbours read) first
and within the first position of every work-item, each group in raster order, for every
tile shape that has the op.  The program is built by __graft_entry__.build()."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "persist_order", "check_persist_order")


def test_persist_order():
    if not os.path.exists(BIN):
        pytest.skip("check_persist_order not built (run __graft_entry__.build())")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "8 shapes, 0 violations" in r.stdout, r.stdout
