"""The dispatch model of vfs_amd.profile_ops (used to attribute rocprofv3 kernel traces to
the ops of a solve) against the library's greedy split of a warp's inner iterations."""
from vfs_amd.profile_ops import inner_ops, program_ops


def test_inner_ops_greedy_split():
    assert inner_ops(50, 1) == ["ITER"] * 50
    assert inner_ops(50, 2) == ["ITER2"] * 25
    assert inner_ops(50, 4) == ["ITER4"] * 12 + ["ITER2"]
    assert inner_ops(50, 8) == ["ITER8"] * 6 + ["ITER2"]
    assert inner_ops(15, 8) == ["ITER8", "ITER4", "ITER2", "ITER"]
    assert inner_ops(11, 4) == ["ITER4", "ITER4", "ITER2", "ITER"]
    assert inner_ops(1, 8) == ["ITER"]


def test_inner_ops_persistent_split():
    # an even count per persistent launch within the 128-step table, the rest one at a time
    assert inner_ops(50, "P") == ["ITERP"]
    assert inner_ops(15, "P") == ["ITERP", "ITER"]
    assert inner_ops(1, "P") == ["ITER"]
    assert inner_ops(130, "P") == ["ITERP", "ITER", "ITER"]
    assert inner_ops(131, "P") == ["ITERP"] + ["ITER"] * 3
    # 848x800 L5 W10 I50 with persistent launches at level 0 only: 10 ITERP dispatches
    ops = program_ops(5, 10, 50, ("P", 4, 8, 8, 8))
    assert sum(1 for k, l in ops if k == "ITERP") == 10
    assert all(l == 0 for k, l in ops if k == "ITERP")


def test_inner_ops_blocked_persistent_split():
    # an even count within the 128-step table (blocks, the last one shortened), the rest
    # split as with 8 per launch
    assert inner_ops(50, "P4") == ["ITERPK"]
    assert inner_ops(50, "P8") == ["ITERPK"]
    assert inner_ops(19, "P8") == ["ITERPK", "ITER"]
    assert inner_ops(3, "P4") == ["ITERPK", "ITER"]
    assert inner_ops(1, "P4") == ["ITER"]
    assert inner_ops(130, "P8") == ["ITERPK", "ITER2"]
    assert inner_ops(130, "P4:1") == ["ITERPK", "ITER", "ITER"]


def test_program_ops_count_matches_traced_solve():
    # a traced 848x800 solve (L5 W10 I50; four iterations per launch at levels 0-2,
    # eight at levels 3-4) issued 651 dispatches with a separate first preparation per
    # level and no energy records: 530 multi-iteration, 121 one-op; the tensor op now also
    # prepares the first warp (5 fewer), and a warp's first multi-iteration launch computes
    # its derivatives (50 fewer)
    ops = program_ops(5, 10, 50, (4, 4, 4, 8, 8), energy=0)
    assert len(ops) == 651 - 5 - 50
    multi = [o for o in ops if o[0] in ("ITER2", "ITER4", "ITER8")]
    assert len(multi) == 530
    assert len(program_ops(5, 10, 50, (4, 4, 4, 4, 4), energy=0)) == 771 - 5 - 50
    # the energy records: one op per level (default) or per warp
    assert len(program_ops(5, 10, 50, (4, 4, 4, 8, 8))) == 596 + 5
    assert len(program_ops(5, 10, 50, (4, 4, 4, 8, 8), energy=2)) == 596 + 50
    # one iteration per launch keeps the derivatives op
    lvl0 = [k for k, l in program_ops(2, 3, 5, (1, 1)) if l == 0][2:]  # after ZERO, CALIB
    assert lvl0[:2] == ["TENSOR", "DERIV"] and lvl0[-2:] == ["ENERGY", "END"]
    lvl0 = [k for k, l in program_ops(2, 3, 5, (4, 1)) if l == 0][2:]
    assert lvl0[:3] == ["TENSOR", "ITER4", "ITER"] and "DERIV" not in lvl0
