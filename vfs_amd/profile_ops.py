"""Per-op breakdown of a solve from a rocprofv3 kernel trace.

Every op of a solve program is one `k_tick` dispatch when the solve is launched without
a graph (option "graph" = 0), in program order.  This module rebuilds the cold masked
program's op list (the same sequence the library issues: zero, calibration warp,
pyramid, then per level the tensor (with the first warp's preparation) and, per warp,
derivatives / inner iterations / end, each end but the level's last fused with the next
warp's prep, the energy partials (ENERGY) before the level's last end (whose extra
workgroup adds them up), then the upscale; a warp whose first inner op is a multi-iteration
or persistent launch that stages at most two positions per work-item computes its
derivatives there and has no DERIV op: `--fold` lists the levels where that holds)
and attributes each dispatch of one solve to its op and level.
Levels that run several inner iterations per launch (`--per-launch`, one count per level
from level 0, as the tuner reports them) issue `k_iterk` dispatches: ITER8 / ITER4 / ITER2 while
they fit the iteration count, the rest one at a time; `P` marks a level whose warps run
their iterations in persistent `k_iterp` launches (ITERP: an even count, at most 128, each;
an odd remainder one at a time), `P2` / `P4` / `P8` one in blocked persistent `k_iterkp`
launches (ITERPK: the same even counts, in blocks, the rest as `--per-launch` 8 would split
it, or as `:n` after it says, e.g. `P4:1`).

    python -m vfs_amd.profile_ops TRACE.csv --levels 5 --warps 10 --iters 50 --per-launch 4,4,4,4,4 --skip-tail 100
    python -m vfs_amd.profile_ops --batch-traffic FETCH.csv WRITE.csv --storage f16   # HBM bytes per px-iteration

`--skip-tail` drops trailing dispatches that follow the solves (e.g. timing launches);
the solve analysed is the last complete one before them.
"""
from __future__ import annotations

import argparse
import collections
import csv


def inner_ops(iters: int, per_launch) -> list[str]:
    """Inner-iteration ops of one warp (the library's greedy split; per_launch "P": the
    persistent op, "P2" / "P4" / "P8" the blocked persistent op)."""
    out, it = [], 0
    if isinstance(per_launch, str) and per_launch[:2] in ("P2", "P4", "P8"):
        # "P4:1": the level's own iterations per launch (for the remainder) is 1
        rest = int(per_launch.split(":")[1]) if ":" in per_launch else 8
        k = min(iters, 128) & ~1
        if k >= 2:
            out.append("ITERPK")
            it = k
        return out + inner_ops(iters - it, rest)
    if per_launch == "P":
        while iters - it >= 2:
            k = min(iters - it, 128 - it) & ~1
            if k < 2:
                break
            out.append("ITERP")
            it += k
        return out + ["ITER"] * (iters - it)
    while it < iters:
        k = next((c for c in (8, 4, 2) if c <= per_launch and it + c <= iters), 1)
        out.append({8: "ITER8", 4: "ITER4", 2: "ITER2", 1: "ITER"}[k])
        it += k
    return out


def program_ops(levels: int, warps: int, iters: int, per_launch=(), energy: int = 1, fold=None) -> list[tuple[str, int]]:
    """Ops of the cold program; `energy` as the solver option (0 none, 1 after each level's
    last warp, 2 after every warp); `fold`: the levels whose warps' first inner op computes
    the derivatives (default: every level whose first inner op is not a one-iteration op)."""
    ops = [("ZERO", 0), ("CALIB", 0)] + [("DOWN", l) for l in range(1, levels)] + [("CINIT", levels - 1)]
    for l in range(levels - 1, -1, -1):
        ops.append(("TENSOR", l))
        k = per_launch[l] if l < len(per_launch) else 1
        inner = [(name, l) for name in inner_ops(iters, k)]
        for wi in range(warps):
            folded = bool(inner) and inner[0][0] != "ITER" and (fold is None or l in fold)
            ops += ([] if folded else [("DERIV", l)]) + inner
            if energy == 2 or (energy == 1 and wi + 1 == warps):
                ops.append(("ENERGY", l))
            ops.append(("ENDPREP" if wi + 1 < warps else "END", l))
        if l > 0:
            ops.append(("UP", l))
    return ops


def summarize(trace_csv: str, levels: int, warps: int, iters: int, skip_tail: int = 0, per_launch=(),
              energy: int = 1, fold=None) -> str:
    rows = [r for r in csv.DictReader(open(trace_csv))
            if r["Kernel_Name"].startswith(("k_tick", "k_iterk", "k_iterp"))]  # k_iterk: also k_iterkp
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    seq = program_ops(levels, warps, iters, per_launch, energy, fold)
    n = len(seq)
    end = len(rows) - skip_tail
    if end < n:
        raise ValueError(f"trace holds {end} tick dispatches, one solve needs {n}")
    solve = rows[end - n:end]
    for (kind, lvl), r in zip(seq, solve):
        name = r["Kernel_Name"]
        if ((kind in ("ITER2", "ITER4", "ITER8")) != (name.startswith("k_iterk") and not name.startswith("k_iterkp"))
                or (kind == "ITERP") != name.startswith("k_iterp")
                or (kind == "ITERPK") != name.startswith("k_iterkp")):
            raise ValueError(f"dispatch {r['Kernel_Name']} does not match op {kind} at level {lvl}: "
                             "check --per-launch")
    agg = collections.OrderedDict()
    for (kind, lvl), r in zip(seq, solve):
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3
        wgs = (int(r["Grid_Size_X"]) // max(int(r["Workgroup_Size_X"]), 1)) * \
              (int(r["Grid_Size_Y"]) // max(int(r["Workgroup_Size_Y"]), 1))
        a = agg.setdefault((kind, lvl), [0, 0.0, wgs])
        a[0] += 1
        a[1] += d
    total = sum(v[1] for v in agg.values())
    span = (int(solve[-1]["End_Timestamp"]) - int(solve[0]["Start_Timestamp"])) * 1e-3
    out = [f"one solve: {n} dispatches, kernel time {total:.0f} us, first start to last end {span:.0f} us", "",
           "| op | level | dispatches | workgroups | total us | avg us | share |",
           "|---|---|---|---|---|---|---|"]
    for (kind, lvl), (c, s, g) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        out.append(f"| {kind} | {lvl} | {c} | {g} | {s:.1f} | {s / c:.2f} | {100 * s / total:.1f}% |")
    by_level = collections.defaultdict(float)
    for (kind, lvl), (c, s, g) in agg.items():
        by_level[lvl] += s
    out += ["", "per level: " + ", ".join(f"L{l} {v:.0f} us" for l, v in sorted(by_level.items()))]
    if skip_tail:
        tail = rows[end:]
        d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3 for r in tail]
        names = sorted({r["Kernel_Name"] for r in tail})
        out += ["", f"the {len(tail)} dispatches after the solves ({', '.join(names)}): "
                    f"average {sum(d) / len(d):.2f} us, min {min(d):.2f}, max {max(d):.2f}"]
    return "\n".join(out)


def _counter_rows(path: str, counter: str):
    """(dispatch id, kernel name, grid size, value) of one counter from a rocprofv3
    counter_collection.csv."""
    out = []
    for r in csv.DictReader(open(path)):
        name = r.get("Counter_Name", counter)
        if name != counter:
            continue
        out.append((int(r["Dispatch_Id"]), r["Kernel_Name"], int(r["Grid_Size"]), float(r["Counter_Value"])))
    return out


def _iterk_k(name: str) -> int:
    """K of a k_iterk<T, shape, K, half> kernel name (0 for other kernels)."""
    if "k_iterk<" not in name:
        return 0
    args = name.split("k_iterk<", 1)[1].split(">", 1)[0].split(",")
    return int(args[2])


_TYPE = {"f32": "float", "f16": "__half", "bf16": "__hip_bfloat16"}


def batch_traffic(fetch_csv: str, write_csv: str, storage: str, pairs: int, px: int, copy_bytes: int,
                  command: str, k: int = 4) -> str:
    """Summary of the HBM bytes per pixel and inner iteration of a batch's level-0
    multi-iteration launches: the k_iterk dispatches of the storage type's instance with
    `k` iterations and the largest grid, FETCH_SIZE and WRITE_SIZE per dispatch (KiB)
    scaled by what the same counters read for the k_copy dispatches that move
    `copy_bytes` each way."""
    import statistics as st
    rows = {c: _counter_rows(p, c) for c, p in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv))}
    med = {}
    scale = {}
    sel = None
    for c, rs in rows.items():
        copies = [v for _, n, _, v in rs if "k_copy" in n]
        scale[c] = copy_bytes / 1024.0 / st.median(copies)
        iterk = [(n, g, v) for _, n, g, v in rs if _iterk_k(n) > 0 and f"k_iterk<{_TYPE[storage]}," in n]
        kmax = k
        gmax = max(g for n, g, _ in iterk if _iterk_k(n) == kmax)
        vals = [v for n, g, v in iterk if _iterk_k(n) == kmax and g == gmax]
        sel = (kmax, gmax, len(vals), sorted({n for n, g, _ in iterk if _iterk_k(n) == kmax and g == gmax}))
        med[c] = (st.median(vals), min(vals), max(vals))
    k, grid, n, names = sel
    kib = med["FETCH_SIZE"][0] * scale["FETCH_SIZE"] + med["WRITE_SIZE"][0] * scale["WRITE_SIZE"]
    per = kib * 1024.0 / (pairs * px * k)
    lines = [f"# HBM traffic of the level-0 inner iterations of a {pairs}-pair batch at 848x800, {storage} storage", "",
             "Command (two passes, one counter each; from the repository root on the GPU machine):", "",
             "    " + command, "",
             f"Dispatches analysed: {n} `k_iterk` launches of {k} iterations with grid {grid} work-items "
             f"({', '.join(names)}); each covers {pairs} pairs of {px} pixels.", "",
             "| counter | scale (k_copy, %d MiB each way) | median per dispatch (KiB) | min / max |" % (copy_bytes >> 20),
             "|---|---|---|---|"]
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        m, lo, hi = med[c]
        lines.append(f"| {c} | x {scale[c]:.3f} | {m:.1f} | {lo:.1f} / {hi:.1f} |")
    lines += ["", f"HBM bytes per launch = {kib * 1024:.0f}; per pixel and iteration:", "",
              f"traffic_bytes_per_px_iter = {per:.2f}", ""]
    return "\n".join(lines)


def kernel_traffic(fetch_csv: str, write_csv: str, kernel: str, copy_bytes: int, command: str, title: str,
                   algorithmic: int) -> str:
    """Summary of the HBM bytes per dispatch of one kernel (name prefix `kernel`, the
    dispatches with the largest grid): FETCH_SIZE and WRITE_SIZE per dispatch (KiB), median
    over the dispatches, scaled by what the same counters read for the k_copy dispatches
    that move `copy_bytes` each way; against `algorithmic` bytes per dispatch."""
    import statistics as st
    rows = {c: _counter_rows(p, c) for c, p in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv))}
    med, scale, sel = {}, {}, None
    for c, rs in rows.items():
        copies = [v for _, n, _, v in rs if "k_copy" in n]
        scale[c] = copy_bytes / 1024.0 / st.median(copies)
        ks = [(n, g, v) for _, n, g, v in rs if n.startswith(kernel)]
        gmax = max(g for _, g, _ in ks)
        vals = [v for _, g, v in ks if g == gmax]
        sel = (gmax, len(vals), sorted({n for n, g, _ in ks if g == gmax}))
        med[c] = (st.median(copies), st.median(vals), min(vals), max(vals))
    grid, n, names = sel
    kib = med["FETCH_SIZE"][1] * scale["FETCH_SIZE"] + med["WRITE_SIZE"][1] * scale["WRITE_SIZE"]
    lines = [f"# {title}", "", "Commands (two passes, one counter each; from the repository root on the GPU machine):",
             "", "    " + command, "",
             f"Dispatches analysed: {n} `{kernel}` launches with grid {grid} work-items ({', '.join(names)}).", "",
             f"| counter | k_copy per dispatch (KiB) | scale to bytes moved | {kernel} median per dispatch (KiB) | "
             "min / max |", "|---|---|---|---|---|"]
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        cm, m, lo, hi = med[c]
        lines.append(f"| {c} | {cm:.1f} | x {scale[c]:.3f} ({copy_bytes >> 20} MiB each way) | {m:.2f} | "
                     f"{lo:.1f} / {hi:.1f} |")
    lines += ["", f"HBM bytes per launch = (FETCH x scale + WRITE x scale):", "",
              f"traffic_bytes_per_launch = {kib * 1024:.0f}", "",
              f"Algorithmic bytes per launch: {algorithmic:,}; measured / algorithmic = "
              f"{kib * 1024 / algorithmic:.4f}.", ""]
    return "\n".join(lines)


def trace_segments(trace_csv: str, gap_ms: float = 20.0, after_last: str = "k_copy") -> str:
    """Segments of a kernel trace: the k_tick / k_iterk dispatches after the last
    `after_last` dispatch, split where the device idles longer than `gap_ms`; per
    segment the dispatch count, span, busy share (kernel time / span), idle gaps between
    dispatches and the average duration of each kernel.  For the served leg the segments
    are the phases of the bench (stream set-up, cold first frames, timed frames)."""
    import statistics as st
    rows = list(csv.DictReader(open(trace_csv)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    last = max([i for i, r in enumerate(rows) if after_last in r["Kernel_Name"]], default=-1)
    sub = [r for r in rows[last + 1:] if r["Kernel_Name"].startswith(("k_tick", "k_iterk")) or
           "k_tick" in r["Kernel_Name"] or "k_iterk" in r["Kernel_Name"]]
    segs, cur = [], [sub[0]]
    for a, b in zip(sub, sub[1:]):
        if int(b["Start_Timestamp"]) - int(a["End_Timestamp"]) > gap_ms * 1e6:
            segs.append(cur)
            cur = []
        cur.append(b)
    segs.append(cur)
    out = ["| segment | dispatches | span ms | busy | idle gap median / p90 / mean us | k_tick avg us | "
           "k_iterk avg us |", "|---|---|---|---|---|---|---|"]
    for i, sg in enumerate(segs):
        s0 = [int(r["Start_Timestamp"]) for r in sg]
        e0 = [int(r["End_Timestamp"]) for r in sg]
        busy = sum(e - s for s, e in zip(s0, e0))
        span = max(e0[-1] - s0[0], 1)
        gaps = sorted([s0[j + 1] - e0[j] for j in range(len(sg) - 1)] or [0])
        dur = lambda k: [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in sg if k in r["Kernel_Name"]]
        tk, ik = dur("k_tick"), dur("k_iterk")
        out.append(f"| {i} | {len(sg)} | {span / 1e6:.1f} | {100 * busy / span:.0f}% | "
                   f"{st.median(gaps) / 1e3:.1f} / {gaps[int(0.9 * (len(gaps) - 1))] / 1e3:.1f} / "
                   f"{st.mean(gaps) / 1e3:.1f} | {st.mean(tk) / 1e3 if tk else 0:.1f} | "
                   f"{st.mean(ik) / 1e3 if ik else 0:.1f} |")
    return "\n".join(out)


def main():
    import sys
    if len(sys.argv) > 1 and sys.argv[1] == "--segments":
        print(trace_segments(sys.argv[2]))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--kernel-traffic":
        kp = argparse.ArgumentParser(prog="profile_ops --kernel-traffic")
        kp.add_argument("fetch_csv")
        kp.add_argument("write_csv")
        kp.add_argument("--kernel", required=True)
        kp.add_argument("--copy-bytes", type=int, default=256 << 20)
        kp.add_argument("--command", default="")
        kp.add_argument("--title", default="")
        kp.add_argument("--algorithmic", type=int, required=True)
        a = kp.parse_args(sys.argv[2:])
        print(kernel_traffic(a.fetch_csv, a.write_csv, a.kernel, a.copy_bytes, a.command, a.title, a.algorithmic))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--batch-traffic":
        bp = argparse.ArgumentParser(prog="profile_ops --batch-traffic")
        bp.add_argument("fetch_csv")
        bp.add_argument("write_csv")
        bp.add_argument("--storage", required=True)
        bp.add_argument("--pairs", type=int, default=16)
        bp.add_argument("--px", type=int, default=848 * 800)
        bp.add_argument("--copy-bytes", type=int, default=512 << 20)
        bp.add_argument("--command", default="")
        bp.add_argument("--k", type=int, default=4, help="iterations per launch of the level-0 launches")
        b = bp.parse_args(sys.argv[2:])
        print(batch_traffic(b.fetch_csv, b.write_csv, b.storage, b.pairs, b.px, b.copy_bytes, b.command, b.k))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--warps", type=int, default=10)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--skip-tail", type=int, default=0)
    ap.add_argument("--per-launch", default="",
                    help="inner iterations per launch of each level, from level 0 (P: persistent launches, "
                         "P2 / P4 / P8: blocked persistent launches)")
    ap.add_argument("--energy", type=int, default=1, help="the solver's energy option (0, 1, 2)")
    ap.add_argument("--fold", default=None,
                    help="levels whose warps compute the derivatives in their first inner launch, e.g. 1,2,3,4 "
                         "(default: every level whose first inner launch runs several iterations)")
    a = ap.parse_args()
    per_launch = [x if x.startswith("P") else int(x) for x in a.per_launch.split(",") if x]
    fold = None if a.fold is None else {int(x) for x in a.fold.split(",") if x}
    print(summarize(a.trace, a.levels, a.warps, a.iters, a.skip_tail, per_launch, a.energy, fold))


if __name__ == "__main__":
    main()
