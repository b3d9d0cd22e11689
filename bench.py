"""Benchmark of the StereoTgv solve on one MI355X.

A step is one masked solve (solveStereoForwardMasked) of one synthetic 848x800 pair
(5 levels, 10 warps, 50 iterations; inputs resident in HBM).  Prints one JSON line:
frame-pairs/s, p50/p99 per-solve latency, mean disparity error against the FP64 oracle,
the roofline of the dominant kernel (the inner TGV iterations at level 0, several per
launch as the tuner chose) and the CPU baseline (the FP64 oracle on the host cores).

    python bench.py [--steps K] [--warmup W] [--workload tgv_848x800] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per pixel and inner iteration (FP32 storage): what one iteration
# moves when each iteration is its own pass: read u, u_, v_(2), p(2), q(4), a, b, c, Iu,
# Iz and write u, u_, v_(2), p(2), q(4) = 25 floats.  A launch of K iterations is
# credited K times that.
B_ITER_FP32 = 100.0
# HBM bytes per launch of the level-0 inner-iteration kernel measured with rocprofv3
# --pmc (FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel of known size), one summary
# per launch choice (kernel: k_iterk or the persistent k_iterp, tile shape, iterations
# per launch); each holds its command and counters.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "pmc_traffic_{kernel}_{tw}x{th}x{k}_{workload}_L0.md")
PERSISTENT = 0x200  # per_launch of a level whose warps run in persistent launches


def profiled_traffic(kernel, tw, th, k, workload):
    """`traffic_bytes_per_launch = N` from the committed PMC summary of this launch choice,
    or None when there is none."""
    import re
    try:
        with open(TRAFFIC_PROFILE.format(kernel=kernel, tw=tw, th=th, k=k, workload=workload)) as f:
            m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9]+)", f.read())
        return int(m.group(1)) if m else None
    except OSError:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="tgv_848x800")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline leg")
    ap.add_argument("--iter-reps", type=int, default=200)
    ap.add_argument("--no-graph", action="store_true", help="launch kernels one by one instead of a HIP graph")
    ap.add_argument("--quick", action="store_true", help="plain form only (no served / batched / reduced-storage legs)")
    ap.add_argument("--legs", default="served,batched,reduced",
                    help="legs beside the plain form, comma-separated: served, batched, reduced")
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--serve-frames", type=int, default=100,
                    help="frames per stream in the serving leg (the first is the stream's cold solve)")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving leg (development aid)")
    ap.add_argument("--no-serve-check", action="store_true",
                    help="skip the FP64 oracle check of the served frames")
    ap.add_argument("--reduced-storage", default="f32,f16,bf16",
                    help="storage types of the reduced-storage leg, comma-separated")
    ap.add_argument("--batch-tuning", default="",
                    help="per-level launch choices of the batched legs' batches, as --tuning (profiling aid)")
    ap.add_argument("--batch-no-autotune", action="store_true",
                    help="no launch tuner for the batched legs' batches (with --batch-tuning for reproducible profiles)")
    ap.add_argument("--shape", type=int, default=-1,
                    help="force one tile shape (0..6) instead of the per-level tuner (profiling aid)")
    ap.add_argument("--tuning", default="",
                    help="per-level launch choices from level 0, comma-separated (profiling aid): 'S[p]' one "
                         "iteration per launch in tile shape S (p = Tp prefetch), 'MxK[h]' K = 2, 4 or 8 iterations "
                         "per launch in tile shape M (h = half workgroups), 'MxP' a warp per persistent launch in "
                         "tile shape M, 'MxPK' the same in blocks of K = 2, 4 or 8 iterations, or 'S[p]/MxK[h]' "
                         "for both")
    ap.add_argument("--no-autotune", action="store_true",
                    help="no launch tuner (fixed defaults; with --tuning for reproducible profiles)")
    ap.add_argument("--iter-only", type=int, default=0,
                    help="profiling aid: run N level-0 inner iterations (after a 256 MiB reference copy) and exit")
    ap.add_argument("--iter-level", type=int, default=0,
                    help="profiling aid: the pyramid level whose inner iterations --iter-only runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); the solver shards by "
              f"frame pair / stream with no cross-GPU exchange, so N GPUs run N independent copies.",
              file=sys.stderr)
        sys.exit(2)

    import vfs_amd
    from vfs_amd import synth as S

    wl = S.WORKLOADS[args.workload]
    P = S.SOLVER_PARAMS
    W, H = wl["W"], wl["H"]
    print(f"bench: workload {args.workload} {W}x{H} L{wl['levels']} W{wl['warps']} I{wl['iters']}", file=sys.stderr,
          flush=True)
    pair = S.make_pair(W, H, wl["seed"], wl["radius"], wl["dmax"])

    g = vfs_amd.StereoTgv()
    g.initialize(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"],
                 wl["levels"], P["fScale"], wl["warps"], wl["iters"])
    g.limitRange = P["limitRange"]
    if args.no_graph:
        g.set_option("graph", 0)
    if args.shape >= 0:
        g.set_option("shape", args.shape)
    g.copyMaskToDevice(pair["mask"])
    g.loadVectorFields(pair["tv"], pair["cv"])
    g.copyImagesToDevice(pair["i0"], pair["i1"])
    if args.no_autotune:
        g.set_option("autotune", 0)
    apply_tuning(g, args.tuning)

    if args.iter_only:
        # A plain device copy of known size (256 MiB read + 256 MiB written per dispatch)
        # calibrates the HBM counters (FETCH_SIZE / WRITE_SIZE), then only the inner
        # iterations of one level run.
        g.time_copy(256 << 20, 3)
        ms, nl = g.time_inner_iterations(args.iter_level, args.iter_only)
        print(f"iter-only: {args.iter_only} level-{args.iter_level} iterations, {ms / args.iter_only * 1e3:.2f} us each "
              f"({nl} iteration(s) per launch)", flush=True)
        return

    for _ in range(args.warmup):
        g.time_solves(True, 1)
    g.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lat.append(g.time_solves(True, 1))
    g.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array(lat)
    pairs_per_s = args.steps / wall
    # Re-solving without new images re-calibrates I1 (the reference's pI1[0] swap), so
    # the disparity error is measured on a solve of freshly copied images.
    g.copyImagesToDevice(pair["i0"], pair["i1"])
    g.solveStereoForwardMasked()
    u_gpu = g.copy1DDisparityToHost()
    print(f"bench: {args.steps} solves in {wall * 1e3:.1f} ms; per-solve p50 {np.percentile(lat, 50):.3f} ms",
          file=sys.stderr, flush=True)
    # the energy / residual monitor (option "energy" 1: one FP64 record per level) against
    # the same solve without it
    g.set_option("energy", 0)
    g.time_solves(True, 2)
    off = [g.time_solves(True, 1) for _ in range(args.steps)]
    g.set_option("energy", 1)
    on = [g.time_solves(True, 1) for _ in range(args.steps)]
    energy_monitor = {"ms_per_solve_on": round(float(np.median(on)), 4),
                             "ms_per_solve_off": round(float(np.median(off)), 4),
                             "overhead_pct": round(100.0 * (np.median(on) / np.median(off) - 1.0), 2),
                             "records": "one per level (after its last warp), FP64 sums on the GPU"}
    print(f"bench: energy monitor {energy_monitor}", file=sys.stderr, flush=True)

    # the drop-in call sequence with host buffers: copy both images in (pageable host
    # memory over PCIe), solve, copy the disparity out
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.copyImagesToDevice(pair["i0"], pair["i1"])
        g.solveStereoForwardMasked()
        g.copy1DDisparityToHost()
    pcie_pairs_per_s = args.steps / (time.perf_counter() - t0)

    # roofline of the dominant kernel: the inner TGV iterations at level 0, timed with HIP
    # events on the solver's stream (launches back to back, as the tuner chose them)
    w0, h0, _ = g.level_dims(0)
    ms_it, per_launch = g.time_inner_iterations(0, args.iter_reps)
    t_iter = ms_it / args.iter_reps * 1e-3
    t_launch = t_iter * per_launch
    bytes_launch = B_ITER_FP32 * w0 * h0 * per_launch
    achieved = bytes_launch / t_launch / 1e9
    (tw0, th0), _, k0 = g.level_tuning(0, True, with_per_launch=True)
    kernel0 = ("k_iterp" if k0 == PERSISTENT else "k_iterkp" if k0 & PERSISTENT else "k_iterk" if per_launch > 1
               else "k_tick_iter")

    # achievable-bandwidth reference: a plain copy kernel, 1 GiB moved per launch
    copy_ms = g.time_copy(512 << 20, 10) / 10
    copy_gbs = 2 * (512 << 20) / (copy_ms * 1e-3) / 1e9

    out = {
        "metric": "frame-pairs/sec + p99 latency @848x800, mean disp err <=0.05px vs FP64",
        "value": round(pairs_per_s, 3),
        "unit": "frame-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: masked solve, {W}x{H}, {wl['levels']} levels, "
                               f"{wl['warps']} warps, {wl['iters']} iterations, one pair per step, FP32",
                   "global_batch": 1},
        "pcie_inclusive_frame_pairs_per_s": round(pcie_pairs_per_s, 3),
        "energy_monitor": energy_monitor,
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                       "p99": round(float(np.percentile(lat, 99)), 4),
                       "mean": round(float(lat.mean()), 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(kernel0, tw0, th0, per_launch, args.workload),
                     "kernel": f"inner TGV iterations at level 0 ({w0}x{h0}): {per_launch} per launch, "
                               f"{tw0}x{th0} tiles" + (" (persistent workgroups: the tiles stay on chip, "
                                                       "borders exchanged every iteration)"
                                                       if k0 == PERSISTENT else
                                                       f" (persistent workgroups, blocks of {k0 & 0xff} iterations)"
                                                       if k0 & PERSISTENT else ""),
                     "us_per_launch": round(t_launch * 1e6, 3), "us_per_iter": round(t_iter * 1e6, 3),
                     "algorithmic_bytes_per_launch": int(bytes_launch),
                     "copy_kernel_GBs": round(copy_gbs, 1)},
    }
    # the whole solve against the same roofline: algorithmic bytes of one pair (valid
    # pixels, one pass per kernel, no credit for reuse), B_pair = sum_l N_l (W (I B_iter
    # + B_warp) + B_level) + 16 N_0 + 40 (sum_l N_l - N_0), B_warp = 104, B_level = 44
    n_l = [g.level_dims(l)[0] * g.level_dims(l)[1] for l in range(wl["levels"])]
    b_pair = (sum(n_l) * (wl["warps"] * (wl["iters"] * B_ITER_FP32 + 104.0) + 44.0) + 16.0 * n_l[0]
              + 40.0 * (sum(n_l) - n_l[0]))
    out["solve_roofline"] = {"algorithmic_bytes_per_pair": int(b_pair),
                             "achieved": round(b_pair * pairs_per_s / 1e9, 1), "unit": "GB/s",
                             "frac": round(b_pair * pairs_per_s / 1e9 / HBM_PEAK_GBS, 4),
                             "ceiling_pairs_per_s": round(HBM_PEAK_GBS * 1e9 / b_pair, 1)}

    # per-level launch choices of the tile-shape tuner (made at the first solve)
    out["tile_tuning"] = {}
    for lvl in range(wl["levels"]):
        (tw, th), pre, k = g.level_tuning(lvl, True, with_per_launch=True)
        lw, lh, _ = g.level_dims(lvl)
        out["tile_tuning"][f"L{lvl} {lw}x{lh}"] = (
            f"{tw}x{th}" + (" +Tp prefetch" if pre else "")
            + (", a warp per launch (persistent workgroups)" if k == PERSISTENT
               else f", a warp per launch in blocks of {k & 0xff} iterations (persistent workgroups)"
               if k & PERSISTENT else f", {k} iterations per launch" if k > 1 else ""))
    print(f"bench: tile tuning {out['tile_tuning']}", file=sys.stderr, flush=True)

    legs = set() if args.quick else ({"served"} if args.serve_only else {l for l in args.legs.split(",") if l})
    if "served" in legs:
        out["served"] = measure_served(args, P, S, vfs_amd, pairs_per_s, W * H)
    if "batched" in legs:
        out["batched"] = measure_batched(args, P, S, vfs_amd)
    if "reduced" in legs:
        out["reduced_storage"] = measure_reduced(args, P, S, vfs_amd, pair_checker=not args.no_cpu)

    if not args.no_cpu:
        base, u_ref = cpu_baseline(wl, pair, P, S)
        m = pair["mask"] > 0
        err = np.abs(u_gpu - u_ref)[m]
        out["mean_disp_err_px"] = float(f"{err.mean():.3e}")
        out["max_disp_err_px"] = float(f"{err.max():.3e}")
        out["cpu_baseline"] = base
    print(json.dumps(out), flush=True)


def apply_tuning(obj, spec):
    """Per-level launch choices from level 0 (see --tuning) on a solver or a batch."""
    for lvl, lspec in enumerate(c for c in spec.split(",") if c):
        for part in lspec.split("/"):
            if "x" in part:
                m, k = part.split("x")
                if k.startswith("P"):
                    obj.set_level_tuning(lvl, int(m), per_launch=PERSISTENT | int(k[1:] or 0))
                else:
                    obj.set_level_tuning(lvl, int(m), per_launch=int(k.rstrip("h")), half_workgroup=k.endswith("h"))
            else:
                obj.set_level_tuning(lvl, int(part.rstrip("p")), preload=part.endswith("p"))


def _oracle_solve_times(pair, wl, P, threads, runs, warm):
    """Median wall time of `runs` masked solves (after `warm`) by the FP64 oracle with
    `threads` OpenMP threads; inputs resident (copied in before each timed solve, so
    every solve is a first solve of fresh images).  Returns (seconds, oracle)."""
    from oracle import oracle as O
    O.set_threads(threads)
    H, W = pair["i0"].shape
    o = O.Oracle(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"],
                 wl["levels"], P["fScale"], wl["warps"], wl["iters"], "f64")
    o.set_limit_range(P["limitRange"])
    o.copyMaskToDevice(pair["mask"])
    o.loadVectorFields(pair["tv"], pair["cv"])
    t = []
    for _ in range(warm + runs):
        o.copyImagesToDevice(pair["i0"], pair["i1"])
        t0 = time.perf_counter()
        o.solveStereoForwardMasked()
        t.append(time.perf_counter() - t0)
    return float(np.median(t[warm:])), o


def cpu_baseline(wl, pair, P, S):
    """The FP64 oracle (OpenMP C restatement of the reference kernels) on the host:
    the measured workload with every host thread (median of 3 after 1 warm-up) and
    with 1 thread (one solve), plus the 424x400 workload both ways (median of 3)."""
    from oracle import oracle as O
    threads = O.get_threads()
    t_all, o = _oracle_solve_times(pair, wl, P, threads, 3, 1)
    u_ref = o.copy1DDisparityToHost()
    t_one, _ = _oracle_solve_times(pair, wl, P, 1, 1, 0)
    w1 = S.WORKLOADS["tgv_424x400"]
    p1 = S.make_pair(w1["W"], w1["H"], w1["seed"], w1["radius"], w1["dmax"])
    c1_all, _ = _oracle_solve_times(p1, w1, P, threads, 3, 1)
    c1_one, _ = _oracle_solve_times(p1, w1, P, 1, 3, 1)
    O.set_threads(threads)
    H, W = pair["i0"].shape
    base = {"value": round(1.0 / t_all, 4), "unit": "frame-pairs/s", "cores": threads, "kind": "port",
            "sample": f"full {W}x{H} L{wl['levels']} W{wl['warps']} I{wl['iters']} masked solves of the measured pair "
                      f"by the FP64 C oracle, inputs resident, {threads} OpenMP threads: median of 3 after 1 warm-up "
                      f"({t_all:.2f} s per solve)",
            "single_thread": {"value": round(1.0 / t_one, 4), "cores": 1, "sample": f"one solve, {t_one:.2f} s"},
            "tgv_424x400": {"threads": threads, "frame_pairs_per_s": round(1.0 / c1_all, 3),
                            "frame_pairs_per_s_1_thread": round(1.0 / c1_one, 3),
                            "sample": "424x400 L4 W5 I20 masked solves, median of 3 after 1 warm-up"}}
    print(f"bench: cpu baseline {base}", file=sys.stderr, flush=True)
    return base, u_ref


def _solver(vfs_amd, P, pair, levels, warps, iters, storage="f32", graph=True):
    H, W = pair["i0"].shape
    g = vfs_amd.StereoTgv()
    g.set_option("storage", storage)
    g.set_option("graph", 1 if graph else 0)
    g.initialize(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], levels,
                 P["fScale"], warps, iters)
    g.limitRange = P["limitRange"]
    g.copyMaskToDevice(pair["mask"])
    g.loadVectorFields(pair["tv"], pair["cv"])
    g.copyImagesToDevice(pair["i0"], pair["i1"])
    return g


def _time_batch(args, b, n_pairs):
    b.time_solves(True, 1)  # warm-up (graph capture)
    b.synchronize()
    t0 = time.perf_counter()
    ms = b.time_solves(True, args.batch_steps)
    b.synchronize()
    wall = time.perf_counter() - t0
    return {"pairs_per_s": round(n_pairs * args.batch_steps / wall, 2),
            "ms_per_batch_solve": round(ms / args.batch_steps, 2)}


def measure_batched(args, P, S, vfs_amd):
    """64 pairs of uneven sizes (424x400, 640x480, 848x800; seeds 100+i), all solved by
    one batch: every kernel launch of the L5 W10 I50 schedule covers all 64 pairs."""
    sizes = [(424, 400), (640, 480), (848, 800)]
    pairs, solvers = [], []
    for i in range(64):
        W, H = sizes[i % 3]
        pair = S.make_pair(W, H, 100 + i, H / 2.0, W / 21.2)
        solvers.append(_solver(vfs_amd, P, pair, 5, 10, 50))
        pairs.append(pair)
    print("bench: batched 64 mixed pairs ready", file=sys.stderr, flush=True)
    if args.batch_no_autotune:
        solvers[0].set_option("autotune", 0)
    b = vfs_amd.Batch(solvers)
    apply_tuning(b, args.batch_tuning)
    res = _time_batch(args, b, 64)
    res["tile_tuning"] = {}
    for l in range(5):
        (tw, th), pre, k = b.level_tuning(l, with_per_launch=True)
        res["tile_tuning"][f"L{l}"] = (f"{tw}x{th}" + (" +Tp prefetch" if pre else "")
                                       + (f", {k} iterations per launch" if k > 1 else ""))
    mpx = sum(p["i0"].size for p in pairs) / 1e6
    res["Mpx_per_s"] = round(res["pairs_per_s"] * mpx / 64, 2)
    res["workload"] = "64 pairs, sizes 424x400 / 640x480 / 848x800 in turn, L5 W10 I50, FP32, one batch"
    b.close()
    for g in solvers:
        g.close()
    return res


# HBM bytes per pixel and inner iteration of the 16-pair batch's level-0 multi-iteration
# launches, from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (one summary per storage type,
# written by vfs_amd.profile_ops --batch-traffic).
BATCH_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "pmc_traffic_batch16_848x800_{st}_L0.md")


def profiled_batch_traffic(st):
    import re
    try:
        with open(BATCH_TRAFFIC_PROFILE.format(st=st)) as f:
            m = re.search(r"traffic_bytes_per_px_iter\s*=\s*([0-9.]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def measure_reduced(args, P, S, vfs_amd, pair_checker=True):
    """16 pairs at 848x800 (seeds 200+i) solved as one batch with FP32, FP16 and BF16
    storage; disparity error of pair 0 against the FP64 oracle per storage type."""
    pairs = [S.make_pair(848, 800, 200 + i, 400.0, 40.0) for i in range(16)]
    ref = None
    if pair_checker:
        from oracle import oracle as O
        ref = O.solve_pair(pairs[0], 5, 10, 50, P, precision="f64").copy1DDisparityToHost()
    m = pairs[0]["mask"] > 0
    out = {"workload": "16 pairs 848x800, L5 W10 I50, one batch per storage type"}
    for st in [x for x in args.reduced_storage.split(",") if x]:
        solvers = [_solver(vfs_amd, P, p, 5, 10, 50, storage=st) for p in pairs]
        if args.batch_no_autotune:
            solvers[0].set_option("autotune", 0)
        b = vfs_amd.Batch(solvers)
        apply_tuning(b, args.batch_tuning)
        res = _time_batch(args, b, 16)
        (tw, th), _, k = b.level_tuning(0, with_per_launch=True)
        res["tile_tuning_L0"] = f"{tw}x{th}, {k} iterations per launch"
        # algorithmic bytes per pixel-iteration (FP32 100; 16-bit storage of p, q, a, b,
        # c, Iu, Iz: 66) and, where profiled, the HBM bytes the counters measured
        res["algorithmic_bytes_per_px_iter"] = 100 if st == "f32" else 66
        res["measured_hbm_bytes_per_px_iter"] = profiled_batch_traffic(st)
        if ref is not None:
            solvers[0].copyImagesToDevice(pairs[0]["i0"], pairs[0]["i1"])
            solvers[0].solveStereoForwardMasked()
            err = np.abs(solvers[0].copy1DDisparityToHost() - ref)[m]
            res["mean_err_px"] = float(f"{err.mean():.3e}")
            res["max_err_px"] = float(f"{err.max():.3e}")
        out[st] = res
        print(f"bench: reduced storage {st}: {res}", file=sys.stderr, flush=True)
        b.close()
        for g in solvers:
            g.close()
    return out


def serving_streams(S, n_streams=32, distinct=4, only=None):
    """Camera streams alternately 424x400 and 848x800 (stream s: seed 300 + s).  Each
    stream has `distinct` synthetic frames along one smooth drift (<= 0.5 px per step);
    frame k of the stream is distinct frame triangle(k), so consecutive frames always
    differ by one drift step.  `only`: the stream indices to build."""
    streams = []
    for sidx in range(n_streams) if only is None else only:
        W, H = (424, 400) if sidx % 2 == 0 else (848, 800)
        base = S.make_pair(W, H, 300 + sidx, H / 2.0, W / 21.2)
        fr = [base] + [S.drift_pair(base, k, 300 + sidx) for k in range(1, distinct)]
        streams.append((W, H, base, fr))
    return streams


def tri(k, n):
    period = 2 * (n - 1)
    m = k % period
    return m if m < n else period - m


def pinned_frames(srv, sids, streams):
    """Every distinct frame of every stream in page-locked host memory (as a camera
    driver would deliver them), submitted to the server without a host copy."""
    return [[srv.pinned_frame(sid, f["i0"], f["i1"]) for f in streams[i][3]] for i, sid in enumerate(sids)]


def prime_streams(srv, sids, streams, frames, track=()):
    """Each stream's first frame (the full cold solve) so that the timed frames are the
    steady state of a stream: warm-started from the stream's previous result.  Returns the
    latencies and, for the streams in `track`, the frame's disparity."""
    lat, u0 = [], {}
    for i, sid in enumerate(sids):
        srv.submit_pinned(sid, frames[i][0], tag=-1 - i, want_result=i in track)
    for r in srv.run_until_idle():
        lat.append(r["latency_ms"])
        if r["u"] is not None:
            u0[-1 - r["tag"]] = r["u"]
    return np.array(lat), u0


def run_served(srv, sids, streams, frames, n_frames, rate, seed=7, backlog=96, track=()):
    """Poisson arrivals at `rate` frames/s, or (rate <= 0, saturation) arrivals faster
    than the server serves them: submissions interleaved with the server's steps keep
    `backlog` frames queued.  Each arrival is the next frame of a uniformly chosen stream
    (frames 1 .. n_frames of the stream, frame 0 was primed).  Returns latencies, their
    stream indices, the wall time from the first arrival to the last result, the number of
    warm-started frames, the largest non-finite count and, for the streams in `track`, the
    disparity of every frame by (stream, frame)."""
    rng = np.random.default_rng(seed)
    n_total = n_frames * len(sids)
    order = []
    left = [n_frames] * len(sids)
    for _ in range(n_total):
        live = [i for i in range(len(sids)) if left[i] > 0]
        i = int(rng.choice(live))
        order.append(i)
        left[i] -= 1
    times = np.zeros(n_total) if rate <= 0 else np.cumsum(rng.exponential(1.0 / rate, n_total))
    next_frame = [1] * len(sids)
    frame_of = {}
    lat, who, warm, nonfinite = [], [], 0, 0
    got = {}
    k = q = 0
    t0 = time.perf_counter()
    while True:
        now = time.perf_counter() - t0
        while k < n_total and (q < backlog if rate <= 0 else times[k] <= now):
            i = order[k]
            frame_of[k] = (i, next_frame[i])
            srv.submit_pinned(sids[i], frames[i][tri(next_frame[i], len(frames[i]))], tag=k, want_result=i in track)
            next_frame[i] += 1
            k += 1
            q += 1
        act, q = srv.step()
        for r in srv.collect():
            lat.append(r["latency_ms"])
            who.append(order[r["tag"]])
            warm += r["warm"]
            nonfinite = max(nonfinite, r["nonfinite"])
            if r["u"] is not None:
                got[frame_of[r["tag"]]] = r["u"]
        if k == n_total and act == 0 and q == 0:
            break
        if time.perf_counter() - t0 > 300:
            raise RuntimeError("serving run did not drain")
    wall = time.perf_counter() - t0
    return np.array(lat), np.array(who), wall, warm, nonfinite, got


def _pct(a, q):
    return round(float(np.percentile(a, q)), 2) if len(a) else None


def _warm_frame_bytes(W, H, P, levels=(1, 0), warps=2, iters=20):
    """Algorithmic bytes of one warm-started frame: the calibration warp and pyramid, the
    downscaled start state, then per level warps x (iters x 100 + 104) + 44 bytes per pixel."""
    dims, w, h = [], W, H
    for _ in range(max(levels) + 1):
        dims.append(w * h)
        w, h = int(np.float32(w) / np.float32(P["fScale"])), int(np.float32(h) / np.float32(P["fScale"]))
    return sum(dims[l] * (warps * (iters * B_ITER_FP32 + 104.0) + 44.0) for l in levels) + 16.0 * dims[0] + \
        40.0 * sum(dims[1:])


def check_served(P, streams, track, u0, got, n_frames, every=10):
    """The FP64 oracle runs each tracked stream's sequence (cold frame 0, then warm W2 I20
    from level 1, as the server) and is compared with the served frames at every
    `every`-th frame and the last: mean and max |u_served - u_fp64| over the mask."""
    from oracle import oracle as O
    res = {}
    for i in track:
        W, H, base, fr = streams[i]
        o = O.Oracle(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], 5,
                     P["fScale"], 10, 50)
        o.set_limit_range(P["limitRange"])
        o.copyMaskToDevice(base["mask"])
        o.loadVectorFields(base["tv"], base["cv"])
        m = base["mask"] > 0
        means, maxes = [], []
        for k in range(n_frames + 1):
            f = fr[tri(k, len(fr))]
            o.copyImagesToDevice(f["i0"], f["i1"])
            if k == 0:
                o.solveStereoForwardMasked()
            else:
                o.solve_warm(1, 2, 20)
            if k % every == 0 or k == n_frames:
                u = u0[i] if k == 0 else got[(i, k)]
                err = np.abs(u - o.copy1DDisparityToHost())[m]
                means.append(float(err.mean()))
                maxes.append(float(err.max()))
        res[f"stream{i}_{W}x{H}"] = {"frames_checked": len(means), "mean_err_px_max_over_frames": float(f"{max(means):.3e}"),
                                     "mean_err_px_avg": float(f"{np.mean(means):.3e}"),
                                     "max_err_px": float(f"{max(maxes):.3e}")}
    return res


def measure_served(args, P, S, vfs_amd, plain_pairs_per_s, plain_px):
    """Serving: 32 streams (424x400 / 848x800), continuous batching, FP16 storage.  Each
    stream's first frame runs the full L5 W10 I50 solve (primed, reported apart); the
    timed frames are warm-started (W2 I20 at levels 1 and 0).  Saturation throughput
    (96 frames kept queued, so its latencies are queueing delays), then Poisson arrivals
    at 50 % and 90 % of it, and a saturation run of the 16 848x800 streams alone.  Streams
    0 and 1 of the saturation run are checked frame by frame against the FP64 oracle."""
    streams = serving_streams(S)
    n_timed = args.serve_frames - 1
    print("bench: serving streams ready", file=sys.stderr, flush=True)
    out = {"workload": f"32 streams (424x400 / 848x800), {args.serve_frames} frames each: one cold frame "
                       f"(L5 W10 I50), then {n_timed} warm-started frames (W2 I20 at levels 1-0), FP16 storage, "
                       f"continuous batching; frames cycle through 4 drift steps per stream and are submitted from "
                       f"page-locked host buffers (no host copy), copied in and out on the server's copy stream"}
    res = {}
    track = (0, 1) if not (args.no_serve_check or args.no_cpu) else ()

    def phase(label, strs, rate, track=()):
        srv = vfs_amd.Server(P, hbm_budget_bytes=64 << 30, max_streams=32, max_active=64, storage="f16")
        sids = [srv.add_stream(b["mask"], b["tv"], b["cv"], 5, 10, 50, 1, 2, 20) for (W, H, b, fr) in strs]
        frames = pinned_frames(srv, sids, strs)
        cold, u0 = prime_streams(srv, sids, strs, frames, track)
        time.sleep(0.05)  # an idle gap that separates the phases in a kernel trace
        r = rate if rate <= 0 else rate * res["saturation"]["frames_per_s"]
        t_before = srv.stats()
        lat, who, wall, warm, nonfinite, got = run_served(srv, sids, strs, frames, n_timed, r, track=track)
        st = srv.stats()
        small = np.array([strs[i][0] == 424 for i in who])
        px = sum(strs[i][0] * strs[i][1] for i in who)
        byts = sum(_warm_frame_bytes(strs[i][0], strs[i][1], P) for i in who)
        entry = {"offered_frames_per_s": round(r, 1) if r > 0 else None,
                 "frames_per_s": round(len(lat) / wall, 1),
                 "Mpx_per_s": round(px / wall / 1e6, 1),
                 "algorithmic_GBs": round(byts / wall / 1e9, 1),
                 "roofline_frac": round(byts / wall / 1e9 / HBM_PEAK_GBS, 3),
                 "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99),
                 "frames": int(len(lat)), "warm_frames": int(warm), "max_nonfinite": int(nonfinite),
                 "ops_per_launch": round((st["entries"] - t_before["entries"])
                                         / max(st["ticks"] - t_before["ticks"], 1), 2)}
        if small.any() and (~small).any():
            entry.update({"p50_ms_424x400": _pct(lat[small], 50), "p99_ms_424x400": _pct(lat[small], 99),
                          "p50_ms_848x800": _pct(lat[~small], 50), "p99_ms_848x800": _pct(lat[~small], 99)})
        if rate <= 0:
            entry["cold_first_frames"] = {"frames": int(len(cold)), "p50_ms": _pct(cold, 50),
                                          "p99_ms": _pct(cold, 99)}
        res[label] = entry
        print(f"bench: served {label}: {entry}", file=sys.stderr, flush=True)
        srv.close()
        time.sleep(0.05)
        return u0, got

    u0, got = phase("saturation", streams, 0.0, track)
    phase("rate_50", streams, 0.5)
    phase("rate_90", streams, 0.9)
    phase("saturation_848x800_only", [st for st in streams if st[0] == 848], 0.0)
    out.update(res)
    sat, big = res["saturation"], res["saturation_848x800_only"]
    out["vs_plain_per_frame_solve"] = round(sat["frames_per_s"] / plain_pairs_per_s, 2)
    out["like_for_like"] = {
        "848x800_frames_vs_plain": round(big["frames_per_s"] / plain_pairs_per_s, 2),
        "Mpx_per_s_vs_plain": round(sat["Mpx_per_s"] / (plain_pairs_per_s * plain_px / 1e6), 2),
        "work_per_848x800_frame_cold_over_warm": round(
            _cold_bytes(848, 800, P) / _warm_frame_bytes(848, 800, P), 2),
        "note": "warm frames do the W2 I20 schedule from level 1 and plain solves the full L5 W10 I50 one"}
    if track:
        out["fp64_check"] = check_served(P, streams, track, u0, got, n_timed)
        print(f"bench: served FP64 check {out['fp64_check']}", file=sys.stderr, flush=True)
    return out


def _cold_bytes(W, H, P, levels=5, warps=10, iters=50):
    dims, w, h = [], W, H
    for _ in range(levels):
        dims.append(w * h)
        w, h = int(np.float32(w) / np.float32(P["fScale"])), int(np.float32(h) / np.float32(P["fScale"]))
    return sum(dims) * (warps * (iters * B_ITER_FP32 + 104.0) + 44.0) + 16.0 * dims[0] + 40.0 * (sum(dims) - dims[0])


if __name__ == "__main__":
    main()
