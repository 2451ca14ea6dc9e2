"""Benchmark: converged sweep points per second on one MI355X.

Workload: Monte Carlo operating points of the 64-OTA MOS1 array (324 MNA
unknowns, 32 live parameters per point), 16,384 points per step, one step =
one esl_op_solve of the whole batch (host-side lead factor, upload, the
Newton kernel, re-pivot rounds, download). Modes:
  fp64_cold     plain FP64 lanes, cold start (INITFIX + junction seeds)
  fp64_nominal  plain FP64 lanes warm-started from the template's nominal
                operating point (the reference's points warm_start,
                src/espice.zig:637-653)
The headline `value` is the best converged points/s among the modes run.

Roofline: the dominant kernel is op_newton, one wavefront per lane with the
lane's factors in LDS, so HBM carries little and the bound is FP64 vector
issue. Algorithmic FLOPs per lane-iteration: 2F + nL (refactor) +
2(nL + nU) + n (solve) + 3n (update and gates) + E_eval (device evaluation),
summed over every lane's iterations and divided by the kernel's HIP-event
time.

CPU baseline: the oracle (the repository's CPU FP64 restatement, same deck,
seeds, pivot tape and start) on the host cores, over a bounded sample.
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

# FLOPs per device evaluation, counted from the restated equations
# (exp/log/sqrt/div count 1 each): MOS1 eval with both junctions and the
# limiter pass; V source; I source; resistor.
E_EVAL = {"mos1": 148, "v": 6, "i": 2, "r": 7}
PEAK_FP64_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=16384)
    ap.add_argument("--modes", default="fp64_cold,fp64_nominal")
    ap.add_argument("--cpu-sample", type=int, default=2048)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tiling", type=int, default=0, help="0 autotune, 1 lane per wavefront, 2 lane per thread")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run one plain device copy of this many MiB (scales the HBM byte counters) and exit")
    return ap.parse_args()


def calibrate_copy(mib):
    import torch
    src = torch.empty(mib << 20, dtype=torch.uint8, device="cuda")
    src.fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": mib << 20, "copy_bytes_written": mib << 20}))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported: sweep points are "
              "independent, so more GPUs would each take a slice of the points", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy_mib:
        calibrate_copy(args.calibrate_copy_mib)
        return
    from espice_amd import circuits, lanes

    fam = circuits.ota_array()
    deck = fam.deck
    n_pts = args.points
    vals = fam.sample(2026, n_pts)
    eng = lanes.Engine(deck.arrays())
    info0 = eng.info()
    _, _, q = eng.pattern()
    n_mos = sum(1 for k in deck.kind if k == 5)
    n_v = sum(1 for k in deck.kind if k == 2)
    n_i = sum(1 for k in deck.kind if k == 3)
    n_r = sum(1 for k in deck.kind if k == 0)
    e_eval = E_EVAL["mos1"] * n_mos + E_EVAL["v"] * n_v + E_EVAL["i"] * n_i + E_EVAL["r"] * n_r

    # the nominal operating point for warm starts: one lane on the GPU
    nom = eng.solve_op(fam.nominal()[None, :])
    x_nom = np.repeat(nom["x"], n_pts, 0)

    results = {}
    for mode in args.modes.split(","):
        x0 = x_nom if mode == "fp64_nominal" else None
        for _ in range(args.warmup):
            eng.solve_op(vals, x0=x0, tiling=args.tiling)
        t_kernel, t_total, conv, iters, lane_it = 0.0, 0.0, 0, 0.0, 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = eng.solve_op(vals, x0=x0, tiling=args.tiling)
            tm = eng.timing()
            t_kernel += tm.kernel_ms
            t_total += tm.total_ms
            conv += int((r["reasons"] == 0).sum())
            iters += float(r["iterations"].mean())
            lane_it += int(tm.lane_iterations)
        wall = time.perf_counter() - t0
        info = eng.info()
        flops_iter = 2 * info.flops + info.nnz_l + 2 * (info.nnz_l + info.nnz_u) + info.n + 3 * info.n + e_eval
        achieved = flops_iter * lane_it / (t_kernel * 1e-3) / 1e12
        results[mode] = {
            "points_per_s": conv / wall,
            "ms_per_step": 1e3 * wall / args.steps,
            "kernel_ms_per_step": t_kernel / args.steps,
            "converged_fraction": conv / (n_pts * args.steps),
            "newton_iterations_per_point": iters / args.steps,
            "tflops_fp64": achieved,
            "flops_per_lane_iteration": flops_iter,
            "repivots": int(tm.repivots),
            "last": r,
        }
    best = max(results, key=lambda m: results[m]["points_per_s"])
    res = results[best]

    # parity on a bounded sample against the oracle (same points, tape, start)
    parity = {}
    cpu = None
    if not args.no_cpu:
        from oracle import orc
        c = orc.Circuit(deck.arrays())
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        k = min(args.cpu_sample, n_pts)
        for mode in results:
            x0 = x_nom[:k] if mode == "fp64_nominal" else None
            t0 = time.perf_counter()
            ref = c.op(vals[:k], q, start_x=x0, threads=threads)
            dt = time.perf_counter() - t0
            got = results[mode]["last"]
            gx = got["x"][:k]
            rel = np.abs(gx - ref["x"]) / np.maximum(np.abs(ref["x"]), 1e-12)
            parity[mode] = {
                "sample": k,
                "bitwise_equal": bool(np.array_equal(gx.view(np.uint64), ref["x"].view(np.uint64))),
                "max_rel_err": float(rel.max()),
                "same_iterations": bool(np.array_equal(got["iterations"][:k], ref["iterations"])),
            }
            if mode == best:
                cpu = {"value": float((ref["reasons"] == 0).sum() / dt), "unit": "points/s", "cores": threads,
                       "kind": "port",
                       "sample": f"{k} points of the same workload ({mode}), oracle on {threads} host threads, "
                                 f"{dt:.2f} s wall"}

    out = {
        "metric": "converged sweep points/sec, 1 MI355X: learned+FP32-refined vs plain FP64 lanes",
        "value": round(res["points_per_s"], 1),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Monte Carlo draws of the 32 live parameters)",
        "config": {
            "workload": f"ota_array_64 .op Monte Carlo, {n_pts} points per step",
            "mode": best,
            "n": info0.n, "nnz": info0.nnz, "nnz_l": res and eng.info().nnz_l, "nnz_u": eng.info().nnz_u,
            "flops_F": eng.info().flops, "live_params": deck.arrays()["n_live"],
            "tiling": "lane-per-wavefront, factors in LDS",
        },
        "roofline": {
            "bound": "mfma",
            "pipe": "FP64 vector ALU (lane factors stay in LDS; peak = MI355X FP64 78.6 TFLOP/s)",
            "kernel": "op_newton",
            "achieved": round(res["tflops_fp64"], 4),
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(res["tflops_fp64"] / PEAK_FP64_TFLOPS, 5),
            "traffic": None,
        },
        "cpu_baseline": cpu,
        "modes": {m: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items() if k != "last"}
                  for m, d in results.items()},
        "parity": parity,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
