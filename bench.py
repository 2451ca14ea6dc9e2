"""Benchmark: converged sweep points per second on one MI355X.

Workload: Monte Carlo operating points of the 64-OTA MOS1 array (324 MNA
unknowns, 32 live parameters per point), 16,384 points per step; one step =
one esl_op_solve of the whole batch, timed from the values on the host to
the results on the host (upload, network inference where used, lead factor,
the Newton kernel, re-pivot and fallback rounds, download). A point's
results are its measures, the 64 OTA output voltages, as the points API
returns per-point measures (include/espice.h:338). Modes:
  fp64_cold      plain FP64 lanes, cold start (INITFIX + junction seeds): the
                 baseline of the headline ratio
  fp64_nominal   plain FP64 lanes warm-started from the template's nominal
                 operating point (the reference's points warm start,
                 src/espice.zig:637-653)
  fp32_cold      FP32 factors + FP64 refinement, cold start
  learned_fp32   the headline: the warm-start network's guesses (trained
                 beforehand, untimed, on 50,000 operating points solved by
                 the engine) + FP32 factors with FP64 refinement, nominal
                 restart for lanes whose residual does not fall
`value` is the headline mode's converged points/s.

Roofline: the dominant kernel is op_newton, one wavefront per lane with the
lane's factors in LDS, so HBM carries little and the bound is vector issue.
Algorithmic FLOPs per lane-iteration: 2F + nL (refactor) + 2(nL + nU) + n
(solve) + 3n (update and gates) + E_eval (device evaluation), summed over
every lane's iterations and divided by the kernel's HIP-event time; the
refinement steps of FP32 lanes are not counted (so `achieved` is a lower
bound there). Peak: MI355X FP64 vector, 78.6 TFLOP/s.

CPU baseline: the oracle (the repository's CPU FP64 restatement: same deck,
seeds, pivot tape and cold start) on the host cores, over a bounded sample.
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
# Gummel-Poon: five junction laws, the base charge qb with its square
# root, the Early and knee terms, the limiter pass and the residual
# corrections.
E_EVAL = {"mos1": 148, "v": 6, "i": 2, "r": 7, "bjt": 190}
# Roofline label of the lane kernels: their factors stay in LDS, so the
# bound is FP64 vector issue (no MFMA instruction runs in them).
BOUND_FP64 = "fp64_vector"
PEAK_FP64_TFLOPS = 78.6
HEADLINE = "learned_fp32"
BASELINE_MODE = "fp64_cold"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=16384)
    ap.add_argument("--modes", default="fp64_cold,fp64_nominal,fp32_cold,learned_fp64,learned_fp32")
    ap.add_argument("--train-points", type=int, default=50000)
    ap.add_argument("--train-epochs", type=int, default=30)
    ap.add_argument("--cpu-sample", type=int, default=16384)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tiling", type=int, default=0, help="0 autotune, 1 lane per wavefront, 2 lane per thread")
    ap.add_argument("--mesh-lanes", type=int, default=4096, help="rc_mesh: points")
    ap.add_argument("--mesh-steps", type=int, default=1000, help="rc_mesh: trapezoid steps of 1 ps")
    ap.add_argument("--mesh-steps-per-launch", type=int, default=10, help="rc_mesh: steps per kernel launch")
    ap.add_argument("--workload", default="ota_array",
                    choices=["ota_array", "mixed_batch", "rc_mesh", "diffpair_array", "inverter_chain"],
                    help="ota_array: the headline; mixed_batch: 10,000 uneven instances of eight families "
                         "in one persistent launch; diffpair_array: 2048 points of 64 Gummel-Poon pairs, "
                         "FP32 factors + FP64 refinement against plain FP64; inverter_chain: 256 points of "
                         "the 16-stage MOS1 chain, plain FP64")
    ap.add_argument("--pair-points", type=int, default=2048, help="diffpair_array: points")
    ap.add_argument("--chain-points", type=int, default=256, help="inverter_chain: points")
    ap.add_argument("--instances", type=int, default=10000)
    ap.add_argument("--families", default="", help="mixed_batch: comma-separated subset of the eight families")
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


def profile_traffic(name):
    """traffic_bytes_per_launch of a committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def measured_traffic(mode, points):
    """HBM bytes per op_newton launch from the committed counter summary for
    this workload, or None when no summary covers it."""
    if mode != "learned_fp32" or points != 16384:
        return None
    return profile_traffic("op_newton_ota_array_64_learned_fp32_hbm_traffic.md")


def train_network(lanes, fam, eng, x_nom, rows, epochs):
    """The warm-start network, trained on operating points the engine solves."""
    t0 = time.perf_counter()
    vals = fam.sample(7, rows)   # a seed disjoint from the benchmark points
    sol = eng.solve_op(vals, x0=np.repeat(x_nom[None, :], rows, 0), tiling=lanes.TILING_WAVE)
    ok = sol["reasons"] == 0
    vals, xs = vals[ok], sol["x"][ok]
    t1 = time.perf_counter()
    net = lanes.Mlp(vals.shape[1], eng.n - 1, n_hidden=256, n_hidden_layers=3, seed=7)
    net.set_scaling(vals.mean(0), vals.std(0) + 1e-30, xs[:, 1:].mean(0), xs[:, 1:].std(0) + 1e-12)
    losses = net.train(vals, xs, epochs=epochs, batch=512)
    t2 = time.perf_counter()
    return net, {"rows": int(ok.sum()), "epochs": epochs, "solve_s": round(t1 - t0, 3),
                 "train_s": round(t2 - t1, 3), "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
                 "arch": "32-256-256-256-323, SiLU, BF16 MFMA, FP32 master, Adam"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported: sweep points are "
              "independent, so more GPUs would each take a slice of the points", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy_mib:
        calibrate_copy(args.calibrate_copy_mib)
        return
    if args.workload == "rc_mesh":
        run_mesh(args)
        return
    if args.workload == "mixed_batch":
        run_mixed(args)
        return
    if args.workload == "diffpair_array":
        run_pairs(args)
        return
    if args.workload == "inverter_chain":
        run_chain(args)
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

    # the nominal operating point: one lane on the GPU
    x_nom = eng.solve_op(fam.nominal()[None, :])["x"][0]
    modes = args.modes.split(",")
    training = None
    net = None
    if any(m.startswith("learned") for m in modes):
        net, training = train_network(lanes, fam, eng, x_nom, args.train_points, args.train_epochs)
        eng.set_warm_start(net, x_nom)

    # measures per point: the 64 OTA outputs (the points API returns
    # per-point measures, include/espice.h:338, not whole state vectors)
    probes = np.array([deck.node_index(f"out{k}") for k in range(64)], dtype=np.uint32)

    def run(mode, measures=True):
        kw = {"tiling": args.tiling}
        if measures:
            kw["probes"] = probes
        if mode == "fp64_nominal":
            kw["x0"] = np.repeat(x_nom[None, :], n_pts, 0)
        if mode.startswith("learned"):
            kw["learned"] = True
            kw["tiling"] = lanes.TILING_WAVE
        if "fp32" in mode:
            kw["precision"] = lanes.FP32_REFINED
            kw["tiling"] = lanes.TILING_WAVE
        return eng.solve_op(vals, **kw)

    results = {}
    for mode in modes:
        for _ in range(args.warmup):
            run(mode)
        t_kernel, t_total, t_warm, conv, iters, lane_it, fb64, fbw = 0.0, 0.0, 0.0, 0, 0.0, 0, 0, 0
        t_lead, t_dl = 0.0, 0.0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = run(mode)
            tm = eng.timing()
            t_kernel += tm.kernel_ms
            t_total += tm.total_ms
            t_warm += tm.warm_ms if mode.startswith("learned") else 0.0
            t_lead += tm.lead_ms
            t_dl += tm.download_ms
            conv += int((r["reasons"] == 0).sum())
            iters += float(r["iterations"].mean())
            lane_it += int(tm.lane_iterations)
            fb64 += int(tm.fp64_lanes)
            fbw += int(tm.warm_fallbacks) if mode.startswith("learned") else 0
        wall = time.perf_counter() - t0
        info = eng.info()
        flops_iter = 2 * info.flops + info.nnz_l + 2 * (info.nnz_l + info.nnz_u) + info.n + 3 * info.n + e_eval
        achieved = flops_iter * lane_it / (t_kernel * 1e-3) / 1e12
        results[mode] = {
            "points_per_s": conv / wall,
            "ms_per_step": 1e3 * wall / args.steps,
            "kernel_ms_per_step": t_kernel / args.steps,
            "library_ms_per_step": t_total / args.steps,
            "inference_ms_per_step": t_warm / args.steps,
            "lead_tape_ms_per_step": t_lead / args.steps,
            "download_ms_per_step": t_dl / args.steps,
            "converged_fraction": conv / (n_pts * args.steps),
            "newton_iterations_per_point": iters / args.steps,
            "fp64_fallback_rate": fb64 / (n_pts * args.steps),
            "nominal_fallback_rate": fbw / (n_pts * args.steps),
            "tflops": achieved,
            "flops_per_lane_iteration": flops_iter,
            "repivots": int(tm.repivots),
            "tiling": int(tm.tiling),
            "last": run(mode, measures=False),
        }
    head = HEADLINE if HEADLINE in results else max(results, key=lambda m: results[m]["points_per_s"])
    res = results[head]

    # parity and error on a bounded sample against the oracle
    parity = {}
    cpu = None
    if not args.no_cpu:
        from oracle import orc
        c = orc.Circuit(deck.arrays())
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        k = min(args.cpu_sample, n_pts)
        ref_cold = None
        if BASELINE_MODE in results or head != BASELINE_MODE:
            t0 = time.perf_counter()
            ref_cold = c.op(vals[:k], q, threads=threads)
            dt = time.perf_counter() - t0
            k1 = min(1024, k)
            t1 = time.perf_counter()
            one = c.op(vals[:k1], q, threads=1)
            dt1 = time.perf_counter() - t1
            cpu_model = ""
            try:
                with open("/proc/cpuinfo") as f:
                    cpu_model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
            except OSError:
                pass
            cpu = {"value": round(float((ref_cold["reasons"] == 0).sum() / dt), 1), "unit": "points/s",
                   "cores": threads, "kind": "port",
                   "sample": f"{k} points of the same workload, plain FP64 cold-start lanes of the oracle "
                             f"(the CPU restatement) on {threads} host threads, {dt:.2f} s wall",
                   "single_thread_points_per_s": round(float((one["reasons"] == 0).sum() / dt1), 1),
                   "host_cpus": os.cpu_count(), "cpu_model": cpu_model}
        nn = deck.arrays()["n_nodes"]
        x0_learned = net.predict(vals[:k]) if net is not None else None
        for mode, got in results.items():
            g = got["last"]
            kw = {}
            if mode == "fp64_nominal":
                kw["start_x"] = np.repeat(x_nom[None, :], k, 0)
            if mode.startswith("learned"):
                kw["start_x"] = x0_learned
                kw["nominal_x"] = x_nom
            same_start_fp64 = c.op(vals[:k], q, threads=threads, **kw)
            gx = g["x"][:k]
            err = np.abs(gx - same_start_fp64["x"]) / np.maximum(np.abs(same_start_fp64["x"]), 1e-6)
            entry = {
                "sample": k,
                "max_rel_err_vs_fp64_oracle_same_start": float(err.max()),
                "bitwise_equal_fp64_oracle": bool(np.array_equal(gx.view(np.uint64),
                                                                 same_start_fp64["x"].view(np.uint64))),
            }
            if "fp32" in mode:
                ref32 = c.op(vals[:k], q, threads=threads, precision=1, **kw)
                entry["bitwise_equal_fp32_oracle"] = bool(np.array_equal(gx.view(np.uint64),
                                                                         ref32["x"].view(np.uint64)))
            if ref_cold is not None:
                ok = (ref_cold["reasons"] == 0) & (g["reasons"][:k] == 0)
                d = np.abs(gx[ok, :nn] - ref_cold["x"][ok, :nn])
                entry["max_abs_node_err_vs_cold_fp64_oracle_V"] = float(d.max()) if ok.any() else None
            parity[mode] = entry

    base = results.get(BASELINE_MODE)
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
        "dtype": "f64+f32" if "fp32" in head else "f64",
        "data": "synthetic (seeded Monte Carlo draws of the 32 live parameters)",
        "config": {
            "workload": f"ota_array_64 .op Monte Carlo, {n_pts} points per step",
            "mode": head,
            "n": info0.n, "nnz": info0.nnz, "nnz_l": eng.info().nnz_l, "nnz_u": eng.info().nnz_u,
            "flops_F": eng.info().flops, "live_params": deck.arrays()["n_live"],
            "tiling": "lane-per-wavefront, factors in LDS" if res["tiling"] == 1 else "lane-per-thread",
        },
        "speedup_vs_plain_fp64": (round(res["points_per_s"] / base["points_per_s"], 3) if base else None),
        "roofline": {
            "bound": BOUND_FP64,
            "pipe": "FP64 vector ALU (lane factors stay in LDS; no MFMA; peak = MI355X FP64 vector 78.6 TFLOP/s)",
            "kernel": "op_newton",
            "achieved": round(res["tflops"], 4),
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(res["tflops"] / PEAK_FP64_TFLOPS, 5),
            "traffic": measured_traffic(head, n_pts),
        },
        "cpu_baseline": cpu,
        "training": training,
        "modes": {m: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items() if k != "last"}
                  for m, d in results.items()},
        "parity": parity,
    }
    print(json.dumps(out))


# Newton iteration limit of the mixed batch's operating points, per family:
# long inverter chains need hundreds of plain-Newton iterations from a cold
# start (about two per stage), the other families converge within 100.
# the uneven batch's operating points (.op items and the transients' initial
# points) run a plain Newton budget, the reference's itl1 (100,
# converger.zig:87) except for the long inverter chains, whose switching
# stages need up to about 1,000 plain iterations (the reference's fourth
# rung, JFNK, which would take many of them, is not built), and climb the
# .op ladder on the device when that fails (dc/op.zig:160-288)
MIXED_OP_MAX_ITER = {"inverter_chain": 1000, "inverter_chain_tran": 1000}
ORC_LADDER = 1   # the oracle's Tol flag for the .op ladder


def run_mesh(args):
    """The 64 x 64 RC mesh with diode clamps (4,099 unknowns), 4,096 points,
    1,000 fixed trapezoid steps of 1 ps after each point's operating point:
    lanes beyond a workgroup's LDS, one wavefront per lane with its factor
    slots and vectors in HBM. One step of the benchmark = one esl_tran_solve
    of every point (the operating points, the transient lead tape, every
    step), host to host. The transient reports its progress every 100 steps
    (the progress callback), which this run logs."""
    from espice_amd import circuits, lanes
    log = lambda m: print(m, file=sys.stderr, flush=True)
    fam = circuits.rc_mesh(64)
    d = fam.deck
    eng = lanes.Engine(d.arrays())
    probes = [d.node_index("n0_0"), d.node_index("n63_63"), d.node_index("n32_32")]
    vals = fam.sample(2024, args.mesh_lanes)
    steps, dt = args.mesh_steps, 1e-12
    t0 = time.perf_counter()
    eng.solve_tran(vals[:64], dt, 2, probes=probes)   # untimed: tapes, workspace
    log(f"rc_mesh 64x64: n {eng.n}, warmup {time.perf_counter() - t0:.1f} s")
    conv, wall, kern, lane_it, res = 0, 0.0, 0.0, 0, None
    for _ in range(args.steps):
        t1 = time.perf_counter()
        prog = lambda done, total: log(f"  {done}/{total} steps, {time.perf_counter() - t1:.1f} s") and False
        res = eng.solve_tran(vals, dt, steps, probes=probes, steps_per_launch=args.mesh_steps_per_launch,
                             progress=prog, progress_steps=100)
        wall += time.perf_counter() - t1
        tm = eng.timing()
        kern += tm.kernel_ms
        lane_it += int(tm.lane_iterations)
        conv += int((res["reasons"] == 0).sum())
        log(f"step: {time.perf_counter() - t1:.1f} s, kernel {tm.kernel_ms:.0f} ms, "
            f"{res['iterations'].mean():.1f} transient iterations per point, repivots {tm.repivots}")
    info = eng.info()
    n, nnz, nl, nu = eng.n, eng.nnz, int(info.nnz_l), int(info.nnz_u)
    # algorithmic HBM bytes per lane-iteration with the factors streamed:
    # A, L and U written then read back by the solves, the pivots, and six
    # n-vectors (x, x_old, dx, rhs, y, w)
    per_it = 8 * (nnz + 2 * (nl + nu) + n + 6 * n)
    achieved = per_it * lane_it / (kern * 1e-3) / 1e9
    parity, cpu = None, None
    if not args.no_cpu:
        from oracle import orc
        k = min(16, len(vals))
        _, _, q = eng.pattern()
        tc = time.perf_counter()
        ref = orc.tran(orc.Circuit(d.arrays()), vals[:k], q, dt, steps, probes=probes, threads=k)
        t_cpu = time.perf_counter() - tc
        ok = (ref["reasons"] == 0) & (res["reasons"][:k] == 0)
        gx, rx = res["probes"][:k][ok], ref["probes"][ok]
        parity = {"checked": int(ok.sum()), "probe_rows_bitwise_equal": bool(np.array_equal(
            gx.view(np.uint64), rx.view(np.uint64))), "x_bitwise_equal": bool(np.array_equal(
            res["x"][:k][ok].view(np.uint64), ref["x"][ok].view(np.uint64))),
            "iterations_equal": bool(np.array_equal(res["iterations"][:k], ref["iterations"]))}
        cpu = {"value": round(int((ref["reasons"] == 0).sum()) / t_cpu, 3), "unit": "points/s", "cores": k,
               "kind": "port", "sample": f"the first {k} points of the same workload, every step, on the oracle "
                                         f"(one thread per point, {k} threads, {t_cpu:.1f} s)"}
        log(f"oracle: {k} points in {t_cpu:.1f} s")
    # measured for launches of one step over 4096 points (two Newton
    # iterations each); a launch here runs steps_per_launch steps
    per_step = profile_traffic("big_lanes_rc_mesh_64_hbm_traffic.md")
    traffic, traffic_note = None, None
    if per_step is not None:
        spl = min(args.mesh_steps_per_launch, steps)
        traffic = int(per_step * spl * len(vals) / 4096)
        traffic_note = (f"counters of one-step launches over 4096 points ({per_step} B), "
                        f"times {spl} steps per launch and {len(vals)}/4096 points")
    out = {
        "metric": "converged sweep points/sec, 1 MI355X, 64x64 RC mesh 1000-step transient",
        "value": round(conv / wall, 3),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": round(1e3 * wall / args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded R, C and IS draws per point)",
        "config": {"workload": f"rc_mesh 64x64 + 5% diode clamps, {len(vals)} points, {steps} trapezoid steps "
                               f"of 1 ps", "n": n, "nnz": nnz, "nnz_l": nl, "nnz_u": nu,
                   "flops_F": int(info.flops), "tiling": "lane per wavefront, factors and vectors in HBM, "
                                                         "replay blocks staged in LDS"},
        "kernel_ms_per_step": round(kern / args.steps, 1),
        "transient_iterations_per_point": round(lane_it / max(conv, 1) / args.steps, 2),
        "roofline": {"bound": "hbm", "kernel": "big_lanes", "achieved": round(achieved, 2), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(achieved / 8000.0, 5), "traffic": traffic,
                     "bytes_per_lane_iteration": per_it, "traffic_note": traffic_note},
        "parity": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    eng.close()


def run_pairs(args):
    """64 Gummel-Poon differential pairs (128 BJTs, 453 unknowns), 2048
    Monte Carlo points per step, FP32 factors with FP64 refinement against
    plain FP64 lanes, both from a cold start."""
    from espice_amd import circuits, lanes
    fam = circuits.diffpair_array(64)
    deck = fam.deck
    n_pts = args.pair_points
    vals = fam.sample(2032, n_pts)
    eng = lanes.Engine(deck.arrays())
    _, _, q = eng.pattern()
    probes = np.array([deck.node_index(f"c{s}{k}") for k in range(64) for s in "pn"], dtype=np.uint32)

    def timed(precision):
        for _ in range(args.warmup):
            eng.solve_op(vals, precision=precision, tiling=lanes.TILING_WAVE, probes=probes)
        t_k, conv, lane_it, fb = 0.0, 0, 0, 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = eng.solve_op(vals, precision=precision, tiling=lanes.TILING_WAVE, probes=probes)
            tm = eng.timing()
            t_k += tm.kernel_ms
            conv += int((r["reasons"] == 0).sum())
            lane_it += int(tm.lane_iterations)
            fb += int(tm.fp64_lanes)
        wall = time.perf_counter() - t0
        return {"points_per_s": round(conv / wall, 1), "ms_per_step": round(1e3 * wall / args.steps, 3),
                "kernel_ms_per_step": round(t_k / args.steps, 3),
                "newton_iterations_per_point": round(lane_it / (n_pts * args.steps), 3),
                "fp64_fallback_rate": fb / (n_pts * args.steps), "lane_iterations": lane_it, "kernel_ms": t_k}

    f64 = timed(lanes.FP64)
    f32 = timed(lanes.FP32_REFINED)
    a = eng.solve_op(vals, tiling=lanes.TILING_WAVE)
    b = eng.solve_op(vals, precision=lanes.FP32_REFINED, tiling=lanes.TILING_WAVE)
    err = float((np.abs(a["x"] - b["x"]) / np.maximum(np.abs(a["x"]), 1e-6)).max())
    info = eng.info()
    # LU, vector and device-evaluation work per lane-iteration (the
    # refinement steps of FP32 lanes are not counted: a lower bound there)
    kinds = list(deck.kind)
    e_eval = (E_EVAL["bjt"] * kinds.count(6) + E_EVAL["r"] * kinds.count(0) + E_EVAL["v"] * kinds.count(2)
              + E_EVAL["i"] * kinds.count(3))
    flops_iter = 2 * info.flops + info.nnz_l + 2 * (info.nnz_l + info.nnz_u) + 4 * info.n + e_eval
    achieved = flops_iter * f32["lane_iterations"] / (f32["kernel_ms"] * 1e-3) / 1e12
    parity, cpu = None, None
    if not args.no_cpu:
        from oracle import orc
        c = orc.Circuit(deck.arrays())
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        pick = np.arange(0, n_pts, 97)
        r64 = c.op(vals[pick], q, threads=threads)
        r32 = c.op(vals[pick], q, threads=threads, precision=1)
        parity = {"sample": len(pick),
                  "fp64_bitwise_equal_oracle": bool(np.array_equal(a["x"][pick].view(np.uint64),
                                                                   r64["x"].view(np.uint64))),
                  "fp32_refined_bitwise_equal_oracle_fp32": bool(np.array_equal(b["x"][pick].view(np.uint64),
                                                                               r32["x"].view(np.uint64))),
                  "max_rel_err_fp32_refined_vs_fp64": err}
        # the sample solved again until at least a second of wall time
        k = min(args.cpu_sample, n_pts)
        t0 = time.perf_counter()
        done, rounds = 0, 0
        while rounds == 0 or time.perf_counter() - t0 < 1.0:
            rc = c.op(vals[:k], q, threads=threads)
            done += int((rc["reasons"] == 0).sum())
            rounds += 1
        t_cpu = time.perf_counter() - t0
        cpu = {"value": round(done / t_cpu, 1), "unit": "points/s", "cores": threads,
               "kind": "port", "sample": f"{k} points of the same workload solved {rounds} times, plain FP64 "
                                         f"cold-start lanes of the oracle on {threads} host threads, "
                                         f"{t_cpu:.2f} s wall"}
    out = {
        "metric": "converged sweep points/sec, 1 MI355X, 64 Gummel-Poon pairs .op, FP32 LU + FP64 refinement",
        "value": f32["points_per_s"],
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded IS, BF and load draws per point)",
        "config": {"workload": f"diffpair_array_64 .op Monte Carlo, {n_pts} points per step, cold start",
                   "n": int(info.n), "nnz_l": int(info.nnz_l), "nnz_u": int(info.nnz_u), "flops_F": int(info.flops),
                   "tiling": "lane-per-wavefront, factors in LDS"},
        "fp32_refined": {k: v for k, v in f32.items() if k not in ("lane_iterations", "kernel_ms")},
        "fp64": {k: v for k, v in f64.items() if k not in ("lane_iterations", "kernel_ms")},
        "fp32_refined_over_fp64": round(f32["points_per_s"] / max(f64["points_per_s"], 1e-9), 3),
        "roofline": {"bound": BOUND_FP64, "pipe": "FP64 vector ALU (peak 78.6 TFLOP/s; LU, vector and "
                     "Gummel-Poon evaluation flops, refinement steps not counted)", "kernel": "op_newton",
                     "achieved": round(achieved, 4), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 5), "traffic": None},
        "parity": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    eng.close()


def run_chain(args):
    """The 16-stage MOS1 inverter chain: plain FP64 Monte Carlo .op over
    --chain-points points (256 by default), cold starts, the autotuned
    tiling; the CPU lanes (the oracle on the host threads) beside it."""
    from espice_amd import circuits, lanes
    fam = circuits.inverter_chain(16)
    deck = fam.deck
    n_pts = args.chain_points
    vals = fam.sample(2025, n_pts)
    eng = lanes.Engine(deck.arrays())
    _, _, q = eng.pattern()
    probes = np.array([deck.node_index(f"s{k}") for k in range(1, 17)], dtype=np.uint32)
    for _ in range(args.warmup):
        eng.solve_op(vals, probes=probes)
    t_k, conv, lane_it = 0.0, 0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = eng.solve_op(vals, probes=probes)
        tm = eng.timing()
        t_k += tm.kernel_ms
        conv += int((r["reasons"] == 0).sum())
        lane_it += int(tm.lane_iterations)
    wall = time.perf_counter() - t0
    tiling = int(eng.timing().tiling)
    info = eng.info()
    kinds = list(deck.kind)
    e_eval = (E_EVAL["mos1"] * kinds.count(5) + E_EVAL["v"] * kinds.count(2) + E_EVAL["r"] * kinds.count(0)
              + E_EVAL["i"] * kinds.count(3))
    flops_iter = 2 * info.flops + info.nnz_l + 2 * (info.nnz_l + info.nnz_u) + 4 * info.n + e_eval
    achieved = flops_iter * lane_it / (t_k * 1e-3) / 1e12
    full = eng.solve_op(vals)
    parity, cpu = None, None
    if not args.no_cpu:
        from oracle import orc
        c = orc.Circuit(deck.arrays())
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        ref = c.op(vals, q, threads=threads)
        err = np.abs(full["x"] - ref["x"]) / np.maximum(np.abs(ref["x"]), 1e-6)
        parity = {"sample": n_pts,
                  "bitwise_equal_fp64_oracle": bool(np.array_equal(full["x"].view(np.uint64),
                                                                   ref["x"].view(np.uint64))),
                  "iterations_equal": bool(np.array_equal(full["iterations"], ref["iterations"])),
                  "max_rel_err_vs_fp64_oracle": float(err.max())}
        t0 = time.perf_counter()
        done, rounds = 0, 0
        while rounds == 0 or time.perf_counter() - t0 < 2.0:
            rc = c.op(vals, q, threads=threads)
            done += int((rc["reasons"] == 0).sum())
            rounds += 1
        t_cpu = time.perf_counter() - t0
        t1 = time.perf_counter()
        one = c.op(vals, q, threads=1)
        t_one = time.perf_counter() - t1
        cpu = {"value": round(done / t_cpu, 1), "unit": "points/s", "cores": threads, "kind": "port",
               "sample": f"the {n_pts} points of the workload solved {rounds} times, plain FP64 cold-start "
                         f"lanes of the oracle on {threads} host threads, {t_cpu:.2f} s wall",
               "single_thread_points_per_s": round(float((one["reasons"] == 0).sum() / t_one), 1)}
    out = {
        "metric": "converged sweep points/sec, 1 MI355X, 16-stage MOS1 inverter chain .op, plain FP64 lanes",
        "value": round(conv / wall, 1),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded VTO, VDD and input-level draws per point)",
        "config": {"workload": f"inverter_chain_16 .op Monte Carlo, {n_pts} points per step, cold start",
                   "n": int(info.n), "nnz": int(info.nnz), "nnz_l": int(info.nnz_l), "nnz_u": int(info.nnz_u),
                   "flops_F": int(info.flops),
                   "tiling": "lane-per-wavefront, factors in LDS" if tiling == 1 else "lane-per-thread"},
        "kernel_ms_per_step": round(t_k / args.steps, 4),
        "newton_iterations_per_point": round(lane_it / (n_pts * args.steps), 3),
        "roofline": {"bound": BOUND_FP64, "pipe": "FP64 vector ALU (peak 78.6 TFLOP/s)",
                     "kernel": "op_newton" if tiling == 1 else "op_newton_lpt",
                     "achieved": round(achieved, 5), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 6), "traffic": None},
        "parity": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    eng.close()


def run_mixed(args):
    """The uneven batch: every instance of eight families, .op and 200-step
    transients mixed, in one esl_batch_solve (one persistent launch, plus
    follow-up launches only for re-pivoted points). The first call tunes the
    tiling per template (untimed warmup)."""
    from espice_amd import circuits, lanes
    groups = circuits.mixed_batch(args.instances, seed=2024,
                                  families=args.families.split(",") if args.families else None)
    jobs, engines, meta = [], [], []
    t0 = time.perf_counter()
    for g in groups:
        fam = g["build"]()
        eng = lanes.Engine(fam.deck.arrays())
        vals = fam.sample_at(2024, g["instances"])
        tol = lanes.default_options().tol
        tol.max_iter = MIXED_OP_MAX_ITER.get(g["family"], 100)
        # operating points the plain Newton leaves unconverged climb the
        # ladder (dc/op.zig:160-288), as the reference's .op always does
        j = {"engine": eng, "values": vals, "tol": tol,
             "analysis": lanes.BATCH_OP if g["analysis"] == "op" else lanes.BATCH_TRAN,
             "ladder": True}
        if g["analysis"] == "tran":
            j.update(dt=g["dt"], steps=g["steps"], probes=[eng.n - 1])
        jobs.append(j)
        engines.append(eng)
        meta.append((g, fam))
    t_build = time.perf_counter() - t0
    n_inst = sum(len(j["values"]) for j in jobs)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    log(f"mixed batch: {n_inst} instances, {len(jobs)} templates built in {t_build:.1f} s")
    _, tune_rep = lanes.solve_batch(jobs, tiling=lanes.TILING_AUTO, tune=True)
    log(f"tuning batch: kernel {tune_rep.kernel_ms:.1f} ms, re-pivot launches {tune_rep.repivot_ms:.1f} ms, "
        f"total {tune_rep.total_ms:.1f} ms")
    for _ in range(max(args.warmup - 1, 0)):
        lanes.solve_batch(jobs)
    conv, wall, kern, rep_ms, res, rep = 0, 0.0, 0.0, 0.0, None, None
    for _ in range(args.steps):
        t1 = time.perf_counter()
        res, rep = lanes.solve_batch(jobs)
        wall += time.perf_counter() - t1
        kern += rep.kernel_ms
        rep_ms += rep.repivot_ms
        conv += sum(int((r["reasons"] == 0).sum()) for r in res)
        log(f"step: {time.perf_counter() - t1:.3f} s, kernel {rep.kernel_ms:.1f} ms, predicted {rep.predicted_ms:.1f} ms, "
            f"re-pivot launches {rep.repivot_ms:.1f} ms in {rep.repivot_launches} rounds ({rep.repivots} re-pivots), "
            f"setup {rep.setup_ms:.1f} ms")
    fams = {}
    for (g, fam), r in zip(meta, res):
        f = fams.setdefault(g["family"], {"instances": 0, "converged": 0, "templates": 0, "iterations": 0,
                                          "max_iterations": 0})
        f["max_iterations"] = max(f["max_iterations"], int(r["iterations"].max()))
        f["instances"] += len(r["reasons"])
        f["converged"] += int((r["reasons"] == 0).sum())
        f["templates"] += 1
        f["iterations"] += int(r["iterations"].sum())
    for (g, fam), j in zip(meta, jobs):
        fams[g["family"]]["tapes"] = fams[g["family"]].get("tapes", 0) + int(j["engine"].info().tapes)
    for f in fams.values():
        f["mean_iterations"] = round(f.pop("iterations") / max(f["instances"], 1), 2)
    # parity: a few points per template against the oracle, bitwise
    parity = {"checked": 0, "bitwise_equal": 0, "max_rel_err": 0.0}
    cpu = None
    if not args.no_cpu:
        from oracle import orc
        t_cpu, n_cpu, conv_cpu, est_s = 0.0, 0, 0, 0.0
        for (g, fam), j, r in zip(meta, jobs, res):
            k = min(2, len(j["values"]))
            vals = j["values"][:k]
            _, _, q = j["engine"].pattern()
            c = orc.Circuit(fam.deck.arrays())
            tc = time.perf_counter()
            if g["analysis"] == "op":
                t = j["tol"]
                ref = c.op(vals, q, threads=1,
                           tol=orc.Tol(t.reltol, t.abstol, t.vntol, t.residual_tol, t.max_iter, ORC_LADDER))

            else:
                t = j["tol"]
                ref = orc.tran(c, vals, q, g["dt"], g["steps"], threads=1,
                               tol_op=orc.Tol(t.reltol, t.abstol, t.vntol, t.residual_tol, t.max_iter, ORC_LADDER))
            dt_t = time.perf_counter() - tc
            t_cpu += dt_t
            est_s += dt_t / k * len(j["values"])   # this template's share of the whole batch
            n_cpu += k
            log(f"oracle {g['family']} {g['size']}: {time.perf_counter() - tc:.2f} s")
            conv_cpu += int((ref["reasons"] == 0).sum())
            ok = (ref["reasons"] == 0) & (r["reasons"][:k] == 0)
            parity["checked"] += int(ok.sum())
            gx, rx = r["x"][:k][ok], ref["x"][ok]
            parity["bitwise_equal"] += int(sum(np.array_equal(a.view(np.uint64), b.view(np.uint64))
                                               for a, b in zip(gx, rx)))
            if ok.any():
                err = np.abs(gx - rx) / np.maximum(np.abs(rx), 1e-6)
                parity["max_rel_err"] = max(parity["max_rel_err"], float(err.max()))
        # the CPU lanes on every host thread: every 8th instance of each
        # template (at least one), independent points spread over the
        # threads as the reference serves sweep points (one point per worker)
        from concurrent.futures import ThreadPoolExecutor
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        tasks = []
        for (g, fam), j in zip(meta, jobs):
            idx = np.arange(0, len(j["values"]), 8)
            t = j["tol"]
            flags = ORC_LADDER
            for i in idx:
                tasks.append((g, fam, j, int(i),
                              orc.Tol(t.reltol, t.abstol, t.vntol, t.residual_tol, t.max_iter, flags)))

        def one(task):
            g, fam, j, i, tol = task
            c = orc.Circuit(fam.deck.arrays())
            _, _, q = j["engine"].pattern()
            v = j["values"][i:i + 1]
            if g["analysis"] == "op":
                r = c.op(v, q, threads=1, tol=tol)
            else:
                r = orc.tran(c, v, q, g["dt"], g["steps"], threads=1, tol_op=tol)
            return int(r["reasons"][0] == 0)

        tasks.sort(key=lambda tk: -tk[0]["size"])   # long points first, for a short makespan
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as ex:
            done_cpu = sum(ex.map(one, tasks))
        t_all = time.perf_counter() - t0
        cpu = {"value": round(done_cpu / t_all, 2), "unit": "points/s", "cores": threads, "kind": "port",
               "sample": f"every 8th instance of each of the {len(jobs)} templates ({len(tasks)} points, "
                         f"{t_all:.1f} s wall) on the oracle, one point per task over {threads} host threads",
               "single_thread_projected_points_per_s": round(n_inst / est_s, 2)}
    out = {
        "metric": "converged sweep points/sec, 1 MI355X, uneven persistent batch",
        "value": round(conv / wall, 1),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded sizes and parameter draws)",
        "config": {"workload": f"mixed_batch: {n_inst} instances, {len(jobs)} templates of 8 families, "
                               f".op and {circuits.MIXED_STEPS}-step transients",
                   "grid": rep.grid, "lds_bytes": rep.lds_bytes},
        "kernel_ms_per_step": round(kern / args.steps, 3),
        "repivot_ms_per_step": round(rep_ms / args.steps, 3),
        "repivot_launches": rep.repivot_launches,
        "repivots": rep.repivots,
        "setup_ms": round(rep.setup_ms, 1),
        "predicted_ms": round(rep.predicted_ms, 3),
        "block_jobs": rep.block_jobs,
        "wave_jobs": rep.wave_jobs,
        "tuning_run": {"kernel_ms": round(tune_rep.kernel_ms, 3), "block_jobs": tune_rep.block_jobs,
                       "wave_jobs": tune_rep.wave_jobs},
        "build_s": round(t_build, 2),
        "families": fams,
        "parity": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    for e in engines:
        e.close()


if __name__ == "__main__":
    main()
