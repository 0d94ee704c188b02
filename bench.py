#!/usr/bin/env python3
"""Benchmark of the NUFFT-driven NLCG hot path on one MI355X.

Workload (GRASP dynamic): 256x256 image x 34 frames, 12 coils, 21 golden-angle
spokes per frame x 512 readout samples, KB width 6, 2x grid, complex64.

A step is one NUFFT pair over the whole workload: forward A of all 34 frames
into 12-coil k-space, then adjoint A^H back (408 coil-frames each way), with
inputs resident in HBM.  ``value`` = coil-frame NUFFT pairs per second.
Also reported: per-kernel HIP-event times and the roofline of the dominant
kernel, the GPU NLCG time to a 1e-4 relative objective change from the
adjoint start (plain) and from the learned U-Net warm start on held-out
volumes (temporal TV 0.02 + spatial TV 0.002), the U-Net's BF16 MFMA rate, and
the CPU oracle timed on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0     # MI355X dense BF16 MFMA
METRIC = "time-to-1e-4 objective per volume: warm-start vs plain NLCG; NUFFT pairs/s"


def kernel_bytes(wl, plan_info):
    """Algorithmic HBM bytes per launch of each NUFFT kernel (every input and output once)."""
    S, G, C, B = wl.S, plan_info["G1"], wl.n_coils, wl.B
    Ktot = int(wl.k_offsets[-1])
    ncf = B * C
    img = 8 * S * S
    inter = 8 * S * G                   # the S x G row-pass intermediate per coil-frame
    return {
        "fwd_rows": ncf * (img + inter) + B * img + 8 * C * Ktot,        # maps, x, T write, y clear
        "fwd_strips": ncf * inter + 8 * C * Ktot + 12 * Ktot,            # T read, y write, traj+dcf
        "adj_strips": 8 * C * Ktot + 12 * Ktot + ncf * inter,            # y read, traj+dcf, T write
        "adj_rows": ncf * (inter + img) + B * img,                       # T read, maps, x write
    }


def grid_pipeline_pair_bytes(wl, G):
    """Per-pair algorithmic bytes of a gridding pipeline that stores the oversampled G x G grid:
    8 S^2 (1 + 1/C) + 16 S G + 16 G^2 + 8 K + 12 K / C per coil-frame and direction."""
    S, C = wl.S, wl.n_coils
    tot = 0
    for b in range(wl.B):
        K = int(wl.k_offsets[b + 1] - wl.k_offsets[b])
        bdir = 8 * S * S * (1 + 1 / C) + 16 * S * G + 16 * G * G + 8 * K + 12 * K / C
        tot += 2 * bdir * C
    return tot


def load_traffic(kernel, strip_cols, coil_chunk):
    """Measured HBM bytes per launch from the committed rocprofv3 PMC summary, for the strip width and
    coils per strip workgroup the plan uses (the strip kernels' traffic depends on both: each coil group
    re-reads the strip's footprint entries), if present."""
    p = os.path.join(ROOT, "profiles", "nufft_traffic_c2.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            v = json.load(f).get(kernel)
        v = v.get(str(int(strip_cols))) if isinstance(v, dict) else v
        return v.get(str(int(coil_chunk))) if isinstance(v, dict) else v
    except Exception:
        return None


def cpu_baseline(gpu_c1=None, plain_iterations=None):
    """The pure-torch CPU port of the operator and the reference NLCG loop (oracle/torch_cpu.py:
    torch.fft, gathers and index_add_, complex64, FP32 reductions), timed on every core this
    process may use: NUFFT pairs/s at c1, c2 and c3 (one warm-up, median of 3); NLCG to the 1e-4
    relative objective change at c1 in full (PR+ with the descent restart, Armijo that evaluates
    A^H at every trial, as the reference loop does); on the GRASP workload c2 the prime step and 3
    full-volume iterations, measured, and the time to tolerance they imply at the GPU's median
    plain iteration count."""
    import numpy as np
    import torch

    from mrirecon_amd import synth
    from oracle.torch_cpu import TorchKBNufft, nlcg

    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    rates, ops = {}, {}
    for name in ("c1", "c2", "c3"):
        wl = getattr(synth, f"config_{name}")()
        op = TorchKBNufft(wl.maps, wl.traj, wl.k_offsets, wl.dcf, wl.W, wl.osf)
        x = torch.from_numpy(wl.x_true)
        op.adjoint(op.forward(x))
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            op.adjoint(op.forward(x))
            ts.append(time.perf_counter() - t0)
        rates[name] = round(wl.B * wl.n_coils / float(np.median(ts)), 2)
        ops[name] = (wl, op)
        print(f"bench: CPU torch port {name}: {rates[name]} coil-frame pairs/s ({cores} threads)",
              file=sys.stderr, flush=True)

    def problem(name):
        wl, op = ops[name]
        y = op.forward(torch.from_numpy(wl.x_true))
        y = y + torch.from_numpy(synth.complex_noise(tuple(y.shape), 0.01 * float(y.abs().max()), wl.noise_seed))
        regs = [(wl.lam_t, (0,))] * (wl.lam_t > 0) + [(wl.lam_s, (1, 2))] * (wl.lam_s > 0)
        return wl, op, y, regs

    wl, op, y1, regs = problem("c1")
    t0 = time.perf_counter()
    _, h1 = nlcg(op.forward, op.adjoint, y1, op.adjoint(y1), regs, max_iter=200, direction="prplus", restart_c=0.0,
                 rel_tol=1e-4)
    t_c1 = time.perf_counter() - t0
    wl, op, y2, regs = problem("c2")
    t0 = time.perf_counter()
    _, h2 = nlcg(op.forward, op.adjoint, y2, op.adjoint(y2), regs, max_iter=3, direction="prplus", restart_c=0.0)
    t_c2 = time.perf_counter() - t0
    per_it = t_c2 / max(1, len(h2))
    nl = {"c1_time_to_tol_s": round(t_c1, 3), "c1_iterations": len(h1),
          "c2_seconds_for_3_iterations": round(t_c2, 3), "c2_seconds_per_iteration": round(per_it, 3)}
    if gpu_c1:
        nl["c1_gpu_time_to_tol_s"] = gpu_c1["seconds"]
        nl["c1_gpu_iterations"] = gpu_c1["iterations"]
    if plain_iterations:
        nl["c2_time_to_tol_estimate_s"] = round(per_it * plain_iterations, 1)
        nl["c2_estimate"] = (f"measured s/iteration (prime included) x the GPU plain run's median "
                             f"{plain_iterations} iterations")
    return {"value": rates["c2"], "unit": "coil-frame NUFFT pairs/s", "cores": cores, "kind": "port",
            "sample": f"whole GRASP problem (408 coil-frames), median of 3 pairs, torch CPU port, {cores} threads",
            "pairs_per_s": rates, "nlcg": nl}


def gpu_c1_time_to_tol(dev):
    """c1 on the GPU from the same data as the CPU port's c1 run (y = the torch port's forward of the
    phantom + 1 % noise, seed 102): adjoint + NLCG to the 1e-4 relative change (PR+ with restart)."""
    import numpy as np
    import torch

    from mrirecon_amd import NLCGEngine, RegManager, TVND, TVParams, synth
    from mrirecon_amd.nufft.backends.hip import HipNUFFTAdapter
    from oracle.torch_cpu import TorchKBNufft

    wl = synth.config_c1()
    op = TorchKBNufft(wl.maps, wl.traj, wl.k_offsets, wl.dcf, wl.W, wl.osf)
    y = op.forward(torch.from_numpy(wl.x_true))
    y = (y + torch.from_numpy(synth.complex_noise(tuple(y.shape), 0.01 * float(y.abs().max()), wl.noise_seed)))
    be = HipNUFFTAdapter(ndim=2, traj_units="rad", kb_width=wl.W, osf=wl.osf)
    be.prepare((wl.n_coils, wl.S, wl.S), torch.from_numpy(wl.traj), torch.from_numpy(wl.dcf),
               torch.from_numpy(wl.maps).to(dev), device=dev, k_offsets=wl.k_offsets)
    yd = y.to(dev)
    eng = NLCGEngine(be.plan, yd, problems=1, frames=1, regm=RegManager([TVND("s", TVParams(weight=wl.lam_s))]),
                     direction="prplus", rel_tol=1e-4, restart_c=0.0, max_hist=201)
    xa = torch.empty((1, wl.S, wl.S), dtype=torch.complex64, device=dev)
    be.plan.adjoint(yd, xa)
    eng.prime(xa)
    eng.capture()
    ts = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        be.plan.adjoint(yd, xa)
        eng.prime(xa)
        it = eng.run(200)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return {"seconds": round(float(np.median(ts)), 5), "iterations": int(eng.status()[1][0])}


def pregen_phantoms(seeds, workers):
    """GRASP phantoms of the held-out seeds, generated in forked worker processes before this
    process touches the GPU."""
    import multiprocessing as mp

    from mrirecon_amd.train_warmstart import _phantom
    if workers <= 1 or len(seeds) < 4:
        return [_phantom(s) for s in seeds]
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(_phantom, seeds, chunksize=4)


def time_to_tolerance(args, dev, seeds, phantoms):
    """Plain vs warm-started NLCG on held-out GRASP volumes, plus the U-Net's MFMA rate.

    Both runs: PR+ with the beta = 0 descent restart, Armijo (c1 1e-4, halving
    from t = 1, 20 trials), temporal TV 0.02 + spatial TV 0.002, inputs in HBM.
    plain = adjoint + NLCG to |f_k - f_(k-1)| <= 1e-4 |f_(k-1)|; warm = adjoint +
    U-Net + fallback check + NLCG to the same rule; matched = the warm pipeline run
    until it reaches the plain run's final objective (within 1e-4)."""
    import numpy as np
    import torch

    from mrirecon_amd.train_warmstart import GraspProblem, _phantom, evaluate, run_warm
    from mrirecon_amd.warmstart import UNet2DT

    weights = os.path.join(ROOT, "mrirecon_amd", "weights", "unet_grasp.npz")
    print(f"bench: time-to-1e-4 on {args.tol_volumes} held-out GRASP volumes (plain vs warm start)",
          file=sys.stderr, flush=True)
    pb = GraspProblem(max_iter=args.tol_iters)
    net = UNet2DT.load(weights, device=dev)
    pb.measure(phantoms[0], noise_seed=seeds[0] + 7)      # warm-up: workspaces, graphs
    run_warm(pb, net)
    res = evaluate(pb, net, seeds, phantoms, log=lambda *a: print(*a, file=sys.stderr, flush=True))
    rows = res.pop("rows")
    ratio = np.array([r["plain_s"] / r["warm_s"] for r in rows])
    per_volume = [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()} for r in rows]
    ttt = {"volumes": len(rows), "seeds": f"{seeds[0]}..{seeds[-1]} (not in the training set)",
           "solver": "PR+ with beta=0 descent restart, Armijo c1=1e-4 halving from t=1 (20 trials), "
                     "temporal TV 0.02 + spatial TV 0.002 (eps 1e-3), FP64 reductions, HIP-graph iteration",
           "stop_rule": "relative objective change <= 1e-4",
           "plain_ms_median": round(res["median_plain_s"] * 1e3, 3),
           "plain_iterations_median": res["median_plain_iters"],
           "warm_ms_median": round(res["median_warm_s"] * 1e3, 3),
           "warm_iterations_median": res["median_warm_iters"],
           "speedup_median": round(res["speedup_median_of_ratios"], 3),
           "matched_objective_ms_median": round(res["median_matched_s"] * 1e3, 3),
           "matched_objective_iterations_median": res["median_matched_iters"],
           "matched_objective_speedup_median": round(res["matched_speedup_median_of_ratios"], 3),
           "matched_objective_reached_fraction": res["matched_reached_frac"],
           "fallback_rate": res["fallback_rate"],
           "warm_final_objective_rel_to_plain_median": round(res["median_rel_f_gap"], 6),
           "warm_start_objective_over_plain_final_median": round(res["median_f_warm_start_over_plain_final"], 4),
           "network_and_fallback_check_ms_median": round(float(np.median([r["warm_start_ms"] for r in rows])), 3),
           "weights": "mrirecon_amd/weights/unet_grasp.npz",
           "speedup_quartiles": [round(float(q), 3) for q in np.percentile(ratio, [25, 50, 75])],
           "speedup_min_max": [round(float(ratio.min()), 3), round(float(ratio.max()), 3)],
           "volumes_below_3x": int((ratio < 3.0).sum()),
           "warm_final_f_above_plain": int(sum(r["f_warm_final"] > r["f_plain_final"] for r in rows)),
           "iterations_saved_median": float(np.median([r["plain_iters"] - r["warm_iters"] for r in rows])),
           "per_volume": per_volume}
    if not args.no_batched_tol:
        nb = min(16, len(seeds))
        ttt["batched"] = batched_grasp(args, dev, net, seeds[:nb], phantoms[:nb])
    # U-Net forward alone: dense MFMA FLOPs over HIP-event time on the launch stream
    x = pb.x_adj
    for _ in range(3):
        net.infer(x, scales=[1.0])
    reps = 10
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(reps):
        net.infer(x, scales=[1.0])
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    flops = net.flops_per_volume(x.shape[0], x.shape[1], x.shape[2])
    unet = {"flops_per_volume": flops, "forward_ms": round(ms, 4),
            "achieved_tflops": round(flops / (ms * 1e-3) / 1e12, 2), "peak_tflops": BF16_PEAK_TFLOPS,
            "mfma_frac": round(flops / (ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4),
            "shape": f"{x.shape[0]} frames x {x.shape[1]}x{x.shape[2]}, levels {net.levels}, base {net.base}, "
                     f"{sum(L.n_weight for L in net.layers)} weights (BF16, FP32 accumulation)",
            "train_steps": net.step}
    return ttt, unet


def batched_grasp(args, dev, net, seeds, phantoms):
    """The time-to-tolerance volumes as one batched problem set: P GRASP volumes (P x 34 frames,
    each with the single-volume trajectory, maps and noise) in the same launches, each with its own
    line search and stop; stopped volumes skip their NUFFT work.  Wall time from the adjoint to the
    last volume's 1e-4 relative-change stop, plain and warm-started (one U-Net pass over all volumes,
    per-volume fallback check)."""
    import math as _m

    import numpy as np
    import torch

    from mrirecon_amd import NLCGEngine, RegManager, TVND, TVParams, synth
    from mrirecon_amd.nufft.backends.hip import HipNUFFTAdapter
    from mrirecon_amd.train_warmstart import GRASP, spokes_for
    from mrirecon_amd.warmstart import ObjectiveEval, launch_warm_start

    g = GRASP
    P, F, S, C = len(seeds), g["F"], g["S"], g["C"]
    print(f"bench: {P} GRASP volumes batched (plain vs warm start)", file=sys.stderr, flush=True)
    traj, dcf, off = synth.radial_frames(spokes_for(g, 0), g["R"])
    K = int(off[-1])
    off_p = np.concatenate([np.asarray(off[:-1]) + v * K for v in range(P)] + [[P * K]]).astype(np.int64)
    maps = synth.gaussian_maps(C, S).astype(np.complex64)
    be = HipNUFFTAdapter(ndim=2, traj_units="rad", kb_width=g["W"], osf=2.0)
    be.prepare((C, S, S), torch.from_numpy(np.tile(traj, (1, P))), torch.from_numpy(np.tile(dcf, P)),
               torch.from_numpy(maps).to(dev), device=dev, k_offsets=off_p)
    plan = be.plan
    x = torch.from_numpy(np.concatenate(phantoms).astype(np.complex64)).to(dev)
    y = torch.empty(plan.n_k, dtype=torch.complex64, device=dev)
    plan.forward(x, y)
    nk = C * K
    for v, seed in enumerate(seeds):                       # the single-volume noise of each volume
        yv = y[v * nk:(v + 1) * nk]
        sigma = g["noise"] * float(yv.abs().max())
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(seed) + 7)
        nz = torch.randn((2, nk), generator=gen, device=dev, dtype=torch.float32)
        yv += torch.complex(nz[0], nz[1]) * (sigma / _m.sqrt(2.0))
    regm = RegManager([TVND("tv_t", TVParams(weight=g["lam_t"], axes="temporal")),
                       TVND("tv_s", TVParams(weight=g["lam_s"], axes="spatial"))])
    eng = NLCGEngine(plan, y, problems=P, frames=F, regm=regm, direction="prplus", rel_tol=1e-4, restart_c=0.0,
                     max_hist=args.tol_iters + 1)
    xa = torch.empty_like(x)
    plan.adjoint(y, xa)
    eng.prime(xa)
    eng.capture()

    def solve(x0, Ax0=None):
        eng.prime(x0, Ax0)
        return eng.run(args.tol_iters, use_graph=True, check_every=1)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.adjoint(y, xa)
    solve(xa)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    _, it_p = eng.status()
    obj = ObjectiveEval(plan, y, regm, problems=P, frames=F)
    obj_adj = ObjectiveEval(plan, y, regm, problems=P, frames=F)
    x0 = torch.empty_like(x)
    net.infer(xa, nvol=P)                                  # workspace allocation outside the timing
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.adjoint(y, xa)
    ch = launch_warm_start(net, obj, obj_adj, xa, nvol=P, out=x0)
    solve(ch.x0, ch.Ax0)
    torch.cuda.synchronize()
    t_warm = time.perf_counter() - t0
    acc, _, _ = ch.read()
    _, it_w = eng.status()
    return {"volumes": P, "frames": P * F, "plain_seconds": round(t_plain, 4), "warm_seconds": round(t_warm, 4),
            "speedup": round(t_plain / t_warm, 3), "plain_volumes_per_second": round(P / t_plain, 2),
            "warm_volumes_per_second": round(P / t_warm, 2),
            "plain_iterations_median": float(np.median(it_p)), "plain_iterations_max": int(it_p.max()),
            "warm_iterations_median": float(np.median(it_w)), "warm_iterations_max": int(it_w.max()),
            "fallback_rate": float(np.mean(~acc)), "strip_cols": int(plan.info()["strip_cols"]),
            "coil_chunk": int(plan.info()["coil_chunk"])}


def solver_variants(args, dev):
    """Milliseconds per NLCG iteration on one held-out GRASP volume for the solver options beyond the
    default path: PR+ with the restart under Armijo, under strong Wolfe (c2 0.9 and 0.1), with the
    regularisers' diagonal majoriser as preconditioner, and with a chain-graph Laplacian term
    (lambda 0.01) added to the TV terms.  20 iterations from the adjoint image, HIP-graph replay,
    median of 3 runs (the objective after 20 iterations is reported beside each)."""
    import numpy as np
    import torch

    from mrirecon_amd import GraphLaplacian, GraphLapParams, NLCGEngine, RegManager
    from mrirecon_amd.train_warmstart import GraspProblem, _phantom

    print("bench: NLCG solver options, ms per iteration on one GRASP volume", file=sys.stderr, flush=True)
    pb = GraspProblem(max_iter=40)
    seed = 520_000
    pb.measure(_phantom(seed), noise_seed=seed + 7)
    pb.adjoint()
    F = pb.spec["F"]
    W = np.zeros((F, F))
    for t in range(F - 1):
        W[t, t + 1] = W[t + 1, t] = 1.0
    reg_gl = RegManager(pb.regm.regs + [GraphLaplacian("gl", W, GraphLapParams(weight=0.01))])
    variants = {"armijo": dict(), "wolfe_c2_0.9": dict(ls_name="wolfe", c2=0.9),
                "wolfe_c2_0.1": dict(ls_name="wolfe", c2=0.1), "armijo_preconditioned": dict(precond=True),
                "armijo_graph_laplacian": dict(regm=reg_gl)}
    out = {}
    iters = 20
    for name, kw in variants.items():
        kw = dict(kw)
        regm = kw.pop("regm", pb.regm)
        eng = NLCGEngine(pb.plan, pb.y, problems=1, frames=F, regm=regm, direction="prplus", restart_c=0.0,
                         max_hist=iters + 1, **kw)
        eng.prime(pb.x_adj)
        eng.capture()
        times = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            eng.prime(pb.x_adj)
            eng.run(iters, use_graph=True, check_every=iters)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        h = eng.history()[0]
        out[name] = {"ms_per_iteration": round(float(np.median(times)) * 1e3 / max(1, len(h)), 4),
                     "iterations": len(h), "objective": round(h[-1]["f_step"], 4) if h else None}
        eng.close()
    return out


def nufft_rate(wl, dev, reps=5):
    """Coil-frame NUFFT pairs/s of one workload on an autotuned plan (inputs in HBM)."""
    import numpy as np
    import torch

    from mrirecon_amd.nufft.backends.hip import HipNUFFTAdapter
    be = HipNUFFTAdapter(ndim=2, traj_units="rad", kb_width=wl.W, osf=wl.osf)
    t0 = time.perf_counter()
    be.prepare((wl.n_coils, wl.S, wl.S), torch.from_numpy(wl.traj), torch.from_numpy(wl.dcf),
               torch.from_numpy(wl.maps).to(dev), device=dev, k_offsets=wl.k_offsets)
    setup = plan_setup(be, time.perf_counter() - t0)
    plan = be.plan
    x = torch.from_numpy(wl.x_true).to(dev)
    y = torch.empty(plan.n_k, dtype=torch.complex64, device=dev)
    plan.forward(x, y)
    plan.adjoint(y, x)
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        plan.forward(x, y)
        plan.adjoint(y, x)
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    ms = float(np.median(ts))
    ncf = wl.B * wl.n_coils
    return be, {"coil_frames": ncf, "ms_per_whole_pair": round(ms, 4),
                "coil_frame_pairs_per_s": round(ncf / (ms * 1e-3), 1),
                "strip_cols": int(plan.info()["strip_cols"]),
                "coil_chunk": plan.tuning["coil_chunk"] if plan.tuning else None,
                "autotune_ms_per_pair": ({str(w): {str(k): round(v, 4) for k, v in t.items()}
                                          for w, t in plan.tuning["ms_per_pair_by_width"].items()}
                                         if plan.tuning and "ms_per_pair_by_width" in plan.tuning else
                                         {str(k): round(v, 4) for k, v in plan.tuning["ms_per_pair"].items()}
                                         if plan.tuning else None),
                "plan_setup": setup}


def plan_setup(be, seconds):
    """Host-side plan cost of an adapter: prepare() wall time, and per strip width the C-ABI plan
    creation (strip binning on host threads + uploads), and the coil-chunk autotune over all widths."""
    plan = be.plan
    t = plan.tuning or {}
    info = plan.info()
    return {"prepare_ms": round(seconds * 1e3, 1), "samples": int(info["n_samples"]),
            "footprint_entries": int(info["n_entries"]),
            "plan_create_ms_by_width": t.get("plan_create_ms_by_width", {int(info["strip_cols"]): round(plan.create_ms, 2)}),
            "autotune_ms": t.get("autotune_ms")}


def nufft3_rate(dev, S=128, C=8, spokes=1024, R=256, reps=5):
    """3-D NUFFT pair on a kooshball trajectory: S^3 volume, C Gaussian-like coil maps, `spokes`
    diametric spokes of R samples with directions on a Fibonacci sphere, KB W = 6, osf 2, no DCF.
    Coil-volume pairs/s with inputs in HBM (HIP events on the launch stream)."""
    import math as _m

    import numpy as np
    import torch

    from mrirecon_amd.nufft.backends.hip import HipNUFFTAdapter
    i = np.arange(spokes) + 0.5
    phi = np.arccos(1.0 - 2.0 * i / spokes)
    theta = _m.pi * (1.0 + 5.0 ** 0.5) * i
    dirs = np.stack([np.sin(phi) * np.cos(theta), np.sin(phi) * np.sin(theta), np.cos(phi)])   # (3, spokes)
    rad = (np.arange(R) - R / 2) / (R / 2) * _m.pi * 0.999
    traj = (dirs[:, :, None] * rad[None, None, :]).reshape(3, -1).astype(np.float32)
    ax = (np.arange(S) - S / 2) / S
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    maps = np.stack([np.exp(-((X - 0.3 * _m.cos(2 * _m.pi * c / C)) ** 2 + (Y - 0.3 * _m.sin(2 * _m.pi * c / C)) ** 2
                              + Z ** 2) / 0.2) for c in range(C)]).astype(np.complex64)
    maps /= np.sqrt((np.abs(maps) ** 2).sum(axis=0, keepdims=True))
    be = HipNUFFTAdapter(ndim=3, traj_units="rad", kb_width=6, osf=2.0, dcf_mode="none")
    be.prepare((C, S, S, S), torch.from_numpy(traj), None, torch.from_numpy(maps).to(dev), device=dev)
    plan = be.plan
    x = torch.ones((1, S, S, S), dtype=torch.complex64, device=dev)
    y = torch.empty(plan.n_k, dtype=torch.complex64, device=dev)
    plan.forward(x, y)
    plan.adjoint(y, x)
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        plan.forward(x, y)
        plan.adjoint(y, x)
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    ms = float(np.median(ts))
    be.close()
    # algorithmic bytes of this grid-in-HBM pipeline, every kernel's inputs and outputs once:
    # forward = volume + maps + pad write + 3 line-FFT passes (read + write) + interpolation read
    # + trajectory + y; the adjoint mirrors it (y, spread write, 3 passes, crop read, maps, volume)
    G = 2 * S
    grid = 8 * C * G ** 3
    K = int(traj.shape[1])
    one_way = 8 * S ** 3 * (1 + C) + grid * (1 + 6 + 1) + 12 * K + 8 * C * K
    bytes_pair = 2 * one_way
    return {"volume": f"{S}^3", "coils": C, "samples": K, "ms_per_pair": round(ms, 3),
            "coil_volume_pairs_per_s": round(C / (ms * 1e-3), 1),
            "grid_model_bytes_per_pair": int(bytes_pair),
            "grid_model_GBps": round(bytes_pair / (ms * 1e-3) / 1e9, 1),
            "grid_model_frac": round(bytes_pair / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def tbasis_rate(dev, T=34, K=8, S=256, vols=1, reps=20):
    """Temporal-basis map u = U V and its adjoint g_V = U^H g_u at GRASP size (T frames, K basis
    images of S^2 for each of ``vols`` volumes sharing U): ms per call (HIP events on the launch
    stream) and GB/s of algorithmic bytes, (K + T) vols S^2 complex64 values per call."""
    import numpy as np
    import torch

    from mrirecon_amd.mapping import TemporalBasisOp
    g = torch.Generator(device="cpu")
    g.manual_seed(3)
    op = TemporalBasisOp(torch.randn((T, K), generator=g, dtype=torch.complex64), device=dev)
    V = torch.randn((K, vols, S, S), generator=g, dtype=torch.complex64).to(dev)
    u = torch.empty((T, vols, S, S), dtype=torch.complex64, device=dev)
    gV = torch.empty_like(V)
    res = {"frames": T, "basis": K, "image": f"{S}^2", "volumes": vols}
    for name, call in (("forward", lambda: op.forward_apply(V, out=u)),
                       ("adjoint", lambda: op.adjoint_apply(u, gV, accumulate=False))):
        call()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            call()
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res[f"{name}_ms"] = round(ms, 4)
        res[f"{name}_GBps"] = round((K + T) * vols * S * S * 8 / (ms * 1e-3) / 1e9, 1)
    return res


def c3_precision(dev, iters=20):
    """Precision study on c3 (384^2, 16 coils, 402 golden-angle spokes x 768, W = 6, x ~ CN(0, 1)
    seed 301): the FP32 NUFFT pair against the FP64 oracle (relative L2, max-abs over max), and
    NLCG (PR+ with the descent restart, Armijo, spatial TV 0.01, y = A x + 1 % noise) for 20
    iterations with the engine's FP64 inner products / beta / line-search values against the
    all-FP32 variant (the reference's precision, graspcg/solvers/cg.py:89-90,
    graspcg/ops/dot.py:54-71), both final objectives evaluated in FP64 on the device."""
    import numpy as np
    import torch

    from mrirecon_amd import NLCGEngine, RegManager, TVND, TVParams, synth
    from mrirecon_amd.nufft.backends.hip import HipNUFFTAdapter
    from mrirecon_amd.warmstart import ObjectiveEval
    from oracle.kbnufft import KBNufft, rel_l2

    print("bench: c3 precision study (operator vs FP64 oracle, NLCG FP64 vs FP32 reductions)", file=sys.stderr,
          flush=True)
    wl = synth.config_c3()
    be = HipNUFFTAdapter(ndim=2, traj_units="rad", kb_width=wl.W, osf=wl.osf)
    be.prepare((wl.n_coils, wl.S, wl.S), torch.from_numpy(wl.traj), torch.from_numpy(wl.dcf),
               torch.from_numpy(wl.maps).to(dev), device=dev, k_offsets=wl.k_offsets)
    plan = be.plan
    x = torch.from_numpy(wl.x_true).to(dev)
    y = torch.empty(plan.n_k, dtype=torch.complex64, device=dev)
    plan.forward(x, y)
    rng = np.random.default_rng(303)
    yr = (rng.standard_normal(plan.n_k) + 1j * rng.standard_normal(plan.n_k)).astype(np.complex64)
    xa = torch.empty_like(x)
    plan.adjoint(torch.from_numpy(yr).to(dev), xa)
    ora = KBNufft(wl.maps, wl.traj, wl.k_offsets, wl.dcf, wl.W, wl.osf)
    yf, xr = ora.forward(wl.x_true), ora.adjoint(yr)
    y_np, xa_np = y.cpu().numpy(), xa.cpu().numpy()
    op = {"forward_rel_l2": float(rel_l2(y_np, yf)), "adjoint_rel_l2": float(rel_l2(xa_np, xr)),
          "forward_max_abs_over_max": float(np.abs(y_np - yf).max() / np.abs(yf).max()),
          "adjoint_max_abs_over_max": float(np.abs(xa_np - xr).max() / np.abs(xr).max())}
    g = torch.Generator(device=dev)
    g.manual_seed(302)
    nz = torch.randn((2, plan.n_k), generator=g, device=dev, dtype=torch.float32)
    y += torch.complex(nz[0], nz[1]) * (0.01 * float(y.abs().max()) / math.sqrt(2.0))
    regm = RegManager([TVND("tv_s", TVParams(weight=0.01, axes="spatial"))])
    x0 = torch.empty_like(x)
    plan.adjoint(y, x0)
    fev = ObjectiveEval(plan, y, regm, problems=1, frames=1)
    res = {}
    for name, fp32 in (("fp64", False), ("fp32", True)):
        eng = NLCGEngine(plan, y, problems=1, frames=1, regm=regm, direction="prplus", restart_c=0.0,
                         max_hist=iters + 1, fp32_reductions=fp32)
        eng.prime(x0)
        eng.capture()
        eng.prime(x0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.run(iters)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        h = eng.history()[0]
        res[name] = {"objective": float(fev(eng.x)[0]), "iterations": len(h), "steps": [r["step"] for r in h],
                     "ms_per_iteration": round(dt * 1e3 / max(1, len(h)), 4)}
        eng.close()
    f64, f32 = res["fp64"]["objective"], res["fp32"]["objective"]
    same = 0
    for a, b in zip(res["fp64"]["steps"], res["fp32"]["steps"]):
        if a != b:
            break
        same += 1
    return {"operator": {k: float(f"{v:.3e}") for k, v in op.items()},
            "nlcg": {"iterations": iters, "objective_fp64_reductions": round(f64, 6),
                     "objective_fp32_reductions": round(f32, 6), "rel_gap_fp32_minus_fp64": float(f"{(f32 - f64) / f64:.3e}"),
                     "identical_leading_steps": same, "ms_per_iteration_fp64": res["fp64"]["ms_per_iteration"],
                     "ms_per_iteration_fp32": res["fp32"]["ms_per_iteration"]}}


def batched_uneven(args, dev):
    """Batched uneven workload: P slices x 20 frames at 256^2, 8 coils, 13..55 golden-angle spokes
    per frame x 512, all slices in the same launches (ragged k-space, per-slice stop), temporal TV 0.02.
    Reports NUFFT throughput and plain / warm-started time to the 1e-4 relative-change stop."""
    import math as _m

    import numpy as np
    import torch

    from mrirecon_amd import NLCGEngine, RegManager, TVND, TVParams, synth
    from mrirecon_amd.warmstart import ObjectiveEval, UNet2DT, launch_warm_start

    print(f"bench: batched uneven workload ({args.c4_slices} slices x 20 frames)", file=sys.stderr, flush=True)
    wl = synth.config_c4(n_slices=args.c4_slices)
    be, rate = nufft_rate(wl, dev, reps=3)
    plan = be.plan
    P, F = wl.n_slices, wl.n_frames
    x = torch.from_numpy(wl.x_true).to(dev)
    y = torch.empty(plan.n_k, dtype=torch.complex64, device=dev)
    plan.forward(x, y)
    sigma = 0.01 * float(y.abs().max())
    gen = torch.Generator(device=dev)
    gen.manual_seed(wl.noise_seed)
    nz = torch.randn((2, y.numel()), generator=gen, device=dev, dtype=torch.float32)
    y += torch.complex(nz[0], nz[1]) * (sigma / _m.sqrt(2.0))
    regm = RegManager([TVND("tv_t", TVParams(weight=wl.lam_t, axes="temporal"))])
    eng = NLCGEngine(plan, y, problems=P, frames=F, regm=regm, direction="prplus", rel_tol=1e-4, restart_c=0.0,
                     max_hist=args.tol_iters + 1)
    xa = torch.empty_like(x)
    plan.adjoint(y, xa)
    eng.prime(xa)
    eng.capture()

    def solve(x0, Ax0=None):
        eng.prime(x0, Ax0)
        return eng.run(args.tol_iters, use_graph=True, check_every=1)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.adjoint(y, xa)
    solve(xa)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    st_p, it_p = eng.status()
    h_p = eng.history()
    f_plain = np.array([rows[-1]["f_step"] for rows in h_p])
    out = {"workload": f"{P} slices x {F} frames, 256x256, 8 coils, 13..55 spokes/frame x 512 (seed 401), "
                       f"temporal TV 0.02", "nufft": rate, "plain_seconds": round(t_plain, 4),
           "plain_iterations_median": float(np.median(it_p)), "plain_iterations_max": int(it_p.max()),
           "plain_slices_per_second": round(P / t_plain, 2)}
    wdir = os.path.join(ROOT, "mrirecon_amd", "weights")
    own = os.path.exists(os.path.join(wdir, "unet_batched.npz"))
    weights = os.path.join(wdir, "unet_batched.npz" if own else "unet_grasp.npz")
    net = UNet2DT.load(weights, device=dev)
    obj = ObjectiveEval(plan, y, regm, problems=P, frames=F)
    obj_adj = ObjectiveEval(plan, y, regm, problems=P, frames=F)
    x0 = torch.empty_like(x)
    net.infer(xa, nvol=P)                                  # workspace allocation outside the timing
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.adjoint(y, xa)
    ch = launch_warm_start(net, obj, obj_adj, xa, nvol=P, out=x0)
    solve(ch.x0, ch.Ax0)
    torch.cuda.synchronize()
    t_warm = time.perf_counter() - t0
    acc, f_w, f_a = ch.read()
    st_w, it_w = eng.status()
    h_w = eng.history()
    f_warm = np.array([rows[-1]["f_step"] for rows in h_w])
    out.update({"warm_seconds": round(t_warm, 4), "warm_iterations_median": float(np.median(it_w)),
                "warm_iterations_max": int(it_w.max()), "warm_slices_per_second": round(P / t_warm, 2),
                "speedup": round(t_plain / t_warm, 3), "fallback_rate": float(np.mean(~acc)),
                "warm_final_objective_rel_to_plain_median": round(float(np.median((f_warm - f_plain) / f_plain)), 6),
                "weights": os.path.relpath(weights, ROOT),
                "note": ("U-Net trained on single slices of this workload (20 frames, 8 coils, 13..55 spokes per "
                         "frame, temporal TV) with spoke patterns and phantoms disjoint from these slices"
                         if own else "U-Net trained on the GRASP workload (34 frames, 12 coils, 21 spokes, "
                                     "spatial+temporal TV), applied unchanged, with the per-slice fallback")})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--strip-cols", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-tol", action="store_true", help="skip the time-to-tolerance NLCG run")
    ap.add_argument("--tol-iters", type=int, default=200)
    ap.add_argument("--tol-volumes", type=int, default=200,
                    help="held-out volumes of the time-to-tolerance runs (seeds 600000 + i)")
    ap.add_argument("--workers", type=int, default=16, help="processes generating the held-out phantoms")
    ap.add_argument("--detail", default="bench_detail.json",
                    help="file for the full results (per-volume rows, autotuner times, ...)")
    ap.add_argument("--no-batched-tol", action="store_true",
                    help="skip the run of the same volumes as one batched problem set")
    ap.add_argument("--c4-slices", type=int, default=48, help="slices of the batched uneven workload")
    ap.add_argument("--no-c4", action="store_true", help="skip the batched uneven workload")
    ap.add_argument("--no-configs", action="store_true", help="skip the c1 / c3 NUFFT rates")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the per-iteration timing of the NLCG solver options")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (reference for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding of "
              "frames / slices is a design note, not built in this tree.", file=sys.stderr)
        sys.exit(2)

    tol_seeds = list(range(600_000, 600_000 + args.tol_volumes))
    phantoms = None
    if not args.no_tol and not args.copy_probe_mib:
        print(f"bench: generating {len(tol_seeds)} held-out GRASP phantoms", file=sys.stderr, flush=True)
        phantoms = pregen_phantoms(tol_seeds, min(args.workers, os.cpu_count() or 1))

    import numpy as np
    import torch

    from mrirecon_amd import synth
    from mrirecon_amd.nufft.backends.hip import HipNUFFTAdapter

    dev = torch.device("cuda:0")
    if args.copy_probe_mib:
        n = args.copy_probe_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": n * 4}), flush=True)
        return
    print("bench: generating the GRASP workload", file=sys.stderr, flush=True)
    wl = synth.config_c2()
    be = HipNUFFTAdapter(ndim=2, traj_units="rad", kb_width=wl.W, osf=wl.osf, strip_cols=args.strip_cols)
    t0 = time.perf_counter()
    be.prepare((wl.n_coils, wl.S, wl.S), torch.from_numpy(wl.traj), torch.from_numpy(wl.dcf),
               torch.from_numpy(wl.maps).to(dev), device=dev, k_offsets=wl.k_offsets)
    setup_c2 = plan_setup(be, time.perf_counter() - t0)
    plan = be.plan
    info = plan.info()
    x = torch.from_numpy(wl.x_true).to(dev)
    y = torch.empty(plan.n_k, dtype=torch.complex64, device=dev)
    xa = torch.empty_like(x)
    ncf = wl.B * wl.n_coils

    def pair(ev=None):
        if ev is None:
            plan.forward(x, y)
            plan.adjoint(y, xa)
            return
        ev[0].record()
        plan.forward_stage(0, x, y)
        ev[1].record()
        plan.forward_stage(1, x, y)
        ev[2].record()
        plan.adjoint_stage(0, y, xa)
        ev[3].record()
        plan.adjoint_stage(1, y, xa)
        ev[4].record()

    print(f"bench: {args.warmup} warm-up + {args.steps} timed NUFFT pairs", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        pair()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        pair(evs[k])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    names = ["fwd_rows", "fwd_strips", "adj_strips", "adj_rows"]
    stage_ms = {n: float(np.mean([e[i].elapsed_time(e[i + 1]) for e in evs])) for i, n in enumerate(names)}
    ms_per_step = wall / args.steps * 1e3
    value = ncf * args.steps / wall

    kb = kernel_bytes(wl, info)
    dom = max(stage_ms, key=stage_ms.get)
    achieved = kb[dom] / (stage_ms[dom] * 1e-3) / 1e9
    traffic = load_traffic(dom, info["strip_cols"], info["coil_chunk"])
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes": int(kb[dom]), "avg_ms": round(stage_ms[dom], 4)}
    per_kernel = {n: {"avg_ms": round(stage_ms[n], 4), "algorithmic_GBps": round(kb[n] / (stage_ms[n] * 1e-3) / 1e9, 1)}
                  for n in names}
    pair_ms = sum(stage_ms.values())
    model = grid_pipeline_pair_bytes(wl, info["G1"])
    pipeline = {"grid_model_bytes_per_pair": int(model), "kernel_sum_ms": round(pair_ms, 4),
                "grid_model_GBps": round(model / (pair_ms * 1e-3) / 1e9, 1),
                "grid_model_frac": round(model / (pair_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "whole_problem_pairs_per_s": round(1e3 / ms_per_step, 2)}

    ttt, unet = None, None
    if not args.no_tol:
        ttt, unet = time_to_tolerance(args, dev, tol_seeds, phantoms)

    configs = None
    if not args.no_configs:
        print("bench: NUFFT rates of the single-frame configurations", file=sys.stderr, flush=True)
        configs = {"c1 (128^2, 8 coils, 64 spokes x 256, W=4)": nufft_rate(synth.config_c1(), dev)[1],
                   "c3 (384^2, 16 coils, 402 spokes x 768, W=6)": nufft_rate(synth.config_c3(), dev)[1],
                   "3-D (128^3, 8 coils, 1024 kooshball spokes x 256, W=6)": nufft3_rate(dev)}
    c3p = None if args.no_configs else c3_precision(dev)
    c4 = None
    if not args.no_c4:
        c4 = batched_uneven(args, dev)
    variants, tbasis = None, None
    if not args.no_variants:
        variants = solver_variants(args, dev)
        tbasis = {"one volume": tbasis_rate(dev), "16 volumes": tbasis_rate(dev, vols=16)}

    cpu = None
    if not args.no_cpu:
        print("bench: CPU baseline (torch CPU port) and c1 NLCG on the GPU", file=sys.stderr, flush=True)
        c1_gpu = gpu_c1_time_to_tol(dev)
        cpu = cpu_baseline(c1_gpu, ttt["plain_iterations_median"] if ttt else None)

    detail = {"roofline": roofline, "kernels": per_kernel, "pipeline": pipeline, "plan_setup": setup_c2,
              "time_to_tol": ttt, "unet": unet,
              "nufft_configs": configs, "c3_precision": c3p, "batched_uneven": c4, "solver_variants": variants,
              "temporal_basis": tbasis, "cpu_baseline": cpu}
    try:
        os.makedirs(os.path.dirname(os.path.abspath(args.detail)), exist_ok=True)
        with open(args.detail, "w") as fh:
            json.dump(detail, fh, indent=1)
    except OSError as e:
        print(f"bench: could not write {args.detail}: {e}", file=sys.stderr)
    print("bench detail: " + json.dumps({k: v for k, v in detail.items() if k != "time_to_tol"}), file=sys.stderr)
    ttt_line = None
    if ttt:
        ttt_line = {k: ttt[k] for k in ("volumes", "plain_ms_median", "plain_iterations_median", "warm_ms_median",
                                        "warm_iterations_median", "speedup_median", "speedup_quartiles",
                                        "speedup_min_max", "volumes_below_3x", "matched_objective_speedup_median",
                                        "matched_objective_reached_fraction", "fallback_rate",
                                        "iterations_saved_median", "warm_final_f_above_plain",
                                        "network_and_fallback_check_ms_median")}
        if ttt.get("batched"):
            ttt_line["batched16_speedup"] = ttt["batched"]["speedup"]
    cpu_line = None
    if cpu:
        cpu_line = dict(cpu)
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "coil-frame NUFFT pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded GRASP phantom, golden-angle radial, Gaussian coil maps)",
        "config": {"workload": "GRASP dynamic 256x256 x 34 frames, 12 coils, 21 golden-angle spokes x 512 "
                               "samples per frame, KB W=6, osf 2, complex64",
                   "coil_frames": ncf, "samples_per_frame": int(wl.uniform_K), "parallelism": "1 GPU"},
        "roofline": {k: roofline[k] for k in ("bound", "kernel", "achieved", "peak", "unit", "frac", "traffic")},
        "cpu_baseline": cpu_line,
        "kernels_ms": {n: per_kernel[n]["avg_ms"] for n in names},
        "time_to_tol": ttt_line,
        "unet_mfma_frac": unet["mfma_frac"] if unet else None,
        "batched_uneven_speedup": c4["speedup"] if c4 else None,
        "c3_fp32_vs_fp64_objective_gap": c3p["nlcg"]["rel_gap_fp32_minus_fp64"] if c3p else None,
        "plan_setup_ms": {"grasp": setup_c2["prepare_ms"],
                          "batched48": c4["nufft"]["plan_setup"]["prepare_ms"] if c4 else None},
        "detail": args.detail,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
