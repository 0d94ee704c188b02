"""Train and evaluate the U-Net warm start on the GRASP or the batched uneven workload.

    python -m mrirecon_amd.train_warmstart --train 2000 --test 200 --epochs 2 \
        --out unet_grasp.npz --log unet_grasp.json [--workload batched]

Data: seeded dynamic phantoms of the GRASP generator (256^2 x 34 frames,
12 coils, 21 golden-angle spokes x 512 samples per frame; or one slice of the
batched workload: 20 frames, 8 coils, 13..55 spokes per frame drawn from a pool
of seeded patterns), complex Gaussian noise of 1 % of max |y|.  Network input: the density-compensated adjoint image
A^H y.  Target: the reconstruction of a fixed-length PR+ NLCG run (150
iterations by default) from that adjoint start (temporal TV 0.02 + spatial TV
0.002), with the beta = 0 restart whenever the next direction would not descend
(without it PR+ stops on a non-descent direction after a few iterations).
Training: Adam with a cosine learning rate, one volume per step, MSE in the
per-volume scaled domain plus, optionally, the NLCG objective at the network
output relative to the target's (its gradient A^H (A x - y) + grad TV comes from
the same kernels as the solver's), every pass on the library's BF16 MFMA kernels.  Evaluation on held-out seeds: wall time and
iterations of NLCG to the relative-change stop (1e-4) from the adjoint start
(plain) and from the fallback-checked warm start (adjoint + network + objective
check included in the warm time), and the time the warm start needs to reach
the plain run's final objective (within 1e-4).
"""
from __future__ import annotations

import argparse
import json
import math
import multiprocessing as mp
import sys
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

# GRASP: 256^2 x 34 frames, 12 coils, 21 spokes x 512 per frame, temporal + spatial TV.
# batched: one slice of the batched uneven workload, 256^2 x 20 frames, 8 coils,
# 13..55 spokes x 512 per frame (drawn per frame), temporal TV only.
WORKLOADS = {
    "grasp": dict(S=256, F=34, C=12, R=512, spokes=(21, 21), W=6, lam_t=0.02, lam_s=0.002, noise=0.01),
    "batched": dict(S=256, F=20, C=8, R=512, spokes=(13, 55), W=6, lam_t=0.02, lam_s=0.0, noise=0.01),
}
GRASP = WORKLOADS["grasp"]


def _phantom(seed: int) -> np.ndarray:
    from . import synth
    return synth.dynamic_phantom(GRASP["S"], GRASP["F"], seed=seed)


def _phantom_sf(args) -> np.ndarray:
    from . import synth
    S, F, seed = args
    return synth.dynamic_phantom(S, F, seed=seed)


def spokes_for(spec: dict, seed: int) -> List[int]:
    """Spokes per frame: fixed, or drawn uniformly from spec['spokes'] (seeded)."""
    lo, hi = spec["spokes"]
    if lo == hi:
        return [lo] * spec["F"]
    return [int(v) for v in np.random.default_rng(seed).integers(lo, hi + 1, size=spec["F"])]


class RadialProblem:
    """One volume of a radial workload on the device: plan, measurement buffer, NLCG engines, objective."""

    def __init__(self, spec: dict, spokes: Sequence[int], device="cuda:0", rel_tol: float = 1e-4,
                 max_iter: int = 200, restart_c: Optional[float] = 0.0, long_iters: int = 150,
                 autotune: bool = True):
        import torch

        from . import NLCGEngine, RegManager, TVND, TVParams, synth
        from .nufft.backends.hip import HipNUFFTAdapter
        from .warmstart import ObjectiveEval

        g = self.spec = spec
        self.torch = torch
        self.dev = torch.device(device)
        traj, dcf, off = synth.radial_frames(list(spokes), g["R"])
        maps = synth.gaussian_maps(g["C"], g["S"]).astype(np.complex64)
        be = HipNUFFTAdapter(ndim=2, traj_units="rad", kb_width=g["W"], osf=2.0, autotune=autotune)
        be.prepare((g["C"], g["S"], g["S"]), torch.from_numpy(traj), torch.from_numpy(dcf),
                   torch.from_numpy(maps).to(self.dev), device=self.dev, k_offsets=off)
        self.adapter = be
        self.plan = be.plan
        self.y = torch.zeros(self.plan.n_k, dtype=torch.complex64, device=self.dev)
        self.x_true = torch.zeros((g["F"], g["S"], g["S"]), dtype=torch.complex64, device=self.dev)
        self.x_adj = torch.zeros_like(self.x_true)
        regs = [TVND("tv_t", TVParams(weight=g["lam_t"], axes="temporal"))]
        if g["lam_s"]:
            regs.append(TVND("tv_s", TVParams(weight=g["lam_s"], axes="spatial")))
        self.regm = RegManager(regs)
        self.max_iter = int(max_iter)
        self.long_iters = int(long_iters)
        # stop rule run (relative objective change) and fixed-length / target run (no change rule)
        self.eng = NLCGEngine(self.plan, self.y, problems=1, frames=g["F"], regm=self.regm, rel_tol=rel_tol,
                              direction="prplus", max_hist=self.max_iter + 1, restart_c=restart_c)
        self.eng_long = NLCGEngine(self.plan, self.y, problems=1, frames=g["F"], regm=self.regm, rel_tol=0.0,
                                   direction="prplus", max_hist=max(self.long_iters, self.max_iter) + 1,
                                   restart_c=restart_c)
        self.objective = ObjectiveEval(self.plan, self.y, self.regm, problems=1, frames=g["F"])
        self.objective_adj = ObjectiveEval(self.plan, self.y, self.regm, problems=1, frames=g["F"])
        self.side = torch.cuda.Stream(device=self.dev)
        self.x0 = torch.zeros_like(self.x_true)
        for e in (self.eng, self.eng_long):
            e.prime(self.x_adj)
            e.capture()

    def measure(self, x_true: np.ndarray, noise_seed: int) -> None:
        """y = A x_true + CN(0, sigma^2), sigma = 1 % of max |A x_true| (seeded)."""
        torch = self.torch
        self.x_true.copy_(torch.from_numpy(x_true))
        self.plan.forward(self.x_true, self.y)
        sigma = self.spec["noise"] * float(self.y.abs().max())
        gen = torch.Generator(device=self.dev)
        gen.manual_seed(int(noise_seed))
        n = torch.randn((2, self.y.numel()), generator=gen, device=self.dev, dtype=torch.float32)
        self.y += torch.complex(n[0], n[1]) * (sigma / math.sqrt(2.0))

    def adjoint(self, out=None):
        out = self.x_adj if out is None else out
        self.plan.adjoint(self.y, out)
        return out

    def solve(self, x0, long: bool = False, iters: Optional[int] = None, Ax0=None) -> int:
        """NLCG from x0: to the relative-change stop (``long`` False) or for ``iters``
        iterations / until the engine's objective target (``long`` True)."""
        eng = self.eng_long if long else self.eng
        eng.prime(x0, Ax0)
        return eng.run(iters or (self.long_iters if long else self.max_iter), use_graph=True, check_every=1)

    def history(self, long: bool = False):
        return (self.eng_long if long else self.eng).history()[0]


class GraspProblem(RadialProblem):
    """One GRASP volume (fixed trajectory: 21 spokes in each of 34 frames)."""

    def __init__(self, device="cuda:0", rel_tol: float = 1e-4, max_iter: int = 200,
                 restart_c: Optional[float] = 0.0, long_iters: int = 150):
        super().__init__(GRASP, spokes_for(GRASP, 0), device=device, rel_tol=rel_tol, max_iter=max_iter,
                         restart_c=restart_c, long_iters=long_iters)


def _sync(torch):
    torch.cuda.synchronize()


def run_plain(pb: GraspProblem) -> Dict[str, float]:
    torch = pb.torch
    _sync(torch)
    t0 = time.perf_counter()
    pb.adjoint()
    pb.solve(pb.x_adj)
    _sync(torch)
    t = time.perf_counter() - t0
    h = pb.history()
    return {"seconds": t, "iterations": len(h), "f_final": h[-1]["f_step"] if h else float("nan"),
            "f_start": h[0]["f"] if h else float("nan"), "f_path": [r["f_step"] for r in h],
            "stop": pb.eng.stop_reasons()[0]}


def run_warm(pb: GraspProblem, net, target: Optional[float] = None) -> Dict[str, float]:
    """Adjoint -> network -> fallback check -> NLCG, all inside the timed region.

    ``target`` None: NLCG to the relative-change stop.  Otherwise: NLCG without
    the change rule until f <= target (or the long-run iteration cap)."""
    from .warmstart import launch_warm_start
    torch = pb.torch
    eng = pb.eng_long if target is not None else pb.eng
    eng.set_targets(None if target is None else [target])
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]      # stage times without host syncs
    _sync(torch)
    t0 = time.perf_counter()
    ev[0].record()
    pb.adjoint()
    ev[1].record()
    ch = launch_warm_start(net, pb.objective, pb.objective_adj, pb.x_adj, out=pb.x0, side_stream=pb.side)
    ev[2].record()
    pb.solve(ch.x0, long=target is not None, iters=pb.max_iter if target is not None else None, Ax0=ch.Ax0)
    _sync(torch)
    t = time.perf_counter() - t0
    acc, f_w, f_a = ch.read()
    h = pb.history(long=target is not None)
    stop = eng.stop_reasons()[0]
    eng.set_targets(None)
    # the fallback check may have decided by a lower bound of f(x_adj); the exact value is
    # evaluated here, outside the timed region, for the report
    f_adj = float(pb.objective_adj(pb.x_adj)[0])
    return {"seconds": t, "iterations": len(h), "f_final": h[-1]["f_step"] if h else float("nan"),
            "f_start": float(f_w[0]), "f_adjoint": f_adj, "f_adjoint_used": float(f_a[0]),
            "decided_by_bound": bool(f_a[0] < f_adj), "accepted": bool(acc[0]),
            "f_path": [r["f_step"] for r in h], "stop": stop,
            "adjoint_ms": ev[0].elapsed_time(ev[1]), "warm_start_ms": ev[1].elapsed_time(ev[2])}


def iterations_to_reach(f_path: Sequence[float], target: float) -> Optional[int]:
    for k, f in enumerate(f_path):
        if f <= target:
            return k + 1
    return None


def evaluate(pb, net, seeds: Sequence[int], phantoms, log=print) -> Dict[str, object]:
    """Plain vs warm-started NLCG on held-out phantoms (same data, same stop rule).
    ``pb``: one RadialProblem for every phantom, or a list with one per phantom."""
    rows = []
    pbs = pb if isinstance(pb, (list, tuple)) else [pb] * len(seeds)
    for pb, seed, x_true in zip(pbs, seeds, phantoms):
        pb.measure(x_true, noise_seed=seed + 7)
        p = run_plain(pb)
        w = run_warm(pb, net)
        target = p["f_final"] * (1 + 1e-4)
        m = run_warm(pb, net, target=target)
        rows.append({"seed": int(seed), "plain_s": p["seconds"], "plain_iters": p["iterations"],
                     "warm_s": w["seconds"], "warm_iters": w["iterations"], "accepted": w["accepted"],
                     "matched_s": m["seconds"], "matched_iters": m["iterations"],
                     "matched_reached": m["stop"] == "target",
                     "f_plain_final": p["f_final"], "f_warm_final": w["f_final"], "f_adjoint": w["f_adjoint"],
                     "f_warm_start": w["f_start"], "warm_start_ms": w["warm_start_ms"],
                     "adjoint_ms": w["adjoint_ms"], "warm_iters_to_plain_f": iterations_to_reach(w["f_path"], target),
                     "plain_stop": p["stop"], "warm_stop": w["stop"]})
        r = rows[-1]
        log(f"eval seed {seed}: plain {r['plain_iters']} it {r['plain_s'] * 1e3:.1f} ms | warm {r['warm_iters']} it "
            f"{r['warm_s'] * 1e3:.1f} ms (network + check {r['warm_start_ms']:.2f} ms, accepted {r['accepted']}) | matched "
            f"{r['matched_iters']} it {r['matched_s'] * 1e3:.1f} ms ({'reached' if r['matched_reached'] else 'NOT reached'}) | "
            f"f plain {r['f_plain_final']:.6g} warm {r['f_warm_final']:.6g} warm-start {r['f_warm_start']:.6g} "
            f"adjoint {r['f_adjoint']:.6g}")
    ps = np.array([r["plain_s"] for r in rows])
    ws = np.array([r["warm_s"] for r in rows])
    ms = np.array([r["matched_s"] for r in rows])
    return {
        "n": len(rows), "median_plain_s": float(np.median(ps)), "median_warm_s": float(np.median(ws)),
        "speedup_median_of_ratios": float(np.median(ps / ws)),
        "speedup_total": float(ps.sum() / ws.sum()),
        "median_matched_s": float(np.median(ms)),
        "matched_speedup_median_of_ratios": float(np.median(ps / ms)),
        "matched_speedup_total": float(ps.sum() / ms.sum()),
        "matched_reached_frac": float(np.mean([r["matched_reached"] for r in rows])),
        "median_matched_iters": float(np.median([r["matched_iters"] for r in rows])),
        "median_f_warm_start_over_plain_final": float(np.median([r["f_warm_start"] / r["f_plain_final"]
                                                                 for r in rows])),
        "median_plain_iters": float(np.median([r["plain_iters"] for r in rows])),
        "median_warm_iters": float(np.median([r["warm_iters"] for r in rows])),
        "fallback_rate": float(np.mean([not r["accepted"] for r in rows])),
        "median_rel_f_gap": float(np.median([(r["f_warm_final"] - r["f_plain_final"]) / r["f_plain_final"]
                                             for r in rows])),
        "frac_warm_reaching_plain_f": float(np.mean([r["warm_iters_to_plain_f"] is not None for r in rows])),
        "rows": rows,
    }


def _chunks(seq, n):
    return [seq[i:i + n] for i in range(0, len(seq), n)]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--train", type=int, default=2000, help="training phantoms")
    ap.add_argument("--test", type=int, default=200, help="held-out phantoms")
    ap.add_argument("--epochs", type=float, default=2.0)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--workers", type=int, default=12, help="CPU processes generating phantoms")
    ap.add_argument("--seed0", type=int, default=10_000, help="first training seed (held-out: seed0 + 500000)")
    ap.add_argument("--init", default=None, help="start from these weights (.npz)")
    ap.add_argument("--out", required=True, help="where to save the trained weights (.npz)")
    ap.add_argument("--log", default=None, help="JSON log")
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--obj-weight", type=float, default=0.0,
                    help="weight of the objective term f(x_warm) / f(target) in the loss (0: MSE only)")
    ap.add_argument("--target-iters", type=int, default=150,
                    help="NLCG iterations (no change rule) of each training target")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="grasp")
    ap.add_argument("--patterns", type=int, default=32,
                    help="random spoke-count patterns cycled through by the training phantoms (varying workloads)")
    args = ap.parse_args(argv)

    def log(*a):
        print(*a, flush=True)

    # phantom workers are forked before this process touches the GPU
    pool = mp.get_context("fork").Pool(max(1, args.workers))
    import torch

    from .warmstart import ObjectiveGrad, UNet2DT

    g = WORKLOADS[args.workload]
    varying = g["spokes"][0] != g["spokes"][1]
    # training trajectories: one (fixed workloads) or a pool of seeded spoke patterns, each with its
    # own plan kept for the objective term; held-out phantoms get patterns outside the pool
    n_pat = args.patterns if varying else 1
    pbs = [RadialProblem(g, spokes_for(g, 1000 + q), long_iters=args.target_iters)
           for q in range(n_pat)]
    pb = pbs[0]
    train_seeds = list(range(args.seed0, args.seed0 + args.train))
    test_seeds = list(range(args.seed0 + 500_000, args.seed0 + 500_000 + args.test))
    shape = (g["F"], g["S"], g["S"])
    X_adj = torch.empty((args.train,) + shape, dtype=torch.complex64, device=pb.dev)
    X_tgt = torch.empty_like(X_adj)
    Y: List[object] = [None] * args.train
    f_norm = np.zeros(args.train)
    pat = np.arange(args.train) % n_pat
    t0 = time.perf_counter()
    chunks = _chunks(train_seeds, 48)
    job = lambda ch: pool.map_async(_phantom_sf, [(g["S"], g["F"], s) for s in ch])
    pending = job(chunks[0]) if chunks else None
    i = 0
    for c, chunk in enumerate(chunks):
        data = pending.get()
        if c + 1 < len(chunks):
            pending = job(chunks[c + 1])
        for seed, x_true in zip(chunk, data):
            pb = pbs[pat[i]]
            pb.measure(x_true, noise_seed=seed + 7)
            pb.adjoint()
            pb.solve(pb.x_adj, long=True)
            X_adj[i].copy_(pb.x_adj)
            X_tgt[i].copy_(pb.eng_long.x)
            f_norm[i] = pb.history(long=True)[-1]["f_step"]
            if args.obj_weight:
                Y[i] = pb.y.clone()
            i += 1
        log(f"targets {i}/{args.train}: {pb.long_iters}-iteration NLCG reconstructions "
            f"({time.perf_counter() - t0:.0f} s)")
    t_data = time.perf_counter() - t0

    net = UNet2DT.load(args.init, device=pb.dev) if args.init else UNet2DT(args.levels, args.base, device=pb.dev,
                                                                              seed=1)
    n_steps = int(round(args.epochs * args.train))
    rng = np.random.default_rng(2)
    order = np.concatenate([rng.permutation(args.train) for _ in range(int(math.ceil(args.epochs)) + 1)])[:n_steps]
    warm = min(100, max(1, n_steps // 20))
    losses, gaps = [], []
    objectives = [ObjectiveGrad(q.plan, q.regm, g["F"]) for q in pbs] if args.obj_weight else None
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    pend = []
    for s, j in enumerate(order):
        j = int(j)
        lr = args.lr * min(1.0, (s + 1) / warm) * 0.5 * (1 + math.cos(math.pi * s / max(1, n_steps)))
        if objectives is not None:
            objective = objectives[pat[j]]
            pend.append(net.train_step(X_adj[j], X_tgt[j], lr=lr, objective=objective, y=Y[j],
                                       obj_weight=args.obj_weight, f_norm=[f_norm[j]]).clone())
            gaps.append(objective.value() / f_norm[j] - 1.0)
        else:
            pend.append(net.train_step(X_adj[j], X_tgt[j], lr=lr).clone())
        if (s + 1) % 50 == 0 or s + 1 == n_steps:
            vals = [float(v.item()) for v in pend]
            pend = []
            losses.extend(vals)
            gap = f" objective gap {np.mean(gaps[-len(vals):]):.4f}" if gaps else ""
            log(f"train step {s + 1}/{n_steps}: mse {np.mean(vals):.5f}{gap} lr {lr:.2e} "
                f"({(time.perf_counter() - t1) / (s + 1) * 1e3:.1f} ms/step)")
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    net.save(args.out)
    log(f"saved {args.out}")
    del X_adj, X_tgt, Y, objectives
    torch.cuda.empty_cache()

    # held-out evaluation
    res = None
    if args.test:
        test_ph = pool.map(_phantom_sf, [(g["S"], g["F"], s) for s in test_seeds])
        if varying:
            del pbs
            torch.cuda.empty_cache()
            pbs_test = [RadialProblem(g, spokes_for(g, 900_000 + s)) for s in test_seeds]
        else:
            pbs_test = pb
        res = evaluate(pbs_test, net, test_seeds, test_ph, log=log)
        log(f"held-out: plain median {res['median_plain_s'] * 1e3:.1f} ms / {res['median_plain_iters']:.0f} it, warm "
            f"median {res['median_warm_s'] * 1e3:.1f} ms / {res['median_warm_iters']:.0f} it, speedup "
            f"{res['speedup_median_of_ratios']:.2f}x, fallback {res['fallback_rate']:.2%}; to the plain objective: "
            f"median {res['median_matched_s'] * 1e3:.1f} ms / {res['median_matched_iters']:.0f} it, speedup "
            f"{res['matched_speedup_median_of_ratios']:.2f}x, reached {res['matched_reached_frac']:.0%}")
    pool.close()
    out = {"train": args.train, "test": args.test, "epochs": args.epochs, "lr": args.lr, "steps": n_steps,
           "data_seconds": t_data, "train_seconds": t_train, "ms_per_step": t_train / max(1, n_steps) * 1e3,
           "target_iters": args.target_iters, "obj_weight": args.obj_weight, "workload": args.workload,
           "patterns": n_pat,
           "gap_curve": [float(np.mean(gaps[k:k + 50])) for k in range(0, len(gaps), 50)],
           "loss_curve": [float(np.mean(losses[k:k + 50])) for k in range(0, len(losses), 50)],
           "eval": res}
    if args.log:
        with open(args.log, "w") as f:
            json.dump(out, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
