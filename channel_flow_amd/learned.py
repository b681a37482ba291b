"""Training and evaluation of the learned initial iterate (host driver over
the C ABI; every kernel runs in libchannel_flow.so).

The reference starts each physical step's inner iterations at U^n
(main.cpp:444-464).  Here a network predicts Q^{n+1} from Q^n, Q^{n-1},
Q^{n-2} (cf_pred_*, include/channel_flow.h), trained on crops of physical
steps that the FP64 solver converges itself:

    python -m channel_flow_amd.learned train --grid 64 64 64 --seeds 1-8 --steps 20 \\
        --out channel_flow_amd/weights/predictor_g64x64x64.npz
    python -m channel_flow_amd.learned evaluate --grid 64 64 64 --weights ... --seed 9

Convergence of a physical step is judged by the reference's running-max
normalisation (main.cpp:387-397): er_k / E_k < tol for k = 1..4, with E_k
the largest er_k of the run so far.  E_k comes from the first steps of the
run, which every start shares, so the test does not depend on the guess.
"""
from __future__ import annotations

import argparse
import json
import math
import time
from pathlib import Path

import numpy as np

from .config import ChannelConfig, FlowParams, poiseuille_state
from .solver import CF_FP32_SWEEPS, CF_FP64, Predictor, Solver

HIGH = 0.005


def channel_cfg(grid, beta=2.0, Re=5600.0, deltaT=1e-8, tol=1e-8, cap=2000, cells_of=None):
    """The channel of the benchmark (2pi x 2 x pi h, tanh beta, Re_b, Mach 0.2,
    laminar driving force) on `grid`, with the physical-step test tol.

    cells_of = (X, Z): a periodic patch of the X x Z channel instead, with
    its cell sizes (Lx = 2 pi h grid[0] / X, Lz = pi h grid[2] / Z).  The
    network sees 3x3x3 neighbourhoods, so weights trained on such a patch
    serve the full grid of the same spacings, Re_b, beta and time step."""
    flow = FlowParams(Re_bulk=Re, Mach=0.2)
    extra = {}
    if cells_of is not None:
        extra = dict(Lx=2 * math.pi * HIGH * grid[0] / cells_of[0], Lz=math.pi * HIGH * grid[2] / cells_of[1])
    cfg = ChannelConfig(*grid, gamma_metric=beta, gamma_grid=beta, deltaT=deltaT,
                        f_drive=flow.driving_force(HIGH), iteration_end_step=cap, tol=tol, **extra)
    return cfg, flow


def _sync():
    import torch
    torch.cuda.synchronize()


def generate(pred: Predictor, cfg, flow, seeds, steps, crops_per_step, precision=CF_FP64, log=print):
    """Runs `steps` physical steps from each seed's noisy Poiseuille state and,
    from the third step on (four levels known), appends crops of every step
    to the predictor's dataset.  Returns per-seed iteration counts."""
    out = {}
    for seed in seeds:
        Q = poiseuille_state(cfg, flow, seed=seed, stream=0)
        with Solver(cfg, precision=precision) as s:
            s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
            s.keep_levels(4)
            s.load_state(Q, Q)
            its = []
            t0 = time.perf_counter()
            for step in range(1, steps + 1):
                it, _ = s.physical_step()
                its.append(it)
                if step >= 3:
                    pred.capture(s, crops_per_step, seed=1000003 * seed + step)
            out[seed] = its
            log(f"seed {seed}: {steps} steps, iterations {its}, {time.perf_counter() - t0:.1f} s, "
                f"dataset {pred.dataset_size()} crops")
    return out


def train(pred: Predictor, steps=2000, lr=1e-3, holdout=0.1, seed=7, log=print):
    """Feature scales from all crops, then Adam on the first (1 - holdout)
    of them; returns the loss history and the held-out errors."""
    n = pred.dataset_size()
    n_train = max(1, int(round(n * (1.0 - holdout))))
    stats = pred.fit_norm(n)
    pred.init_params(seed)
    t0 = time.perf_counter()
    losses = pred.train(n_train, steps, lr=lr, seed=seed)
    _sync()
    dt = time.perf_counter() - t0
    held = pred.eval(n_train, n - n_train) if n > n_train else (float("nan"), float("nan"))
    tr = pred.eval(0, min(n_train, 4 * pred.batch_crops))
    log(f"trained {steps} steps in {dt:.1f} s; loss {losses[0]:.4g} -> {losses[-1]:.4g}; "
        f"held-out {held[0]:.4g} (base guess {held[1]:.4g}); train {tr[0]:.4g} (base {tr[1]:.4g})")
    return {"losses": losses, "held_out_loss": held[0], "held_out_base_loss": held[1], "train_loss": tr[0],
            "train_base_loss": tr[1], "n_crops": n, "n_train": n_train, "seconds": dt,
            "stats": stats.tolist()}


def save(pred: Predictor, path, meta):
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(path, params=pred.get_params(), norm=pred.get_norm(),
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))


def load(pred: Predictor, path):
    z = np.load(path)
    pred.set_params(z["params"])
    pred.set_norm(z["norm"])
    return json.loads(z["meta"].tobytes().decode())


def run_steps(cfg, flow, seed, steps, order, pred=None, precision=CF_FP64, graph=False, extrapolation=2, log=None):
    """Physical steps from the seed's state with initial guess `order` (0 U^n,
    2 extrapolation, 3 the attached predictor; `extrapolation` 2 quadratic or
    3 cubic, for order 2 and the predictor's fallback): per-step inner
    iterations and wall seconds, the predictor's fallback count, the final
    level-n state.  `log`, if given, receives one line per step."""
    Q = poiseuille_state(cfg, flow, seed=seed, stream=0)
    with Solver(cfg, precision=precision) as s:
        s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
        s.set_graph(graph)
        if pred is not None:
            s.set_predictor(pred)
        s.set_extrapolation_order(extrapolation)
        s.set_initial_guess(order)
        s.load_state(Q, Q)
        its, wall = [], []
        for _ in range(steps):
            _sync()
            t0 = time.perf_counter()
            it, _ = s.physical_step()
            _sync()
            its.append(it)
            wall.append(time.perf_counter() - t0)
            if log is not None:
                log(f"step {len(its)}: {it} inner iterations, {wall[-1]:.3f} s")
        stats = s.predictor_stats()
        qn = s.read_state()[0]
    return {"iterations": its, "wall_s": wall, "predictor_steps": stats[0], "fallbacks": stats[1]}, qn


def compare(cfg, flow, pred, seed, steps, skip=3, precision=CF_FP64, log=print, graph=False):
    """The reference's start (U^n), the quadratic extrapolation and the
    predictor (with its fallback) on a held-out seed: iterations and wall
    time per physical step over steps skip+1..steps, and the predictor run's
    final state against the U^n run's (both converged to the same tol)."""
    res, states = {}, {}
    for name, order in (("start_Un", 0), ("extrapolated", 2), ("predictor", 3)):
        r, qn = run_steps(cfg, flow, seed, steps, order, pred if order == 3 else None, precision, graph)
        r["mean_iterations"] = float(np.mean(r["iterations"][skip:]))
        r["mean_wall_s"] = float(np.mean(r["wall_s"][skip:]))
        res[name] = r
        states[name] = qn
        log(f"{name}: iterations {r['iterations']}, mean over steps {skip + 1}-{steps} "
            f"{r['mean_iterations']:.1f} its / {r['mean_wall_s']:.3f} s; fallbacks {r['fallbacks']} of "
            f"{r['predictor_steps']}")
    base = states["start_Un"]
    change = None
    for name in ("extrapolated", "predictor"):
        d = float(np.linalg.norm(states[name] - base) / np.linalg.norm(base))
        res[name]["state_rel_l2_vs_start_Un"] = d
    res["gain_iterations_vs_extrapolated"] = res["extrapolated"]["mean_iterations"] / res["predictor"]["mean_iterations"]
    res["gain_iterations_vs_start_Un"] = res["start_Un"]["mean_iterations"] / res["predictor"]["mean_iterations"]
    res["gain_wall_vs_start_Un"] = res["start_Un"]["mean_wall_s"] / res["predictor"]["mean_wall_s"]
    res["fallback_rate"] = (res["predictor"]["fallbacks"] / res["predictor"]["predictor_steps"]
                            if res["predictor"]["predictor_steps"] else None)
    del change
    return res


def _seeds(text):
    a, _, b = text.partition("-")
    return list(range(int(a), int(b) + 1)) if b else [int(a)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["train", "evaluate"])
    ap.add_argument("--grid", type=int, nargs=3, default=[64, 64, 64])
    ap.add_argument("--beta", type=float, default=2.0)
    ap.add_argument("--re", type=float, default=5600.0)
    ap.add_argument("--dt", type=float, default=1e-8)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--seeds", default="1-8")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--crops", type=int, default=16)
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--fp32-sweeps", action="store_true")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--out", default=None)
    ap.add_argument("--seed", type=int, default=9, help="held-out seed of the evaluation")
    ap.add_argument("--eval-steps", type=int, default=12)
    ap.add_argument("--skip", type=int, default=3, help="steps left out of the evaluation means")
    ap.add_argument("--cells-of", type=int, nargs=2, default=None, metavar=("X", "Z"),
                    help="train on a periodic patch with the cell sizes of the X x Z channel")
    a = ap.parse_args()
    import torch
    torch.cuda.set_device(0)
    cfg, flow = channel_cfg(tuple(a.grid), a.beta, a.re, a.dt, a.tol, cells_of=a.cells_of)
    precision = CF_FP32_SWEEPS if a.fp32_sweeps else CF_FP64
    seeds = _seeds(a.seeds)
    if a.mode == "train":
        n_data = len(seeds) * max(a.steps - 2, 0) * a.crops
        with Predictor(dataset_crops=max(n_data, 1), batch_crops=a.batch) as pred:
            its = generate(pred, cfg, flow, seeds, a.steps, a.crops, precision)
            info = train(pred, a.train_steps, a.lr)
            meta = {"grid": a.grid, "cells_of": a.cells_of, "beta": a.beta, "Re_bulk": a.re, "deltaT": a.dt, "tol": a.tol, "seeds": seeds,
                    "steps_per_seed": a.steps, "crops_per_step": a.crops, "snapshots": len(seeds) * (a.steps - 2),
                    "train_steps": a.train_steps, "lr": a.lr, "batch_crops": a.batch,
                    "data_iterations": {str(k): v for k, v in its.items()},
                    **{k: v for k, v in info.items() if k != "losses"},
                    "loss_first_last": [float(info["losses"][0]), float(info["losses"][-1])]}
            if a.out:
                save(pred, a.out, meta)
            if a.eval_steps:
                res = compare(cfg, flow, pred, a.seed, a.eval_steps, skip=a.skip, precision=precision)
                print(json.dumps({"train": meta, "evaluate": res}, default=float))
    else:
        with Predictor(dataset_crops=1, batch_crops=a.batch) as pred:
            meta = load(pred, a.weights)
            res = compare(cfg, flow, pred, a.seed, a.eval_steps, skip=a.skip, precision=precision)
            print(json.dumps({"weights": meta, "evaluate": res}, default=float))


if __name__ == "__main__":
    main()
