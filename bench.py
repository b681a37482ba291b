"""Benchmark of the dual-time inner iteration on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

One step = one inner iteration of the dual-time path (ghosts, Flux_X/Y/Z,
Viscous_terms, Dual_time_stepping with its LU-SGS sweeps, update and norms)
on the 256x192x256 channel (FP64, tanh stretching beta=2.0, Re_b=5600,
Mach 0.2, seeded laminar Poiseuille + 5% noise), state resident in HBM.
value = Mcell-iter/s = cells * K / wall seconds.  Also reported: per-stage
HIP-event times with the roofline of the dominant stage, the wall time of
one physical step run to a 1e-8 residual drop, and the reference CPU code
timed on this host (bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

GRID = (256, 192, 256)
HBM_PEAK_GBS = 8000.0
# algorithmic FP64 fields read + written per interior cell and inner
# iteration (geometry excluded, each field counted once): fluxes U_ 5 -> F 5;
# viscous U_ 5 -> vF 4; RHS U_,U,Uq,inF x3,vF -> MR (39: the first NL =
# D^-1 MR is formed in the sweep order, never written in the natural layout);
# forward sweep U_,MR,NL -> NL and backward sweep U_,NL,NR -> NR, 20 each per
# sweep; update U_,NR -> U_ 15.  The summary pre-pass and the sync reset are
# implementation work with no algorithmic bytes of their own.
def stage_fields(n_sweeps):
    return {"ghosts": 0, "flux_x": 10, "flux_y": 10, "flux_z": 10, "viscous": 9, "rhs": 39, "sweep_prep": 0,
            "sweep_sync": 0, "lusgs_fwd": 20 * n_sweeps, "lusgs_bwd": 20 * n_sweeps, "update_norms": 15}


# stages that are one kernel per launch, and that kernel's name in rocprofv3
STAGE_KERNEL = {"flux_x": "cf::k_flux_col<0>", "flux_y": "cf::k_flux_col<1>", "flux_z": "cf::k_flux_zs<2>",
                "rhs": "cf::k_rhs", "viscous": "cf::k_visc_fused",
                "lusgs_fwd": "cf::k_sweep_pipe<true, 3, 2, double>", "lusgs_bwd": "cf::k_sweep_pipe<false, 3, 2, double>",
                "sweep_prep": "cf::k_sweep_prep_rhs<double>"}
TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_256x192x256.md"
VALU_FILE = ROOT / "profiles" / "valu_inner_iteration_256x192x256.md"
WEIGHTS = ROOT / "channel_flow_amd" / "weights" / "predictor_g256x192x256.npz"
CONFIG5_DT = 2.5e-9
# trained on a 64x384x64 periodic patch with the 512x384x512 cell sizes,
# Re_b 10000, beta 2.2 and the time step above
CONFIG5_WEIGHTS = ROOT / "channel_flow_amd" / "weights" / "predictor_g512x384x512_re10000.npz"
MFMA_PEAK_TFLOPS = 2500.0  # dense bf16
PHYS_SEED = 9  # seed of the physical-step legs: held out of the predictor's training (seeds 1-8)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (copy-calibrated), or None."""
    try:
        for ln in TRAFFIC_FILE.read_text().splitlines():
            cols = [c.strip().strip("`") for c in ln.strip().strip("|").split("|")]
            if len(cols) >= 2 and cols[0] == kernel:
                return int(float(cols[-1]))
    except (OSError, ValueError):
        return None
    return None


def valu_floor_us(kernel):
    """VALU-issue floor of one launch of `kernel` in microseconds (its VALU
    instructions / (256 CUs x 2.4 GHz)) from the committed rocprofv3 SQ
    counter summary of this workload, or None."""
    try:
        lines = VALU_FILE.read_text().splitlines()
        head = next(ln for ln in lines if ln.startswith("| kernel"))
        names = [c.strip() for c in head.strip().strip("|").split("|")]
        col = names.index("VALU-bound us")
        for ln in lines:
            cols = [c.strip().strip("`") for c in ln.strip().strip("|").split("|")]
            if len(cols) == len(names) and cols[0] == kernel:
                return float(cols[col])
    except (OSError, ValueError, StopIteration):
        return None
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cores():
    """Cores this process may run on: its CPU affinity set, capped by the
    cgroup's CPU quota where one is set (cpu.max), with the CPU model."""
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(period)))
    except (OSError, ValueError):
        pass
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    cores = min(affinity, quota) if quota else affinity
    return cores, {"model": model, "online_cpus": os.cpu_count(), "affinity_cpus": affinity,
                   "cgroup_cpu_quota": quota, "threads_used": cores}


def cpu_baseline(cfg, seconds_budget=30.0):
    """The reference solver (oracle/_ref, built from its own sources) timed on
    this host's cores for a few inner iterations of the same grid."""
    exe = ROOT / "oracle" / "_ref" / f"g{GRID[0]}x{GRID[1]}x{GRID[2]}" / "ref_driver"
    if not exe.exists():
        return None
    from channel_flow_amd.config import FlowParams, poiseuille_state, write_initial_bin
    threads, cpu_info = host_cores()
    nit = 3
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as td:
        Q = poiseuille_state(cfg, FlowParams(), seed=1, stream=0)
        write_initial_bin(Path(td) / "initial.bin", Q, Q)
        del Q
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
        log(f"cpu baseline: reference ref_driver, {threads} threads, {nit} inner iterations")
        r = subprocess.run([str(exe), "time", td, "1e-8", str(nit)], env=env, capture_output=True, text=True,
                           timeout=600)
        if r.returncode != 0:
            log(f"cpu baseline failed: {r.stderr[-500:]}")
            return None
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("{")][-1]
        rec = json.loads(line)
    mcell = cfg.cells * rec["inner_iterations"] / rec["seconds"] / 1e6
    return {"value": round(mcell, 4), "unit": "Mcell-iter/s", "cores": rec["threads"], "kind": "reference",
            "host": cpu_info,
            "sample": f"{rec['inner_iterations']} inner iterations of the reference Dual_time_stepping path at "
                      f"{GRID[0]}x{GRID[1]}x{GRID[2]} (OpenMP, {rec['threads']} threads, OMP_PROC_BIND=close, "
                      f"{rec['seconds']:.1f} s, {cpu_info['model']})"}


def timed(s, steps, warmup):
    """ms per inner iteration over `steps` timed iterations after `warmup`."""
    import torch
    for _ in range(warmup):
        s.inner_iteration()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    er = None
    for _ in range(steps):
        er = s.inner_iteration()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3, er


def mean_profiles(q):
    """Mean velocity and Reynolds-stress profiles over x and z of a state Q."""
    rho = q[0]
    u, v, w = q[1] / rho, q[2] / rho, q[3] / rho
    m = lambda a: a.mean(axis=(0, 2))  # noqa: E731
    U, V, W = m(u), m(v), m(w)
    return {"U": U, "V": V, "W": W, "uu": m(u * u) - U * U, "vv": m(v * v) - V * V, "ww": m(w * w) - W * W,
            "uv": m(u * v) - U * V}


def profile_errors(q, q64):
    p, p64 = mean_profiles(q), mean_profiles(q64)
    return {k: float(np.abs(p[k] - p64[k]).max() / np.abs(p64[k]).max()) for k in p64}


def stage_profile(s, cfg, pairs):
    """Per-stage HIP-event times of the eager inner iteration and the
    roofline of the dominant single-kernel stage."""
    fields = stage_fields(pairs)
    stages = {}
    for _ in range(3):
        _, st = s.profile_inner_iteration()
        for k, (ms, nl) in st.items():
            stages.setdefault(k, []).append((ms, nl))
    stage_ms = {k: float(np.median([m for m, _ in v])) for k, v in stages.items()}
    launches = {k: v[0][1] for k, v in stages.items()}
    if not launches.get("rhs"):
        # the RHS runs inside the sweep pre-pass (k_sweep_prep_rhs)
        fields["sweep_prep"], fields["rhs"] = fields["rhs"], 0
    dom = max((k for k in STAGE_KERNEL if launches.get(k)), key=stage_ms.get)
    per_launch_ms = stage_ms[dom] / launches[dom]
    bytes_per_launch = fields[dom] * 8 * cfg.cells / launches[dom]
    achieved = bytes_per_launch / (per_launch_ms * 1e-3) / 1e9
    traffic = measured_traffic(STAGE_KERNEL[dom])
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": STAGE_KERNEL[dom],
                "stage": dom, "launches_per_iteration": launches[dom], "avg_launch_us": round(per_launch_ms * 1e3, 3),
                "algorithmic_bytes_per_launch": int(bytes_per_launch),
                "algorithmic_fields_per_cell": fields[dom] // launches[dom],
                "traffic_source": str(TRAFFIC_FILE.relative_to(ROOT)) if traffic else None}
    per_stage = {}
    for k, v in stage_ms.items():
        gbs = fields[k] * 8 * cfg.cells / (v * 1e-3) / 1e9 if fields[k] and v > 0 else None
        # the whole stage's VALU floor over its measured time (stages of one kernel)
        floor = valu_floor_us(STAGE_KERNEL[k]) if k in STAGE_KERNEL and cfg.cells == GRID[0] * GRID[1] * GRID[2] \
            else None
        per_stage[k] = {"ms": round(v, 4), "launches": launches[k],
                        "GBps_algorithmic": round(gbs, 1) if gbs else None,
                        "hbm_frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None,
                        "valu_floor_frac": (round(floor * launches[k] * 1e-3 / v, 4)
                                            if floor is not None and v > 0 else None)}
    return per_stage, roofline


def physical_steps(cfg, flow, precision, order, pred, steps, extrapolation=3):
    """Physical steps from the seeded state (seed PHYS_SEED, which the
    predictor's training did not see) with the reference's convergence test
    at tol 1e-8: er_k / E_k < 1e-8 for k = 1..4, E_k the running maxima of
    the run (main.cpp:387-406); inner iterations run as graph replays.  The
    extrapolated start and the predictor's fallback are cubic once four
    levels are known."""
    from channel_flow_amd.learned import run_steps
    r, qn = run_steps(cfg, flow, PHYS_SEED, steps, order, pred, precision, graph=True, extrapolation=extrapolation)
    return r, qn


def predictor_leg(cfg, pred, meta):
    """The predictor's kernels timed with HIP events: inference over the whole
    256x192x256 domain and one training batch, with their MFMA rates."""
    prof = pred.profile(*GRID)
    cells = cfg.cells
    # algorithmic FLOPs per cell: 27-point conv 2*27*64*64, pointwise 2*16*64
    # and 2*64*5; training: forward (pointwise + conv) on the inner cells,
    # backward data on every crop cell, backward weight on the inner cells
    conv = 2 * 27 * 64 * 64
    infer_flop = cells * (conv + 2 * 16 * 64 + 2 * 64 * 5)
    B, inner, vol = pred.batch_crops, 14 ** 3, 16 ** 3
    t = {"forward": B * inner * conv + B * vol * 2 * 16 * 64, "bwd_data": B * vol * conv, "bwd_weight": B * inner * conv}
    mfma = {"infer": {"ms": round(prof["infer"], 4), "TFLOPs": round(infer_flop / prof["infer"] * 1e-9, 1),
                      "frac_of_2.5_PFLOPs_bf16": round(infer_flop / prof["infer"] * 1e-9 / MFMA_PEAK_TFLOPS, 4),
                      "kernel": "cfp::k_fwd<true, false>"},
            "train_forward": {"ms": round(prof["train_forward"], 4),
                              "TFLOPs": round(t["forward"] / prof["train_forward"] * 1e-9, 1)},
            "train_bwd_data": {"ms": round(prof["train_bwd_data"], 4),
                               "TFLOPs": round(t["bwd_data"] / prof["train_bwd_data"] * 1e-9, 1)},
            "train_bwd_weight": {"ms": round(prof["train_bwd_weight"], 4),
                                 "TFLOPs": round(t["bwd_weight"] / prof["train_bwd_weight"] * 1e-9, 1)},
            "train_head_ms": round(prof["train_head"], 4), "train_reduce_ms": round(prof["train_reduce"], 4)}
    for k in ("train_forward", "train_bwd_data", "train_bwd_weight"):
        mfma[k]["frac_of_2.5_PFLOPs_bf16"] = round(mfma[k]["TFLOPs"] / MFMA_PEAK_TFLOPS, 4)
    keep = ("grid", "seeds", "snapshots", "steps_per_seed", "crops_per_step", "n_crops", "train_steps",
            "held_out_loss", "held_out_base_loss", "seconds")
    return {"weights": str(pred.path), "training": {k: meta.get(k) for k in keep},
            "kernels": mfma}


def config5_leg(args):
    """512x384x512, Re_b 10000, Mach 0.2, beta 2.2: FP32 sweeps with the
    graph-captured inner iteration against plain FP64, throughput, device
    memory in use, and the converged state / mean profiles against FP64."""
    import torch
    from channel_flow_amd.learned import channel_cfg, run_steps
    from channel_flow_amd.solver import CF_FP32_SWEEPS, CF_FP64, Predictor, Solver
    from channel_flow_amd.config import poiseuille_state
    # the wall cells are 2.7x thinner than at 256x192x256, beta 2.0: 1e-8 s
    # diverges in the first step, 5e-9 s converges without a downward trend
    # in the iteration count, 2.5e-9 s converges steadily
    cfg, flow = channel_cfg((512, 384, 512), beta=2.2, Re=10000.0, deltaT=CONFIG5_DT, tol=1e-8)
    Q = poiseuille_state(cfg, flow, seed=1, stream=0)
    out = {"grid": "512x384x512", "cells": cfg.cells, "Re_bulk": 10000.0, "beta": 2.2, "deltaT": CONFIG5_DT}
    for name, prec, graph in (("fp64_eager", CF_FP64, False), ("fp32_sweeps_graph", CF_FP32_SWEEPS, True)):
        free0, total = torch.cuda.mem_get_info()
        with Solver(cfg, precision=prec) as s:
            s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
            s.set_graph(graph)
            s.load_state(Q, Q)
            free1, _ = torch.cuda.mem_get_info()
            ms, _ = timed(s, 5, 2)
        out[name] = {"ms_per_iteration": round(ms, 3), "Mcell_iter_s": round(cfg.cells / ms * 1e-3, 1),
                     "device_memory_GB": round((free0 - free1) / 1e9, 2), "device_total_GB": round(total / 1e9, 1)}
        log(f"config 5 {name}: {out[name]}")
    steps = args.config5_steps
    pred, meta = None, None
    wpath = CONFIG5_WEIGHTS if CONFIG5_WEIGHTS.exists() else WEIGHTS
    if wpath.exists():
        from channel_flow_amd.learned import load
        pred = Predictor(dataset_crops=1, batch_crops=8)
        meta = load(pred, wpath)
    r64, q64 = run_steps(cfg, flow, 1, steps, 0, None, CF_FP64, graph=True, log=lambda m: log("config 5 fp64 " + m))
    r32, q32 = run_steps(cfg, flow, 1, steps, 3 if pred else 2, pred, CF_FP32_SWEEPS, graph=True, extrapolation=3,
                         log=lambda m: log("config 5 fp32 sweeps + predictor " + m))
    out["phys_steps_fp64_start_Un"] = {k: r64[k] for k in ("iterations", "wall_s")}
    out["phys_steps_fp32_sweeps_" + ("predictor" if pred else "extrapolated")] = {
        k: r32[k] for k in ("iterations", "wall_s", "predictor_steps", "fallbacks")}
    out["predictor_weights"] = ({"file": str(wpath.relative_to(ROOT)), "grid": meta.get("grid"),
                                 "cells_of": meta.get("cells_of"), "Re_bulk": meta.get("Re_bulk"),
                                 "deltaT": meta.get("deltaT"), "seeds": meta.get("seeds")} if pred else None)
    skip = min(2, steps - 1)  # the first steps have no third time level yet
    m64 = float(np.mean(r64["wall_s"][skip:]))
    m32 = float(np.mean(r32["wall_s"][skip:]))
    out["mean_over_steps"] = f"{skip + 1}-{steps}"
    out["mean_iterations_fp64"] = float(np.mean(r64["iterations"][skip:]))
    out["mean_iterations_fp32_sweeps_predictor"] = float(np.mean(r32["iterations"][skip:]))
    out["speedup_wall_vs_fp64"] = round(m64 / m32, 3)
    out["state_rel_l2_vs_fp64"] = float(np.linalg.norm(q32 - q64) / np.linalg.norm(q64))
    out["profile_max_rel_err_vs_fp64"] = profile_errors(q32, q64)
    if pred:
        pred.close()
    return out


def phys_long_leg(args):
    """Physical steps to er_k / E_k < 1e-8 at 256x192x256 past the seeded
    transient: FP64 from U^n (the reference's start), FP64 and FP32 sweeps
    from the cubic extrapolation, and FP32 sweeps with the learned start
    (kept where its first residuals beat the quadratic extrapolation's,
    otherwise the step restarts from the cubic one), each `--phys-long-steps`
    steps from seed PHYS_SEED, with iterations and wall time per window of
    steps and the speedup over FP64 from U^n."""
    from channel_flow_amd.learned import channel_cfg, load, run_steps
    from channel_flow_amd.solver import CF_FP32_SWEEPS, CF_FP64, Predictor
    n = args.phys_long_steps
    cfg, flow = channel_cfg(GRID, beta=2.0, deltaT=1e-8, tol=1e-8, cap=2000)
    wpath = Path(args.weights).resolve() if args.weights else WEIGHTS
    pred = Predictor(dataset_crops=1, batch_crops=8)
    load(pred, wpath)
    edges = [e for e in (3, 20, 40, 60, 100, 200) if e < n] + [n]
    windows = list(zip(edges[:-1], edges[1:]))
    out = {"steps": n, "seed": PHYS_SEED,
           "weights": str(wpath.relative_to(ROOT) if wpath.is_relative_to(ROOT) else wpath),
           "windows": [f"{a + 1}-{b}" for a, b in windows]}
    runs = (("fp64_start_Un", CF_FP64, 0), ("fp64_cubic", CF_FP64, 2), ("fp32_sweeps_cubic", CF_FP32_SWEEPS, 2),
            ("fp32_sweeps_predictor", CF_FP32_SWEEPS, 3))
    for name, prec, order in runs:
        r, _ = run_steps(cfg, flow, PHYS_SEED, n, order, pred if order == 3 else None, prec, graph=True,
                         extrapolation=3, log=(lambda m, nm=name: log(f"phys-long {nm} {m}")))
        its, wall = np.array(r["iterations"], float), np.array(r["wall_s"])
        out[name] = {"iterations": r["iterations"], "fallbacks": r["fallbacks"], "predictor_steps": r["predictor_steps"],
                     "mean_iterations": [round(float(its[a:b].mean()), 1) for a, b in windows],
                     "mean_wall_s": [round(float(wall[a:b].mean()), 4) for a, b in windows],
                     "total_wall_s_after_step_3": round(float(wall[3:].sum()), 2)}
        log(f"phys-long {name}: {out[name]['mean_iterations']} its, {out[name]['mean_wall_s']} s per window")
    base = out["fp64_start_Un"]
    for name, _, _ in runs[1:]:
        out[name]["speedup_wall_vs_fp64_start_Un"] = [round(x / y, 3) for x, y in
                                                      zip(base["mean_wall_s"], out[name]["mean_wall_s"])]
        out[name]["speedup_total_after_step_3"] = round(base["total_wall_s_after_step_3"] /
                                                        out[name]["total_wall_s_after_step_3"], 3)
    pred.close()
    return out


def read_st(path, ny):
    """st%05d.dat (Statistic.cpp:1117-1146): seven blocks U, V, W, UU, VV,
    WW, UV of ny - 1 lines (j = 2..ny)."""
    v = np.array([float(x) for x in Path(path).read_text().split()])
    return v.reshape(7, ny - 1)


def long_runs(args):
    """Config 2: 100 FP64 physical steps to a 1e-8 drop of er_k / E_k at
    256x192x256 (iterations and wall time per step).  Config 3: 1000
    physical steps under the reference's own test (er_k / E_k < 1e-3, at most
    200 inner iterations, main.cpp:387-406) in FP64 and with FP32 sweeps,
    with cf_statistic after every step (statistic_step 1000, the reference's
    st file), and the L2 error of mean u(y), u'v'(y), u'u'(y) of the FP32 run
    against the FP64 run."""
    import torch
    from channel_flow_amd.learned import channel_cfg
    from channel_flow_amd.config import poiseuille_state
    from channel_flow_amd.solver import CF_FP32_SWEEPS, CF_FP64, Solver
    out = {}
    cfg, flow = channel_cfg(GRID, beta=2.0, deltaT=1e-8, tol=1e-8, cap=2000)
    Q = poiseuille_state(cfg, flow, seed=1, stream=0)
    with Solver(cfg) as s:
        s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
        s.set_graph(True)
        s.load_state(Q, Q)
        its, wall = [], []
        for step in range(args.config2_steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            it, _ = s.physical_step()
            torch.cuda.synchronize()
            its.append(it)
            wall.append(round(time.perf_counter() - t0, 4))
            if step % 10 == 9:
                log(f"config 2: {step + 1} steps, last {it} inner iterations {wall[-1]} s")
    out["config2_fp64_tol_1e-8"] = {"steps": len(its), "iterations": its, "wall_s": wall,
                                    "total_inner_iterations": int(sum(its)), "total_wall_s": round(sum(wall), 2),
                                    "mean_wall_s_per_step": round(float(np.mean(wall)), 4)}
    del Q
    cfg3, _ = channel_cfg(GRID, beta=2.0, deltaT=1e-8, tol=1e-3, cap=200)
    Q = poiseuille_state(cfg3, flow, seed=1, stream=0)
    n = args.config3_steps
    prof = {}
    for name, prec in (("fp64", CF_FP64), ("fp32_sweeps", CF_FP32_SWEEPS)):
        with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as td, \
                Solver(cfg3, precision=prec) as s:
            s.Jacobian_all(cfg3.gamma_metric, want_arrays=False)
            s.set_graph(True)
            s.load_state(Q, Q)
            its = []
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for step in range(1, n + 1):
                it, _ = s.physical_step()
                its.append(it)
                s.Statistic(step, n, None, td)
                if step % 100 == 0:
                    log(f"config 3 {name}: {step} steps")
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            prof[name] = read_st(Path(td) / f"st{n:05d}.dat", cfg3.ny)
            shear = (Path(td) / "Shear_stress.dat").read_text().splitlines()
        out[f"config3_{name}"] = {"steps": n, "total_inner_iterations": int(sum(its)), "wall_s": round(dt, 2),
                                  "mean_inner_iterations": round(float(np.mean(its)), 2),
                                  "wall_s_per_step": round(dt / n, 4),
                                  "wall_shear_last_step": shear[-1].split("\t")}
        log(f"config 3 {name}: {out[f'config3_{name}']}")
    a, b = prof["fp32_sweeps"], prof["fp64"]
    l2 = lambda x, y: float(np.linalg.norm(x - y) / np.linalg.norm(y))  # noqa: E731
    U = (a[0], b[0])
    uv = (a[6] - a[0] * a[1], b[6] - b[0] * b[1])
    uu = (a[3] - a[0] * a[0], b[3] - b[0] * b[0])
    out["config3_fp32_vs_fp64_profile_rel_l2"] = {"U": l2(*U), "u'v'": l2(*uv), "u'u'": l2(*uu)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-converge", action="store_true", help="skip the physical steps to 1e-8")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-sweep leg")
    ap.add_argument("--no-predictor", action="store_true", help="skip the learned-predictor legs")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every kernel eagerly (for counter collection); the timed value then has no graph")
    ap.add_argument("--weights", default=None,
                    help="predictor weights for the 256x192x256 legs (default: the committed ones)")
    ap.add_argument("--serial-stencils", action="store_true",
                    help="run the fluxes and the viscous terms one after another on one stream (per-kernel traces)")
    ap.add_argument("--phys-steps", type=int, default=8, help="physical steps per run of the 1e-8 legs")
    ap.add_argument("--config5", action="store_true", help="also run the 512x384x512 leg (config 5)")
    ap.add_argument("--config5-steps", type=int, default=6)
    ap.add_argument("--long-runs", action="store_true", help="also run configs 2 and 3 (100 and 1000 steps)")
    ap.add_argument("--config2-steps", type=int, default=100)
    ap.add_argument("--config3-steps", type=int, default=1000)
    ap.add_argument("--phys-long-steps", type=int, default=60, help="physical steps per run of --only phys-long")
    ap.add_argument("--only", choices=["long-runs", "config5", "phys-long"], default=None,
                    help="run only that leg and print its JSON")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain 1 GiB device copy (the reference kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the solver has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import CF_FP32_SWEEPS, CF_FP64, Predictor, Solver

    torch.cuda.set_device(0)
    if args.only:
        leg = {"long-runs": long_runs, "config5": config5_leg, "phys-long": phys_long_leg}[args.only]
        print(json.dumps(leg(args)), flush=True)
        return
    if args.copy_calibration:
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB written")
        del a, b
    flow = FlowParams(Re_bulk=5600.0, Mach=0.2)
    cfg = ChannelConfig(*GRID, gamma_metric=2.0, gamma_grid=2.0, deltaT=1e-8,
                        f_drive=flow.driving_force(0.005), x_periodic=False, tol=1e-8, iteration_end_step=2000)
    Q = poiseuille_state(cfg, flow, seed=1, stream=0)
    log(f"grid {GRID}, cells {cfg.cells}")

    # the inner iteration: graph replays, one executed sweep pair (the
    # reference's second pair reproduces the first bit for bit in one domain)
    s = Solver(cfg)
    s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
    s.set_graph(not args.no_graph)
    s.set_concurrent_stencils(not args.serial_stencils)
    s.load_state(Q, Q)
    ms_per_step, er = timed(s, args.steps, args.warmup)
    value = cfg.cells / ms_per_step * 1e-3
    log(f"timed {args.steps} inner iterations (graph, one sweep pair): {ms_per_step:.3f} ms/iter, "
        f"{value:.1f} Mcell-iter/s, er={er}")
    s.set_graph(False)
    s.load_state(Q, Q)
    ms_eager, _ = timed(s, args.steps, args.warmup)
    s.set_sweep_reuse(False)
    s.load_state(Q, Q)
    ms_two, _ = timed(s, args.steps, args.warmup)
    s.set_sweep_reuse(True)
    log(f"eager launches: {ms_eager:.3f} ms/iter; both sweep pairs executed: {ms_two:.3f} ms/iter")
    s.load_state(Q, Q)
    per_stage, roofline = stage_profile(s, cfg, 1)
    log(f"stages: {json.dumps(per_stage)}")

    # variant kernels off the dual-time path (Filter sensor, limiter fluxes) on
    # the same resident state; algorithmic fields per cell: Filter reads rho
    # and the momenta (4) and writes 3 sensors, each limiter flux reads U_ 5
    # and its sensor, writes 5
    s.Grid(2.8)
    vt = {}
    for _ in range(3):
        for k, v in s.profile_variants(0.007).items():
            vt.setdefault(k, []).append(v)
    vfields = {"filter": 7, "flux_x_limiter": 11, "flux_y_limiter": 11, "flux_z_limiter": 11}
    variants = {k: {"ms": round(float(np.median(v)), 4),
                    "GBps_algorithmic": round(vfields[k] * 8 * cfg.cells / (float(np.median(v)) * 1e-3) / 1e9, 1)}
                for k, v in vt.items()}
    log(f"variants: {json.dumps(variants)}")
    s.close()

    fp32 = None
    if not args.no_fp32:
        # FP32 sweeps: the same inner iteration
        s = Solver(cfg, precision=CF_FP32_SWEEPS)
        s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
        s.set_graph(not args.no_graph)
        s.set_concurrent_stencils(not args.serial_stencils)
        s.load_state(Q, Q)
        ms32, _ = timed(s, args.steps, args.warmup)
        s.load_state(Q, Q)
        stages32, _ = stage_profile(s, cfg, 1)
        s.close()
        fp32 = {"value": round(cfg.cells / ms32 * 1e-3, 3), "unit": "Mcell-iter/s", "ms_per_step": round(ms32, 4),
                "stages_ms": {k: v["ms"] for k, v in stages32.items()}}
        log(f"FP32 sweeps: {json.dumps(fp32)}")

    pred, meta, predictor = None, None, None
    wpath = Path(args.weights).resolve() if args.weights else WEIGHTS
    if wpath.exists() and not args.no_predictor:
        from channel_flow_amd.learned import load
        pred = Predictor(dataset_crops=1, batch_crops=32)
        meta = load(pred, wpath)
        pred.path = wpath.relative_to(ROOT) if wpath.is_relative_to(ROOT) else wpath
        predictor = predictor_leg(cfg, pred, meta)
        log(f"predictor: {json.dumps(predictor)}")

    phys = None
    if not args.no_converge:
        # wall time per physical step to a 1e-8 drop of er_k / E_k: the plain
        # FP64 path (start U^n, the reference's) against FP32 sweeps with the
        # learned start; mean over steps 4.. (the first ones are the seeded
        # transient, and the predictor needs three levels)
        n = args.phys_steps
        runs = [("fp64_start_Un", CF_FP64, 0), ("fp64_cubic", CF_FP64, 2)]
        if not args.no_fp32:
            runs.append(("fp32_sweeps_start_Un", CF_FP32_SWEEPS, 0))
        if pred is not None:
            runs.append(("fp64_predictor", CF_FP64, 3))
            if not args.no_fp32:
                runs.append(("fp32_sweeps_predictor", CF_FP32_SWEEPS, 3))
        phys, states = {}, {}
        for name, prec, order in runs:
            r, qn = physical_steps(cfg, flow, prec, order, pred, n)
            r["mean_iterations_steps_4_on"] = round(float(np.mean(r["iterations"][3:])), 2)
            r["mean_wall_s_steps_4_on"] = round(float(np.mean(r["wall_s"][3:])), 4)
            r["wall_s"] = [round(w, 4) for w in r["wall_s"]]
            states[name] = qn
            phys[name] = r
            log(f"physical steps {name}: {r}")
        base = states["fp64_start_Un"]
        for name in phys:
            if name != "fp64_start_Un":
                phys[name]["state_rel_l2_vs_fp64_start_Un"] = float(np.linalg.norm(states[name] - base) /
                                                                    np.linalg.norm(base))
                phys[name]["profile_max_rel_err_vs_fp64_start_Un"] = profile_errors(states[name], base)
        best = [k for k in ("fp32_sweeps_predictor", "fp64_predictor", "fp32_sweeps_start_Un") if k in phys]
        if best:
            phys["speedup_" + best[0] + "_vs_fp64_start_Un"] = round(
                phys["fp64_start_Un"]["mean_wall_s_steps_4_on"] / phys[best[0]]["mean_wall_s_steps_4_on"], 3)
        if pred is not None:
            p = phys["fp64_predictor"]
            phys["predictor_fallback_rate"] = p["fallbacks"] / p["predictor_steps"] if p["predictor_steps"] else None
        del states
    if pred is not None:
        pred.close()
    config5 = config5_leg(args) if args.config5 else None
    longr = long_runs(args) if args.long_runs else None
    del Q

    cpu = None if args.no_cpu else cpu_baseline(cfg)
    out = {
        "metric": "Mcell-iter/s at 256x192x256 + wall-time/phys step to 1e-8 residual drop",
        "value": round(value, 3), "unit": "Mcell-iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded laminar Poiseuille + 5% uniform velocity noise",
        "config": {"workload": "plain FP64 dual-time inner iteration, 256x192x256 channel "
                               "(2pi x 2 x pi, tanh beta=2.0, Re_b=5600, Mach 0.2, dt=1e-8)",
                   "grid": "x".join(map(str, GRID)), "cells": cfg.cells, "n_sweeps": cfg.n_sweeps,
                   "sweeps_executed": "1 of 2 (the second pair reproduces the first bit for bit in one domain: "
                                      "tests/test_gpu_graph_sweeps.py::test_gpu_one_sweep_pair_is_two, "
                                      "tests/test_gpu_sizes.py::test_gpu_benchmark_grid_one_sweep_pair_and_graph, "
                                      "and bit for bit against the reference's two pairs in "
                                      "tests/test_gpu_sizes.py::test_gpu_benchmark_grid_matches_reference)",
                   "launch": ("eager launches" if args.no_graph else
                              "hipGraph replay of each inner iteration (same bits as eager launches)") +
                             ("; fluxes and viscous terms one after another" if args.serial_stencils else
                              "; fluxes and viscous terms on four streams"),
                   "parallelism": "single GPU"},
        "ms_per_step_eager_launches": round(ms_eager, 4),
        "ms_per_step_both_sweep_pairs_executed": round(ms_two, 4),
        "Mcell_iter_s_both_sweep_pairs_executed": round(cfg.cells / ms_two * 1e-3, 3),
        "phys_step_to_1e-8": phys,
        "fp32_sweeps": fp32,
        "predictor": predictor,
        "config5": config5,
        "long_runs": longr,
        "stages_ms": per_stage,
        "variants_ms": variants,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
