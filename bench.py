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
                "rhs": "cf::k_rhs",
                "lusgs_fwd": "cf::k_sweep_pipe<true, 3, 2, double>", "lusgs_bwd": "cf::k_sweep_pipe<false, 3, 2, double>",
                "sweep_prep": "cf::k_sweep_prep_rhs<double>"}
TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_256x192x256.md"


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


def physical_steps(s, Q, steps=2, cap=800):
    """From the seeded state (Q at both levels): physical steps, each with
    inner iterations until max_k er_k/er_k(1) <= 1e-8 (k = 1..4), then the
    level shift.  The second step, the first with a BDF2 history, is the one
    reported; the iterate after it is returned for comparisons."""
    import torch
    s.load_state(Q, Q)
    recs = []
    for step in range(steps):
        if step:
            s.advance_levels()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        first, its, ratio = None, 0, 1.0
        while its < cap:
            e = s.inner_iteration()
            its += 1
            if first is None:
                first = e.copy()
            ratio = float(np.max(e[:4] / first[:4]))
            if ratio <= 1e-8:
                break
        torch.cuda.synchronize()
        recs.append({"wall_s": round(time.perf_counter() - t1, 4), "inner_iterations": its, "residual_drop": ratio,
                     "reached_1e-8": ratio <= 1e-8})
    conv = dict(recs[-1], cap=cap, step=steps, earlier_steps=recs[:-1])
    return conv, s.read_state()[2]


def reference_test_steps(cfg, Q, precision, order, steps=6):
    """Physical steps under the reference's own convergence test
    (cf_physical_step: er1..er4 over their running maxima below 1e-3, at most
    200 inner iterations) from the seeded state, with the given sweep
    precision and initial-guess order; per-step iterations and wall times."""
    import torch
    from channel_flow_amd.solver import Solver
    s = Solver(cfg, precision=precision)
    s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
    s.set_initial_guess(order)
    s.load_state(Q, Q)
    its, wall = [], []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it, _ = s.physical_step()
        torch.cuda.synchronize()
        its.append(it)
        wall.append(time.perf_counter() - t0)
    qn = s.read_state()[0]
    s.close()
    return its, wall, qn


def mean_profiles(q):
    """Mean velocity and Reynolds-stress profiles over x and z of a state Q."""
    rho = q[0]
    u, v, w = q[1] / rho, q[2] / rho, q[3] / rho
    m = lambda a: a.mean(axis=(0, 2))  # noqa: E731
    U, V, W = m(u), m(v), m(w)
    return {"U": U, "V": V, "W": W, "uu": m(u * u) - U * U, "vv": m(v * v) - V * V, "ww": m(w * w) - W * W,
            "uv": m(u * v) - U * V}


def fp32_sweeps_leg(cfg, Q, args, q64):
    """The same measurements with CF_FP32_SWEEPS (LU-SGS sweeps and split
    Jacobians in float inside the FP64 loop), and the converged state and
    mean profiles against the FP64 run's."""
    import torch
    from channel_flow_amd.solver import CF_FP32_SWEEPS, Solver
    s = Solver(cfg, precision=CF_FP32_SWEEPS)
    s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
    s.load_state(Q, Q)
    for _ in range(args.warmup):
        s.inner_iteration()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.inner_iteration()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    sts = [s.profile_inner_iteration()[1] for _ in range(3)]
    stage_ms = {k: round(float(np.median([x[k][0] for x in sts])), 4) for k in sts[0]}
    out = {"value": round(cfg.cells * args.steps / dt / 1e6, 3), "unit": "Mcell-iter/s",
           "ms_per_step": round(dt / args.steps * 1e3, 4), "stages_ms": stage_ms}
    if q64 is not None:
        conv, q32 = physical_steps(s, Q)
        out["phys_step_to_1e-8"] = conv
        change = float(np.linalg.norm(q64 - Q))
        diff = float(np.linalg.norm(q32 - q64))
        p64, p32 = mean_profiles(q64), mean_profiles(q32)
        out["vs_fp64"] = {
            "state_rel_l2": diff / float(np.linalg.norm(q64)),
            "state_diff_over_change": diff / change if change > 0 else None,
            "profile_max_rel_err": {k: float(np.abs(p32[k] - p64[k]).max() / np.abs(p64[k]).max()) for k in p64}}
    s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-converge", action="store_true", help="skip the physical steps to 1e-8")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-sweep leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain 1 GiB device copy (the reference kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the solver has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver

    torch.cuda.set_device(0)
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
                        f_drive=flow.driving_force(0.005), x_periodic=False)
    Q = poiseuille_state(cfg, flow, seed=1, stream=0)
    log(f"grid {GRID}, cells {cfg.cells}")
    s = Solver(cfg)
    s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
    s.load_state(Q, Q)
    for _ in range(args.warmup):
        s.inner_iteration()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    er = None
    for _ in range(args.steps):
        er = s.inner_iteration()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    value = cfg.cells * args.steps / dt / 1e6
    log(f"timed {args.steps} inner iterations: {ms_per_step:.3f} ms/iter, {value:.1f} Mcell-iter/s, er={er}")

    # per-stage HIP-event times (events on the context's stream)
    fields = stage_fields(cfg.n_sweeps)
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
    # dominant kernel: the single-kernel stage with the largest time per iteration
    dom = max((k for k in STAGE_KERNEL if launches.get(k)), key=stage_ms.get)
    per_launch_ms = stage_ms[dom] / launches[dom]
    bytes_per_launch = fields[dom] * 8 * cfg.cells / launches[dom]
    achieved = bytes_per_launch / (per_launch_ms * 1e-3) / 1e9
    traffic = measured_traffic(STAGE_KERNEL[dom])
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": STAGE_KERNEL[dom],
                "stage": dom, "launches_per_iteration": launches[dom], "avg_launch_us": round(per_launch_ms * 1e3, 3),
                "algorithmic_bytes_per_launch": int(bytes_per_launch),
                "traffic_source": str(TRAFFIC_FILE.relative_to(ROOT)) if traffic else None}
    per_stage = {k: {"ms": round(v, 4), "launches": launches[k],
                     "GBps_algorithmic": (round(fields[k] * 8 * cfg.cells / (v * 1e-3) / 1e9, 1)
                                          if fields[k] and v > 0 else None)} for k, v in stage_ms.items()}
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

    conv = q64 = None
    if not args.no_converge:
        conv, q64 = physical_steps(s, Q)
        log(f"physical steps to 1e-8 (FP64): {conv}")
    s.close()

    fp32 = None
    if not args.no_fp32:
        fp32 = fp32_sweeps_leg(cfg, Q, args, q64)
        log(f"FP32 sweeps: {json.dumps(fp32)}")

    ref_test = None
    if not args.no_converge:
        # the reference's convergence test over six physical steps: its own
        # configuration (FP64, steps start at U^n), the FP32 sweeps, and the
        # linearly extrapolated start
        from channel_flow_amd.solver import CF_FP32_SWEEPS, CF_FP64
        ref_test, base = {}, None
        for name, prec, order in (("fp64_start_Un", CF_FP64, 0), ("fp32_sweeps_start_Un", CF_FP32_SWEEPS, 0),
                                  ("fp64_start_linear", CF_FP64, 1)):
            its, wall, qn = reference_test_steps(cfg, Q, prec, order)
            rec = {"inner_iterations": its, "wall_s": [round(w, 4) for w in wall],
                   "wall_s_per_step_3_to_6": round(float(np.mean(wall[2:])), 4)}
            if base is None:
                base = qn
            else:
                rec["state_rel_l2_vs_fp64_start_Un"] = float(np.linalg.norm(qn - base) / np.linalg.norm(base))
            ref_test[name] = rec
            log(f"reference convergence test, {name}: {rec}")
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
                   "parallelism": "single GPU"},
        "phys_step_to_1e-8": conv,
        "fp32_sweeps": fp32,
        "phys_steps_reference_test": ref_test,
        "stages_ms": per_stage,
        "variants_ms": variants,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
