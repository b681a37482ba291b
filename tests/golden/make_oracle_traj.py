"""Writes oracle-made trajectory fixtures of tests/golden/ for runs the
reference cannot be configured for.

The reference hard-codes the metric stretching gamma = 1.5 in Jacobian*.cpp
(Jacobian.cpp:50) and the driving force in prm.h:13, so a run at tanh
beta = 2.0 with the laminar driving force 3 mu U_b / h^2 has no reference
counterpart.  This script runs it through the CPU oracle (oracle/cf_oracle.c),
whose every kernel and whose 16x12x16 trajectory match the reference bit for
bit at gamma = 1.5 (tests/test_oracle_golden.py); the stretching enters only
through the closed-form metrics, which the same tests pin.

    python tests/golden/make_oracle_traj.py            # config 1 at beta 2.0 (about 3 minutes)

Output: otraj_b2_g64x64x64.npz with the per-iteration residual rows
(step, iteration, er1..er5, er1..er5 over their running maxima), the sums of
the ten LES blocks after the last step and the LES plane i = 32.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent
# config 1: 64^3, tanh beta = 2.0, Re_b = 5600, Mach 0.2, laminar driving
# force, seeded Poiseuille + 5% noise, 10 physical steps x 20 inner iterations
CONFIG1_B2 = dict(grid=(64, 64, 64), beta=2.0, seed=1, deltaT=5e-8, steps=10, iteration_end_step=20)


def config1_b2_cfg():
    flow = FlowParams(Re_bulk=5600.0, Mach=0.2)
    X, Y, Z = CONFIG1_B2["grid"]
    cfg = ChannelConfig(X, Y, Z, gamma_metric=CONFIG1_B2["beta"], gamma_grid=CONFIG1_B2["beta"],
                        deltaT=CONFIG1_B2["deltaT"], f_drive=flow.driving_force(0.005),
                        iteration_end_step=CONFIG1_B2["iteration_end_step"])
    return cfg, flow


def oracle_trajectory(cfg, Qn, Qm, steps):
    """main.cpp:143-498 with the dual-time path (convergence test
    main.cpp:387-406, level shift main.cpp:444-464), through the oracle."""
    met = O.metrics(cfg)
    U = O.state_from_Q(cfg, Qn, met["J"])
    Uq = O.state_from_Q(cfg, Qm, met["J"])
    Ui = [a.copy() for a in U]
    E = np.zeros(5)
    rows = []
    g, ny, nz = cfg.gend, cfg.ny, cfg.nz
    sl = (slice(3, g + 1), slice(2, ny + 1), slice(2, nz + 1))
    for step in range(1, steps + 1):
        for it in range(1, 50000):
            er = O.inner_iteration(cfg, Ui, U, Uq, met)
            E = np.maximum(E, er)
            ern = er / E
            rows.append([step, it, *er, *ern])
            if (ern[:4] < cfg.tol).all() or it == cfg.iteration_end_step:
                for v in range(5):
                    Uq[v][sl] = U[v][sl]
                    U[v][sl] = Ui[v][sl]
                break
        print(f"step {step}: {it} inner iterations, er1 {rows[-1][2]:.3e}", flush=True)
    les = np.stack([U[v][sl] * met["J"][sl] for v in range(5)] + [Uq[v][sl] * met["J"][sl] for v in range(5)])
    return np.array(rows), les


def main():
    cfg, flow = config1_b2_cfg()
    Qn = poiseuille_state(cfg, flow, seed=CONFIG1_B2["seed"], stream=0)
    rows, les = oracle_trajectory(cfg, Qn, Qn, CONFIG1_B2["steps"])
    X = cfg.X_out
    np.savez_compressed(OUT / f"otraj_b2_g{X}x{cfg.Y_out}x{cfg.Z_out}.npz", er=rows,
                        les_sum=les.reshape(10, -1).sum(axis=1), les_plane=les[:, X // 2])
    man = json.loads((OUT / "manifest.json").read_text())
    man["fixtures"][f"otraj_b2_g{X}x{cfg.Y_out}x{cfg.Z_out}"] = dict(
        CONFIG1_B2, generator="tests/golden/make_oracle_traj.py (CPU oracle)", f_drive=cfg.f_drive,
        Re_bulk=flow.Re_bulk, Mach=flow.Mach,
        driver="main.cpp:143-498 with Viscous_terms + Dual_time_stepping enabled")
    (OUT / "manifest.json").write_text(json.dumps(man, indent=1) + "\n")


if __name__ == "__main__":
    main()
