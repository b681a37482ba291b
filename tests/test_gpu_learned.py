"""The learned initial iterate inside the solver (cf_set_predictor,
cf_set_initial_guess(ctx, 3)).

- The whole pipeline at a small grid: crops captured from the solver's own
  converged steps, feature scales, training on the GPU, the predictor
  attached; the converged state of physical steps started by it equals the
  one started at U^n (the reference's start) to 1e-10 of the change over the
  steps: the network changes the iteration count, never the answer.
- What the trained network gains: with the committed weights (trained on the
  GPU by channel_flow_amd.learned on 8 seeds at 64^3), on held-out seed 9,
  the inner iterations per physical step to er_k / E_k < 1e-8 (the
  reference's running-max normalisation, main.cpp:387-397) against the U^n
  start and the quadratic extrapolation (the fallback), with the fallback
  rate reported.
- The same for the 512x384x512 weights (Re_b 10000, beta 2.2, dt 2.5e-9 s),
  trained on a 64x384x64 periodic patch with that grid's cell sizes, on the
  patch and held-out seed 1.
"""
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT, gpu_available

pytestmark = pytest.mark.gpu

WEIGHTS64 = ROOT / "channel_flow_amd" / "weights" / "predictor_g64x64x64_dt5e-8.npz"
WEIGHTS512 = ROOT / "channel_flow_amd" / "weights" / "predictor_g512x384x512_re10000.npz"


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def test_gpu_learned_pipeline_same_answer():
    from channel_flow_amd import learned as L
    from channel_flow_amd.config import poiseuille_state
    from channel_flow_amd.solver import Predictor, Solver
    cfg, flow = L.channel_cfg((32, 24, 40), beta=2.0, deltaT=5e-8, tol=1e-8, cap=3000)
    with Predictor(dataset_crops=2 * 4 * 8, batch_crops=8) as pred:
        its = L.generate(pred, cfg, flow, seeds=[1, 2], steps=6, crops_per_step=8)
        assert pred.dataset_size() == 2 * 4 * 8
        info = L.train(pred, steps=150, lr=2e-3)
        assert np.all(np.isfinite(info["losses"])) and info["losses"][-1] < info["losses"][0]
        # the same steps, converged tightly, from U^n and from the predictor
        # er is a 4th-power norm of the increments (Dual_time_stepping.cpp:1318-1352):
        # er / E < 1e-24 leaves increments ~1e-12 of the first ones
        tight, _ = L.channel_cfg((32, 24, 40), beta=2.0, deltaT=5e-8, tol=1e-24, cap=6000)
        Q = poiseuille_state(tight, flow, seed=3, stream=0)
        finals = {}
        # U^n; the predictor with the quadratic fallback; the predictor with
        # the cubic restart (a loss to the quadratic check restarts the step
        # from the cubic extrapolation once four levels are known)
        for name, order, extrapolation in (("Un", 0, 2), ("pred", 3, 2), ("pred_cubic", 3, 3)):
            with Solver(tight) as s:
                s.Jacobian_all(2.0, want_arrays=False)
                s.set_predictor(pred)
                s.set_extrapolation_order(extrapolation)
                s.set_initial_guess(order)
                s.load_state(Q, Q)
                n_its = [s.physical_step()[0] for _ in range(5)]
                steps, falls = s.predictor_stats()
                finals[name] = s.read_state()[0]
            print(f"{name}: iterations {n_its}, predictor steps {steps}, fallbacks {falls}", flush=True)
            # the first step has no U^{n-2} yet and starts from the extrapolation
            assert steps == (4 if order == 3 else 0)
    change = np.linalg.norm(finals["Un"] - np.asarray(Q))
    for name in ("pred", "pred_cubic"):
        diff = np.linalg.norm(finals[name] - finals["Un"])
        print(f"{name}: state difference {diff / change:.2e} of the change over 5 steps", flush=True)
        assert diff <= 1e-10 * change
    del its


def test_gpu_learned_gain_64cubed():
    from channel_flow_amd import learned as L
    from channel_flow_amd.solver import Predictor
    assert WEIGHTS64.exists()
    cfg, flow = L.channel_cfg((64, 64, 64), beta=2.0, deltaT=5e-8, tol=1e-8)
    with Predictor(dataset_crops=1, batch_crops=8) as pred:
        meta = L.load(pred, WEIGHTS64)
        assert 9 not in meta["seeds"]
        res = L.compare(cfg, flow, pred, seed=9, steps=8, skip=3)
    its = {k: res[k]["mean_iterations"] for k in ("start_Un", "extrapolated", "predictor")}
    print(f"held-out seed 9, steps 4-8, inner iterations per physical step: {its}; fallback rate "
          f"{res['fallback_rate']}; wall gain vs U^n {res['gain_wall_vs_start_Un']:.2f}", flush=True)
    assert its["predictor"] < 0.85 * min(its["start_Un"], its["extrapolated"])
    assert res["fallback_rate"] is not None and res["fallback_rate"] <= 0.5


def test_gpu_learned_gain_512_patch():
    from channel_flow_amd import learned as L
    from channel_flow_amd.solver import Predictor
    assert WEIGHTS512.exists()
    cfg, flow = L.channel_cfg((64, 384, 64), beta=2.2, Re=10000.0, deltaT=2.5e-9, tol=1e-8, cells_of=(512, 512))
    with Predictor(dataset_crops=1, batch_crops=8) as pred:
        meta = L.load(pred, WEIGHTS512)
        assert 1 not in meta["seeds"] and meta["cells_of"] == [512, 512]
        res = L.compare(cfg, flow, pred, seed=1, steps=8, skip=3)
    its = {k: res[k]["mean_iterations"] for k in ("start_Un", "extrapolated", "predictor")}
    print(f"patch of 512x384x512, held-out seed 1, steps 4-8: {its}; fallback rate {res['fallback_rate']}",
          flush=True)
    assert its["predictor"] < 0.85 * min(its["start_Un"], its["extrapolated"])
    assert res["fallback_rate"] is not None and res["fallback_rate"] <= 0.5


def test_gpu_predictor_close_detaches():
    """Closing a predictor that a context still holds detaches it first: the
    next physical steps start from the extrapolation (the context keeps no
    pointer to freed memory)."""
    from channel_flow_amd import learned as L
    from channel_flow_amd.config import poiseuille_state
    from channel_flow_amd.solver import Predictor, Solver
    cfg, flow = L.channel_cfg((24, 20, 24), beta=2.0, deltaT=5e-8, tol=1e-6, cap=400)
    Q = poiseuille_state(cfg, flow, seed=5, stream=0)
    pred = Predictor(dataset_crops=1, batch_crops=8)
    L.load(pred, WEIGHTS64)
    with Solver(cfg) as s:
        s.Jacobian_all(2.0, want_arrays=False)
        s.set_predictor(pred)
        s.set_initial_guess(3)
        s.load_state(Q, Q)
        for _ in range(3):
            s.physical_step()
        assert s.predictor_stats()[0] == 2
        pred.close()
        assert s._pred is None
        its = [s.physical_step()[0] for _ in range(2)]
        assert all(0 < i <= 400 for i in its)
        assert np.isfinite(np.concatenate(s.read_state())).all()
