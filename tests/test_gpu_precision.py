"""CF_FP32_SWEEPS: the LU-SGS sweeps and their split Jacobians in float,
inside the FP64 defect-correction loop.

The right-hand side (fluxes, viscous terms, the BDF2 term), D^-1's inputs
and the state update stay FP64, so the converged state is the FP64 one: the
sweeps only shape the path to it.  The tests check that the float path is
the one that runs (its increments differ from FP64's at float precision),
that one iteration's increment equals the CPU oracle's FP64 one to float
precision, that physical steps converged to a 1e-8 drop of er land on the
FP64 state and mean profiles, and that the unsupported schedules are
refused.
"""
import numpy as np
import pytest

from conftest import gpu_available

from channel_flow_amd.solver import CF_FP32_SWEEPS as FP32_SWEEPS, CF_FP64 as FP64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def _cfg(grid=(32, 24, 40)):
    from channel_flow_amd.config import ChannelConfig, FlowParams
    fl = FlowParams()
    return ChannelConfig(*grid, gamma_metric=2.0, deltaT=1e-8, f_drive=fl.driving_force(0.005)), fl


def _profiles(q):
    """Mean velocity and Reynolds-stress profiles over x and z."""
    rho = q[0]
    u, v, w = q[1] / rho, q[2] / rho, q[3] / rho
    m = lambda a: a.mean(axis=(0, 2))  # noqa: E731
    U, V, W = m(u), m(v), m(w)
    return {"U": U, "V": V, "W": W, "uu": m(u * u) - U * U, "vv": m(v * v) - V * V, "uv": m(u * v) - U * V}


def _converged_steps(precision, steps=3, cap=800):
    """Physical steps, each iterated until max_k er_k/er_k(1) <= 1e-8."""
    from channel_flow_amd.config import poiseuille_state
    from channel_flow_amd.solver import Solver
    cfg, fl = _cfg()
    Q = poiseuille_state(cfg, fl, seed=1)
    its = []
    with Solver(cfg, precision=precision) as s:
        s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
        s.load_state(Q, Q)
        for step in range(steps):
            if step:
                s.advance_levels()
            first = None
            for it in range(1, cap + 1):
                e = s.inner_iteration()
                first = e.copy() if first is None else first
                if np.max(e[:4] / first[:4]) <= 1e-8:
                    break
            its.append(it)
        qi = s.read_state()[2]
    return np.array(its), qi, Q


def test_gpu_fp32_sweeps_run_in_float():
    """One inner iteration: the FP32 increments NR agree with the FP64 ones
    to float precision and are not the FP64 bits."""
    from channel_flow_amd.config import poiseuille_state
    from channel_flow_amd.solver import Solver
    cfg, fl = _cfg()
    Qn = poiseuille_state(cfg, fl, seed=3, stream=0)
    Qm = poiseuille_state(cfg, fl, seed=3, stream=1)
    out = {}
    for prec in (FP64, FP32_SWEEPS):
        with Solver(cfg, precision=prec) as s:
            s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
            s.load_state(Qn, Qm)
            er = s.inner_iteration()
            out[prec] = (er, s.read_state()[2] - Qn)
    (e64, d64), (e32, d32) = out[FP64], out[FP32_SWEEPS]
    rel = np.linalg.norm(d32 - d64) / np.linalg.norm(d64)
    assert 0 < rel <= 1e-5, rel
    np.testing.assert_allclose(e32, e64, rtol=1e-4)
    assert not np.array_equal(e32, e64)


def test_gpu_fp32_sweeps_against_the_oracle():
    """One inner iteration with the FP32 sweeps on the GPU against the CPU
    oracle's FP64 inner iteration (Dual_time_stepping.cpp) from the same
    state: the state increments agree to 1e-6 relative (measured 7e-8: the
    float rounding of the sweeps' inputs and arithmetic).  The bound separates
    sweeps: on the CPU, the oracle's DPLUSGS Jacobi sweeps in place of the
    LU-SGS ones (a plausible mix-up of the two implicit solvers) move the
    increment by 2.8e-2 relative."""
    from oracle import oracle as O
    from channel_flow_amd.config import poiseuille_state
    from channel_flow_amd.solver import Solver
    cfg, fl = _cfg()
    Qn = poiseuille_state(cfg, fl, seed=3, stream=0)
    Qm = poiseuille_state(cfg, fl, seed=3, stream=1)
    met = O.metrics(cfg)
    U, Uq = O.state_from_Q(cfg, Qn, met["J"]), O.state_from_Q(cfg, Qm, met["J"])
    sl = (slice(3, cfg.gend + 1), slice(2, cfg.ny + 1), slice(2, cfg.nz + 1))

    def oracle_increment(dplusgs):
        Ui = [a.copy() for a in U]
        O.inner_iteration(cfg, Ui, U, Uq, met, dplusgs=dplusgs)
        return np.stack([Ui[v][sl] * met["J"][sl] for v in range(5)]) - Qn

    d_o = oracle_increment(False)
    with Solver(cfg, precision=FP32_SWEEPS) as s:
        s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
        s.load_state(Qn, Qm)
        s.inner_iteration()
        d32 = s.read_state()[2] - Qn
    rel = np.linalg.norm(d32 - d_o) / np.linalg.norm(d_o)
    print(f"FP32 sweeps against the oracle: increment relative difference {rel:.2e}", flush=True)
    assert 0 < rel <= 1e-6, rel
    planted = np.linalg.norm(oracle_increment(True) - d_o) / np.linalg.norm(d_o)
    assert planted > 1e-2, planted


def test_gpu_fp32_sweeps_converge_to_the_fp64_state():
    """Three physical steps to a 1e-8 drop of er: the same inner-iteration
    counts (within one), and the state and mean profiles of FP64 to 1e-9 of
    the change over the steps (measured: about 1e-11)."""
    its64, q64, Q = _converged_steps(FP64)
    its32, q32, _ = _converged_steps(FP32_SWEEPS)
    assert np.abs(its32 - its64).max() <= 1
    change = np.linalg.norm(q64 - Q)
    assert np.linalg.norm(q32 - q64) <= 1e-9 * change
    p64, p32 = _profiles(q64), _profiles(q32)
    for k in p64:
        scale = np.abs(p64[k]).max()
        assert np.abs(p32[k] - p64[k]).max() <= 1e-9 * scale, k


def test_gpu_fp32_sweeps_refuse_other_schedules():
    """FP32 runs on the pipelined tiled schedule only: selecting another one
    is an error, and nothing falls back to FP64."""
    from channel_flow_amd.solver import Solver
    cfg, _ = _cfg((16, 12, 16))
    with Solver(cfg, precision=FP32_SWEEPS) as s:
        for schedule in (1, 2):
            with pytest.raises(RuntimeError, match="pipelined"):
                s.set_sweep_schedule(schedule)
        s.set_sweep_schedule(0)


def test_gpu_fp32_sweeps_benchmark_grid():
    """256x192x256, stretching 2.0, dt 1e-8: one physical step from the
    seeded state to a 1e-8 drop of er in both precisions.  Same iteration
    count within one, and state and mean profiles of FP64 to 1e-9 of the
    step's change (measured in the benchmark: about 1e-11)."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    fl = FlowParams()
    cfg = ChannelConfig(256, 192, 256, gamma_metric=2.0, deltaT=1e-8, f_drive=fl.driving_force(0.005))
    Q = poiseuille_state(cfg, fl, seed=1)
    res = {}
    for prec in (FP64, FP32_SWEEPS):
        with Solver(cfg, precision=prec) as s:
            s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
            s.load_state(Q, Q)
            first = None
            for it in range(1, 801):
                e = s.inner_iteration()
                first = e.copy() if first is None else first
                if np.max(e[:4] / first[:4]) <= 1e-8:
                    break
            res[prec] = (it, s.read_state()[2])
            print(f"precision {prec}: {it} inner iterations", flush=True)
    (i64, q64), (i32, q32) = res[FP64], res[FP32_SWEEPS]
    assert abs(i32 - i64) <= 1
    change = np.linalg.norm(q64 - Q)
    assert np.linalg.norm(q32 - q64) <= 1e-9 * change
    p64, p32 = _profiles(q64), _profiles(q32)
    for k in p64:
        assert np.abs(p32[k] - p64[k]).max() <= 1e-9 * np.abs(p64[k]).max(), k
