"""cf_set_initial_guess: the iterate each physical step starts from.

The reference starts every physical step's inner iterations at U^n.  Orders
1 and 2 start them at the linear or quadratic extrapolation of the time
levels instead; the converged state is the same to the convergence
tolerance, and under the reference's running-max test (main.cpp:387-406)
the steps take fewer inner iterations.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def _cfg(**kw):
    from channel_flow_amd.config import ChannelConfig, FlowParams
    fl = FlowParams()
    return ChannelConfig(32, 24, 40, gamma_metric=2.0, deltaT=1e-8, f_drive=fl.driving_force(0.005), **kw), fl


def _run(order, steps, cfg, Qn, Qm, checkpoint=None, restart=None, extrapolation=2):
    from channel_flow_amd.solver import Solver
    its = []
    with Solver(cfg) as s:
        s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
        s.set_extrapolation_order(extrapolation)
        s.set_initial_guess(order)
        if restart:
            s.load_checkpoint(restart)
        else:
            s.load_state(Qn, Qm)
        for step in range(steps):
            it, _ = s.physical_step()
            its.append(it)
            if checkpoint and step == checkpoint[0]:
                s.save_checkpoint(checkpoint[1])
        return np.array(its), np.concatenate(s.read_state())


def test_gpu_initial_guess_is_exact_on_equal_levels():
    """With U^n = U^{n-1} both extrapolations are U^n exactly (2a - a and
    3a - 3a + a round to a), so the first physical step is bit for bit the
    reference's."""
    from channel_flow_amd.config import poiseuille_state
    cfg, fl = _cfg(iteration_end_step=5)
    Q = poiseuille_state(cfg, fl, seed=4)
    base = _run(0, 1, cfg, Q, Q)
    for order in (1, 2):
        got = _run(order, 1, cfg, Q, Q)
        np.testing.assert_array_equal(got[0], base[0])
        np.testing.assert_array_equal(got[1], base[1])


def test_gpu_initial_guess_fewer_iterations_closer_state():
    """Eight physical steps under the reference's convergence test (tol 1e-3
    on er normalised by its running maximum): the linear and quadratic
    extrapolations need at most half the inner iterations of the reference's
    start at U^n, the cubic one (from the fourth step) at most 0.6 of them,
    and all end at least as close to the fully converged steps (300 inner
    iterations each).  Measured at this grid: 178 / 61 / 75 / 93 iterations,
    and 0.11 / 0.034 / 0.069 / 0.025 of the change over the steps away from
    the converged state."""
    from channel_flow_amd.config import poiseuille_state
    cfg_strict, fl = _cfg(tol=1e-300, iteration_end_step=300)
    Q = poiseuille_state(cfg_strict, fl, seed=1)
    _, q_strict = _run(0, 8, cfg_strict, Q, Q)
    cfg, _ = _cfg()
    runs = {order: _run(order, 8, cfg, Q, Q) for order in (0, 1, 2)}
    runs["cubic"] = _run(2, 8, cfg, Q, Q, extrapolation=3)
    change = np.linalg.norm(q_strict[:5] - Q)
    err = {o: np.linalg.norm(q[:5] - q_strict[:5]) / change for o, (_, q) in runs.items()}
    its = {o: int(i.sum()) for o, (i, _) in runs.items()}
    print("inner iterations over 8 steps:", its, "distance to the converged steps / change:",
          {o: f"{e:.3e}" for o, e in err.items()})
    for order in (1, 2, "cubic"):
        assert its[order] <= (0.6 if order == "cubic" else 0.5) * its[0]
        assert err[order] <= err[0]


@pytest.mark.parametrize("extrapolation", [2, 3])
def test_gpu_initial_guess_checkpoint_restart_is_exact(tmp_path, extrapolation):
    """The checkpoint keeps U^{n-2} (and U^{n-3} for the cubic start): a run
    restarted after step 3 continues the quadratic or cubic guess bit for
    bit."""
    from channel_flow_amd.config import poiseuille_state
    cfg, fl = _cfg()
    Q = poiseuille_state(cfg, fl, seed=2)
    ck = tmp_path / "guess.ckpt"
    its_full, q_full = _run(2, 6, cfg, Q, Q, checkpoint=(2, ck), extrapolation=extrapolation)
    its_rest, q_rest = _run(2, 3, cfg, None, None, restart=ck, extrapolation=extrapolation)
    np.testing.assert_array_equal(its_rest, its_full[3:])
    np.testing.assert_array_equal(q_rest, q_full)
