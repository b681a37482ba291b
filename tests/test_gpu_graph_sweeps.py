"""One executed LU-SGS sweep pair against the reference's n_sweeps = 2 pairs,
and hipGraph replays of the inner iteration against launches issued one by
one: identical bits in both cases.

Dual_time_stepping.cpp:514-1277 runs two forward/backward pairs with
NL <- NR between them; in one domain the second pair reads only MR and
values its own pass has formed, so it reproduces the first.  The graph test
replays several times with eager work (state reads, the bulk-flux kernel) and
synchronisations between the replays, and checks every replay.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

GRIDS = [(21, 13, 18), (8, 8, 8), (40, 33, 9), (70, 45, 37)]


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def _run(cfg, Q, iters, precision=0, reuse=True, graph=False, schedule=0, between=None, concurrent=True):
    from channel_flow_amd.solver import Solver
    with Solver(cfg, precision=precision) as s:
        s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
        s.set_sweep_schedule(schedule)
        s.set_sweep_reuse(reuse)
        s.set_graph(graph)
        s.set_concurrent_stencils(concurrent)
        s.load_state(Q, Q)
        ers = []
        for it in range(iters):
            ers.append(s.inner_iteration())
            if between is not None:
                between(s, it)
        builds = s.graph_builds()
        state = np.concatenate(s.read_state())
    return np.array(ers), state, builds


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("precision", [0, 1])
def test_gpu_one_sweep_pair_is_two(grid, precision):
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    print(f"sweep reuse {grid} precision {precision}", flush=True)
    cfg = ChannelConfig(*grid, gamma_metric=2.0, deltaT=5e-8)
    Q = poiseuille_state(cfg, FlowParams(), seed=4, stream=0)
    e1, s1, _ = _run(cfg, Q, 3, precision, reuse=True)
    e2, s2, _ = _run(cfg, Q, 3, precision, reuse=False)
    np.testing.assert_array_equal(e1, e2)
    np.testing.assert_array_equal(s1, s2)


@pytest.mark.parametrize("grid", [(21, 13, 18), (16, 12, 16)])
def test_gpu_one_sweep_pair_is_two_per_plane_and_host_entry(grid, golden):
    """The per-hyperplane schedule, and cf_dual_time_stepping's MR, NL, NR,
    U_ and er on host arrays."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    cfg = ChannelConfig(*grid, gamma_metric=1.5, deltaT=5e-8)
    Q = poiseuille_state(cfg, FlowParams(), seed=6, stream=0)
    e1, s1, _ = _run(cfg, Q, 2, reuse=True, schedule=1)
    e2, s2, _ = _run(cfg, Q, 2, reuse=False, schedule=1)
    np.testing.assert_array_equal(e1, e2)
    np.testing.assert_array_equal(s1, s2)
    outs = []
    for reuse in (True, False):
        with Solver(cfg) as s:
            met = s.Jacobian_all(1.5)
            s.set_sweep_reuse(reuse)
            from oracle import oracle as O
            U = O.state_from_Q(cfg, Q, met["J"])
            U_ = [a.copy() for a in U]
            s.boundary_conditions(U_, met["J"])
            F = [s.Flux_X(U_, met["J"], met["J_u"], met["m_u"]), s.Flux_Y(U_, met["J"], met["J_v"], met["m_v"]),
                 s.Flux_Z(U_, met["J"], met["J_w"], met["m_w"])]
            vF = s.Viscous_terms(U_, met)
            Ui = [a.copy() for a in U_]
            er, MR, NR, NL = s.Dual_time_stepping(Ui, U, U, met, *F, vF)
            outs.append((er, np.stack(Ui), np.stack(MR), np.stack(NR), np.stack(NL)))
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("grid", [(21, 13, 18), (40, 33, 9)])
@pytest.mark.parametrize("precision", [0, 1])
def test_gpu_graph_replays_equal_eager(grid, precision):
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    print(f"graph {grid} precision {precision}", flush=True)
    cfg = ChannelConfig(*grid, gamma_metric=2.0, deltaT=5e-8)
    Q = poiseuille_state(cfg, FlowParams(), seed=8, stream=0)
    seen = []

    def between(s, it):
        # eager work and synchronisations between the replays
        seen.append(s.read_state()[2].sum())
        s.Driven_force()
    e1, s1, b1 = _run(cfg, Q, 5, precision, graph=False)
    e2, s2, b2 = _run(cfg, Q, 5, precision, graph=True, between=between)
    np.testing.assert_array_equal(e1, e2)
    np.testing.assert_array_equal(s1, s2)
    assert b1 == 0 and b2 == 1
    assert len(set(seen)) == len(seen)  # the iterate moved on every replay


def test_gpu_graph_recaptures_on_change():
    """A changed setting (here the sweep reuse) is captured anew, and the
    replays still equal the eager launches."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    cfg = ChannelConfig(21, 13, 18, gamma_metric=2.0, deltaT=5e-8)
    Q = poiseuille_state(cfg, FlowParams(), seed=9, stream=0)
    runs = []
    for graph in (False, True):
        with Solver(cfg) as s:
            s.Jacobian_all(2.0, want_arrays=False)
            s.set_graph(graph)
            s.load_state(Q, Q)
            ers = [s.inner_iteration()]
            s.set_sweep_reuse(False)
            ers.append(s.inner_iteration())
            s.set_sweep_reuse(True)
            ers.append(s.inner_iteration())
            runs.append((np.array(ers), s.read_state()[2], s.graph_builds()))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    assert runs[1][2] == 3


@pytest.mark.parametrize("grid", [(21, 13, 18), (70, 45, 37)])
@pytest.mark.parametrize("graph", [False, True])
def test_gpu_concurrent_stencils_same_bits(grid, graph):
    """Flux_X/Y/Z and Viscous_terms on four streams (the default) against one
    after another, eagerly and inside the graph."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    cfg = ChannelConfig(*grid, gamma_metric=2.0, deltaT=5e-8)
    Q = poiseuille_state(cfg, FlowParams(), seed=10, stream=0)
    e1, s1, _ = _run(cfg, Q, 4, graph=graph, concurrent=True)
    e2, s2, _ = _run(cfg, Q, 4, graph=graph, concurrent=False)
    np.testing.assert_array_equal(e1, e2)
    np.testing.assert_array_equal(s1, s2)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("graph", [False, True])
def test_gpu_sweep_gives_up_then_recovers(precision, graph):
    """A tiled sweep that gives up, forced from the host by a one-poll bound
    on its halo wave's wait (every kernel still runs to its end), reports
    CF_ERR_TIMEOUT.  With the state reloaded and the bound restored, the next
    iterations give a fresh context's bits: the cross-tile hand-off slots,
    which the abandoned sweeps may have left without the sentinel, are seeded
    again before the forward sweep reads them."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import CfError, Solver
    print(f"sweep gives up: precision {precision} graph {graph}", flush=True)
    cfg = ChannelConfig(64, 48, 32, gamma_metric=2.0, deltaT=1e-8)
    Qn = poiseuille_state(cfg, FlowParams(), seed=6, stream=0)
    Qm = poiseuille_state(cfg, FlowParams(), seed=6, stream=1)

    def start(s):
        s.Jacobian_all(2.0, want_arrays=False)
        s.set_graph(graph)
        s.load_state(Qn, Qm)

    with Solver(cfg, precision=precision) as s:
        start(s)
        e_ref = np.array([s.inner_iteration() for _ in range(3)])
        q_ref = s.read_state()[2]
    with Solver(cfg, precision=precision) as s:
        start(s)
        s.set_spin_limit(1)
        with pytest.raises(CfError, match="TIMEOUT"):
            s.inner_iteration()
        s.set_spin_limit(1 << 22)
        s.load_state(Qn, Qm)
        e = np.array([s.inner_iteration() for _ in range(3)])
        np.testing.assert_array_equal(e, e_ref)
        np.testing.assert_array_equal(s.read_state()[2], q_ref)
