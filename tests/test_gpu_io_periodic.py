"""Device-path file entry points, true-periodic x, and host-array entry points
on a context with a resident state.

- cf_load_initial reads initial.bin (Initial_condition.cpp:64-79) and
  cf_write_output writes LES%05d.bin (Output.cpp:160-190); after the
  reference's three physical steps the file equals the reference's own
  LES00003.bin byte for byte.
- x_periodic = 1 is the np >= 2 semantics of main.cpp:205-240: ghost plane 2
  holds plane gend and ghost plane gend+1 holds plane 3 (closed form), and a
  short trajectory in that mode matches the oracle run in the same mode.
- Per-function entry points given host arrays leave a resident state as it
  was.
- Config 1 at its stated parameters (64^3, tanh beta = 2.0, laminar driving
  force), 10 physical steps x 20 inner iterations against the oracle run.
"""
import numpy as np
import pytest

from conftest import gpu_available, rel_l2

pytestmark = pytest.mark.gpu

TOL_STATE = 1e-10


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def test_gpu_load_initial_write_output_bytes(golden, tmp_path):
    """initial.bin in, three physical steps, LES00003.bin out: the reference's
    file (traj_g16x12x16) byte for byte."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state, write_initial_bin
    from channel_flow_amd.solver import Solver
    z = golden("traj_g16x12x16")
    cfg = ChannelConfig(16, 12, 16, gamma_metric=1.5, deltaT=5e-8, iteration_end_step=6)
    Qn = poiseuille_state(cfg, FlowParams(), seed=7, stream=0)
    write_initial_bin(tmp_path / "initial.bin", Qn, Qn)
    with Solver(cfg) as s:
        s.Jacobian_all(1.5, want_arrays=False)
        s.Initial_condition(tmp_path / "initial.bin")
        for _ in range(3):
            s.physical_step()
        s.Output(3, tmp_path)
    got = np.fromfile(tmp_path / "LES00003.bin", dtype="<f8")
    want = np.ascontiguousarray(z["les"], dtype="<f8").ravel()
    assert got.size == want.size
    assert rel_l2(got, want) <= TOL_STATE, rel_l2(got, want)
    assert (tmp_path / "LES00003.bin").read_bytes() == want.tobytes()


def test_gpu_load_initial_short_file(tmp_path):
    """A truncated initial.bin is an error (CF_ERR_IO), not a partial load
    (the reference reads it unchecked, Initial_condition.cpp:64-79)."""
    from channel_flow_amd.config import ChannelConfig
    from channel_flow_amd.solver import CfError, Solver
    cfg = ChannelConfig(16, 12, 16, gamma_metric=1.5)
    (tmp_path / "initial.bin").write_bytes(b"\0" * (8 * 10 * cfg.cells - 8))
    with Solver(cfg) as s:
        s.Jacobian_all(1.5, want_arrays=False)
        with pytest.raises(CfError, match="CF_ERR_IO"):
            s.Initial_condition(tmp_path / "initial.bin")
        with pytest.raises(CfError, match="CF_ERR_IO"):
            s.Initial_condition(tmp_path / "missing.bin")


@pytest.mark.parametrize("grid", [(16, 12, 16), (9, 10, 13)])
def test_gpu_x_periodic_ghosts(grid):
    """x_periodic = 1: after the ghosts, plane 2 equals plane gend and plane
    gend+1 equals plane 3 over the planes' filled region, for all five
    fields; x_periodic = 0 copies the neighbours (planes 3 and gend)."""
    from channel_flow_amd.config import ChannelConfig
    from channel_flow_amd.solver import Solver
    from oracle import oracle as O
    rng = np.random.default_rng(5)
    for periodic in (True, False):
        cfg = ChannelConfig(*grid, gamma_metric=1.5, x_periodic=periodic)
        met = O.metrics(cfg)
        U = [np.ascontiguousarray(1.0 + 0.1 * rng.standard_normal(cfg.shape)) for _ in range(5)]
        U[4] += 2.5e5  # rho E well above the kinetic part: positive pressure
        with Solver(cfg) as s:
            s.Jacobian_all(1.5, want_arrays=False)
            got = [a.copy() for a in U]
            s.boundary_conditions(got, met["J"])
        g, ny, nz = cfg.gend, cfg.ny, cfg.nz
        lo, hi = (g, 3) if periodic else (3, g)
        reg = (slice(1, ny + 2), slice(1, nz + 2))
        for v in range(5):
            np.testing.assert_array_equal(got[v][2][reg], got[v][lo][reg])
            np.testing.assert_array_equal(got[v][g + 1][reg], got[v][hi][reg])
        # and the oracle's ghosts in the same mode, bit for bit
        ref = [a.copy() for a in U]
        O.halo_bc(cfg, ref, met["J"])
        for v in range(5):
            np.testing.assert_array_equal(got[v][2:g + 2][(slice(None),) + reg], ref[v][2:g + 2][(slice(None),) + reg])


def test_gpu_x_periodic_trajectory():
    """Three physical steps (at most 6 inner iterations each) with true
    periodic x against the oracle in the same mode: the iteration history,
    er to 1e-10 and both levels of the state to 1e-10."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    from oracle import oracle as O
    cfg = ChannelConfig(16, 12, 16, gamma_metric=1.5, deltaT=5e-8, iteration_end_step=6, x_periodic=True)
    Qn = poiseuille_state(cfg, FlowParams(), seed=9, stream=0)
    # oracle: main.cpp:143-498 with the dual-time path
    met = O.metrics(cfg)
    U = O.state_from_Q(cfg, Qn, met["J"])
    Uq = [a.copy() for a in U]
    Ui = [a.copy() for a in U]
    E = np.zeros(5)
    rows = []
    g, ny, nz = cfg.gend, cfg.ny, cfg.nz
    sl = (slice(3, g + 1), slice(2, ny + 1), slice(2, nz + 1))
    for step in range(1, 4):
        for it in range(1, 100):
            er = O.inner_iteration(cfg, Ui, U, Uq, met)
            E = np.maximum(E, er)
            rows.append([step, it, *er])
            if ((er / E)[:4] < cfg.tol).all() or it == cfg.iteration_end_step:
                for v in range(5):
                    Uq[v][sl] = U[v][sl]
                    U[v][sl] = Ui[v][sl]
                break
    ref_q = np.stack([U[v][sl] * met["J"][sl] for v in range(5)] + [Uq[v][sl] * met["J"][sl] for v in range(5)])
    ref = np.array(rows)
    got = []
    with Solver(cfg) as s:
        s.Jacobian_all(1.5, want_arrays=False)
        s.load_state(Qn, Qn)
        E = np.zeros(5)
        for step in range(1, 4):
            for it in range(1, 100):
                er = s.inner_iteration()
                E = np.maximum(E, er)
                got.append([step, it, *er])
                if ((er / E)[:4] < cfg.tol).all() or it == cfg.iteration_end_step:
                    break
            s.advance_levels()
        qn, qm, _ = s.read_state()
    got = np.array(got)
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got[:, :2], ref[:, :2])
    np.testing.assert_allclose(got[:, 2:], ref[:, 2:], rtol=1e-10, atol=0)
    assert rel_l2(np.concatenate([qn, qm]), ref_q) <= TOL_STATE
    # the mode matters: the zero-gradient run ends elsewhere
    cfg0 = ChannelConfig(16, 12, 16, gamma_metric=1.5, deltaT=5e-8, iteration_end_step=6, x_periodic=False)
    with Solver(cfg0) as s:
        s.Jacobian_all(1.5, want_arrays=False)
        s.load_state(Qn, Qn)
        s.inner_iteration()
        q0 = s.read_state()[2]
    assert rel_l2(q0, s_first_iterate(cfg, Qn)) > 1e-12


def s_first_iterate(cfg, Qn):
    from channel_flow_amd.solver import Solver
    with Solver(cfg) as s:
        s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
        s.load_state(Qn, Qn)
        s.inner_iteration()
        return s.read_state()[2]


def test_gpu_host_arrays_keep_resident_state(tmp_path):
    """With a state loaded, the per-function entry points given host arrays
    (fluxes, viscous terms, Dual_time_stepping, Driven_force, Statistic) leave
    the resident iterate and levels as they were: the read-back state is
    unchanged and the next inner iteration gives the same bits as in a
    context that never made those calls."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    cfg = ChannelConfig(16, 12, 16, gamma_metric=1.5, deltaT=5e-8)
    Qn = poiseuille_state(cfg, FlowParams(), seed=21, stream=0)
    Qm = poiseuille_state(cfg, FlowParams(), seed=21, stream=1)
    other = poiseuille_state(cfg, FlowParams(), seed=22, stream=0)
    with Solver(cfg) as ref:
        ref.Jacobian_all(1.5, want_arrays=False)
        ref.load_state(Qn, Qm)
        ref.inner_iteration()
        er_ref = ref.inner_iteration()
        st_ref = ref.read_state()
    with Solver(cfg) as s:
        met = s.Jacobian_all(1.5)
        s.load_state(Qn, Qm)
        s.inner_iteration()
        before = s.read_state()
        from oracle import oracle as O
        Uo = O.state_from_Q(cfg, other, met["J"])
        U_ = [a.copy() for a in Uo]
        s.boundary_conditions(U_, met["J"])
        Fx = s.Flux_X(U_, met["J"], met["J_u"], met["m_u"])
        Fy = s.Flux_Y(U_, met["J"], met["J_v"], met["m_v"])
        Fz = s.Flux_Z(U_, met["J"], met["J_w"], met["m_w"])
        vF = s.Viscous_terms(U_, met)
        s.Dual_time_stepping([a.copy() for a in U_], U_, U_, met, Fx, Fy, Fz, vF)
        s.Driven_force(U2_=U_[1], J=met["J"])
        s.Statistic(1, 4, U_, tmp_path)
        after = s.read_state()
        for a, b in zip(before, after):
            np.testing.assert_array_equal(a, b)
        er = s.inner_iteration()
        st = s.read_state()
    np.testing.assert_array_equal(er, er_ref)
    for a, b in zip(st, st_ref):
        np.testing.assert_array_equal(a, b)


def test_gpu_trajectory_64cubed_config1_beta2(golden):
    """Config 1 as stated: 64^3, tanh beta = 2.0, Re_b 5600, Mach 0.2, the
    laminar driving force, 10 physical steps x 20 inner iterations, against
    the oracle run (otraj_b2_g64x64x64): iteration history, er to 1e-10, the
    sums of the ten LES blocks to 1e-12 and the plane i = 32 to 1e-10."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
    from make_oracle_traj import config1_b2_cfg
    from channel_flow_amd.config import poiseuille_state
    from channel_flow_amd.solver import Solver
    z = golden("otraj_b2_g64x64x64")
    cfg, flow = config1_b2_cfg()
    Qn = poiseuille_state(cfg, flow, seed=1, stream=0)
    rows = []
    with Solver(cfg) as s:
        s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
        s.load_state(Qn, Qn)
        E = np.zeros(5)
        for step in range(1, 11):
            for it in range(1, 100):
                er = s.inner_iteration()
                E = np.maximum(E, er)
                rows.append([step, it, *er, *(er / E)])
                if ((er / E)[:4] < cfg.tol).all() or it == cfg.iteration_end_step:
                    break
            s.advance_levels()
        qn, qm, _ = s.read_state()
    rows = np.array(rows)
    les = np.concatenate([qn, qm])
    ref = z["er"]
    assert rows.shape == ref.shape
    np.testing.assert_array_equal(rows[:, :2], ref[:, :2])
    np.testing.assert_allclose(rows[:, 2:], ref[:, 2:], rtol=1e-10, atol=0)
    np.testing.assert_allclose(les.reshape(10, -1).sum(axis=1), z["les_sum"], rtol=1e-12)
    assert rel_l2(les[:, 32], z["les_plane"]) <= TOL_STATE
