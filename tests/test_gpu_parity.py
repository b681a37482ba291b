"""HIP path (libchannel_flow.so through the C ABI) against the oracle and the
reference's golden vectors.

The library is compiled with -ffp-contract=off, so every FP64 operation is
rounded once in the reference's order; the per-function outputs are expected
to match the reference bit for bit.  The residual norms are reduced in a
different (tree) order on the GPU, so er and the states downstream of a
physical step are compared with a relative tolerance of 1e-12 (relative L2
<= 1e-12 per kernel, state after 10 steps <= 1e-10).
"""
import numpy as np
import pytest

from conftest import (UNIT_GRIDS, VAR_GRIDS, fields, fixture_metrics, gpu_available, limiter_mask, masks, rel_l2,
                      unit_cfg)

pytestmark = pytest.mark.gpu

TOL_KERNEL = 1e-12
TOL_STATE = 1e-10


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def solver(cfg):
    from channel_flow_amd.solver import Solver
    return Solver(cfg)


@pytest.mark.parametrize("grid", UNIT_GRIDS)
def test_gpu_metrics_and_hyperplane(grid, golden):
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    with solver(cfg) as s:
        met = s.Jacobian_all(1.5)
        ref = fixture_metrics(z)
        for key in ("J", "J_u", "J_v", "J_w"):
            mk = {"J": "bc", "J_u": "x", "J_v": "y", "J_w": "z"}[key]
            m = masks(cfg)[mk]
            np.testing.assert_array_equal(met[key][m], ref[key][m], err_msg=key)
        counts = s.Hyperplane()
        np.testing.assert_array_equal(counts, z["hyperplane"][:, 2])


@pytest.mark.parametrize("grid", UNIT_GRIDS)
def test_gpu_boundary_conditions(grid, golden):
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    U_ = [a.copy() for a in fields(z, "U_pre")]
    with solver(cfg) as s:
        s.boundary_conditions(U_, np.ascontiguousarray(z["J"]))
    m = masks(cfg)["bc"]
    for v, ref in enumerate(fields(z, "U_bc")):
        np.testing.assert_array_equal(U_[v][m], ref[m])


@pytest.mark.parametrize("grid", UNIT_GRIDS)
def test_gpu_fluxes(grid, golden):
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    met = fixture_metrics(z)
    U_ = fields(z, "U_bc")
    with solver(cfg) as s:
        outs = [s.Flux_X(U_, met["J"], met["J_u"], met["m_u"]), s.Flux_Y(U_, met["J"], met["J_v"], met["m_v"]),
                s.Flux_Z(U_, met["J"], met["J_w"], met["m_w"])]
    for d in range(3):
        m = masks(cfg)["xyz"[d]]
        for v, ref in enumerate(fields(z, "inF" + "xyz"[d])):
            assert rel_l2(outs[d][v][m], ref[m]) <= TOL_KERNEL
            np.testing.assert_array_equal(outs[d][v][m], ref[m], err_msg=f"inF{'xyz'[d]}{v + 1}")


def _limiter_outputs(cfg, z, met, sensor):
    U_ = fields(z, "U_bc")
    sfx = "_s" if sensor == "S" else ""
    Ep = [np.ascontiguousarray(z[f"Ep{a}{sfx}"]) for a in "XYZ"]
    with solver(cfg) as s:
        return [s.Flux_X_limiter(U_, met["J"], met["J_u"], met["m_u"], Ep[0]),
                s.Flux_Y_limiter(U_, met["J"], met["J_v"], met["m_v"], Ep[1]),
                s.Flux_Z_limiter(U_, met["J"], met["J_w"], met["m_w"], Ep[2])], U_, Ep


@pytest.mark.parametrize("grid", VAR_GRIDS)
@pytest.mark.parametrize("sensor", ["L", "S"])
def test_gpu_limiter_fluxes(grid, sensor, golden):
    """Flux_X/Y/Z_limiter on the perturbed state: bit for bit against the
    reference on every face it defines, and against the oracle on every face
    (the near-wall y faces follow the documented face-J definition)."""
    from oracle import oracle as O
    z = golden(f"var_g{grid[0]}x{grid[1]}x{grid[2]}")
    met = fixture_metrics(golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}"))
    cfg = unit_cfg(grid)
    outs, U_, Ep = _limiter_outputs(cfg, z, met, sensor)
    Jf = [met["J_u"], met["J_v"], met["J_w"]]
    mf = [met["m_u"], met["m_v"], met["m_w"]]
    for d in range(3):
        ref_o = O.flux_limiter(cfg, d, [u.copy() for u in U_], met["J"], Jf[d], mf[d], Ep[d])
        m = limiter_mask(cfg, d)
        mall = masks(cfg)["xyz"[d]]
        for v, ref in enumerate(fields(z, "inF" + "xyz"[d] + sensor)):
            np.testing.assert_array_equal(outs[d][v][m], ref[m], err_msg=f"inF{'xyz'[d]}{sensor}{v + 1}")
            np.testing.assert_array_equal(outs[d][v][mall], ref_o[v][mall], err_msg=f"oracle {'xyz'[d]}{v + 1}")


@pytest.mark.parametrize("grid", UNIT_GRIDS)
def test_gpu_limiter_nan_on_flat_density(grid, golden):
    """Constant density along x and z: 0/0 slope ratios, NaN fluxes on the
    same faces as the reference (x and z)."""
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    met = fixture_metrics(z)
    outs, _, _ = _limiter_outputs(cfg, z, met, "S")
    for d in (0, 2):
        m = limiter_mask(cfg, d)
        for v, ref in enumerate(fields(z, "inF" + "xyz"[d] + "S")):
            np.testing.assert_array_equal(outs[d][v][m], ref[m], err_msg=f"inF{'xyz'[d]}S{v + 1}")


@pytest.mark.parametrize("grid", UNIT_GRIDS)
def test_gpu_viscous(grid, golden):
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    with solver(cfg) as s:
        out = s.Viscous_terms(fields(z, "U_bc"), fixture_metrics(z))
    m = masks(cfg)["int"]
    for v, ref in enumerate(fields(z, "vF", 4)):
        assert rel_l2(out[v][m], ref[m]) <= TOL_KERNEL
        np.testing.assert_array_equal(out[v][m], ref[m], err_msg=f"vF{v + 2}")


@pytest.mark.parametrize("grid", UNIT_GRIDS)
def test_gpu_dual_time(grid, golden):
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    U_ = [a.copy() for a in fields(z, "U_bc")]
    with solver(cfg) as s:
        er, MR, NR, NL = s.Dual_time_stepping(U_, fields(z, "U"), fields(z, "Uq"), fixture_metrics(z),
                                              fields(z, "inFx"), fields(z, "inFy"), fields(z, "inFz"),
                                              fields(z, "vF", 4), deltaT=5e-8)
    m = masks(cfg)["int"]
    for v in range(5):
        np.testing.assert_array_equal(MR[v][m], z[f"MR_dt{v + 1}"][m])
        np.testing.assert_array_equal(NR[v][m], z[f"NR_dt{v + 1}"][m])
        np.testing.assert_array_equal(NL[v][m], z[f"NL_dt{v + 1}"][m])
        np.testing.assert_array_equal(U_[v][m], z[f"U_dt{v + 1}"][m])
    np.testing.assert_allclose(er, z["er_dt"], rtol=TOL_KERNEL, atol=0)


@pytest.mark.parametrize("grid", UNIT_GRIDS)
def test_gpu_dplusgs(grid, golden):
    """DPLUSGS: MR bit for bit against the reference's golden RHS; the four
    Jacobi sweeps, NL/NR and the update bit for bit against the oracle; the
    norms to 1e-12 (tree reduction on the GPU)."""
    from oracle import oracle as O
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    args = (fields(z, "U"), fields(z, "Uq"), fixture_metrics(z), fields(z, "inFx"), fields(z, "inFy"),
            fields(z, "inFz"), fields(z, "vF", 4))
    U_o = [a.copy() for a in fields(z, "U_bc")]
    er_o, MR_o, NL_o, NR_o = O.dual_time(cfg, U_o, *args, deltaT=5e-8, dplusgs=True)
    U_ = [a.copy() for a in fields(z, "U_bc")]
    with solver(cfg) as s:
        er, MR, NR, NL = s.DPLUSGS(U_, *args, deltaT=5e-8)
    m = masks(cfg)["int"]
    for v in range(5):
        np.testing.assert_array_equal(MR[v][m], z[f"MR_dt{v + 1}"][m])
        np.testing.assert_array_equal(NR[v][m], NR_o[v][m])
        np.testing.assert_array_equal(NL[v][m], NL_o[v][m])
        np.testing.assert_array_equal(U_[v][m], U_o[v][m])
    np.testing.assert_allclose(er, er_o, rtol=TOL_KERNEL, atol=0)


def test_gpu_inner_iteration_dplusgs_16x12x16():
    """The device-resident inner iteration with the DPLUSGS solver selected,
    three iterations against the oracle: the iterate bit for bit, the norms
    to 1e-12."""
    from oracle import oracle as O
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    cfg = ChannelConfig(16, 12, 16, gamma_metric=1.5, deltaT=5e-8)
    Qn = poiseuille_state(cfg, FlowParams(), seed=7, stream=0)
    Qm = poiseuille_state(cfg, FlowParams(), seed=7, stream=1)
    met = O.metrics(cfg)
    U, Uq = O.state_from_Q(cfg, Qn, met["J"]), O.state_from_Q(cfg, Qm, met["J"])
    Ui = [a.copy() for a in U]
    er_o = [O.inner_iteration(cfg, Ui, U, Uq, met, dplusgs=True) for _ in range(3)]
    sl = (slice(3, cfg.gend + 1), slice(2, cfg.ny + 1), slice(2, cfg.nz + 1))
    qi_o = np.stack([Ui[v][sl] * met["J"][sl] for v in range(5)])
    with Solver(cfg) as s:
        s.set_implicit_solver(1)
        s.Jacobian_all(1.5, want_arrays=False)
        s.load_state(Qn, Qm)
        er = [s.inner_iteration() for _ in range(3)]
        qi = s.read_state()[2]
    np.testing.assert_array_equal(qi, qi_o)
    np.testing.assert_allclose(np.array(er), np.array(er_o), rtol=TOL_KERNEL, atol=0)


@pytest.mark.parametrize("grid", UNIT_GRIDS)
@pytest.mark.parametrize("prefix", ["unit", "var"])
def test_gpu_rk3nodt(grid, prefix, golden):
    """RK3NODT stages 1-3 bit for bit against the reference; MR = dt Rk
    bit for bit against the oracle."""
    from oracle import oracle as O
    z = golden(f"{prefix}_g{grid[0]}x{grid[1]}x{grid[2]}")
    # the perturbed-state fixtures share the flat-state fixture's grid metrics
    met = fixture_metrics(golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}"))
    cfg = unit_cfg(grid)
    m = masks(cfg)["int"]
    with solver(cfg) as s:
        for RK in (1, 2, 3):
            U_ = [a.copy() for a in fields(z, "U_bc")]
            MR = s.RK3NODT(RK, U_, fields(z, "U"), fields(z, "Uq"), met, fields(z, "inFx"),
                           fields(z, "inFy"), fields(z, "inFz"), fields(z, "vF", 4), deltaT=5e-8)
            Uo = [a.copy() for a in fields(z, "U_bc")]
            MRo = O.rk3nodt(cfg, RK, Uo, fields(z, "U"), np.ascontiguousarray(z["J"]), fields(z, "inFx"),
                            fields(z, "inFy"), fields(z, "inFz"), fields(z, "vF", 4), deltaT=5e-8)
            for v in range(5):
                np.testing.assert_array_equal(U_[v][m], z[f"U_rk{RK}_{v + 1}"][m])
                np.testing.assert_array_equal(MR[v][m], MRo[v][m])


@pytest.mark.parametrize("viscous", [True, False])
def test_gpu_rk3_step_16x12x16(viscous):
    """Two passes of the device-resident RK3NODT path (three stages of
    ghosts, fluxes, viscous terms or none, and the update) against the
    oracle, bit for bit."""
    from oracle import oracle as O
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    cfg = ChannelConfig(16, 12, 16, gamma_metric=1.5, deltaT=5e-8)
    Qn = poiseuille_state(cfg, FlowParams(), seed=11, stream=0)
    met = O.metrics(cfg)
    U = O.state_from_Q(cfg, Qn, met["J"])
    Ui = [a.copy() for a in U]
    for _ in range(2):
        O.rk3_step(cfg, Ui, U, met, with_viscous=viscous)
    sl = (slice(3, cfg.gend + 1), slice(2, cfg.ny + 1), slice(2, cfg.nz + 1))
    qi_o = np.stack([Ui[v][sl] * met["J"][sl] for v in range(5)])
    with Solver(cfg) as s:
        s.Jacobian_all(1.5, want_arrays=False)
        s.load_state(Qn, Qn)
        for _ in range(2):
            s.rk3_step(viscous)
        qi = s.read_state()[2]
    np.testing.assert_array_equal(qi, qi_o)


def _st_values(text):
    return np.array([float(x) for x in text.split()])


@pytest.mark.parametrize("grid", VAR_GRIDS)
def test_gpu_statistic(grid, golden, tmp_path):
    """Statistic's mean profiles on the GPU against the reference's
    st00002.dat / st00003.dat: same lines and format, values to 1e-13 of the
    velocity scale (the per-j sums are combined in a fixed tree, not in the
    reference's single running sum)."""
    z = golden(f"var_g{grid[0]}x{grid[1]}x{grid[2]}")
    st = golden(f"stat_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    with solver(cfg) as s:
        s.Statistic(1, 2, fields(z, "U_bc"), tmp_path)
        assert not (tmp_path / "st00001.dat").exists()
        s.Statistic(2, 2, fields(z, "U"), tmp_path)
        s.Statistic(3, 1, fields(z, "U_bc"), tmp_path)
    for n in (2, 3):
        ref = st[f"st{n}"].tobytes().decode()
        got = (tmp_path / f"st0000{n}.dat").read_text()
        assert got.count("\n") == ref.count("\n") == 7 * (cfg.ny - 1)
        assert all(len(x.split(".")[1]) == 16 for x in got.split())
        a, b = _st_values(got), _st_values(ref)
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-13 * np.abs(b).max())


def test_gpu_statistic_device_state(tmp_path):
    """The device iterate (U_ = NULL) gives the oracle's profiles."""
    from oracle import oracle as O
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    cfg = ChannelConfig(16, 12, 16, gamma_metric=1.5, deltaT=5e-8)
    Qn = poiseuille_state(cfg, FlowParams(), seed=5, stream=0)
    met = O.metrics(cfg)
    ref = O.Statistic(cfg)(4, 1, O.state_from_Q(cfg, Qn, met["J"]))
    with Solver(cfg) as s:
        s.Jacobian_all(1.5, want_arrays=False)
        s.load_state(Qn, Qn)
        s.Statistic(4, 1, None, tmp_path)
    a, b = _st_values((tmp_path / "st00004.dat").read_text()), _st_values(ref)
    np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-13 * np.abs(b).max())


@pytest.mark.parametrize("grid", VAR_GRIDS)
def test_gpu_driven_force(grid, golden):
    """Driven_force on the GPU: Qall to 1e-13 of the oracle's (the planes are
    summed in a fixed tree), ER_Q to what that allows, and the reference's
    printed line."""
    from oracle import oracle as O
    z = golden(f"var_g{grid[0]}x{grid[1]}x{grid[2]}")
    want = golden(f"force_g{grid[0]}x{grid[1]}x{grid[2]}")["lines"].tobytes().decode().splitlines(keepends=True)
    cfg = unit_cfg(grid)
    with solver(cfg) as s:
        for name, line in (("U_bc2", want[0]), ("U2", want[1])):
            Q, E = s.Driven_force(O.Q0_REFERENCE, z[name], z["J"])
            Qo, Eo, _ = O.driven_force(cfg, z[name], z["J"])
            np.testing.assert_allclose(Q, Qo, rtol=1e-13, atol=0)
            # ER_Q = |Q0 - Q| / Q0 cancels: bounded through Q's own tolerance
            assert abs(E - Eo) <= 1e-13 * Qo / O.Q0_REFERENCE
            assert "Q0=%f\tQall=%f\tER_Q=%f\n" % (O.Q0_REFERENCE, Q, E) == line


def test_gpu_checkpoint_restart_is_exact(tmp_path):
    """Two physical steps, a checkpoint, two more steps with statistics;
    a new context restarted from the checkpoint repeats the last two steps
    bit for bit: iteration counts, normalised residuals, state and st file."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    cfg = ChannelConfig(16, 12, 16, gamma_metric=1.5, deltaT=5e-8, iteration_end_step=6)
    Qn = poiseuille_state(cfg, FlowParams(), seed=13, stream=0)
    ck = tmp_path / "run.ckpt"
    runs = []
    for restart in (False, True):
        out = tmp_path / ("b" if restart else "a")
        out.mkdir()
        with Solver(cfg) as s:
            s.Jacobian_all(1.5, want_arrays=False)
            if restart:
                s.load_checkpoint(ck)
            else:
                s.load_state(Qn, Qn)
                for step in (1, 2):
                    s.physical_step()
                    s.Statistic(step, 4, None, out)
                s.save_checkpoint(ck)
            hist = []
            for step in (3, 4):
                it, ern = s.physical_step()
                hist.append((it, *ern))
                s.Statistic(step, 4, None, out)
            runs.append((np.array(hist), np.concatenate(s.read_state()), (out / "st00004.dat").read_text()))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    assert runs[0][2] == runs[1][2]


def gpu_trajectory(cfg, Qn, Qm, steps):
    rows = []
    with solver(cfg) as s:
        s.Jacobian_all(cfg.gamma_metric, want_arrays=False)
        s.load_state(Qn, Qm)
        E = np.zeros(5)
        for step in range(1, steps + 1):
            for it in range(1, 50000):
                er = s.inner_iteration()
                E = np.maximum(E, er)
                ern = er / E
                rows.append([step, it, *er, *ern])
                if (ern[:4] < cfg.tol).all() or it == cfg.iteration_end_step:
                    break
            s.advance_levels()
        qn, qm, _ = s.read_state()
    return np.array(rows), np.concatenate([qn, qm])


def test_gpu_trajectory_16x12x16(golden):
    """3 physical steps of the dual-time loop against the reference run."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    z = golden("traj_g16x12x16")
    cfg = ChannelConfig(16, 12, 16, gamma_metric=1.5, deltaT=5e-8, iteration_end_step=6)
    Qn = poiseuille_state(cfg, FlowParams(), seed=7, stream=0)
    rows, les = gpu_trajectory(cfg, Qn, Qn, steps=3)
    ref = z["er"]
    assert rows.shape == ref.shape
    np.testing.assert_array_equal(rows[:, :2], ref[:, :2])
    np.testing.assert_allclose(rows[:, 2:], ref[:, 2:], rtol=1e-10, atol=0)
    assert rel_l2(les, z["les"]) <= TOL_STATE


def test_gpu_physical_step_api_16x12x16(golden):
    """cf_physical_step (device-side shift, running maxima) against the reference."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    z = golden("traj_g16x12x16")
    cfg = ChannelConfig(16, 12, 16, gamma_metric=1.5, deltaT=5e-8, iteration_end_step=6)
    Qn = poiseuille_state(cfg, FlowParams(), seed=7, stream=0)
    ref = z["er"]
    with Solver(cfg) as s:
        s.Jacobian_all(1.5, want_arrays=False)
        s.load_state(Qn, Qn)
        for step in range(1, 4):
            it, ern = s.physical_step()
            last = ref[ref[:, 0] == step][-1]
            assert it == int(last[1])
            np.testing.assert_allclose(ern, last[7:], rtol=1e-10)
        qn, qm, _ = s.read_state()
    assert rel_l2(np.concatenate([qn, qm]), z["les"]) <= TOL_STATE


def test_gpu_trajectory_64cubed_config1(golden):
    """Config 1: 64^3, 10 physical steps x 20 inner iterations, residual
    history and state after 10 steps against the reference (<= 1e-10)."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    z = golden("traj_g64x64x64")
    cfg = ChannelConfig(64, 64, 64, gamma_metric=1.5, deltaT=5e-8, iteration_end_step=20)
    Qn = poiseuille_state(cfg, FlowParams(), seed=1, stream=0)
    rows, les = gpu_trajectory(cfg, Qn, Qn, steps=10)
    ref = z["er"]
    assert rows.shape == ref.shape
    np.testing.assert_array_equal(rows[:, :2], ref[:, :2])
    np.testing.assert_allclose(rows[:, 2:], ref[:, 2:], rtol=1e-10, atol=0)
    np.testing.assert_allclose(les.reshape(10, -1).sum(axis=1), z["les_sum"], rtol=1e-12)
    assert rel_l2(les[:, 32], z["les_plane"]) <= TOL_STATE


def test_gpu_repeatable_bits():
    """Two runs of the same inner iterations give identical bits (the
    residual reduction has a fixed order)."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    cfg = ChannelConfig(32, 24, 40, gamma_metric=2.0, deltaT=5e-8)
    Qn = poiseuille_state(cfg, FlowParams(), seed=3, stream=0)
    outs = []
    for _ in range(2):
        with Solver(cfg) as s:
            s.Jacobian_all(2.0, want_arrays=False)
            s.load_state(Qn, Qn)
            ers = [s.inner_iteration() for _ in range(3)]
            outs.append((np.array(ers), s.read_state()[2]))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("grid", [(21, 13, 18), (8, 8, 8), (40, 33, 9), (70, 45, 37)])
def test_gpu_sweep_schedules_identical(grid):
    """Pipelined tiles (0), one launch per hyperplane (1) and single-wave tiles
    (2) give the same bits (the DAG fixes the arithmetic), on grids whose
    sizes are not multiples of the 8x8 tile."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    cfg = ChannelConfig(*grid, gamma_metric=2.0, deltaT=1e-8)
    Qn = poiseuille_state(cfg, FlowParams(), seed=9, stream=0)
    Qm = poiseuille_state(cfg, FlowParams(), seed=9, stream=1)
    outs = []
    for schedule in (0, 1, 2):
        with Solver(cfg) as s:
            s.set_sweep_schedule(schedule)
            s.Jacobian_all(2.0, want_arrays=False)
            s.load_state(Qn, Qm)
            ers = [s.inner_iteration() for _ in range(3)]
            outs.append((np.array(ers), s.read_state()[2]))
    # the state is identical; the per-hyperplane norms are summed in a
    # different block order
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    np.testing.assert_array_equal(outs[0][1], outs[2][1])
    np.testing.assert_array_equal(outs[0][0], outs[2][0])
    np.testing.assert_allclose(outs[0][0], outs[1][0], rtol=1e-14, atol=0)


@pytest.mark.parametrize("grid", [(21, 13, 18), (8, 8, 8), (40, 33, 9), (70, 45, 37)])
def test_gpu_viscous_schedules_identical(grid):
    """The fused viscous kernel (0) and the three staged kernels (1) give the
    same bits over three inner iterations, on grids whose sizes are not
    multiples of the fused kernel's 8x32 tile or its 16-plane run."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    cfg = ChannelConfig(*grid, gamma_metric=2.0, deltaT=1e-8)
    Qn = poiseuille_state(cfg, FlowParams(), seed=4, stream=0)
    Qm = poiseuille_state(cfg, FlowParams(), seed=4, stream=1)
    outs = []
    for schedule in (0, 1):
        with Solver(cfg) as s:
            s.set_viscous_schedule(schedule)
            s.Jacobian_all(2.0, want_arrays=False)
            s.load_state(Qn, Qm)
            ers = [s.inner_iteration() for _ in range(3)]
            outs.append((np.array(ers), s.read_state()[2]))
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    np.testing.assert_array_equal(outs[0][0], outs[1][0])


@pytest.mark.parametrize("grid", VAR_GRIDS)
def test_gpu_grid_and_filter(grid, golden):
    """Grid (Y_point, stretching 2.8) and Filter(0.007, 1, 0) bit for bit
    against the reference, on both the flat-density and the perturbed state,
    and the sensor patterns at the pinning thresholds (each sample face
    flips between a ratio r and the next double, so the GPU's ratio there is
    the reference's bit for bit)."""
    cfg = unit_cfg(grid)
    pin = golden(f"fpin_g{grid[0]}x{grid[1]}x{grid[2]}")
    sh = tuple(pin["pattern_shape"])
    pats = np.unpackbits(pin["patterns"], axis=-1)[..., :sh[-1]].reshape(sh).astype(np.float64)
    with solver(cfg) as s:
        for prefix in ("unit", "var"):
            z = golden(f"{prefix}_g{grid[0]}x{grid[1]}x{grid[2]}")
            _, Yp = s.Grid(2.8)
            m = masks(cfg)["bc"]
            np.testing.assert_array_equal(Yp[m], z["Y_point"][m])
            Ep = s.Filter(0.007, 1.0, 0.0, fields(z, "U_bc"), np.ascontiguousarray(z["Y_point"]),
                          np.ascontiguousarray(z["J"]))
            for a in range(3):
                mk = masks(cfg)["xyz"[a]]
                np.testing.assert_array_equal(Ep[a][mk], z["Ep" + "XYZ"[a]][mk], err_msg=f"{prefix} Ep{'XYZ'[a]}")
        z = golden(f"var_g{grid[0]}x{grid[1]}x{grid[2]}")
        for r, crit in enumerate(pin["crits"]):
            Ep = s.Filter(float(crit), 1.0, 0.0, fields(z, "U_bc"), np.ascontiguousarray(z["Y_point"]),
                          np.ascontiguousarray(z["J"]))
            for a in range(3):
                mk = masks(cfg)["xyz"[a]]
                np.testing.assert_array_equal(Ep[a][mk], pats[r][a][mk], err_msg=f"run {r} Ep{'XYZ'[a]}")
