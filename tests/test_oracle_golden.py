"""The CPU oracle against golden vectors of the reference (bit-exact).

The fixtures in tests/golden/ were written by tests/golden/make_golden.py,
which runs the reference solver compiled from its own sources
(oracle/ref_build.sh).  Each reference function's inputs are fed to the
oracle's restatement and the outputs must agree bit for bit.
"""
import numpy as np
import pytest

from conftest import UNIT_GRIDS, VAR_GRIDS, fields, fixture_metrics, limiter_mask, masks, unit_cfg
from oracle import oracle as O


@pytest.mark.parametrize("grid", UNIT_GRIDS)
def test_metrics_match_reference(grid, golden):
    """Jacobian.cpp, Jacobian_u/_v/_w.cpp incl. ghost copies and wall faces."""
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    met = O.metrics(cfg)
    ref = fixture_metrics(z)
    for key in ("J", "J_u", "J_v", "J_w"):
        np.testing.assert_array_equal(met[key], ref[key], err_msg=key)
    for key in ("m", "m_u", "m_v", "m_w"):
        for c in range(9):
            np.testing.assert_array_equal(met[key][c], ref[key][c], err_msg=f"{key}[{c}]")


@pytest.mark.parametrize("grid", UNIT_GRIDS)
def test_hyperplane_counts_match_reference(grid, golden):
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    hp = z["hyperplane"]
    counts = O.hyperplane_counts(unit_cfg(grid))
    np.testing.assert_array_equal(counts, hp[:, 2])
    np.testing.assert_array_equal(np.cumsum(counts), hp[:, 1])


@pytest.mark.parametrize("grid", UNIT_GRIDS)
def test_boundary_conditions_match_reference(grid, golden):
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    U_ = [a.copy() for a in fields(z, "U_pre")]
    O.halo_bc(cfg, U_, np.ascontiguousarray(z["J"]))
    m = masks(cfg)["bc"]
    for v, ref in enumerate(fields(z, "U_bc")):
        np.testing.assert_array_equal(U_[v][m], ref[m])


@pytest.mark.parametrize("grid", UNIT_GRIDS)
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_fluxes_match_reference(grid, direction, golden):
    """Flux_X/Y/Z: MUSCL5 with boundary fall-backs + Roe (ROE == 1)."""
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    met = fixture_metrics(z)
    Jf = [met["J_u"], met["J_v"], met["J_w"]][direction]
    mf = [met["m_u"], met["m_v"], met["m_w"]][direction]
    out = O.flux(cfg, direction, fields(z, "U_bc"), met["J"], Jf, mf)
    name = "inF" + "xyz"[direction]
    m = masks(cfg)["xyz"[direction]]
    for v, ref in enumerate(fields(z, name)):
        np.testing.assert_array_equal(out[v][m], ref[m], err_msg=f"{name}{v + 1}")


@pytest.mark.parametrize("grid", VAR_GRIDS)
@pytest.mark.parametrize("direction", [0, 1, 2])
@pytest.mark.parametrize("sensor", ["L", "S"])
def test_limiter_fluxes_match_reference(grid, direction, sensor, golden):
    """Flux_X/Y/Z_limiter on a state with every variable perturbed, with the
    Filter sensor (L) and with a smooth sensor field (S): every defined face
    bit for bit (limiter_mask)."""
    z = golden(f"var_g{grid[0]}x{grid[1]}x{grid[2]}")
    met = fixture_metrics(golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}"))
    cfg = unit_cfg(grid)
    Jf = [met["J_u"], met["J_v"], met["J_w"]][direction]
    mf = [met["m_u"], met["m_v"], met["m_w"]][direction]
    ep = "Ep" + "XYZ"[direction] + ("_s" if sensor == "S" else "")
    out = O.flux_limiter(cfg, direction, fields(z, "U_bc"), met["J"], Jf, mf, np.ascontiguousarray(z[ep]))
    name = "inF" + "xyz"[direction] + sensor
    m = limiter_mask(cfg, direction)
    assert m.sum() > 0
    for v, ref in enumerate(fields(z, name)):
        assert not np.isnan(ref[m]).any()
        np.testing.assert_array_equal(out[v][m], ref[m], err_msg=f"{name}{v + 1}")


@pytest.mark.parametrize("grid", UNIT_GRIDS)
@pytest.mark.parametrize("direction", [0, 2])
def test_limiter_nan_on_flat_density_matches_reference(grid, direction, golden):
    """With a density that is constant along x and z the reference's slope
    ratios are 0/0 and its min/max macros pass the NaN on: the x and z
    limiter fluxes are NaN on exactly the same faces."""
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    met = fixture_metrics(z)
    Jf = [met["J_u"], met["J_v"], met["J_w"]][direction]
    mf = [met["m_u"], met["m_v"], met["m_w"]][direction]
    ep = "Ep" + "XYZ"[direction] + "_s"
    out = O.flux_limiter(cfg, direction, fields(z, "U_bc"), met["J"], Jf, mf, np.ascontiguousarray(z[ep]))
    name = "inF" + "xyz"[direction] + "S"
    m = limiter_mask(cfg, direction)
    for v, ref in enumerate(fields(z, name)):
        assert np.isnan(ref[m]).any()
        np.testing.assert_array_equal(out[v][m], ref[m], err_msg=f"{name}{v + 1}")


@pytest.mark.parametrize("grid", UNIT_GRIDS)
def test_viscous_matches_reference(grid, golden):
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    out = O.viscous(cfg, fields(z, "U_bc"), fixture_metrics(z))
    m = masks(cfg)["int"]
    for v, ref in enumerate(fields(z, "vF", 4)):
        np.testing.assert_array_equal(out[v][m], ref[m], err_msg=f"vF{v + 2}")


@pytest.mark.parametrize("grid", UNIT_GRIDS)
@pytest.mark.parametrize("hyperplane", [False, True])
def test_dual_time_matches_reference(grid, hyperplane, golden):
    """Dual_time_stepping: RHS, two symmetric LU-SGS sweeps, update, norms."""
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    U_ = [a.copy() for a in fields(z, "U_bc")]
    er, MR, NL, NR = O.dual_time(cfg, U_, fields(z, "U"), fields(z, "Uq"), fixture_metrics(z), fields(z, "inFx"),
                                 fields(z, "inFy"), fields(z, "inFz"), fields(z, "vF", 4), hyperplane=hyperplane,
                                 deltaT=5e-8)
    m = masks(cfg)["int"]
    for v in range(5):
        np.testing.assert_array_equal(U_[v][m], z[f"U_dt{v + 1}"][m])
        np.testing.assert_array_equal(MR[v][m], z[f"MR_dt{v + 1}"][m])
        np.testing.assert_array_equal(NR[v][m], z[f"NR_dt{v + 1}"][m])
        np.testing.assert_array_equal(NL[v][m], z[f"NL_dt{v + 1}"][m])
    np.testing.assert_array_equal(er, z["er_dt"])


def run_oracle_trajectory(cfg, Qn, Qm, steps, cap, deltaT):
    """main.cpp:143-498 with the dual-time path, driven through the oracle."""
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
            if (ern[:4] < 0.001).all() or it == cap:
                for v in range(5):
                    Uq[v][sl] = U[v][sl]
                    U[v][sl] = Ui[v][sl]
                break
    les = np.stack([U[v][sl] * met["J"][sl] for v in range(5)] + [Uq[v][sl] * met["J"][sl] for v in range(5)])
    return np.array(rows), les


def test_trajectory_matches_reference(golden):
    """3 physical steps x up to 6 inner iterations, residual history and
    LES00003.bin state bit for bit (Output.cpp layout)."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    z = golden("traj_g16x12x16")
    cfg = ChannelConfig(16, 12, 16, gamma_metric=1.5, deltaT=5e-8)
    Qn = poiseuille_state(cfg, FlowParams(), seed=7, stream=0)
    rows, les = run_oracle_trajectory(cfg, Qn, Qn, steps=3, cap=6, deltaT=5e-8)
    np.testing.assert_array_equal(rows, z["er"])
    np.testing.assert_array_equal(les, z["les"])


def test_first_iteration_64cubed_matches_reference(golden):
    """First inner iteration of the 64^3 run (config 1) in the oracle."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    z = golden("traj_g64x64x64")
    cfg = ChannelConfig(64, 64, 64, gamma_metric=1.5, deltaT=5e-8)
    Qn = poiseuille_state(cfg, FlowParams(), seed=1, stream=0)
    met = O.metrics(cfg)
    U = O.state_from_Q(cfg, Qn, met["J"])
    Ui = [a.copy() for a in U]
    er = O.inner_iteration(cfg, Ui, U, [a.copy() for a in U], met)
    np.testing.assert_array_equal(er, z["er"][0, 2:7])


def _fpin(z):
    sh = tuple(z["pattern_shape"])
    return np.unpackbits(z["patterns"], axis=-1)[..., :sh[-1]].reshape(sh).astype(np.float64)


@pytest.mark.parametrize("prefix", ["unit", "var"])
@pytest.mark.parametrize("grid", VAR_GRIDS)
def test_grid_y_point_matches_reference(prefix, grid, golden):
    """Grid.cpp:50-141 with the hard-coded stretching 2.8 (Y_point and its
    ghost planes, the cells Filter reads)."""
    z = golden(f"{prefix}_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    _, Y = O.grid_points(cfg, 2.8)
    m = masks(cfg)["bc"]
    np.testing.assert_array_equal(Y[m], z["Y_point"][m])


@pytest.mark.parametrize("prefix", ["unit", "var"])
@pytest.mark.parametrize("grid", VAR_GRIDS)
def test_filter_sensor_matches_reference(prefix, grid, golden):
    """Filter(0.007, 1, 0): EpX/EpY/EpZ on their face layouts."""
    z = golden(f"{prefix}_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    Ep = O.filter_sensor(cfg, fields(z, "U_bc"), np.ascontiguousarray(z["Y_point"]), np.ascontiguousarray(z["J"]))
    for a in range(3):
        m = masks(cfg)["xyz"[a]]
        np.testing.assert_array_equal(Ep[a][m], z["Ep" + "XYZ"[a]][m])


@pytest.mark.parametrize("grid", VAR_GRIDS)
def test_filter_ratios_pinned_by_thresholds(grid, golden):
    """The reference's sensor patterns at thresholds r and next(r) for sample
    faces: every face agrees at every threshold, and each sample face flips
    between the two, so the reference's ratio there is r bit for bit."""
    z = golden(f"var_g{grid[0]}x{grid[1]}x{grid[2]}")
    pin = golden(f"fpin_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    pats = _fpin(pin)
    U_, Yp, J = fields(z, "U_bc"), np.ascontiguousarray(z["Y_point"]), np.ascontiguousarray(z["J"])
    for r, crit in enumerate(pin["crits"]):
        Ep = O.filter_sensor(cfg, U_, Yp, J, crit=float(crit))
        for a in range(3):
            m = masks(cfg)["xyz"[a]]
            np.testing.assert_array_equal(Ep[a][m], pats[r][a][m], err_msg=f"run {r} Ep{'XYZ'[a]}")
    for s, (a, i, j, k) in enumerate(pin["faces"]):
        assert pats[2 * s][a][i, j, k] == 1.0 and pats[2 * s + 1][a][i, j, k] == 0.0


@pytest.mark.parametrize("grid", UNIT_GRIDS)
@pytest.mark.parametrize("prefix", ["unit", "var"])
def test_rk3nodt_matches_reference(grid, prefix, golden):
    """RK3NODT.cpp:115-183, stages 1-3 from the same post-BC state, fluxes
    and viscous terms, bit for bit against the reference's U_."""
    z = golden(f"{prefix}_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    m = masks(cfg)["int"]
    for RK in (1, 2, 3):
        U_ = [a.copy() for a in fields(z, "U_bc")]
        MR = O.rk3nodt(cfg, RK, U_, fields(z, "U"), np.ascontiguousarray(z["J"]), fields(z, "inFx"),
                       fields(z, "inFy"), fields(z, "inFz"), fields(z, "vF", 4), deltaT=5e-8)
        for v in range(5):
            np.testing.assert_array_equal(U_[v][m], z[f"U_rk{RK}_{v + 1}"][m], err_msg=f"stage {RK} U{v + 1}_")
        # stage 1 is U + dt Rk exactly as formed: MR = U_ - U there
        if RK == 1:
            for v in range(5):
                np.testing.assert_array_equal(U_[v][m], (1.0 * fields(z, "U")[v] + 0.0 * z[f"U_bc{v + 1}"]
                                                         + 1.0 * MR[v])[m])


def test_rk3_step_stage_structure():
    """The oracle's three-stage pass: stage 1 from U_ = U is U + dt R(U), and
    with the viscous terms off the pass differs from the viscous one only
    through vF."""
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    cfg = ChannelConfig(10, 8, 9, gamma_metric=1.5, deltaT=5e-8)
    met = O.metrics(cfg)
    U = O.state_from_Q(cfg, poiseuille_state(cfg, FlowParams(), seed=2, stream=0), met["J"])
    a, b = [x.copy() for x in U], [x.copy() for x in U]
    O.rk3_step(cfg, a, U, met, with_viscous=True)
    O.rk3_step(cfg, b, U, met, with_viscous=False)
    m = masks(cfg)["int"]
    for v in range(5):
        assert np.isfinite(a[v][m]).all() and np.isfinite(b[v][m]).all()
    assert any(not np.array_equal(a[v][m], b[v][m]) for v in range(1, 5))
    # within one stage the mass equation carries no viscous term
    O.halo_bc(cfg, U, met["J"])
    F = [O.flux(cfg, d, U, met["J"], met[f"J_{x}"], met[f"m_{x}"]) for d, x in enumerate("uvw")]
    rng = np.random.default_rng(4)
    c, d = [x.copy() for x in U], [x.copy() for x in U]
    O.rk3nodt(cfg, 1, c, U, met["J"], *F, [rng.standard_normal(U[0].shape) for _ in range(4)])
    O.rk3nodt(cfg, 1, d, U, met["J"], *F, O.zeros(cfg, 4))
    np.testing.assert_array_equal(c[0][m], d[0][m])
    assert all(not np.array_equal(c[v][m], d[v][m]) for v in range(1, 5))


@pytest.mark.parametrize("grid", VAR_GRIDS)
def test_statistic_matches_reference(grid, golden):
    """Statistic.cpp:1043-1154: st00002.dat (iterate and level n averaged
    over two steps) and st00003.dat (the iterate alone) byte for byte, so
    the sum order, the scaling and the "%.16f" format are the reference's."""
    z = golden(f"var_g{grid[0]}x{grid[1]}x{grid[2]}")
    st = golden(f"stat_g{grid[0]}x{grid[1]}x{grid[2]}")
    S = O.Statistic(unit_cfg(grid))
    assert S(1, 2, fields(z, "U_bc")) is None
    assert S(2, 2, fields(z, "U")) == st["st2"].tobytes().decode()
    assert S(3, 1, fields(z, "U_bc")) == st["st3"].tobytes().decode()


@pytest.mark.parametrize("grid", VAR_GRIDS)
def test_driven_force_matches_reference(grid, golden):
    """Driven_force.cpp:38-73: the lines the reference prints for the
    iterate and for level n (its only output)."""
    z = golden(f"var_g{grid[0]}x{grid[1]}x{grid[2]}")
    want = golden(f"force_g{grid[0]}x{grid[1]}x{grid[2]}")["lines"].tobytes().decode()
    cfg = unit_cfg(grid)
    got = O.driven_force(cfg, z["U_bc2"], z["J"])[2] + O.driven_force(cfg, z["U2"], z["J"])[2]
    assert got == want


def test_driven_force_closed_form():
    """A uniform rho*u gives Qall = rho*u (the weights 1/J / Jsum sum to one);
    a plane Poiseuille profile gives the discrete bulk flux."""
    cfg = unit_cfg((12, 10, 16))
    met = O.metrics(cfg)
    J = met["J"]
    U2 = np.zeros(cfg.shape)
    sl = (slice(3, cfg.gend + 1), slice(2, cfg.ny + 1), slice(2, cfg.nz + 1))
    U2[sl] = 0.3 / J[sl]
    Q, E, _ = O.driven_force(cfg, U2, J, Q0=0.3)
    assert abs(Q - 0.3) <= 1e-14 and E <= 1e-13
    w = 1.0 / J[3, 2:cfg.ny + 1, 3]
    prof = np.linspace(0.1, 1.0, cfg.ny - 1)
    U2[sl] = (prof[None, :, None] / J[sl])
    Q, _, _ = O.driven_force(cfg, U2, J)
    assert abs(Q - (prof * w).sum() / w.sum()) <= 1e-13


@pytest.mark.parametrize("grid", [(16, 12, 16), (9, 10, 13)])
def test_oracle_x_ghosts_closed_form(grid):
    """main.cpp:205-240: with np >= 2 (x_periodic) ghost plane 2 holds plane
    gend and gend+1 holds plane 3; with one rank the self-Sendrecv copies the
    neighbours (plane 3 and plane gend).  The walls and the z ghosts of the
    ghost planes follow from those planes' values, so whole planes agree."""
    from channel_flow_amd.config import ChannelConfig
    rng = np.random.default_rng(3)
    for periodic in (True, False):
        cfg = ChannelConfig(*grid, gamma_metric=1.5, x_periodic=periodic)
        met = O.metrics(cfg)
        U = [np.ascontiguousarray(1.0 + 0.1 * rng.standard_normal(cfg.shape)) for _ in range(5)]
        U[4] += 2.5e5
        O.halo_bc(cfg, U, met["J"])
        g, ny, nz = cfg.gend, cfg.ny, cfg.nz
        lo, hi = (g, 3) if periodic else (3, g)
        reg = (slice(1, ny + 2), slice(1, nz + 2))
        for v in range(5):
            np.testing.assert_array_equal(U[v][2][reg], U[v][lo][reg])
            np.testing.assert_array_equal(U[v][g + 1][reg], U[v][hi][reg])
        # a planted mistake (source planes swapped) would break it
        assert not np.array_equal(U[0][2][reg], U[0][hi][reg])


@pytest.mark.parametrize("grid", VAR_GRIDS)
def test_shear_stress_matches_reference(grid, golden):
    """Shear_stress.dat (Statistic.cpp:552-911) over four steps on post-BC
    iterates with statistic_step 3, byte for byte against the reference's
    file (the fourth line follows the reset of Statistic.cpp:1008-1018)."""
    import sys
    from conftest import GOLDEN
    sys.path.insert(0, str(GOLDEN))
    from make_golden import perturbed_state
    cfg = unit_cfg(grid)
    met = O.metrics(cfg)
    seed = 21 + VAR_GRIDS.index(grid)
    iterates = [perturbed_state(cfg, seed, 2), perturbed_state(cfg, seed, 3)]
    sh = O.ShearStress(cfg)
    text = ""
    for step in range(1, 5):
        U_ = O.state_from_Q(cfg, iterates[(step - 1) % 2], met["J"])
        O.halo_bc(cfg, U_, met["J"])
        text += sh(step, 3, U_, met)
    want = golden(f"shear_g{grid[0]}x{grid[1]}x{grid[2]}")["lines"].tobytes().decode()
    assert text == want


def test_shear_stress_laminar_closed_form():
    """For the noise-free laminar Poiseuille state u = 1.5 U_b (1 - (y/h - 1)^2)
    (uniform density and temperature, v = w = 0) the wall values are
    mu (u_2 - u_ghost) / (deltaET y_eta) with the no-slip ghost u_ghost = -u_2
    and y_eta the stretching's analytic derivative at the first cell centre
    (the wall-face metrics are that cell's, Jacobian_v.cpp:124-141), the same
    on every face and opposite on the two walls; and as the grid is refined
    they tend to the exact wall stress 3 mu U_b / h of the parabola."""
    import math
    from channel_flow_amd.config import MU_L, ChannelConfig, FlowParams, poiseuille_state
    flow = FlowParams(noise=0.0)
    errs = []
    for Y in (16, 32, 64):
        cfg = ChannelConfig(8, Y, 8, gamma_metric=2.0)
        met = O.metrics(cfg)
        U_ = O.state_from_Q(cfg, poiseuille_state(cfg, flow, seed=1), met["J"])
        O.halo_bc(cfg, U_, met["J"])
        vals = O.wall_shear(cfg, U_, met)
        # analytic: y(eta) = h (1 - tanh(g (1 - 2 eta)) / tanh g), eta_wall = 0,
        # first cell centre at eta = deltaET / 2
        h, g = cfg.high / 2, 2.0
        Ub, _ = flow.bulk_state(cfg.high)
        y2 = h * (1 - math.tanh(g * (1 - cfg.deltaET)) / math.tanh(g))
        u2 = 1.5 * Ub * (1 - (y2 / h - 1) ** 2)
        y_eta = 2 * h * g / math.tanh(g) / math.cosh(g * (1 - cfg.deltaET)) ** 2
        want = MU_L * 2 * u2 / (cfg.deltaET * y_eta)
        np.testing.assert_allclose(vals[0], want, rtol=1e-9)
        np.testing.assert_allclose(vals[1], -want, rtol=1e-9)
        errs.append(abs(vals[0].mean() - 3 * MU_L * Ub / h) / (3 * MU_L * Ub / h))
    assert errs[2] < errs[1] < errs[0] and errs[2] < 0.05
