"""Shear_stress.dat on the GPU (cf_statistic; Statistic.cpp:552-911).

- The reference's file (shear_g*: four steps on post-BC iterates with
  statistic_step 3, the fourth after the reset of Statistic.cpp:1008-1018)
  byte for byte: the wall values are the reference's operations in its order
  and the running sums are added on the host in its order.
- The device iterate after physical steps against the oracle's file from the
  same iterates, and a checkpoint restart that continues the file exactly.
"""
import sys

import numpy as np
import pytest

from conftest import GOLDEN, VAR_GRIDS, gpu_available, unit_cfg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


@pytest.mark.parametrize("grid", VAR_GRIDS)
def test_gpu_shear_stress_matches_reference(grid, golden, tmp_path):
    from oracle import oracle as O
    from channel_flow_amd.solver import Solver
    sys.path.insert(0, str(GOLDEN))
    from make_golden import perturbed_state
    cfg = unit_cfg(grid)
    seed = 21 + VAR_GRIDS.index(grid)
    iterates = [perturbed_state(cfg, seed, 2), perturbed_state(cfg, seed, 3)]
    with Solver(cfg) as s:
        met = s.Jacobian_all(1.5)
        for step in range(1, 5):
            U_ = O.state_from_Q(cfg, iterates[(step - 1) % 2], met["J"])
            s.boundary_conditions(U_, met["J"])
            s.Statistic(step, 3, U_, tmp_path)
    want = golden(f"shear_g{grid[0]}x{grid[1]}x{grid[2]}")["lines"].tobytes().decode()
    assert (tmp_path / "Shear_stress.dat").read_text() == want


def test_gpu_shear_stress_device_state_and_restart(tmp_path):
    """Four physical steps with the statistic of the device iterate after
    each and a checkpoint after the second: a context restarted from the
    checkpoint writes the last two lines again, byte for byte; and on host
    arrays the device's line equals the oracle's."""
    from oracle import oracle as O
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    cfg = ChannelConfig(16, 12, 16, gamma_metric=2.0, deltaT=5e-8, iteration_end_step=4)
    Qn = poiseuille_state(cfg, FlowParams(), seed=17, stream=0)
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    ck = tmp_path / "c.ckpt"
    sh = O.ShearStress(cfg)
    with Solver(cfg) as s:
        s.Jacobian_all(2.0, want_arrays=False)
        s.load_state(Qn, Qn)
        for step in range(1, 5):
            s.physical_step()
            s.Statistic(step, 5, None, a)
            if step == 2:
                s.save_checkpoint(ck)
    lines = (a / "Shear_stress.dat").read_text().splitlines()
    assert len(lines) == 4 and all(np.isfinite(float(x)) for ln in lines for x in ln.split("\t"))
    with Solver(cfg) as s:
        s.Jacobian_all(2.0, want_arrays=False)
        s.load_checkpoint(ck)
        for step in (3, 4):
            s.physical_step()
            s.Statistic(step, 5, None, b)
    assert (b / "Shear_stress.dat").read_text().splitlines() == lines[2:]
    # the oracle's wall values from host arrays equal the device's on the same input
    with Solver(cfg) as s:
        mt = s.Jacobian_all(2.0)
        U_ = O.state_from_Q(cfg, Qn, mt["J"])
        s.boundary_conditions(U_, mt["J"])
        s.Statistic(1, 5, U_, b)
        text = sh(1, 5, U_, mt)
    assert (b / "Shear_stress.dat").read_text().splitlines()[-1] == text.strip("\n")
