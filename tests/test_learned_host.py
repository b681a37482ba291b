"""Host logic of channel_flow_amd.learned without a GPU: a periodic patch made
with cells_of has the full channel's cell sizes, wall-normal grid, metrics and
driving force, which is what lets weights trained on the patch serve the full
grid (the network sees 3x3x3 neighbourhoods)."""
import numpy as np
import pytest

from channel_flow_amd.config import cell_y
from channel_flow_amd.learned import _seeds, channel_cfg


@pytest.mark.parametrize("full, patch, beta, re, dt", [((512, 384, 512), (64, 384, 64), 2.2, 10000.0, 2.5e-9),
                                                        ((256, 192, 256), (32, 192, 32), 2.0, 5600.0, 1e-8)])
def test_patch_has_the_full_grids_cells(full, patch, beta, re, dt):
    cf, ff = channel_cfg(full, beta, re, dt)
    cp, fp = channel_cfg(patch, beta, re, dt, cells_of=(full[0], full[2]))
    assert cp.X_out == patch[0] and cp.Z_out == patch[2]
    np.testing.assert_allclose([cp.deltaXI, cp.deltaET, cp.deltaZT], [cf.deltaXI, cf.deltaET, cf.deltaZT],
                               rtol=1e-15)
    np.testing.assert_array_equal(cell_y(cp), cell_y(cf))
    assert cp.f_drive == cf.f_drive and cp.gamma_metric == cf.gamma_metric and cp.deltaT == cf.deltaT
    assert fp.Re_bulk == ff.Re_bulk == re


def test_patch_metrics_equal_the_full_grids():
    """The j-profiles of the metrics (cell centres and the three face sets) of
    a 16x48x16 patch of the 64x48x64 channel equal the channel's."""
    from oracle import oracle as O
    cf, _ = channel_cfg((64, 48, 64), 2.2, 10000.0, 2.5e-9)
    cp, _ = channel_cfg((16, 48, 16), 2.2, 10000.0, 2.5e-9, cells_of=(64, 64))
    mp, mf = O.metrics(cp), O.metrics(cf)
    for key in ("J", "J_u", "J_v", "J_w"):
        np.testing.assert_allclose(mp[key][4, :, 4], mf[key][4, :, 4], rtol=1e-14, err_msg=key)


def test_patch_without_cells_of_is_the_whole_channel():
    c, _ = channel_cfg((32, 24, 40), 2.0, 5600.0, 5e-8)
    assert c.Lx == pytest.approx(2 * np.pi * 0.005) and c.Lz == pytest.approx(np.pi * 0.005)


def test_seed_ranges():
    assert _seeds("1-8") == list(range(1, 9)) and _seeds("9") == [9]
