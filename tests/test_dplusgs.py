"""DPLUSGS (DPLUSGS.cpp:189-1361) in the oracle.

Its right-hand side, D^-1 and update are those of Dual_time_stepping, so MR
is pinned by the reference's own golden vectors.  Its four point-Jacobi
sweeps are pinned through the LU-SGS sweeps, which the golden vectors pin
bit for bit: the LU-SGS solve is NR = M^-1 MR with M = (D + L) D^-1 (D + U),
so probing it with unit vectors gives M.  Because L and U couple only face
neighbours and L D^-1 U only adds to the diagonal blocks and to cells that
are not face neighbours, the face-neighbour blocks of M are L and U, and
the diagonal blocks give D in lexicographic order.  Four Jacobi steps
x <- D^-1 (MR - (L + U) x) from x = D^-1 MR, done as dense linear algebra
on those matrices, must then equal the oracle's sweeps to rounding.  A
dropped neighbour, a wrong sign or a neighbour read at the wrong index in
the Jacobi restatement fails that check.

The reference reads the upper neighbours' iterate at indices it never
assigns (DPLUSGS.cpp:995-1051, undefined behaviour). The restatement reads
them at j+1 and k+1, where the reference reads their states and metrics.
No reference run fixes that choice.
"""
import numpy as np
import pytest

from conftest import UNIT_GRIDS, fields, fixture_metrics, masks, unit_cfg

from oracle import oracle as O


@pytest.mark.parametrize("grid", UNIT_GRIDS)
def test_dplusgs_rhs_matches_reference(grid, golden):
    """MR (the preconditioned unsteady residual) bit for bit against the
    reference's Dual_time_stepping outputs; the norms are finite and the
    update changes the state."""
    z = golden(f"unit_g{grid[0]}x{grid[1]}x{grid[2]}")
    cfg = unit_cfg(grid)
    U_ = [a.copy() for a in fields(z, "U_bc")]
    er, MR, NL, NR = O.dual_time(cfg, U_, fields(z, "U"), fields(z, "Uq"), fixture_metrics(z), fields(z, "inFx"),
                                 fields(z, "inFy"), fields(z, "inFz"), fields(z, "vF", 4), deltaT=5e-8,
                                 dplusgs=True)
    m = masks(cfg)["int"]
    for v in range(5):
        np.testing.assert_array_equal(MR[v][m], z[f"MR_dt{v + 1}"][m])
        np.testing.assert_array_equal(NL[v][m], NR[v][m])
    assert np.isfinite(er).all() and (er > 0).all()
    assert any((U_[v][m] != z[f"U_bc{v + 1}"][m]).any() for v in range(5))


def _probe_setup():
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    cfg = ChannelConfig(X_out=8, Y_out=8, Z_out=8, gamma_metric=1.5)
    met = O.metrics(cfg)
    rng = np.random.default_rng(11)
    Q = np.asarray(poiseuille_state(cfg, FlowParams(), seed=5))
    Q = Q * (1 + 0.02 * rng.standard_normal(Q.shape))
    U_ = O.state_from_Q(cfg, Q, met["J"])
    O.halo_bc(cfg, U_, met["J"])
    cells = [(i, j, k) for i in range(3, cfg.gend + 1) for j in range(2, cfg.ny + 1) for k in range(2, cfg.nz + 1)]
    return cfg, met, U_, cells, rng


def _apply(cfg, mode, U_, met, cells, vec, deltaT):
    MR = O.zeros(cfg, 5)
    for c, (i, j, k) in enumerate(cells):
        for q in range(5):
            MR[q][i, j, k] = vec[5 * c + q]
    NR = O.sweeps_apply(cfg, mode, U_, met, MR, deltaT=deltaT)
    return np.array([NR[q][i, j, k] for (i, j, k) in cells for q in range(5)])


@pytest.mark.parametrize("deltaT", [5e-8, 1e-6])
def test_dplusgs_sweeps_are_jacobi_on_the_lusgs_operators(deltaT):
    cfg, met, U_, cells, rng = _probe_setup()
    n = 5 * len(cells)
    Minv = np.stack([_apply(cfg, 0, U_, met, cells, e, deltaT) for e in np.eye(n)], axis=1)
    M = np.linalg.inv(Minv)
    idx = {c: t for t, c in enumerate(cells)}

    def blk(A, a, b):
        return A[5 * a:5 * a + 5, 5 * b:5 * b + 5]

    L, U, D = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    steps = ((1, 0, 0), (0, 1, 0), (0, 0, 1))
    for t, (i, j, k) in enumerate(cells):
        for e in steps:
            lo, hi = (i - e[0], j - e[1], k - e[2]), (i + e[0], j + e[1], k + e[2])
            if lo in idx:
                blk(L, t, idx[lo])[:] = blk(M, t, idx[lo])
            if hi in idx:
                blk(U, t, idx[hi])[:] = blk(M, t, idx[hi])
    for t, (i, j, k) in enumerate(cells):
        acc = blk(M, t, t).copy()
        for e in steps:
            lo = (i - e[0], j - e[1], k - e[2])
            if lo in idx:
                s = idx[lo]
                acc -= blk(L, t, s) @ np.linalg.inv(blk(D, s, s)) @ blk(U, s, t)
        blk(D, t, t)[:] = acc
    Dinv = np.linalg.inv(D)
    # the factorisation read off M reproduces all of M (the LU-SGS structure)
    R = M - (D + L + U + L @ Dinv @ U)
    assert np.abs(R).max() <= 1e-12 * np.abs(M).max()
    # D^-1 couples only (1,1), (k,k) and (5,1) (Dual_time_stepping.cpp:480-507)
    pattern = np.zeros((5, 5), bool)
    pattern[[0, 1, 2, 3, 4, 4], [0, 1, 2, 3, 0, 4]] = True
    for t in range(len(cells)):
        b = blk(Dinv, t, t)
        assert np.abs(b[~pattern]).max() <= 1e-12 * np.abs(b).max()
    for trial in range(2):
        mr = rng.standard_normal(n)
        x = Dinv @ mr
        for _ in range(4):
            x = Dinv @ (mr - (L + U) @ x)
        y = _apply(cfg, 2, U_, met, cells, mr, deltaT)
        np.testing.assert_allclose(y, x, rtol=0, atol=1e-12 * np.abs(x).max())
