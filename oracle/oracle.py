"""ctypes bindings of the CPU oracle (oracle/cf_oracle.c) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  Arrays are numpy float64 in the
reference layout [X_np][Y_m][Z_m] (Array.h), k fastest.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libcforacle.so"


class Grid(C.Structure):
    _fields_ = [
        ("X_out", C.c_int), ("Y_out", C.c_int), ("Z_out", C.c_int),
        ("deltaXI", C.c_double), ("deltaET", C.c_double), ("deltaZT", C.c_double),
        ("high", C.c_double), ("f_drive", C.c_double),
        ("x_periodic", C.c_int), ("n_sweeps", C.c_int),
    ]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        P = C.POINTER(C.c_double)
        PP = C.POINTER(P)
        _lib.cfo_metrics.argtypes = [C.POINTER(Grid), C.c_double, P, PP, P, PP, P, PP, P, PP]
        _lib.cfo_hyperplane_counts.argtypes = [C.POINTER(Grid), C.POINTER(C.c_int), C.c_int]
        _lib.cfo_hyperplane_counts.restype = C.c_int
        _lib.cfo_halo_bc.argtypes = [C.POINTER(Grid), PP, P]
        _lib.cfo_flux.argtypes = [C.POINTER(Grid), C.c_int, PP, P, P, PP, PP]
        _lib.cfo_viscous.argtypes = [C.POINTER(Grid), PP, P, P, PP, P, PP, P, PP, PP]
        _lib.cfo_flux_limiter.argtypes = [C.POINTER(Grid), C.c_int, PP, P, P, PP, P, PP]
        _lib.cfo_grid_points.argtypes = [C.POINTER(Grid), C.c_double, P, P]
        _lib.cfo_filter.argtypes = [C.POINTER(Grid), C.c_double, C.c_double, C.c_double, PP, P, P, P, P, P]
        _lib.cfo_filter_ratios.argtypes = [C.POINTER(Grid), PP, P, P, P, P, P]
        dt_args = [C.POINTER(Grid), C.c_double, PP, PP, PP, P, PP, PP, PP, PP, PP, PP, PP, PP, P]
        _lib.cfo_dual_time.argtypes = dt_args
        _lib.cfo_dual_time_hyperplane.argtypes = dt_args
        _lib.cfo_dplusgs.argtypes = dt_args
        _lib.cfo_sweeps_apply.argtypes = [C.POINTER(Grid), C.c_double, C.c_int, PP, P, PP, PP, PP, PP]
        _lib.cfo_statistic_accum.argtypes = [C.POINTER(Grid), PP, P]
        _lib.cfo_wall_shear.argtypes = [C.POINTER(Grid), PP, P, PP, P]
        _lib.cfo_driven_force.argtypes = [C.POINTER(Grid), P, P, C.c_double, P]
        _lib.cfo_rk3nodt.argtypes = [C.POINTER(Grid), C.c_int, C.c_double, PP, PP, P, PP, PP, PP, PP, PP]
    return _lib


def grid_of(cfg) -> Grid:
    return Grid(cfg.X_out, cfg.Y_out, cfg.Z_out, cfg.deltaXI, cfg.deltaET, cfg.deltaZT,
                cfg.high, cfg.f_drive, int(cfg.x_periodic), int(cfg.n_sweeps))


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pp(arrs):
    arr = (C.POINTER(C.c_double) * len(arrs))(*[_p(a) for a in arrs])
    return arr


def zeros(cfg, n=None):
    if n is None:
        return np.zeros(cfg.shape)
    return [np.zeros(cfg.shape) for _ in range(n)]


def metrics(cfg):
    """Return dict with J, m(9), J_u, m_u, J_v, m_v, J_w, m_w."""
    g = grid_of(cfg)
    J, Ju, Jv, Jw = zeros(cfg, 4)
    m, mu, mv, mw = zeros(cfg, 9), zeros(cfg, 9), zeros(cfg, 9), zeros(cfg, 9)
    lib().cfo_metrics(C.byref(g), cfg.gamma_metric, _p(J), _pp(m), _p(Ju), _pp(mu), _p(Jv), _pp(mv),
                      _p(Jw), _pp(mw))
    return dict(J=J, m=m, J_u=Ju, m_u=mu, J_v=Jv, m_v=mv, J_w=Jw, m_w=mw)


def hyperplane_counts(cfg):
    g = grid_of(cfg)
    n = cfg.gend + cfg.ny + cfg.nz - 6
    out = (C.c_int * n)()
    r = lib().cfo_hyperplane_counts(C.byref(g), out, n)
    assert r == n
    return np.array(out[:], dtype=np.int64)


def halo_bc(cfg, U_, J):
    lib().cfo_halo_bc(C.byref(grid_of(cfg)), _pp(U_), _p(J))


def flux(cfg, direction, U_, J, Jf, mf):
    out = zeros(cfg, 5)
    lib().cfo_flux(C.byref(grid_of(cfg)), int(direction), _pp(U_), _p(J), _p(Jf), _pp(mf), _pp(out))
    return out


def flux_limiter(cfg, direction, U_, J, Jf, mf, Ep):
    out = zeros(cfg, 5)
    lib().cfo_flux_limiter(C.byref(grid_of(cfg)), int(direction), _pp(U_), _p(J), _p(Jf), _pp(mf), _p(Ep),
                           _pp(out))
    return out


def grid_points(cfg, gamma=2.8):
    X = np.zeros(cfg.shape)
    Y = np.zeros(cfg.shape)
    lib().cfo_grid_points(C.byref(grid_of(cfg)), float(gamma), _p(X), _p(Y))
    return X, Y


def filter_sensor(cfg, U_, Y_point, J, crit=0.007, E_high=1.0, E_low=0.0):
    Ep = [np.zeros(cfg.shape) for _ in range(3)]
    lib().cfo_filter(C.byref(grid_of(cfg)), float(crit), float(E_high), float(E_low), _pp(U_), _p(Y_point),
                     _p(Ep[0]), _p(Ep[1]), _p(Ep[2]), _p(J))
    return Ep


def filter_ratios(cfg, U_, Y_point, J):
    R = [np.zeros(cfg.shape) for _ in range(3)]
    lib().cfo_filter_ratios(C.byref(grid_of(cfg)), _pp(U_), _p(Y_point), _p(J), _p(R[0]), _p(R[1]), _p(R[2]))
    return R


def viscous(cfg, U_, met):
    out = zeros(cfg, 4)
    lib().cfo_viscous(C.byref(grid_of(cfg)), _pp(U_), _p(met["J"]), _p(met["J_u"]), _pp(met["m_u"]),
                      _p(met["J_v"]), _pp(met["m_v"]), _p(met["J_w"]), _pp(met["m_w"]), _pp(out))
    return out


def dual_time(cfg, U_, U, Uq, met, inFx, inFy, inFz, vF, hyperplane=False, deltaT=None, dplusgs=False):
    """Updates U_ in place; returns (er[5], MR, NL, NR).  dplusgs: the
    DPLUSGS Jacobi sweeps instead of LU-SGS."""
    MR, NL, NR = zeros(cfg, 5), zeros(cfg, 5), zeros(cfg, 5)
    er = (C.c_double * 5)()
    fn = lib().cfo_dplusgs if dplusgs else (lib().cfo_dual_time_hyperplane if hyperplane else lib().cfo_dual_time)
    fn(C.byref(grid_of(cfg)), cfg.deltaT if deltaT is None else deltaT, _pp(U_), _pp(U), _pp(Uq),
       _p(met["J"]), _pp(met["m"]), _pp(inFx), _pp(inFy), _pp(inFz), _pp(vF), _pp(MR), _pp(NL), _pp(NR), er)
    return np.array(er[:]), MR, NL, NR


def sweeps_apply(cfg, mode, U_, met, MR, deltaT=None):
    """The implicit sweeps alone on a given MR (0: LU-SGS lexicographic,
    1: LU-SGS by planes, 2: DPLUSGS Jacobi); returns NR."""
    MR = [np.ascontiguousarray(a, dtype=np.float64) for a in MR]
    NL, NR = zeros(cfg, 5), zeros(cfg, 5)
    lib().cfo_sweeps_apply(C.byref(grid_of(cfg)), cfg.deltaT if deltaT is None else deltaT, int(mode), _pp(U_),
                           _p(met["J"]), _pp(met["m"]), _pp(MR), _pp(NL), _pp(NR))
    return NR


Q0_REFERENCE = 20.464969232814912  # prm.h:15


def driven_force(cfg, U2_, J, Q0=Q0_REFERENCE):
    """Driven_force.cpp:38-73: (Qall, ER_Q) and the line the reference prints."""
    out = (C.c_double * 2)()
    lib().cfo_driven_force(C.byref(grid_of(cfg)), _p(np.ascontiguousarray(U2_)), _p(np.ascontiguousarray(J)), Q0, out)
    return out[0], out[1], "Q0=%f\tQall=%f\tER_Q=%f\n" % (Q0, out[0], out[1])


class Statistic:
    """Statistic.cpp:1043-1154, the mean profiles: per-j sums accumulated over
    the calls of statistic_step steps, then averaged (1/statistic_step/
    X_out/Z_out) and written as st%05d.dat (seven blocks of "%.16f" lines,
    j = 2..ny: U, V, W, UU, VV, WW, UV), then reset."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.sums = np.zeros(7 * cfg.shape[1])

    def __call__(self, step, statistic_step, U_):
        lib().cfo_statistic_accum(C.byref(grid_of(self.cfg)), _pp(U_), _p(self.sums))
        if step % statistic_step != 0:
            return None
        c = self.cfg
        Ym = c.shape[1]
        inv = 1. / statistic_step / c.X_out / c.Z_out
        lines = []
        for q in range(7):
            for j in range(2, c.ny + 1):
                lines.append("%.16f\n" % (self.sums[q * Ym + j] * inv))
        self.sums[:] = 0
        return "".join(lines)


def wall_shear(cfg, U_, met):
    """Statistic.cpp:552-828: ML2 at the bottom and top wall faces, shape
    (2, X_out, Z_out) (i = 3..gend, k = 2..nz)."""
    out = np.zeros((2, cfg.X_out, cfg.Z_out))
    lib().cfo_wall_shear(C.byref(grid_of(cfg)), _pp(U_), _p(met["J_v"]), _pp(met["m_v"]), _p(out))
    return out


class ShearStress:
    """Statistic.cpp:879-911 and :1008-1018: the wall values are added to
    running sums (j, then i, then k, as the reference's SXYm[j]); each call
    writes "%f\t%f\n" of the change of the running averages SXYm[1] /
    X_out / Z_out and SXYm[ny] / X_out / Z_out, and the statistic step
    resets the sums and the previous averages."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.cum = [0.0, 0.0]
        self.prev = [0.0, 0.0]

    def __call__(self, step, statistic_step, U_, met):
        vals = wall_shear(self.cfg, U_, met)
        for w in range(2):
            s = self.cum[w]
            for v in vals[w].ravel().tolist():
                s = s + v
            self.cum[w] = s
        X, Z = self.cfg.X_out, self.cfg.Z_out
        avg = [self.cum[0] / X / Z, self.cum[1] / X / Z]
        line = "%f\t%f\n" % (avg[0] - self.prev[0], avg[1] - self.prev[1])
        self.prev = avg
        if step % statistic_step == 0:
            self.cum = [0.0, 0.0]
            self.prev = [0.0, 0.0]
        return line


def rk3nodt(cfg, RK, U_, U, J, inFx, inFy, inFz, vF, deltaT=None):
    """RK3NODT stage RK on U_ in place; returns MR = dt Rk."""
    MR = zeros(cfg, 5)
    lib().cfo_rk3nodt(C.byref(grid_of(cfg)), int(RK), cfg.deltaT if deltaT is None else deltaT, _pp(U_), _pp(U),
                      _p(J), _pp(inFx), _pp(inFy), _pp(inFz), _pp(vF), _pp(MR))
    return MR


def rk3_step(cfg, U_, U, met, with_viscous=True, deltaT=None):
    """One pass of main.cpp:152-377 with the RK3NODT path: three stages of
    ghosts, fluxes, viscous terms (vF = 0 when with_viscous is False, as when
    main.cpp leaves Viscous_terms commented out) and RK3NODT."""
    for RK in (1, 2, 3):
        halo_bc(cfg, U_, met["J"])
        inFx = flux(cfg, 0, U_, met["J"], met["J_u"], met["m_u"])
        inFy = flux(cfg, 1, U_, met["J"], met["J_v"], met["m_v"])
        inFz = flux(cfg, 2, U_, met["J"], met["J_w"], met["m_w"])
        vF = viscous(cfg, U_, met) if with_viscous else zeros(cfg, 4)
        rk3nodt(cfg, RK, U_, U, met["J"], inFx, inFy, inFz, vF, deltaT=deltaT)


def inner_iteration(cfg, U_, U, Uq, met, dplusgs=False):
    """One inner iteration of the dual-time path (main.cpp:156-353 with the DT
    calls enabled): ghosts, three fluxes, viscous terms, dual time step
    (dplusgs: DPLUSGS instead of Dual_time_stepping)."""
    halo_bc(cfg, U_, met["J"])
    inFx = flux(cfg, 0, U_, met["J"], met["J_u"], met["m_u"])
    inFy = flux(cfg, 1, U_, met["J"], met["J_v"], met["m_v"])
    inFz = flux(cfg, 2, U_, met["J"], met["J_w"], met["m_w"])
    vF = viscous(cfg, U_, met)
    er, MR, NL, NR = dual_time(cfg, U_, U, Uq, met, inFx, inFy, inFz, vF, dplusgs=dplusgs)
    return er


def state_from_Q(cfg, Q, J):
    """Interior Q (5,X,Y,Z) -> reference-layout U = Q/J (Initial_condition.cpp:100-110)."""
    U = zeros(cfg, 5)
    for v in range(5):
        U[v][3:cfg.gend + 1, 2:cfg.ny + 1, 2:cfg.nz + 1] = Q[v] / J[3:cfg.gend + 1, 2:cfg.ny + 1, 2:cfg.nz + 1]
    return U
