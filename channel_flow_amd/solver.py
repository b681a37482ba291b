"""Python mirror of the reference call surface over the C ABI.

`Solver` owns one `cf_ctx` (include/channel_flow.h) and exposes the
reference's functions under their own names (main.h:10-1069): `Jacobian_all`,
`Hyperplane`, `boundary_conditions`, `Flux_X/Y/Z`, `Viscous_terms`,
`Dual_time_stepping`, plus the device-resident driver (`inner_iteration`,
`physical_step`, `Output`).  Arrays are numpy float64 in the reference layout
[X_np][Y_m][Z_m].  All compute runs in libchannel_flow.so (HIP, gfx950); if
the library is missing, or no GPU is visible, construction raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .config import ChannelConfig

LIB_PATH = Path(__file__).resolve().parent / "libchannel_flow.so"

CF_OK = 0
# enum cf_precision
CF_FP64 = 0          # everything FP64, bit-identical to the reference
CF_FP32_SWEEPS = 1   # LU-SGS sweeps and split Jacobians in float (pipelined sweep schedule)
ERRORS = {-1: "CF_ERR_ARG", -2: "CF_ERR_HIP", -3: "CF_ERR_NONFINITE", -4: "CF_ERR_METRICS",
          -5: "CF_ERR_IO", -6: "CF_ERR_STATE", -7: "CF_ERR_TIMEOUT"}

EXPORTS = [
    "cf_create", "cf_destroy", "cf_last_error", "cf_field_elems", "cf_jacobian_all", "cf_set_metrics",
    "cf_hyperplane", "cf_boundary_conditions", "cf_flux_x", "cf_flux_y", "cf_flux_z", "cf_viscous_terms",
    "cf_dual_time_stepping", "cf_load_state", "cf_load_initial", "cf_set_iterate", "cf_read_state",
    "cf_inner_iteration", "cf_physical_step", "cf_advance_levels", "cf_write_output", "cf_synchronize",
    "cf_profile_inner_iteration", "cf_set_sweep_schedule", "cf_flux_x_limiter", "cf_flux_y_limiter",
    "cf_flux_z_limiter", "cf_grid", "cf_filter", "cf_profile_variants", "cf_dplusgs", "cf_set_implicit_solver",
    "cf_set_viscous_schedule", "cf_rk3nodt", "cf_rk3_step",
    "cf_statistic", "cf_driven_force", "cf_save_checkpoint", "cf_load_checkpoint", "cf_set_initial_guess",
    "cf_driven_force_noalloc",
]


class CfParams(C.Structure):
    _fields_ = [
        ("X_out", C.c_int), ("Y_out", C.c_int), ("Z_out", C.c_int),
        ("deltaXI", C.c_double), ("deltaET", C.c_double), ("deltaZT", C.c_double),
        ("high", C.c_double), ("f_drive", C.c_double), ("deltaT", C.c_double),
        ("x_periodic", C.c_int), ("n_sweeps", C.c_int), ("iteration_end_step", C.c_int),
        ("tol", C.c_double), ("precision", C.c_int),
    ]


class CfError(RuntimeError):
    pass


_lib = None


def load_library() -> C.CDLL:
    """Loads libchannel_flow.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise CfError(f"{LIB_PATH} is missing: build it with `python -m channel_flow_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    P = C.POINTER(C.c_double)
    PP = C.POINTER(P)
    vp = C.c_void_p
    lib.cf_create.argtypes = [C.POINTER(CfParams), C.POINTER(vp)]
    lib.cf_destroy.argtypes = [vp]
    lib.cf_destroy.restype = None
    lib.cf_last_error.argtypes = [vp]
    lib.cf_last_error.restype = C.c_char_p
    lib.cf_field_elems.argtypes = [vp]
    lib.cf_field_elems.restype = C.c_longlong
    lib.cf_jacobian_all.argtypes = [vp, C.c_double, P, PP, P, PP, P, PP, P, PP]
    lib.cf_set_metrics.argtypes = [vp, P, PP, P, PP, P, PP, P, PP]
    lib.cf_hyperplane.argtypes = [vp, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]
    lib.cf_boundary_conditions.argtypes = [vp, PP, P]
    for n in ("cf_flux_x", "cf_flux_y", "cf_flux_z"):
        getattr(lib, n).argtypes = [vp, PP, P, P, PP, PP]
    for n in ("cf_flux_x_limiter", "cf_flux_y_limiter", "cf_flux_z_limiter"):
        getattr(lib, n).argtypes = [vp, PP, P, P, PP, PP, P]
    lib.cf_grid.argtypes = [vp, C.c_double, P, P]
    lib.cf_profile_variants.argtypes = [vp, C.c_double, P]
    lib.cf_set_implicit_solver.argtypes = [vp, C.c_int]
    lib.cf_filter.argtypes = [vp, C.c_double, C.c_double, C.c_double, PP, P, P, P, P, P]
    lib.cf_viscous_terms.argtypes = [vp, PP, PP, P, PP, P, PP, P, PP, P, PP]
    lib.cf_dual_time_stepping.argtypes = [vp, C.c_double, P, PP, PP, PP, P, PP, PP, PP, PP, PP, PP, PP, PP]
    lib.cf_dplusgs.argtypes = lib.cf_dual_time_stepping.argtypes
    lib.cf_load_state.argtypes = [vp, PP, PP]
    lib.cf_load_initial.argtypes = [vp, C.c_char_p]
    lib.cf_set_iterate.argtypes = [vp, PP]
    lib.cf_read_state.argtypes = [vp, PP, PP, PP]
    lib.cf_inner_iteration.argtypes = [vp, P]
    lib.cf_physical_step.argtypes = [vp, C.POINTER(C.c_int), P]
    lib.cf_advance_levels.argtypes = [vp]
    lib.cf_profile_inner_iteration.argtypes = [vp, P, P, C.POINTER(C.c_int)]
    lib.cf_set_sweep_schedule.argtypes = [vp, C.c_int]
    lib.cf_set_viscous_schedule.argtypes = [vp, C.c_int]
    lib.cf_set_initial_guess.argtypes = [vp, C.c_int]
    lib.cf_rk3nodt.argtypes = [vp, C.c_int, C.c_double, P, PP, PP, PP, P, PP, PP, PP, PP, PP, PP]
    lib.cf_rk3_step.argtypes = [vp, C.c_int]
    lib.cf_statistic.argtypes = [vp, C.c_int, C.c_int, PP, C.c_char_p]
    lib.cf_driven_force.argtypes = [vp, C.c_double, P, P, P, P]
    lib.cf_driven_force_noalloc.argtypes = [vp, C.c_double, vp, P, P]
    lib.cf_save_checkpoint.argtypes = [vp, C.c_char_p]
    lib.cf_load_checkpoint.argtypes = [vp, C.c_char_p]
    lib.cf_write_output.argtypes = [vp, C.c_int, C.c_char_p]
    lib.cf_synchronize.argtypes = [vp]
    _lib = lib
    return lib


def _p(a):
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise CfError("arrays must be C-contiguous float64")
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pp(arrs):
    if arrs is None:
        return None
    out = (C.POINTER(C.c_double) * len(arrs))(*[_p(a) for a in arrs])
    out._keep = list(arrs)  # the pointer array alone does not keep the arrays alive
    return out


class Solver:
    """One GPU context for one grid (replaces main.cpp's driver state).

    precision: CF_FP64 (default) or CF_FP32_SWEEPS.  The latter runs the
    LU-SGS sweeps in float inside the FP64 defect-correction loop; the
    converged state is the FP64 one, the per-iteration increments are not.
    """

    def __init__(self, cfg: ChannelConfig, precision: int = CF_FP64):
        self.cfg = cfg
        self.lib = load_library()
        p = CfParams(cfg.X_out, cfg.Y_out, cfg.Z_out, cfg.deltaXI, cfg.deltaET, cfg.deltaZT, cfg.high,
                     cfg.f_drive, cfg.deltaT, int(cfg.x_periodic), int(cfg.n_sweeps),
                     int(cfg.iteration_end_step), cfg.tol, int(precision))
        h = C.c_void_p()
        r = self.lib.cf_create(C.byref(p), C.byref(h))
        if r != CF_OK:
            raise CfError(f"cf_create failed: {ERRORS.get(r, r)}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self.lib.cf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, r):
        if r != CF_OK:
            msg = self.lib.cf_last_error(self._h).decode()
            raise CfError(f"{ERRORS.get(r, r)}: {msg}")

    def zeros(self, n=None):
        if n is None:
            return np.zeros(self.cfg.shape)
        return [np.zeros(self.cfg.shape) for _ in range(n)]

    # --- geometry -----------------------------------------------------------
    def Jacobian_all(self, gamma=None, want_arrays=True):
        """Jacobian + Jacobian_u/_v/_w (Jacobian*.cpp); returns the metric dict."""
        gamma = self.cfg.gamma_metric if gamma is None else gamma
        if not want_arrays:
            self._chk(self.lib.cf_jacobian_all(self._h, gamma, None, None, None, None, None, None, None, None))
            return None
        J, Ju, Jv, Jw = self.zeros(4)
        m, mu, mv, mw = self.zeros(9), self.zeros(9), self.zeros(9), self.zeros(9)
        self._chk(self.lib.cf_jacobian_all(self._h, gamma, _p(J), _pp(m), _p(Ju), _pp(mu), _p(Jv), _pp(mv),
                                           _p(Jw), _pp(mw)))
        return dict(J=J, m=m, J_u=Ju, m_u=mu, J_v=Jv, m_v=mv, J_w=Jw, m_w=mw)

    def set_metrics(self, met):
        self._chk(self.lib.cf_set_metrics(self._h, _p(met["J"]), _pp(met["m"]), _p(met["J_u"]), _pp(met["m_u"]),
                                          _p(met["J_v"]), _pp(met["m_v"]), _p(met["J_w"]), _pp(met["m_w"])))

    def Hyperplane(self):
        cap = self.cfg.gend + self.cfg.ny + self.cfg.nz
        out = (C.c_int * cap)()
        n = C.c_int()
        self._chk(self.lib.cf_hyperplane(self._h, out, cap, C.byref(n)))
        return np.array(out[: n.value], dtype=np.int64)

    # --- per-function entry points -------------------------------------------
    def boundary_conditions(self, U_, J):
        """Single-rank x halo + Y_boundary_condition + Z_boundary_condition, in place."""
        self._chk(self.lib.cf_boundary_conditions(self._h, _pp(U_), _p(J)))

    def Flux_X(self, U_, J, J_u, m_u):
        out = self.zeros(5)
        self._chk(self.lib.cf_flux_x(self._h, _pp(U_), _p(J), _p(J_u), _pp(m_u), _pp(out)))
        return out

    def Flux_Y(self, U_, J, J_v, m_v):
        out = self.zeros(5)
        self._chk(self.lib.cf_flux_y(self._h, _pp(U_), _p(J), _p(J_v), _pp(m_v), _pp(out)))
        return out

    def Flux_Z(self, U_, J, J_w, m_w):
        out = self.zeros(5)
        self._chk(self.lib.cf_flux_z(self._h, _pp(U_), _p(J), _p(J_w), _pp(m_w), _pp(out)))
        return out

    VARIANTS = ["filter", "flux_x_limiter", "flux_y_limiter", "flux_z_limiter"]

    def profile_variants(self, Roe_criterion=0.007):
        """HIP-event times (ms) of the variant kernels on the resident iterate."""
        ms = (C.c_double * len(self.VARIANTS))()
        self._chk(self.lib.cf_profile_variants(self._h, float(Roe_criterion), ms))
        return {n: ms[q] for q, n in enumerate(self.VARIANTS)}

    def Grid(self, gamma=2.8):
        """Grid.cpp:50-141: (X_point, Y_point) in the reference layout."""
        X, Y = self.zeros(1)[0], self.zeros(1)[0]
        self._chk(self.lib.cf_grid(self._h, float(gamma), _p(X), _p(Y)))
        return X, Y

    def Filter(self, Roe_criterion, E_high, E_low, U_, Y_point, J, Ep=None):
        """Filter.cpp:12-1720: (EpX, EpY, EpZ); Ep gives the arrays to update."""
        Ep = [self.zeros(1)[0] for _ in range(3)] if Ep is None else Ep
        self._chk(self.lib.cf_filter(self._h, float(Roe_criterion), float(E_high), float(E_low), _pp(U_), _p(Y_point),
                                     _p(Ep[0]), _p(Ep[1]), _p(Ep[2]), _p(J)))
        return Ep

    def Flux_X_limiter(self, U_, J, J_u, m_u, EpX):
        """Flux_X_limiter.cpp:104-547 (sensor EpX on the face layout)."""
        out = self.zeros(5)
        self._chk(self.lib.cf_flux_x_limiter(self._h, _pp(U_), _p(J), _p(J_u), _pp(m_u), _pp(out), _p(EpX)))
        return out

    def Flux_Y_limiter(self, U_, J, J_v, m_v, EpY):
        """Flux_Y_limiter.cpp:106-685."""
        out = self.zeros(5)
        self._chk(self.lib.cf_flux_y_limiter(self._h, _pp(U_), _p(J), _p(J_v), _pp(m_v), _pp(out), _p(EpY)))
        return out

    def Flux_Z_limiter(self, U_, J, J_w, m_w, EpZ):
        """Flux_Z_limiter.cpp:113-710."""
        out = self.zeros(5)
        self._chk(self.lib.cf_flux_z_limiter(self._h, _pp(U_), _p(J), _p(J_w), _pp(m_w), _pp(out), _p(EpZ)))
        return out

    def Viscous_terms(self, U_, met):
        out = self.zeros(4)
        self._chk(self.lib.cf_viscous_terms(self._h, _pp(U_), _pp(out), _p(met["J"]), _pp(met["m"]),
                                            _p(met["J_u"]), _pp(met["m_u"]), _p(met["J_v"]), _pp(met["m_v"]),
                                            _p(met["J_w"]), _pp(met["m_w"])))
        return out

    def Dual_time_stepping(self, U_, U, Uq, met, inFx, inFy, inFz, vF, deltaT=None):
        """Updates U_ in place; returns (er[5], MR, NR, NL)."""
        er = (C.c_double * 5)()
        MR, NR, NL = self.zeros(5), self.zeros(5), self.zeros(5)
        dt = self.cfg.deltaT if deltaT is None else deltaT
        self._chk(self.lib.cf_dual_time_stepping(self._h, dt, er, _pp(U_), _pp(U), _pp(Uq), _p(met["J"]),
                                                 _pp(met["m"]), _pp(inFx), _pp(inFy), _pp(inFz), _pp(vF), _pp(MR),
                                                 _pp(NR), _pp(NL)))
        return np.array(er[:]), MR, NR, NL

    def DPLUSGS(self, U_, U, Uq, met, inFx, inFy, inFz, vF, deltaT=None):
        """DPLUSGS.cpp:189-1361 (Jacobi sweeps); as Dual_time_stepping."""
        er = (C.c_double * 5)()
        MR, NR, NL = self.zeros(5), self.zeros(5), self.zeros(5)
        dt = self.cfg.deltaT if deltaT is None else deltaT
        self._chk(self.lib.cf_dplusgs(self._h, dt, er, _pp(U_), _pp(U), _pp(Uq), _p(met["J"]), _pp(met["m"]),
                                      _pp(inFx), _pp(inFy), _pp(inFz), _pp(vF), _pp(MR), _pp(NR), _pp(NL)))
        return np.array(er[:]), MR, NR, NL

    def RK3NODT(self, RK, U_, U, Uq, met, inFx, inFy, inFz, vF, deltaT=None):
        """RK3NODT.cpp:115-183, stage RK; updates U_ in place, returns MR = dt Rk."""
        er = (C.c_double * 5)()
        MR = self.zeros(5)
        dt = self.cfg.deltaT if deltaT is None else deltaT
        self._chk(self.lib.cf_rk3nodt(self._h, int(RK), dt, er, _pp(U_), _pp(U), _pp(Uq), _p(met["J"]),
                                      _pp(met["m"]), _pp(inFx), _pp(inFy), _pp(inFz), _pp(vF), _pp(MR)))
        return MR

    def Statistic(self, step, statistic_step, U_=None, directory="."):
        """Statistic.cpp:1043-1154 mean profiles; writes st%05d.dat when
        step % statistic_step == 0.  U_: host arrays, or None for the
        device iterate."""
        # the converted arrays must outlive the call: the pointer array does
        # not keep them alive
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in U_] if U_ is not None else None
        ptr = _pp(arrs)
        self._chk(self.lib.cf_statistic(self._h, int(step), int(statistic_step), ptr, str(directory).encode()))
        del arrs

    def Driven_force(self, Q0=20.464969232814912, U2_=None, J=None):
        """Driven_force.cpp:38-73: (Qall, ER_Q) of U2_ (host array) or of the
        device iterate; Q0 defaults to prm.h's."""
        Q, E = C.c_double(), C.c_double()
        u = _p(np.ascontiguousarray(U2_, dtype=np.float64)) if U2_ is not None else None
        jj = _p(np.ascontiguousarray(J, dtype=np.float64)) if J is not None else None
        self._chk(self.lib.cf_driven_force(self._h, float(Q0), u, jj, C.byref(Q), C.byref(E)))
        return Q.value, E.value

    def Driven_force_noalloc(self, Q0=20.464969232814912, U2_dev=None):
        """cf_driven_force_noalloc: Driven_force with no device allocation.
        U2_dev: device address (int) of a reference-layout rho u field, or
        None for the resident iterate."""
        Q, E = C.c_double(), C.c_double()
        self._chk(self.lib.cf_driven_force_noalloc(self._h, float(Q0), U2_dev, C.byref(Q), C.byref(E)))
        return Q.value, E.value

    def save_checkpoint(self, path):
        """Exact restart file: state as stored, E_k, statistic sums."""
        self._chk(self.lib.cf_save_checkpoint(self._h, str(path).encode()))

    def load_checkpoint(self, path):
        self._chk(self.lib.cf_load_checkpoint(self._h, str(path).encode()))

    def rk3_step(self, viscous: bool = True):
        """Three RK3NODT stages on the device state (main.cpp:152-377)."""
        self._chk(self.lib.cf_rk3_step(self._h, int(bool(viscous))))

    def set_implicit_solver(self, solver: int = 0):
        """0: LU-SGS (Dual_time_stepping, default); 1: DPLUSGS."""
        self._chk(self.lib.cf_set_implicit_solver(self._h, int(solver)))

    # --- device-resident path --------------------------------------------------
    def load_state(self, Qn, Qnm1):
        Qn = [np.ascontiguousarray(q, dtype=np.float64) for q in Qn]
        Qnm1 = [np.ascontiguousarray(q, dtype=np.float64) for q in Qnm1]
        self._chk(self.lib.cf_load_state(self._h, _pp(Qn), _pp(Qnm1)))

    def load_initial(self, path):
        """Initial_condition.cpp:64-116: levels n and n-1 from an initial.bin /
        LES%05d.bin file (Q = U J, ten raw FP64 blocks)."""
        self._chk(self.lib.cf_load_initial(self._h, str(path).encode()))

    Initial_condition = load_initial

    def set_iterate(self, Q):
        Q = [np.ascontiguousarray(q, dtype=np.float64) for q in Q]
        self._chk(self.lib.cf_set_iterate(self._h, _pp(Q)))

    def read_state(self):
        c = self.cfg
        shp = (c.X_out, c.Y_out, c.Z_out)
        Qn = [np.zeros(shp) for _ in range(5)]
        Qm = [np.zeros(shp) for _ in range(5)]
        Qi = [np.zeros(shp) for _ in range(5)]
        self._chk(self.lib.cf_read_state(self._h, _pp(Qn), _pp(Qm), _pp(Qi)))
        return np.stack(Qn), np.stack(Qm), np.stack(Qi)

    def inner_iteration(self):
        er = (C.c_double * 5)()
        self._chk(self.lib.cf_inner_iteration(self._h, er))
        return np.array(er[:])

    STAGES = ["ghosts", "flux_x", "flux_y", "flux_z", "viscous", "rhs", "sweep_prep", "sweep_sync", "lusgs_fwd",
              "lusgs_bwd", "update_norms"]

    def profile_inner_iteration(self):
        """One inner iteration with per-stage HIP-event times: (er, {stage: (ms, launches)})."""
        er = (C.c_double * 5)()
        ms = (C.c_double * len(self.STAGES))()
        nl = (C.c_int * len(self.STAGES))()
        self._chk(self.lib.cf_profile_inner_iteration(self._h, er, ms, nl))
        return np.array(er[:]), {n: (ms[q], nl[q]) for q, n in enumerate(self.STAGES)}

    def physical_step(self):
        er = (C.c_double * 5)()
        it = C.c_int()
        self._chk(self.lib.cf_physical_step(self._h, C.byref(it), er))
        return it.value, np.array(er[:])

    def advance_levels(self):
        """U^{n-1} <- U^n <- U_ (main.cpp:444-464)."""
        self._chk(self.lib.cf_advance_levels(self._h))

    def Output(self, step, directory):
        self._chk(self.lib.cf_write_output(self._h, int(step), str(directory).encode()))

    def set_sweep_schedule(self, schedule: int = 0):
        """0: pipelined tiled sweep (default); 1: one launch per hyperplane;
        2: single-wave tiles.  All three give the same bits."""
        self._chk(self.lib.cf_set_sweep_schedule(self._h, int(schedule)))

    def set_viscous_schedule(self, schedule: int = 0):
        """0: one fused viscous kernel (default); 1: three kernels through
        staging fields.  Both give the same bits."""
        self._chk(self.lib.cf_set_viscous_schedule(self._h, int(schedule)))

    def set_initial_guess(self, order: int = 0):
        """Initial iterate of each physical_step(): 0 U^n (default, the
        reference's); 1 2U^n - U^{n-1}; 2 3U^n - 3U^{n-1} + U^{n-2}."""
        self._chk(self.lib.cf_set_initial_guess(self._h, int(order)))

    def synchronize(self):
        self._chk(self.lib.cf_synchronize(self._h))
