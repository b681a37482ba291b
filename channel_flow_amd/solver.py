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
import weakref
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
    "cf_pred_create", "cf_pred_destroy", "cf_pred_last_error", "cf_pred_init_params", "cf_pred_set_params",
    "cf_pred_get_params", "cf_pred_set_norm", "cf_pred_get_norm", "cf_pred_capture", "cf_pred_dataset_size",
    "cf_pred_fit_norm", "cf_pred_load_dataset", "cf_pred_train", "cf_pred_eval", "cf_pred_forward", "cf_pred_gradient", "cf_pred_profile",
    "cf_keep_levels", "cf_set_predictor", "cf_predictor_stats", "cf_set_sweep_reuse", "cf_set_graph",
    "cf_graph_builds", "cf_set_concurrent_stencils", "cf_set_extrapolation_order", "cf_set_spin_limit",
]
CF_PRED_NPARAM = 112005
CF_PRED_NNORM = 22
CF_PRED_NSTATS = 36
PRED_PROFILE_STAGES = ["infer", "train_forward", "train_head", "train_bwd_data", "train_bwd_weight", "train_reduce"]


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
    lib.cf_set_extrapolation_order.argtypes = [vp, C.c_int]
    lib.cf_rk3nodt.argtypes = [vp, C.c_int, C.c_double, P, PP, PP, PP, P, PP, PP, PP, PP, PP, PP]
    lib.cf_rk3_step.argtypes = [vp, C.c_int]
    lib.cf_statistic.argtypes = [vp, C.c_int, C.c_int, PP, C.c_char_p]
    lib.cf_driven_force.argtypes = [vp, C.c_double, P, P, P, P]
    lib.cf_driven_force_noalloc.argtypes = [vp, C.c_double, vp, P, P]
    lib.cf_save_checkpoint.argtypes = [vp, C.c_char_p]
    lib.cf_load_checkpoint.argtypes = [vp, C.c_char_p]
    lib.cf_write_output.argtypes = [vp, C.c_int, C.c_char_p]
    lib.cf_synchronize.argtypes = [vp]
    F = C.POINTER(C.c_float)
    u64 = C.c_ulonglong
    lib.cf_pred_create.argtypes = [C.c_int, C.c_int, C.POINTER(vp)]
    lib.cf_pred_destroy.argtypes = [vp]
    lib.cf_pred_destroy.restype = None
    lib.cf_pred_last_error.argtypes = [vp]
    lib.cf_pred_last_error.restype = C.c_char_p
    lib.cf_pred_init_params.argtypes = [vp, u64]
    lib.cf_pred_set_params.argtypes = [vp, F, C.c_int]
    lib.cf_pred_get_params.argtypes = [vp, F, C.c_int]
    lib.cf_pred_set_norm.argtypes = [vp, F]
    lib.cf_pred_get_norm.argtypes = [vp, F]
    lib.cf_pred_capture.argtypes = [vp, vp, C.c_int, u64]
    lib.cf_pred_dataset_size.argtypes = [vp]
    lib.cf_pred_fit_norm.argtypes = [vp, C.c_int, P]
    lib.cf_pred_load_dataset.argtypes = [vp, F, C.c_int]
    lib.cf_pred_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, u64, F]
    lib.cf_pred_eval.argtypes = [vp, C.c_int, C.c_int, P, P]
    lib.cf_pred_forward.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, F, F]
    lib.cf_pred_gradient.argtypes = [vp, C.c_int, F, F, F, P]
    lib.cf_pred_profile.argtypes = [vp, C.c_int, C.c_int, C.c_int, P]
    lib.cf_keep_levels.argtypes = [vp, C.c_int]
    lib.cf_set_sweep_reuse.argtypes = [vp, C.c_int]
    lib.cf_set_spin_limit.argtypes = [vp, C.c_longlong]
    lib.cf_set_graph.argtypes = [vp, C.c_int]
    lib.cf_graph_builds.argtypes = [vp, C.POINTER(C.c_longlong)]
    lib.cf_set_concurrent_stencils.argtypes = [vp, C.c_int]
    lib.cf_set_predictor.argtypes = [vp, vp]
    lib.cf_predictor_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
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

    def set_extrapolation_order(self, order: int = 2):
        """Order of the extrapolated start (initial guess 2 and the predictor's
        fallback): 2 quadratic (default), 3 cubic once four levels are known."""
        self._chk(self.lib.cf_set_extrapolation_order(self._h, int(order)))

    def synchronize(self):
        self._chk(self.lib.cf_synchronize(self._h))

    def set_sweep_reuse(self, enabled: bool = True):
        """One LU-SGS sweep pair for any n_sweeps (bit-identical in one domain)."""
        self._chk(self.lib.cf_set_sweep_reuse(self._h, int(bool(enabled))))

    def set_spin_limit(self, polls: int):
        """Polls a tiled sweep waits for one upstream tile's value before it
        gives up (CF_ERR_TIMEOUT; the iterate must then be reloaded)."""
        self._chk(self.lib.cf_set_spin_limit(self._h, int(polls)))

    def set_graph(self, enabled: bool = True):
        """Replay each inner iteration's kernels as a captured hipGraph."""
        self._chk(self.lib.cf_set_graph(self._h, int(bool(enabled))))

    def set_concurrent_stencils(self, enabled: bool = True):
        """Fluxes and viscous terms on four streams (same bits)."""
        self._chk(self.lib.cf_set_concurrent_stencils(self._h, int(bool(enabled))))

    def graph_builds(self) -> int:
        n = C.c_longlong()
        self._chk(self.lib.cf_graph_builds(self._h, C.byref(n)))
        return n.value

    def keep_levels(self, levels: int):
        """Time levels kept: 2 (U^n, U^{n-1}), 3 (+U^{n-2}) or 4 (+U^{n-3})."""
        self._chk(self.lib.cf_keep_levels(self._h, int(levels)))

    def set_predictor(self, pred):
        """Attaches a Predictor (None detaches); set_initial_guess(3) uses it."""
        self._chk(self.lib.cf_set_predictor(self._h, pred._h if pred is not None else None))
        old = getattr(self, "_pred", None)
        if old is not None and old is not pred:
            old._attached.discard(self)
        if pred is not None:
            pred._attached.add(self)
        self._pred = pred  # keeps the predictor alive while attached

    def predictor_stats(self):
        """(physical steps started by the network, fallbacks to the extrapolation)."""
        a, b = C.c_longlong(), C.c_longlong()
        self._chk(self.lib.cf_predictor_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value


def _f(a):
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous):
        raise CfError("arrays must be C-contiguous float32")
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Predictor:
    """The learned initial iterate (cf_pred_*): parameters, crop dataset,
    normalisation and training, all on the GPU."""

    CROP = 16  # crop edge in cells (inner 14)

    def __init__(self, dataset_crops: int = 1024, batch_crops: int = 32):
        self.lib = load_library()
        h = C.c_void_p()
        r = self.lib.cf_pred_create(int(dataset_crops), int(batch_crops), C.byref(h))
        if r != CF_OK:
            raise CfError(f"cf_pred_create failed: {ERRORS.get(r, r)}")
        self._h = h
        self.batch_crops = batch_crops
        self._attached = weakref.WeakSet()  # Solvers holding this predictor

    def close(self):
        """Destroys the predictor, after detaching it from every open Solver
        it is attached to (a context keeps the pointer)."""
        if getattr(self, "_h", None):
            for s in list(getattr(self, "_attached", ())):
                if getattr(s, "_h", None):
                    s.lib.cf_set_predictor(s._h, None)
                s._pred = None
            self.lib.cf_pred_destroy(self._h)
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
            raise CfError(f"{ERRORS.get(r, r)}: {self.lib.cf_pred_last_error(self._h).decode()}")

    def init_params(self, seed: int = 1):
        self._chk(self.lib.cf_pred_init_params(self._h, int(seed)))

    def set_params(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32)
        self._chk(self.lib.cf_pred_set_params(self._h, _f(a), a.size))

    def get_params(self):
        a = np.zeros(CF_PRED_NPARAM, np.float32)
        self._chk(self.lib.cf_pred_get_params(self._h, _f(a), a.size))
        return a

    def set_norm(self, norm):
        a = np.ascontiguousarray(norm, dtype=np.float32)
        assert a.size == CF_PRED_NNORM
        self._chk(self.lib.cf_pred_set_norm(self._h, _f(a)))

    def get_norm(self):
        a = np.zeros(CF_PRED_NNORM, np.float32)
        self._chk(self.lib.cf_pred_get_norm(self._h, _f(a)))
        return a

    def capture(self, solver, ncrops: int, seed: int):
        """Crops of the solver's last physical step (it must keep 4 levels)."""
        r = self.lib.cf_pred_capture(self._h, solver._h, int(ncrops), int(seed))
        if r != CF_OK:
            raise CfError(f"{ERRORS.get(r, r)}: {self.lib.cf_pred_last_error(self._h).decode()} / "
                          f"{self.lib.cf_last_error(solver._h).decode()}")

    def _load_dataset(self, crops):
        """Appends crops [n][16][16][16][24] (the layout capture writes)."""
        a = np.ascontiguousarray(crops, dtype=np.float32)
        self._chk(self.lib.cf_pred_load_dataset(self._h, _f(a), a.shape[0]))

    def dataset_size(self) -> int:
        return int(self.lib.cf_pred_dataset_size(self._h))

    def fit_norm(self, n=None):
        n = self.dataset_size() if n is None else n
        st = np.zeros(CF_PRED_NSTATS)
        self._chk(self.lib.cf_pred_fit_norm(self._h, int(n), _p(st)))
        return st

    def train(self, n_train: int, steps: int, lr: float = 1e-3, seed: int = 1):
        losses = np.zeros(max(steps, 1), np.float32)
        self._chk(self.lib.cf_pred_train(self._h, int(n_train), int(steps), float(lr), int(seed), _f(losses)))
        return losses[:steps]

    def eval(self, first: int, n: int):
        """(mean squared error of the network, of the base guess) on crops [first, first + n)."""
        a, b = C.c_double(), C.c_double()
        self._chk(self.lib.cf_pred_eval(self._h, int(first), int(n), C.byref(a), C.byref(b)))
        return a.value, b.value

    def forward(self, x, margin: int = 1):
        """y on the inner region of volumes x [n][VI][VJ][VK][16] (float32)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        n, VI, VJ, VK, _ = x.shape
        y = np.zeros((n, VI - 2 * margin, VJ - 2 * margin, VK - 2 * margin, 5), np.float32)
        self._chk(self.lib.cf_pred_forward(self._h, n, VI, VJ, VK, int(margin), _f(x), _f(y)))
        return y

    def gradient(self, x, target):
        """(gradient, loss) of crops x [n][16][16][16][16], targets [n][14][14][14][5]."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(CF_PRED_NPARAM, np.float32)
        loss = C.c_double()
        self._chk(self.lib.cf_pred_gradient(self._h, x.shape[0], _f(x), _f(t), _f(g), C.byref(loss)))
        return g, loss.value

    def profile(self, X: int, Y: int, Z: int):
        ms = np.zeros(len(PRED_PROFILE_STAGES))
        self._chk(self.lib.cf_pred_profile(self._h, int(X), int(Y), int(Z), _p(ms)))
        return dict(zip(PRED_PROFILE_STAGES, ms.tolist()))
