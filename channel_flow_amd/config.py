"""Run parameters of the channel solver and the seeded initial condition.

The reference keeps its parameters as compile-time macros and globals:
grid sizes and spacings in Resolution.h:1-27, gas constants and the driving
force in prm.h:7-20, transport coefficients in Viscous_terms.h:3-9 and the run
constants as locals of main.cpp:47-57.  `ChannelConfig` holds the same values
as runtime fields (plus the stretching factor that Jacobian.cpp:50 and
Grid.cpp:41 hard-code), and `parse_reference_headers` reads them back from a
Resolution.h / prm.h pair so a reference user's configuration carries over.

The initial condition is new (the reference only reads `initial.bin`,
Initial_condition.cpp:64-79): laminar Poiseuille flow with seeded uniform
noise, written in the same raw layout.
"""
from __future__ import annotations

import dataclasses
import math
import re
from pathlib import Path

import numpy as np

# physical constants fixed by the reference
GAMMA_GAS = 1.4            # prm.h:9 (K)
R_GAS = 287.0              # prm.h:10 (R)
CV = 717.5                 # Viscous_terms.h:3
MU_L = 0.0000185           # Viscous_terms.h:4
PR_L = 0.72                # Viscous_terms.h:5
T_WALL = 298.0592          # Y_boundary_condition.cpp:87 (2*298.0592 - T)
PRECOND_FLOOR = 0.0001     # prm.h:7 (the macro `e`)
F_DRIVE_REF = 599.29876710  # prm.h:13
HIGH_REF = 0.005           # prm.h:18 (channel height 2h)


@dataclasses.dataclass
class ChannelConfig:
    """One run's parameters (reference names in the comments)."""

    X_out: int = 64
    Y_out: int = 64
    Z_out: int = 64
    high: float = HIGH_REF            # prm.h:18
    Lx: float | None = None           # default 2*pi*high (Resolution.h:13)
    Lz: float | None = None           # default pi*high (Resolution.h:15)
    gamma_metric: float = 1.5         # Jacobian*.cpp hard-code
    gamma_grid: float = 2.8           # Grid.cpp:41 hard-code (only Filter reads it)
    deltaT: float = 5e-8              # main.cpp:53
    f_drive: float = F_DRIVE_REF      # prm.h:13
    iteration_end_step: int = 200     # main.cpp:48
    tol: float = 0.001                # main.cpp:406
    x_periodic: bool = False          # single rank: zero-gradient x ghosts (main.cpp:207-238)
    n_sweeps: int = 2                 # Dual_time_stepping.cpp:514
    Roe_criterion: float = 0.007      # main.cpp:55
    E_high: float = 1.0               # main.cpp:56
    E_low: float = 0.0                # main.cpp:57

    def __post_init__(self):
        if self.Lx is None:
            self.Lx = 2 * math.pi * self.high
        if self.Lz is None:
            self.Lz = math.pi * self.high
        for n in ("X_out", "Y_out", "Z_out"):
            if getattr(self, n) < 8:
                raise ValueError(f"{n} must be >= 8 (the 5th-order stencils need it)")

    # Resolution.h derived sizes (single rank: X_np = gcount + 6, gend = gcount + 2)
    @property
    def X_np(self) -> int:
        return self.X_out + 6

    @property
    def Y_m(self) -> int:
        return self.Y_out + 4

    @property
    def Z_m(self) -> int:
        return self.Z_out + 4

    @property
    def gend(self) -> int:
        return self.X_out + 2

    @property
    def ny(self) -> int:
        return self.Y_out + 1

    @property
    def nz(self) -> int:
        return self.Z_out + 1

    @property
    def deltaXI(self) -> float:
        return self.Lx / self.X_out

    @property
    def deltaET(self) -> float:
        return 1.0 / self.Y_out

    @property
    def deltaZT(self) -> float:
        return self.Lz / self.Z_out

    @property
    def shape(self) -> tuple[int, int, int]:
        """Reference array shape [X_np][Y_m][Z_m] (Array.h)."""
        return (self.X_np, self.Y_m, self.Z_m)

    @property
    def cells(self) -> int:
        return self.X_out * self.Y_out * self.Z_out


def parse_reference_headers(resolution_h: str, prm_h: str | None = None) -> dict:
    """Read the grid (Resolution.h) and driving force/height (prm.h) macros."""
    out: dict = {}
    txt = Path(resolution_h).read_text()
    for name in ("X_out", "Y_out", "Z_out"):
        m = re.search(rf"#define\s+{name}\s+(\d+)", txt)
        if m:
            out[name] = int(m.group(1))
    for name in ("deltaXI", "deltaZT"):
        m = re.search(rf"#define\s+{name}\s+([0-9.eE+-]+)", txt)
        if m:
            out[name] = float(m.group(1))
    if "deltaXI" in out and "X_out" in out:
        out["Lx"] = out.pop("deltaXI") * out["X_out"]
    if "deltaZT" in out and "Z_out" in out:
        out["Lz"] = out.pop("deltaZT") * out["Z_out"]
    if prm_h:
        ptxt = Path(prm_h).read_text()
        m = re.search(r"double\s+f\s*=\s*([0-9.eE+-]+)", ptxt)
        if m:
            out["f_drive"] = float(m.group(1))
        m = re.search(r"double\s+high\s*=\s*([0-9.eE+-]+)", ptxt)
        if m:
            out["high"] = float(m.group(1))
    return out


# ---------------------------------------------------------------------------
# seeded initial condition
# ---------------------------------------------------------------------------

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) on uint64 counters."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """n uniforms in [-0.5, 0.5) from counter (seed, stream, index)."""
    base = (np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream) * np.uint64(0x1000193)) & _M64
    with np.errstate(over="ignore"):
        ctr = (np.arange(n, dtype=np.uint64) + base * np.uint64(0x9E3779B97F4A7C15)) & _M64
    bits = splitmix64(ctr) >> np.uint64(11)
    return bits.astype(np.float64) * (1.0 / 9007199254740992.0) - 0.5


def cell_y(cfg: ChannelConfig) -> np.ndarray:
    """Wall-normal position of interior cell centres j=2..ny (Jacobian.cpp:74-78 mapping)."""
    g = cfg.gamma_metric
    eta = (np.arange(2, cfg.ny + 1) - 1.5) * cfg.deltaET
    return 0.5 * cfg.high * (1.0 - np.tanh(g * (1.0 - 2.0 * eta)) / math.tanh(g))


@dataclasses.dataclass
class FlowParams:
    Re_bulk: float = 5600.0
    Mach: float = 0.2
    noise: float = 0.05

    def bulk_state(self, high: float):
        c_w = math.sqrt(GAMMA_GAS * R_GAS * T_WALL)
        Ub = self.Mach * c_w
        rho = MU_L * self.Re_bulk / (Ub * high)
        return Ub, rho

    def driving_force(self, high: float) -> float:
        """Laminar balance f = 3 mu U_b / h^2 (h = high/2)."""
        Ub, _ = self.bulk_state(high)
        h = 0.5 * high
        return 3.0 * MU_L * Ub / (h * h)


def poiseuille_state(cfg: ChannelConfig, flow: FlowParams, seed: int, stream: int = 0) -> np.ndarray:
    """Q = (rho, rho u, rho v, rho w, rho E) on the interior, shape (5, X, Y, Z)."""
    Ub, rho0 = flow.bulk_state(cfg.high)
    h = 0.5 * cfg.high
    y = cell_y(cfg)
    u_lam = 1.5 * Ub * (1.0 - (y / h - 1.0) ** 2)
    X, Y, Z = cfg.X_out, cfg.Y_out, cfg.Z_out
    n = X * Y * Z
    a = flow.noise * 2.0  # uniform(-0.5,0.5)*2*noise = U(-noise, noise)
    u = u_lam[None, :, None] * (1.0 + a * uniform(seed, 3 * stream + 0, n).reshape(X, Y, Z))
    v = a * Ub * uniform(seed, 3 * stream + 1, n).reshape(X, Y, Z)
    w = a * Ub * uniform(seed, 3 * stream + 2, n).reshape(X, Y, Z)
    rho = np.full((X, Y, Z), rho0)
    P = rho0 * R_GAS * T_WALL
    E = P / (GAMMA_GAS - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, E])


def write_initial_bin(path, Qn: np.ndarray, Qnm1: np.ndarray) -> None:
    """Raw little-endian FP64, 10 blocks [x][y][z]: Q^n then Q^{n-1} (Initial_condition.cpp:64-79)."""
    with open(path, "wb") as f:
        for Q in (Qn, Qnm1):
            np.ascontiguousarray(Q, dtype="<f8").tofile(f)


def read_les_bin(path, cfg: ChannelConfig) -> tuple[np.ndarray, np.ndarray]:
    """Read an LES%05d.bin / initial.bin file (Output.cpp:160-190 layout)."""
    a = np.fromfile(path, dtype="<f8")
    n = cfg.cells
    if a.size != 10 * n:
        raise ValueError(f"{path}: expected {10 * n} doubles, found {a.size}")
    a = a.reshape(10, cfg.X_out, cfg.Y_out, cfg.Z_out)
    return a[:5], a[5:]
