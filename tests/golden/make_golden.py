"""Writes the golden fixtures of tests/golden/ from the reference solver.

Runs only in a container that has the reference sources (/root/reference)
and MPICH: oracle/ref_build.sh compiles the reference where it lies for each
grid, oracle/_ref/<grid>/ref_driver runs it, and this script keeps the arrays
the tests compare against.  Inputs are this project's seeded initial
condition (channel_flow_amd.config); outputs are whatever the reference
computes from them.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state, write_initial_bin  # noqa: E402

OUT = Path(__file__).resolve().parent
DT_UNIT = 5e-8          # main.cpp:53
UNIT_GRIDS = [(12, 10, 16), (8, 12, 9)]
TRAJ = dict(grid=(16, 12, 16), seed=7, deltaT=5e-8, steps=3, iteration_end_step=6)
TRAJ64 = dict(grid=(64, 64, 64), seed=1, deltaT=5e-8, steps=10, iteration_end_step=20)


def ref_driver(grid) -> Path:
    X, Y, Z = grid
    exe = ROOT / "oracle" / "_ref" / f"g{X}x{Y}x{Z}" / "ref_driver"
    subprocess.run([str(ROOT / "oracle" / "ref_build.sh"), str(X), str(Y), str(Z)], check=True,
                   stdout=subprocess.DEVNULL)
    return exe


def write_inputs(d: Path, cfg: ChannelConfig, seed: int, iterate: bool):
    fl = FlowParams()
    Qn = poiseuille_state(cfg, fl, seed=seed, stream=0)
    Qm = poiseuille_state(cfg, fl, seed=seed, stream=1) if iterate else Qn
    write_initial_bin(d / "initial.bin", Qn, Qm)
    if iterate:
        np.ascontiguousarray(poiseuille_state(cfg, fl, seed=seed, stream=2)).tofile(d / "iterate.bin")


def perturbed_state(cfg: ChannelConfig, seed: int, stream: int) -> np.ndarray:
    """Poiseuille state with every conserved variable perturbed: density and
    total energy by +-2 %, momenta with the density (velocity noise as in
    poiseuille_state).  No two neighbouring cells then hold equal values, so
    the limiters' slope ratios are finite everywhere."""
    Q = np.array(poiseuille_state(cfg, FlowParams(), seed=seed, stream=stream))
    rng = np.random.default_rng(1000 * seed + stream)
    f_rho = 1.0 + 0.02 * rng.uniform(-1.0, 1.0, Q[0].shape)
    f_e = 1.0 + 0.02 * rng.uniform(-1.0, 1.0, Q[0].shape)
    for v in range(4):
        Q[v] *= f_rho
    Q[4] *= f_e
    return Q


def load(d: Path, name: str, cfg: ChannelConfig) -> np.ndarray:
    return np.fromfile(d / f"{name}.bin").reshape(cfg.shape)


def unit_fixture(grid, seed):
    cfg = ChannelConfig(X_out=grid[0], Y_out=grid[1], Z_out=grid[2])
    exe = ref_driver(grid)
    with tempfile.TemporaryDirectory() as td:
        d = Path(td)
        write_inputs(d, cfg, seed, iterate=True)
        subprocess.run([str(exe), "unit", str(d), repr(DT_UNIT)], check=True, stdout=subprocess.DEVNULL)
        arrs = {}
        names = ["J", "J_u", "J_v", "J_w", "Y_point", "EpX", "EpY", "EpZ", "EpX_s", "EpY_s", "EpZ_s"]
        for m in ["xidx", "xidy", "xidz", "etdx", "etdy", "etdz", "ztdx", "ztdy", "ztdz"]:
            names += [m, m + "_u", m + "_v", m + "_w"]
        for base, n in [("U_pre", 5), ("U", 5), ("Uq", 5), ("U_bc", 5), ("inFx", 5), ("inFy", 5), ("inFz", 5),
                        ("vF", 4), ("U_dt", 5), ("MR_dt", 5), ("NL_dt", 5), ("NR_dt", 5), ("inFxL", 5),
                        ("inFyL", 5), ("inFzL", 5), ("inFxS", 5), ("inFyS", 5), ("inFzS", 5), ("U_rk1_", 5),
                        ("U_rk2_", 5), ("U_rk3_", 5)]:
            names += [f"{base}{v + 1}" for v in range(n)]
        for n in names:
            arrs[n] = load(d, n, cfg)
        er = [float(x) for x in (d / "scalars.txt").read_text().split()[1:6]]
        hp = np.loadtxt(d / "hyperplane.txt").astype(np.int64)
        arrs["hyperplane"] = hp
        arrs["er_dt"] = np.array(er)
    meta = dict(grid=list(grid), seed=seed, deltaT=DT_UNIT, gamma_metric=1.5, f_drive=599.29876710,
                reference_calls=["halo (main.cpp:156-240)", "Y_boundary_condition", "Z_boundary_condition",
                                 "Flux_X", "Flux_Y", "Flux_Z", "Viscous_terms", "Dual_time_stepping",
                                 "Filter", "Flux_X/Y/Z_limiter", "RK3NODT"])
    name = f"unit_g{grid[0]}x{grid[1]}x{grid[2]}"
    np.savez_compressed(OUT / f"{name}.npz", **arrs)
    return name, meta


VAR_GRIDS = [(12, 10, 16), (8, 12, 9)]


def variant_fixture(grid, seed):
    """Limiter fluxes, Filter sensor and RK3NODT stages on a perturbed state
    (the unit fixtures' state has a uniform density along x and z, where the
    limiters' slope ratios are 0/0)."""
    cfg = ChannelConfig(X_out=grid[0], Y_out=grid[1], Z_out=grid[2])
    exe = ref_driver(grid)
    with tempfile.TemporaryDirectory() as td:
        d = Path(td)
        Qn, Qm = perturbed_state(cfg, seed, 0), perturbed_state(cfg, seed, 1)
        write_initial_bin(d / "initial.bin", Qn, Qm)
        np.ascontiguousarray(perturbed_state(cfg, seed, 2)).tofile(d / "iterate.bin")
        subprocess.run([str(exe), "unit", str(d), repr(DT_UNIT)], check=True, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
        names = ["J", "J_u", "J_v", "J_w", "Y_point", "EpX", "EpY", "EpZ", "EpX_s", "EpY_s", "EpZ_s"]
        for base, n in [("U", 5), ("Uq", 5), ("U_bc", 5), ("inFx", 5), ("inFy", 5), ("inFz", 5), ("vF", 4),
                        ("inFxL", 5), ("inFyL", 5), ("inFzL", 5), ("inFxS", 5), ("inFyS", 5), ("inFzS", 5),
                        ("U_rk1_", 5), ("U_rk2_", 5), ("U_rk3_", 5)]:
            names += [f"{base}{v + 1}" for v in range(n)]
        arrs = {n: load(d, n, cfg) for n in names}
    meta = dict(grid=list(grid), seed=seed, deltaT=DT_UNIT, gamma_metric=1.5, f_drive=599.29876710,
                state="perturbed_state (density, energy +-2 %)", omp_threads=1,
                reference_calls=["halo", "Y/Z_boundary_condition", "Flux_X/Y/Z", "Viscous_terms",
                                 "Filter(0.007, 1, 0)", "Flux_X/Y/Z_limiter (Filter sensor, then a smooth "
                                 "sensor field)", "RK3NODT stages 1-3"])
    name = f"var_g{grid[0]}x{grid[1]}x{grid[2]}"
    np.savez_compressed(OUT / f"{name}.npz", **arrs)
    return name, meta


def filter_pin_fixture(grid):
    """The reference Filter's sensor patterns for thresholds placed at sensor
    ratios: for sample faces the threshold is the ratio r that the oracle
    computes there and the next double above it.  The reference marks the
    face E_high at r and E_low at next(r) exactly when its own ratio is r, so
    each pair pins one ratio bit for bit; the full patterns (bit-packed) pin
    every face against every threshold."""
    from oracle import oracle as O
    cfg = ChannelConfig(X_out=grid[0], Y_out=grid[1], Z_out=grid[2])
    z = np.load(OUT / f"var_g{grid[0]}x{grid[1]}x{grid[2]}.npz")
    U = [np.ascontiguousarray(z[f"U_bc{v + 1}"]) for v in range(5)]
    R = O.filter_ratios(cfg, U, np.ascontiguousarray(z["Y_point"]), np.ascontiguousarray(z["J"]))
    g, ny, nz = cfg.gend, cfg.ny, cfg.nz
    regions = [(slice(2, g + 1), slice(1, ny), slice(1, nz)), (slice(2, g), slice(1, ny + 1), slice(1, nz)),
               (slice(2, g), slice(1, ny), slice(1, nz + 1))]
    crits, faces = [], []
    for a in range(3):
        vals = R[a][regions[a]]
        flat = np.sort(np.unique(vals[np.isfinite(vals)]))
        for q in (0.02, 0.3, 0.7, 0.98):
            r = flat[int(q * (len(flat) - 1))]
            idx = np.argwhere(R[a] == r)
            idx = [t for t in idx.tolist() if all(regions[a][d].start <= t[d] < regions[a][d].stop for d in range(3))][0]
            faces.append([a] + idx)
            crits += [float(r), float(np.nextafter(r, np.inf))]
    exe = ref_driver(grid)
    with tempfile.TemporaryDirectory() as td:
        d = Path(td)
        write_initial_bin(d / "initial.bin", perturbed_state(cfg, 21 + VAR_GRIDS.index(tuple(grid)), 0),
                          perturbed_state(cfg, 21 + VAR_GRIDS.index(tuple(grid)), 1))
        np.ascontiguousarray(perturbed_state(cfg, 21 + VAR_GRIDS.index(tuple(grid)), 2)).tofile(d / "iterate.bin")
        (d / "crits.txt").write_text("".join(f"{c.hex()}\n" for c in crits))
        subprocess.run([str(exe), "filter", str(d), repr(DT_UNIT), str(d / "crits.txt")], check=True,
                       stdout=subprocess.DEVNULL, env=dict(os.environ, OMP_NUM_THREADS="1"))
        pats = np.stack([np.stack([load(d, f"Ep{a}_c{r}", cfg) for a in "XYZ"]) for r in range(len(crits))])
    assert np.array_equal(np.unique(pats), [0.0, 1.0]) or set(np.unique(pats)) <= {0.0, 1.0}
    arrs = dict(crits=np.array(crits), faces=np.array(faces, dtype=np.int64),
                patterns=np.packbits(pats.astype(bool), axis=-1), pattern_shape=np.array(pats.shape))
    name = f"fpin_g{grid[0]}x{grid[1]}x{grid[2]}"
    np.savez_compressed(OUT / f"{name}.npz", **arrs)
    meta = dict(grid=list(grid), state=f"var_g{grid[0]}x{grid[1]}x{grid[2]} iterate", omp_threads=1,
                reference_calls=["Filter(crit, 1, 0) for each listed crit"], runs=len(crits))
    return name, meta


def traj_fixture(spec, keep_state: bool):
    grid = spec["grid"]
    cfg = ChannelConfig(X_out=grid[0], Y_out=grid[1], Z_out=grid[2])
    exe = ref_driver(grid)
    with tempfile.TemporaryDirectory() as td:
        d = Path(td)
        write_inputs(d, cfg, spec["seed"], iterate=False)
        subprocess.run([str(exe), "traj", str(d), repr(spec["deltaT"]), str(spec["steps"]),
                        str(spec["iteration_end_step"])], check=True, stdout=subprocess.DEVNULL)
        er = np.loadtxt(d / "er.txt")
        les = np.fromfile(d / f"LES{spec['steps']:05d}.bin").reshape(10, *grid)
    arrs = dict(er=er)
    if keep_state:
        arrs["les"] = les
    else:
        # per-variable sums and one x-plane of both levels
        arrs["les_sum"] = les.reshape(10, -1).sum(axis=1)
        arrs["les_plane"] = les[:, grid[0] // 2].copy()
    name = f"traj_g{grid[0]}x{grid[1]}x{grid[2]}"
    np.savez_compressed(OUT / f"{name}.npz", **arrs)
    meta = dict(spec, grid=list(grid), gamma_metric=1.5, f_drive=599.29876710,
                driver="main.cpp:143-498 with Viscous_terms + Dual_time_stepping enabled")
    return name, meta


def stat_fixture(grid, seed):
    """Statistic's mean profiles (st%05d.dat, Statistic.cpp:1043-1154) from
    the reference: the variant fixture's post-BC iterate (step 1) and
    level-n state (step 2) averaged over statistic_step = 2, then the
    iterate alone with statistic_step = 1.  The files are kept as text."""
    cfg = ChannelConfig(X_out=grid[0], Y_out=grid[1], Z_out=grid[2])
    exe = ref_driver(grid)
    with tempfile.TemporaryDirectory() as td:
        d = Path(td)
        Qn, Qm = perturbed_state(cfg, seed, 0), perturbed_state(cfg, seed, 1)
        write_initial_bin(d / "initial.bin", Qn, Qm)
        np.ascontiguousarray(perturbed_state(cfg, seed, 2)).tofile(d / "iterate.bin")
        subprocess.run([str(exe), "stat", str(d), repr(DT_UNIT)], check=True, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
        arrs = {f"st{n}": np.frombuffer((d / f"st0000{n}.dat").read_bytes(), dtype=np.uint8) for n in (2, 3)}
    meta = dict(grid=list(grid), seed=seed, inputs=f"var_g{grid[0]}x{grid[1]}x{grid[2]} U_bc (iterate) and U",
                reference_calls=["Statistic(step 1, statistic_step 2, iterate)",
                                 "Statistic(step 2, statistic_step 2, level n) -> st00002.dat",
                                 "Statistic(step 3, statistic_step 1, iterate) -> st00003.dat"])
    name = f"stat_g{grid[0]}x{grid[1]}x{grid[2]}"
    np.savez_compressed(OUT / f"{name}.npz", **arrs)
    return name, meta


def force_fixture(grid, seed):
    """Driven_force's printed lines (Driven_force.cpp:72) for the variant
    fixture's post-BC iterate and level-n state."""
    cfg = ChannelConfig(X_out=grid[0], Y_out=grid[1], Z_out=grid[2])
    exe = ref_driver(grid)
    with tempfile.TemporaryDirectory() as td:
        d = Path(td)
        Qn, Qm = perturbed_state(cfg, seed, 0), perturbed_state(cfg, seed, 1)
        write_initial_bin(d / "initial.bin", Qn, Qm)
        np.ascontiguousarray(perturbed_state(cfg, seed, 2)).tofile(d / "iterate.bin")
        out = subprocess.run([str(exe), "force", str(d), repr(DT_UNIT)], check=True, capture_output=True,
                             env=dict(os.environ, OMP_NUM_THREADS="1")).stdout
    lines = [ln for ln in out.decode().splitlines(keepends=True) if ln.startswith("Q0=")]
    assert len(lines) == 2, out
    meta = dict(grid=list(grid), seed=seed, inputs=f"var_g{grid[0]}x{grid[1]}x{grid[2]} U_bc (iterate) and U",
                reference_calls=["Driven_force(iterate)", "Driven_force(level n)"])
    name = f"force_g{grid[0]}x{grid[1]}x{grid[2]}"
    np.savez_compressed(OUT / f"{name}.npz", lines=np.frombuffer("".join(lines).encode(), dtype=np.uint8))
    return name, meta


def shear_fixture(grid, seed):
    """Statistic's Shear_stress.dat lines (Statistic.cpp:552-911: the mean
    y-face viscous x-momentum flux on each wall, written per step as the
    change of its running average) over four steps on post-BC iterates
    (iterate, iterate2, iterate, iterate2) with statistic_step 3, so that the
    fourth line follows the reset of Statistic.cpp:1008-1018."""
    cfg = ChannelConfig(X_out=grid[0], Y_out=grid[1], Z_out=grid[2])
    exe = ref_driver(grid)
    with tempfile.TemporaryDirectory() as td:
        d = Path(td)
        Qn, Qm = perturbed_state(cfg, seed, 0), perturbed_state(cfg, seed, 1)
        write_initial_bin(d / "initial.bin", Qn, Qm)
        np.ascontiguousarray(perturbed_state(cfg, seed, 2)).tofile(d / "iterate.bin")
        np.ascontiguousarray(perturbed_state(cfg, seed, 3)).tofile(d / "iterate2.bin")
        subprocess.run([str(exe), "shear", str(d), repr(DT_UNIT)], check=True, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
        text = (d / "Shear_stress.dat").read_bytes()
    assert text.count(b"\n") == 4, text
    meta = dict(grid=list(grid), seed=seed,
                inputs="perturbed_state(seed, 2) and (seed, 3) as iterates, post-BC, steps 1-4",
                reference_calls=["Statistic(step 1..4, statistic_step 3) -> Shear_stress.dat"])
    name = f"shear_g{grid[0]}x{grid[1]}x{grid[2]}"
    np.savez_compressed(OUT / f"{name}.npz", lines=np.frombuffer(text, dtype=np.uint8))
    return name, meta


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "shear":
        manifest = json.loads((OUT / "manifest.json").read_text())
        for i, g in enumerate(VAR_GRIDS):
            n, m = shear_fixture(g, seed=21 + i)
            manifest["fixtures"][n] = m
        (OUT / "manifest.json").write_text(json.dumps(manifest, indent=1))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "variants":
        # regenerate only the variant fixtures and merge them into the manifest
        manifest = json.loads((OUT / "manifest.json").read_text())
        for i, g in enumerate(VAR_GRIDS):
            n, m = variant_fixture(g, seed=21 + i)
            manifest["fixtures"][n] = m
        (OUT / "manifest.json").write_text(json.dumps(manifest, indent=1))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "force":
        manifest = json.loads((OUT / "manifest.json").read_text())
        for i, g in enumerate(VAR_GRIDS):
            n, m = force_fixture(g, seed=21 + i)
            manifest["fixtures"][n] = m
        (OUT / "manifest.json").write_text(json.dumps(manifest, indent=1))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "stat":
        manifest = json.loads((OUT / "manifest.json").read_text())
        for i, g in enumerate(VAR_GRIDS):
            n, m = stat_fixture(g, seed=21 + i)
            manifest["fixtures"][n] = m
        (OUT / "manifest.json").write_text(json.dumps(manifest, indent=1))
        return
    if len(sys.argv) > 1 and sys.argv[1] == "filterpin":
        manifest = json.loads((OUT / "manifest.json").read_text())
        for g in VAR_GRIDS:
            n, m = filter_pin_fixture(g)
            manifest["fixtures"][n] = m
        (OUT / "manifest.json").write_text(json.dumps(manifest, indent=1))
        return
    manifest = {"generator": "tests/golden/make_golden.py",
                "reference": "ricklee209/Channel_Flow built by oracle/ref_build.sh (np=1, MPICH 3.3.2)",
                "fixtures": {}}
    for i, g in enumerate(UNIT_GRIDS):
        n, m = unit_fixture(g, seed=11 + i)
        manifest["fixtures"][n] = m
    for i, g in enumerate(VAR_GRIDS):
        n, m = variant_fixture(g, seed=21 + i)
        manifest["fixtures"][n] = m
    for g in VAR_GRIDS:
        n, m = filter_pin_fixture(g)
        manifest["fixtures"][n] = m
    for i, g in enumerate(VAR_GRIDS):
        n, m = stat_fixture(g, seed=21 + i)
        manifest["fixtures"][n] = m
        n, m = force_fixture(g, seed=21 + i)
        manifest["fixtures"][n] = m
        n, m = shear_fixture(g, seed=21 + i)
        manifest["fixtures"][n] = m
    n, m = traj_fixture(TRAJ, keep_state=True)
    manifest["fixtures"][n] = m
    n, m = traj_fixture(TRAJ64, keep_state=False)
    manifest["fixtures"][n] = m
    (OUT / "manifest.json").write_text(json.dumps(manifest, indent=1))
    for f in sorted(OUT.glob("*.npz")):
        print(f"{f.name}: {f.stat().st_size / 1e6:.2f} MB")


if __name__ == "__main__":
    main()
