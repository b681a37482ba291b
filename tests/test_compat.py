"""The C++ compatibility header (include/channel_flow_compat.hpp) keeps the
reference's function names and parameter lists: a driver that takes the
address of every one of them compiles and links against libchannel_flow.so.
Its host-only parts run here without a GPU: Hyperplane reproduces the
reference's plane table (golden fixture of Hyperplane.cpp:50-78), and
Initial_condition / Output read and write the initial.bin / LES%05d.bin
layout (Initial_condition.cpp:64-116, Output.cpp:56-190)."""
import subprocess
import textwrap
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "channel_flow_amd" / "libchannel_flow.so"

DRIVER = textwrap.dedent(r"""
    #include <cstdio>
    #include <cstring>
    #include <vector>
    #define Y_m (YOUT + 4)
    #define Z_m (ZOUT + 4)
    #include "channel_flow_compat.hpp"

    template <class F> static void keep(F f) { volatile const void* p = (const void*)f; (void)p; }

    int main(int argc, char** argv) {
      keep(&Jacobian); keep(&Jacobian_u); keep(&Jacobian_v); keep(&Jacobian_w); keep(&Hyperplane);
      keep(&Y_boundary_condition); keep(&Z_boundary_condition); keep(&Flux_X); keep(&Flux_Y); keep(&Flux_Z);
      keep(&Flux_X_limiter); keep(&Flux_Y_limiter); keep(&Flux_Z_limiter); keep(&Grid); keep(&Filter);
      keep(&DPLUSGS); keep(&RK3NODT); keep(&Statistic);
      keep(&Driven_force);
      keep(&Viscous_terms); keep(&Dual_time_stepping); keep(&Initial_condition); keep(&Output);
      cf_params p{};
      p.X_out = XOUT; p.Y_out = YOUT; p.Z_out = ZOUT;
      // host-only functions: the context pointer is bound but never used
      cf_compat::attach(reinterpret_cast<cf_ctx*>(&p), p);
      const int X_np = XOUT + 6, ny = YOUT + 1, nz = ZOUT + 1, gend = XOUT + 2;
      if (argc > 1 && !std::strcmp(argv[1], "hyperplane")) {
        std::vector<int> N((X_np - 3 + ny + nz + 7) * 2);
        std::vector<int> L((size_t)(X_np - 6) * YOUT * ZOUT * 3 + 3);
        Hyperplane(0, (int (*)[2])N.data(), (int (*)[3])L.data());
        for (int m = 7; m <= gend + ny + nz; m++) std::printf("%d %d %d\n", m, N[2 * m], N[2 * m + 1]);
        // every cell once, on its plane
        std::vector<int> seen((size_t)X_np * Y_m * Z_m, 0);
        for (int q = 1; q <= N[2 * (gend + ny + nz)]; q++) {
          const int i = L[3 * q], j = L[3 * q + 1], k = L[3 * q + 2];
          seen[((size_t)i * Y_m + j) * Z_m + k]++;
        }
        long bad = 0;
        for (int i = 3; i <= gend; i++)
          for (int j = 2; j <= ny; j++)
            for (int k = 2; k <= nz; k++) bad += seen[((size_t)i * Y_m + j) * Z_m + k] != 1;
        std::printf("bad %ld\n", bad);
        return 0;
      }
      // initial.bin -> arrays -> LES00007.bin
      std::vector<std::vector<double>> a(16, std::vector<double>((size_t)X_np * Y_m * Z_m, 0.0));
      auto F = [&](int v) { return (cf_compat::field)a[v].data(); };
      for (size_t c = 0; c < a[15].size(); c++) a[15][c] = 1.0 + 0.001 * (double)(c % 97);
      Initial_condition(0, F(0), F(1), F(2), F(3), F(4), F(5), F(6), F(7), F(8), F(9), F(10), F(11), F(12), F(13),
                        F(14), F(15));
      for (int v = 0; v < 5; v++)
        if (a[v] != a[5 + v]) { std::printf("iterate differs from level n\n"); return 1; }
      Output(7, 0, F(5), F(6), F(7), F(8), F(9), F(10), F(11), F(12), F(13), F(14), F(15));
      std::printf("ok\n");
      return 0;
    }
""")


def build_driver(tmp_path, grid):
    if not LIB.exists():
        pytest.fail(f"{LIB} not built (python -m channel_flow_amd.build)")
    src = tmp_path / "driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "driver"
    X, Y, Z = grid
    subprocess.run(["g++", "-std=c++17", "-O1", f"-DXOUT={X}", f"-DYOUT={Y}", f"-DZOUT={Z}",
                    f"-I{ROOT / 'include'}", str(src), "-o", str(exe), f"-L{LIB.parent}", "-lchannel_flow",
                    f"-Wl,-rpath,{LIB.parent}"], check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("grid", [(12, 10, 16), (8, 12, 9)])
def test_compat_hyperplane_matches_reference(tmp_path, grid, golden):
    exe = build_driver(tmp_path, grid)
    out = subprocess.run([str(exe), "hyperplane"], capture_output=True, text=True, check=True).stdout.split("\n")
    rows = np.array([list(map(int, ln.split())) for ln in out if ln and not ln.startswith("bad")])
    assert out[-2] == "bad 0"
    hp = golden("unit_g%dx%dx%d" % grid)["hyperplane"]
    np.testing.assert_array_equal(rows, hp)


def test_compat_initial_and_output_layout(tmp_path):
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state, write_initial_bin
    grid = (9, 8, 10)
    exe = build_driver(tmp_path, grid)
    cfg = ChannelConfig(*grid)
    Qn = poiseuille_state(cfg, FlowParams(), seed=3, stream=0)
    Qm = poiseuille_state(cfg, FlowParams(), seed=3, stream=1)
    write_initial_bin(tmp_path / "initial.bin", Qn, Qm)
    r = subprocess.run([str(exe)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
    les = np.fromfile(tmp_path / "LES00007.bin", dtype="<f8").reshape(10, *grid)
    # the driver's J: 1 + 0.001 (c mod 97) over the padded [X_np][Y_m][Z_m] index c
    X, Y, Z = grid
    c = np.arange((X + 6) * (Y + 4) * (Z + 4)).reshape(X + 6, Y + 4, Z + 4)
    J = (1.0 + 0.001 * (c % 97))[3:3 + X, 2:2 + Y, 2:2 + Z]
    q = np.concatenate([np.asarray(Qn), np.asarray(Qm)]).reshape(10, *grid)
    # Output stores (Q/J)*J, the round trip of the reference's own arithmetic
    np.testing.assert_array_equal(les, (q / J) * J)
