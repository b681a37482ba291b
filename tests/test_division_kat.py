"""The sweep kernels divide by the grid spacings with a residual-corrected
reciprocal (Markstein, IBM J. Res. Dev. 34(1), 1990):

    q0 = x * r,  r = RN(1/d);   q = fma(fma(-q0, d, x), r, q0)

which must equal the IEEE quotient x/d bit for bit for the reference's
division `(...)/deltaXI` to be reproduced.  This checks the identity with C's
correctly rounded fma() on the spacings of every grid used here, over operands
spanning 2^-40 .. 2^23 of both signs.
"""
import subprocess
import textwrap

import pytest

SRC = textwrap.dedent(r"""
    #include <math.h>
    #include <stdint.h>
    #include <stdio.h>
    static uint64_t s = 88172645463325252ULL;
    static uint64_t xr(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
    int main(void) {
      const double hi = 0.005, pi = 3.14159265358979323846;
      const int grids[][3] = {{256,192,256},{64,64,64},{12,10,16},{8,12,9},{16,12,16},{512,384,512},{21,13,18}};
      long bad = 0, n = 0;
      for (unsigned g = 0; g < sizeof grids / sizeof grids[0]; g++) {
        double ds[3] = {2 * pi * hi / grids[g][0], 1.0 / grids[g][1], pi * hi / grids[g][2]};
        for (int t = 0; t < 3; t++) {
          double d = ds[t], r = 1.0 / d;
          for (long i = 0; i < 1000000; i++) {
            uint64_t u = xr();
            double m = (double)(u >> 11) * (1.0 / 9007199254740992.0) + 0.5;
            int ex = (int)((xr() >> 58) & 63) - 40;
            double x = ldexp(m, ex) * ((u & 1) ? -1 : 1);
            double q0 = x * r, e = fma(-q0, d, x), q = fma(e, r, q0);
            n++;
            if (q != x / d) bad++;
          }
        }
      }
      printf("%ld %ld\n", n, bad);
      return bad != 0;
    }
""")


def test_residual_corrected_division_is_exact(tmp_path):
    c = tmp_path / "kat.c"
    c.write_text(SRC)
    exe = tmp_path / "kat"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(c), "-lm"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    n, bad = map(int, r.stdout.split())
    assert n == 21_000_000 and bad == 0, r.stdout
