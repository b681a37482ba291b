#!/usr/bin/env bash
# Build the reference solver (ricklee209/Channel_Flow) from its own sources,
# where they lie, into oracle/_ref/g<X>x<Y>x<Z>/ref_driver.
#
# Test infrastructure only: the binary is the golden-vector generator and the
# timed CPU baseline; nothing in the product links or runs it.
#
# The reference is configured the way its authors configure it: grid sizes in
# Resolution.h and the rank count in MPI_prm.h.  Both are generated into
# oracle/_ref/<grid>/cfg/, which `-I-` puts ahead of the reference directory
# for "quoted" includes; every .cpp and every other header is read from the
# reference tree unchanged.  MPI is the MPICH 3.3.2 that ships in /opt/conda
# (real mpi.h + libmpi.so); the reference runs as a single rank (np=1).
#
# usage: oracle/ref_build.sh X_out Y_out Z_out [refdir]
set -euo pipefail
X=${1:?X_out}; Y=${2:?Y_out}; Z=${3:?Z_out}
REF=${4:-/root/reference}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref/g${X}x${Y}x${Z}"
MPI_PREFIX=${MPI_PREFIX:-/opt/conda}
[ -f "$REF/Flux_X.cpp" ] || { echo "reference sources not found at $REF" >&2; exit 2; }
[ -f "$MPI_PREFIX/include/mpi.h" ] || { echo "MPICH not found under $MPI_PREFIX" >&2; exit 2; }
mkdir -p "$OUT/cfg" "$OUT/obj"

# Domain of the reference run: Lx = 2*pi*high, Lz = pi*high with high = 0.005
# (Resolution.h:13,15 give deltaXI = Lx/96, deltaZT = Lz/72); eta in [0,1].
python3 - "$X" "$Y" "$Z" > "$OUT/cfg/Resolution.h" <<'EOF'
import math, sys
X, Y, Z = map(int, sys.argv[1:4])
high = 0.005
print(f"#define X_out {X}\n#define Y_out {Y}\n#define Z_out {Z}\n")
print(f"#define X_m {X+4}\n#define Y_m {Y+4}\n#define Z_m {Z+4}\n")
# the macros stay unparenthesised, as the reference's own expressions expect
print("#define nx X_out+1\n#define ny Y_out+1\n#define nz Z_out+1\n")
print(f"#define deltaXI {2*math.pi*high/X!r}\n#define deltaET {1.0/Y!r}\n#define deltaZT {math.pi*high/Z!r}\n")
print("#define nxx nx+1\n#define nyy ny+1\n#define nzz nz+1\n")
print("#define nxxx nxx+1\n#define nyyy nyy+1\n#define nzzz nzz+1\n")
EOF
# single rank: the only edit is the rank count on MPI_prm.h:5
sed 's/^#define np [0-9]*/#define np 1/' "$REF/MPI_prm.h" > "$OUT/cfg/MPI_prm.h"
grep -q '^#define np 1' "$OUT/cfg/MPI_prm.h"

CXX=${CXX:-g++}
CXXFLAGS="-O2 -fopenmp -w"
INC="-I$OUT/cfg -I- -I$REF -I$MPI_PREFIX/include"
SRCS="Flux_X Flux_Y Flux_Z Flux_X_limiter Flux_Y_limiter Flux_Z_limiter Filter
      Viscous_terms Dual_time_stepping DPLUSGS RK3NODT Y_boundary_condition
      Z_boundary_condition Jacobian Jacobian_u Jacobian_v Jacobian_w Hyperplane
      Grid Initial_condition Output Statistic Driven_force"
pids=()
for s in $SRCS; do
  # stdio.h is force-included: Flux_Y_limiter.cpp:171 and Driven_force.cpp:73
  # call printf without including it (MPICH's mpi.h does not pull it in)
  $CXX $CXXFLAGS -include stdio.h $INC -c "$REF/$s.cpp" -o "$OUT/obj/$s.o" 2> "$OUT/obj/$s.err" &
  pids+=($!)
done
for p in "${pids[@]}"; do wait "$p"; done
$CXX $CXXFLAGS -include stdio.h $INC -c "$HERE/ref_driver.cpp" -o "$OUT/obj/ref_driver.o"
$CXX $CXXFLAGS -o "$OUT/ref_driver" "$OUT"/obj/*.o -L"$MPI_PREFIX/lib" -Wl,-rpath,"$MPI_PREFIX/lib" -lmpi
echo "$OUT/ref_driver"
